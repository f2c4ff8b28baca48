// graph_builder_internal.hpp -- host helpers of graph_builder.cpp shared
// with the device CSR builder (b40c_csr_build.hip).  Not installed.
#pragma once

#include <stdint.h>

#include <vector>

// The seeded vertex permutation that scrambles R-MAT ids (b40c_graph_rmat).
void b40c_rmat_permutation(int32_t scale, uint64_t seed, std::vector<int32_t> *perm);
