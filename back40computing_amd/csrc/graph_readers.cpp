// graph_readers.cpp -- DIMACS, METIS and Matrix Market file readers
// (C ABI in include/b40c_graph.h, b40c_read_*).
//
// Each reader returns the file's directed (src, dst) pairs, 0-based, in file
// order; b40c_graph_from_coo or b40c_device_csr_from_coo turns them into a
// CSR.  The accepted grammar, the 1-based ids, the count checks and the
// undirected doubling follow the reference's readers
// (b40c/graph/builder/dimacs.cuh:53-225, metis.cuh:51-176,
// market.cuh:49-164); the parsing itself is a single pass over the file
// loaded into memory.
#include "b40c_graph.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// Whole file in memory, NUL-terminated.
int slurp(const char *path, std::vector<char> *buf)
{
    if (!path) return 1;
    FILE *f = std::fopen(path, "rb");
    if (!f) return 5;
    std::vector<char> &b = *buf;
    b.clear();
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) b.insert(b.end(), chunk, chunk + got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return 5;
    b.push_back('\0');
    return 0;
}

// Cursor over one line: integers separated by blanks.
struct Line {
    const char *p, *end;
    void skip_blanks()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
    }
    bool done()
    {
        skip_blanks();
        return p >= end;
    }
    bool integer(long long *out)
    {
        skip_blanks();
        bool neg = false;
        if (p < end && (*p == '-' || *p == '+')) neg = *p++ == '-';
        if (p >= end || *p < '0' || *p > '9') return false;
        long long v = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            if (v > (INT64_MAX - 9) / 10) return false;
            v = v * 10 + (*p++ - '0');
        }
        // a number ends at a blank or the end of the line
        if (p < end && *p != ' ' && *p != '\t' && *p != '\r') return false;
        *out = neg ? -v : v;
        return true;
    }
    bool word(std::string *out)
    {
        skip_blanks();
        const char *b = p;
        while (p < end && *p != ' ' && *p != '\t' && *p != '\r') p++;
        out->assign(b, p);
        return p > b;
    }
};

// Calls fn(line) for every line of the buffer (without its newline).
template <typename Fn>
int for_each_line(const std::vector<char> &buf, Fn &&fn)
{
    const char *p = buf.data(), *end = buf.data() + buf.size() - 1;
    while (p < end) {
        const char *nl = static_cast<const char *>(std::memchr(p, '\n', (size_t) (end - p)));
        const char *le = nl ? nl : end;
        Line l{p, le};
        int rc = fn(l);
        if (rc) return rc;
        p = nl ? nl + 1 : end;
    }
    return 0;
}

int finish(int32_t nodes, std::vector<int32_t> &src, std::vector<int32_t> &dst, b40c_coo *out)
{
    out->nodes = nodes;
    out->pairs = (int64_t) src.size();
    const size_t n = std::max<size_t>(src.size(), 1);
    out->src = (int32_t *) std::malloc(sizeof(int32_t) * n);
    out->dst = (int32_t *) std::malloc(sizeof(int32_t) * n);
    if (!out->src || !out->dst) {
        b40c_coo_free(out);
        return 2;
    }
    if (!src.empty()) {
        std::memcpy(out->src, src.data(), sizeof(int32_t) * src.size());
        std::memcpy(out->dst, dst.data(), sizeof(int32_t) * dst.size());
    }
    return 0;
}

bool vertex_id(long long one_based, int32_t nodes, int32_t *out)
{
    if (one_based < 1 || one_based > nodes) return false;
    *out = (int32_t) (one_based - 1);
    return true;
}

}  // namespace

extern "C" {

// DIMACS shortest-path format (dimacs.cuh:53-225): "p <type> <n> <m>" once,
// then m arc lines "a <u> <v> [<w>]"; other lines (comments "c ...") are
// skipped.  With undirected, each arc also contributes its reverse, placed
// before it as the reference inserts it.  Exactly m arcs are required.
int b40c_read_dimacs(const char *path, int undirected, b40c_coo *out)
{
    if (!out) return 1;
    *out = b40c_coo{};
    std::vector<char> buf;
    if (int rc = slurp(path, &buf)) return rc;
    long long nodes = -1, arcs = -1;
    std::vector<int32_t> src, dst;
    int rc = for_each_line(buf, [&](Line &l) -> int {
        if (l.p >= l.end) return 0;
        const char tag = *l.p;
        if (tag == 'p') {
            l.p++;
            std::string type;
            if (nodes >= 0 || !l.word(&type) || !l.integer(&nodes) || !l.integer(&arcs)) return 1;
            if (nodes < 1 || nodes > INT32_MAX || arcs < 0 || (undirected ? 2 : 1) * arcs > INT32_MAX) return 1;
            src.reserve((size_t) arcs * (undirected ? 2 : 1));
            dst.reserve((size_t) arcs * (undirected ? 2 : 1));
        } else if (tag == 'a') {
            l.p++;
            long long u, v, w;
            int32_t a, b;
            if (nodes < 0 || !l.integer(&u) || !l.integer(&v)) return 1;
            if (!l.done() && !l.integer(&w)) return 1;
            if (!vertex_id(u, (int32_t) nodes, &a) || !vertex_id(v, (int32_t) nodes, &b)) return 1;
            if ((long long) src.size() >= (undirected ? 2 : 1) * arcs) return 1;  // more arcs than declared
            if (undirected) {
                src.push_back(b);
                dst.push_back(a);
            }
            src.push_back(a);
            dst.push_back(b);
        }
        return 0;
    });
    if (rc) return rc;
    if (nodes < 0 || (long long) src.size() != (undirected ? 2 : 1) * arcs) return 1;
    return finish((int32_t) nodes, src, dst, out);
}

// METIS graph format (metis.cuh:51-176): a header "<n> <m> [<fmt>]", then
// line k (k = 1..n) lists the 1-based neighbours of vertex k - 1; lines
// starting with '%' are comments.  The file holds each undirected edge in
// both rows, 2m entries in all; a file whose header counts all m entries is
// also accepted, as by the reference.  Weighted formats (fmt != 0) are
// rejected.
int b40c_read_metis(const char *path, b40c_coo *out)
{
    if (!out) return 1;
    *out = b40c_coo{};
    std::vector<char> buf;
    if (int rc = slurp(path, &buf)) return rc;
    long long nodes = -1, edges = -1;
    long long row = -1;  // vertex of the current line
    std::vector<int32_t> src, dst;
    int rc = for_each_line(buf, [&](Line &l) -> int {
        Line peek = l;
        peek.skip_blanks();
        if (peek.p < peek.end && *peek.p == '%') return 0;
        if (nodes < 0) {
            if (peek.p >= peek.end) return 0;  // blank lines before the header
            long long fmt = 0;
            if (!l.integer(&nodes) || !l.integer(&edges)) return 1;
            if (!l.done() && (!l.integer(&fmt) || fmt != 0)) return 1;
            if (nodes < 1 || nodes > INT32_MAX || edges < 0 || 2 * edges > INT32_MAX) return 1;
            src.reserve((size_t) (2 * edges));
            dst.reserve((size_t) (2 * edges));
            row = 0;
            return 0;
        }
        if (row >= nodes) return l.done() ? 0 : 1;  // trailing blank lines only
        long long v;
        int32_t b;
        while (!l.done()) {
            if (!l.integer(&v) || !vertex_id(v, (int32_t) nodes, &b)) return 1;
            if ((long long) src.size() >= 2 * edges) return 1;
            src.push_back((int32_t) row);
            dst.push_back(b);
        }
        row++;
        return 0;
    });
    if (rc) return rc;
    if (nodes < 0) return 1;
    const long long got = (long long) src.size();
    if (got != 2 * edges && got != edges) return 1;
    return finish((int32_t) nodes, src, dst, out);
}

// Matrix Market coordinate format (market.cuh:49-164): '%' lines (the
// banner included) are comments; "<rows> <cols> <nnz>" with rows == cols;
// then nnz lines "<i> <j> [value]".  As in the reference, an entry (i, j)
// becomes the arc j -> i (its first number is read as the column), and with
// undirected each entry adds i -> j right after it.
int b40c_read_market(const char *path, int undirected, b40c_coo *out)
{
    if (!out) return 1;
    *out = b40c_coo{};
    std::vector<char> buf;
    if (int rc = slurp(path, &buf)) return rc;
    long long nodes = -1, nnz = -1;
    std::vector<int32_t> src, dst;
    int rc = for_each_line(buf, [&](Line &l) -> int {
        Line peek = l;
        peek.skip_blanks();
        if (peek.p >= peek.end || *peek.p == '%') return 0;
        if (nodes < 0) {
            long long cols;
            if (!l.integer(&nodes) || !l.integer(&cols) || !l.integer(&nnz) || !l.done()) return 1;
            if (nodes != cols || nodes < 1 || nodes > INT32_MAX || nnz < 0 || (undirected ? 2 : 1) * nnz > INT32_MAX)
                return 1;
            src.reserve((size_t) nnz * (undirected ? 2 : 1));
            dst.reserve((size_t) nnz * (undirected ? 2 : 1));
            return 0;
        }
        long long i, j;
        int32_t a, b;
        if (!l.integer(&i) || !l.integer(&j)) return 1;
        if (!vertex_id(i, (int32_t) nodes, &a) || !vertex_id(j, (int32_t) nodes, &b)) return 1;
        if ((long long) src.size() >= (undirected ? 2 : 1) * nnz) return 1;
        src.push_back(b);
        dst.push_back(a);
        if (undirected) {
            src.push_back(a);
            dst.push_back(b);
        }
        return 0;
    });
    if (rc) return rc;
    if (nodes < 0 || (long long) src.size() != (undirected ? 2 : 1) * nnz) return 1;
    return finish((int32_t) nodes, src, dst, out);
}

void b40c_coo_free(b40c_coo *c)
{
    if (!c) return;
    std::free(c->src);
    std::free(c->dst);
    c->src = c->dst = nullptr;
    c->pairs = 0;
}

}  // extern "C"
