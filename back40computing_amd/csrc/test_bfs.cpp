// test_bfs.cpp -- command-line BFS driver over libb40c_hip.so.
//
// Mirrors the reference driver (test/bfs/test_bfs.cu:97-190 usage, 584-783
// RunTests, 903-1137 main): build a synthetic graph, run the requested
// strategies from a fixed or randomized source, validate every GPU result
// against a serial CPU traversal, and print per-iteration and running
// statistics.  Differences from the reference: validation failures set a
// nonzero exit status, graph builders follow include/b40c_graph.h (no
// self-loops in lattices; Graph500-style R-MAT), and the CPU time is
// wall-clock.
#include "b40c_bfs.hpp"
#include "b40c_graph.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace b40c::graph::bfs;

namespace {

struct Args {
    std::vector<std::string> pos;
    std::vector<std::pair<std::string, std::string>> kv;
    Args(int argc, char **argv)
    {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a.rfind("--", 0) == 0) {
                size_t eq = a.find('=');
                if (eq == std::string::npos) kv.emplace_back(a.substr(2), "");
                else kv.emplace_back(a.substr(2, eq - 2), a.substr(eq + 1));
            } else {
                pos.push_back(a);
            }
        }
    }
    bool flag(const char *k) const
    {
        for (auto &p : kv)
            if (p.first == k) return true;
        return false;
    }
    std::string get(const char *k, const std::string &def = "") const
    {
        for (auto &p : kv)
            if (p.first == k) return p.second;
        return def;
    }
};

// Welford running statistic (test_bfs.cu:214-235).
struct Statistic {
    double mean = 0, m2 = 0;
    int count = 0;
    double Update(double x)
    {
        count++;
        double d = x - mean;
        mean += d / count;
        m2 += d * (x - mean);
        return count > 1 ? m2 / (count - 1) : 0.0;
    }
};

// Serial FIFO traversal used by the driver to validate GPU results, as the
// reference driver does with its own CPU BFS (test_bfs.cu:514-554).
double cpu_bfs(const b40c_csr &g, int src, std::vector<int> &labels)
{
    labels.assign(g.nodes, -1);
    std::vector<int> queue(g.nodes);
    size_t head = 0, tail = 0;
    labels[src] = 0;
    queue[tail++] = src;
    auto t0 = std::chrono::steady_clock::now();
    while (head < tail) {
        int u = queue[head++];
        int d = labels[u] + 1;
        for (int e = g.row_offsets[u]; e < g.row_offsets[u + 1]; e++) {
            int v = g.column_indices[e];
            if (labels[v] == -1) {
                labels[v] = d;
                queue[tail++] = v;
            }
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

struct SearchResult {
    const char *name = nullptr;  // null: unknown strategy
    bfs_status status = BFS_OK;
    long long total_queued = 0;
    int search_depth = 0;
    double avg_duty = 0, ms = 0;
};

// One search with strategy s (test_bfs.cu:687-729): reset the problem for the
// enactor's frontier type, enact, and collect the enactor's statistics.
template <bool INSTRUMENT, typename Problem>
SearchResult run_strategy(Problem &problem, int s, int src, bool iterative, int max_ctas, double qs, bool verbose)
{
    SearchResult r;
    EnactorExpandContract<INSTRUMENT> ec(verbose);
    EnactorContractExpand<INSTRUMENT> ce(verbose);
    EnactorTwoPhase<INSTRUMENT> tp(verbose);
    EnactorHybrid<INSTRUMENT> hy(verbose);
    EnactorDirectionOptimizing<INSTRUMENT> dopt(verbose);
    EnactorBase *base = nullptr;
    switch (s) {
    case 0:
        r.name = "Expand-contract GPU BFS";
        if ((r.status = problem.Reset(ec.GetFrontierType(), qs))) return r;
        r.status = iterative ? ec.EnactIterativeSearch(problem, src, max_ctas) : ec.EnactFusedSearch(problem, src, max_ctas);
        base = &ec;
        break;
    case 1:
        r.name = "Contract-expand GPU BFS";
        if ((r.status = problem.Reset(ce.GetFrontierType(), qs))) return r;
        r.status = iterative ? ce.EnactIterativeSearch(problem, src, max_ctas) : ce.EnactFusedSearch(problem, src, max_ctas);
        base = &ce;
        break;
    case 2:
        r.name = "Two-phase GPU BFS";
        if ((r.status = problem.Reset(tp.GetFrontierType(), qs))) return r;
        r.status = iterative ? tp.EnactIterativeSearch(problem, src, max_ctas) : tp.EnactFusedSearch(problem, src, max_ctas);
        base = &tp;
        break;
    case 3:
        r.name = "Hybrid contract-expand + two-phase GPU BFS";
        if ((r.status = problem.Reset(hy.GetFrontierType(), qs))) return r;
        r.status = hy.EnactSearch(problem, src, max_ctas);
        base = &hy;
        break;
    case 4:
        r.name = "Direction-optimizing GPU BFS";
        if ((r.status = problem.Reset(dopt.GetFrontierType(), qs))) return r;
        r.status = iterative ? dopt.EnactIterativeSearch(problem, src, max_ctas) : dopt.EnactFusedSearch(problem, src, max_ctas);
        base = &dopt;
        break;
    default: return r;
    }
    if (r.status == BFS_OK) {
        base->GetStatistics(r.total_queued, r.search_depth, r.avg_duty);
        r.ms = base->ElapsedMillis();
    }
    return r;
}

void usage()
{
    std::printf(
        "\ntest_bfs <graph type> <graph type args> [--device=<i>] [--i=<iterations>]\n"
        "         [--src=<id|randomize>] [--queue-sizing=<f>] [--mark-pred]\n"
        "         [--strategy=<s>[,<s>]*] [--iterative] [--max-ctas=<n>] [--instrumented]\n"
        "         [--quick] [--v]\n\n"
        "Graph types: grid2d <width> | grid3d <width> | rmat <scale> <edge-factor> |\n"
        "             er <n> <m> | random <n> <m> [--undirected]\n"
        "Strategies: 0 expand-contract, 1 contract-expand, 2 two-phase, 3 hybrid (default 3),\n"
        "            4 direction-optimizing (addition)\n\n");
}

}  // namespace

int main(int argc, char **argv)
{
    Args args(argc, argv);
    if (args.pos.empty() || args.flag("help")) {
        usage();
        return 1;
    }
    const std::string type = args.pos[0];
    b40c_csr g{};
    int rc = 1;
    auto arg = [&](size_t i) { return i < args.pos.size() ? std::atoll(args.pos[i].c_str()) : -1LL; };
    if (type == "grid2d") rc = b40c_graph_grid2d((int) arg(1), &g);
    else if (type == "grid3d") rc = b40c_graph_grid3d((int) arg(1), &g);
    else if (type == "rmat") rc = b40c_graph_rmat((int) arg(1), (int) arg(2), 0.57, 0.19, 0.19, 0x5EED0022ull, 1, &g);
    else if (type == "er") rc = b40c_graph_erdos_renyi((int) arg(1), arg(2), 0xE5D05ull, &g);
    else if (type == "random") rc = b40c_graph_random((int) arg(1), arg(2), 0, args.flag("undirected"), &g);
    else {
        std::fprintf(stderr, "Unspecified graph type\n");
        return 1;
    }
    if (rc != 0) {
        std::fprintf(stderr, "graph construction failed (%d)\n", rc);
        return 1;
    }
    std::printf("Graph %s: %d vertices, %lld directed edges\n", type.c_str(), g.nodes, (long long) g.edges);

    const int iterations = std::atoi(args.get("i", "1").c_str());
    const double qs = std::atof(args.get("queue-sizing", "1.3").c_str());
    const int max_ctas = std::atoi(args.get("max-ctas", "0").c_str());
    const bool quick = args.flag("quick"), verbose = args.flag("v"), mark_pred = args.flag("mark-pred");
    const bool iterative = args.flag("iterative");
    const bool instrumented = args.flag("instrumented");
    std::vector<int> strategies;
    {
        std::string s = args.get("strategy", "3");
        size_t p = 0;
        while (p < s.size()) {
            size_t q = s.find(',', p);
            strategies.push_back(std::atoi(s.substr(p, q - p).c_str()));
            if (q == std::string::npos) break;
            p = q + 1;
        }
    }
    std::string src_str = args.get("src", "randomize");
    bool randomize = src_str == "randomize";
    std::vector<int> sources(std::max(iterations + 1, 1));
    if (randomize) {
        if (b40c_graph_sample_sources(&g, (int) sources.size(), 0x50C4CE5ull, sources.data()) != 0) {
            std::fprintf(stderr, "not enough vertices of nonzero degree\n");
            return 1;
        }
    } else {
        for (auto &s : sources) s = std::atoi(src_str.c_str());
    }

    int failures = 0;
    auto run = [&](auto &problem) -> int {
        problem.device = std::atoi(args.get("device", "0").c_str());
        if (problem.FromHostProblem(false, g.nodes, (int) g.edges, g.column_indices, g.row_offsets, 1)) return 1;
        std::vector<int> ref, labels(g.nodes), depths(g.nodes);
        Statistic cpu_rate, rates[5];
        for (int it = 0; it <= iterations; it++) {
            int src = sources[it];
            std::printf("---------------------------------------------------------------\n");
            double cpu_ms = cpu_bfs(g, src, ref);
            long long ev = 0, nv = 0;
            for (int v = 0; v < g.nodes; v++)
                if (ref[v] >= 0) {
                    nv++;
                    ev += g.row_offsets[v + 1] - g.row_offsets[v];
                }
            if (it > 0) cpu_rate.Update(ev / (cpu_ms * 1000.0));
            std::printf("CPU BFS from %d: %.3f ms, %.3f MiEdges/s\n", src, cpu_ms, ev / (cpu_ms * 1000.0));
            for (int s : strategies) {
                SearchResult r = instrumented
                                     ? run_strategy<true>(problem, s, src, iterative, max_ctas, qs, verbose)
                                     : run_strategy<false>(problem, s, src, iterative, max_ctas, qs, verbose);
                if (!r.name) {
                    std::fprintf(stderr, "unknown strategy %d\n", s);
                    return 1;
                }
                if (r.status != BFS_OK) return 1;
                const char *name = r.name;
                const long long queued = r.total_queued;
                const int depth = r.search_depth;
                const double duty = r.avg_duty, ms = r.ms;
                const char *validity = "skipped";
                if (!quick) {
                    if (problem.ExtractResults(labels.data())) return 1;
                    bool ok = true;
                    if (mark_pred) {
                        bfs_extract_depths(problem.handle, depths.data());
                        for (int v = 0; v < g.nodes && ok; v++) {
                            int p = labels[v];
                            if (depths[v] != ref[v]) ok = false;
                            else if (p < 0) ok = (p == -2) == (ref[v] == 0) && (p == -1) == (ref[v] == -1);
                            else {
                                bool found = false;
                                for (int e = g.row_offsets[p]; e < g.row_offsets[p + 1] && !found; e++)
                                    found = g.column_indices[e] == v;
                                ok = found && ref[p] + 1 == ref[v];
                            }
                        }
                    } else {
                        ok = labels == ref;
                    }
                    validity = ok ? "CORRECT" : "INCORRECT";
                    if (!ok) failures++;
                }
                if (it == 0) {
                    std::printf("[%s] warmup iteration: %.3f ms, Validity: %s\n", name, ms, validity);
                } else {
                    double rate = ev / (ms * 1000.0);
                    double sd = std::sqrt(rates[s].Update(rate));
                    // duplicates put through the queues, as a share of what an exact
                    // search queues: the traversed edges for edge frontiers (test_bfs.cu:
                    // 369-373), the visited vertices for vertex frontiers
                    const bool vertex_frontier = s == 0 || s == 4;
                    const double unit = vertex_frontier ? (double) nv : (double) ev;
                    const double redundant = queued > 0 ? 100.0 * ((double) queued - unit) / unit : 0.0;
                    std::printf("[%s] finished. Validity: %s\n  elapsed: %.3f ms, rate: %.3f MiEdges/s, "
                                "search_depth: %d\n", name, validity, ms, rate, depth);
                    if (duty != 0) std::printf("  avg cta duty: %.2f%%\n", duty * 100);
                    std::printf("  src: %d, nodes visited: %lld, edges visited: %lld, total queued: %lld, "
                                "redundant work: %.2f%%\n  [Rate MiEdges/s]: u: %.3f, s: %.3f\n",
                                src, nv, ev, queued, redundant, rates[s].mean, sd);
                }
            }
        }
        if (cpu_rate.count) std::printf("CPU mean rate: %.3f MiEdges/s\n", cpu_rate.mean);
        return 0;
    };

    int status;
    if (mark_pred) {
        CsrProblem<int, int, true> p;
        status = run(p);
    } else {
        CsrProblem<int, int, false> p;
        status = run(p);
    }
    b40c_graph_free(&g);
    if (status) return status;
    if (failures) {
        std::printf("%d validation failure(s)\n", failures);
        return 2;
    }
    return 0;
}
