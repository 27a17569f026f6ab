// sc-bin: the reference's command line (sc-main.cpp:1556-1805) on MI355X.
//
//   sc-bin <graph> <template> <iterations> <threads> <load_binary> <write_binary>
//          [isPruned=1] [useSPMM=1] [vtuneStart=-1] [benchItr=1] [seed] [precision]
//
// The first ten arguments mean what they mean for the reference.  Two are
// added: seed (uint64 color stream; default: time-based, like the
// reference's srand(time(0))) and precision ("fp32" = reference numerics,
// "fp64" = exact counts).  <graph> may also be "rmat:<scale>:<edge_factor>[:seed]"
// for a synthetic R-MAT graph (a=.57 b=.19 c=.19).  The reference's start-up
// memory-bandwidth probe (sc-main.cpp:1606) is not run.
//
// The reference switches its estimate report on by rebuilding (the VERBOSE
// macro, and isEstimate at sc-main.cpp:1581-1584); here the environment does:
//   PG_VERBOSE=1   print the cost-model estimates before counting and the
//                  per-kernel device times and model bandwidth after it
//   PG_ESTIMATE=1  stop after the estimates (compute(t, isEstimate = true))
//   PG_AUTOMORPHISMS=1  divide the final count by the template's automorphism
//                  count (the reference's hard-coded calculate_automorphism,
//                  sc-main.cpp:1568, which it leaves false)
// and one switch has no reference counterpart:
//   PG_AUTOTUNE=1  let the library's autotuner pick kernel variants, degree
//                  buckets and colorings per pass (one-time, before the first
//                  pass of the template)
// isPruned = 0 runs the reference's unpruned mode with its numerics (FP32
// only, slow by design); useSPMM = 0 forms the same sums as useSPMM = 1.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "countmat.hpp"

using namespace pgbsc;

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int load_rmat(const std::string& spec, CSCGraph* g)
{
    int scale = 0, ef = 0;
    unsigned long long seed = 1;
    if (sscanf(spec.c_str(), "rmat:%d:%d:%llu", &scale, &ef, &seed) < 2) return PG_ERR_INVALID;
    int64_t m = 0;
    int rc = pg_rmat_edges(scale, ef, 0.57, 0.19, 0.19, seed, 0, nullptr, nullptr, &m);
    if (rc) return rc;
    std::vector<int32_t> s(m), d(m);
    rc = pg_rmat_edges(scale, ef, 0.57, 0.19, 0.19, seed, m, s.data(), d.data(), &m);
    if (rc) return rc;
    g->createFromEdgeListFile((int32_t)1 << scale, m, s.data(), d.data());
    return PG_OK;
}

int main(int argc, char** argv)
{
    if (argc < 7) {
        fprintf(stderr,
                "usage: %s <graph> <template> <iterations> <threads> <load_binary> <write_binary> "
                "[isPruned=1] [useSPMM=1] [vtuneStart=-1] [benchItr=1] [seed] [fp32|fp64]\n",
                argv[0]);
        return 2;
    }
    const std::string graph_name = argv[1], template_name = argv[2];
    const int iterations = atoi(argv[3]);
    const int comp_thds = atoi(argv[4]);
    const int load_binary = atoi(argv[5]);
    const int write_binary = atoi(argv[6]);
    const int isPruned = argc > 7 ? atoi(argv[7]) : 1;
    const int useSPMM = argc > 8 ? atoi(argv[8]) : 1;
    const int vtuneStart = argc > 9 ? atoi(argv[9]) : -1;
    const uint64_t seed = argc > 11 ? strtoull(argv[11], nullptr, 10) : (uint64_t)time(nullptr);
    int precision = PG_FP32;
    if (argc > 12) {
        if (!strcmp(argv[12], "fp64")) precision = PG_FP64;
        else if (strcmp(argv[12], "fp32")) {
            fprintf(stderr, "precision must be fp32 or fp64\n");
            return 2;
        }
    }
    if (iterations <= 0 || comp_thds <= 0) {
        fprintf(stderr, "iterations and threads must be positive\n");
        return 2;
    }

    printf("Start loading datasets\n");
    fflush(stdout);
    const double t0 = now_s();
    CSCGraph g;
    std::string err;
    int rc;
    if (graph_name.rfind("rmat:", 0) == 0) {
        rc = load_rmat(graph_name, &g);
        if (rc) err = "bad rmat spec " + graph_name;
    } else if (load_binary) {
        rc = g.deserialize(graph_name, &err);
    } else {
        EdgeList el;
        rc = el.read(graph_name, &err);
        if (!rc) g.createFromEdgeListFile(el.getNumVertices(), el.getNumEdges(), el.getSrcList(), el.getDstList());
    }
    if (rc) {
        fprintf(stderr, "error: %s\n", err.c_str());
        return 1;
    }
    if (write_binary) {
        rc = g.serialize("graph.data", &err);
        if (rc) {
            fprintf(stderr, "error: %s\n", err.c_str());
            return 1;
        }
    }
    printf("CSC Format Total vertices is : %d\n", g.getNumVertices());
    printf("Prepare Datasets using %f secs\n", now_s() - t0);
    fflush(stdout);

    Graph tmpl;
    rc = tmpl.read_enlist(template_name, &err);
    if (rc) {
        fprintf(stderr, "error: %s\n", err.c_str());
        return 1;
    }
    CountMat executor;
    const char* autom = getenv("PG_AUTOMORPHISMS");
    const bool calculate_automorphism = autom && atoi(autom) != 0;
    rc = executor.initialization(&g, comp_thds, iterations, isPruned, useSPMM, vtuneStart, calculate_automorphism, 0,
                                 precision, seed);
    if (rc) {
        fprintf(stderr, "error: %s\n", executor.error().c_str());
        return 1;
    }
    const char* autotune = getenv("PG_AUTOTUNE");
    if (autotune && atoi(autotune) != 0 && pg_set_option(executor.context(), "autotune", 1) != PG_OK) {
        fprintf(stderr, "error: %s\n", pg_last_error(executor.context()));
        return 1;
    }
    const char* verbose = getenv("PG_VERBOSE");
    const char* estimate = getenv("PG_ESTIMATE");
    const bool is_estimate = estimate && atoi(estimate) != 0;
    executor.set_verbose(is_estimate || (verbose && atoi(verbose) != 0));
    double res = executor.compute(tmpl, is_estimate);
    if (res < 0) {
        fprintf(stderr, "error: %s\n", executor.error().c_str());
        return 1;
    }
    return 0;
}
