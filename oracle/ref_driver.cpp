// Test-only driver for the compiled reference (dsc-nec/subgraph, -DNEC build).
//
// This file is part of the parity oracle, not of the product.  It links the
// reference's own objects (CountMat.cpp, DataTableColMajor.cpp, IndexSys.cpp,
// DivideTemplates.cpp, Graph.cpp, EdgeList.cpp, CSRGraph.cpp, Helper.cpp),
// compiled from their sources by oracle/Makefile, and drives them exactly as
// sc-main.cpp:1692-1795 does (text edge list -> CSCGraph -> splitCSC(4*thds)
// -> Graph::read_enlist -> CountMat::initialization/compute), with two
// differences that only make runs reproducible and per-coloring visible:
//
//  * colors come from a file (or a seeded hash) instead of srand(time(0)):
//    this file defines srand()/rand(), which the reference's colorInit
//    (CountMat.cpp:979-991) resolves to at link time.  srand() records the
//    calling OpenMP thread's static-schedule chunk so rand() hands each
//    vertex its own color whatever the thread count;
//  * CountMat::compute(tmpl, isEstimate=true) performs the setup
//    (CountMat.cpp:79-125) and returns before counting; the driver then
//    allocates _bufVecLeaf as CountMat.cpp:174-197 does and calls the private
//    colorCounting() (CountMat.cpp:268-322) once per coloring, printing its
//    raw double with full precision and its wall time (the scope of the
//    reference's "Time for count per iter", CountMat.cpp:200-213).
//
// Usage: pg_ref <graph.el> <template.fascia> <colors.u8|-> <ncolorings> <threads> [seed]
//   colors.u8: ncolorings*n bytes, coloring-major; "-" draws colors from the
//   library's color stream (seed, coloring, vertex) documented in pgbsc.h.
// Output lines: "PGREF n <n> nnz <nnz> k <k>" then per coloring
//   "PGREF coloring <i> raw <%.17g> seconds <t>".
//
//        pg_ref dump <graph.el> <template.fascia>
//   prints the reference's CSC ("PGREF colptr ...", "PGREF rowidx ...") and
//   its subtemplate chain and index tables ("PGREF sub s size main aux",
//   "PGREF c2c s ...", "PGREF smain s ...", "PGREF saux s ...") with leaves'
//   child links as -1.
//
//        pg_ref estimate <graph.el> <template.fascia>
//   runs the reference's cost models on the compute(t, true) setup:
//   estimatePeakMemUsage (its own "Peak memory usage estimated" line),
//   estimate{Flops,MemComm}{PGBSC,Fascia,PrunedFascia}, then
//   "PGREF est <name> <%.17g>" per figure; degreeDistribution writes
//   degDistri.data in the working directory.
//
//        pg_ref estimate_template <graph.el> <template.fascia>
//   runs estimateTemplate (CountMat.cpp:1448-1496) on the same setup.  That
//   function has no return statement; built at -O3 it runs on past its two
//   lines ("Workload N: ...", "Memload N: ...") and does not return, so the
//   caller reads those two lines and then ends the process.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <iostream>
#include <fstream>
#include <algorithm>
#include <omp.h>
#include <sys/time.h>

#define private public
#include "CountMat.hpp"
#include "EdgeList.hpp"
#undef private

static const uint8_t* g_colors = nullptr;   // current coloring, n bytes
static int64_t g_n = 0;
static uint64_t g_seed = 0;
static int64_t g_coloring = 0;
static bool g_hash = false;
static uint64_t g_k = 1;
static thread_local int64_t t_pos = 0;

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// libgomp's schedule(static) without a chunk: thread i gets q+1 items from
// (q+1)*i when i < n%T, else q items from q*i + n%T.
extern "C" void srand(unsigned) noexcept
{
    int64_t T = omp_get_num_threads();
    int64_t i = omp_get_thread_num();
    int64_t q = g_n / T, r = g_n % T;
    t_pos = (i < r) ? (q + 1) * i : q * i + r;
}

// In hash mode the color is the library's documented color stream
// (pgbsc.h, pg_colors_fill): rand() % k then returns it unchanged.
extern "C" int rand(void) noexcept
{
    int64_t v = t_pos++;
    if (g_hash) {
        uint64_t h = splitmix64(g_seed ^ splitmix64((uint64_t)g_coloring * 0x9E3779B97F4A7C15ull + (uint64_t)v));
        return (int)(((h >> 32) * (uint64_t)g_k) >> 32);
    }
    return g_colors[v];
}

static double now_s()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static void print_ints(const char* tag, int s, const int* a, long long len)
{
    printf("PGREF %s %d", tag, s);
    for (long long i = 0; i < len; ++i) printf(" %d", a[i]);
    printf("\n");
}

static int dump(const std::string& graph_name, const std::string& tmpl_name)
{
    EdgeList elist(graph_name);
    CSCGraph<int32_t, float>* g = new CSCGraph<int32_t, float>();
    g->createFromEdgeListFile(elist.getNumVertices(), elist.getNumEdges(), elist.getSrcList(), elist.getDstList(), false);
    int n = g->getNumVertices();
    print_ints("colptr", n, g->getIndexCol(), (long long)n + 1);
    print_ints("rowidx", n, g->getIndexRow(), (long long)g->getNNZ());
    g->splitCSC(4);
    Graph tmpl;
    tmpl.read_enlist(tmpl_name);
    int k = tmpl.get_vert_num();
    CountMat ex;
    ex.initialization(nullptr, g, 1, 1, 1, 1, -1, false);
    ex.compute(tmpl, true);
    int S = ex._total_sub_num;
    for (int s = 0; s < S; ++s) {
        int sz = ex._subtmp_array[s].get_vert_num();
        int mi = ex.div_tp.get_main_node_idx(s), ai = ex.div_tp.get_aux_node_idx(s);
        printf("PGREF sub %d %d %d %d\n", s, sz, mi == DUMMY_VAL ? -1 : mi, ai == DUMMY_VAL ? -1 : ai);
        int nc = ex.indexer.comb_calc(k, sz);
        print_ints("c2c", s, ex.indexer.getCombToCountTable()[s], nc);
        if (sz > 1) {
            int ns = ex.indexer.comb_calc(sz, ex.div_tp.get_main_node_vert_num(s));
            std::vector<int> sm, sa;
            for (int i = 0; i < nc; ++i)
                for (int j = 0; j < ns; ++j) {
                    sm.push_back(ex.indexer.getSplitToCountTable()[0][s][i][j]);
                    sa.push_back(ex.indexer.getSplitToCountTable()[1][s][i][j]);
                }
            print_ints("smain", s, sm.data(), (long long)sm.size());
            print_ints("saux", s, sa.data(), (long long)sa.size());
        }
    }
    fflush(stdout);
    delete g;
    return 0;
}

static int estimate(const std::string& graph_name, const std::string& tmpl_name)
{
    EdgeList elist(graph_name);
    CSCGraph<int32_t, float>* g = new CSCGraph<int32_t, float>();
    g->createFromEdgeListFile(elist.getNumVertices(), elist.getNumEdges(), elist.getSrcList(), elist.getDstList(), false);
    g->splitCSC(4);
    Graph tmpl;
    tmpl.read_enlist(tmpl_name);
    CountMat ex;
    ex.initialization(nullptr, g, 1, 1, 1, 1, -1, false);
    ex.compute(tmpl, true);
    printf("PGREF n %d nnz %lld\n", g->getNumVertices(), (long long)g->getNNZ());
    fflush(stdout);
    ex.estimatePeakMemUsage();
    const double pgbsc_flop = ex.estimateFlopsPGBSC();
    const double pgbsc_bytes = ex.estimateMemCommPGBSC();
    const double fascia_flop = ex.estimateFlopsFascia();
    const double fascia_bytes = ex.estimateMemCommFascia();
    const double pruned_flop = ex.estimateFlopsPrunedFascia();
    const double pruned_bytes = ex.estimateMemCommPrunedFascia();
    printf("PGREF est pgbsc_spmv_gflop %.17g\n", ex._spmvFlops);
    printf("PGREF est pgbsc_fma_gflop %.17g\n", ex._fmaFlops);
    printf("PGREF est pgbsc_spmv_gbytes %.17g\n", ex._spmvMemBytes);
    printf("PGREF est pgbsc_fma_gbytes %.17g\n", ex._fmaMemBytes);
    printf("PGREF est pgbsc_gflop %.17g\n", pgbsc_flop);
    printf("PGREF est pgbsc_gbytes %.17g\n", pgbsc_bytes);
    printf("PGREF est fascia_gflop %.17g\n", fascia_flop);
    printf("PGREF est fascia_gbytes %.17g\n", fascia_bytes);
    printf("PGREF est pruned_fascia_gflop %.17g\n", pruned_flop);
    printf("PGREF est pruned_fascia_gbytes %.17g\n", pruned_bytes);
    fflush(stdout);
    ex.degreeDistribution();
    fflush(stdout);
    return 0;
}

static int estimate_template(const std::string& graph_name, const std::string& tmpl_name)
{
    EdgeList elist(graph_name);
    CSCGraph<int32_t, float>* g = new CSCGraph<int32_t, float>();
    g->createFromEdgeListFile(elist.getNumVertices(), elist.getNumEdges(), elist.getSrcList(), elist.getDstList(), false);
    g->splitCSC(4);
    Graph tmpl;
    tmpl.read_enlist(tmpl_name);
    CountMat ex;
    ex.initialization(nullptr, g, 1, 1, 1, 1, -1, false);
    ex.compute(tmpl, true);
    fflush(stdout);
    ex.estimateTemplate();
    fflush(stdout);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "dump")
        return dump(argv[2], argv[3]);
    if (argc == 4 && std::string(argv[1]) == "estimate")
        return estimate(argv[2], argv[3]);
    if (argc == 4 && std::string(argv[1]) == "estimate_template")
        return estimate_template(argv[2], argv[3]);
    if (argc < 6) {
        fprintf(stderr, "usage: %s graph.el template.fascia colors.u8|- ncolorings threads [seed]\n", argv[0]);
        return 2;
    }
    std::string graph_name = argv[1], tmpl_name = argv[2], colors_name = argv[3];
    int ncol = atoi(argv[4]);
    int thds = atoi(argv[5]);
    g_seed = argc > 6 ? strtoull(argv[6], nullptr, 10) : 1;

    EdgeList elist(graph_name);
    CSCGraph<int32_t, float>* g = new CSCGraph<int32_t, float>();
    g->createFromEdgeListFile(elist.getNumVertices(), elist.getNumEdges(), elist.getSrcList(), elist.getDstList(), false);
    g->splitCSC(4 * thds);
    g_n = g->getNumVertices();

    Graph tmpl;
    tmpl.read_enlist(tmpl_name);
    int k = tmpl.get_vert_num();
    g_k = (uint64_t)k;

    std::vector<uint8_t> colors;
    if (colors_name == "-") {
        g_hash = true;
    } else {
        std::ifstream f(colors_name, std::ios::binary);
        colors.resize((size_t)ncol * g_n);
        f.read((char*)colors.data(), colors.size());
        if (!f) { fprintf(stderr, "short colors file\n"); return 2; }
    }

    printf("PGREF n %lld nnz %lld k %d\n", (long long)g_n, (long long)g->getNNZ(), k);
    fflush(stdout);

    CountMat ex;
    ex.initialization(nullptr, g, thds, ncol, 1, 1, -1, false);
    ex.compute(tmpl, true);
    // CountMat.cpp:174-197 (the useSPMM == 1 branch)
    ex._bufVecLeaf = (float**)malloc(k * sizeof(float*));
    ex._bufVecLeaf[0] = (float*)aligned_alloc(64, (size_t)g_n * k * sizeof(float));
    for (int i = 1; i < k; ++i) ex._bufVecLeaf[i] = ex._bufVecLeaf[0] + (size_t)i * g_n;

    for (int i = 0; i < ncol; ++i) {
        g_coloring = i;
        if (!g_hash) g_colors = colors.data() + (size_t)i * g_n;
        double t0 = now_s();
        double raw = ex.colorCounting();
        double t1 = now_s();
        printf("PGREF coloring %d raw %.17g seconds %.6f\n", i, raw, t1 - t0);
        fflush(stdout);
    }
    delete g;
    return 0;
}
