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
// Usage: pg_ref <graph.el> <template.fascia> <colors.u8|-> <ncolorings> <threads> [seed] [isPruned] [useSPMM]
//   colors.u8: ncolorings*n bytes, coloring-major; "-" draws colors from the
//   library's color stream (seed, coloring, vertex) documented in pgbsc.h.
//   isPruned / useSPMM (default 1 1) are CountMat::initialization's mode
//   flags: 1 0 runs countNonBottomePruned (CountMat.cpp:325-554), 0 x the
//   unpruned countNonBottomeOriginal (CountMat.cpp:851-977).
// Output lines: "PGREF n <n> nnz <nnz> k <k>" then per coloring
//   "PGREF coloring <i> raw <%.17g> seconds <t>".
//
//        pg_ref compute <graph.el> <template.fascia> <ncolorings> <threads> <seed> <isPruned> <useSPMM> <autom>
//   runs the whole CountMat::compute(tmpl, false) (CountMat.cpp:79-266), as
//   sc-main.cpp:1793-1795 does, with the color stream: each thread's n-th
//   srand() call (one per colorInit) selects coloring n-1.  The reference
//   prints its own "Final raw count is", "Prob is" and "Final count is"
//   lines; the driver adds "PGREF final <%.17g>" with compute()'s return.
//
//        pg_ref serialize <graph.el> <out.data>
//   writes the graph through the reference's CSCGraph::serialize
//   (CSCGraph.hpp:426-434), the graph.data of sc-main's write_binary.
//
// Every <graph> argument may also be a binary graph.data (a path ending in
// ".data"), read with the reference's CSCGraph::deserialize.
//
//        pg_ref autom <template.fascia>
//   prints "PGREF autom <n>", CountMat::automorphismNum() (CountMat.cpp:
//   1498-1572) of the template.
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
#include <fcntl.h>
#include <unistd.h>

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
static bool g_auto = false;                 // compute mode: coloring = this thread's srand() calls - 1
static thread_local int64_t t_calls = 0;
static thread_local int64_t t_coloring = 0;

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
    t_coloring = g_auto ? t_calls++ : g_coloring;
}

// In hash mode the color is the library's documented color stream
// (pgbsc.h, pg_colors_fill): rand() % k then returns it unchanged.
extern "C" int rand(void) noexcept
{
    int64_t v = t_pos++;
    if (g_hash) {
        uint64_t h = splitmix64(g_seed ^ splitmix64((uint64_t)t_coloring * 0x9E3779B97F4A7C15ull + (uint64_t)v));
        return (int)(((h >> 32) * (uint64_t)g_k) >> 32);
    }
    return g_colors[v];
}

// A text edge list (EdgeList -> createFromEdgeListFile, sc-main.cpp:1692,
// 1762) or, for a path ending in ".data", the binary CSC through the
// reference's own CSCGraph::deserialize (sc-main.cpp:1659-1675).
static CSCGraph<int32_t, float>* load_graph(const std::string& path)
{
    CSCGraph<int32_t, float>* g = new CSCGraph<int32_t, float>();
    if (path.size() > 5 && path.compare(path.size() - 5, 5, ".data") == 0) {
        int fd = open(path.c_str(), O_RDONLY);
        if (fd < 0) { perror(path.c_str()); exit(2); }
        g->deserialize(fd);
        close(fd);
        return g;
    }
    EdgeList elist(path);
    g->createFromEdgeListFile(elist.getNumVertices(), elist.getNumEdges(), elist.getSrcList(), elist.getDstList(), false);
    return g;
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
    CSCGraph<int32_t, float>* g = load_graph(graph_name);
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
    CSCGraph<int32_t, float>* g = load_graph(graph_name);
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
    CSCGraph<int32_t, float>* g = load_graph(graph_name);
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

// CountMat::compute with the color stream (see the usage above)
static int compute_mode(int argc, char** argv)
{
    if (argc != 10) {
        fprintf(stderr, "usage: %s compute graph.el template.fascia ncolorings threads seed isPruned useSPMM autom\n",
                argv[0]);
        return 2;
    }
    const int ncol = atoi(argv[4]), thds = atoi(argv[5]);
    g_seed = strtoull(argv[6], nullptr, 10);
    const int isPruned = atoi(argv[7]), useSPMM = atoi(argv[8]), autom = atoi(argv[9]);
    CSCGraph<int32_t, float>* g = load_graph(argv[2]);
    g->splitCSC(4 * thds);
    g_n = g->getNumVertices();
    Graph tmpl;
    tmpl.read_enlist(argv[3]);
    g_k = (uint64_t)tmpl.get_vert_num();
    g_hash = true;
    g_auto = true;
    CountMat ex;
    ex.initialization(nullptr, g, thds, ncol, isPruned, useSPMM, -1, autom != 0);
    const double fin = ex.compute(tmpl, false);
    printf("PGREF final %.17g\n", fin);
    fflush(stdout);
    delete g;
    return 0;
}

static int autom_mode(const std::string& tmpl_name)
{
    Graph tmpl;
    tmpl.read_enlist(tmpl_name);
    // never destroyed: without initialization() the destructor would free
    // buffers it never allocated
    CountMat* ex = new CountMat();
    ex->_templates = &tmpl;
    printf("PGREF autom %d\n", ex->automorphismNum());
    fflush(stdout);
    return 0;
}

// the reference's own CSCGraph::serialize of a text edge list (CSCGraph.hpp:426-434)
static int serialize_mode(const std::string& graph_name, const std::string& out_name)
{
    CSCGraph<int32_t, float>* g = load_graph(graph_name);
    ofstream out(out_name.c_str(), ios::binary);
    g->serialize(out);
    out.close();
    printf("PGREF serialized n %d nnz %lld\n", g->getNumVertices(), (long long)g->getNNZ());
    delete g;
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "serialize")
        return serialize_mode(argv[2], argv[3]);
    if (argc >= 2 && std::string(argv[1]) == "compute")
        return compute_mode(argc, argv);
    if (argc == 3 && std::string(argv[1]) == "autom")
        return autom_mode(argv[2]);
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
    const int isPruned = argc > 7 ? atoi(argv[7]) : 1;
    const int useSPMM = argc > 8 ? atoi(argv[8]) : 1;

    CSCGraph<int32_t, float>* g = load_graph(graph_name);
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
    ex.initialization(nullptr, g, thds, ncol, isPruned, useSPMM, -1, false);
    ex.compute(tmpl, true);
    // CountMat.cpp:174-197: one buffer per color (useSPMM == 0) or one block
    ex._bufVecLeaf = (float**)malloc(k * sizeof(float*));
    if (useSPMM == 0) {
        for (int i = 0; i < k; ++i) ex._bufVecLeaf[i] = (float*)aligned_alloc(64, (size_t)g_n * sizeof(float));
    } else {
        ex._bufVecLeaf[0] = (float*)aligned_alloc(64, (size_t)g_n * k * sizeof(float));
        for (int i = 1; i < k; ++i) ex._bufVecLeaf[i] = ex._bufVecLeaf[0] + (size_t)i * g_n;
    }

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
