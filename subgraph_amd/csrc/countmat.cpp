// CountMat driver over the libpgbsc C ABI (reference: CountMat.cpp:21-266,
// 979-999, 1067-1572).
#include "countmat.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>

namespace pgbsc {

CountMat::~CountMat()
{
    if (ctx_) pg_destroy(ctx_);
}

int CountMat::initialization(const CSCGraph* g, int thd_num, int itr_num, int isPruned, int useSPMM, int,
                             bool calculate_automorphisms, int device, int precision, uint64_t seed)
{
    if (!g || itr_num <= 0) {
        err_ = "initialization needs a graph and itr_num > 0";
        return PG_ERR_INVALID;
    }
    // isPruned = 0 is the reference's unpruned mode, with its own numerics
    // (PG_FLAG_UNPRUNED).  useSPMM = 0 (countNonBottomePruned, CountMat.cpp:
    // 325-554) forms the same FP32 sums in the same order as the SpMM path
    // (spmvNaiveSplit adds 1.0f * x, CSCGraph.hpp:220-239), so it runs the
    // same kernels.
    if (isPruned == 0 && precision != PG_FP32) {
        err_ = "isPruned = 0 has the reference's FP32 numerics only";
        return PG_ERR_INVALID;
    }
    unpruned_ = isPruned == 0;
    (void)useSPMM;
    graph_ = g;
    itr_ = itr_num;
    thd_ = thd_num;
    autom_ = calculate_automorphisms;
    precision_ = precision;
    seed_ = seed;
    if (ctx_) pg_destroy(ctx_);
    ctx_ = nullptr;
    int rc = pg_create(&ctx_, device);
    if (rc) {
        err_ = "no usable HIP device";
        return rc;
    }
    rc = pg_graph_upload(ctx_, g->getNumVertices(), g->getNNZ(), g->getIndexCol(), g->getIndexRow(), nullptr);
    if (rc) err_ = pg_last_error(ctx_);
    return rc;
}

// CountMat.cpp:993-999 computes k! in int, which overflows for k >= 13
// (13! > 2^31); the product here is exact in double for every k <= 14.
static double factorial(int n) { return n <= 1 ? 1.0 : n * factorial(n - 1); }

double CountMat::compute(const Graph& t, bool isEstimate)
{
    tmpl_ = &t;
    if (!ctx_) {
        err_ = "compute before initialization";
        return PG_ERR_STATE;
    }
    FlatPlan& fp = fp_;
    int rc = flatten_plan(t, &fp, &err_);
    if (rc) return rc;
    pg_plan* plan = nullptr;
    rc = pg_plan_create(ctx_, fp.k, fp.nsub, fp.sub_size.data(), fp.main_idx.data(), fp.aux_idx.data(),
                        fp.comb_to_count.data(), fp.split_main.data(), fp.split_aux.data(), &plan);
    if (rc) {
        err_ = pg_last_error(ctx_);
        return rc;
    }
    pg_estimate est{};
    if (verbose_) {
        // CountMat.cpp:132-162
        estimatePeakMemUsage();
        int64_t dev_bytes = 0;
        double alg = 0;
        if (pg_plan_footprint(ctx_, plan, 1, precision_, &dev_bytes, &alg) == PG_OK)
            printf("Device memory of a one-coloring pass : %f GB; algorithmic HBM bytes per coloring : %f GB\n",
                   dev_bytes / 1073741824.0, alg / 1073741824.0);
        const double flops = estimateFlopsPGBSC(), bytes = estimateMemCommPGBSC();
        printf("PGBSC Arith Intensity: %f\n", flops / bytes);
        const double ff = estimateFlopsFascia(), fb = estimateMemCommFascia();
        printf("Fascia Arith Intensity: %f\n", ff / fb);
        const double pf = estimateFlopsPrunedFascia(), pb = estimateMemCommPrunedFascia();
        printf("Pruned Fascia Arith Intensity: %f\n", pf / pb);
        degreeDistribution();
        estimateTemplate();
        fflush(stdout);
        model(&est);
    }
    if (isEstimate) {
        pg_plan_destroy(plan);
        return 0.0;
    }
    raw_.assign(itr_, 0.0);
    if (verbose_) pg_set_option(ctx_, "timing", 1);
    auto t0 = std::chrono::steady_clock::now();
    rc = pg_count(ctx_, plan, itr_, nullptr, seed_, precision_, unpruned_ ? PG_FLAG_UNPRUNED : 0u, raw_.data());
    auto t1 = std::chrono::steady_clock::now();
    if (verbose_) pg_set_option(ctx_, "timing", 0);
    if (rc) {
        err_ = pg_last_error(ctx_);
        pg_plan_destroy(plan);
        return rc;
    }
    sec_per_col_ = std::chrono::duration<double>(t1 - t0).count() / itr_;
    // CountMat.cpp:306-311, once per subtemplate of every coloring: the
    // memory in use (here the device bytes of the call's pass plan) and the
    // value colorCounting holds after the step, which is 0 until the root
    // (subtemplate 0) returns the raw count
    {
        int64_t dev = 0, tile = 1;
        pg_get_option(ctx_, "last_tile", &tile);
        pg_plan_footprint(ctx_, plan, (int)std::max<int64_t>(tile, 1), precision_, &dev, nullptr);
        for (int i = 0; i < itr_; ++i)
            for (int s = fp_.nsub - 1; s >= 0; --s)
                printf("Mem utilization compute step sub %d: %9.6lf GB, count val: %e\n", s, dev / 1073741824.0,
                       s == 0 ? raw_[i] : 0.0);
    }
    printf("\nTime for count per iter: %9.6lf seconds\n", sec_per_col_);
    if (verbose_) {
        // CountMat.cpp:216-242.  SpMM and eMA run fused in one kernel, so the
        // model's two parts are reported over the one device time.
        double total_ms = 0;
        pg_last_timing(ctx_, &total_ms, nullptr, nullptr, nullptr);
        const double sec = total_ms * 1e-3;
        printf("Device kernel time per iter : %f seconds\n", sec / itr_);
        char name[128];
        double ms = 0, kb = 0;
        int64_t nl = 0;
        const int nk = pg_last_timing_entry(ctx_, 0, nullptr, 0, nullptr, nullptr, nullptr);
        for (int i = 0; i < nk; ++i)
            if (pg_last_timing_entry(ctx_, i, name, (int)sizeof name, &ms, &nl, &kb) >= 0)
                printf("Kernel %s : %f ms in %lld launches, %f GB/s algorithmic\n", name, ms, (long long)nl,
                       ms > 0 ? kb / (ms * 1e-3) / 1e9 : 0.0);
        // CountMat.cpp:833-840 per subtemplate, averaged over the colorings:
        // its device time, then each kernel of it.  SpMM and eMA run fused,
        // so the split the reference prints (Spmv / Mul) is by launch: the
        // neighbour-summing launches against the rest.
        const int ne = pg_last_timing_sub(ctx_, 0, nullptr, nullptr, 0, nullptr, nullptr, nullptr);
        struct Row { std::string name; double ms, bytes; int64_t launches; };
        std::map<int, std::vector<Row>> by_sub;
        int sub = 0;
        for (int i = 0; i < ne; ++i)
            if (pg_last_timing_sub(ctx_, i, &sub, name, (int)sizeof name, &ms, &nl, &kb) >= 0)
                by_sub[sub].push_back({name, ms, kb, nl});
        for (auto it = by_sub.rbegin(); it != by_sub.rend(); ++it) {
            double tot = 0, spmv = 0;
            for (const Row& r : it->second) {
                tot += r.ms;
                // aux modes of k_sub<T,team,aux,main>: 2 reads the histogram, 3 sums hub partials
                const size_t c1 = r.name.find(',');
                const size_t c2 = c1 == std::string::npos ? c1 : r.name.find(',', c1 + 1);
                const int aux = c2 == std::string::npos ? -1 : atoi(r.name.c_str() + c2 + 1);
                if (r.name.rfind("k_sub<", 0) == 0 && aux != 2 && aux != 3) spmv += r.ms;
            }
            const double sec = tot * 1e-3 / itr_, ssec = spmv * 1e-3 / itr_;
            if (it->first < 0)
                printf("Pass-level launches (colors, neighbour-color histogram, root sums), counting time %f\n", sec);
            else
                printf("Sub %d, counting time %f, Spmv time %f: ratio: %f%%,  Mul time %f: ratio: %f%% \n", it->first,
                       sec, ssec, sec > 0 ? 100 * ssec / sec : 0.0, sec - ssec, sec > 0 ? 100 * (sec - ssec) / sec : 0.0);
            for (const Row& r : it->second)
                printf("    %s : %f ms per iter in %lld launches, %f GB/s algorithmic\n", r.name.c_str(), r.ms / itr_,
                       (long long)(r.launches / itr_), r.ms > 0 ? r.bytes / (r.ms * 1e-3) / 1e9 : 0.0);
        }
        if (sec > 0) {
            printf("Total Memory bandwidth is : %f GBytes per second\n",
                   ((est.pgbsc_spmv_gbytes + est.pgbsc_fma_gbytes) * itr_) / sec);
            printf("Total Throughput is : %f Gflops per second\n",
                   ((est.pgbsc_spmv_gflop + est.pgbsc_fma_gflop) * itr_) / sec);
        }
    }
    double mean = 0;
    for (double r : raw_) mean += r;
    mean /= itr_;
    const int k = t.get_vert_num();
    // CountMat.cpp:248-259: colorful probability k!/k^k, times |Aut| if asked
    const double prob = factorial(k) / (factorial(0) * std::pow((double)k, k));
    printf("Final raw count is %e\n", mean);
    printf("Prob is %f\n", prob);
    const int aut = autom_ ? automorphismNum() : 1;
    const double fin = std::floor(mean / (prob * (double)aut) + 0.5);
    printf("Final count is %e\n", fin);
    fflush(stdout);
    pg_plan_destroy(plan);
    return fin;
}

// CountMat.cpp:1498-1572, the reference's acceptance rule
// (pg_template_automorphisms)
int CountMat::automorphismNum() const
{
    if (!tmpl_) return 0;
    const int64_t n = pg_template_automorphisms(tmpl_->get_vert_num(), (int)tmpl_->src().size(),
                                                tmpl_->src().data(), tmpl_->dst().data());
    return n < 0 ? 0 : (int)n;
}

bool CountMat::model(pg_estimate* e)
{
    if (!graph_ || fp_.nsub == 0) {
        err_ = "the estimators need initialization() and compute() first";
        return false;
    }
    if (pg_estimate_model(graph_->getNumVertices(), graph_->getNNZ(), fp_.k, fp_.nsub, fp_.sub_size.data(),
                          fp_.main_idx.data(), fp_.aux_idx.data(), e) != PG_OK) {
        err_ = "invalid subtemplate chain";
        return false;
    }
    return true;
}

void CountMat::estimatePeakMemUsage()
{
    pg_estimate e;
    if (!model(&e)) return;
    printf("Peak memory usage estimated : %f GB \n", e.peak_mem_gb);
    fflush(stdout);
}

double CountMat::estimateFlopsPGBSC()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("|V| is: %d, |E| nnz is: %lld \n", graph_->getNumVertices(), (long long)graph_->getNNZ());
    printf("Flops PGBSC estimated : SpMV %f Gflop, FMA %f Gflop \n", e.pgbsc_spmv_gflop, e.pgbsc_fma_gflop);
    fflush(stdout);
    return e.pgbsc_spmv_gflop + e.pgbsc_fma_gflop;
}

double CountMat::estimateMemCommPGBSC()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("Comm Bytes estimated PGBSC : SpMV %f GBytes, FMA %f GBytes \n", e.pgbsc_spmv_gbytes, e.pgbsc_fma_gbytes);
    fflush(stdout);
    return e.pgbsc_spmv_gbytes + e.pgbsc_fma_gbytes;
}

double CountMat::estimateFlopsFascia()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("|V| is: %d, |E| nnz is: %lld \n", graph_->getNumVertices(), (long long)graph_->getNNZ());
    printf("Flops estimated fascia: %f Gflop\n", e.fascia_gflop);
    fflush(stdout);
    return e.fascia_gflop;
}

double CountMat::estimateMemCommFascia()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("Comm Bytes Fascia estimated:  %f GBytes\n", e.fascia_gbytes);
    fflush(stdout);
    return e.fascia_gbytes;
}

double CountMat::estimateFlopsPrunedFascia()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("|V| is: %d, |E| nnz is: %lld \n", graph_->getNumVertices(), (long long)graph_->getNNZ());
    printf("Flops estimated Pruned Fascia: %f Gflop\n", e.pruned_fascia_gflop);
    fflush(stdout);
    return e.pruned_fascia_gflop;
}

double CountMat::estimateMemCommPrunedFascia()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("Comm Bytes Pruned Fascia estimated:  %f GBytes\n", e.pruned_fascia_gbytes);
    fflush(stdout);
    return e.pruned_fascia_gbytes;
}

void CountMat::degreeDistribution()
{
    if (!graph_) return;
    const int32_t n = graph_->getNumVertices();
    const int32_t* deg = graph_->getDegList();
    int64_t max_deg = 0;
    for (int32_t i = 0; i < n; ++i) max_deg = std::max<int64_t>(max_deg, deg[i]);
    printf("maxDeg: %lld\n", (long long)max_deg);
    fflush(stdout);
    std::vector<int64_t> hist((size_t)max_deg + 1, 0);
    for (int32_t i = 0; i < n; ++i) hist[deg[i]]++;
    std::ofstream out("degDistri.data");
    for (size_t d = 1; d < hist.size(); ++d)
        if (hist[d] > 0) out << d << " " << hist[d] << "\n";
}

double CountMat::estimateTemplate()
{
    pg_estimate e;
    if (!model(&e)) return 0.0;
    printf("Workload N: %f, NNZ: %f \n", e.workload_n, e.workload_nnz);
    printf("Memload N: %f, NNZ: %f \n", e.memload_n, e.memload_nnz);
    fflush(stdout);
    return 0.0;
}

}  // namespace pgbsc
