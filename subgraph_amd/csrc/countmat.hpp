// CountMat: the DP driver of the reference (CountMat.hpp:19-191), rebuilt on
// the libpgbsc C ABI.  Same entry points and printed results; the
// per-subtemplate body runs as HIP kernels on one MI355X.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pg_host.hpp"
#include "pgbsc.h"

namespace pgbsc {

class CountMat {
public:
    ~CountMat();
    // CountMat.cpp:21-77.  thd_num sizes the host-side OpenMP work only.
    // isPruned = 0 selects the reference's unpruned mode with its numerics
    // (PG_FLAG_UNPRUNED, FP32 only); useSPMM = 0 forms the same sums in the
    // same order as the SpMM path and runs its kernels.  Returns 0 or a
    // PG_ERR_* code.
    int initialization(const CSCGraph* graphCSC, int thd_num, int itr_num, int isPruned, int useSPMM,
                       int vtuneStart = -1, bool calculate_automorphisms = false, int device = 0,
                       int precision = PG_FP32, uint64_t seed = 0);
    // CountMat.cpp:79-266: decomposes the template, counts itr_num colorings
    // and returns floor(mean_raw / (k!/k^k * |Aut|) + 0.5); < 0 on error.
    double compute(const Graph& templates, bool isEstimate = false);
    // CountMat.cpp:1498-1572: number of automorphisms of the template.
    int automorphismNum() const;

    // The reference's analytic estimators (CountMat.cpp:1067-1496), printing
    // its lines.  They need the subtemplate chain, so they work after
    // compute() (compute(t, true) sets it up without counting), like the
    // reference's, which read the tables compute() builds.  GB / GBytes /
    // Gflop are 2^30.  Figures come from pg_estimate_model.
    void estimatePeakMemUsage();
    double estimateFlopsPGBSC();
    double estimateMemCommPGBSC();
    double estimateFlopsFascia();
    double estimateMemCommFascia();
    double estimateFlopsPrunedFascia();
    double estimateMemCommPrunedFascia();
    // Writes "degree count" lines to degDistri.data in the working directory
    // and prints the largest degree (CountMat.cpp:1374-1400).
    void degreeDistribution();
    // Prints the raw workload / memload sums (CountMat.cpp:1448-1496; the
    // reference's version has no return statement, this one returns 0).
    double estimateTemplate();

    // VERBOSE-style report (the reference's compile-time VERBOSE blocks,
    // CountMat.cpp:132-162, 216-242): with it on, compute() prints every
    // estimate before counting, then the per-kernel breakdown of the device
    // time and the model's bandwidth and throughput over the measured time.
    void set_verbose(bool on) { verbose_ = on; }

    const std::vector<double>& raw_counts() const { return raw_; }
    double seconds_per_coloring() const { return sec_per_col_; }
    const std::string& error() const { return err_; }
    pg_ctx* context() { return ctx_; }

private:
    bool model(pg_estimate* e);

    pg_ctx* ctx_ = nullptr;
    const CSCGraph* graph_ = nullptr;
    const Graph* tmpl_ = nullptr;
    FlatPlan fp_;
    bool verbose_ = false;
    int itr_ = 1, thd_ = 1, precision_ = PG_FP32;
    bool autom_ = false;
    bool unpruned_ = false;
    uint64_t seed_ = 0;
    std::vector<double> raw_;
    double sec_per_col_ = 0;
    std::string err_;
};

}  // namespace pgbsc
