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
    // CountMat.cpp:21-77.  thd_num sizes the host-side OpenMP work only;
    // isPruned / useSPMM select reference modes that all map onto the same
    // kernels (identical results).  Returns 0 or a PG_ERR_* code.
    int initialization(const CSCGraph* graphCSC, int thd_num, int itr_num, int isPruned, int useSPMM,
                       int vtuneStart = -1, bool calculate_automorphisms = false, int device = 0,
                       int precision = PG_FP32, uint64_t seed = 0);
    // CountMat.cpp:79-266: decomposes the template, counts itr_num colorings
    // and returns floor(mean_raw / (k!/k^k * |Aut|) + 0.5); < 0 on error.
    double compute(const Graph& templates, bool isEstimate = false);
    // CountMat.cpp:1498-1572: number of automorphisms of the template.
    int automorphismNum() const;

    const std::vector<double>& raw_counts() const { return raw_; }
    double seconds_per_coloring() const { return sec_per_col_; }
    const std::string& error() const { return err_; }
    pg_ctx* context() { return ctx_; }

private:
    pg_ctx* ctx_ = nullptr;
    const Graph* tmpl_ = nullptr;
    int itr_ = 1, thd_ = 1, precision_ = PG_FP32;
    bool autom_ = false;
    uint64_t seed_ = 0;
    std::vector<double> raw_;
    double sec_per_col_ = 0;
    std::string err_;
};

}  // namespace pgbsc
