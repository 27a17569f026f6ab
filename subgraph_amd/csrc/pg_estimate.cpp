// The reference's analytic cost models of the pruned color-coding DP and of
// the two Fascia variants (CountMat.cpp:1067-1446), restated on the flattened
// subtemplate chain.  Host only.
//
// Every model walks the non-leaf subtemplates s (|s| > 1) from the last index
// down to 0, with a = the aux child, len(x) = C(k, |x|) table columns and
// P(s) = C(|s|, |a|) splits per column:
//   estimateFlopsPGBSC       spmv += nnz*len(a) [|a|>1];  fma += len(s)*P(s)*2n
//   estimateMemCommPGBSC     spmv += (4*2nnz + 4*2nnz/16)*len(a) [|a|>1];  fma += len(s)*P(s)*12n
//   estimateFlopsFascia      += len(s)*P(s)*(nnz + 2n)
//   estimateMemCommFascia    += len(s)*8n + len(s)*P(s)*(4(n+nnz) + 8n)
//   estimateFlopsPrunedFascia  += nnz*len(a) [|a|>1] + len(s)*P(s)*2n
//   estimateMemCommPrunedFascia += 4nnz*len(a) [|a|>1] + len(s)*8n + len(s)*P(s)*16n
// and all of these are divided by 2^30.  estimateTemplate (CountMat.cpp:
// 1448-1496) keeps raw sums: workload_nnz += 2 len(a), memload_nnz += 12 len(a),
// memload_n += 8 len(a) [|a|>1]; workload_n += 2 len(s)P(s), memload_n +=
// 16 len(s)P(s).  estimatePeakMemUsage (CountMat.cpp:1067-1115) tracks live
// table columns over s = S-1..1 (the root excluded): start from 2 + 16 + k
// buffer columns, add len(s), take the maximum, then drop the non-leaf
// children; the peak is that maximum times 4n bytes plus the graph's
// 4(n + 2nnz) bytes, in 2^30-byte units.
#include "pg_host.hpp"
#include "pgbsc.h"

namespace {

constexpr double kGiga = 1024.0 * 1024.0 * 1024.0;
constexpr int kBufMatCols = 16;  // CountMat.cpp:62, the SpMM's 16-column batches

}  // namespace

extern "C" int pg_estimate_model(int64_t n, int64_t nnz, int k, int nsub, const int32_t* sub_size,
                                 const int32_t* main_idx, const int32_t* aux_idx, pg_estimate* out)
{
    if (!out || !sub_size || !main_idx || !aux_idx || n < 0 || nnz < 0 || k < 1 || k > 30 || nsub < 1)
        return PG_ERR_INVALID;
    for (int s = 0; s < nsub; ++s) {
        if (sub_size[s] < 1 || sub_size[s] > k) return PG_ERR_INVALID;
        if (sub_size[s] > 1) {
            const int m = main_idx[s], a = aux_idx[s];
            if (m <= s || a <= s || m >= nsub || a >= nsub || sub_size[m] + sub_size[a] != sub_size[s])
                return PG_ERR_INVALID;
        }
    }
    using pgbsc::IndexSys;
    auto len = [&](int s) { return IndexSys::comb_calc(k, sub_size[s]); };
    const double dn = (double)n, dz = (double)nnz;
    pg_estimate e{};

    // estimatePeakMemUsage
    const double per_col = dn * 4.0 / kGiga;
    double live = (2 + kBufMatCols + k) * per_col, peak = 0.0;
    for (int s = nsub - 1; s > 0; --s) {
        if (sub_size[s] <= 1) continue;
        live += len(s) * per_col;
        peak = live > peak ? live : peak;
        if (sub_size[main_idx[s]] > 1) live -= len(main_idx[s]) * per_col;
        if (sub_size[aux_idx[s]] > 1) live -= len(aux_idx[s]) * per_col;
    }
    e.peak_mem_gb = peak + (dn + 2.0 * dz) * 4.0 / kGiga;

    const double spmv_bytes = 4.0 * (2.0 * dz) + 4.0 * (2.0 * dz / 16.0);
    const double fma_bytes = 4.0 * (3.0 * dn);
    const double comb_bytes = 4.0 * dn + 4.0 * dn;
    const double split_bytes_fascia = 4.0 * (dn + dz) + 4.0 * 2.0 * dn;
    const double split_bytes_pruned = 4.0 * (2.0 * dn) + 4.0 * 2.0 * dn;
    double fascia_bytes = 0, pruned_bytes = 0, fascia_flop = 0, pruned_flop = 0;
    for (int s = nsub - 1; s >= 0; --s) {
        if (sub_size[s] <= 1) continue;
        const int a = aux_idx[s];
        const bool aux_table = sub_size[a] > 1;
        const double ls = len(s);
        const double lsP = (double)(len(s) * IndexSys::comb_calc(sub_size[s], sub_size[a]));
        if (aux_table) {
            const double la = len(a);
            e.pgbsc_spmv_gflop += dz * la;
            e.pgbsc_spmv_gbytes += spmv_bytes * la;
            pruned_flop += dz * la;
            pruned_bytes += 4.0 * dz * la;
            e.workload_nnz += 2.0 * la;
            e.memload_nnz += 12.0 * la;
            e.memload_n += 8.0 * la;
        }
        e.pgbsc_fma_gflop += lsP * 2.0 * dn;
        e.pgbsc_fma_gbytes += lsP * fma_bytes;
        fascia_flop += lsP * (dz + 2.0 * dn);
        fascia_bytes += ls * comb_bytes + lsP * split_bytes_fascia;
        pruned_flop += lsP * 2.0 * dn;
        pruned_bytes += ls * comb_bytes + lsP * split_bytes_pruned;
        e.workload_n += lsP * 2.0;
        e.memload_n += 16.0 * lsP;
    }
    e.pgbsc_spmv_gflop /= kGiga;
    e.pgbsc_fma_gflop /= kGiga;
    e.pgbsc_spmv_gbytes /= kGiga;
    e.pgbsc_fma_gbytes /= kGiga;
    e.fascia_gflop = fascia_flop / kGiga;
    e.fascia_gbytes = fascia_bytes / kGiga;
    e.pruned_fascia_gflop = pruned_flop / kGiga;
    e.pruned_fascia_gbytes = pruned_bytes / kGiga;
    *out = e;
    return PG_OK;
}
