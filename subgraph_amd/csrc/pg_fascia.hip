// The reference's unpruned color-coding mode ("Fascia", isPruned = 0):
// CountMat::countNonBottomeOriginal, CountMat.cpp:851-977, with its exact
// FP32 numerics.  It is a compatibility mode for the drop-in CLI, not a fast
// path: the pruned DP (pg_kernels.hip) computes the same counts with far
// less work.
//
// What the reference does per non-leaf subtemplate s, for every output set i
// (countCombNum, in IndexSys order) and every split j of it:
//   _bufVec += A * aux[split_aux[i][j]]        (spmvNaiveSplit, y[row] += 1.0f * x[col],
//                                               CSCGraph.hpp:220-239; _bufVec is never
//                                               cleared: CountMat.cpp:52-59 zero it once,
//                                               at initialization)
//   obj[i] = obj[i] + _bufVec * main[split_main[i][j]]   (arrayWiseFMANaive, FP32,
//                                               DataTableColMajor.cpp:502-510)
// with unscaled FP32 tables (no float(1e-12), no 1e12), and at the root
// obj = bufLastSub, a float vector summed in FP64 in vertex order
// (CountMat.cpp:884-955).  Every operation on _bufVec[v] and obj[v] involves
// only vertex v and its neighbours' child-table values, so one thread per
// vertex replays the reference's sequence of FP32 operations exactly: the
// neighbours of v in ascending caller id (the order of spmvNaiveSplit's row
// owner lists), pair after pair.
//
// Tables are the reference's own layout here: column-major, C(k, |s|)
// columns of n_act floats (internal vertex order), absolute color sets at
// their colex rank.  Leaf tables are one-hot and read from the colors.
#include "pg_kernels.hpp"

#include <algorithm>

namespace pgbsc {

namespace {

__global__ __launch_bounds__(256) void k_fascia_step(const FasciaArgs a)
{
    const int64_t n = a.n_act;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int e0 = a.rowptr[v], e1 = a.rowptr[v + 1];
        const uint32_t cv = a.colors[v];
        float bv = a.acc[v];
        for (int i = 0; i < a.Cs; ++i) {
            float obj = 0.0f;  // the table (and bufLastSub) start zeroed
            for (int j = 0; j < a.P; ++j) {
                const uint32_t sp = a.split[(int64_t)i * a.P + j];
                const uint32_t mi = sp >> 16, ai = sp & 0xffffu;
                for (int e = e0; e < e1; ++e) {
                    const int u = a.col[e];
                    const float x = a.aux_leaf ? (a.colors[u] == ai ? 1.0f : 0.0f) : a.ta[(int64_t)ai * n + u];
                    bv = bv + x;
                }
                const float m = a.main_leaf ? (cv == mi ? 1.0f : 0.0f) : a.tm[(int64_t)mi * n + v];
                obj = obj + bv * m;
            }
            a.out[(int64_t)a.c2c[i] * n + v] = obj;
        }
        a.acc[v] = bv;
    }
}

// vals (internal order) into dst (caller order): dst[perm[v]] = vals[v]
__global__ void k_fascia_scatter(float* dst, const float* vals, const int32_t* perm, int32_t n_act)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n_act; v += (int64_t)gridDim.x * blockDim.x)
        dst[perm[v]] = vals[v];
}

// out = sum over v = 0..n-1 of (double) vals[v], strictly in order
// (CountMat.cpp:952-955).  One workgroup: 256 values at a time are loaded
// into LDS (the next chunk while the current one is added) and thread 0 adds
// them one after another.
__global__ __launch_bounds__(256) void k_fascia_sum(const float* vals, int64_t n, double* out)
{
    __shared__ float buf[2][256];
    double sum = 0.0;
    int cur = 0;
    if (threadIdx.x < n) buf[0][threadIdx.x] = vals[threadIdx.x];
    __syncthreads();
    for (int64_t base = 0; base < n; base += 256) {
        const int64_t nx = base + 256 + threadIdx.x;
        if (nx < n) buf[cur ^ 1][threadIdx.x] = vals[nx];
        if (threadIdx.x == 0) {
            const int64_t m = n - base < 256 ? n - base : 256;
            for (int64_t t = 0; t < m; ++t) sum = sum + (double)buf[cur][t];
        }
        __syncthreads();
        cur ^= 1;
    }
    if (threadIdx.x == 0) *out = sum;
}

}  // namespace

hipError_t launch_fascia_step(const FasciaArgs& a, int num_cus, hipStream_t st)
{
    if (a.n_act == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(((int64_t)a.n_act + 255) / 256, (int64_t)num_cus * 16);
    hipLaunchKernelGGL(k_fascia_step, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_fascia_root_sum(const float* vals, const int32_t* perm, int32_t n_act, float* scratch, int64_t n,
                                  double* out, hipStream_t st)
{
    if (n > 0) {
        hipError_t e = hipMemsetAsync(scratch, 0, (size_t)n * sizeof(float), st);
        if (e != hipSuccess) return e;
    }
    if (n_act > 0) {
        const int64_t blocks = std::min<int64_t>(((int64_t)n_act + 255) / 256, 1 << 16);
        hipLaunchKernelGGL(k_fascia_scatter, dim3((unsigned)blocks), dim3(256), 0, st, scratch, vals, perm, n_act);
    }
    hipLaunchKernelGGL(k_fascia_sum, dim3(1), dim3(256), 0, st, scratch, n, out);
    return hipGetLastError();
}

}  // namespace pgbsc
