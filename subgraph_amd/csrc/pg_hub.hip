// Hub rows of the forward neighbour sum in the reference's order.
//
// The reference adds a vertex's neighbour rows one after another, in column
// order, starting from zero (CSCGraph::spmmSplit, CSCGraph.hpp:254-307).  To
// reproduce its FP32 rounding each Y value must be one running sum over the
// whole neighbour list, so a hub of degree 163k (R-MAT scale 22) is a chain of
// 163k dependent adds per value.  The adds are cheap; what bounds the chain is
// the loads feeding it.  Here a wavefront owns 64 Y values of one (hub,
// coloring) and takes the neighbours in batches of NB: lane j of the batch
// loads entry j's id and colour (one coalesced load each), every lane then
// reads those through v_readlane, issues its NB row loads back to back and
// adds them in order while the next batch's ids load.  NB loads in flight per
// lane instead of the four of the general gathering kernel.
//
// The result is the hub's Y row, stored as its partial row (one per hub); the
// finishing launch of the gathering kernel combines it with the main table.
#include "pg_kernels.hpp"

#include <map>

namespace pgbsc {

namespace {

__device__ __forceinline__ uint32_t hub_expand_bit(uint32_t m, uint32_t c)
{
    return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u));
}

__device__ __forceinline__ uint32_t hub_squeeze_bit(uint32_t m, uint32_t c)
{
    return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c);
}

__host__ __device__ constexpr size_t hub_rank_bytes(int K) { return (((size_t)2 << K) + 15) & ~(size_t)15; }

template <typename T>
__device__ __forceinline__ T lane_value(T x, int j);

template <>
__device__ __forceinline__ float lane_value<float>(float x, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}

template <>
__device__ __forceinline__ double lane_value<double>(double x, int j)
{
    const long long bits = __double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bits, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Work unit = (hub h, coloring b, chunk of 64 Y values), hub-major, so the
// biggest hubs (the lowest internal ids) start first.  Y value i of coloring
// b: set expand(ymask[i], col(v)); neighbour u of colour cu adds its row at
// the rank of (set - cu) relative to cu when cu is in the set (GATHER).
template <typename T, int NB>
__global__ __launch_bounds__(kBlock) void k_hub_seq(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int16_t* s_rank = reinterpret_cast<int16_t*>(smem_raw);
    for (int i = threadIdx.x; i < (1 << a.K); i += kBlock) s_rank[i] = a.rank_tab[i];
    __syncthreads();
    constexpr int WAVES = kBlock / 64;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // per wavefront: for each neighbour colour cu, the row index the lane's
    // Y value reads (-1: none), [cu][lane]
    int16_t* s_tab = reinterpret_cast<int16_t*>(smem_raw + hub_rank_bytes(a.K)) + (size_t)wave * kMaxK * 64;
    const int B = a.B, Wy = a.Wy;
    const int cpb = (Wy + 63) >> 6;
    const int64_t per_hub = (int64_t)B * cpb;
    const int64_t n_work = (int64_t)a.n_items * per_hub;
    for (int64_t w = (int64_t)blockIdx.x * WAVES + wave; w < n_work; w += (int64_t)gridDim.x * WAVES) {
        const int h = (int)(w / per_hub);
        const int rem = (int)(w - (int64_t)h * per_hub);
        const int b = rem / cpb;
        const int i = (rem - b * cpb) * 64 + lane;
        const int v = a.item_base + h;
        const uint32_t cv = a.colors[(int64_t)v * B + b];
        // a lane past the row's end has the empty set: it adds nothing
        const uint32_t amask = i < Wy ? hub_expand_bit(a.ymask[i], cv) : 0u;
        // the lane's index for each neighbour colour, once per work unit (a
        // lane writes and reads only its own column: no barrier)
        for (int cu = 0; cu < a.k; ++cu)
            s_tab[cu * 64 + lane] = (amask >> cu) & 1u
                                        ? s_rank[hub_squeeze_bit(amask & ~(1u << (uint32_t)cu), (uint32_t)cu)]
                                        : (int16_t)-1;
        const int e0 = a.rowptr[v], e1 = a.rowptr[v + 1];
        const T* rows = a.aux + (int64_t)b * a.Wa;
        T acc = T(0);
        // lane j < NB holds entry e + j of the batch; an entry past the end
        // reads the list's first entry with colour col(v), which no set holds
        int um = 0;
        uint32_t cm = cv;
        T wm = T(1);
        auto meta = [&](int e) {
            const int ee = e + lane;
            const bool ok = lane < NB && ee < e1;
            const int ec = ok ? ee : e0;
            um = a.col[ec];
            cm = ok ? (uint32_t)a.ecol[(int64_t)ec * B + b] : cv;
            wm = a.w ? a.w[ec] : T(1);
        };
        if (e0 < e1) meta(e0);
        for (int e = e0; e < e1; e += NB) {
            // the batch's row indices from the lane's table (-1: cu is not
            // in the set)
            int idx[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const uint32_t cu = (uint32_t)__builtin_amdgcn_readlane((int)cm, j);
                idx[j] = s_tab[cu * 64 + lane];
            }
            T x[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int u = __builtin_amdgcn_readlane(um, j);
                // every lane loads (index 0 for "no value") and selects
                const T val = rows[(int64_t)u * a.lda + (idx[j] < 0 ? 0 : idx[j])];
                x[j] = idx[j] < 0 ? T(0) : val;
            }
            T wv[NB];
            if (a.w) {
#pragma unroll
                for (int j = 0; j < NB; ++j) wv[j] = lane_value<T>(wm, j);
            }
            if (e + NB < e1) meta(e + NB);  // the next batch's ids load while these rows arrive
            if (a.w) {
#pragma unroll
                for (int j = 0; j < NB; ++j) acc = acc + wv[j] * x[j];
            } else {
#pragma unroll
                for (int j = 0; j < NB; ++j) acc = acc + x[j];
            }
        }
        if (i < Wy) a.out[(int64_t)h * a.ldo + (int64_t)b * Wy + i] = acc;
    }
}

template <typename T, int NB>
hipError_t launch_hub_seq_t(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    const size_t lds = hub_seq_lds_bytes(a.K);
    const int64_t waves = (int64_t)a.n_items * a.B * ((a.Wy + 63) / 64);
    const int64_t need = (waves + kBlock / 64 - 1) / (kBlock / 64);
    if (need <= 0) return hipSuccess;
    static thread_local std::map<size_t, int> per_cu_cache;
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        auto it = per_cu_cache.find(lds);
        if (it != per_cu_cache.end()) {
            per_cu = it->second;
        } else {
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hub_seq<T, NB>, kBlock, lds);
            if (e != hipSuccess) return e;
            if (per_cu < 1) return hipErrorInvalidConfiguration;
            per_cu_cache[lds] = per_cu;
        }
    }
    const int64_t cap = (int64_t)per_cu * num_cus;
    const int grid = (int)(need < cap ? need : cap);
    hipLaunchKernelGGL((k_hub_seq<T, NB>), dim3(grid), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

}  // namespace

size_t hub_seq_lds_bytes(int K) { return hub_rank_bytes(K) + (size_t)(kBlock / 64) * kMaxK * 64 * sizeof(int16_t); }

template <typename T>
hipError_t launch_hub_seq(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    if (a.n_items <= 0) return hipSuccess;
    if (a.K < 1 || a.K > kMaxK - 1 || a.k > 27 || !a.ymask || !a.rank_tab || a.out_mode != OUT_PARTIAL)
        return hipErrorInvalidValue;
    // rows of up to 64 values (a wavefront per hub and coloring): the largest
    // hub's chain bounds the launch, so a deeper batch
    if constexpr (sizeof(T) == 4) {
        if (a.Wy <= 64) return launch_hub_seq_t<T, kHubBatchNarrow>(a, num_cus, blocks_per_cu, st);
    }
    return launch_hub_seq_t<T, kHubBatch>(a, num_cus, blocks_per_cu, st);
}

template hipError_t launch_hub_seq<float>(const SubArgs<float>&, int, int, hipStream_t);
template hipError_t launch_hub_seq<double>(const SubArgs<double>&, int, int, hipStream_t);

}  // namespace pgbsc
