// Neighbour sums over wide count-table rows, read whole and coalesced, one
// neighbour colour at a time ("colour runs").
//
// The reference's SpMM (CSCGraph::spmmSplit, CSCGraph.hpp:243-310) adds the
// aux child's table columns of every neighbour u into the vertex's row.  In
// the colour-relative layout (pg_kernels.hpp) the position of a colour set in
// u's row depends on u's colour, so the gathering kernel (k_sub) reads, per
// neighbour, the few values each lane needs at scattered positions.  Here a
// team of 64 lanes owns one (vertex, colouring) pair and takes the neighbours
// colour by colour: for colour c it adds the whole rows of the neighbours of
// colour c (lane l holds positions l, l+64, ...: coalesced loads, no index
// arithmetic per neighbour), then moves the sum, once, to the positions of
// the vertex's row it belongs to.  That map depends only on (col(v), c):
//   forward (RUNS_FWD): position j of a row of colour c is the set
//     R_j + {c}; it adds to Y set R_j + {c} when that set avoids col(v);
//   transposed (RUNS_T, the adjoint's G_aux): position t of a neighbour's dY
//     row of colour c is the set A_t (A_t avoids c); it adds to the vertex's
//     row entry A_t - {col(v)} when A_t holds col(v).
// Neighbour sums are therefore formed per colour, in row order, and the
// colours are added in increasing order: deterministic, and a different
// FP32 rounding order than the reference's (within the parity tolerance).
// After the sum the team finishes like k_sub: the eMA combine into the
// output row (or the root value), a hub chunk's partial row, or the row
// itself.
#include "pg_kernels.hpp"

#include <map>

namespace pgbsc {

namespace {

__device__ __forceinline__ uint32_t expand_bit_r(uint32_t m, uint32_t c)
{
    return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u));
}

__device__ __forceinline__ uint32_t squeeze_bit_r(uint32_t m, uint32_t c)
{
    return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c);
}

// RV consecutive int16 values (RV * 2 bytes, aligned) with one load
template <int RV>
__device__ __forceinline__ void load_i16(const int16_t* p, int16_t (&v)[RV])
{
    if constexpr (RV == 8) {
        const int4 w = *reinterpret_cast<const int4*>(p);
        const int ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[2 * q] = (int16_t)(ws[q] & 0xffff);
            v[2 * q + 1] = (int16_t)(ws[q] >> 16);
        }
    } else if constexpr (RV == 4) {
        const int2 w = *reinterpret_cast<const int2*>(p);
        v[0] = (int16_t)(w.x & 0xffff);
        v[1] = (int16_t)(w.x >> 16);
        v[2] = (int16_t)(w.y & 0xffff);
        v[3] = (int16_t)(w.y >> 16);
    } else if constexpr (RV == 2) {
        const int w = *reinterpret_cast<const int*>(p);
        v[0] = (int16_t)(w & 0xffff);
        v[1] = (int16_t)(w >> 16);
    } else {
        v[0] = p[0];
    }
}

// lane l's value of x, in every lane (l uniform)
__device__ __forceinline__ float read_lane(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

__device__ __forceinline__ double read_lane(double x, int l)
{
    const int64_t b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__host__ __device__ constexpr size_t rank_bytes(int K) { return (((size_t)2 << K) + 15) & ~(size_t)15; }

// LDS: rank table, optional split table [P][Wo], then per team the summed
// row (Wy values) and, for the combine of a main table, the main row
template <typename T>
__host__ __device__ size_t runs_head_bytes(const SubArgs<T>& a)
{
    const bool combine = a.out_mode == OUT_EMA || a.out_mode == OUT_EMA_SCALED;
    return rank_bytes(a.K) + (a.split_lds && combine ? (((size_t)4 * a.P * a.Wo + 15) & ~(size_t)15) : 0);
}

template <typename T>
__host__ __device__ int runs_main_cols(int main_mode, const SubArgs<T>& a)
{
    const bool comb = a.out_mode == OUT_EMA || a.out_mode == OUT_EMA_SCALED || a.out_mode == OUT_ROOT;
    return main_mode == MAIN_TABLE && comb ? a.Wm : 0;
}

// one output of the combine, the splits in the reference's order j (the same
// operations as k_sub's combine)
template <typename T, typename Split>
__device__ __forceinline__ T runs_combine(Split split_at, int P, const T* y, const T* m, int mode, float scale)
{
    constexpr int U = 8;
    if (mode == OUT_EMA_SCALED) {
        T dst = T(0);
        int j = 0;
        for (; j + U <= P; j += U) {
            uint32_t p[U];
#pragma unroll
            for (int q = 0; q < U; ++q) p[q] = split_at(j + q);
            T yy[U], mm[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                yy[q] = y[p[q] & 0xffffu];
                mm[q] = m[p[q] >> 16];
            }
#pragma unroll
            for (int q = 0; q < U; ++q) dst = (T)((double)dst + ((double)yy[q] * (double)mm[q]) * (double)scale);
        }
        for (; j < P; ++j) {
            const uint32_t p = split_at(j);
            dst = (T)((double)dst + ((double)y[p & 0xffffu] * (double)m[p >> 16]) * (double)scale);
        }
        return dst;
    }
    T acc = T(0);
    int j = 0;
    for (; j + U <= P; j += U) {
        uint32_t p[U];
#pragma unroll
        for (int q = 0; q < U; ++q) p[q] = split_at(j + q);
        T yy[U], mm[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            yy[q] = y[p[q] & 0xffffu];
            mm[q] = m[p[q] >> 16];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const T prod = yy[q] * mm[q];
            acc = acc + prod;
        }
    }
    for (; j < P; ++j) {
        const uint32_t p = split_at(j);
        const T prod = y[p & 0xffffu] * m[p >> 16];
        acc = acc + prod;
    }
    return acc;
}

// The team-wide sums of U per-lane values (U a power of two, <= 64) by a
// butterfly reduce-scatter: at the step of lane bit s < U each lane keeps
// half of its values (the half its bit selects) and adds the partner's copy
// of them, so after log2(U) steps a lane holds one value, index *q (its low
// lane bits reversed), summed over its U-lane group; xor steps over the
// higher lane bits finish the sums in every lane.  (U - 1) + 6 - log2(U)
// shuffles for U sums instead of 6 each; a fixed order, the same in every
// lane holding a given sum.
template <typename T, int U>
__device__ __forceinline__ T team_sums(const T (&d)[U], int lane, int* q)
{
    T v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) v[j] = d[j];
    int idx = 0;
#pragma unroll
    for (int s = 1, n = U; s < U; s <<= 1, n >>= 1) {
        const bool h = (lane & s) != 0;
#pragma unroll
        for (int j = 0; j < n / 2; ++j) {
            const T keep = h ? v[j + n / 2] : v[j];
            const T send = h ? v[j] : v[j + n / 2];
            v[j] = keep + __shfl_xor(send, s, 64);
        }
        idx += h ? n / 2 : 0;
    }
#pragma unroll
    for (int s = U; s < 64; s <<= 1) v[0] = v[0] + __shfl_xor(v[0], s, 64);
    *q = idx;
    return v[0];
}

}  // namespace

template <typename T>
size_t runs_lds_bytes(int main_mode, const SubArgs<T>& a)
{
    return runs_head_bytes(a) + (size_t)(kBlock / 64) * (a.Wy + runs_main_cols(main_mode, a)) * sizeof(T);
}

template <typename T>
size_t runs_g_lds_bytes(const SubArgs<T>& a)
{
    return rank_bytes(a.K) + (size_t)(kBlock / 64) * 2 * a.Wy * sizeof(T);
}

template size_t runs_g_lds_bytes<float>(const SubArgs<float>&);
template size_t runs_g_lds_bytes<double>(const SubArgs<double>&);

template size_t runs_lds_bytes<float>(int, const SubArgs<float>&);
template size_t runs_lds_bytes<double>(int, const SubArgs<double>&);

namespace {

// RV: row values per lane (a gathered row has at most 64 * RV values); U:
// neighbour rows in flight per team
template <typename T, int RV, int U, int MODE, int MAIN>
__global__ __launch_bounds__(kBlock) void k_runs(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TEAM = 64;
    constexpr int TEAMS = kBlock / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    const int B = a.B;
    const int mode = a.out_mode;
    const int Wg = a.Wa;  // values per colouring of a gathered row
    const int Wt = a.Wy;  // values per colouring of the summed row
    const int nM = runs_main_cols(MAIN, a);
    int16_t* s_rank = reinterpret_cast<int16_t*>(smem_raw);
    uint32_t* s_split = reinterpret_cast<uint32_t*>(smem_raw + rank_bytes(a.K));
    const bool use_split_lds = a.split_lds && MAIN == MAIN_TABLE && (mode == OUT_EMA || mode == OUT_EMA_SCALED);
    T* sY = reinterpret_cast<T*>(smem_raw + runs_head_bytes(a)) + (size_t)team * (Wt + nM);
    T* sM = sY + Wt;
    for (int i = threadIdx.x; i < (1 << a.K); i += kBlock) s_rank[i] = a.rank_tab[i];
    if (use_split_lds)
        for (int i = threadIdx.x; i < a.P * a.Wo; i += kBlock) {
            const int r = i / a.P;
            s_split[(i - r * a.P) * a.Wo + r] = a.split[i];
        }
    // set mask of each row position this lane gathers
    uint32_t xm[RV];
#pragma unroll
    for (int t = 0; t < RV; ++t) {
        const int j = lane + TEAM * t;
        xm[t] = j < Wg ? (uint32_t)a.xmask[j] : 0u;
    }
    __syncthreads();

    const int64_t n_work = (int64_t)a.n_items * B;
    for (int64_t base = (int64_t)blockIdx.x * TEAMS; base < n_work; base += (int64_t)gridDim.x * TEAMS) {
        const int64_t work = base + team;
        const bool active = work < n_work;
        int v = 0, e0 = 0, e1 = 0, slot = 0, b = 0;
        if (active) {
            const int64_t item = work / B;
            b = (int)(work - item * B);
            if (a.items) {
                const int4 it = a.items[item];
                v = it.x; e0 = it.y; e1 = it.z; slot = it.w;
            } else {
                v = a.item_base + (int)item;
                e0 = a.rowptr[v];
                e1 = a.rowptr[v + 1];
            }
        }
        for (int t = lane; t < Wt; t += TEAM) sY[t] = T(0);
        wave_sync();
        if (active) {
            const uint32_t cv = a.colors[(int64_t)v * B + b];
            const T* rows = a.aux + (int64_t)b * Wg;
            // colours present among the neighbours (one bit each)
            uint32_t present = 0;
            for (int e = e0 + lane; e < e1; e += TEAM) present |= 1u << a.ecol[(int64_t)e * B + b];
            for (int sh = 1; sh < TEAM; sh <<= 1) present |= __shfl_xor(present, sh, TEAM);
            present &= ~(1u << cv);  // a neighbour of the vertex's own colour adds nothing
            while (present) {
                const uint32_t c = (uint32_t)(__ffs(present) - 1);
                present &= present - 1u;
                // the entry of the vertex's row each of the lane's row
                // positions adds to for colour c (-1: none; such a position
                // is never read)
                int idx[RV];
                T acc[RV];
#pragma unroll
                for (int t = 0; t < RV; ++t) {
                    const int j = lane + TEAM * t;
                    acc[t] = T(0);
                    idx[t] = -1;
                    if (j >= Wg) continue;
                    if (MODE == RUNS_FWD) {
                        const uint32_t S = expand_bit_r(xm[t], c) | (1u << c);
                        if (!((S >> cv) & 1u)) idx[t] = s_rank[squeeze_bit_r(S, cv)];
                    } else {
                        const uint32_t A = expand_bit_r(xm[t], c);
                        if ((A >> cv) & 1u) idx[t] = s_rank[squeeze_bit_r(A & ~(1u << cv), cv)];
                    }
                }
                for (int e = e0; e < e1; e += TEAM) {
                    const int ee = e + lane;
                    const bool ok = ee < e1;
                    const uint32_t cc = ok ? (uint32_t)a.ecol[(int64_t)ee * B + b] : 0xffu;
                    const int uid = ok ? a.col[ee] : 0;
                    const T wv = ok && a.w ? a.w[ee] : T(1);
                    uint64_t m = __ballot(ok && cc == c);
                    while (m) {
                        int src[U];
#pragma unroll
                        for (int q = 0; q < U; ++q) {
                            src[q] = m ? __ffsll((unsigned long long)m) - 1 : -1;
                            m &= m ? m - 1 : 0;
                        }
                        T x[U][RV];
                        T wq[U];
#pragma unroll
                        for (int q = 0; q < U; ++q) {
                            // a slot past the last match reads the first match's row, weight 0
                            const int s = src[q] >= 0 ? src[q] : src[0];
                            const int u = __shfl(uid, s, TEAM);
                            wq[q] = src[q] >= 0 ? __shfl(wv, s, TEAM) : T(0);
                            const T* row = rows + (int64_t)u * a.lda;
#pragma unroll
                            for (int t = 0; t < RV; ++t) {
                                const int j = lane + TEAM * t;
                                x[q][t] = idx[t] >= 0 ? row[j] : T(0);
                            }
                        }
#pragma unroll
                        for (int q = 0; q < U; ++q) {
                            if (src[q] < 0) break;
#pragma unroll
                            for (int t = 0; t < RV; ++t) acc[t] = a.w ? acc[t] + wq[q] * x[q][t] : acc[t] + x[q][t];
                        }
                    }
                }
                // move the colour's sum to the vertex's row: distinct targets
                // per lane and position, one wavefront per team (in order)
#pragma unroll
                for (int t = 0; t < RV; ++t)
                    if (idx[t] >= 0) sY[idx[t]] = sY[idx[t]] + acc[t];
            }
        }
        wave_sync();
        if (mode == OUT_Y || mode == OUT_PARTIAL) {
            if (active) {
                T* dst = a.out + (int64_t)(mode == OUT_Y ? v : slot) * a.ldo + (int64_t)b * Wt;
                for (int t = lane; t < Wt; t += TEAM) dst[t] = sY[t];
            }
            wave_sync();
            continue;
        }
        if (active && nM)
            for (int t = lane; t < nM; t += TEAM) sM[t] = a.main[(int64_t)v * a.ldm + (int64_t)b * nM + t];
        if (active && a.ykeep)  // Y kept for the adjoint
            for (int t = lane; t < Wt; t += TEAM) a.ykeep[(int64_t)v * a.ldy + (int64_t)b * Wt + t] = sY[t];
        wave_sync();
        if (active) {
            const int P = a.P;
            if (mode == OUT_ROOT) {
                if (lane == 0) {
                    double val = 0.0;
                    if (MAIN == MAIN_LEAF) {
                        val = (double)sY[0];  // f32 product y * 1, widened (DataTableColMajor.cpp:396)
                    } else {
                        for (int j = 0; j < P; ++j) {
                            const uint32_t p = a.split[j];
                            const T prod = sY[p & 0xffffu] * sM[p >> 16];
                            val = val + (double)prod;
                        }
                    }
                    a.root_vals[(int64_t)b * a.n_root + v] = val;
                }
            } else {
                const int Wo = a.Wo;
                T* dst = a.out + (int64_t)v * a.ldo + (int64_t)b * Wo;
                for (int r = lane; r < Wo; r += TEAM) {
                    T res;
                    if (MAIN == MAIN_LEAF) {
                        res = mode == OUT_EMA_SCALED ? (T)((double)sY[r] * (double)a.scale) : sY[r];
                    } else if (use_split_lds) {
                        res = runs_combine<T>([&](int j) { return s_split[j * Wo + r]; }, P, sY, sM, mode, a.scale);
                    } else {
                        res = runs_combine<T>([&](int j) { return a.split[(size_t)r * P + j]; }, P, sY, sM, mode,
                                              a.scale);
                    }
                    dst[r] = res;
                }
            }
        }
        wave_sync();
    }
}

// The adjoint's gradient gather (RUNS_G): the forward sum of Y (written
// out, OUT_Y or a hub chunk's OUT_PARTIAL) and, for every entry e = (v, u),
// grad[e] += dY[v] . x_u, the neighbour's row mapped to Y positions
// (d Y[v] / d w_e = x_u).  The FP32 dot product is summed over the team's
// lanes by a fixed xor tree and added to the FP64 gradient with one atomic.
// A team owns an item for every colouring, in order, so each entry's adds
// happen in program order and the results are deterministic.
template <typename T, int RV, int U>
__global__ __launch_bounds__(kBlock) void k_runs_g(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TEAM = 64;
    constexpr int TEAMS = kBlock / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    const int B = a.B;
    const int Wg = a.Wa, Wt = a.Wy;
    int16_t* s_rank = reinterpret_cast<int16_t*>(smem_raw);
    T* sY = reinterpret_cast<T*>(smem_raw + rank_bytes(a.K)) + (size_t)team * 2 * Wt;
    T* sD = sY + Wt;
    for (int i = threadIdx.x; i < (1 << a.K); i += kBlock) s_rank[i] = a.rank_tab[i];
    uint32_t xm[RV];
#pragma unroll
    for (int t = 0; t < RV; ++t) {
        const int j = lane + TEAM * t;
        xm[t] = j < Wg ? (uint32_t)a.xmask[j] : 0u;
    }
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * TEAMS; base < a.n_items; base += (int64_t)gridDim.x * TEAMS) {
        const int64_t item = base + team;
        const bool active = item < a.n_items;
        int v = 0, e0 = 0, e1 = 0, slot = 0;
        if (active) {
            if (a.items) {
                const int4 it = a.items[item];
                v = it.x; e0 = it.y; e1 = it.z; slot = it.w;
            } else {
                v = a.item_base + (int)item;
                e0 = a.rowptr[v];
                e1 = a.rowptr[v + 1];
            }
        }
        for (int b = 0; b < B; ++b) {
            for (int t = lane; t < Wt; t += TEAM) {
                sY[t] = T(0);
                sD[t] = active ? a.dy[(int64_t)v * a.ldd + (int64_t)b * Wt + t] : T(0);
            }
            wave_sync();
            if (active) {
                const uint32_t cv = a.colors[(int64_t)v * B + b];
                const T* rows = a.aux + (int64_t)b * Wg;
                uint32_t present = 0;
                for (int e = e0 + lane; e < e1; e += TEAM) present |= 1u << a.ecol[(int64_t)e * B + b];
                for (int sh = 1; sh < TEAM; sh <<= 1) present |= __shfl_xor(present, sh, TEAM);
                present &= ~(1u << cv);
                while (present) {
                    const uint32_t c = (uint32_t)(__ffs(present) - 1);
                    present &= present - 1u;
                    // where each of the lane's row positions lands for colour c
                    int idx[RV];
                    T dyc[RV];
                    T acc[RV];
#pragma unroll
                    for (int t = 0; t < RV; ++t) {
                        const int j = lane + TEAM * t;
                        idx[t] = -1;
                        if (j < Wg) {
                            const uint32_t S = expand_bit_r(xm[t], c) | (1u << c);
                            if (!((S >> cv) & 1u)) idx[t] = s_rank[squeeze_bit_r(S, cv)];
                        }
                        dyc[t] = idx[t] >= 0 ? sD[idx[t]] : T(0);
                        acc[t] = T(0);
                    }
                    for (int e = e0; e < e1; e += TEAM) {
                        const int ee = e + lane;
                        const bool ok = ee < e1;
                        const uint32_t cc = ok ? (uint32_t)a.ecol[(int64_t)ee * B + b] : 0xffu;
                        const int uid = ok ? a.col[ee] : 0;
                        const T wv = ok && a.w ? a.w[ee] : T(1);
                        uint64_t m = __ballot(ok && cc == c);
                        while (m) {
                            int src[U];
#pragma unroll
                            for (int q = 0; q < U; ++q) {
                                src[q] = m ? __ffsll((unsigned long long)m) - 1 : -1;
                                m &= m ? m - 1 : 0;
                            }
                            T x[U][RV];
                            T wq[U];
#pragma unroll
                            for (int q = 0; q < U; ++q) {
                                const int s = src[q] >= 0 ? src[q] : src[0];
                                const int u = __shfl(uid, s, TEAM);
                                wq[q] = __shfl(wv, s, TEAM);
                                const T* row = rows + (int64_t)u * a.lda;
                                // positions that land in no Y value are not read
#pragma unroll
                                for (int t = 0; t < RV; ++t) {
                                    const int j = lane + TEAM * t;
                                    x[q][t] = idx[t] >= 0 ? row[j] : T(0);
                                }
                            }
                            T dot[U];
#pragma unroll
                            for (int q = 0; q < U; ++q) {
                                dot[q] = T(0);
#pragma unroll
                                for (int t = 0; t < RV; ++t) dot[q] = dot[q] + dyc[t] * x[q][t];
                                if (src[q] < 0) continue;
#pragma unroll
                                for (int t = 0; t < RV; ++t)
                                    acc[t] = a.w ? acc[t] + wq[q] * x[q][t] : acc[t] + x[q][t];
                            }
                            int qs;
                            const T sum = team_sums<T, U>(dot, lane, &qs);
                            int sq = -1;
#pragma unroll
                            for (int q = 0; q < U; ++q)
                                if (q == qs) sq = src[q];
                            if (lane < U && sq >= 0) (void)unsafeAtomicAdd(&a.grad[e + sq], (double)sum);
                        }
                    }
#pragma unroll
                    for (int t = 0; t < RV; ++t)
                        if (idx[t] >= 0) sY[idx[t]] = sY[idx[t]] + acc[t];
                }
            }
            wave_sync();
            if (active) {
                T* dst = a.out + (int64_t)(a.out_mode == OUT_Y ? v : slot) * a.ldo + (int64_t)b * Wt;
                for (int t = lane; t < Wt; t += TEAM) dst[t] = sY[t];
            }
            wave_sync();
        }
    }
}

// The adjoint's transposed gather fused with the gradient (RUNS_TG).  As
// RUNS_T, a team sums, colour by colour, the dY rows of the vertex's
// neighbours into its G_aux row (position t of a dY row of colour c is the
// set A_t; it adds to the vertex's entry A_t - {col(v)} when A_t holds
// col(v)).  That entry is also where the vertex's own aux-child row enters
// the neighbour's Y value A_t, so for each neighbour u the team dots u's dY
// row with its own row mapped the same way: d Y[u] / d w_(u,v) . dY[u], the
// gradient of the transposed entry (u, v), added to grad[tr[e]] in FP64.
// With the forward's Y kept, this replaces the gradient gather (RUNS_G) and
// its second read of the aux rows.
//
// The item's entries are taken 64 at a time (a window).  The team sorts a
// window by neighbour colour into an LDS queue (ballots give each entry its
// place: colours ascending, entry order within a colour; entries of col(v)
// add nothing and are dropped), then streams the queue U rows at a time
// regardless of colour boundaries: a row load never waits on a colour's
// setup, and the colour's map (one load of the lane's indices from the
// table SubArgs::ttab, and the own-row values they select, staged in LDS
// once per coloring) is set up when the queue reaches a new colour, where
// the previous colour's sum moves into the vertex's row.  For an item
// of at most 64 entries the sums are formed exactly as colour runs form
// them; a longer item moves each colour's sum once per window.  A team owns
// an item for every coloring, in order, and each entry belongs to one item,
// so the gradient adds are deterministic.
template <typename T>
__host__ __device__ constexpr size_t runs_tg_team_bytes(int Wt)
{
    // queue (int4 per entry), its weights, the summed row and the own row
    return (16 * 64 + 64 * sizeof(T) + 2 * (size_t)Wt * sizeof(T) + 15) & ~(size_t)15;
}

// FP32 rows of up to 512 values at 5 wavefronts per SIMD: 96 VGPRs and a
// 48-byte spill, against 107 VGPRs at 4 (the 12-tree's fused gathers at
// R-MAT s22: 1,105 ms against 1,159).  FP64 keeps the compiler's choice (it
// would spill 316 bytes).
template <typename T, int RV, int U>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && RV == 8 ? 5 : 1)))
void k_runs_tg(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TEAM = 64;
    constexpr int TEAMS = kBlock / TEAM;
    constexpr uint32_t kNone = 0xffu;  // queue slot past the window's last entry
    const int lane = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    const int B = a.B;
    const int Wg = a.Wa;  // values per colouring of a neighbour's dY row
    const int Wt = a.Wy;  // values per colouring of the G_aux row (and of the own row)
    unsigned char* tb = smem_raw + (size_t)team * runs_tg_team_bytes<T>(Wt);
    int4* sQ = reinterpret_cast<int4*>(tb);  // (neighbour, transposed entry, colour, -)
    T* sW = reinterpret_cast<T*>(tb + 16 * TEAM);
    T* sY = sW + TEAM;
    T* sO = sY + Wt;
    for (int64_t base = (int64_t)blockIdx.x * TEAMS; base < a.n_items; base += (int64_t)gridDim.x * TEAMS) {
        const int64_t item = base + team;
        const bool active = item < a.n_items;
        int v = 0, e0 = 0, e1 = 0, slot = 0;
        if (active) {
            if (a.items) {
                const int4 it = a.items[item];
                v = it.x; e0 = it.y; e1 = it.z; slot = it.w;
            } else {
                v = a.item_base + (int)item;
                e0 = a.rowptr[v];
                e1 = a.rowptr[v + 1];
            }
        }
        for (int b = 0; b < B; ++b) {
            const T* own = a.own + (int64_t)v * a.ldown + (int64_t)b * Wt;
            for (int t = lane; t < Wt; t += TEAM) {
                sY[t] = T(0);
                sO[t] = active ? own[t] : T(0);
            }
            wave_sync();
            if (active) {
                const uint32_t cv = a.colors[(int64_t)v * B + b];
                const T* rows = a.aux + (int64_t)b * Wg;
                for (int w0 = e0; w0 < e1; w0 += TEAM) {
                    const int ee = w0 + lane;
                    const bool ok = ee < e1;
                    const uint32_t cc = ok ? (uint32_t)a.ecol[(int64_t)ee * B + b] : kNone;
                    // the entry's neighbour, transposed entry and weight load
                    // beside its colour
                    const int uid = ok ? a.col[ee] : 0;
                    const int tre = ok ? a.tr[ee] : 0;
                    const T wv = ok && a.w ? a.w[ee] : T(1);
                    const bool use = ok && cc != cv;
                    int pos = 0, n = 0;
                    for (int c = 0; c < a.k; ++c) {
                        const bool mine = use && cc == (uint32_t)c;
                        const uint64_t m = __ballot(mine);
                        if (mine) pos = n + __popcll(m & ((1ull << lane) - 1ull));
                        n += __popcll(m);
                    }
                    if (use) {
                        sQ[pos] = make_int4(uid, tre, (int)cc, 0);
                        sW[pos] = wv;
                    }
                    wave_sync();
                    // the current colour: each of the lane's dY positions'
                    // entry of the vertex's row (-1: none), the own value
                    // there, and the colour's running sum
                    uint32_t cur = kNone;
                    int idx[RV];
                    T xo[RV];
                    T acc[RV];
#pragma unroll
                    for (int t = 0; t < RV; ++t) {
                        idx[t] = -1;
                        xo[t] = T(0);
                        acc[t] = T(0);
                    }
                    auto flush = [&]() {
                        // the colour's sum into the vertex's row: distinct
                        // targets per lane and position, in order
#pragma unroll
                        for (int t = 0; t < RV; ++t) {
                            if (idx[t] >= 0) sY[idx[t]] = sY[idx[t]] + acc[t];
                            acc[t] = T(0);
                        }
                    };
                    auto enter = [&](uint32_t c) {
                        // the lane's RV indices for (cv, c): one load from the table
                        int16_t iv[RV];
                        load_i16<RV>(a.ttab + (((size_t)cv * a.k + c) * TEAM + lane) * RV, iv);
#pragma unroll
                        for (int t = 0; t < RV; ++t) {
                            idx[t] = iv[t];
                            xo[t] = idx[t] >= 0 ? sO[idx[t]] : T(0);
                        }
                    };
                    for (int i = 0; i < n; i += U) {
                        uint32_t cq[U];
                        int tq[U];
                        T wq[U];
                        T x[U][RV];
#pragma unroll
                        for (int q = 0; q < U; ++q) {
                            // a slot past the queue's end reads its first entry's row (ignored)
                            const int s = i + q < n ? i + q : i;
                            const int4 e = sQ[s];
                            const int u = __builtin_amdgcn_readfirstlane(e.x);
                            tq[q] = __builtin_amdgcn_readfirstlane(e.y);
                            cq[q] = i + q < n ? (uint32_t)__builtin_amdgcn_readfirstlane(e.z) : kNone;
                            wq[q] = sW[s];
                            const T* row = rows + (int64_t)u * a.lda;
#pragma unroll
                            for (int t = 0; t < RV; ++t) {
                                const int j = lane + TEAM * t;
                                x[q][t] = j < Wg ? row[j] : T(0);
                            }
                        }
                        T dot[U];
#pragma unroll
                        for (int q = 0; q < U; ++q) {
                            dot[q] = T(0);
                            if (cq[q] == kNone) continue;
                            if (cq[q] != cur) {
                                if (cur != kNone) flush();
                                cur = cq[q];
                                enter(cur);
                            }
#pragma unroll
                            for (int t = 0; t < RV; ++t) dot[q] = dot[q] + xo[t] * x[q][t];
#pragma unroll
                            for (int t = 0; t < RV; ++t)
                                acc[t] = a.w ? acc[t] + wq[q] * x[q][t] : acc[t] + x[q][t];
                        }
                        int qs;
                        const T sum = team_sums<T, U>(dot, lane, &qs);
                        bool live = false;
                        int tgt = 0;
#pragma unroll
                        for (int q = 0; q < U; ++q)
                            if (q == qs) {
                                live = cq[q] != kNone;
                                tgt = tq[q];
                            }
                        if (lane < U && live) (void)unsafeAtomicAdd(&a.grad[tgt], (double)sum);
                    }
                    if (cur != kNone) flush();
                    wave_sync();  // the next window rewrites the queue
                }
            }
            wave_sync();
            if (active) {
                T* dst = a.out + (int64_t)(a.out_mode == OUT_Y ? v : slot) * a.ldo + (int64_t)b * Wt;
                for (int t = lane; t < Wt; t += TEAM) dst[t] = sY[t];
            }
            wave_sync();
        }
    }
}

template <typename T, int RV>
hipError_t launch_runs_tg_t(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    constexpr int U = RV <= 2 ? 8 : 4;
    const size_t lds = runs_tg_lds_bytes<T>(a.K, a.Wy);
    const int64_t need = ((int64_t)a.n_items + (kBlock / 64) - 1) / (kBlock / 64);
    if (need <= 0) return hipSuccess;
    static thread_local std::map<size_t, int> per_cu_cache;
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        auto it = per_cu_cache.find(lds);
        if (it != per_cu_cache.end()) {
            per_cu = it->second;
        } else {
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_runs_tg<T, RV, U>, kBlock, lds);
            if (e != hipSuccess) return e;
            if (per_cu < 1) return hipErrorInvalidConfiguration;
            per_cu_cache[lds] = per_cu;
        }
    }
    const int64_t cap = (int64_t)per_cu * num_cus;
    const int grid = (int)(need < cap ? need : cap);
    hipLaunchKernelGGL((k_runs_tg<T, RV, U>), dim3(grid), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

template <typename T, int RV>
hipError_t launch_runs_g_t(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    constexpr int U = RV <= 2 ? 8 : 4;
    const size_t lds = rank_bytes(a.K) + (size_t)(kBlock / 64) * 2 * a.Wy * sizeof(T);
    const int64_t need = ((int64_t)a.n_items + (kBlock / 64) - 1) / (kBlock / 64);
    if (need <= 0) return hipSuccess;
    static thread_local std::map<size_t, int> per_cu_cache;
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        auto it = per_cu_cache.find(lds);
        if (it != per_cu_cache.end()) {
            per_cu = it->second;
        } else {
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_runs_g<T, RV, U>, kBlock, lds);
            if (e != hipSuccess) return e;
            if (per_cu < 1) return hipErrorInvalidConfiguration;
            per_cu_cache[lds] = per_cu;
        }
    }
    const int64_t cap = (int64_t)per_cu * num_cus;
    const int grid = (int)(need < cap ? need : cap);
    hipLaunchKernelGGL((k_runs_g<T, RV, U>), dim3(grid), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

// RUNS_SEQ: the forward sum in the reference's order.  A team per (vertex,
// colouring) reads each neighbour's row whole (coalesced), in the row's
// column order, and adds every value into the Y value it belongs to (LDS),
// neighbour after neighbour: each Y value is a running sum over the
// neighbours in the reference's order, starting from zero (CSCGraph.hpp:
// 254-307).  The target of row position j of a neighbour of colour cu is
// xidx[(cv * k + cu) * Wa4 + j] (-1 for none).  Hubs need no partial rows: a
// team takes a vertex's whole neighbour list.  Then it finishes as k_runs.
template <typename T, int RV, int U, int MAIN>
__global__ __launch_bounds__(kBlock) void k_runs_seq(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TEAM = 64;
    constexpr int TEAMS = kBlock / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    const int B = a.B;
    const int mode = a.out_mode;
    const int Wg = a.Wa, Wt = a.Wy;
    const int nM = runs_main_cols(MAIN, a);
    uint32_t* s_split = reinterpret_cast<uint32_t*>(smem_raw + rank_bytes(a.K));
    const bool use_split_lds = a.split_lds && MAIN == MAIN_TABLE && (mode == OUT_EMA || mode == OUT_EMA_SCALED);
    T* sY = reinterpret_cast<T*>(smem_raw + runs_head_bytes(a)) + (size_t)team * (Wt + nM);
    T* sM = sY + Wt;
    if (use_split_lds)
        for (int i = threadIdx.x; i < a.P * a.Wo; i += kBlock) {
            const int r = i / a.P;
            s_split[(i - r * a.P) * a.Wo + r] = a.split[i];
        }
    __syncthreads();
    const int64_t n_work = (int64_t)a.n_items * B;
    for (int64_t base = (int64_t)blockIdx.x * TEAMS; base < n_work; base += (int64_t)gridDim.x * TEAMS) {
        const int64_t work = base + team;
        const bool active = work < n_work;
        int v = 0, e0 = 0, e1 = 0, b = 0;
        if (active) {
            const int64_t item = work / B;
            b = (int)(work - item * B);
            v = a.item_base + (int)item;
            e0 = a.rowptr[v];
            e1 = a.rowptr[v + 1];
        }
        for (int t = lane; t < Wt; t += TEAM) sY[t] = T(0);
        wave_sync();
        if (active) {
            const uint32_t cv = a.colors[(int64_t)v * B + b];
            const T* rows = a.aux + (int64_t)b * Wg;
            // the vertex colour's rows of the per-lane index table
            const int16_t* xrow = a.xidx_g + (size_t)cv * a.k * TEAM * RV;
            // entries 64 at a time: lane l loads entry w0 + l's neighbour,
            // colour and weight once, and every lane reads a batch's from
            // there (v_readlane), so a batch waits only on its rows
            for (int w0 = e0; w0 < e1; w0 += TEAM) {
              const int nw = min(TEAM, e1 - w0);
              const bool okw = lane < nw;
              const int uw = okw ? a.col[w0 + lane] : 0;
              const uint32_t cw = okw ? (uint32_t)a.ecol[(int64_t)(w0 + lane) * B + b] : cv;
              const T ww = okw && a.w ? a.w[w0 + lane] : T(1);
              for (int i = 0; i < nw; i += U) {
                int u[U];
                uint32_t cu[U];
                T wq[U];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const bool ok = i + q < nw;
                    const int s = ok ? i + q : i;  // a slot past the end repeats the first (ignored)
                    u[q] = __builtin_amdgcn_readlane(uw, s);
                    cu[q] = ok ? (uint32_t)__builtin_amdgcn_readlane((int)cw, s) : cv;
                    wq[q] = read_lane(ww, s);
                }
                T x[U][RV];
                int16_t id[U][RV];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const T* row = rows + (int64_t)u[q] * a.lda;
                    // the lane's RV targets for this neighbour colour, one load
                    load_i16<RV>(xrow + ((size_t)cu[q] * TEAM + lane) * RV, id[q]);
#pragma unroll
                    for (int t = 0; t < RV; ++t) {
                        const int j = lane + TEAM * t;
                        x[q][t] = j < Wg ? row[j] : T(0);
                    }
                }
                // in the neighbours' order: every value into its Y value
#pragma unroll
                for (int q = 0; q < U; ++q) {
#pragma unroll
                    for (int t = 0; t < RV; ++t) {
                        const int yi = id[q][t];
                        if (yi >= 0) sY[yi] = a.w ? sY[yi] + wq[q] * x[q][t] : sY[yi] + x[q][t];
                    }
                }
              }
            }
        }
        wave_sync();
        if (mode == OUT_Y || mode == OUT_PARTIAL) {
            if (active) {
                T* dst = a.out + (int64_t)v * a.ldo + (int64_t)b * Wt;
                for (int t = lane; t < Wt; t += TEAM) dst[t] = sY[t];
            }
            wave_sync();
            continue;
        }
        if (active && nM)
            for (int t = lane; t < nM; t += TEAM) sM[t] = a.main[(int64_t)v * a.ldm + (int64_t)b * nM + t];
        if (active && a.ykeep)  // Y kept for the adjoint
            for (int t = lane; t < Wt; t += TEAM) a.ykeep[(int64_t)v * a.ldy + (int64_t)b * Wt + t] = sY[t];
        wave_sync();
        if (active) {
            const int P = a.P;
            if (mode == OUT_ROOT) {
                if (lane == 0) {
                    double val = 0.0;
                    if (MAIN == MAIN_LEAF) {
                        val = (double)sY[0];
                    } else {
                        for (int j = 0; j < P; ++j) {
                            const uint32_t p = a.split[j];
                            const T prod = sY[p & 0xffffu] * sM[p >> 16];
                            val = val + (double)prod;
                        }
                    }
                    a.root_vals[(int64_t)b * a.n_root + v] = val;
                }
            } else {
                const int Wo = a.Wo;
                T* dst = a.out + (int64_t)v * a.ldo + (int64_t)b * Wo;
                for (int r = lane; r < Wo; r += TEAM) {
                    T res;
                    if (MAIN == MAIN_LEAF) {
                        res = mode == OUT_EMA_SCALED ? (T)((double)sY[r] * (double)a.scale) : sY[r];
                    } else if (use_split_lds) {
                        res = runs_combine<T>([&](int j) { return s_split[j * Wo + r]; }, P, sY, sM, mode, a.scale);
                    } else {
                        res = runs_combine<T>([&](int j) { return a.split[(size_t)r * P + j]; }, P, sY, sM, mode,
                                              a.scale);
                    }
                    dst[r] = res;
                }
            }
        }
        wave_sync();
    }
}

template <typename T, int RV, int MAIN>
hipError_t launch_runs_seq_t(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    constexpr int U = RV <= 2 ? 8 : 4;
    const size_t lds = runs_lds_bytes<T>(MAIN, a);
    const int64_t need = ((int64_t)a.n_items * a.B + (kBlock / 64) - 1) / (kBlock / 64);
    if (need <= 0) return hipSuccess;
    static thread_local std::map<size_t, int> per_cu_cache;
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        auto it = per_cu_cache.find(lds);
        if (it != per_cu_cache.end()) {
            per_cu = it->second;
        } else {
            hipError_t e =
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_runs_seq<T, RV, U, MAIN>, kBlock, lds);
            if (e != hipSuccess) return e;
            if (per_cu < 1) return hipErrorInvalidConfiguration;
            per_cu_cache[lds] = per_cu;
        }
    }
    const int64_t cap = (int64_t)per_cu * num_cus;
    const int grid = (int)(need < cap ? need : cap);
    hipLaunchKernelGGL((k_runs_seq<T, RV, U, MAIN>), dim3(grid), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

template <typename T, int MAIN>
hipError_t launch_runs_seq_rv(int rv, const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    switch (rv) {
    case 1: return launch_runs_seq_t<T, 1, MAIN>(a, num_cus, blocks_per_cu, st);
    case 2: return launch_runs_seq_t<T, 2, MAIN>(a, num_cus, blocks_per_cu, st);
    case 4: return launch_runs_seq_t<T, 4, MAIN>(a, num_cus, blocks_per_cu, st);
    case 8: return launch_runs_seq_t<T, 8, MAIN>(a, num_cus, blocks_per_cu, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int RV, int MODE, int MAIN>
hipError_t launch_runs_t(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    constexpr int U = RV <= 2 ? 8 : 4;
    const size_t lds = runs_lds_bytes<T>(MAIN, a);
    const int64_t need = ((int64_t)a.n_items * a.B + (kBlock / 64) - 1) / (kBlock / 64);
    if (need <= 0) return hipSuccess;
    static thread_local std::map<size_t, int> per_cu_cache;
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        auto it = per_cu_cache.find(lds);
        if (it != per_cu_cache.end()) {
            per_cu = it->second;
        } else {
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_runs<T, RV, U, MODE, MAIN>,
                                                                         kBlock, lds);
            if (e != hipSuccess) return e;
            if (per_cu < 1) return hipErrorInvalidConfiguration;
            per_cu_cache[lds] = per_cu;
        }
    }
    const int64_t cap = (int64_t)per_cu * num_cus;
    const int grid = (int)(need < cap ? need : cap);
    hipLaunchKernelGGL((k_runs<T, RV, U, MODE, MAIN>), dim3(grid), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

template <typename T, int MODE, int MAIN>
hipError_t launch_runs_rv(int rv, const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    switch (rv) {
    case 1: return launch_runs_t<T, 1, MODE, MAIN>(a, num_cus, blocks_per_cu, st);
    case 2: return launch_runs_t<T, 2, MODE, MAIN>(a, num_cus, blocks_per_cu, st);
    case 4: return launch_runs_t<T, 4, MODE, MAIN>(a, num_cus, blocks_per_cu, st);
    case 8: return launch_runs_t<T, 8, MODE, MAIN>(a, num_cus, blocks_per_cu, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int runs_values_per_lane(int wg)
{
    const int rv = (wg + 63) / 64;
    return rv <= 1 ? 1 : rv <= 2 ? 2 : rv <= 4 ? 4 : rv <= 8 ? 8 : 0;
}

template <typename T>
hipError_t launch_runs(int mode, int main_mode, const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    if (a.n_items <= 0) return hipSuccess;
    const int rv = runs_values_per_lane(a.Wa);
    if (rv == 0 || a.K < 1 || a.K > kMaxK - 1 || a.k > 27) return hipErrorInvalidValue;
    if (mode == RUNS_SEQ) {
        if (a.items || !a.xidx_g || a.out_mode == OUT_PARTIAL) return hipErrorInvalidValue;
        return main_mode == MAIN_LEAF ? launch_runs_seq_rv<T, MAIN_LEAF>(rv, a, num_cus, blocks_per_cu, st)
                                      : launch_runs_seq_rv<T, MAIN_TABLE>(rv, a, num_cus, blocks_per_cu, st);
    }
    if (mode == RUNS_G) {
        if (main_mode != MAIN_LEAF || (a.out_mode != OUT_Y && a.out_mode != OUT_PARTIAL) || !a.dy || !a.grad)
            return hipErrorInvalidValue;
        switch (rv) {
        case 1: return launch_runs_g_t<T, 1>(a, num_cus, blocks_per_cu, st);
        case 2: return launch_runs_g_t<T, 2>(a, num_cus, blocks_per_cu, st);
        case 4: return launch_runs_g_t<T, 4>(a, num_cus, blocks_per_cu, st);
        case 8: return launch_runs_g_t<T, 8>(a, num_cus, blocks_per_cu, st);
        default: return hipErrorInvalidValue;
        }
    }
    if (mode == RUNS_T) {
        if (main_mode != MAIN_LEAF || (a.out_mode != OUT_Y && a.out_mode != OUT_PARTIAL)) return hipErrorInvalidValue;
        return launch_runs_rv<T, RUNS_T, MAIN_LEAF>(rv, a, num_cus, blocks_per_cu, st);
    }
    if (mode == RUNS_TG) {
        if (main_mode != MAIN_LEAF || (a.out_mode != OUT_Y && a.out_mode != OUT_PARTIAL) || !a.own || !a.tr ||
            !a.grad || !a.ttab)
            return hipErrorInvalidValue;
        switch (rv) {
        case 1: return launch_runs_tg_t<T, 1>(a, num_cus, blocks_per_cu, st);
        case 2: return launch_runs_tg_t<T, 2>(a, num_cus, blocks_per_cu, st);
        case 4: return launch_runs_tg_t<T, 4>(a, num_cus, blocks_per_cu, st);
        case 8: return launch_runs_tg_t<T, 8>(a, num_cus, blocks_per_cu, st);
        default: return hipErrorInvalidValue;
        }
    }
    if (mode != RUNS_FWD) return hipErrorInvalidValue;
    return main_mode == MAIN_LEAF ? launch_runs_rv<T, RUNS_FWD, MAIN_LEAF>(rv, a, num_cus, blocks_per_cu, st)
                                  : launch_runs_rv<T, RUNS_FWD, MAIN_TABLE>(rv, a, num_cus, blocks_per_cu, st);
}

template <typename T>
size_t runs_tg_lds_bytes(int K, int Wt)
{
    (void)K;  // the index table is read from memory (SubArgs::ttab)
    return (size_t)(kBlock / 64) * runs_tg_team_bytes<T>(Wt);
}

template size_t runs_tg_lds_bytes<float>(int, int);
template size_t runs_tg_lds_bytes<double>(int, int);

template hipError_t launch_runs<float>(int, int, const SubArgs<float>&, int, int, hipStream_t);
template hipError_t launch_runs<double>(int, int, const SubArgs<double>&, int, int, hipStream_t);

}  // namespace pgbsc
