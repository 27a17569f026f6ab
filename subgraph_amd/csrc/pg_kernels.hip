// HIP kernels (gfx950) for one subtemplate of the pruned color-coding DP.
//
// One kernel does the whole per-subtemplate body of the reference's
// countNonBottomePrunedSPMM (CountMat.cpp:556-849): the SpMM of the adjacency
// against the aux child's table (CSCGraph::spmmSplit, CSCGraph.hpp:243-310)
// and the element-wise multiply-add over color-set splits
// (DataTableColMajor::arrayWiseFMA*, DataTableColMajor.cpp:346-500), fused so
// the SpMM result never leaves the workgroup.
//
// Layout: tables are vertex-major, row v holds B colorings x C color-set
// columns (column = colex rank), so a neighbour gather reads one contiguous
// row segment.  A "team" of TEAM lanes (8..64) owns one work item; its lanes
// split the row's columns.  Teams of a workgroup move through the phases in
// lockstep: (1) gather Y into registers, then LDS; (2) stage the main row in
// LDS; (3) every lane produces output sets, reading the split table and LDS.
//
// FP32 numerics follow the reference: neighbours are summed in the order of
// the row (ascending caller id), splits in the reference's j order, with
// separate multiply and add (this file is compiled with -ffp-contract=off);
// the scaled subtemplate and the root use FP64 exactly as
// DataTableColMajor.cpp:367 and :396 do.
#include "pg_kernels.hpp"

namespace pgbsc {

namespace {

__device__ __forceinline__ uint64_t d_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

template <typename T, int AUX>
__device__ __forceinline__ T aux_value(const SubArgs<T>& a, int u, int c, int64_t lda)
{
    if (AUX == AUX_GATHER) {
        return a.aux[(int64_t)u * lda + c];
    } else {
        // one-hot leaf: column c = (coloring b, color c % k)
        const int b = c / a.k;
        return a.colors[(int64_t)u * a.B + b] == (uint8_t)(c - b * a.k) ? T(1) : T(0);
    }
}

}  // namespace

template <typename T, int TEAM, int AUX, int MAIN>
__global__ __launch_bounds__(kBlock) void k_sub(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TEAMS = kBlock / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    const int item = blockIdx.x * TEAMS + team;
    const bool active = item < a.n_items;
    const int B = a.B;
    const int nY = B * a.Ca;
    const int nM = (MAIN == MAIN_TABLE) ? B * a.Cm : 0;
    const int per_team = nY + nM;
    T* sY = reinterpret_cast<T*>(smem_raw) + (size_t)team * per_team;
    T* sM = sY + nY;

    int v = 0, e0 = 0, e1 = 0, slot = 0;
    if (active) {
        if (a.items) {
            const int4 it = a.items[item];
            v = it.x; e0 = it.y; e1 = it.z; slot = it.w;
        } else {
            v = a.item_base + item;
            if (AUX == AUX_GATHER || AUX == AUX_COLORS) { e0 = a.rowptr[v]; e1 = a.rowptr[v + 1]; }
        }
    }
    const int mode = a.out_mode;

    // ---------------------------------------------------------- phase 1: Y
    if (AUX == AUX_GATHER || AUX == AUX_COLORS) {
        for (int c0 = 0; c0 < nY; c0 += TEAM * kRegCols) {
            T acc[kRegCols];
            int cc[kRegCols];
#pragma unroll
            for (int r = 0; r < kRegCols; ++r) {
                acc[r] = T(0);
                cc[r] = c0 + lane + r * TEAM;
            }
            if (active) {
                int e = e0;
                for (; e + 4 <= e1; e += 4) {
                    int u[4];
                    T wv[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        u[q] = a.col[e + q];
                        wv[q] = a.w ? a.w[e + q] : T(1);
                    }
                    T x[4][kRegCols];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int r = 0; r < kRegCols; ++r)
                            x[q][r] = cc[r] < nY ? aux_value<T, AUX>(a, u[q], cc[r], a.lda) : T(0);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int r = 0; r < kRegCols; ++r)
                            acc[r] = a.w ? acc[r] + wv[q] * x[q][r] : acc[r] + x[q][r];
                }
                for (; e < e1; ++e) {
                    const int u = a.col[e];
                    const T wv = a.w ? a.w[e] : T(1);
#pragma unroll
                    for (int r = 0; r < kRegCols; ++r) {
                        const T x = cc[r] < nY ? aux_value<T, AUX>(a, u, cc[r], a.lda) : T(0);
                        acc[r] = a.w ? acc[r] + wv * x : acc[r] + x;
                    }
                }
            }
            if (active) {
#pragma unroll
                for (int r = 0; r < kRegCols; ++r) {
                    if (cc[r] >= nY) continue;
                    if (mode == OUT_Y) a.out[(int64_t)v * a.ldo + cc[r]] = acc[r];
                    else if (mode == OUT_PARTIAL) a.out[(int64_t)slot * a.ldo + cc[r]] = acc[r];
                    else sY[cc[r]] = acc[r];
                }
            }
        }
    } else if (AUX == AUX_READ) {
        if (active)
            for (int c = lane; c < nY; c += TEAM) sY[c] = a.aux[(int64_t)v * a.lda + c];
    } else {  // AUX_PARTIALS: item = hub index, v = hub vertex
        if (active) {
            const int s0 = a.slot_begin[item], s1 = a.slot_begin[item + 1];
            for (int c = lane; c < nY; c += TEAM) {
                T acc = T(0);
                for (int sl = s0; sl < s1; ++sl) acc = acc + a.aux[(int64_t)sl * a.lda + c];
                if (mode == OUT_Y) a.out[(int64_t)v * a.ldo + c] = acc;
                else sY[c] = acc;
            }
        }
    }
    if (mode == OUT_Y || mode == OUT_PARTIAL) return;  // uniform over the grid

    // ------------------------------------------------------ phase 2: main
    if (MAIN == MAIN_TABLE && active)
        for (int c = lane; c < nM; c += TEAM) sM[c] = a.main[(int64_t)v * a.ldm + c];
    __syncthreads();

    // --------------------------------------------------- phase 3: combine
    const int Cs = a.Cs, P = a.P, Ca = a.Ca, Cm = a.Cm;
    const int nO = B * Cs;
    if (mode == OUT_ROOT) {
        // Cs == 1: lane b owns coloring b of this vertex; the per-vertex FP64
        // value goes to root_vals[b][v] and is summed later in a fixed order
        if (active) {
            for (int b = lane; b < B; b += TEAM) {
                double val = 0.0;
                const T* y = sY + b * Ca;
                if (MAIN == MAIN_LEAF) {
                    const int c = a.colors[(int64_t)v * B + b];
                    const int ai = a.drop[c];
                    // f32 product y * 1, widened (DataTableColMajor.cpp:396)
                    val = ai >= 0 ? (double)y[ai] : 0.0;
                } else {
                    const T* m = sM + b * Cm;
                    for (int j = 0; j < P; ++j) {
                        const uint32_t p = a.split[j];
                        const T prod = y[p & 0xffffu] * m[p >> 16];
                        val = val + (double)prod;
                    }
                }
                a.root_vals[(int64_t)b * a.n_root + v] = val;
            }
        }
        return;
    }
    if (active) {
        for (int o = lane; o < nO; o += TEAM) {
            const int b = o / Cs;
            const int r = o - b * Cs;
            const T* y = sY + b * Ca;
            T res;
            if (MAIN == MAIN_LEAF) {
                // only the split whose main set is {col(v)} has a non-zero
                // main value (1.0): res = 0 + 1 * y exactly
                const int c = a.colors[(int64_t)v * B + b];
                const int ai = a.drop[r * a.k + c];
                const T yv = ai >= 0 ? y[ai] : T(0);
                if (mode == OUT_EMA_SCALED) res = (T)((double)yv * (double)a.scale);
                else res = yv;
            } else {
                const T* m = sM + b * Cm;
                const uint32_t* sp = a.split + (size_t)r * P;
                if (mode == OUT_EMA_SCALED) {
                    // DataTableColMajor.cpp:367: dst + (a * (double)b) * scale
                    T dst = T(0);
                    for (int j = 0; j < P; ++j) {
                        const uint32_t p = sp[j];
                        const double t = ((double)y[p & 0xffffu] * (double)m[p >> 16]) * (double)a.scale;
                        dst = (T)((double)dst + t);
                    }
                    res = dst;
                } else {
                    T acc = T(0);
                    for (int j = 0; j < P; ++j) {
                        const uint32_t p = sp[j];
                        const T prod = y[p & 0xffffu] * m[p >> 16];
                        acc = acc + prod;
                    }
                    res = acc;
                }
            }
            a.out[(int64_t)v * a.ldo + o] = res;
        }
    }
}

template <typename T, int TEAM>
static hipError_t launch_team(int aux, int main_mode, const SubArgs<T>& a, int nblocks, size_t lds, hipStream_t st)
{
#define PG_LAUNCH(AUXV, MAINV) \
    hipLaunchKernelGGL((k_sub<T, TEAM, AUXV, MAINV>), dim3(nblocks), dim3(kBlock), lds, st, a)
    if (main_mode == MAIN_LEAF) {
        switch (aux) {
        case AUX_GATHER: PG_LAUNCH(AUX_GATHER, MAIN_LEAF); break;
        case AUX_COLORS: PG_LAUNCH(AUX_COLORS, MAIN_LEAF); break;
        case AUX_READ: PG_LAUNCH(AUX_READ, MAIN_LEAF); break;
        case AUX_PARTIALS: PG_LAUNCH(AUX_PARTIALS, MAIN_LEAF); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (aux) {
        case AUX_GATHER: PG_LAUNCH(AUX_GATHER, MAIN_TABLE); break;
        case AUX_READ: PG_LAUNCH(AUX_READ, MAIN_TABLE); break;
        case AUX_PARTIALS: PG_LAUNCH(AUX_PARTIALS, MAIN_TABLE); break;
        default: return hipErrorInvalidValue;
        }
    }
#undef PG_LAUNCH
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sub(int team, int aux, int main_mode, const SubArgs<T>& a, int nblocks, size_t lds, hipStream_t st)
{
    if (nblocks <= 0) return hipSuccess;
    switch (team) {
    case 8: return launch_team<T, 8>(aux, main_mode, a, nblocks, lds, st);
    case 16: return launch_team<T, 16>(aux, main_mode, a, nblocks, lds, st);
    case 32: return launch_team<T, 32>(aux, main_mode, a, nblocks, lds, st);
    case 64: return launch_team<T, 64>(aux, main_mode, a, nblocks, lds, st);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_sub<float>(int, int, int, const SubArgs<float>&, int, size_t, hipStream_t);
template hipError_t launch_sub<double>(int, int, int, const SubArgs<double>&, int, size_t, hipStream_t);

// ------------------------------------------------------------ utilities ---

__global__ void k_colors_rng(uint8_t* out, const int32_t* perm, int32_t n_act, int B, uint64_t seed,
                             int64_t first, int k)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n_act * B) return;
    const int64_t v = i / B;
    const int b = (int)(i - v * B);
    const uint64_t h = d_splitmix64(seed ^ d_splitmix64((uint64_t)(first + b) * 0x9E3779B97F4A7C15ull +
                                                         (uint64_t)perm[v]));
    out[i] = (uint8_t)(((h >> 32) * (uint64_t)k) >> 32);
}

hipError_t launch_colors_rng(uint8_t* out, const int32_t* perm, int32_t n_act, int B, uint64_t seed,
                             int64_t first, int k, hipStream_t st)
{
    const int64_t total = (int64_t)n_act * B;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_colors_rng, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, out, perm, n_act, B,
                       seed, first, k);
    return hipGetLastError();
}

__global__ void k_colors_gather(uint8_t* out, const uint8_t* in, const int32_t* perm, int32_t n_act, int B,
                                int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n_act * B) return;
    const int64_t v = i / B;
    const int b = (int)(i - v * B);
    out[i] = in[(int64_t)b * n + perm[v]];
}

hipError_t launch_colors_gather(uint8_t* out, const uint8_t* in, const int32_t* perm, int32_t n_act, int B,
                                int64_t n, hipStream_t st)
{
    const int64_t total = (int64_t)n_act * B;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_colors_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, out, in, perm,
                       n_act, B, n);
    return hipGetLastError();
}

// One workgroup per coloring; thread t sums vertices t, t+1024, ... in order,
// then a fixed-shape tree: the result depends neither on timing nor on how
// the root values were produced (batch tile, team size, hub split).
__global__ __launch_bounds__(1024) void k_root_reduce(const double* vals, int64_t n, int B, double* out)
{
    __shared__ double s[1024];
    const int b = blockIdx.x;
    const double* row = vals + (int64_t)b * n;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) acc += row[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = s[0];
}

hipError_t launch_root_reduce(const double* vals, int64_t n, int B, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_root_reduce, dim3(B), dim3(1024), 0, st, vals, n, B, out);
    return hipGetLastError();
}

template <typename T>
__global__ void k_gather_weights(T* dst, const float* src, const int32_t* ent, int64_t nnz)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) dst[i] = (T)src[ent[i]];
}

hipError_t launch_gather_weights_f32(float* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st)
{
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_weights<float>, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, dst, src,
                       ent, nnz);
    return hipGetLastError();
}

hipError_t launch_gather_weights_f64(double* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st)
{
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_weights<double>, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, dst, src,
                       ent, nnz);
    return hipGetLastError();
}

__global__ void k_fill_zero(uint32_t* dst, int64_t words)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) dst[i] = 0u;
}

hipError_t launch_fill_zero(void* dst, size_t bytes, hipStream_t st)
{
    const int64_t words = (int64_t)(bytes / 4);
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, (uint32_t*)dst, words);
    return hipGetLastError();
}

}  // namespace pgbsc
