// HIP kernels (gfx950) for one subtemplate of the pruned color-coding DP.
//
// One kernel does the whole per-subtemplate body of the reference's
// countNonBottomePrunedSPMM (CountMat.cpp:556-849): the SpMM of the adjacency
// against the aux child's table (CSCGraph::spmmSplit, CSCGraph.hpp:243-310)
// and the element-wise multiply-add over color-set splits
// (DataTableColMajor::arrayWiseFMA*, DataTableColMajor.cpp:346-500), fused so
// the SpMM result never leaves the workgroup.
//
// Work decomposition: a team of TEAM lanes owns one vertex (or one hub
// chunk).  The team's lanes are `ng` neighbour groups x (TEAM/ng) column
// lanes: group g sums neighbours g, g+ng, ... in row order, column lanes
// split the Y values, and the groups are folded with an xor butterfly, so
// the result is fixed by the launch shape.  Workgroups are persistent: they
// stride over items, and stage the colex rank table in LDS once.
//
// Phases per item: (1) Y into registers, then LDS; (2) the main row into
// LDS; (3) each lane produces output values from the split table and LDS.
//
// FP32 numerics follow the reference: separate multiply and add (this file
// is compiled with -ffp-contract=off), splits in the reference's j order; the
// scaled subtemplate and the root use FP64 exactly as
// DataTableColMajor.cpp:367 and :396 do.
#include "pg_kernels.hpp"

#include <algorithm>
#include <map>

namespace pgbsc {

namespace {

__device__ __forceinline__ uint64_t d_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// insert a zero bit at position c (K-bit reduced mask -> k-bit color mask)
__device__ __forceinline__ uint32_t expand_bit(uint32_t m, uint32_t c)
{
    return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u));
}

// remove bit position c (k-bit color mask -> K-bit mask relative to color c)
__device__ __forceinline__ uint32_t squeeze_bit(uint32_t m, uint32_t c)
{
    return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c);
}

// one neighbour's contribution to a value: GATHER reads the neighbour's
// color-relative row at the rank of (Y set minus the neighbour's color);
// GATHER_T reads the neighbour's dY at the rank of the vertex's row set in
// the neighbour's frame, when the neighbour's color is outside the set;
// COLORS is the one-hot leaf (amask holds the value's color)
template <typename T, int AUX>
__device__ __forceinline__ T aux_term(const SubArgs<T>& a, const int16_t* s_rank, uint32_t amask, int aoff, int u,
                                      uint32_t cu)
{
    if (AUX == AUX_GATHER || AUX == AUX_GATHER_G) {
        if ((amask >> cu) & 1u) {
            const int idx = s_rank[squeeze_bit(amask & ~(1u << cu), cu)];
            return a.aux[(int64_t)u * a.lda + aoff + idx];
        }
        return T(0);
    }
    if (AUX == AUX_GATHER_T) {
        if (!((amask >> cu) & 1u)) {
            const int idx = s_rank[squeeze_bit(amask, cu)];
            return a.aux[(int64_t)u * a.lda + aoff + idx];
        }
        return T(0);
    }
    return cu == amask ? T(1) : T(0);
}

__host__ __device__ constexpr size_t rank_lds_bytes(int K) { return (((size_t)2 << K) + 15) & ~(size_t)15; }

// per-lane gathers of neighbour rows (through the rank table or the gather index table)
__host__ __device__ constexpr bool gathers_rows(int aux)
{
    return aux == AUX_GATHER || aux == AUX_GATHER_T || aux == AUX_STAGED || aux == AUX_GATHER_G;
}
// launches that stage the rank table (and, row-wise, the aux row masks) in LDS
__host__ __device__ constexpr bool uses_rank(int aux) { return gathers_rows(aux) || aux == AUX_ROWS; }
// launches that walk a vertex's entries
__host__ __device__ constexpr bool walks_entries(int aux) { return uses_rank(aux) || aux == AUX_COLORS; }

// Index tables before the per-team rows: the gather index table (AUX_GATHER
// with gidx), else the rank table, then (AUX_ROWS) the Wa row masks
template <typename T>
__host__ __device__ size_t index_lds_bytes(int aux, const SubArgs<T>& a)
{
    if ((aux == AUX_GATHER || aux == AUX_STAGED || aux == AUX_GATHER_T || aux == AUX_GATHER_G) && a.gidx)
        return ((size_t)2 * (a.k + 1) * a.k * a.Wy4 + 15) & ~(size_t)15;
    if (aux == AUX_ROWS && a.xidx) return ((size_t)2 * a.k * a.k * a.Wa4 + 15) & ~(size_t)15;
    return !uses_rank(aux) ? 0
                           : rank_lds_bytes(a.K) + (aux == AUX_ROWS ? (((size_t)2 * a.Wa + 15) & ~(size_t)15) : 0);
}

// the split table staged for the combine, [P][Wo] uint32, after the index tables
template <typename T>
__host__ __device__ size_t split_lds_bytes(const SubArgs<T>& a)
{
    const bool combine = a.out_mode == OUT_EMA || a.out_mode == OUT_EMA_SCALED;
    return a.split_lds && combine ? (((size_t)4 * a.P * a.Wo + 15) & ~(size_t)15) : 0;
}

// LDS before the per-team rows
template <typename T>
__host__ __device__ size_t head_lds_bytes(int aux, const SubArgs<T>& a)
{
    return index_lds_bytes(aux, a) + split_lds_bytes(a);
}

}  // namespace

template <typename T>
size_t sub_lds_bytes(int team, int aux, int main_mode, const SubArgs<T>& a)
{
    const int teams = kBlock / team;
    const size_t bytes = head_lds_bytes(aux, a);
    const size_t ycopies = aux == AUX_ROWS ? (size_t)a.ng : 1;  // one Y accumulator per neighbour group
    const size_t stage = aux == AUX_STAGED ? (size_t)a.ng * stage_group_stride(a.segpad, team / a.ng) : 0;
    const bool y_in_lds = !(a.out_mode == OUT_Y || a.out_mode == OUT_PARTIAL) || aux == AUX_ROWS;
    const bool keeps_main = main_mode == MAIN_TABLE && a.out_mode != OUT_Y && a.out_mode != OUT_PARTIAL;
    const size_t per = (y_in_lds ? ycopies * a.B * a.Wy : 0) + (keeps_main ? (size_t)a.B * a.Wm : 0) + stage;
    return bytes + (size_t)teams * per * sizeof(T);
}

template size_t sub_lds_bytes<float>(int, int, int, const SubArgs<float>&);
template size_t sub_lds_bytes<double>(int, int, int, const SubArgs<double>&);

// One output value of the combine: the P splits in the reference's order j
// (DataTableColMajor.cpp:402-500 add them one after another; :367 for the
// scaled subtemplate, in FP64 with float(1e-12)).  The loads of 8 splits are
// issued before their products are added, so the split table and LDS
// latencies overlap; the additions keep their order.
template <typename T, typename Split>
__device__ __forceinline__ T combine_splits(Split split_at, int P, const T* y, const T* m, int mode, float scale)
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

template <typename T, int TEAM, int AUX, int MAIN, int RC>
__global__ __launch_bounds__(kBlock) void k_sub(const SubArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TEAMS = kBlock / TEAM;
    const int lane = threadIdx.x % TEAM;
    const int team = threadIdx.x / TEAM;
    const int B = a.B;
    const int mode = a.out_mode;
    const int nY = B * a.Wy;
    const int nM = (MAIN == MAIN_TABLE && mode != OUT_Y && mode != OUT_PARTIAL) ? B * a.Wm : 0;
    const int ng = a.ng;
    const int cl = TEAM / ng;   // column lanes per group
    const int grp = lane / cl;
    const int cln = lane - grp * cl;
    const int ycopies = AUX == AUX_ROWS ? ng : 1;
    const size_t head = head_lds_bytes(AUX, a);
    const bool use_gidx =
        (AUX == AUX_GATHER || AUX == AUX_STAGED || AUX == AUX_GATHER_T || AUX == AUX_GATHER_G) && a.gidx != nullptr;
    constexpr bool staged = AUX == AUX_STAGED;
    static_assert(RC == kRegCols || gathers_rows(AUX), "one value per lane is a gather layout");
    const bool y_in_lds = !(mode == OUT_Y || mode == OUT_PARTIAL) || AUX == AUX_ROWS;
    const int gstride = staged ? stage_group_stride(a.segpad, TEAM / ng) : 0;
    const int nstage = ng * gstride;
    const int per_team = (y_in_lds ? ycopies * nY : 0) + nM + nstage;
    int16_t* s_rank = reinterpret_cast<int16_t*>(smem_raw);
    int16_t* s_gidx = reinterpret_cast<int16_t*>(smem_raw);
    uint16_t* s_xmask = reinterpret_cast<uint16_t*>(smem_raw + rank_lds_bytes(a.K));
    T* sY = reinterpret_cast<T*>(smem_raw + head) + (size_t)team * per_team;
    T* sM = sY + (y_in_lds ? ycopies * nY : 0);
    T* s_stage = sM + nM;
    // split table of the combine, transposed to [P][Wo] so that lanes on
    // consecutive outputs read consecutive words
    uint32_t* s_split = reinterpret_cast<uint32_t*>(smem_raw + index_lds_bytes(AUX, a));
    const bool use_split_lds = a.split_lds != 0 && MAIN == MAIN_TABLE && (mode == OUT_EMA || mode == OUT_EMA_SCALED);
    if (use_split_lds) {
        const int words = a.P * a.Wo;
        for (int i = threadIdx.x; i < words; i += kBlock) {
            const int r = i / a.P;
            s_split[(i - r * a.P) * a.Wo + r] = a.split[i];
        }
    }

    const bool use_xidx = AUX == AUX_ROWS && a.xidx != nullptr;
    if (use_xidx) {
        // k*k rows of Wa4 int16: a whole number of 8-byte words
        const int words = a.k * a.k * a.Wa4 / 4;
        const int2* src = reinterpret_cast<const int2*>(a.xidx);
        int2* dst = reinterpret_cast<int2*>(s_gidx);
        for (int i = threadIdx.x; i < words; i += kBlock) dst[i] = src[i];
    } else if (use_gidx) {
        // (k+1)*k rows of Wy4 int16: a whole number of 8-byte words
        const int words = (a.k + 1) * a.k * a.Wy4 / 4;
        const int2* src = reinterpret_cast<const int2*>(a.gidx);
        int2* dst = reinterpret_cast<int2*>(s_gidx);
        for (int i = threadIdx.x; i < words; i += kBlock) dst[i] = src[i];
    } else if (uses_rank(AUX)) {
        for (int i = threadIdx.x; i < (1 << a.K); i += kBlock) s_rank[i] = a.rank_tab[i];
        if (AUX == AUX_ROWS)
            for (int i = threadIdx.x; i < a.Wa; i += kBlock) s_xmask[i] = a.xmask[i];
    }
    // the shared tables are complete; from here on each team touches only
    // its own LDS region and syncs within its wavefront
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
                if (walks_entries(AUX)) { e0 = a.rowptr[v]; e1 = a.rowptr[v + 1]; }
            }
        }

        // ------------------------------------------------------ phase 1: Y
        if (AUX == AUX_ROWS) {
            // Value q = b*Wa + j of neighbour row u is the aux child at set R_j
            // (relative to col(u)); it belongs to the Y set S = R_j + {col(u)}
            // (relative to col(v)) when col(v) is outside S, and to no other.
            // Lane cln of group grp owns values c0 + cln of each chunk and adds
            // them, neighbour after neighbour, into the group's own LDS copy of
            // Y (LDS atomics of one wavefront apply in program order, and a
            // neighbour adds to distinct sets, so the sums are deterministic);
            // the copies then fold in group order.
            T* acc = sY + grp * nY;
            for (int t = cln; t < nY; t += cl) acc[t] = T(0);
            wave_sync();
            if (active) {
                const int Wa = a.Wa;
                const int nq = B * Wa;
                for (int c0 = 0; c0 < nq; c0 += cl) {
                    const int q = c0 + cln;
                    const bool qv = q < nq;
                    const int b = qv ? q / Wa : 0;
                    const uint32_t xm = qv && !use_xidx ? (uint32_t)s_xmask[q - b * Wa] : 0u;
                    const uint32_t cv = a.colors[(int64_t)v * B + b];
                    // xidx row of this value for neighbour color 0 (past the end: -1 rows)
                    const int xbase = qv ? (int)cv * a.k * a.Wa4 + (q - b * Wa) : 0;
                    T* accb = acc + b * a.Wy;
                    const int step = ng;
                    int e = e0 + grp;
                    // a neighbour past the end has col(v) (its set holds col(v): adds nothing)
                    int un[4];
                    uint32_t cn[4];
                    T wn[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int ee = e + r * step;
                        const bool ok = ee < e1;
                        un[r] = ok ? a.col[ee] : v;
                        cn[r] = ok ? (uint32_t)a.ecol[(int64_t)ee * B + b] : cv;
                        wn[r] = ok && a.w ? a.w[ee] : T(1);
                    }
                    for (; e < e1; e += 4 * step) {
                        int u[4];
                        uint32_t cu[4];
                        T wv[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            u[r] = un[r];
                            cu[r] = cn[r];
                            wv[r] = wn[r];
                            const int ee = e + (4 + r) * step;
                            const bool ok = ee < e1;
                            un[r] = ok ? a.col[ee] : v;
                            cn[r] = ok ? (uint32_t)a.ecol[(int64_t)ee * B + b] : cv;
                            wn[r] = ok && a.w ? a.w[ee] : T(1);
                        }
                        T x[4];
                        int t[4];
                        bool ok[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            if (use_xidx) {
                                // cu == cv (also a past-the-end neighbour) reads a -1 row
                                const int tt = s_gidx[xbase + (int)__umul24(cu[r], a.Wa4)];
                                ok[r] = qv && tt >= 0;
                                t[r] = ok[r] ? tt : 0;
                            } else {
                                const uint32_t S = expand_bit(xm, cu[r]) | (1u << cu[r]);
                                ok[r] = qv && !((S >> cv) & 1u);
                                t[r] = ok[r] ? s_rank[squeeze_bit(S, cv)] : 0;
                            }
                            x[r] = ok[r] ? a.aux[(int64_t)u[r] * a.lda + q] : T(0);
                        }
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (ok[r]) atomicAdd(accb + t[r], a.w ? wv[r] * x[r] : x[r]);
                    }
                }
            }
            wave_sync();
            for (int t = lane; t < nY; t += TEAM) {
                T sum = sY[t];
                for (int g = 1; g < ng; ++g) sum = sum + sY[g * nY + t];
                if (!active) continue;
                if (mode == OUT_Y) a.out[(int64_t)v * a.ldo + t] = sum;
                else if (mode == OUT_PARTIAL) a.out[(int64_t)slot * a.ldo + t] = sum;
                else sY[t] = sum;
            }
        } else if (gathers_rows(AUX) || AUX == AUX_COLORS) {
            // a slot is RC consecutive Y values of one coloring; column
            // lane cln of every group owns slot s0 + cln of each chunk
            const int W = gathers_rows(AUX) ? a.Wy : a.k;
            const int spb = (W + RC - 1) / RC;
            const int nslots = B * spb;
            for (int s0 = 0; s0 < nslots; s0 += cl) {
                const int sid = s0 + cln;
                const bool sv = active && sid < nslots;
                const int b = sv ? sid / spb : 0;
                const int i0 = (sid - b * spb) * RC;
                T acc[RC];
                // k-bit color set of each value (GATHER: avoids col(v), 0 = unused;
                // GATHER_T: holds col(v), all ones = unused); COLORS: the value's color
                uint32_t amask[RC];
                const uint32_t cv = sv && gathers_rows(AUX) ? a.colors[(int64_t)v * B + b] : 0u;
                // gidx row of this slot's RC values for neighbour color 0
                // (a slot past the end reads the all -1 rows)
                const int gbase = sv ? ((int)cv * a.k) * a.Wy4 + i0 : a.k * a.k * a.Wy4;
#pragma unroll
                for (int r = 0; r < RC; ++r) {
                    acc[r] = T(0);
                    const int i = i0 + r;
                    const bool ok = sv && i < W;
                    if (AUX == AUX_GATHER || AUX == AUX_GATHER_G) amask[r] = ok ? expand_bit(a.ymask[i], cv) : 0u;
                    else if (AUX == AUX_STAGED) amask[r] = 0u;  // the gather index table replaces it
                    else if (AUX == AUX_GATHER_T) amask[r] = ok ? expand_bit(a.ymask[i], cv) | (1u << cv) : ~0u;
                    else amask[r] = ok ? (uint32_t)i : 0xffu;
                }
                const int aoff = b * a.Wa;
                // adjoint: this slot's dY values (zero past the row's end)
                T dyv[RC];
                constexpr bool edge_grad = AUX == AUX_GATHER_G;
                if (edge_grad)
#pragma unroll
                    for (int r = 0; r < RC; ++r)
                        dyv[r] = sv && i0 + r < W ? a.dy[(int64_t)v * a.ldd + b * W + i0 + r] : T(0);
                // a neighbour past the end reads as the vertex itself with a
                // color that adds nothing (col(v) is never in a GATHER set and
                // always in a GATHER_T set)
                const uint32_t dead = gathers_rows(AUX) ? cv : 0xffu;
                if constexpr (AUX == AUX_STAGED) {
                    static_assert(RC == 4, "the staged path reads a slot's 4 indices at once");
                    if (active) {
                    // Rows staged in LDS: the group copies, per neighbour, the
                    // segment of its row that holds this chunk's colorings
                    // (coalesced), then each lane reads its values there.
                    // Pipeline: ids two batches ahead, rows one batch ahead.
                    const int step = ng;
                    const int b_first = s0 / spb;
                    const int b_last = min(B - 1, (s0 + cl - 1) / spb);
                    const int seglen = (b_last - b_first + 1) * a.Wa;
                    const int nt = (seglen + cl - 1) / cl;  // <= kStageRegs (host)
                    const int spad = a.segpad;
                    T* stg = s_stage + grp * gstride;
                    const int lb = sv ? (b - b_first) * a.Wa : 0;
                    const T* tab = a.aux + b_first * a.Wa;
                    const int sbase = a.k * a.k * a.Wy4;  // gidx rows of all -1
                    int uA[4], uB[4], uC[4];
                    uint32_t cA[4], cB[4], cC[4];
                    T wA[4], wB[4], wC[4];
                    bool oA[4], oB[4], oC[4];
                    auto fetch4 = [&](int eb, int* uo, uint32_t* co, T* wo, bool* oo) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int ee = eb + q * step;
                            oo[q] = ee < e1;
                            const int ec = oo[q] ? ee : e1 - 1;  // a real entry of this item
                            uo[q] = a.col[ec];
                            co[q] = a.ecol[(int64_t)ec * B + b];
                            wo[q] = a.w ? a.w[ec] : T(1);
                        }
                    };
                    T rv[4][kStageRegs];
                    auto load_rows = [&](const int* uo) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const T* src = tab + (int64_t)uo[q] * a.lda;
#pragma unroll
                            for (int t = 0; t < kStageRegs; ++t)
                                if (t < nt) rv[q][t] = src[min(cln + t * cl, seglen - 1)];
                        }
                    };
                    int e = e0 + grp;
                    fetch4(e, uA, cA, wA, oA);
                    fetch4(e + 4 * step, uB, cB, wB, oB);
                    load_rows(uA);
                    for (; e < e1; e += 4 * step) {
                        fetch4(e + 8 * step, uC, cC, wC, oC);
#pragma unroll
                        for (int q = 0; q < 4; ++q)
#pragma unroll
                            for (int t = 0; t < kStageRegs; ++t)
                                if (t < nt) stg[q * spad + cln + t * cl] = rv[q][t];
                        load_rows(uB);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int gb = oA[q] ? gbase : sbase;
                            const int2 ids =
                                *reinterpret_cast<const int2*>(s_gidx + gb + (int)__umul24(cA[q], a.Wy4));
                            const int id[4] = {(int)(int16_t)(ids.x & 0xffff), ids.x >> 16,
                                               (int)(int16_t)(ids.y & 0xffff), ids.y >> 16};
                            const T* row = stg + q * spad + lb;
#pragma unroll
                            for (int r = 0; r < kRegCols; ++r) {
                                const T val = row[id[r] < 0 ? 0 : id[r]];
                                const T xv = id[r] < 0 ? T(0) : val;
                                acc[r] = a.w ? acc[r] + wA[q] * xv : acc[r] + xv;
                            }
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            uA[q] = uB[q]; cA[q] = cB[q]; wA[q] = wB[q]; oA[q] = oB[q];
                            uB[q] = uC[q]; cB[q] = cC[q]; wB[q] = wC[q]; oB[q] = oC[q];
                        }
                    }
                    }
                } else if (active) {
                    const int step = ng;
                    int e = e0 + grp;
                    // software pipeline: the ids, colors and weights of the
                    // next 4 neighbours load while the current rows are read.
                    // Loads are unconditional: an entry past the end reads the
                    // item's last entry (which exists) and is flagged.  With
                    // gidx the flag only moves the table base to the all -1
                    // rows, so no loaded value feeds a select (which would let
                    // the compiler sink the loads under branches).
                    int un[4];
                    uint32_t cn[4];
                    T wn[4];
                    bool okn[4];
                    auto fetch = [&](int ee, int q) {
                        okn[q] = ee < e1;
                        const int ec = okn[q] ? ee : e1 - 1;
                        un[q] = a.col[ec];
                        cn[q] = a.ecol[(int64_t)ec * B + b];
                        wn[q] = a.w ? a.w[ec] : T(1);
                    };
#pragma unroll
                    for (int q = 0; q < 4; ++q) fetch(e + q * step, q);
                    const T* rowbase = a.aux + aoff;
                    const int sbase = a.k * a.k * a.Wy4;  // gidx rows of all -1
                    for (; e < e1; e += 4 * step) {
                        int u[4];
                        uint32_t cu[4];
                        T wv[4];
                        bool ok[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            u[q] = un[q];
                            cu[q] = cn[q];
                            wv[q] = wn[q];
                            ok[q] = okn[q];
                            fetch(e + (4 + q) * step, q);
                        }
                        T x[4][RC];
                        if (use_gidx) {
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const int gb = ok[q] ? gbase : sbase;
                                const int16_t* gp = s_gidx + gb + (int)__umul24(cu[q], a.Wy4);
                                int id[RC];
                                if constexpr (RC == 4) {  // one 8-byte LDS read holds the slot's indices
                                    const int2 ids = *reinterpret_cast<const int2*>(gp);
                                    id[0] = (int)(int16_t)(ids.x & 0xffff);
                                    id[1] = ids.x >> 16;
                                    id[2] = (int)(int16_t)(ids.y & 0xffff);
                                    id[3] = ids.y >> 16;
                                } else {
#pragma unroll
                                    for (int r = 0; r < RC; ++r) id[r] = gp[r];
                                }
                                const T* row = rowbase + (int64_t)u[q] * a.lda;
                                // every lane reads (index 0 for "no value") and selects
#pragma unroll
                                for (int r = 0; r < RC; ++r) {
                                    const T val = row[(unsigned)(id[r] < 0 ? 0 : id[r])];
                                    x[q][r] = id[r] < 0 ? T(0) : val;
                                }
                            }
                        } else {
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const int uq = ok[q] ? u[q] : v;
                                const uint32_t cq = ok[q] ? cu[q] : dead;
#pragma unroll
                                for (int r = 0; r < RC; ++r)
                                    x[q][r] = aux_term<T, AUX>(a, s_rank, amask[r], aoff, uq, cq);
                            }
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q)
#pragma unroll
                            for (int r = 0; r < RC; ++r)
                                acc[r] = a.w ? acc[r] + wv[q] * x[q][r] : acc[r] + x[q][r];
                        if (edge_grad) {
                            // d Y[v] / d w_e = x_e: contract with dY over the group's
                            // values in the table precision, one FP64 add per entry.
                            // The 4 neighbours' dots reduce together (fixed order):
                            // xor 1 and xor 2 halve the values a lane carries, so lane
                            // bits (0,1) end up naming the neighbour 2*b0 + b1, then
                            // the higher xor steps sum across the group.  The FP64 add
                            // to the entry's gradient is a hardware atomic that returns
                            // nothing, so no wait for it holds up the rows already in
                            // flight; an entry has one owning lane, whose adds (one per
                            // column chunk) land in program order, as a plain
                            // read-modify-write would.
                            T dot[4];
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                dot[q] = T(0);
#pragma unroll
                                for (int r = 0; r < RC; ++r) dot[q] = dot[q] + dyv[r] * x[q][r];
                            }
                            if (cl >= 4) {
                                const bool h1 = cln & 1, h2 = (cln >> 1) & 1;
                                const T pa = (h1 ? dot[2] : dot[0]) + __shfl_xor(h1 ? dot[0] : dot[2], 1, TEAM);
                                const T pb = (h1 ? dot[3] : dot[1]) + __shfl_xor(h1 ? dot[1] : dot[3], 1, TEAM);
                                T sum = (h2 ? pb : pa) + __shfl_xor(h2 ? pa : pb, 2, TEAM);
                                for (int sh = 4; sh < cl; sh <<= 1) sum = sum + __shfl_xor(sum, sh, TEAM);
                                const int q = 2 * (int)h1 + (int)h2;
                                if (cln < 4 && ok[q]) (void)unsafeAtomicAdd(&a.grad[e + q * step], (double)sum);
                            } else {
#pragma unroll
                                for (int q = 0; q < 4; ++q) {
                                    T sum = dot[q];
                                    for (int sh = 1; sh < cl; sh <<= 1) sum = sum + __shfl_xor(sum, sh, TEAM);
                                    if (cln == 0 && ok[q]) (void)unsafeAtomicAdd(&a.grad[e + q * step], (double)sum);
                                }
                            }
                        }
                    }
                }
                // fold the neighbour groups: fixed xor tree, identical in every lane
                for (int st = cl; st < TEAM; st <<= 1)
#pragma unroll
                    for (int r = 0; r < RC; ++r) acc[r] = acc[r] + __shfl_xor(acc[r], st, TEAM);
                if (sv && grp == 0) {
#pragma unroll
                    for (int r = 0; r < RC; ++r) {
                        const int i = i0 + r;
                        if (i >= W) continue;
                        const int cc = b * W + i;
                        if (mode == OUT_Y) a.out[(int64_t)v * a.ldo + cc] = acc[r];
                        else if (mode == OUT_PARTIAL) a.out[(int64_t)slot * a.ldo + cc] = acc[r];
                        else sY[cc] = acc[r];
                    }
                }
            }
        } else if (AUX == AUX_READ) {
            // aux child is a leaf: Y value i is H at the i-th color other than col(v)
            if (active)
                for (int c = lane; c < nY; c += TEAM) {
                    const int b = c / a.Wy;
                    const int i = c - b * a.Wy;
                    const int cv = a.colors[(int64_t)v * B + b];
                    sY[c] = a.aux[(int64_t)v * a.lda + b * a.k + i + (i >= cv ? 1 : 0)];
                }
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
        if (mode == OUT_Y || mode == OUT_PARTIAL) continue;  // uniform over the grid

        // -------------------------------------------------- phase 2: main
        if (MAIN == MAIN_TABLE && active)
            for (int c = lane; c < nM; c += TEAM) sM[c] = a.main[(int64_t)v * a.ldm + c];
        wave_sync();
        if constexpr (AUX != AUX_STAGED) {  // (the host never stages a step whose Y is kept)
            if (a.ykeep && active)  // Y kept for the adjoint
                for (int c = lane; c < nY; c += TEAM) a.ykeep[(int64_t)v * a.ldy + c] = sY[c];
        }

        // ----------------------------------------------- phase 3: combine
        const int Wy = a.Wy, Wm = a.Wm, Wo = a.Wo, P = a.P;
        if (mode == OUT_ROOT) {
            if (active) {
                for (int b = lane; b < B; b += TEAM) {
                    double val = 0.0;
                    const T* y = sY + b * Wy;
                    if (MAIN == MAIN_LEAF) {
                        val = (double)y[0];  // f32 product y * 1, widened (DataTableColMajor.cpp:396)
                    } else {
                        const T* m = sM + b * Wm;
                        for (int j = 0; j < P; ++j) {
                            const uint32_t p = a.split[j];
                            const T prod = y[p & 0xffffu] * m[p >> 16];
                            val = val + (double)prod;
                        }
                    }
                    a.root_vals[(int64_t)b * a.n_root + v] = val;
                }
            }
        } else if (active) {
            const int nO = B * Wo;
            for (int o = lane; o < nO; o += TEAM) {
                const int b = o / Wo;
                const int r = o - b * Wo;
                const T* y = sY + b * Wy;
                T res;
                if (MAIN == MAIN_LEAF) {
                    // main = {col(v)}: the set's other colors are the Y set itself
                    if (mode == OUT_EMA_SCALED) res = (T)((double)y[r] * (double)a.scale);
                    else res = y[r];
                } else {
                    const T* m = sM + b * Wm;
                    // split j of output r: s_split[j * Wo + r] (LDS) or a.split[r * P + j]
                    if (use_split_lds)
                        res = combine_splits<T>([&](int j) { return s_split[j * Wo + r]; }, P, y, m, mode, a.scale);
                    else
                        res = combine_splits<T>([&](int j) { return a.split[(size_t)r * P + j]; }, P, y, m, mode,
                                                a.scale);
                }
                a.out[(int64_t)v * a.ldo + o] = res;
            }
        }
        wave_sync();  // LDS rows are rewritten by the next item
    }
}

// Persistent grid: as many workgroups as can be resident at once (occupancy
// for this instantiation and LDS size, or the caller's blocks_per_cu, times
// the CU count), never more than there are items.
template <typename T, int TEAM, int AUX, int MAIN, int RC>
static hipError_t launch_one(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st)
{
    const size_t lds = sub_lds_bytes<T>(TEAM, AUX, MAIN, a);
    const int64_t need = (a.n_items + (kBlock / TEAM) - 1) / (kBlock / TEAM);
    if (need <= 0) return hipSuccess;
    static thread_local std::map<size_t, int> per_cu_cache;
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        auto it = per_cu_cache.find(lds);
        if (it != per_cu_cache.end()) {
            per_cu = it->second;
        } else {
            hipError_t e =
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sub<T, TEAM, AUX, MAIN, RC>, kBlock, lds);
            if (e != hipSuccess) return e;
            if (per_cu < 1) return hipErrorInvalidConfiguration;
            per_cu_cache[lds] = per_cu;
        }
    }
    const int64_t cap = (int64_t)per_cu * num_cus;
    const int grid = (int)(need < cap ? need : cap);
    hipLaunchKernelGGL((k_sub<T, TEAM, AUX, MAIN, RC>), dim3(grid), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

// gathers come with one or kRegCols values per lane, every other mode with kRegCols
template <typename T, int TEAM, int AUX, int MAIN>
static hipError_t launch_rc(int rc, const SubArgs<T>& a, int num_cus, int per_cu, hipStream_t st)
{
    if constexpr (AUX == AUX_GATHER || AUX == AUX_GATHER_G || AUX == AUX_GATHER_T) {
        if (rc == 1) return launch_one<T, TEAM, AUX, MAIN, 1>(a, num_cus, per_cu, st);
    }
    if (rc != kRegCols) return hipErrorInvalidValue;
    return launch_one<T, TEAM, AUX, MAIN, kRegCols>(a, num_cus, per_cu, st);
}

template <typename T, int TEAM>
static hipError_t launch_team(int aux, int main_mode, int rc, const SubArgs<T>& a, int num_cus, int per_cu,
                              hipStream_t st)
{
    if (main_mode == MAIN_LEAF) {
        switch (aux) {
        case AUX_GATHER: return launch_rc<T, TEAM, AUX_GATHER, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_ROWS: return launch_rc<T, TEAM, AUX_ROWS, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_STAGED: return launch_rc<T, TEAM, AUX_STAGED, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_GATHER_G: return launch_rc<T, TEAM, AUX_GATHER_G, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_GATHER_T: return launch_rc<T, TEAM, AUX_GATHER_T, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_COLORS: return launch_rc<T, TEAM, AUX_COLORS, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_READ: return launch_rc<T, TEAM, AUX_READ, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        case AUX_PARTIALS: return launch_rc<T, TEAM, AUX_PARTIALS, MAIN_LEAF>(rc, a, num_cus, per_cu, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (aux) {
    case AUX_GATHER: return launch_rc<T, TEAM, AUX_GATHER, MAIN_TABLE>(rc, a, num_cus, per_cu, st);
    case AUX_ROWS: return launch_rc<T, TEAM, AUX_ROWS, MAIN_TABLE>(rc, a, num_cus, per_cu, st);
    case AUX_STAGED: return launch_rc<T, TEAM, AUX_STAGED, MAIN_TABLE>(rc, a, num_cus, per_cu, st);
    case AUX_READ: return launch_rc<T, TEAM, AUX_READ, MAIN_TABLE>(rc, a, num_cus, per_cu, st);
    case AUX_PARTIALS: return launch_rc<T, TEAM, AUX_PARTIALS, MAIN_TABLE>(rc, a, num_cus, per_cu, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_sub(int team, int aux, int main_mode, int rc, const SubArgs<T>& a, int num_cus, int blocks_per_cu,
                      hipStream_t st)
{
    if (a.n_items <= 0) return hipSuccess;
    if (a.ng < 1 || team % a.ng != 0 || (a.ng & (a.ng - 1)) != 0) return hipErrorInvalidValue;
    if (uses_rank(aux) && (a.K < 1 || a.K > kMaxK - 1)) return hipErrorInvalidValue;
    switch (team) {
    case 8: return launch_team<T, 8>(aux, main_mode, rc, a, num_cus, blocks_per_cu, st);
    case 16: return launch_team<T, 16>(aux, main_mode, rc, a, num_cus, blocks_per_cu, st);
    case 32: return launch_team<T, 32>(aux, main_mode, rc, a, num_cus, blocks_per_cu, st);
    case 64: return launch_team<T, 64>(aux, main_mode, rc, a, num_cus, blocks_per_cu, st);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_sub<float>(int, int, int, int, const SubArgs<float>&, int, int, hipStream_t);
template hipError_t launch_sub<double>(int, int, int, int, const SubArgs<double>&, int, int, hipStream_t);

// ------------------------------------------------------------ utilities ---

// Flat launches cover `total` work items with a grid-stride loop over at most
// kFlatBlocks workgroups: a dispatch holds fewer than 2^32 work-items, and
// n * B or nnz * B exceed that at the largest shapes.
constexpr int64_t kFlatBlocks = 1 << 20;
inline dim3 flat_grid(int64_t total) { return dim3((unsigned)std::min<int64_t>((total + 255) / 256, kFlatBlocks)); }
#define FLAT_LOOP(i, total) \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (total); i += (int64_t)gridDim.x * blockDim.x)

__global__ void k_colors_rng(uint8_t* out, const int32_t* perm, int32_t n_act, int B, const PassParams* pp, int k)
{
    FLAT_LOOP(i, (int64_t)n_act * B)
    {
        const int64_t v = i / B;
        const int b = (int)(i - v * B);
        const uint64_t h = d_splitmix64(pp->seed ^ d_splitmix64((uint64_t)(pp->first + b) * 0x9E3779B97F4A7C15ull +
                                                                 (uint64_t)perm[v]));
        out[i] = (uint8_t)(((h >> 32) * (uint64_t)k) >> 32);
    }
}

hipError_t launch_colors_rng(uint8_t* out, const int32_t* perm, int32_t n_act, int B, const PassParams* pp, int k,
                             hipStream_t st)
{
    const int64_t total = (int64_t)n_act * B;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_colors_rng, flat_grid(total), dim3(256), 0, st, out, perm, n_act, B, pp, k);
    return hipGetLastError();
}

__global__ void k_colors_gather(uint8_t* out, const uint8_t* in, const int32_t* perm, int32_t n_act, int B, int64_t n,
                                const PassParams* pp)
{
    FLAT_LOOP(i, (int64_t)n_act * B)
    {
        const int64_t v = i / B;
        const int b = (int)(i - v * B);
        out[i] = in[(pp->first + b) * n + perm[v]];
    }
}

hipError_t launch_colors_gather(uint8_t* out, const uint8_t* in, const int32_t* perm, int32_t n_act, int B,
                                int64_t n, const PassParams* pp, hipStream_t st)
{
    const int64_t total = (int64_t)n_act * B;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_colors_gather, flat_grid(total), dim3(256), 0, st, out, in, perm, n_act, B, n, pp);
    return hipGetLastError();
}

__global__ void k_edge_colors(uint8_t* ecol, const int32_t* col, const uint8_t* colors, int64_t nnz, int B)
{
    FLAT_LOOP(i, nnz * B)
    {
        const int64_t e = i / B;
        const int b = (int)(i - e * B);
        ecol[i] = colors[(int64_t)col[e] * B + b];
    }
}

hipError_t launch_edge_colors(uint8_t* ecol, const int32_t* col, const uint8_t* colors, int64_t nnz, int B,
                              hipStream_t st)
{
    const int64_t total = nnz * B;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_edge_colors, flat_grid(total), dim3(256), 0, st, ecol, col, colors, nnz, B);
    return hipGetLastError();
}

// Two fixed-shape passes: block j of coloring b sums the contiguous slice j
// of vals[b] (thread t takes every 256th element of it, then a tree), and one
// block per coloring folds the kReduceBlocks slice sums with the same tree.
// The result depends neither on timing nor on how the values were produced.
__global__ __launch_bounds__(256) void k_reduce_slices(const double* vals, int64_t n, double* tmp)
{
    __shared__ double s[256];
    const int b = blockIdx.y;
    const int64_t per = (n + kReduceBlocks - 1) / kReduceBlocks;
    const int64_t lo = (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < n ? lo + per : n;
    const double* row = vals + (int64_t)b * n;
    double acc = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) acc += row[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) tmp[(int64_t)b * kReduceBlocks + blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* tmp, double* out)
{
    __shared__ double s[256];
    const int b = blockIdx.x;
    s[threadIdx.x] = tmp[(int64_t)b * kReduceBlocks + threadIdx.x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = s[0];
}

static_assert(kReduceBlocks == 256, "k_reduce_final folds exactly 256 slices");

hipError_t launch_root_reduce(const double* vals, int64_t n, int B, double* tmp, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_slices, dim3(kReduceBlocks, B), dim3(256), 0, st, vals, n, tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(B), dim3(256), 0, st, tmp, out);
    return hipGetLastError();
}

template <typename T>
__global__ void k_gather_weights(T* dst, const float* src, const int32_t* ent, int64_t nnz)
{
    FLAT_LOOP(i, nnz) dst[i] = (T)src[ent[i]];
}

hipError_t launch_gather_weights_f32(float* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st)
{
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_weights<float>, flat_grid(nnz), dim3(256), 0, st, dst, src,
                       ent, nnz);
    return hipGetLastError();
}

hipError_t launch_gather_weights_f64(double* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st)
{
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_weights<double>, flat_grid(nnz), dim3(256), 0, st, dst, src,
                       ent, nnz);
    return hipGetLastError();
}

// ------------------------------------------------------------- adjoint ---

// One term of a split sum: FP64 multiply and add, or (FP32 sums of FP32
// tables) one fused multiply-add.
__device__ __forceinline__ double node_term(double acc, double g, double r) { return acc + g * r; }
__device__ __forceinline__ float node_term(float acc, float g, float r) { return __builtin_fmaf(g, r, acc); }

// One thread per (vertex, coloring, value): the first n_act*B*Wy threads
// form dY, the next n_act*B*Wm form G_main.  Sums run over the transposed
// split lists in their fixed order, in Acc (FP64, or FP32 with fused
// multiply-adds), then round once to T.
template <typename T, typename Acc>
__device__ __forceinline__ void bwd_node_item(const NodeArgs<T>& a, int64_t i, int64_t ny)
{
    const bool is_dy = i < ny;
    const int64_t j = is_dy ? i : i - ny;
    const int W = is_dy ? a.Wy : a.Wm;
    const int64_t vb = j / W;            // v * B + b
    const int x = (int)(j - vb * W);     // value index
    const int64_t v = vb / a.B;
    const int b = (int)(vb - v * a.B);
    const T* G = a.G ? a.G + vb * a.Wo : nullptr;
    const Acc g0 = (Acc)a.root_cot[b];  // the root's single "table" value
    Acc acc = Acc(0);
    if (is_dy) {
        if (!a.Tm) {
            acc = G ? (Acc)G[x] : g0;  // main leaf: T_s[o] = Y[o]
        } else {
            const T* m = a.Tm + vb * a.Wm;
            const uint32_t* lst = a.by_aux + (size_t)x * a.Pa;
            for (int p = 0; p < a.Pa; ++p) {
                const uint32_t e = lst[p];
                acc = node_term(acc, G ? (Acc)G[e >> 16] : g0, (Acc)m[e & 0xffffu]);
            }
        }
        a.dY[vb * a.Wy + x] = (T)(acc * (Acc)a.scale);
    } else {
        const int cv = a.colors[vb];
        const uint32_t* lst = a.by_main + (size_t)x * a.Pm;
        for (int p = 0; p < a.Pm; ++p) {
            const uint32_t e = lst[p];
            const int ra = (int)(e & 0xffffu);
            Acc y;
            if (a.y_from_h) y = (Acc)a.Y[vb * a.k + ra + (ra >= cv ? 1 : 0)];
            else y = (Acc)a.Y[vb * a.Wy + ra];
            acc = node_term(acc, G ? (Acc)G[e >> 16] : g0, y);
        }
        a.Gm[vb * a.Wm + x] = (T)(acc * (Acc)a.scale);
    }
}

template <typename T, typename Acc>
__global__ void k_bwd_node(const NodeArgs<T> a)
{
    const int64_t ny = (int64_t)a.n_act * a.B * a.Wy;
    const int64_t nm = a.Tm ? (int64_t)a.n_act * a.B * a.Wm : 0;
    const int64_t lo = a.part == 2 ? ny : 0;        // part 2: the G_main items only
    const int64_t hi = a.part == 1 ? ny : ny + nm;  // part 1: the dY items only
    FLAT_LOOP(i, hi - lo) bwd_node_item<T, Acc>(a, lo + i, ny);
}

// The same values with the rows staged in LDS: one workgroup per group of
// `vpb` consecutive (vertex, coloring) rows copies their G row, main row (read
// by the dY values) and Y row (read by the G_main values) into LDS with
// coalesced loads, then every thread forms values from LDS.  The split lists
// stay in global memory, where every workgroup reads the same few KB through
// the cache.  Per value the same FP64 sum in the same order as bwd_node_item,
// so the results are identical.
// sum over a split list, in list order, of G[o] (or g0 at the root) times the
// row value the entry names: row[r] (skip < 0), or the histogram row read
// past the vertex's own color, row[r + (r >= skip)].  Entries are (o << 16 |
// r).  The loads of 8 entries are issued before their terms are added.
template <typename Acc, typename T, typename List>
__device__ __forceinline__ Acc node_sum(List entry_at, int n, const T* G, Acc g0, const T* row, int skip)
{
    constexpr int U = 8;
    Acc acc = Acc(0);
    int p = 0;
    for (; p + U <= n; p += U) {
        uint32_t e[U];
#pragma unroll
        for (int q = 0; q < U; ++q) e[q] = entry_at(p + q);
        Acc gv[U], rv[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int r = (int)(e[q] & 0xffffu);
            gv[q] = G ? (Acc)G[e[q] >> 16] : g0;
            rv[q] = (Acc)row[r + (skip >= 0 && r >= skip ? 1 : 0)];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) acc = node_term(acc, gv[q], rv[q]);
    }
    for (; p < n; ++p) {
        const uint32_t e = entry_at(p);
        const int r = (int)(e & 0xffffu);
        acc = node_term(acc, G ? (Acc)G[e >> 16] : g0, (Acc)row[r + (skip >= 0 && r >= skip ? 1 : 0)]);
    }
    return acc;
}

// dst[at(i)] = src[i] for i < nvalid and T(0) for nvalid <= i < n
// (1 <= nvalid <= n): each thread issues 8 global loads before it writes
// them to LDS (a copy loop of single load-wait-write steps waits out the
// memory latency once per element).  Indexes past nvalid load
// src[nvalid - 1], so the loads need no branches.
template <typename T, typename At>
__device__ __forceinline__ void stage_copy(T* dst, const T* src, int n, int nvalid, At at)
{
    constexpr int SU = 8;
    for (int i0 = threadIdx.x; i0 < n; i0 += SU * (int)blockDim.x) {
        T v[SU];
#pragma unroll
        for (int q = 0; q < SU; ++q) {
            const int i = i0 + q * (int)blockDim.x;
            v[q] = src[i < nvalid ? i : nvalid - 1];
        }
#pragma unroll
        for (int q = 0; q < SU; ++q) {
            const int i = i0 + q * (int)blockDim.x;
            if (i < n) dst[at(i)] = i < nvalid ? v[q] : T(0);
        }
    }
}

template <typename T, typename Acc>
__global__ __launch_bounds__(256) void k_bwd_node_lds(const NodeArgs<T> a, int vpb, int lists_lds, int t_dy)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const bool do_dy = a.part != 2;
    const bool do_gm = a.part != 1 && a.Tm != nullptr;
    const int wg = a.G ? a.Wo : 0;
    const int wm = (do_dy && a.Tm) ? a.Wm : 0;
    const int yw = a.y_from_h ? a.k : a.Wy;  // Y row length in global memory
    const int wy = do_gm ? yw : 0;
    // lists_lds: the split lists the values read (by_auxT for dY, by_mainT
    // for G_main) are copied to LDS once per workgroup, ahead of the rows
    const int la = lists_lds && do_dy && a.Tm ? a.Pa * a.Wy : 0;
    const int lm = lists_lds && do_gm ? a.Pm * a.Wm : 0;
    uint32_t* s_la = reinterpret_cast<uint32_t*>(smem_raw);
    uint32_t* s_lm = s_la + la;
    for (int i = threadIdx.x; i < la; i += blockDim.x) s_la[i] = a.by_auxT[i];
    for (int i = threadIdx.x; i < lm; i += blockDim.x) s_lm[i] = a.by_mainT[i];
    T* sG = reinterpret_cast<T*>(smem_raw + (((size_t)4 * (la + lm) + 15) & ~(size_t)15));  // [vpb][wg]
    T* sM = sG + (size_t)vpb * wg;           // [vpb][wm]
    T* sY = sM + (size_t)vpb * wm;           // [vpb][wy]
    const int ny = do_dy ? a.Wy : 0;
    const int per = ny + (do_gm ? a.Wm : 0);
    const int64_t nvb = (int64_t)a.n_act * a.B;
    const int64_t ngroups = (nvb + vpb - 1) / vpb;
    for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int64_t vb0 = g * vpb;
        const int nv = (int)(nvb - vb0 < vpb ? nvb - vb0 : vpb);
        __syncthreads();  // the previous group's reads are done
        const auto same = [](int i) { return i; };
        if (wg) stage_copy(sG, a.G + vb0 * wg, nv * wg, nv * wg, same);
        if (wm) stage_copy(sM, a.Tm + vb0 * wm, nv * wm, nv * wm, same);
        if (wy) stage_copy(sY, a.Y + vb0 * yw, nv * wy, nv * wy, same);
        __syncthreads();
        // items by kind: threads [0, t_dy) take the group's dY values and
        // threads [t_dy, blockDim) its G_main values (t_dy a whole number of
        // wavefronts), so the lanes of a wavefront sum lists of one length
        // (Pa or Pm terms; 14 against 210 in the (6;2,4) step of the
        // 12-vertex tree) and the host sizes each kind's share of threads by
        // its longest thread's terms
        const int n_dy = nv * ny;
        const bool take_dy = (int)threadIdx.x < t_dy;
        const int n_kind = take_dy ? n_dy : nv * (per - ny);
        const int step = take_dy ? t_dy : (int)blockDim.x - t_dy;
        for (int jt = take_dy ? (int)threadIdx.x : (int)threadIdx.x - t_dy; jt < n_kind; jt += step) {
            const int it = take_dy ? jt : n_dy + jt;
            int lv, x0;
            if (it < n_dy) {
                lv = it / ny;
                x0 = it - lv * ny;
            } else {
                const int j = it - n_dy;
                const int nmv = per - ny;
                lv = j / nmv;
                x0 = ny + (j - lv * nmv);
            }
            const int64_t vb = vb0 + lv;
            const int b = (int)(vb % a.B);
            const T* G = wg ? sG + lv * wg : nullptr;
            const Acc g0 = (Acc)a.root_cot[b];
            Acc acc = Acc(0);
            if (x0 < ny) {
                const int x = x0;
                if (!a.Tm) {
                    acc = G ? (Acc)G[x] : g0;
                } else {
                    const T* m = sM + lv * wm;
                    if (la)
                        acc = node_sum<Acc>([&](int p) { return s_la[p * a.Wy + x]; }, a.Pa, G, g0, m, -1);
                    else
                        acc = node_sum<Acc>([&](int p) { return a.by_auxT[(size_t)p * a.Wy + x]; }, a.Pa, G, g0, m,
                                            -1);
                }
                a.dY[vb * a.Wy + x] = (T)(acc * (Acc)a.scale);
            } else {
                const int x = x0 - ny;
                const int cv = a.colors[vb];
                const T* y = sY + lv * wy;
                const int skip = a.y_from_h ? cv : -1;
                if (lm)
                    acc = node_sum<Acc>([&](int p) { return s_lm[p * a.Wm + x]; }, a.Pm, G, g0, y, skip);
                else
                    acc = node_sum<Acc>([&](int p) { return a.by_mainT[(size_t)p * a.Wm + x]; }, a.Pm, G, g0, y,
                                        skip);
                a.Gm[vb * a.Wm + x] = (T)(acc * (Acc)a.scale);
            }
        }
    }
}

// The staged kernel with the loops the other way round: a thread takes one
// value index for every staged row of the group, so each split-list entry is
// read once per group instead of once per row, and the group's rows share it.
// The staged rows are interleaved, value-major ([value][R] with R rows per
// group): the R values one term reads from a table sit side by side, so a
// lane takes them with R * sizeof(T) / 16 16-byte LDS reads instead of R
// 4- or 8-byte ones, and lanes that name different values of one bank group
// collide once per 16-byte read instead of once per row.  Rows past the
// group's last are zero-filled and their sums not stored.  Per row the same
// sum, in the same order, as k_bwd_node_lds.  Steps below the root (a G
// table) whose Y is not the histogram.
constexpr int kNodeRows = 8;

// R consecutive values of T from 16-byte aligned LDS
template <typename T, int R>
__device__ __forceinline__ void lds_take(const T* p, T (&v)[R])
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    static_assert((R * sizeof(T)) % 16 == 0, "R values of T span whole 16-byte words");
    constexpr int N = (int)(R * sizeof(T) / 16);
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const f4 w = reinterpret_cast<const f4*>(p)[q];
        __builtin_memcpy(reinterpret_cast<char*>(v) + 16 * q, &w, 16);
    }
}

// rows [vb0, vb0 + nv) of a row-major global table of width w into LDS as
// [w][R], zeros past row nv
template <typename T, int R>
__device__ __forceinline__ void stage_interleaved(T* dst, const T* src, int64_t vb0, int nv, int w)
{
    stage_copy(dst, src + vb0 * w, R * w, nv * w, [w](int i) {
        const int r = i / w;
        return (i - r * w) * R + r;
    });
}

// acc[r] += G_r[o] * row_r[i] over a split list's entries (o << 16 | i) in
// list order, for the R interleaved rows; the LDS reads of U entries are
// issued before their terms are added
template <typename T, typename Acc, int R, typename List>
__device__ __forceinline__ void rows_sum(List entry_at, int n, const T* sG, const T* rows, Acc (&acc)[R])
{
    constexpr int U = R * sizeof(T) <= 32 ? 4 : 2;
    int p = 0;
    for (; p + U <= n; p += U) {
        uint32_t e[U];
#pragma unroll
        for (int q = 0; q < U; ++q) e[q] = entry_at(p + q);
        T gv[U][R], rv[U][R];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            lds_take<T, R>(sG + (int)(e[q] >> 16) * R, gv[q]);
            lds_take<T, R>(rows + (int)(e[q] & 0xffffu) * R, rv[q]);
        }
#pragma unroll
        for (int q = 0; q < U; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = node_term(acc[r], (Acc)gv[q][r], (Acc)rv[q][r]);
    }
    for (; p < n; ++p) {
        const uint32_t e = entry_at(p);
        T gv[R], rv[R];
        lds_take<T, R>(sG + (int)(e >> 16) * R, gv);
        lds_take<T, R>(rows + (int)(e & 0xffffu) * R, rv);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = node_term(acc[r], (Acc)gv[r], (Acc)rv[r]);
    }
}

template <typename T, typename Acc, int R>
__global__ __launch_bounds__(256) void k_bwd_node_rows(const NodeArgs<T> a, int lists_lds)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const bool do_dy = a.part != 2;
    const bool do_gm = a.part != 1 && a.Tm != nullptr;
    const int wg = a.Wo;
    const int wm = (do_dy && a.Tm) ? a.Wm : 0;
    const int wy = do_gm ? a.Wy : 0;
    const int la = lists_lds && do_dy && a.Tm ? a.Pa * a.Wy : 0;
    const int lm = lists_lds && do_gm ? a.Pm * a.Wm : 0;
    uint32_t* s_la = reinterpret_cast<uint32_t*>(smem_raw);
    uint32_t* s_lm = s_la + la;
    for (int i = threadIdx.x; i < la; i += blockDim.x) s_la[i] = a.by_auxT[i];
    for (int i = threadIdx.x; i < lm; i += blockDim.x) s_lm[i] = a.by_mainT[i];
    T* sG = reinterpret_cast<T*>(smem_raw + (((size_t)4 * (la + lm) + 15) & ~(size_t)15));  // [wg][R]
    T* sM = sG + (size_t)R * wg;            // [wm][R]
    T* sY = sM + (size_t)R * wm;            // [wy][R]
    const int ny = do_dy ? a.Wy : 0;
    const int per = ny + (do_gm ? a.Wm : 0);
    const int64_t nvb = (int64_t)a.n_act * a.B;
    const int64_t ngroups = (nvb + R - 1) / R;
    for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int64_t vb0 = g * R;
        const int nv = (int)(nvb - vb0 < R ? nvb - vb0 : R);
        __syncthreads();
        stage_interleaved<T, R>(sG, a.G, vb0, nv, wg);
        if (wm) stage_interleaved<T, R>(sM, a.Tm, vb0, nv, wm);
        if (wy) stage_interleaved<T, R>(sY, a.Y, vb0, nv, wy);
        __syncthreads();
        for (int x0 = threadIdx.x; x0 < per; x0 += blockDim.x) {
            const bool is_dy = x0 < ny;
            const int x = is_dy ? x0 : x0 - ny;
            Acc acc[R];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = Acc(0);
            if (is_dy && !a.Tm) {
                T gv[R];
                lds_take<T, R>(sG + x * R, gv);
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = (Acc)gv[r];
            } else if (is_dy) {
                if (la) rows_sum<T, Acc, R>([&](int p) { return s_la[p * a.Wy + x]; }, a.Pa, sG, sM, acc);
                else rows_sum<T, Acc, R>([&](int p) { return a.by_auxT[(size_t)p * a.Wy + x]; }, a.Pa, sG, sM, acc);
            } else {
                if (lm) rows_sum<T, Acc, R>([&](int p) { return s_lm[p * a.Wm + x]; }, a.Pm, sG, sY, acc);
                else rows_sum<T, Acc, R>([&](int p) { return a.by_mainT[(size_t)p * a.Wm + x]; }, a.Pm, sG, sY, acc);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r >= nv) break;
                const int64_t vb = vb0 + r;
                if (is_dy) a.dY[vb * a.Wy + x] = (T)(acc[r] * (Acc)a.scale);
                else a.Gm[vb * a.Wm + x] = (T)(acc[r] * (Acc)a.scale);
            }
        }
    }
}

// LDS per workgroup of the staged kernel, the largest row set it takes, and
// the split reads per staged value from which it pays.  Measured in FP32,
// forward+backward with every main-table launch staged: the 12-vertex tree
// at R-MAT scale 22 (7.4-21 reads per staged value) ran 8% faster, the
// 10-vertex tree (2.6-5.4) 4% slower, the 5-star at scale 18 (0.9-1.6) 16%
// slower.  The split lists go to LDS too when together they fit
// kBwdListsLds (rows then get kBwdLdsBudget - kBwdListsTrim).
constexpr size_t kBwdLdsBudget = 32 * 1024;
constexpr size_t kBwdLdsMax = 48 * 1024;
constexpr int kBwdStageReuse = 6;
constexpr size_t kBwdListsLds = 40 * 1024;
constexpr size_t kBwdListsTrim = 8 * 1024;

template <typename T, typename Acc>
hipError_t launch_bwd_node_t(const NodeArgs<T>& a, int num_cus, hipStream_t st)
{
    const int64_t ny = (int64_t)a.n_act * a.B * a.Wy;
    const int64_t nm = a.Tm ? (int64_t)a.n_act * a.B * a.Wm : 0;
    const int64_t total = a.part == 1 ? ny : a.part == 2 ? nm : ny + nm;
    if (total == 0) return hipSuccess;
    const bool do_dy = a.part != 2, do_gm = a.part != 1 && a.Tm != nullptr;
    const size_t row = sizeof(T) * ((a.G ? a.Wo : 0) + ((do_dy && a.Tm) ? a.Wm : 0) +
                                    (do_gm ? (a.y_from_h ? a.k : a.Wy) : 0));
    // split reads per vertex-coloring: Wo*P each for dY and for G_main (only
    // main tables have split lists)
    const int64_t reads = a.Tm ? (int64_t)a.Wo * a.P * ((do_dy ? 1 : 0) + (do_gm ? 1 : 0)) : 0;
    if (row > kBwdLdsMax || reads < (int64_t)kBwdStageReuse * (int64_t)(row / sizeof(T))) {
        hipLaunchKernelGGL((k_bwd_node<T, Acc>), flat_grid(total), dim3(256), 0, st, a);
        return hipGetLastError();
    }
    const size_t lists = (size_t)4 * (((do_dy && a.Tm) ? (size_t)a.Pa * a.Wy : 0) + (do_gm ? (size_t)a.Pm * a.Wm : 0));
    const size_t lists_head = (lists + 15) & ~(size_t)15;
    auto vpb_for = [&](size_t budget) {
        return (int)std::max<size_t>(1, std::min<size_t>(64, budget / std::max<size_t>(row, 1)));
    };
    // workgroups stay within 64 KB of LDS
    const int lists_lds = lists > 0 && lists <= kBwdListsLds &&
                          lists_head + (size_t)vpb_for(kBwdLdsBudget - kBwdListsTrim) * row <= 64 * 1024;
    const size_t head = lists_lds ? lists_head : 0;
    int vpb = vpb_for(lists_lds ? kBwdLdsBudget - kBwdListsTrim : kBwdLdsBudget);
    // rows inner from 32 values per row.  Measured on the 12-vertex tree at
    // R-MAT scale 22, forward+backward of 16 colorings: the (10;7,3) step
    // (627 values) 118 ms rows inner against 316 per (row, value); the
    // (6;2,4) step's launches (176 values of 14-210 terms, and narrower)
    // 205 ms with none rows inner, 133 ms from 128 values, 103 ms from 32
    const int per_row = (do_dy ? a.Wy : 0) + (do_gm ? a.Wm : 0);
    // groups of kNodeRows rows, or 4 where the LDS budget holds fewer
    const int rows_r = vpb >= kNodeRows ? kNodeRows : vpb >= 4 ? 4 : 0;
    const bool rows_inner = a.rows_inner && a.G && !a.y_from_h && per_row >= 32 && rows_r > 0;
    if (rows_inner) vpb = rows_r;
    const int64_t ngroups = ((int64_t)a.n_act * a.B + vpb - 1) / vpb;
    const unsigned grid = (unsigned)std::min<int64_t>(ngroups, (int64_t)std::max(num_cus, 1) * 8);
    if (rows_inner) {
        if (rows_r == kNodeRows)
            hipLaunchKernelGGL((k_bwd_node_rows<T, Acc, kNodeRows>), dim3(grid), dim3(256), head + (size_t)vpb * row,
                               st, a, lists_lds);
        else
            hipLaunchKernelGGL((k_bwd_node_rows<T, Acc, 4>), dim3(grid), dim3(256), head + (size_t)vpb * row, st, a,
                               lists_lds);
        return hipGetLastError();
    }
    // threads on dY values: the whole-wavefront share whose longer kind's
    // thread sums the fewest terms (all threads when one kind is absent)
    const int64_t n_dy = do_dy ? (int64_t)vpb * a.Wy : 0, n_gm = do_gm ? (int64_t)vpb * a.Wm : 0;
    const int64_t len_dy = a.Tm ? a.Pa : 1, len_gm = a.Pm;
    int t_dy = n_gm == 0 ? 256 : n_dy == 0 ? 0 : 128;
    if (n_dy > 0 && n_gm > 0) {
        int64_t best = -1;
        for (int t = 64; t < 256; t += 64) {
            const int64_t worst = std::max((n_dy + t - 1) / t * len_dy, (n_gm + (256 - t) - 1) / (256 - t) * len_gm);
            if (best < 0 || worst < best) best = worst, t_dy = t;
        }
    }
    hipLaunchKernelGGL((k_bwd_node_lds<T, Acc>), dim3(grid), dim3(256), head + (size_t)vpb * row, st, a, vpb,
                       lists_lds, t_dy);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_bwd_node(const NodeArgs<T>& a, int num_cus, hipStream_t st)
{
    if constexpr (sizeof(T) == 4) {
        if (a.sums32) return launch_bwd_node_t<T, float>(a, num_cus, st);
    }
    return launch_bwd_node_t<T, double>(a, num_cus, st);
}

template hipError_t launch_bwd_node<float>(const NodeArgs<float>&, int, hipStream_t);
template hipError_t launch_bwd_node<double>(const NodeArgs<double>&, int, hipStream_t);

// One thread per entry e = (row v, column u), all colorings: the derivative of
// Y[v] = sum_e w_e * T_a-term(u) with respect to w_e, contracted with dY[v].
template <typename T>
__device__ __forceinline__ void edge_grad_item(const EdgeGradArgs<T>& a, int64_t e)
{
    const int64_t v = a.erow[e];
    const int64_t u = a.col[e];
    double acc = 0.0;
    for (int b = 0; b < a.B; ++b) {
        const uint32_t cv = a.colors[v * a.B + b];
        const uint32_t cu = a.ecol[e * a.B + b];
        if (cu == cv) continue;
        const T* dy = a.dY + (v * a.B + b) * a.Wy;
        if (!a.Ta) {
            acc += (double)dy[cu - (cu > cv ? 1u : 0u)];  // aux leaf: Y value of color cu
            continue;
        }
        const T* ta = a.Ta + (u * a.B + b) * a.Wa;
        for (int i = 0; i < a.Wy; ++i) {
            const uint32_t A = expand_bit(a.ymask[i], cv);
            if (!((A >> cu) & 1u)) continue;
            const int idx = a.rank_tab[squeeze_bit(A & ~(1u << cu), cu)];
            acc += (double)dy[i] * (double)ta[idx];
        }
    }
    a.grad[e] += acc;
}

template <typename T>
__global__ void k_edge_grad(const EdgeGradArgs<T> a)
{
    FLAT_LOOP(e, a.nnz) edge_grad_item(a, e);
}

template <typename T>
hipError_t launch_edge_grad(const EdgeGradArgs<T>& a, hipStream_t st)
{
    if (a.nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_edge_grad<T>, flat_grid(a.nnz), dim3(256), 0, st, a);
    return hipGetLastError();
}

template hipError_t launch_edge_grad<float>(const EdgeGradArgs<float>&, hipStream_t);
template hipError_t launch_edge_grad<double>(const EdgeGradArgs<double>&, hipStream_t);

template <typename TO>
__global__ void k_scatter_grad(TO* out, const double* grad, const int32_t* ent, int64_t nnz, double factor)
{
    FLAT_LOOP(i, nnz) out[ent[i]] = (TO)(factor * grad[i]);
}

hipError_t launch_scatter_grad_f32(float* out, const double* grad, const int32_t* ent, int64_t nnz, double factor,
                                   hipStream_t st)
{
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_grad<float>, flat_grid(nnz), dim3(256), 0, st, out, grad, ent,
                       nnz, factor);
    return hipGetLastError();
}

hipError_t launch_scatter_grad_f64(double* out, const double* grad, const int32_t* ent, int64_t nnz, double factor,
                                   hipStream_t st)
{
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_grad<double>, flat_grid(nnz), dim3(256), 0, st, out, grad, ent,
                       nnz, factor);
    return hipGetLastError();
}

__global__ void k_zero_f64(double* p, int64_t n)
{
    FLAT_LOOP(i, n) p[i] = 0.0;
}

hipError_t launch_zero_f64(double* p, int64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_f64, flat_grid(n), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace pgbsc
