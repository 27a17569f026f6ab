// Device kernels of the color-coding DP (gfx950).  Declarations shared by the
// C-ABI layer (pg_capi.hip) and the kernel definitions (pg_kernels.hip).
//
// Table layout ("color-relative rows").  The reference's table for a
// subtemplate s has C(k,|s|) columns, one per color set S, and T_s[S][v] is
// zero unless S contains col(v) (the subtemplate's root sits on v).  The
// device therefore stores, per vertex v and coloring b, only the sets that
// contain col(v): entry o is the set {col(v)} + U, where U is the o-th
// (|s|-1)-subset, in colex order, of the K = k-1 colors other than col(v)
// renumbered 0..K-1.  Row width Wo = C(K, |s|-1) instead of C(k, |s|); the
// dropped entries are exact zeros, so every sum the reference forms is
// reproduced term for term.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pgbsc {

// Where a work item's aux values Y come from.
enum AuxMode : int {
    AUX_GATHER = 0,    // Y = sum over neighbours u of A[v][u] * T_aux[u][.] (color-relative rows)
    AUX_COLORS = 1,    // the same with the one-hot leaf table: neighbour-color histogram H
    AUX_READ = 2,      // Y read from H[v][.] (aux child is a leaf)
    AUX_PARTIALS = 3,  // Y = ordered sum of a hub's partial rows
    AUX_GATHER_T = 4,  // adjoint of AUX_GATHER: the vertex's row entry S (a set holding its color)
                       // sums the neighbours' dY at S, over neighbours whose color is not in S
    AUX_ROWS = 5,      // AUX_GATHER computed row-wise: each neighbour row is read whole (coalesced)
                       // and every value is added to the one Y set it belongs to
    AUX_STAGED = 6,    // AUX_GATHER through the gather index table with neighbour rows staged in
                       // LDS (needs SubArgs::gidx and segpad)
    AUX_GATHER_G = 7,  // AUX_GATHER of the adjoint: also grad[e] += dY . x_e per entry (needs dy, grad)
};
// Where the main-child values come from.
enum MainMode : int {
    MAIN_TABLE = 0,  // T_main[v][.] (color-relative row)
    MAIN_LEAF = 1,   // main child is a leaf: the only non-zero split is main = {col(v)}
};
// What the work item produces.
enum OutMode : int {
    OUT_Y = 0,           // store Y itself (histogram H, dense k columns per coloring)
    OUT_EMA = 1,         // T_s[v][o] = sum_j main * aux (reference order of j)
    OUT_EMA_SCALED = 2,  // first processed subtemplate: FP64 terms times float(1e-12)
    OUT_ROOT = 3,        // root: FP64 sum over splits into a per-vertex value
    OUT_PARTIAL = 4,     // hub chunk: store the partial Y row to a slot
};

constexpr int kBlock = 256;  // threads per workgroup (4 wavefronts of 64)
constexpr int kRegCols = 4;  // Y values a lane accumulates per neighbour sweep
constexpr int kMaxK = 14;      // colors supported on the device (rank table 2^(k-1) int16 in LDS)
constexpr int kStageRegs = 4;  // staged row values per lane per neighbour

// Orders a team's LDS accesses where the team lies within one wavefront and
// its LDS region is its own: a wavefront's LDS operations execute in issue
// order, so only the compiler must keep accesses on their side of this point.
// Teams then run independently instead of in step with the workgroup's
// slowest item.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A work item is (vertex, first entry, end entry, partial slot).  Items may
// be explicit (hub chunks) or implicit: item i is vertex item_base + i with
// its whole neighbour list.  A team of TEAM lanes owns an item; its lanes
// form `ng` neighbour groups of TEAM/ng column lanes.
template <typename T>
struct SubArgs {
    int32_t n_items;
    int32_t item_base;
    const int4* items;
    const int32_t* rowptr;   // [n_act+1] internal CSR
    const int32_t* col;      // [nnz] neighbour (internal id)
    const T* w;              // [nnz] entry weights, nullptr = 1
    const uint8_t* colors;   // [n_act][B]
    const uint8_t* ecol;     // [nnz][B] color of each entry's neighbour (colors[col[e]])
    int32_t B;               // colorings in this pass
    int32_t k;               // colors
    int32_t K;               // k - 1
    int32_t ng;              // neighbour groups per team (power of two)
    // aux side
    const T* aux;            // aux table / H / partial rows
    int32_t lda;             // aux row stride (B * Wa, or B * k for H)
    int32_t Wa;              // aux values per coloring in a row of `aux`
    int32_t Wy;              // Y values per coloring
    const uint16_t* ymask;   // [Wy] K-bit mask of each Y set (AUX_GATHER)
    const uint16_t* xmask;   // [Wa] K-bit mask of each aux-child row set (AUX_ROWS)
    const int16_t* rank_tab; // [2^K] colex rank of a K-bit mask among masks of its popcount
    // AUX_GATHER (and AUX_GATHER_T), optional: gidx[(cv*k + cu)*Wy4 + i] =
    // index in a neighbour's row (color cu) of the value Y value i (vertex
    // color cv) takes from it, or -1; rows k*k .. k*k+k-1 are all -1 (a slot
    // past the end reads there)
    const int16_t* gidx;
    int32_t Wy4;             // Wy rounded up to kRegCols
    // AUX_GATHER with gidx, optional (0 = off): neighbour rows are staged in
    // LDS, 4 per neighbour group, each a segment of segpad values (the
    // colorings of one chunk of slots); segpad <= kStageRegs * column lanes
    int32_t segpad;
    // AUX_ROWS, optional: xidx[(cv*k + cu)*Wa4 + j] = the Y value (vertex color
    // cv) that value j of a neighbour row (color cu) adds to, or -1
    const int16_t* xidx;
    int32_t Wa4;             // Wa rounded up to 4
    // RUNS_SEQ: the same targets for rows of up to 512 values, per lane:
    // xidx_g[((cv * k + cu) * 64 + lane) * RV + t] for value lane + 64 t,
    // RV = runs_values_per_lane(Wa) (-1 past the row's end)
    const int16_t* xidx_g;
    const int32_t* slot_begin;  // hub h's partial slots [slot_begin[h], slot_begin[h+1])
    // AUX_GATHER_G: each neighbour's contribution x to the vertex's Y row is
    // also contracted with dY there, and grad[e] += sum over the row of dY * x
    // (in FP64) for the entry e
    const T* dy;             // [n_act][ldd]
    int32_t ldd;             // dy row stride (B * Wy)
    double* grad;            // [nnz] internal entry order
    // RUNS_TG (the transposed gather fused with the gradient): the vertex's
    // own row of the aux child's table, and the transposed entry of each
    // entry, whose gradient the neighbour's dY row dotted with it gives
    const T* own;            // [n_act][ldown]
    int32_t ldown;           // B * (aux child's row width)
    const int32_t* tr;       // [nnz] entry (u, v) of entry (v, u)
    // ttab[((cv * k + c) * 64 + lane) * RV + t]: the entry of the vertex's
    // row (colour cv) that position lane + 64 t of a neighbour's dY row
    // (colour c) adds to, or -1; RV = runs_values_per_lane(dY width)
    const int16_t* ttab;
    // forward steps, optional: Y of every vertex written out as well
    // (ykeep[v * ldy + b * Wy + i]) for the adjoint
    T* ykeep;
    int32_t ldy;
    // main side
    const T* main;
    int32_t ldm;
    int32_t Wm;
    // output
    T* out;
    int32_t ldo;
    int32_t Wo;              // output values per coloring
    int32_t P;               // non-zero splits per output set
    const uint32_t* split;   // [Wo][P] (main_rank << 16 | y_rank), reference order
    int32_t split_lds;       // 1: the workgroup first copies split into LDS as [P][Wo] (combine reads)
    int32_t out_mode;
    double* root_vals;       // [B][n_root] per-vertex root values
    int64_t n_root;
    float scale;
};

// LDS stride between neighbour groups' staged rows (4 rows of segpad values,
// segpad a multiple of cl): for cl < 32 one more cl, an odd multiple of cl,
// so the groups' lanes of one wavefront spread over the 32 banks
__host__ __device__ constexpr int stage_group_stride(int segpad, int cl)
{
    return cl >= 32 ? 4 * segpad : 4 * segpad + cl;
}

template <typename T>
size_t sub_lds_bytes(int team, int aux, int main_mode, const SubArgs<T>& a);

// Launches k_sub<T, team, aux, main, rc> as a persistent grid of
// blocks_per_cu x num_cus workgroups (blocks_per_cu = 0: the occupancy
// query's resident count), never more than there are items; rc = values per
// lane (1 for gathers whose row fits one chunk, else kRegCols); returns
// hipErrorInvalidValue for an unsupported combination.
template <typename T>
hipError_t launch_sub(int team, int aux, int main_mode, int rc, const SubArgs<T>& a, int num_cus, int blocks_per_cu,
                      hipStream_t st);

// Neighbour sums over wide rows, read whole, one neighbour colour at a time
// (pg_runs.hip): a team of 64 lanes per (item, colouring).  RUNS_FWD sums
// the aux child's rows into Y (SubArgs::Wa values per row, their set masks in
// xmask; Y has Wy values) and finishes as k_sub does for out_mode; RUNS_T is
// the adjoint's transposed gather: it sums the neighbours' dY rows (Wa values,
// masks in xmask) into the vertex's G_aux row (Wy values), OUT_Y or
// OUT_PARTIAL.  Rows up to 512 values per colouring.
// RUNS_G is the adjoint's gradient gather: RUNS_FWD's sum written out
// (OUT_Y / OUT_PARTIAL) and grad[e] += dY[v] . x_u for every entry.
// RUNS_SEQ is the forward sum in the reference's order: each neighbour row
// read whole and coalesced, its values added into Y (LDS) neighbour after
// neighbour through xidx (every (vertex colour, neighbour colour) row of it),
// a team per (vertex, colouring) with the whole neighbour list.
// RUNS_TG is RUNS_T plus, for every entry (v, u) of the vertex's list, the
// gradient of its transpose: grad[tr[e]] += dY[u] . (own row of v mapped to
// u's Y positions), a team per item with every coloring in order.
enum RunsMode : int { RUNS_FWD = 0, RUNS_T = 1, RUNS_G = 2, RUNS_SEQ = 3, RUNS_TG = 4 };
// values per lane for a gathered row of wg values (0: too wide)
int runs_values_per_lane(int wg);
template <typename T>
size_t runs_lds_bytes(int main_mode, const SubArgs<T>& a);
template <typename T>
size_t runs_g_lds_bytes(const SubArgs<T>& a);
// LDS of the fused transposed gather (RUNS_TG) for rank tables of K bits and
// G_aux rows of Wt values
template <typename T>
size_t runs_tg_lds_bytes(int K, int Wt);
template <typename T>
hipError_t launch_runs(int mode, int main_mode, const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st);

// Hub rows of the forward sum in the reference's order (GATHER, OUT_PARTIAL
// into out[h][B*Wy] for hubs item_base + h, h < n_items): a wavefront per
// (hub, coloring, 64 Y values) walks the whole neighbour list with kHubBatch
// row loads in flight per lane.  LDS: the rank table.
constexpr int kHubBatch = 32;
constexpr int kHubBatchNarrow = 56;  // FP32 rows of up to 64 values
size_t hub_seq_lds_bytes(int K);
template <typename T>
hipError_t launch_hub_seq(const SubArgs<T>& a, int num_cus, int blocks_per_cu, hipStream_t st);

// Per-pass values read from device memory, so that one captured graph serves
// every pass of the same shape (the host writes them before each replay).
struct PassParams {
    uint64_t seed;
    int64_t first;  // index of the pass's first coloring
};

// colors[v][b] for the pass: from the counter RNG keyed by the caller id
hipError_t launch_colors_rng(uint8_t* out, const int32_t* perm, int32_t n_act, int B, const PassParams* pp, int k,
                             hipStream_t st);
// colors[v][b] gathered from caller colors in[first + b][caller id]
hipError_t launch_colors_gather(uint8_t* out, const uint8_t* in, const int32_t* perm, int32_t n_act, int B,
                                int64_t n, const PassParams* pp, hipStream_t st);
// ecol[e][b] = colors[col[e]][b]: the neighbour colors streamed beside the ids
hipError_t launch_edge_colors(uint8_t* ecol, const int32_t* col, const uint8_t* colors, int64_t nnz, int B,
                              hipStream_t st);
// out[b] = FP64 sum of vals[b][0..n) in a fixed order (tmp: [B][kReduceBlocks])
constexpr int kReduceBlocks = 256;
hipError_t launch_root_reduce(const double* vals, int64_t n, int B, double* tmp, double* out, hipStream_t st);
// dst[i] = (T) src[ent[i]] (weights into internal entry order)
hipError_t launch_gather_weights_f32(float* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st);
hipError_t launch_gather_weights_f64(double* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st);

// ---- adjoint DP -------------------------------------------------------
// Per vertex and coloring, from G = dR/dT_s (ones at the root), the main row
// and Y: dY[i] = sum over splits using aux set i of G[o] * T_m[rm], and
// G_m[r] = sum over splits using main set r of G[o] * Y[ra] (reference split
// list transposed by aux set / by main set).  main leaf: dY = G.
template <typename T>
struct NodeArgs {
    int32_t n_act, B, k;
    const uint8_t* colors;   // [n_act][B]
    const T* G;              // [n_act][B*Wo], nullptr at the root (G = root_cot[b])
    const double* root_cot;  // [B] cotangent of each coloring's raw count
    const T* Tm;             // [n_act][B*Wm] (main not a leaf)
    const T* Y;              // aux sums: Ybuf [n_act][B*Wy], or H [n_act][B*k] (aux leaf)
    int32_t y_from_h;        // 1: Y value i = H at the i-th color other than col(v)
    T* dY;                   // [n_act][B*Wy]
    T* Gm;                   // [n_act][B*Wm] (main not a leaf)
    int32_t Wo, Wy, Wm, Pa, Pm;
    int32_t P;                // splits per output value
    const uint32_t* by_aux;  // [Wy][Pa] (o << 16 | main rank)
    const uint32_t* by_main; // [Wm][Pm] (o << 16 | y rank)
    const uint32_t* by_auxT; // the same, split-major: [Pa][Wy]
    const uint32_t* by_mainT;// [Pm][Wm]
    double scale;            // 1, or float(1e-12) for the scaled subtemplate
    int32_t part;            // 0: dY and G_main, 1: dY only, 2: G_main only
    int32_t sums32;          // FP32 tables: 1 = split sums in FP32 (fused multiply-add), 0 = in FP64
    int32_t rows_inner;      // staged launches: a thread per value index over the group's rows
};
template <typename T>
hipError_t launch_bwd_node(const NodeArgs<T>& a, int num_cus, hipStream_t st);

// Per entry e = (row v, column u): grad[e] += sum_b sum_i dY[v][b][i] * T_a-term(u)
// (aux table) or dY[v][b][i(col u)] (aux leaf).  One thread per entry.
template <typename T>
struct EdgeGradArgs {
    int64_t nnz;
    int32_t B, k, K;
    const int32_t* erow;     // [nnz] row (internal vertex) of each entry
    const int32_t* col;
    const uint8_t* colors;   // [n_act][B]
    const uint8_t* ecol;     // [nnz][B]
    const T* dY;             // [n_act][B*Wy]
    int32_t Wy;
    const T* Ta;             // [n_act][B*Wa] (aux not a leaf), nullptr for a leaf aux
    int32_t Wa;
    const uint16_t* ymask;   // [Wy]
    const int16_t* rank_tab;
    double* grad;            // [nnz] internal order, accumulated
};
template <typename T>
hipError_t launch_edge_grad(const EdgeGradArgs<T>& a, hipStream_t st);

// out[ent[i]] (=/+=) (TO) (factor * grad[i])
hipError_t launch_scatter_grad_f32(float* out, const double* grad, const int32_t* ent, int64_t nnz, double factor,
                                   hipStream_t st);
hipError_t launch_scatter_grad_f64(double* out, const double* grad, const int32_t* ent, int64_t nnz, double factor,
                                   hipStream_t st);
hipError_t launch_zero_f64(double* p, int64_t n, hipStream_t st);

// ---- the reference's unpruned mode (isPruned = 0, CountMat.cpp:851-977) ----
// One non-leaf subtemplate for one coloring, thread per vertex, exactly the
// reference's FP32 operation sequence (pg_fascia.hip).  Tables are
// column-major [C(k,|x|)][n_act] over absolute color sets; leaf children
// read the colors.
struct FasciaArgs {
    int32_t n_act;
    const int32_t* rowptr;   // internal CSR, neighbours in ascending caller id
    const int32_t* col;
    const uint8_t* colors;   // [n_act]
    float* acc;              // [n_act] the reference's _bufVec, carried across calls
    int32_t Cs, P;
    const int32_t* c2c;      // [Cs] output column of the i-th set (IndexSys combToCount)
    const uint32_t* split;   // [Cs][P] (split_main << 16 | split_aux)
    int32_t main_leaf, aux_leaf;
    const float* tm;         // [Cm][n_act] (main not a leaf)
    const float* ta;         // [Ca][n_act] (aux not a leaf)
    float* out;              // [Cs][n_act]; at the root [n_act] (bufLastSub)
};
hipError_t launch_fascia_step(const FasciaArgs& a, int num_cus, hipStream_t st);
// *out = sum over caller vertices 0..n-1, in order, of (double) the root
// buffer (internal order, scattered through perm into scratch[n] first)
hipError_t launch_fascia_root_sum(const float* vals, const int32_t* perm, int32_t n_act, float* scratch, int64_t n,
                                  double* out, hipStream_t st);

}  // namespace pgbsc
