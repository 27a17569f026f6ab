// Device kernels of the color-coding DP (gfx950).  Declarations shared by the
// C-ABI layer (pg_capi.hip) and the kernel definitions (pg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pgbsc {

// Where a work item's aux rows Y[v][.] come from.
enum AuxMode : int {
    AUX_GATHER = 0,    // Y = sum over neighbours u of A[v][u] * T_aux[u][.]
    AUX_COLORS = 1,    // the same with T_aux = one-hot leaf table built from colors
    AUX_READ = 2,      // Y = H[v][.] read directly (aux child is a leaf)
    AUX_PARTIALS = 3,  // Y = ordered sum of a hub's partial rows
};
// Where the main-child values come from.
enum MainMode : int {
    MAIN_TABLE = 0,  // T_main[v][.]
    MAIN_LEAF = 1,   // one-hot from colors: only the split whose main set is {col(v)} is non-zero
};
// What the work item produces.
enum OutMode : int {
    OUT_Y = 0,           // store Y itself (neighbour-color histogram H)
    OUT_EMA = 1,         // T_s[v][S] = sum_j main * aux (reference order of j)
    OUT_EMA_SCALED = 2,  // first processed subtemplate: FP64 terms times float(1e-12)
    OUT_ROOT = 3,        // root: FP64 sum over splits into a per-vertex value
    OUT_PARTIAL = 4,     // hub chunk: store partial Y row to a slot
};

constexpr int kBlock = 256;  // threads per workgroup (4 wavefronts of 64)
constexpr int kRegCols = 4;  // Y columns a lane accumulates per neighbour sweep

// A work item is (vertex, first entry, end entry, partial slot).  Items may
// be explicit (hub chunks) or implicit: item i is vertex item_base + i with
// its whole neighbour list.
template <typename T>
struct SubArgs {
    int32_t n_items;
    int32_t item_base;
    const int4* items;
    const int32_t* rowptr;   // [n_act+1] internal CSR
    const int32_t* col;      // [nnz] neighbour (internal id)
    const T* w;              // [nnz] entry weights, nullptr = 1
    const uint8_t* colors;   // [n_act][B]
    int32_t B;               // colorings in this pass
    int32_t k;               // colors
    const T* aux;            // aux table / H / partial rows
    int64_t lda;
    int32_t Ca;              // aux columns per coloring
    const int32_t* slot_begin;  // hub h's partial slots [slot_begin[h], slot_begin[h+1])
    const T* main;
    int64_t ldm;
    int32_t Cm;
    T* out;
    int64_t ldo;
    int32_t Cs;              // output columns per coloring
    int32_t P;               // splits per output set
    const uint32_t* split;   // [Cs][P] (main_rank << 16 | aux_rank), row = output colex rank
    const int16_t* drop;     // [Cs][k] aux rank of S \ {c}, -1 if c not in S
    int32_t out_mode;
    double* root_vals;       // [B][n_root] per-vertex root values
    int64_t n_root;
    float scale;
};

// Launches k_sub<T, team, aux, main>; returns hipErrorInvalidValue for an
// unsupported combination.
template <typename T>
hipError_t launch_sub(int team, int aux, int main_mode, const SubArgs<T>& a, int nblocks, size_t lds,
                      hipStream_t st);

// colors[v][b] for the pass: from the counter RNG keyed by the caller id
hipError_t launch_colors_rng(uint8_t* out, const int32_t* perm, int32_t n_act, int B, uint64_t seed,
                             int64_t first, int k, hipStream_t st);
// colors[v][b] gathered from caller colors in[c][caller id]
hipError_t launch_colors_gather(uint8_t* out, const uint8_t* in, const int32_t* perm, int32_t n_act, int B,
                                int64_t n, hipStream_t st);
// out[b] = FP64 sum of vals[b][0..n) in a fixed order
hipError_t launch_root_reduce(const double* vals, int64_t n, int B, double* out, hipStream_t st);
// dst[i] = (T) src[ent[i]] (weights into internal entry order)
hipError_t launch_gather_weights_f32(float* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st);
hipError_t launch_gather_weights_f64(double* dst, const float* src, const int32_t* ent, int64_t nnz, hipStream_t st);
// dst[i] = 0 (graph-capture-safe clear)
hipError_t launch_fill_zero(void* dst, size_t bytes, hipStream_t st);

}  // namespace pgbsc
