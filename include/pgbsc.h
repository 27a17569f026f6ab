/*
 * pgbsc.h -- C ABI of the MI355X color-coding subgraph-counting library
 * (libpgbsc.so).
 *
 * The reference (dsc-nec/subgraph) has no FFI: its hot path sits behind the
 * C++ classes CountMat, CSCGraph, DataTableColMajor, IndexSys and
 * DivideTemplates.  Each entry point below names the reference interface it
 * replaces.  Conventions:
 *   - every function returns PG_OK (0) or a negative PG_ERR_* code; the text of
 *     the last error is available from pg_last_error(ctx); nothing throws;
 *   - host arrays are owned by the caller and only read (or written, for
 *     outputs) during the call; device memory is owned by the context or plan;
 *   - one context per host thread; every call is stream-ordered on the
 *     context's stream and synchronous at return;
 *   - vertex ids and CSC entry indices are the caller's (the reference's)
 *     numbering; the library's internal relabelling never leaks out.
 */
#ifndef PGBSC_H
#define PGBSC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGBSC_ABI_VERSION 1

#define PG_OK 0
#define PG_ERR_INVALID (-1) /* bad argument or malformed input */
#define PG_ERR_NOMEM (-2)   /* device or host allocation failed */
#define PG_ERR_HIP (-3)     /* a HIP runtime call failed */
#define PG_ERR_STATE (-4)   /* call out of order (e.g. no graph uploaded) */
#define PG_ERR_IO (-5)      /* file could not be opened or is truncated */
#define PG_ERR_RANGE (-6)   /* size beyond what the library supports */

/* precision of the count tables */
#define PG_FP32 0 /* reference numerics: FP32 tables, float(1e-12) scaling, FP64 root */
#define PG_FP64 1 /* FP64 tables, no scaling: exact integer counts below 2^53 */

/* flags */
#define PG_FLAG_FIX_2VERTEX 1u /* do not multiply two-vertex templates by 1e12
                                  (the reference does, CountMat.cpp:811-824) */
#define PG_FLAG_UNPRUNED 2u    /* pg_count only, PG_FP32 only: the reference's unpruned
                                  color coding (isPruned = 0, countNonBottomeOriginal,
                                  CountMat.cpp:851-977) with its numerics: unscaled
                                  FP32 tables, one SpMV per (color set, split) pair
                                  added onto an accumulator that is zeroed only at
                                  pg_graph_upload and carried across pairs,
                                  subtemplates, colorings and calls (the reference
                                  clears _bufVec once, CountMat.cpp:52-59), a float
                                  root buffer summed in FP64 in vertex order, no 1e12.
                                  Colorings run one after another, slowly: it is the
                                  reference's baseline mode, kept for CLI parity */

typedef struct pg_ctx pg_ctx;
typedef struct pg_plan pg_plan;

/* ---- context (replaces CountMat::initialization, CountMat.cpp:21-77) ---- */
int pg_create(pg_ctx** out, int device);
void pg_destroy(pg_ctx* ctx);
const char* pg_last_error(const pg_ctx* ctx);
/* Tuning knobs (no reference counterpart; the reference hard-codes
 * _bufMatCols = 16, CountMat.cpp:62):
 *   "hub_degree"   vertices above this degree are split across teams (default 512)
 *   "hub_chunk"    neighbours per hub work item (default 1024)
 *   "bucket_high"  vertices of at least this degree get 64-lane teams (default 512)
 *   "bucket_mid"   ... at least this degree 16-lane teams, the rest 8 (default 64)
 *   "batch_tile"   colorings per pass, 0 = automatic (default 0)
 *   "mem_budget_mb" device bytes the tables may use, 0 = 90% of free memory
 *   "autotune"     1 = before the first pass of each (plan, batch tile, precision), time
 *                  one pass per candidate of reg_cols, stage_rows, aux_rows, bucket_high,
 *                  bucket_mid, hub_degree and hub_chunk and keep
 *                  the fastest (a second sweep if the first changed one), then the
 *                  column tile: 3/4, 1/2 and 1/4 of the colorings per pass, kept when
 *                  a coloring costs less (unless batch_tile fixes it);
 *                  pg_get_option then reports the choice, "last_tile" the tile.  A
 *                  gradient call applies what forward passes of its own tile tuned,
 *                  including a smaller tile (default 0)
 *   "timing"       1 = event pair around every launch, passes run eagerly (default 0)
 *   "graphs"       1 = each pass replays a captured hipGraph (default 1)
 *   "blocks_per_cu" workgroups per CU of persistent launches, 0 = occupancy query
 *   "stage_rows"   1 = gathers stage neighbour rows in LDS, 0 = direct loads (default 0)
 *   "split_lds"    1 = the eMA combine reads the split table from LDS when it fits (default),
 *                  0 = from memory
 *   "reg_cols"     values per lane of gathers: 0 = one for Y rows narrower than four values
 *                  (the root), else four (default); 1 or 4 = that for every gather
 *   "aux_rows"     neighbour sums row-wise with LDS atomics: 1 always, 0 never (default),
 *                  -1 when most of a neighbour row is used
 *   "exact_order"  1 = the forward neighbour sums in the reference's order, so FP32 counts
 *                  reproduce the reference's rounding (default); 0 = reordered sums (hub
 *                  chunks, neighbour groups, colour runs), faster, as accurate, but a
 *                  different rounding
 *   "hub_concurrent" 1 = a gathering step's hub launch runs on a second stream beside the
 *                  other vertices' launch (default); 0 = one after the other (same results)
 *   "color_runs"   whole-row neighbour sums: 1 always (rows up to 512 values), 0 never,
 *                  -1 by row width and used share (default)
 *   "fuse_grad"    pg_count_grad with kept tables: 1 = the forward keeps Y of the steps whose
 *                  transposed gather runs by colour runs, and that gather also takes the
 *                  entries' gradients (default); 0 = a gradient gather recomputes Y
 *   "fuse_min_width" with fuse_grad: also fuse steps whose dY rows have at least this many
 *                  values, by colour runs whatever share of a row a vertex uses (default 128)
 *   "node_fp32"    pg_count_grad with FP32 tables: 1 = the adjoint's split sums in FP32
 *                  (fused multiply-adds, default), 0 = in FP64
 *   "node_rows"    pg_count_grad: 1 = staged node launches over rows of 32+ values take a
 *                  thread per value index over a group's rows (default), 0 = per (row, value)
 *   "runs_g_min"   pg_count_grad: gradient gathers by colour runs from rows of this many values
 *                  (default 128; narrower rows use the per-lane gathers)
 *   "seq_max_width" exact_order 1, color_runs -1: forward gathers of rows of 32 up to this many
 *                  values (at most 512), half or more of each used, read whole rows in the
 *                  reference's order (default 512; wider rows use the per-lane gathers)
 *   "vertex_order" internal vertex order set at the next pg_graph_upload: 0 = descending
 *                  degree (default); 1 = degree classes aligned with the bucket thresholds,
 *                  breadth-first within a class (one-coloring passes of smaller graphs gain
 *                  a few percent, tuned batched passes of larger ones lose; results equal)
 *   "adjoint_recompute" pg_count_grad: 0 keeps every forward table for the reverse
 *                  sweep, 1 rebuilds each table from its subtree when the sweep needs it,
 *                  -1 rebuilds only when that fits four times as many colorings per pass,
 *                  or when kept tables do not fit one coloring (default)
 * Read-only: "last_tile" (colorings per pass of the last count / gradient call) and
 * "last_recompute" (1 if that gradient call rebuilt tables). */
int pg_set_option(pg_ctx* ctx, const char* key, int64_t value);
int pg_get_option(const pg_ctx* ctx, const char* key, int64_t* value);

/* ---- graph (replaces CSCGraph::createFromEdgeListFile + splitCSC,
 *      CSCGraph.hpp:96-217, and the binary deserialize, CSCGraph.hpp:439-463) ----
 * colptr[n+1], rowidx[nnz]: the reference's CSC (column u lists the rows v
 * adjacent to u).  vals[nnz] (NULL = all 1.0) weights the SpMM
 * Y[v] += A[v][u] * T[u], A[v][u] being the entry (row v, column u); the
 * reference stores 1.0 and ignores the values (CSCGraph.hpp:153, 283). */
int pg_graph_upload(pg_ctx* ctx, int32_t n, int64_t nnz, const int32_t* colptr,
                    const int32_t* rowidx, const float* vals);
/* Replace the entry weights: host array (on_device = 0) or a device pointer
 * on the context's device (on_device = 1); NULL restores unit weights. */
int pg_graph_set_weights(pg_ctx* ctx, const float* vals, int on_device);
int pg_graph_stats(const pg_ctx* ctx, int64_t* n, int64_t* nnz, int64_t* n_active,
                   int64_t* max_degree);

/* ---- plan (replaces DivideTemplates::DivideTp/sort_tps,
 *      DivideTemplates.cpp:6-327, and IndexSys::initialization,
 *      IndexSys.cpp:9-57, as consumed by CountMat::compute) ----
 * sub_size[nsub]; main_idx/aux_idx[nsub] (-1 for leaves);
 * comb_to_count: for s = 0..nsub-1 in order, C(k, sub_size[s]) ints
 *   (IndexSys::getCombToCountTable);
 * split_main/split_aux: for every non-leaf s in order, C(k,size) rows of
 *   C(size, main size) ints, row i = i-th lexicographic color set
 *   (IndexSys::getSplitToCountTable()[0/1][s]). */
int pg_plan_create(pg_ctx* ctx, int k, int nsub, const int32_t* sub_size,
                   const int32_t* main_idx, const int32_t* aux_idx,
                   const int32_t* comb_to_count, const int32_t* split_main,
                   const int32_t* split_aux, pg_plan** out);
/* Convenience: run the library's DivideTemplates + IndexSys on a template
 * tree given as an edge list over vertices 0..nv-1 (Graph::build_graph order,
 * Graph.cpp:82-129), then pg_plan_create. */
int pg_plan_from_edges(pg_ctx* ctx, int nv, int ne, const int32_t* src, const int32_t* dst,
                       pg_plan** out);
void pg_plan_destroy(pg_plan* plan);
int pg_plan_info(const pg_plan* plan, int* k, int* nsub, int* n_nonleaf, int* scaled_sub);
/* Host only (no device needed): the DivideTemplates + IndexSys tables of a
 * template tree, flattened exactly as pg_plan_create takes them.  First
 * pg_template_sizes for the lengths (nsub, total comb_to_count ints, total
 * split ints per array), then pg_template_tables to fill caller arrays. */
int pg_template_sizes(int nv, int ne, const int32_t* src, const int32_t* dst, int32_t* nsub,
                      int64_t* n_comb, int64_t* n_split);
int pg_template_tables(int nv, int ne, const int32_t* src, const int32_t* dst, int32_t* sub_size,
                       int32_t* main_idx, int32_t* aux_idx, int32_t* comb_to_count,
                       int32_t* split_main, int32_t* split_aux);

/* ---- analytic estimators (replace CountMat::estimatePeakMemUsage,
 * estimateFlops{PGBSC,Fascia,PrunedFascia}, estimateMemComm{PGBSC,Fascia,
 * PrunedFascia} and estimateTemplate, CountMat.cpp:1067-1446) ----
 * The reference's models, in its units: "GB", "GBytes" and "Gflop" are
 * 2^30, and the estimateTemplate figures are unscaled.  Inputs are the
 * graph's vertex count (holes removed, as the reference's CSC holds it) and
 * directed nnz, and the subtemplate chain as pg_plan_create takes it
 * (sub_size, main_idx, aux_idx; leaves -1).  Host only. */
typedef struct pg_estimate {
    double peak_mem_gb;                            /* estimatePeakMemUsage */
    double pgbsc_spmv_gflop, pgbsc_fma_gflop;      /* estimateFlopsPGBSC */
    double pgbsc_spmv_gbytes, pgbsc_fma_gbytes;    /* estimateMemCommPGBSC */
    double fascia_gflop, fascia_gbytes;            /* estimate{Flops,MemComm}Fascia */
    double pruned_fascia_gflop, pruned_fascia_gbytes; /* estimate{Flops,MemComm}PrunedFascia */
    double workload_n, workload_nnz;               /* estimateTemplate */
    double memload_n, memload_nnz;
} pg_estimate;
int pg_estimate_model(int64_t n, int64_t nnz, int k, int nsub, const int32_t* sub_size,
                      const int32_t* main_idx, const int32_t* aux_idx, pg_estimate* out);
/* This library's own figures for a pass of `colorings_per_pass` colorings of
 * `plan` on the uploaded graph (no reference counterpart; the device side of
 * estimatePeakMemUsage): device bytes of the pass's memory plan (count
 * tables placed over their lifetimes, per-pass buffers, graph), and the
 * algorithmic HBM bytes per coloring the roofline figures use (each table
 * written once and read once, the CSC streamed once per gathering launch). */
int pg_plan_footprint(pg_ctx* ctx, const pg_plan* plan, int colorings_per_pass, int precision,
                      int64_t* device_bytes, double* algorithmic_bytes_per_coloring);

/* ---- colorings (replaces CountMat::colorInit, CountMat.cpp:979-991) ----
 * Color of vertex v (caller numbering) in coloring c of stream `seed`:
 *   h = splitmix64(seed ^ splitmix64(c * 0x9E3779B97F4A7C15 + v));
 *   color = ((h >> 32) * k) >> 32.
 * The device generates the same colors when pg_count gets colors == NULL. */
void pg_colors_fill(uint64_t seed, int64_t first_coloring, int ncolorings, int32_t n, int k,
                    uint8_t* out /* [ncolorings][n] */);

/* ---- counting (replaces the loop of CountMat::colorCounting,
 *      CountMat.cpp:203-205 and 268-322, whose per-subtemplate body is
 *      countNonBottomePrunedSPMM, CountMat.cpp:556-849) ----
 * colors: [ncolorings][n] uint8 in caller numbering, or NULL to draw them on
 * the device from (seed, coloring index).  raw_out[ncolorings] receives each
 * coloring's raw colorful count exactly as colorCounting returns it (FP32:
 * the reference's 1e-12 / 1e12 round trip; FP64: exact). */
int pg_count(pg_ctx* ctx, const pg_plan* plan, int ncolorings, const uint8_t* colors,
             uint64_t seed, int precision, uint32_t flags, double* raw_out);

/* No reference counterpart: pg_count for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as pg_count (PG_FLAG_UNPRUNED
 * excluded), but every buffer of a pass lives in a static device array of
 * 64 MiB that the library holds on each device it runs on until the process
 * ends: the call makes no hipMalloc, hipFree or other allocation, captures
 * no graph and records no event.  A graph and template whose one-coloring
 * pass (plus the caller's colors) needs more than 64 MiB are rejected with
 * PG_ERR_RANGE; up to 64 colorings per pass are used when they fit.  The
 * graph upload and plan creation that precede it allocate as usual, and the
 * degree-bucket options must not have changed since the last allocating call
 * (PG_ERR_STATE).  Every call on a device shares that device's array, so
 * calls that overlap on one GPU (from several contexts or threads) race;
 * each call writes every buffer it reads, so nothing depends on what an
 * earlier call left there. */
int pg_count_noalloc(pg_ctx* ctx, const pg_plan* plan, int ncolorings, const uint8_t* colors,
                     uint64_t seed, int precision, uint32_t flags, double* raw_out);

/* Mixed batch: plans[p] is counted for ncolorings[p] colorings drawn on the
 * device from seeds[p]; raw_out is the concatenation of the per-plan results,
 * equal to what pg_count gives each plan with the same seed.  With graphs on
 * and timing off, each round of passes (one pass of every plan with
 * colorings left) is one launch of a captured graph with a branch per plan,
 * when the plans fit side by side in memory and share bucket settings;
 * otherwise the plans run one after another.  No reference counterpart (the
 * reference counts one template per run). */
int pg_count_multi(pg_ctx* ctx, int nplans, const pg_plan* const* plans,
                   const int32_t* ncolorings, const uint64_t* seeds, int precision,
                   uint32_t flags, double* raw_out);

/* ---- differentiation (no reference counterpart: the reference ignores edge
 *      values, CSCGraph.hpp:283) ----
 * Forward as pg_count with the current weights, then the adjoint DP:
 * grad[e] = sum over colorings c of cotangent[c] * d raw_c / d A-entry e
 * (caller CSC entry order, nnz values; cotangent NULL = all ones).  At unit
 * weights this is the number of colorful maps using entry e.
 * grad_on_device = 1: grad is a device pointer (float for PG_FP32, double
 * for PG_FP64), else a host double array.  Needs a symmetric pattern. */
int pg_count_grad(pg_ctx* ctx, const pg_plan* plan, int ncolorings, const uint8_t* colors,
                  uint64_t seed, int precision, uint32_t flags, const double* cotangent,
                  double* raw_out, void* grad, int grad_on_device);

/* ---- timing of the last pg_count / pg_count_grad call ----
 * total device milliseconds, and the dominant kernel's launches and summed
 * milliseconds (HIP events on the context's stream). */
int pg_last_timing(const pg_ctx* ctx, double* total_ms, double* spmm_ms, int64_t* spmm_launches,
                   double* spmm_bytes);
/* name of that dominant kernel ("k_sub<float,64,0,1>": dtype, team lanes,
 * aux mode, main mode); timing is collected when option "timing" is 1 */
int pg_last_timing_kernel(const pg_ctx* ctx, char* buf, int len);
/* per-kernel breakdown of the same call (the VERBOSE-style per-phase report
 * of CountMat::colorCounting, CountMat.cpp:268-322, as device timings):
 * entry idx (0-based, kernels in name order) -> name, summed ms, launches and
 * algorithmic bytes.  Returns the number of kernels, or PG_ERR_RANGE for an
 * idx past the end (name/ms/launches/bytes may be NULL). */
int pg_last_timing_entry(const pg_ctx* ctx, int idx, char* name, int len, double* ms,
                         int64_t* launches, double* bytes);

/* the same launches broken down by (subtemplate, kernel): the per-subtemplate
 * "Sub %d, counting time ..." report of the reference's VERBOSE build
 * (CountMat.cpp:833-840, 967-970).  *sub is the subtemplate index (forward
 * and adjoint steps of it alike), -1 for launches outside any step (colors,
 * the neighbour-color histogram, root reductions).  Entries are ordered by
 * (sub, name).  Returns the number of entries, or PG_ERR_RANGE past the end. */
int pg_last_timing_sub(const pg_ctx* ctx, int idx, int* sub, char* name, int len, double* ms,
                       int64_t* launches, double* bytes);

/* ---- host surface helpers used by the CLI and the Python layer ---- */
/* R-MAT generator: n = 2^scale, edge_factor*n draws with quadrant
 * probabilities a,b,c (d = 1-a-b-c), self loops dropped, (min,max)
 * canonical, duplicates dropped, vertex labels permuted by a seeded shuffle.
 * Writes at most cap undirected edges; *m_out = count (call with cap = 0 to
 * size).  No reference counterpart (graph500-1.2/ is not built). */
int pg_rmat_edges(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                  int64_t cap, int32_t* src, int32_t* dst, int64_t* m_out);
/* CSC build exactly as CSCGraph::createFromEdgeListFile (CSCGraph.hpp:96-178):
 * colptr[n+1], rowidx[2m] */
int pg_csc_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                      int32_t* colptr, int32_t* rowidx);
/* Replaces CountMat::automorphismNum (CountMat.cpp:1498-1572): the number of
 * permutations s of the template's vertices that the reference's
 * calcAutomorphismZero (CountMat.cpp:1510-1544) accepts, i.e. deg(i) =
 * deg(s(i)) and N(i) a subset of s(N(s(i))) for every vertex i; for most
 * trees this is the automorphism group's order.  Host only.  Returns the
 * count, or PG_ERR_INVALID. */
int64_t pg_template_automorphisms(int nv, int ne, const int32_t* src, const int32_t* dst);
/* The text loader's "hole" rule (EdgeList.cpp:70-112) on an in-memory list:
 * if the largest id is not n-1, the ids that occur are renumbered 0.. in
 * ascending order.  Returns the resulting n (or a negative PG_ERR_*). */
int64_t pg_edges_remove_holes(int64_t n, int64_t m, int32_t* src, int32_t* dst);

#ifdef __cplusplus
}
#endif
#endif
