// C ABI of libpgbsc.so: context, graph upload (relabelling and degree
// buckets), plans (color-relative index tables), the per-pass DP schedule and
// its memory plan.
//
// Reference counterparts: CountMat::initialization/compute/colorCounting
// (CountMat.cpp:21-322), DataTableColMajor table lifetimes
// (DataTableColMajor.cpp:63-249), CSCGraph (CSCGraph.hpp:96-217),
// IndexSys (IndexSys.cpp:315-436) for the split order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "pg_host.hpp"
#include "pg_kernels.hpp"
#include "pgbsc.h"

using namespace pgbsc;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

struct LaunchRec {
    std::string name;
    double bytes;
    int sub;  // subtemplate index, -1 for pass-level launches (colors, histogram, reductions)
    hipEvent_t beg, end;
};

struct KernelTotals {
    double ms = 0, bytes = 0;
    int64_t launches = 0;
};

struct GraphKey {
    const pg_plan* plan;
    int B, tsize;
    const void* colors_in;  // caller colors buffer, nullptr = device RNG
    const void* arena;
    uint64_t generation;    // graph / weights / bucket version
    int kind = 0;           // 0 count pass, 1 gradient pass (kept tables), 2 gradient pass (rebuilt tables)
    const void* grad = nullptr;  // gradient accumulator of a gradient pass
    bool operator<(const GraphKey& o) const
    {
        return std::tie(plan, B, tsize, colors_in, arena, generation, kind, grad) <
               std::tie(o.plan, o.B, o.tsize, o.colors_in, o.arena, o.generation, o.kind, o.grad);
    }
};

struct PassGraph {
    hipGraphExec_t exec = nullptr;
};

// One non-leaf subtemplate as the device schedule sees it.  Widths are per
// coloring in color-relative rows (K = k - 1):
//   Wo = C(K, size-1) output values, Wy = C(K, aa) aux sums,
//   Wa = C(K, aa-1) values per row of the aux child, Wm = C(K, mm-1) of the main child,
//   P  = C(size-1, mm-1) non-zero splits per output set.
// Full (reference) widths Cs, Ca, Cm are kept for the algorithmic byte count.
struct DevSub {
    int s, size, mi, ai, mm, aa;
    int Wo, Wy, Wa, Wm, P;
    int Cs, Ca, Cm;
    bool root, scaled;
    DevBuf split;  // [Wo][P] uint32 (main_rank << 16 | y_rank)
    DevBuf ymask;  // [Wy] uint16
    // adjoint: the split list transposed by aux set and by main set, and the
    // row masks of the aux child (sets of size aa-1)
    int Pa = 0, Pm = 0;
    DevBuf by_aux;   // [Wy][Pa] uint32 (o << 16 | main rank)
    DevBuf by_main;  // [Wm][Pm] uint32 (o << 16 | y rank)
    DevBuf by_auxT;  // the same lists split-major: [Pa][Wy], [Pm][Wm]
    DevBuf by_mainT;
    DevBuf amask;    // [Wa] uint16
    // gather index table of AUX_GATHER (SubArgs::gidx), when it fits kGidxMaxBytes
    DevBuf gidx;     // [(k+1)*k][Wy4] int16
    int Wy4 = 0;
    // row-wise index table of AUX_ROWS (SubArgs::xidx), when it fits kGidxMaxBytes
    DevBuf xidx;     // [k*k][Wa4] int16
    int Wa4 = 0;
    DevBuf xidx_g;   // the same targets per lane for rows up to 512 values (RUNS_SEQ): [k*k][64][RV]
    // the adjoint's transposed gather (AUX_GATHER_T) through its own index table
    DevBuf gidxT;    // [(k+1)*k][Wa4] int16
    // the fused transposed gather's per-lane index table (SubArgs::ttab)
    DevBuf ttab;     // [k*k][64][RV] int16, RV = runs_values_per_lane(Wy)
};

constexpr size_t kGidxMaxBytes = 24 * 1024;

}  // namespace

struct pg_plan {
    pg_ctx* ctx = nullptr;
    int device = 0;
    int k = 0, K = 0, nsub = 0, nonleaf = 0, scaled_sub = -1;
    std::vector<int> sizes, mains, auxs;
    std::vector<std::unique_ptr<DevSub>> subs;  // processing order (descending s)
    DevBuf rank_tab;                            // [2^K] int16
    // the unpruned mode's tables over absolute color sets, per subtemplate
    // index (non-leaf only): combToCount [C(k,size)] and the split pairs
    // [C(k,size)][C(size,main)] as (split_main << 16 | split_aux), uploaded
    // on first use
    std::vector<std::vector<int32_t>> h_c2c;
    std::vector<std::vector<uint32_t>> h_fsplit;
    std::vector<std::unique_ptr<DevBuf>> d_c2c, d_fsplit;
};

struct pg_ctx {
    int device = 0, num_cus = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // options
    // bucket defaults: what the autotuner settles on for one coloring per pass
    // on power-law graphs (small teams for low degrees); wide batched rows run
    // 64-lane teams whatever the thresholds
    int64_t hub_degree = 512, hub_chunk = 1024, batch_tile = 0, mem_budget_mb = 0, timing = 0, autotune = 0;
    int64_t bucket_high = 512, bucket_mid = 64;
    int64_t blocks_per_cu = 0;  // 0: resident workgroups from the occupancy query
    int64_t aux_rows = 0;       // neighbour sums row-wise: 1 always, 0 never, -1 per subtemplate
    int64_t stage_rows = 0;     // gathers read neighbour rows staged in LDS (1) or directly (0)
    int64_t split_lds = 1;      // combines read the split table from LDS (1) when it fits, or from memory (0)
    int64_t color_runs = -1;    // wide-row neighbour sums by colour runs: 1 always, 0 never, -1 rows of 32+ values
    int64_t exact_order = 1;    // forward neighbour sums in the reference's order (one group, hubs split by columns)
    int64_t fuse_grad = 1;      // adjoint: forward Y kept and the gradient taken in the transposed gather
    int64_t node_fp32 = 1;      // adjoint node sums of FP32 tables: 1 FP32 fused multiply-adds, 0 FP64
    int64_t runs_g_min = 128;   // gradient gathers by colour runs from this row width (color_runs -1)
    int64_t fuse_min_width = 128;   // fused gradient also for dY rows of at least this many values
    int64_t seq_max_width = 512;    // exact-order gathers by whole rows (RUNS_SEQ) up to this row width
    int64_t node_rows = 1;      // adjoint node launches: a thread per value over a group's rows (1) or per (row, value)
    int64_t reg_cols = 0;       // values per lane of gathers: 0 one for rows under kRegCols, else 1 or 4
    int64_t adjoint_recompute = -1;  // adjoint tables: 0 kept, 1 recomputed, -1 recomputed when that fits 4x more
    int64_t vertex_order = 0;   // internal order at the next upload: 0 descending degree, 1 degree classes + BFS
    int64_t last_tile = 0, last_recompute = 0;  // of the last count / gradient call (read-only options)
    // graph
    bool has_graph = false, weighted = false;
    int32_t n = 0, n_act = 0, max_deg = 0;
    int64_t nnz = 0;
    std::vector<int32_t> h_rowptr, h_col, h_perm;
    // adjoint support, built on first use: row of each entry, the entry of
    // the transposed pair, and weights in transposed order
    bool has_transpose = false, symmetric = false;
    DevBuf d_erow, d_tr, d_wT32, d_wT64, d_grad, d_grad_out;
    DevBuf d_rowptr, d_col, d_ent, d_perm, d_w32, d_w64, d_wsrc;
    // degree buckets and hub work items for the current options
    int64_t built_deg = -1, built_chunk = -1, built_high = -1, built_mid = -1;
    int32_t n_hub = 0, hi_end = 0, mid_end = 0;  // [0,n_hub) hubs, [n_hub,hi_end) high, [hi_end,mid_end) mid
    int32_t n_hub_items = 0, n_slots = 0;
    DevBuf d_hub_items, d_slot_begin;
    DevBuf d_hub_iota;  // [n_hub+1] h -> h: one partial row per hub (column-split hub launches)
    // per-call workspace
    DevBuf d_arena, d_colors_in;
    // pg_count_noalloc: the device address of this device's static scratch
    // array, and (during such a call) the base the pass buffers are placed at
    char* scratch = nullptr;
    char* arena_override = nullptr;
    // captured passes, and what invalidates them
    bool use_graphs = true;
    uint64_t generation = 0;
    std::map<GraphKey, PassGraph> graphs;
    // mixed batches (pg_count_multi): one graph per set of (plan, colorings,
    // settings) with a branch per plan, the side streams its branches are
    // captured on, and the fork / join events
    std::map<std::vector<int64_t>, PassGraph> multi_graphs;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev_join;
    hipEvent_t ev_fork = nullptr;
    // hub launches of a gathering step run on hub_stream beside the other
    // vertices' launch (fork and join events), option "hub_concurrent"
    hipStream_t hub_stream = nullptr;
    hipEvent_t ev_hub_fork = nullptr, ev_hub_join = nullptr;
    int64_t hub_concurrent = 1;
    bool in_branches = false;  // enqueueing the branches of a mixed-batch capture
    std::map<std::tuple<const pg_plan*, int, int>, std::vector<int64_t>> tuned;  // (plan, B, tsize) -> settings, tile
    // the unpruned mode (PG_FLAG_UNPRUNED): the reference's _bufVec, zeroed
    // at graph upload and then only ever added to, and its tables
    DevBuf d_fascia_acc, d_fascia;
    bool fascia_acc_ready = false;
    // timing of the last call
    std::vector<LaunchRec>* cur_recs = nullptr;
    std::map<std::string, KernelTotals> agg;
    // the same launches keyed by (subtemplate, kernel); cur_sub tags new records
    std::map<std::pair<int, std::string>, KernelTotals> agg_sub;
    int cur_sub = -1;
    double last_total_ms = 0;
    std::string dom_name;
    double dom_ms = 0, dom_bytes = 0;
    int64_t dom_launches = 0;
};

namespace {

int set_err(pg_ctx* ctx, int code, const char* fmt, ...)
{
    if (!ctx) return code;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return set_err(ctx, e_ == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "%s: %s (%s:%d)", \
                           #expr, hipGetErrorString(e_), __FILE__, __LINE__);                          \
    } while (0)

int pow2_at_least(int x)
{
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// column lanes for `cols` value slots (1..64)
int cl_for(int cols) { return std::min(64, pow2_at_least(std::max(cols, 1))); }
// value slots of a neighbour-summing launch: rc values of one coloring each
int slots_for(int B, int W, int rc) { return B * ((W + rc - 1) / rc); }
// lanes of a team that only streams rows (no neighbour loop)
int team_for(int cols) { return std::max(8, cl_for(cols)); }

// (Re)build degree buckets and hub chunk items for the current options.
void drop_graphs(pg_ctx* c, const pg_plan* only = nullptr);

int build_buckets(pg_ctx* c)
{
    if (c->built_deg == c->hub_degree && c->built_chunk == c->hub_chunk && c->built_high == c->bucket_high &&
        c->built_mid == c->bucket_mid)
        return PG_OK;
    auto deg = [&](int32_t v) { return (int64_t)c->h_rowptr[v + 1] - c->h_rowptr[v]; };
    int32_t nh = 0;
    while (nh < c->n_act && deg(nh) > c->hub_degree) ++nh;
    int32_t hi = nh;
    while (hi < c->n_act && deg(hi) >= c->bucket_high) ++hi;
    int32_t mid = hi;
    while (mid < c->n_act && deg(mid) >= c->bucket_mid) ++mid;
    std::vector<int4> items;
    std::vector<int32_t> slot_begin(nh + 1, 0);
    for (int32_t h = 0; h < nh; ++h) {
        slot_begin[h] = (int32_t)items.size();
        for (int32_t e = c->h_rowptr[h]; e < c->h_rowptr[h + 1]; e += (int32_t)c->hub_chunk) {
            int32_t e1 = (int32_t)std::min<int64_t>((int64_t)e + c->hub_chunk, c->h_rowptr[h + 1]);
            items.push_back(make_int4(h, e, e1, (int)items.size()));
        }
    }
    slot_begin[nh] = (int32_t)items.size();
    c->n_hub = nh;
    c->hi_end = hi;
    c->mid_end = mid;
    c->n_hub_items = (int32_t)items.size();
    c->n_slots = (int32_t)items.size();
    HIP_TRY(c, c->d_hub_items.alloc(std::max<size_t>(items.size(), 1) * sizeof(int4)));
    HIP_TRY(c, c->d_slot_begin.alloc(slot_begin.size() * sizeof(int32_t)));
    if (!items.empty())
        HIP_TRY(c, hipMemcpy(c->d_hub_items.p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_slot_begin.p, slot_begin.data(), slot_begin.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    {
        std::vector<int32_t> iota((size_t)nh + 1);
        std::iota(iota.begin(), iota.end(), 0);
        HIP_TRY(c, c->d_hub_iota.alloc(iota.size() * sizeof(int32_t)));
        HIP_TRY(c, hipMemcpy(c->d_hub_iota.p, iota.data(), iota.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    drop_graphs(c);  // captured passes hold the old bucket ranges
    c->generation++;
    c->built_deg = c->hub_degree;
    c->built_chunk = c->hub_chunk;
    c->built_high = c->bucket_high;
    c->built_mid = c->bucket_mid;
    return PG_OK;
}

// ---- memory plan for one pass: first-fit offsets over table lifetimes ----
struct MemPlan {
    size_t total = 0;
    size_t off_H = 0, off_part = 0, off_root = 0, off_colors = 0, off_ecol = 0, off_rawdev = 0, off_tmp = 0;
    size_t off_params = 0;
    std::vector<size_t> off_tab;  // per subtemplate index (tables of non-root non-leaf)
    // per subtemplate index, 0 = none: Y kept by the forward step for the
    // adjoint (gradient calls with every table kept)
    std::vector<size_t> off_ykeep;
};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

MemPlan plan_memory(const pg_ctx* c, const pg_plan* p, int B, size_t tsz)
{
    MemPlan m;
    const int S = p->nsub;
    m.off_tab.assign(S, 0);
    struct Iv { size_t bytes; int beg, end; size_t* off; };
    std::vector<Iv> ivs;
    const size_t na = (size_t)c->n_act;
    const int nsteps = (int)p->subs.size();
    // pass-long buffers: H, hub partial rows, root values, colors, results
    ivs.push_back({align_up(na * B * p->k * tsz, 256), -1, nsteps, &m.off_H});
    int maxY = p->k;
    for (auto& d : p->subs)
        if (d->aa > 1) maxY = std::max(maxY, std::max(d->Wy, d->Wa));  // Wa: adjoint gathers
    ivs.push_back({align_up((size_t)std::max(c->n_slots, 1) * B * maxY * tsz, 256), -1, nsteps, &m.off_part});
    ivs.push_back({align_up(na * B * sizeof(double), 256), -1, nsteps, &m.off_root});
    ivs.push_back({align_up(na * B, 256), -1, nsteps, &m.off_colors});
    ivs.push_back({align_up((size_t)std::max<int64_t>(c->nnz, 1) * B, 256), -1, nsteps, &m.off_ecol});
    ivs.push_back({align_up((size_t)B * sizeof(double), 256), -1, nsteps, &m.off_rawdev});
    ivs.push_back({align_up((size_t)B * kReduceBlocks * sizeof(double), 256), -1, nsteps, &m.off_tmp});
    ivs.push_back({align_up(sizeof(PassParams), 256), -1, nsteps, &m.off_params});
    // table of step i lives until its parent's step consumes it
    for (int i = 0; i < nsteps; ++i) {
        const DevSub* d = p->subs[i].get();
        if (d->root) continue;
        int parent_step = nsteps;
        for (int j = i + 1; j < nsteps; ++j)
            if (p->subs[j]->mi == d->s || p->subs[j]->ai == d->s) { parent_step = j; break; }
        ivs.push_back({align_up(na * B * d->Wo * tsz, 256), i, parent_step, &m.off_tab[d->s]});
    }
    struct Placed { size_t off, bytes; int beg, end; };
    std::vector<Placed> placed;
    for (auto& iv : ivs) {
        std::vector<std::pair<size_t, size_t>> busy;
        for (auto& pl : placed)
            if (!(pl.end < iv.beg || iv.end < pl.beg)) busy.push_back({pl.off, pl.off + pl.bytes});
        std::sort(busy.begin(), busy.end());
        size_t off = 0;
        for (auto& bz : busy) {
            if (off + iv.bytes <= bz.first) break;
            off = std::max(off, bz.second);
        }
        *iv.off = off;
        placed.push_back({off, iv.bytes, iv.beg, iv.end});
        m.total = std::max(m.total, off + iv.bytes);
    }
    return m;
}

// How a subtemplate with a non-leaf aux child sums over neighbours.  Row-wise
// (AUX_ROWS) reads each neighbour row whole, which pays when most of it is
// used: a neighbour of another color adds C(K-1, aa-1) of its C(K, aa-1)
// values, a fraction (K - aa + 1) / K.  Option "aux_rows": 1 always, 0 never,
// -1 by that fraction.
int aux_mode_for(const pg_ctx* c, const pg_plan* p, const DevSub& d)
{
    if (c->aux_rows == 0 || c->exact_order) return AUX_GATHER;
    if (c->aux_rows == 1) return AUX_ROWS;
    return 2 * (p->K - d.aa + 1) >= p->K ? AUX_ROWS : AUX_GATHER;
}

// Values per lane of a neighbour-summing launch: kRegCols, except for gathers
// whose Y row per coloring is narrower than that (the root's single value,
// where kRegCols per lane would load three unused values per neighbour), which
// take one.  Option "reg_cols": 0 that rule, 1 or kRegCols for every gather.
int rc_for(const pg_ctx* c, int aux, int B, int Wy)
{
    (void)B;
    const bool gather = aux == AUX_GATHER || aux == AUX_GATHER_G || aux == AUX_GATHER_T;
    if (!gather) return kRegCols;
    if (c->reg_cols != 0) return (int)c->reg_cols;
    return Wy < kRegCols ? 1 : kRegCols;
}

// column lanes of a neighbour-summing launch
int cl_of(int aux, int B, int k, int Wy, int Wa, int rc)
{
    if (aux == AUX_ROWS) return cl_for(B * Wa);
    return cl_for(slots_for(B, aux == AUX_COLORS ? k : Wy, rc));
}

// LDS row staging of a gathering launch (AUX_GATHER with a gather index
// table): the longest row segment one chunk of cl slots needs, padded to a
// multiple of cl; 0 (direct gathers) when a segment needs more than
// kStageRegs values per lane or option "stage_rows" is 0
constexpr size_t kStageLdsMax = 32 * 1024;  // staged rows per workgroup

int stage_pad(const pg_ctx* c, int aux, bool has_gidx, int B, int Wy, int Wa, int cl, size_t tsz)
{
    if (aux != AUX_GATHER || !has_gidx || c->stage_rows == 0) return 0;
    const int spb = (Wy + kRegCols - 1) / kRegCols;
    const int nslots = B * spb;
    int mx = 0;
    for (int s0 = 0; s0 < nslots; s0 += cl) {
        const int bf = s0 / spb, bl = std::min(B - 1, (s0 + cl - 1) / spb);
        mx = std::max(mx, (bl - bf + 1) * Wa);
    }
    if (mx > kStageRegs * cl) return 0;
    const int pad = (mx + cl - 1) / cl * cl;
    // every workgroup stages kBlock / cl groups' rows, whatever the team shape
    if ((size_t)(kBlock / cl) * stage_group_stride(pad, cl) * tsz > kStageLdsMax) return 0;
    return pad;
}

// The combine may read the split table from LDS (option "split_lds"): a
// main-table step's outputs read it Wo * P times per vertex and coloring.
// Taken when the table fits kSplitLdsMax and the step's launches still fit
// kLdsLimit with it.
constexpr size_t kSplitLdsMax = 24 * 1024;
constexpr size_t kLdsLimit = 64 * 1024;
bool split_lds_candidate(const pg_ctx* c, const DevSub& d)
{
    return c->split_lds && d.mm > 1 && !d.root && (size_t)4 * d.P * d.Wo <= kSplitLdsMax;
}

template <typename T>
size_t lds_of(int team, int aux, int main_mode, int ng, int out_mode, int B, const pg_plan* p, const DevSub& d,
              int segpad, int split_lds)
{
    SubArgs<T> a{};
    a.P = d.P;
    a.Wo = d.Wo;
    a.split_lds = split_lds;
    a.B = B;
    a.k = p->k;
    a.K = p->K;
    a.Wy = d.Wy;
    a.Wa = d.Wa;
    a.Wm = d.Wm;
    a.ng = ng;
    a.out_mode = out_mode;
    a.gidx = (const int16_t*)d.gidx.p;
    a.Wy4 = d.Wy4;
    a.xidx = (const int16_t*)d.xidx.p;
    a.Wa4 = d.Wa4;
    a.segpad = segpad;
    return sub_lds_bytes<T>(team, aux, main_mode, a);
}

int runs_mode_for(const pg_ctx* c, int aux, int wg, int used_per_mille);
int runs_used_per_mille(const pg_plan* p, const DevSub& d, int aux);
bool runs_seq_for(const pg_ctx* c, int wa, int used_per_mille);

// LDS of the largest launch of subtemplate d for B colorings, over every
// launch shape launch_gathering (or the aux-leaf read) would use, with or
// without the split table staged
template <typename T>
size_t sub_lds_max_t(const pg_ctx* c, const pg_plan* p, const DevSub& d, int B, int split)
{
    const int mm = d.mm > 1 ? MAIN_TABLE : MAIN_LEAF;
    const int out = d.root ? OUT_ROOT : OUT_EMA;
    const int fin_team = team_for(std::max(B * d.Wy, B * d.Wo));
    size_t m = lds_of<T>(fin_team, d.aa == 1 ? AUX_READ : AUX_PARTIALS, mm, 1, out, B, p, d, 0, split);
    if (d.aa == 1) return m;
    const int aux = aux_mode_for(c, p, d);
    if (runs_mode_for(c, aux, d.Wa, runs_used_per_mille(p, d, AUX_GATHER)) >= 0) {
        // colour runs: a team per (vertex, colouring), LDS independent of B
        SubArgs<T> a{};
        a.B = B;
        a.K = p->K;
        a.Wy = d.Wy;
        a.Wm = d.Wm;
        a.Wo = d.Wo;
        a.P = d.P;
        a.split_lds = split;
        for (int o : {out, (int)OUT_PARTIAL}) {
            a.out_mode = o;
            m = std::max(m, runs_lds_bytes<T>(mm, a));
        }
        return m;
    }
    const int rc = rc_for(c, aux, B, d.Wy);
    if (c->exact_order && aux == AUX_GATHER) {
        // launch_gathering_exact: hub rows (k_hub_seq), the other vertices by
        // whole rows (RUNS_SEQ, LDS independent of B) or by teams of
        // max(8, cl) lanes, one group
        const int cl = std::max(8, cl_of(aux, B, p->k, d.Wy, d.Wa, rc));
        m = std::max(m, hub_seq_lds_bytes(p->K));
        if (d.xidx_g.p && runs_seq_for(c, d.Wa, runs_used_per_mille(p, d, AUX_GATHER))) {
            SubArgs<T> a{};
            a.B = B;
            a.K = p->K;
            a.Wy = d.Wy;
            a.Wm = d.Wm;
            a.Wo = d.Wo;
            a.P = d.P;
            a.split_lds = split;
            a.out_mode = out;
            return std::max(m, runs_lds_bytes<T>(mm, a));
        }
        const int sp = rc == kRegCols ? stage_pad(c, aux, d.gidx.p != nullptr, B, d.Wy, d.Wa, cl, sizeof(T)) : 0;
        return std::max(m, lds_of<T>(cl, sp > 0 ? AUX_STAGED : aux, mm, 1, out, B, p, d, sp, split));
    }
    const int cl = cl_of(aux, B, p->k, d.Wy, d.Wa, rc);
    const int sp = rc == kRegCols ? stage_pad(c, aux, d.gidx.p != nullptr, B, d.Wy, d.Wa, cl, sizeof(T)) : 0;
    const int aux_l = sp > 0 ? AUX_STAGED : aux;
    m = std::max(m, lds_of<T>(64, aux_l, mm, 64 / cl, OUT_PARTIAL, B, p, d, sp, split));
    for (int lanes : {64, 16, 8}) {
        const int team = std::max(lanes, cl);
        m = std::max(m, lds_of<T>(team, aux_l, mm, team / cl, out, B, p, d, sp, split));
    }
    return m;
}

// whether subtemplate d's combine reads the split table from LDS at B colorings
int step_split_lds(const pg_ctx* c, const pg_plan* p, const DevSub& d, int B, size_t tsz)
{
    if (!split_lds_candidate(c, d)) return 0;
    const size_t m = tsz == 4 ? sub_lds_max_t<float>(c, p, d, B, 1) : sub_lds_max_t<double>(c, p, d, B, 1);
    return m <= kLdsLimit;
}

size_t sub_lds_max(const pg_ctx* c, const pg_plan* p, const DevSub& d, int B, size_t tsz)
{
    const int split = step_split_lds(c, p, d, B, tsz);
    return tsz == 4 ? sub_lds_max_t<float>(c, p, d, B, split) : sub_lds_max_t<double>(c, p, d, B, split);
}

// Timing: an event pair around each launch, appended to c->cur_recs.  Passes
// run eagerly while timing is on: event-record nodes inside an instantiated
// graph give no usable elapsed times on this runtime, so graphs carry none.
void record(pg_ctx* c, const std::string& name, double bytes, bool begin)
{
    if (!c->timing || !c->cur_recs) return;
    if (begin) {
        LaunchRec r{name, bytes, c->cur_sub, nullptr, nullptr};
        (void)hipEventCreate(&r.beg);
        (void)hipEventCreate(&r.end);
        (void)hipEventRecord(r.beg, c->stream);
        c->cur_recs->push_back(r);
    } else {
        (void)hipEventRecord(c->cur_recs->back().end, c->stream);
    }
}

// adds the elapsed times of `recs` (all complete) to the per-kernel totals
void accumulate_timing(pg_ctx* c, const std::vector<LaunchRec>& recs)
{
    for (auto& r : recs) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.beg, r.end) != hipSuccess) continue;
        for (KernelTotals* a : {&c->agg[r.name], &c->agg_sub[{r.sub, r.name}]}) {
            a->ms += ms;
            a->bytes += r.bytes;
            a->launches += 1;
        }
    }
}

void destroy_recs(std::vector<LaunchRec>& recs)
{
    for (auto& r : recs) {
        (void)hipEventDestroy(r.beg);
        (void)hipEventDestroy(r.end);
    }
    recs.clear();
}

template <typename T>
std::string kname(int team, int aux, int main_mode, int rc = kRegCols)
{
    char buf[96];
    snprintf(buf, sizeof buf, "k_sub<%s,%d,%d,%d%s>", sizeof(T) == 4 ? "float" : "double", team, aux, main_mode,
             rc == kRegCols ? "" : ",1");
    return buf;
}

// A step's algorithmic bytes by what they scale with: the CSC (rowptr per
// vertex, one column index and weight per entry), bytes gathered per entry
// (the aux child's rows; the adjoint's per-entry gradient), and rows read or
// written per vertex (main child, output, Y, dY).
struct StepBytes {
    double csc = 0, per_entry = 0, per_vertex = 0;
};

// The neighbour-summing launches of one subtemplate (or of H): explicit hub
// chunks first (partial rows), then the degree buckets of the remaining
// vertices with team shapes fitted to their degree, then the hub rows from
// their partials.  `a` carries everything but the item range and team shape.
// The step's algorithmic bytes are shared among its launches by the entries
// and vertices each covers: the CSC by entries + vertices, per-entry bytes by
// entries, per-vertex bytes by vertices, so per-kernel sums stay exact.
// Colour-run neighbour sums (pg_runs.hip) for a gathering launch over rows
// of wg values: RUNS_FWD for the forward gather, RUNS_T for the adjoint's
// transposed one, -1 for the k_sub kernels.  Option "color_runs": 1 always
// (rows up to 512 values), 0 never, -1 rows of at least kRunsMinWidth values
// (the 12-vertex tree's 55- to 462-value rows; narrow rows fit k_sub's
// per-lane gathers).
constexpr int kRunsMinWidth = 32;
int runs_mode_for(const pg_ctx* c, int aux, int wg, int used_per_mille)
{
    if (c->color_runs == 0 || runs_values_per_lane(wg) == 0) return -1;
    if (c->color_runs < 0 && (wg < kRunsMinWidth || used_per_mille < 500)) return -1;
    if (aux == AUX_GATHER || aux == AUX_ROWS || aux == AUX_STAGED) return c->exact_order ? -1 : RUNS_FWD;
    if (aux == AUX_GATHER_T) return RUNS_T;
    if (aux == AUX_GATHER_G) return c->color_runs > 0 || wg >= c->runs_g_min ? RUNS_G : -1;
    return -1;
}

// Share (per mille) of a gathered row's values a vertex uses: forward and
// gradient gathers read an aux row at the sets that avoid the vertex's colour,
// (K - aa + 1) / K of them; the transposed gather reads a dY row at the sets
// that hold it, aa / K.  Colour runs read whole rows, so by default they are
// taken when at least half of a row is used.
int runs_used_per_mille(const pg_plan* p, const DevSub& d, int aux)
{
    const int K = p->K;
    return aux == AUX_GATHER_T ? 1000 * d.aa / K : 1000 * (K - d.aa + 1) / K;
}

// The exact-order forward sum by whole rows (RUNS_SEQ): with colour runs on
// (1), for rows up to 512 values; by default (-1), for rows of 32 to
// seq_max_width values (option, default 512) of which a vertex uses at least
// half.  A team streams each neighbour's whole row once and takes a window's
// entry ids with one load per lane, so a row costs one round trip; the
// per-lane gathers re-walk the entry list for every chunk of 256 Y values.
// Measured on the 12-vertex tree at R-MAT s22, 16 colorings in passes of 4:
// the (7;1,6) step's 462-value rows 431 ms by rows, 523 ms by gathers; its
// (10;7,3) step's 55-value rows 227 ms by rows.
bool runs_seq_for(const pg_ctx* c, int wa, int used_per_mille)
{
    const int rv = runs_values_per_lane(wa);
    if (c->color_runs == 0 || rv == 0) return false;
    return c->color_runs > 0 || (wa >= kRunsMinWidth && wa <= c->seq_max_width && used_per_mille >= 500);
}

std::string runs_name(size_t tsz, int mode, int main_mode)
{
    char buf[64];
    snprintf(buf, sizeof buf, "k_runs<%s,%s,%d>", tsz == 4 ? "float" : "double",
             mode == RUNS_T ? "t" : mode == RUNS_G ? "g" : mode == RUNS_SEQ ? "seq" : mode == RUNS_TG ? "tg" : "fwd",
             main_mode);
    return buf;
}

// A gathering step's hub launch (hubs(stream)) runs on the hub stream beside
// the other vertices' launches (rest(stream)) on the context's stream, which
// then waits for it: the hub rows' long neighbour lists overlap the rest of
// the step instead of leaving the device to a few wavefronts.  Sequential
// with option hub_concurrent 0, while timing (per-launch events on one
// stream), without a hub stream, and inside the branches of a mixed-batch
// capture (a third stream joined from several capturing branches took the
// process down in the runtime there; those branches already run side by
// side).
template <typename H, typename R>
int fork_hubs(pg_ctx* c, bool have_hubs, H hubs, R rest)
{
    if (!have_hubs) return rest(c->stream);
    if (!c->hub_concurrent || c->timing || !c->hub_stream || c->in_branches) {
        const int rc = hubs(c->stream);
        return rc ? rc : rest(c->stream);
    }
    HIP_TRY(c, hipEventRecord(c->ev_hub_fork, c->stream));
    HIP_TRY(c, hipStreamWaitEvent(c->hub_stream, c->ev_hub_fork, 0));
    int rc = hubs(c->hub_stream);
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(c->ev_hub_join, c->hub_stream));
    rc = rest(c->stream);
    if (rc) return rc;
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_hub_join, 0));
    return PG_OK;
}

// launch_gathering with colour runs: the hub chunks' partial rows, all other
// vertices in one launch, then the hub rows from their partials
template <typename T>
int launch_gathering_runs(pg_ctx* c, int rmode, int main_mode, SubArgs<T> a, T* part, int y_cols, StepBytes bytes)
{
    const hipStream_t st = c->stream;
    const int cus = c->num_cus;
    const double Z = (double)c->nnz, N = (double)c->n_act;
    auto share = [&](double ents, double verts) {
        double b = 0;
        if (Z + N > 0) b += bytes.csc * (ents + verts) / (Z + N);
        if (Z > 0) b += bytes.per_entry * ents / Z;
        if (N > 0) b += bytes.per_vertex * verts / N;
        return b;
    };
    auto hubs = [&](hipStream_t hs) -> int {
        SubArgs<T> h = a;
        h.out = part;
        h.ldo = y_cols;
        h.out_mode = OUT_PARTIAL;
        h.items = c->d_hub_items.as<int4>();
        h.n_items = c->n_hub_items;
        h.item_base = 0;
        record(c, runs_name(sizeof(T), rmode, main_mode), share(c->h_rowptr[c->n_hub], 0), true);
        HIP_TRY(c, launch_runs<T>(rmode, main_mode, h, cus, c->blocks_per_cu, hs));
        record(c, "", 0, false);
        return PG_OK;
    };
    auto rest = [&](hipStream_t rs) -> int {
        if (c->n_hub >= c->n_act) return PG_OK;
        SubArgs<T> r = a;
        r.items = nullptr;
        r.item_base = c->n_hub;
        r.n_items = c->n_act - c->n_hub;
        record(c, runs_name(sizeof(T), rmode, main_mode),
               share((double)c->nnz - c->h_rowptr[c->n_hub], c->n_act - c->n_hub), true);
        HIP_TRY(c, launch_runs<T>(rmode, main_mode, r, cus, c->blocks_per_cu, rs));
        record(c, "", 0, false);
        return PG_OK;
    };
    const int frc = fork_hubs(c, c->n_hub > 0, hubs, rest);
    if (frc) return frc;
    if (c->n_hub > 0) {
        SubArgs<T> f = a;
        f.aux = part;
        f.lda = y_cols;
        f.items = nullptr;
        f.item_base = 0;
        f.n_items = c->n_hub;
        f.ng = 1;
        f.Wa = a.Wy;
        const int team = team_for(std::max(y_cols, a.B * a.Wo));
        record(c, kname<T>(team, AUX_PARTIALS, main_mode), share(0, c->n_hub), true);
        HIP_TRY(c, launch_sub<T>(team, AUX_PARTIALS, main_mode, kRegCols, f, cus, c->blocks_per_cu, st));
        record(c, "", 0, false);
    }
    return PG_OK;
}

template <typename T>
int launch_gathering_exact(pg_ctx* c, int aux, int main_mode, SubArgs<T> a, T* part, int y_cols, StepBytes bytes,
                           bool rows);

template <typename T>
int launch_gathering(pg_ctx* c, int aux, int main_mode, SubArgs<T> a, T* part, int y_cols, StepBytes bytes,
                     int used_per_mille = 0)
{
    {
        if (a.tr) {  // the fused transposed gather and gradient (fused_grad_step)
            if (runs_values_per_lane(a.Wa) == 0) return set_err(c, PG_ERR_STATE, "fused gradient gather too wide");
            return launch_gathering_runs<T>(c, RUNS_TG, main_mode, a, part, y_cols, bytes);
        }
        const int rmode = runs_mode_for(c, aux, a.Wa, used_per_mille);
        if (rmode >= 0) return launch_gathering_runs<T>(c, rmode, main_mode, a, part, y_cols, bytes);
    }
    if (c->exact_order && (aux == AUX_GATHER || aux == AUX_STAGED))
        return launch_gathering_exact<T>(c, AUX_GATHER, main_mode, a, part, y_cols, bytes,
                                         a.xidx_g && runs_seq_for(c, a.Wa, used_per_mille));
    const int rc = rc_for(c, aux, a.B, a.Wy);
    const int cl = cl_of(aux, a.B, a.k, a.Wy, a.Wa, rc);
    // a step whose Y is kept for the adjoint gathers unstaged (the staged
    // kernel carries no Y store, which keeps its registers at 3 waves per SIMD)
    a.segpad = rc == kRegCols && !a.ykeep ? stage_pad(c, aux, a.gidx != nullptr, a.B, a.Wy, a.Wa, cl, sizeof(T)) : 0;
    if (a.segpad > 0) aux = AUX_STAGED;
    const hipStream_t st = c->stream;
    const int cus = c->num_cus;
    const double Z = (double)c->nnz, N = (double)c->n_act;
    auto share = [&](double ents, double verts) {
        double b = 0;
        if (Z + N > 0) b += bytes.csc * (ents + verts) / (Z + N);
        if (Z > 0) b += bytes.per_entry * ents / Z;
        if (N > 0) b += bytes.per_vertex * verts / N;
        return b;
    };
    auto hubs = [&](hipStream_t hs) -> int {
        SubArgs<T> h = a;
        h.out = part;
        h.ldo = y_cols;
        h.out_mode = OUT_PARTIAL;
        h.items = c->d_hub_items.as<int4>();
        h.n_items = c->n_hub_items;
        h.item_base = 0;
        h.ng = 64 / cl;
        record(c, kname<T>(64, aux, main_mode, rc), share(c->h_rowptr[c->n_hub], 0), true);
        HIP_TRY(c, launch_sub<T>(64, aux, main_mode, rc, h, cus, c->blocks_per_cu, hs));
        record(c, "", 0, false);
        return PG_OK;
    };
    auto rest = [&](hipStream_t rs) -> int {
        struct Bucket { int32_t lo, hi; int lanes; };
        const Bucket buckets[3] = {{c->n_hub, c->hi_end, 64}, {c->hi_end, c->mid_end, 16}, {c->mid_end, c->n_act, 8}};
        for (const Bucket& bk : buckets) {
            if (bk.hi <= bk.lo) continue;
            const int team = std::max(bk.lanes, cl);
            SubArgs<T> r = a;
            r.items = nullptr;
            r.item_base = bk.lo;
            r.n_items = bk.hi - bk.lo;
            r.ng = team / cl;
            record(c, kname<T>(team, aux, main_mode, rc),
                   share((double)c->h_rowptr[bk.hi] - c->h_rowptr[bk.lo], bk.hi - bk.lo), true);
            HIP_TRY(c, launch_sub<T>(team, aux, main_mode, rc, r, cus, c->blocks_per_cu, rs));
            record(c, "", 0, false);
        }
        return PG_OK;
    };
    const int frc = fork_hubs(c, c->n_hub > 0, hubs, rest);
    if (frc) return frc;
    if (c->n_hub > 0) {
        SubArgs<T> f = a;
        f.aux = part;
        f.lda = y_cols;
        f.items = nullptr;
        f.item_base = 0;
        f.n_items = c->n_hub;
        f.ng = 1;
        const int team = team_for(std::max(y_cols, a.B * a.Wo));
        record(c, kname<T>(team, AUX_PARTIALS, main_mode), share(0, c->n_hub), true);
        HIP_TRY(c, launch_sub<T>(team, AUX_PARTIALS, main_mode, kRegCols, f, cus, c->blocks_per_cu, st));
        record(c, "", 0, false);
    }
    return PG_OK;
}

// The forward neighbour sums in the reference's order (CSCGraph.hpp:254-307:
// every Y value adds its neighbours' values one after another, in ascending
// column order, starting from zero), so the FP32 sums are the reference's,
// bit for bit:
//  - hubs: one partial row per hub, each value one running sum over the whole
//    neighbour list, by k_hub_seq (a wavefront per 64 values with 32-56 row
//    loads in flight per lane);
//  - the other vertices: a team with the whole neighbour list, one
//    neighbour group, by whole rows (RUNS_SEQ, when `rows`) or by gathers;
//  - the finishing launch adds each hub's single partial row to zero and
//    combines it with the main table.
template <typename T>
int launch_gathering_exact(pg_ctx* c, int aux, int main_mode, SubArgs<T> a, T* part, int y_cols, StepBytes bytes,
                           bool rows)
{
    const int rc = rc_for(c, aux, a.B, a.Wy);
    const int cl = std::max(8, cl_of(aux, a.B, a.k, a.Wy, a.Wa, rc));
    a.segpad = rc == kRegCols && !a.ykeep ? stage_pad(c, aux, a.gidx != nullptr, a.B, a.Wy, a.Wa, cl, sizeof(T)) : 0;
    const int aux_b = a.segpad > 0 ? AUX_STAGED : aux;
    const hipStream_t st = c->stream;
    const int cus = c->num_cus;
    const double Z = (double)c->nnz, N = (double)c->n_act;
    auto share = [&](double ents, double verts) {
        double b = 0;
        if (Z + N > 0) b += bytes.csc * (ents + verts) / (Z + N);
        if (Z > 0) b += bytes.per_entry * ents / Z;
        if (N > 0) b += bytes.per_vertex * verts / N;
        return b;
    };
    a.ng = 1;
    auto hubs = [&](hipStream_t hs) -> int {
        SubArgs<T> h = a;
        h.segpad = 0;
        h.out = part;
        h.ldo = y_cols;
        h.out_mode = OUT_PARTIAL;
        h.items = nullptr;
        h.item_base = 0;
        h.n_items = c->n_hub;
        record(c, sizeof(T) == 4 ? "k_hub_seq<float>" : "k_hub_seq<double>", share(c->h_rowptr[c->n_hub], 0), true);
        HIP_TRY(c, launch_hub_seq<T>(h, cus, c->blocks_per_cu, hs));
        record(c, "", 0, false);
        return PG_OK;
    };
    auto rest = [&](hipStream_t rs) -> int {
        if (c->n_hub >= c->n_act) return PG_OK;
        SubArgs<T> r = a;
        r.items = nullptr;
        r.item_base = c->n_hub;
        r.n_items = c->n_act - c->n_hub;
        const double sh = share((double)c->nnz - c->h_rowptr[c->n_hub], c->n_act - c->n_hub);
        if (rows) {
            record(c, runs_name(sizeof(T), RUNS_SEQ, main_mode), sh, true);
            HIP_TRY(c, launch_runs<T>(RUNS_SEQ, main_mode, r, cus, c->blocks_per_cu, rs));
        } else {
            record(c, kname<T>(cl, aux_b, main_mode, rc), sh, true);
            HIP_TRY(c, launch_sub<T>(cl, aux_b, main_mode, rc, r, cus, c->blocks_per_cu, rs));
        }
        record(c, "", 0, false);
        return PG_OK;
    };
    const int frc = fork_hubs(c, c->n_hub > 0, hubs, rest);
    if (frc) return frc;
    if (c->n_hub > 0) {
        SubArgs<T> f = a;
        f.aux = part;
        f.lda = y_cols;
        f.items = nullptr;
        f.item_base = 0;
        f.n_items = c->n_hub;
        f.slot_begin = c->d_hub_iota.as<int32_t>();
        f.segpad = 0;
        const int team = team_for(std::max(y_cols, a.B * a.Wo));
        record(c, kname<T>(team, AUX_PARTIALS, main_mode), share(0, c->n_hub), true);
        HIP_TRY(c, launch_sub<T>(team, AUX_PARTIALS, main_mode, kRegCols, f, cus, c->blocks_per_cu, st));
        record(c, "", 0, false);
    }
    return PG_OK;
}

// Device pointers of one pass's buffers (colors, neighbour colors, the
// histogram H, hub partial rows, root values, results, parameters).
template <typename T>
struct PassBufs {
    const PassParams* pp;
    uint8_t* colors;
    uint8_t* ecol;
    T* H;
    T* part;
    double* rootv;
    double* rawd;
    double* tmp;
};

// the device memory a pass's buffers are placed in: the context's arena, or
// the static scratch during pg_count_noalloc
char* arena_base(const pg_ctx* c) { return c->arena_override ? c->arena_override : c->d_arena.as<char>(); }

template <typename T>
PassBufs<T> pass_bufs(pg_ctx* c, const MemPlan& mp)
{
    char* base = arena_base(c);
    return {(const PassParams*)(base + mp.off_params), (uint8_t*)(base + mp.off_colors),
            (uint8_t*)(base + mp.off_ecol),          (T*)(base + mp.off_H),
            (T*)(base + mp.off_part),                (double*)(base + mp.off_root),
            (double*)(base + mp.off_rawdev),         (double*)(base + mp.off_tmp)};
}

// The arguments every launch of a pass shares.
template <typename T>
SubArgs<T> pass_args(const pg_ctx* c, const pg_plan* p, int B, const PassBufs<T>& pb)
{
    SubArgs<T> a{};
    a.rowptr = c->d_rowptr.as<int32_t>();
    a.col = c->d_col.as<int32_t>();
    if (c->weighted) a.w = sizeof(T) == 4 ? (const T*)c->d_w32.p : (const T*)c->d_w64.p;
    a.colors = pb.colors;
    a.ecol = pb.ecol;
    a.B = B;
    a.k = p->k;
    a.K = p->K;
    a.ng = 1;
    a.rank_tab = p->rank_tab.as<int16_t>();
    a.slot_begin = c->d_slot_begin.as<int32_t>();
    a.scale = 1.0e-12f;
    return a;
}

// survey byte model: CSC streamed once per gathering launch, full-width tables
double csc_bytes(const pg_ctx* c) { return 4.0 * (c->n_act + 1) + 4.0 * c->nnz * (c->weighted ? 2 : 1); }

// One forward step: subtemplate d of B colorings from its children's tables
// (nullptr for a leaf child) into out_tab or, at the root, into the
// per-vertex root values and the pass results.
// Tags the timing records of one subtemplate's launches with its index.
struct SubTag {
    pg_ctx* c;
    SubTag(pg_ctx* ctx, int s) : c(ctx) { c->cur_sub = s; }
    ~SubTag() { c->cur_sub = -1; }
};

template <typename T>
int enqueue_step(pg_ctx* c, const pg_plan* p, const DevSub& d, int B, const PassBufs<T>& pb, const T* main_tab,
                 const T* aux_tab, T* out_tab, T* ykeep)
{
    const SubTag tag(c, d.s);
    const hipStream_t st = c->stream;
    const double na = (double)c->n_act;
    SubArgs<T> a = pass_args<T>(c, p, B, pb);
    if (ykeep && d.aa > 1) {
        a.ykeep = ykeep;
        a.ldy = B * d.Wy;
    }
    a.Wy = d.Wy;
    a.Wm = d.Wm;
    a.Wo = d.Wo;
    a.P = d.P;
    a.split = d.split.as<uint32_t>();
    a.split_lds = step_split_lds(c, p, d, B, sizeof(T));
    a.ymask = d.ymask.as<uint16_t>();
    const int main_mode = d.mm == 1 ? MAIN_LEAF : MAIN_TABLE;
    if (d.mm > 1) {
        a.main = main_tab;
        a.ldm = B * d.Wm;
    }
    if (d.root) {
        a.out_mode = OUT_ROOT;
        a.root_vals = pb.rootv;
        a.n_root = c->n_act;
    } else {
        a.out = out_tab;
        a.ldo = B * d.Wo;
        a.out_mode = (d.scaled && sizeof(T) == 4) ? OUT_EMA_SCALED : OUT_EMA;
    }
    const double tab_bytes = 4.0 * na * B * ((d.aa > 1 ? d.Ca : 0) + (d.mm > 1 ? d.Cm : 0) + (d.root ? 0 : d.Cs));
    StepBytes sb;  // aux rows gathered by entries; main and output rows per vertex
    sb.csc = csc_bytes(c);
    sb.per_entry = 4.0 * na * B * (d.aa > 1 ? d.Ca : 0);
    sb.per_vertex = 4.0 * na * B * ((d.mm > 1 ? d.Cm : 0) + (d.root ? 0 : d.Cs));
    if (d.aa == 1) {
        const int team = team_for(std::max(B * d.Wy, B * d.Wo));
        a.aux = pb.H;
        a.lda = B * p->k;
        a.Wa = p->k;
        a.n_items = c->n_act;
        a.item_base = 0;
        record(c, kname<T>(team, AUX_READ, main_mode), tab_bytes, true);
        HIP_TRY(c, launch_sub<T>(team, AUX_READ, main_mode, kRegCols, a, c->num_cus, c->blocks_per_cu, st));
        record(c, "", 0, false);
    } else {
        a.aux = aux_tab;
        a.Wa = d.Wa;
        a.lda = B * d.Wa;
        a.xmask = d.amask.as<uint16_t>();
        a.gidx = d.gidx.as<int16_t>();
        a.Wy4 = d.Wy4;
        a.xidx = d.xidx.as<int16_t>();
        a.xidx_g = d.xidx_g.as<int16_t>();
        a.Wa4 = d.Wa4;
        int rc = launch_gathering<T>(c, aux_mode_for(c, p, d), main_mode, a, pb.part, B * d.Wy, sb,
                                     runs_used_per_mille(p, d, AUX_GATHER));
        if (rc) return rc;
    }
    if (d.root) HIP_TRY(c, launch_root_reduce(pb.rootv, c->n_act, B, pb.tmp, pb.rawd, st));
    return PG_OK;
}

// Enqueues one pass of B colorings through the whole DP on c->stream.  The
// pass's seed and first coloring are read on the device from mp.off_params,
// so the same launches (captured once as a graph) serve every pass of this
// shape.  Results land in mp.off_rawdev ([B] doubles, before the 1e12 factor).
template <typename T>
int enqueue_pass(pg_ctx* c, const pg_plan* p, int B, const uint8_t* d_colors_in, const MemPlan& mp)
{
    char* base = arena_base(c);
    const PassBufs<T> pb = pass_bufs<T>(c, mp);
    const hipStream_t st = c->stream;
    if (d_colors_in)
        HIP_TRY(c, launch_colors_gather(pb.colors, d_colors_in, c->d_perm.as<int32_t>(), c->n_act, B, c->n, pb.pp, st));
    else
        HIP_TRY(c, launch_colors_rng(pb.colors, c->d_perm.as<int32_t>(), c->n_act, B, pb.pp, p->k, st));
    HIP_TRY(c, launch_edge_colors(pb.ecol, c->d_col.as<int32_t>(), pb.colors, c->nnz, B, st));

    // ---- neighbour-color histogram H = A * leaf (shared by every leaf aux)
    bool need_H = false;
    for (auto& d : p->subs) need_H |= d.get()->aa == 1;
    if (need_H) {
        SubArgs<T> a = pass_args<T>(c, p, B, pb);
        a.Wy = p->k;
        a.Wo = p->k;
        a.out = pb.H;
        a.ldo = B * p->k;
        a.out_mode = OUT_Y;
        int rc = launch_gathering<T>(c, AUX_COLORS, MAIN_LEAF, a, pb.part, B * p->k,
                                     StepBytes{csc_bytes(c), 0.0, (double)c->n_act * B});
        if (rc) return rc;
    }

    // ---- subtemplates, children before parents
    auto tab = [&](int s) { return (T*)(base + mp.off_tab[s]); };
    auto ykeep = [&](int s) {
        return (size_t)s < mp.off_ykeep.size() && mp.off_ykeep[s] ? (T*)(base + mp.off_ykeep[s]) : nullptr;
    };
    for (auto& dp : p->subs) {
        const DevSub& d = *dp;
        int rc = enqueue_step<T>(c, p, d, B, pb, d.mm > 1 ? tab(d.mi) : nullptr, d.aa > 1 ? tab(d.ai) : nullptr,
                                 d.root ? nullptr : tab(d.s), ykeep(d.s));
        if (rc) return rc;
    }
    return PG_OK;
}

// CountMat.cpp:815-824: the root sum is multiplied by 1e12 unconditionally
// (it undoes the float(1e-12) of the scaled subtemplate; a template whose
// only non-leaf is the root was never scaled, and the reference still
// multiplies).  FP64 tables are unscaled: only that quirk is reproduced.
double finish_raw(const pg_plan* p, bool fp32, uint32_t flags, double r)
{
    const bool two_vertex = p->nonleaf == 1;
    const bool fix = (flags & PG_FLAG_FIX_2VERTEX) != 0;
    if (fp32) return (two_vertex && fix) ? r : r * 1.0e12;
    return (two_vertex && !fix) ? r * 1.0e12 : r;
}

int choose_tile(pg_ctx* c, const pg_plan* p, int ncol, size_t tsz, int* tile, MemPlan* mp)
{
    size_t budget;
    if (c->mem_budget_mb > 0) {
        budget = (size_t)c->mem_budget_mb << 20;
    } else {
        size_t fr = 0, tot = 0;
        HIP_TRY(c, hipMemGetInfo(&fr, &tot));
        budget = (size_t)(0.9 * (double)(fr + c->d_arena.bytes));
    }
    const int cap = c->batch_tile > 0 ? (int)std::min<int64_t>(c->batch_tile, ncol) : std::min(ncol, 64);
    for (int B = cap; B >= 1; --B) {
        size_t lds = 0;
        for (auto& d : p->subs) lds = std::max(lds, sub_lds_max(c, p, *d, B, tsz));
        if (lds > kLdsLimit) {
            if (c->batch_tile > 0) break;
            continue;
        }
        MemPlan m = plan_memory(c, p, B, tsz);
        if (m.total > budget) {
            if (c->batch_tile > 0) break;
            continue;
        }
        *tile = B;
        *mp = m;
        return PG_OK;
    }
    return set_err(c, PG_ERR_NOMEM, "no batch tile fits (LDS %zu KB per block, device budget %.1f GB)",
                   kLdsLimit >> 10, budget / 1e9);
}

void finish_timing(pg_ctx* c)
{
    c->dom_name.clear();
    c->dom_ms = c->dom_bytes = 0;
    c->dom_launches = 0;
    for (auto& kv : c->agg) {
        if (kv.second.ms > c->dom_ms) {
            c->dom_name = kv.first;
            c->dom_ms = kv.second.ms;
            c->dom_bytes = kv.second.bytes;
            c->dom_launches = kv.second.launches;
        }
    }
}

void drop_graphs(pg_ctx* c, const pg_plan* only)
{
    for (auto it = c->graphs.begin(); it != c->graphs.end();) {
        if (only && it->first.plan != only) {
            ++it;
            continue;
        }
        (void)hipGraphExecDestroy(it->second.exec);
        it = c->graphs.erase(it);
    }
    for (auto it = c->multi_graphs.begin(); it != c->multi_graphs.end();) {
        const auto& k = it->first;
        if (only && std::find(k.begin(), k.end(), (int64_t)(intptr_t)only) == k.end()) {
            ++it;
            continue;
        }
        (void)hipGraphExecDestroy(it->second.exec);
        it = c->multi_graphs.erase(it);
    }
}

// One pass: through its cached graph when graphs are on and timing is off
// (capturing it on first use), else eagerly.  Writes B raw counts (before
// finish_raw).
template <typename T>
int run_pass(pg_ctx* c, const pg_plan* p, int B, int64_t first, const uint8_t* d_in, uint64_t seed,
             const MemPlan& mp, double* raw)
{
    char* base = c->d_arena.as<char>();
    const PassParams hp{seed, first};
    // pageable source: the copy completes before the call returns
    HIP_TRY(c, hipMemcpyAsync(base + mp.off_params, &hp, sizeof hp, hipMemcpyHostToDevice, c->stream));
    if (!c->use_graphs || c->timing) {
        std::vector<LaunchRec> recs;
        c->cur_recs = &recs;
        int rc = enqueue_pass<T>(c, p, B, d_in, mp);
        c->cur_recs = nullptr;
        if (rc) {
            destroy_recs(recs);
            return rc;
        }
        HIP_TRY(c, hipMemcpyAsync(raw, base + mp.off_rawdev, sizeof(double) * B, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        accumulate_timing(c, recs);
        destroy_recs(recs);
        return PG_OK;
    }
    const GraphKey key{p, B, (int)sizeof(T), d_in, c->d_arena.p, c->generation};
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        PassGraph g;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_pass<T>(c, p, B, d_in, mp);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc) return rc;
            return set_err(c, PG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        }
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return set_err(c, PG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        it = c->graphs.emplace(key, std::move(g)).first;
    }
    HIP_TRY(c, hipGraphLaunch(it->second.exec, c->stream));
    HIP_TRY(c, hipMemcpyAsync(raw, base + mp.off_rawdev, sizeof(double) * B, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return PG_OK;
}

// ---------------------------------------------------------------- autotuner
//
// Option "autotune" = 1: before the first pass of a (plan, batch tile,
// precision) the library times one pass per candidate setting of the kernel
// variant and degree-bucket knobs (coordinate descent over the knobs below,
// a second sweep when the first changed one, each candidate kept only if at
// least 2% faster), applies the
// fastest and remembers it for that key.  Tuning passes draw their colorings
// on the device from a fixed seed and discard the counts.

struct Knob {
    const char* name;
    int64_t pg_ctx::*field;
    std::vector<int64_t> values;
};

const std::vector<Knob>& tune_knobs()
{
    static const std::vector<Knob> knobs = {
        {"reg_cols", &pg_ctx::reg_cols, {0, 1}},
        {"stage_rows", &pg_ctx::stage_rows, {0, 1}},
        {"split_lds", &pg_ctx::split_lds, {1, 0}},
        {"color_runs", &pg_ctx::color_runs, {-1, 0}},
        {"aux_rows", &pg_ctx::aux_rows, {0, -1}},
        {"bucket_high", &pg_ctx::bucket_high, {16, 32, 64, 128, 256, 512}},
        {"bucket_mid", &pg_ctx::bucket_mid, {4, 8, 16, 32, 64, 128}},
        {"hub_degree", &pg_ctx::hub_degree, {128, 256, 512, 1024}},
        {"hub_chunk", &pg_ctx::hub_chunk, {256, 1024, 4096}},
    };
    return knobs;
}

std::vector<int64_t> current_settings(const pg_ctx* c)
{
    std::vector<int64_t> v;
    for (const Knob& k : tune_knobs()) v.push_back(c->*k.field);
    return v;
}

void apply_settings(pg_ctx* c, const std::vector<int64_t>& v)
{
    const auto& knobs = tune_knobs();
    for (size_t i = 0; i < knobs.size(); ++i) {
        if (c->*knobs[i].field != v[i]) drop_graphs(c);
        c->*knobs[i].field = v[i];
    }
}

// Device time of one eager pass of B device-drawn colorings with the current
// settings (best of two after a warmup; a first pass over 50 ms is taken as
// is), or +inf when they do not fit the LDS limit or the arena as already
// sized.
template <typename T>
int time_pass(pg_ctx* c, const pg_plan* p, int B, double* ms)
{
    *ms = INFINITY;
    int rc = build_buckets(c);
    if (rc) return rc;
    size_t lds = 0;
    for (auto& d : p->subs) lds = std::max(lds, sub_lds_max(c, p, *d, B, sizeof(T)));
    const MemPlan mp = plan_memory(c, p, B, sizeof(T));
    if (lds > kLdsLimit || mp.total > c->d_arena.bytes) return PG_OK;
    std::vector<double> raw(B);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    for (int rep = 0; rep < 3 && rc == PG_OK; ++rep) {
        (void)hipEventRecord(e0, c->stream);
        rc = run_pass<T>(c, p, B, 0, nullptr, 0x7475ULL, mp, raw.data());
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float t = 0;
        (void)hipEventElapsedTime(&t, e0, e1);
        if (rep > 0 || t > 50.0f) *ms = std::min(*ms, (double)t);
        if (rep == 0 && t > 50.0f) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Coordinate descent over the knobs at the largest tile B that fits, then
// the column tile itself: 3B/4, B/2 and B/4 colorings per pass are timed with the
// chosen knobs and kept when a coloring costs 2% less (wide templates can
// run better with fewer, narrower passes).  A tile fixed by option
// batch_tile is kept.  *tile receives the tile to run.
template <typename T>
int autotune(pg_ctx* c, const pg_plan* p, int B, int* tile)
{
    *tile = B;
    const auto key = std::make_tuple(p, B, (int)sizeof(T));
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) {
        apply_settings(c, it->second);
        *tile = (int)it->second.back();
        return build_buckets(c);
    }
    const int64_t timing = c->timing;
    const bool graphs = c->use_graphs;
    c->timing = 0;
    c->use_graphs = false;
    std::vector<int64_t> best = current_settings(c);
    double best_ms = INFINITY;
    int rc = time_pass<T>(c, p, B, &best_ms);
    const auto& knobs = tune_knobs();
    // a second sweep when the first changed a knob: the variant knobs come
    // first, and their best value can depend on the bucket knobs after them
    // (R-MAT scale 20, 7-vertex tree, one coloring per pass: staged rows pay
    // only with 64-lane teams from degree 64 and hubs from degree 1024)
    for (int sweep = 0; sweep < 2 && rc == PG_OK; ++sweep) {
        bool changed = false;
        for (size_t i = 0; i < knobs.size() && rc == PG_OK; ++i) {
            for (int64_t v : knobs[i].values) {
                if (v == best[i]) continue;
                std::vector<int64_t> cand = best;
                cand[i] = v;
                apply_settings(c, cand);
                double ms = INFINITY;
                rc = time_pass<T>(c, p, B, &ms);
                if (rc) break;
                if (ms < 0.98 * best_ms) {
                    best = cand;
                    best_ms = ms;
                    changed = true;
                }
            }
        }
        if (!changed) break;
    }
    int best_tile = B;
    if (rc == PG_OK && c->batch_tile <= 0 && best_ms < INFINITY) {
        apply_settings(c, best);
        double best_per = best_ms / B;
        int tried = B;
        for (const int t : {B * 3 / 4, B / 2, B / 4}) {
            if (t < 1 || t >= tried) continue;  // distinct, decreasing candidates
            tried = t;
            double ms = INFINITY;
            rc = time_pass<T>(c, p, t, &ms);
            if (rc) break;
            if (ms / t < 0.98 * best_per) {
                best_per = ms / t;
                best_tile = t;
            }
        }
    }
    c->timing = timing;
    c->use_graphs = graphs;
    apply_settings(c, best);
    if (rc) return rc;
    std::vector<int64_t> rec = best;
    rec.push_back(best_tile);
    c->tuned[key] = rec;
    *tile = best_tile;
    return build_buckets(c);
}

template <typename T>
int count_impl(pg_ctx* c, const pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed, uint32_t flags,
               double* raw_out)
{
    int rc = build_buckets(c);
    if (rc) return rc;
    int tile = 1;
    MemPlan mp;
    rc = choose_tile(c, p, ncol, sizeof(T), &tile, &mp);
    if (rc) return rc;
    if (mp.total > c->d_arena.bytes) {
        drop_graphs(c);  // graphs hold arena addresses
        HIP_TRY(c, c->d_arena.alloc(mp.total));
    }
    c->last_tile = tile;
    c->last_recompute = 0;
    if (c->autotune) {
        int tuned_tile = tile;
        rc = autotune<T>(c, p, tile, &tuned_tile);
        if (rc) return rc;
        // the chosen settings fit this tile (tuning skipped those that did not)
        tile = tuned_tile;
        c->last_tile = tile;
        mp = plan_memory(c, p, tile, sizeof(T));
    }
    const uint8_t* d_in = nullptr;
    if (colors) {
        if ((size_t)ncol * c->n > c->d_colors_in.bytes) {
            drop_graphs(c);
            HIP_TRY(c, c->d_colors_in.alloc((size_t)ncol * c->n));
        }
        HIP_TRY(c, hipMemcpyAsync(c->d_colors_in.p, colors, (size_t)ncol * c->n, hipMemcpyHostToDevice, c->stream));
        d_in = c->d_colors_in.as<uint8_t>();
    }
    c->agg.clear();
    c->agg_sub.clear();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(c, hipEventCreate(&t0));
    HIP_TRY(c, hipEventCreate(&t1));
    HIP_TRY(c, hipEventRecord(t0, c->stream));
    std::vector<double> raw;
    for (int first = 0; first < ncol; first += tile) {
        const int B = std::min(tile, ncol - first);
        const MemPlan m = B == tile ? mp : plan_memory(c, p, B, sizeof(T));
        raw.assign(B, 0.0);
        rc = run_pass<T>(c, p, B, first, d_in, seed, m, raw.data());
        if (rc) break;
        for (int b = 0; b < B; ++b) raw_out[first + b] = finish_raw(p, sizeof(T) == 4, flags, raw[b]);
    }
    (void)hipEventRecord(t1, c->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    c->last_total_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    finish_timing(c);
    return rc;
}

// ------------------------------------------------------------ mixed batch
//
// pg_count_multi with graphs on: each plan gets its tile (and, with option
// autotune, its tuned kernel variants and tile), its own region of the arena,
// and a branch of ONE captured graph per round of passes.  The branches are
// captured on side streams forked from and joined back into the context's
// stream, so a single graph launch runs a pass of every template and the
// device may overlap them.  The branches share the degree-bucket work lists,
// so every plan must use the same bucket settings; otherwise (and with timing
// on, or graphs off) the plans run one after another through count_impl.

// the knob values without dropping captured graphs (the bucket knobs, which
// size the shared work lists, are equal for every plan of a mixed graph)
void put_settings(pg_ctx* c, const std::vector<int64_t>& v)
{
    const auto& knobs = tune_knobs();
    for (size_t i = 0; i < knobs.size(); ++i) c->*knobs[i].field = v[i];
}

bool same_buckets(const std::vector<int64_t>& a, const std::vector<int64_t>& b)
{
    const auto& knobs = tune_knobs();
    for (size_t i = 0; i < knobs.size(); ++i) {
        const std::string name = knobs[i].name;
        if ((name == "bucket_high" || name == "bucket_mid" || name == "hub_degree") && a[i] != b[i]) return false;
    }
    return true;
}

MemPlan shifted(MemPlan m, size_t d)
{
    for (size_t* o : {&m.off_H, &m.off_part, &m.off_root, &m.off_colors, &m.off_ecol, &m.off_rawdev, &m.off_tmp,
                      &m.off_params})
        *o += d;
    for (auto& o : m.off_tab) o += d;
    m.total += d;
    return m;
}

int ensure_branches(pg_ctx* c, int n)
{
    if (!c->ev_fork) HIP_TRY(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    while ((int)c->side.size() < n) {
        hipStream_t s = nullptr;
        HIP_TRY(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        c->side.push_back(s);
        hipEvent_t e = nullptr;
        HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_join.push_back(e);
    }
    return PG_OK;
}

// *done = false: not handled here (the caller runs the plans one by one)
template <typename T>
int count_multi_impl(pg_ctx* c, int np, const pg_plan* const* plans, const int32_t* ncol, const uint64_t* seeds,
                     uint32_t flags, double* raw_out, bool* done)
{
    *done = false;
    if (!c->use_graphs || c->timing || np < 2) return PG_OK;
    int rc = build_buckets(c);
    if (rc) return rc;
    const size_t tsz = sizeof(T);
    const std::vector<int64_t> entry = current_settings(c);
    std::vector<int> tile(np);
    std::vector<std::vector<int64_t>> knob(np, entry);
    for (int i = 0; i < np; ++i) {
        MemPlan mp;
        rc = choose_tile(c, plans[i], ncol[i], tsz, &tile[i], &mp);
        if (rc) return rc;
        if (!c->autotune) continue;
        const auto key = std::make_tuple(plans[i], tile[i], (int)tsz);
        auto it = c->tuned.find(key);
        if (it == c->tuned.end()) {
            if (mp.total > c->d_arena.bytes) {
                drop_graphs(c);
                HIP_TRY(c, c->d_arena.alloc(mp.total));
            }
            int t = tile[i];
            rc = autotune<T>(c, plans[i], tile[i], &t);
            if (rc) return rc;
            it = c->tuned.find(key);
        }
        knob[i].assign(it->second.begin(), it->second.begin() + (long)tune_knobs().size());
        tile[i] = (int)it->second.back();
    }
    if (c->autotune) {  // without it every plan runs the caller's settings as they are
        for (int i = 1; i < np; ++i)
            if (!same_buckets(knob[i], knob[0])) {
                apply_settings(c, entry);
                return build_buckets(c);
            }
        apply_settings(c, knob[0]);  // the shared bucket settings
        rc = build_buckets(c);
        if (rc) return rc;
    }
    // arena regions, one per plan
    size_t budget;
    if (c->mem_budget_mb > 0) {
        budget = (size_t)c->mem_budget_mb << 20;
    } else {
        size_t fr = 0, tot = 0;
        HIP_TRY(c, hipMemGetInfo(&fr, &tot));
        budget = (size_t)(0.9 * (double)(fr + c->d_arena.bytes));
    }
    std::vector<MemPlan> mp(np);
    std::vector<size_t> base(np);
    size_t total = 0;
    for (int i = 0; i < np; ++i) {
        mp[i] = plan_memory(c, plans[i], tile[i], tsz);
        base[i] = total;
        total += align_up(mp[i].total, 256);
    }
    if (total > budget) return PG_OK;  // the plans do not fit side by side: one after another
    if (total > c->d_arena.bytes) {
        drop_graphs(c);
        HIP_TRY(c, c->d_arena.alloc(total));
    }
    rc = ensure_branches(c, np);
    if (rc) return rc;
    *done = true;
    char* arena = c->d_arena.as<char>();
    std::vector<int64_t> first(np, 0), out_off(np, 0);
    for (int i = 1; i < np; ++i) out_off[i] = out_off[i - 1] + ncol[i - 1];
    c->agg.clear();
    c->agg_sub.clear();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(c, hipEventCreate(&t0));
    HIP_TRY(c, hipEventCreate(&t1));
    HIP_TRY(c, hipEventRecord(t0, c->stream));
    std::vector<double> raw;
    for (;;) {
        std::vector<int> B(np, 0);
        std::vector<MemPlan> m(np);
        std::vector<int64_t> key = {(int64_t)tsz, (int64_t)(intptr_t)c->d_arena.p, (int64_t)c->generation};
        bool any = false;
        for (int i = 0; i < np; ++i) {
            B[i] = (int)std::min<int64_t>(tile[i], ncol[i] - first[i]);
            if (B[i] <= 0) continue;
            any = true;
            m[i] = shifted(B[i] == tile[i] ? mp[i] : plan_memory(c, plans[i], B[i], tsz), base[i]);
            key.push_back((int64_t)(intptr_t)plans[i]);
            key.push_back(B[i]);
            key.insert(key.end(), knob[i].begin(), knob[i].end());
            const PassParams hp{seeds[i], first[i]};
            // pageable source: the copy completes before the call returns
            HIP_TRY(c, hipMemcpyAsync(arena + m[i].off_params, &hp, sizeof hp, hipMemcpyHostToDevice, c->stream));
        }
        if (!any) break;
        auto it = c->multi_graphs.find(key);
        if (it == c->multi_graphs.end()) {
            const hipStream_t main = c->stream;
            HIP_TRY(c, hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
            c->in_branches = true;
            hipError_t e = hipEventRecord(c->ev_fork, main);
            for (int i = 0; i < np && rc == PG_OK && e == hipSuccess; ++i) {
                if (B[i] <= 0) continue;
                e = hipStreamWaitEvent(c->side[i], c->ev_fork, 0);
                if (e != hipSuccess) break;
                put_settings(c, knob[i]);
                c->stream = c->side[i];
                rc = enqueue_pass<T>(c, plans[i], B[i], nullptr, m[i]);
                c->stream = main;
                if (rc) break;
                e = hipEventRecord(c->ev_join[i], c->side[i]);
                if (e == hipSuccess) e = hipStreamWaitEvent(main, c->ev_join[i], 0);
            }
            put_settings(c, knob[0]);
            c->in_branches = false;
            hipGraph_t graph = nullptr;
            const hipError_t ee = hipStreamEndCapture(main, &graph);
            if (rc || e != hipSuccess || ee != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                if (rc) return rc;
                return set_err(c, PG_ERR_HIP, "mixed-batch graph capture failed: %s",
                               hipGetErrorString(e != hipSuccess ? e : ee));
            }
            PassGraph g;
            e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess)
                return set_err(c, PG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
            it = c->multi_graphs.emplace(key, std::move(g)).first;
        }
        HIP_TRY(c, hipGraphLaunch(it->second.exec, c->stream));
        raw.assign(std::accumulate(B.begin(), B.end(), 0), 0.0);
        size_t o = 0;
        for (int i = 0; i < np; ++i) {
            if (B[i] <= 0) continue;
            HIP_TRY(c, hipMemcpyAsync(raw.data() + o, arena + m[i].off_rawdev, sizeof(double) * B[i],
                                      hipMemcpyDeviceToHost, c->stream));
            o += B[i];
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        o = 0;
        for (int i = 0; i < np; ++i) {
            if (B[i] <= 0) continue;
            for (int b = 0; b < B[i]; ++b)
                raw_out[out_off[i] + first[i] + b] = finish_raw(plans[i], sizeof(T) == 4, flags, raw[o + b]);
            o += B[i];
            first[i] += B[i];
        }
    }
    (void)hipEventRecord(t1, c->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    c->last_total_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    c->last_tile = tile[0];
    return PG_OK;
}

// ---------------------------------------------------------------- adjoint

// Row of every entry and, for entry (row v, column u), the index of its
// transposed entry (row u, column v): the k-th duplicate pairs with the k-th.
// Rows list columns in ascending caller id, so the partner is a binary search.
int build_transpose(pg_ctx* c)
{
    if (c->has_transpose) return PG_OK;
    const int64_t nnz = c->nnz;
    std::vector<int32_t> erow(std::max<int64_t>(nnz, 1)), tr(std::max<int64_t>(nnz, 1), -1);
    const auto& rp = c->h_rowptr;
    const auto& col = c->h_col;
    const auto& perm = c->h_perm;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : bad)
    for (int32_t v = 0; v < c->n_act; ++v) {
        const int32_t t = perm[v];
        for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
            erow[e] = v;
            const int32_t u = col[e];
            const int32_t cu = perm[u];
            int32_t dup = 0;
            while (e - dup - 1 >= rp[v] && perm[col[e - dup - 1]] == cu) ++dup;
            int32_t lo = rp[u], hi = rp[u + 1];
            while (lo < hi) {
                const int32_t mid = lo + (hi - lo) / 2;
                if (perm[col[mid]] < t) lo = mid + 1;
                else hi = mid;
            }
            const int32_t p = lo + dup;
            if (p < rp[u + 1] && perm[col[p]] == t) tr[e] = p;
            else ++bad;
        }
    }
    c->symmetric = bad == 0;
    HIP_TRY(c, c->d_erow.alloc(erow.size() * 4));
    HIP_TRY(c, c->d_tr.alloc(tr.size() * 4));
    HIP_TRY(c, hipMemcpy(c->d_erow.p, erow.data(), erow.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_tr.p, tr.data(), tr.size() * 4, hipMemcpyHostToDevice));
    c->has_transpose = true;
    return PG_OK;
}

// Buffers of the adjoint beside the forward pass's: the recomputed aux sums
// Y, dY (both sized for the widest subtemplate) and the cotangents.  With
// keep_all every forward table and every G_s has its own slot for the whole
// pass; otherwise (recompute) the tables live in a pool above the pass-long
// buffers, and off_pool / pool_bytes describe it.
struct GradPlan {
    MemPlan fwd;
    bool keep_all = true;
    std::vector<size_t> off_G;
    size_t off_Y = 0, off_dY = 0, off_cot = 0, off_pool = 0, pool_bytes = 0, total = 0;
};

size_t table_bytes(const pg_ctx* c, const DevSub& d, int B, size_t tsz)
{
    return align_up((size_t)c->n_act * B * d.Wo * tsz, 256);
}

// Whether the adjoint of step d takes its entries' gradients in the
// transposed gather (RUNS_TG, from the forward's kept Y) instead of a
// gradient gather that recomputes Y, with option fuse_grad on: when the
// transposed gather of d runs by colour runs, or its dY rows have at least
// fuse_min_width values (then by colour runs too, whatever share of a row a
// vertex uses: one whole-row read per entry replaces the gradient gather's
// read of the aux rows and the transposed gather's partial one).
bool fused_grad_step(const pg_ctx* c, const pg_plan* p, const DevSub& d)
{
    if (!c->fuse_grad || d.aa <= 1 || runs_values_per_lane(d.Wy) == 0) return false;
    // its LDS (the G_aux and own rows per team) must fit at either precision
    if (runs_tg_lds_bytes<double>(p->K, d.Wa) > kLdsLimit) return false;
    return runs_mode_for(c, AUX_GATHER_T, d.Wy, runs_used_per_mille(p, d, AUX_GATHER_T)) == RUNS_T ||
           (c->color_runs != 0 && d.Wy >= c->fuse_min_width);
}

GradPlan plan_memory_grad(const pg_ctx* c, const pg_plan* p, int B, size_t tsz)
{
    GradPlan g;
    g.fwd = plan_memory(c, p, B, tsz);
    // forward plan reuses table space; give every table its own slot instead
    size_t off = g.fwd.total;
    for (auto& d : p->subs) {
        if (d->root) continue;
        g.fwd.off_tab[d->s] = off;
        off += table_bytes(c, *d, B, tsz);
    }
    g.off_G.assign(p->nsub, 0);
    for (auto& d : p->subs) {
        if (d->root) continue;
        g.off_G[d->s] = off;
        off += table_bytes(c, *d, B, tsz);
    }
    g.fwd.off_ykeep.assign(p->nsub, 0);
    for (auto& d : p->subs) {
        if (!fused_grad_step(c, p, *d)) continue;
        g.fwd.off_ykeep[d->s] = off;
        off += align_up((size_t)c->n_act * B * d->Wy * tsz, 256);
    }
    int maxY = 1;
    for (auto& d : p->subs) maxY = std::max(maxY, d->Wy);
    g.off_Y = off;
    off += align_up((size_t)c->n_act * B * maxY * tsz, 256);
    g.off_dY = off;
    off += align_up((size_t)c->n_act * B * maxY * tsz, 256);
    g.off_cot = off;
    off += align_up((size_t)B * sizeof(double), 256);
    g.total = off;
    return g;
}

// The backward step of one subtemplate d: from G_s (nullptr at the root,
// whose cotangents are cot), the main and aux children's tables Tm / Ta,
// write G_main into Gm and G_aux into Ga (nullptr for leaf children), and add
// the step's share of every entry's gradient.  Order: dY (from G and the
// main row), one gather pass that recomputes Y and adds dY . T_aux-term to
// every entry's gradient, G_main (from G and Y), then G_aux by the
// transposed gather of the neighbours' dY.
//
// With Y kept by the forward step (Yk, fused_grad_step), the node launch
// forms dY and G_main together from it, and the transposed gather also takes
// the entries' gradients (RUNS_TG): no gradient gather, no second read of
// the aux rows.
template <typename T>
int enqueue_sub_backward(pg_ctx* c, const pg_plan* p, const DevSub& d, int B, const PassBufs<T>& pb, T* Ybuf,
                         T* dYbuf, const double* cot, const T* Tm, const T* Ta, const T* G, T* Gm, T* Ga,
                         const T* Yk = nullptr)
{
    const SubTag tag(c, d.s);
    const hipStream_t st = c->stream;
    const bool fp32 = sizeof(T) == 4;
    const SubArgs<T> base_args = pass_args<T>(c, p, B, pb);
    NodeArgs<T> na{};
    na.n_act = c->n_act;
    na.B = B;
    na.k = p->k;
    na.colors = pb.colors;
    na.G = G;
    na.root_cot = cot;
    na.Tm = d.mm > 1 ? Tm : nullptr;
    na.Y = d.aa > 1 ? Ybuf : pb.H;
    na.y_from_h = d.aa > 1 ? 0 : 1;
    na.dY = dYbuf;
    na.Gm = d.mm > 1 ? Gm : nullptr;
    na.Wo = d.Wo;
    na.Wy = d.Wy;
    na.Wm = d.Wm;
    na.Pa = d.Pa;
    na.Pm = d.Pm;
    na.by_aux = d.by_aux.as<uint32_t>();
    na.by_main = d.by_main.as<uint32_t>();
    na.by_auxT = d.by_auxT.as<uint32_t>();
    na.by_mainT = d.by_mainT.as<uint32_t>();
    na.P = d.P;
    na.scale = (d.scaled && fp32) ? (double)1.0e-12f : 1.0;
    na.sums32 = fp32 && c->node_fp32 ? 1 : 0;
    na.rows_inner = c->node_rows ? 1 : 0;
    // A main leaf's node values are dY = G * scale.  Where G is a table of
    // dY's width and the scale is 1 they equal G bit for bit, so no node
    // launch copies them and the launches below read G as dY.
    const bool dy_is_g = na.Tm == nullptr && G != nullptr && d.Wo == d.Wy && na.scale == 1.0;
    const T* dYin = dy_is_g ? G : dYbuf;
    const std::string node_name = std::string("k_bwd_node<") + (fp32 ? "float>" : "double>");
    // Algorithmic bytes, in the forward's units (full-width FP32 columns of
    // n_act x B values; Y and dY are as wide as the aux child's table; the
    // root's G, its cotangents, counts nothing): every buffer a launch reads
    // or writes, once, the CSC once per gathering launch, and the FP64
    // per-entry gradient read and written by each launch that adds to it.
    const double col = 4.0 * (double)c->n_act * B, Z = (double)c->nnz;
    const double Cs = d.root ? 0.0 : d.Cs, Cm = d.mm > 1 ? d.Cm : 0.0, Ca = d.Ca;
    if (d.aa > 1 && Yk) {
        // dY and G_main from G, the main row and the kept Y
        na.Y = Yk;
        na.part = 0;
        if (!dy_is_g) {
            // reads G, the main row and Y, writes dY and G_main
            record(c, node_name, col * (Cs + 2.0 * Ca + 2.0 * Cm), true);
            HIP_TRY(c, launch_bwd_node<T>(na, c->num_cus, st));
            record(c, "", 0, false);
        }
        SubArgs<T> t = base_args;
        t.out_mode = OUT_Y;
        if (c->weighted) t.w = fp32 ? (const T*)c->d_wT32.p : (const T*)c->d_wT64.p;
        t.aux = dYin;
        t.lda = B * d.Wy;
        t.Wa = d.Wy;  // per-coloring stride of the neighbours' dY rows
        t.Wy = d.Wa;  // values per coloring: the aux child's row
        t.Wo = d.Wa;
        t.ymask = d.amask.as<uint16_t>();
        t.xmask = d.ymask.as<uint16_t>();  // the set of each position of a neighbour's dY row
        t.gidx = d.gidxT.as<int16_t>();
        t.Wy4 = d.Wa4;
        t.out = Ga;
        t.ldo = B * d.Wa;
        t.own = Ta;
        t.ldown = B * d.Wa;
        t.tr = c->d_tr.as<int32_t>();
        t.ttab = d.ttab.as<int16_t>();
        t.grad = c->d_grad.as<double>();
        // reads the neighbours' dY, the own aux rows and the transposed
        // entry ids, writes G_aux, reads and writes each entry's gradient
        return launch_gathering<T>(c, AUX_GATHER_T, MAIN_LEAF, t, pb.part, B * d.Wa,
                                   StepBytes{csc_bytes(c), col * Ca + 20.0 * Z, col * 2.0 * Ca},
                                   runs_used_per_mille(p, d, AUX_GATHER_T));
    }
    if (d.aa > 1) {
        na.part = 1;
        if (!dy_is_g) {
            record(c, node_name, col * (Cs + Cm + Ca), true);
            HIP_TRY(c, launch_bwd_node<T>(na, c->num_cus, st));
            record(c, "", 0, false);
        }
        SubArgs<T> a = base_args;
        a.out_mode = OUT_Y;
        a.aux = Ta;
        a.Wa = d.Wa;
        a.lda = B * d.Wa;
        a.Wy = d.Wy;
        a.Wo = d.Wy;
        a.ymask = d.ymask.as<uint16_t>();
        a.xmask = d.amask.as<uint16_t>();  // the set of each position of an aux row (colour runs)
        a.gidx = d.gidx.as<int16_t>();
        a.Wy4 = d.Wy4;
        a.out = Ybuf;
        a.ldo = B * d.Wy;
        a.dy = dYin;
        a.ldd = B * d.Wy;
        a.grad = c->d_grad.as<double>();
        // reads T_aux and dY, writes Y
        int rc = launch_gathering<T>(c, AUX_GATHER_G, MAIN_LEAF, a, pb.part, B * d.Wy,
                                     StepBytes{csc_bytes(c), col * Ca + 16.0 * Z, col * 2.0 * Ca},
                                     runs_used_per_mille(p, d, AUX_GATHER_G));
        if (rc) return rc;
        if (na.Tm) {
            na.part = 2;
            record(c, node_name, col * (Cs + Ca + Cm), true);
            HIP_TRY(c, launch_bwd_node<T>(na, c->num_cus, st));
            record(c, "", 0, false);
        }
        SubArgs<T> t = base_args;
        t.out_mode = OUT_Y;
        if (c->weighted) t.w = fp32 ? (const T*)c->d_wT32.p : (const T*)c->d_wT64.p;
        t.aux = dYin;
        t.lda = B * d.Wy;
        t.Wa = d.Wy;  // per-coloring stride of the neighbours' dY rows
        t.Wy = d.Wa;  // values per coloring: the aux child's row
        t.Wo = d.Wa;
        t.ymask = d.amask.as<uint16_t>();
        t.xmask = d.ymask.as<uint16_t>();  // the set of each position of a neighbour's dY row (colour runs)
        t.gidx = d.gidxT.as<int16_t>();
        t.Wy4 = d.Wa4;
        t.out = Ga;
        t.ldo = B * d.Wa;
        // reads the neighbours' dY, writes G_aux
        rc = launch_gathering<T>(c, AUX_GATHER_T, MAIN_LEAF, t, pb.part, B * d.Wa,
                                 StepBytes{csc_bytes(c), col * Ca, col * Ca}, runs_used_per_mille(p, d, AUX_GATHER_T));
        if (rc) return rc;
    } else {
        // aux child a leaf: Y is the histogram H; the entry's derivative is
        // dY at the neighbour's color.  The node launch reads G and the main
        // row, writes dY, and with a main table also reads H and writes G_main.
        if (!dy_is_g) {
            record(c, node_name, col * (Cs + 2.0 * Cm + (d.mm > 1 ? Ca : 0.0) + Ca), true);
            HIP_TRY(c, launch_bwd_node<T>(na, c->num_cus, st));
            record(c, "", 0, false);
        }
        EdgeGradArgs<T> eg{};
        eg.nnz = c->nnz;
        eg.B = B;
        eg.k = p->k;
        eg.K = p->K;
        eg.erow = c->d_erow.as<int32_t>();
        eg.col = c->d_col.as<int32_t>();
        eg.colors = pb.colors;
        eg.ecol = pb.ecol;
        eg.dY = dYin;
        eg.Wy = d.Wy;
        eg.Ta = nullptr;
        eg.Wa = d.Wa;
        eg.ymask = d.ymask.as<uint16_t>();
        eg.rank_tab = p->rank_tab.as<int16_t>();
        eg.grad = c->d_grad.as<double>();
        // entry rows and columns, the neighbour colors (a byte per entry and
        // coloring), the vertex colors, dY, and the gradient
        record(c, std::string("k_edge_grad<") + (fp32 ? "float>" : "double>"),
               8.0 * Z + Z * B + (double)c->n_act * B + col * Ca + 16.0 * Z, true);
        HIP_TRY(c, launch_edge_grad<T>(eg, st));
        record(c, "", 0, false);
    }
    return PG_OK;
}

// Reverse sweep with every forward table kept (root first).
template <typename T>
int enqueue_backward_kept(pg_ctx* c, const pg_plan* p, int B, const GradPlan& gp)
{
    char* base = c->d_arena.as<char>();
    const PassBufs<T> pb = pass_bufs<T>(c, gp.fwd);
    auto tab = [&](int s) { return (T*)(base + gp.fwd.off_tab[s]); };
    auto gad = [&](int s) { return (T*)(base + gp.off_G[s]); };
    for (auto it = p->subs.rbegin(); it != p->subs.rend(); ++it) {
        const DevSub& d = **it;
        const size_t yk = (size_t)d.s < gp.fwd.off_ykeep.size() ? gp.fwd.off_ykeep[d.s] : 0;
        int rc = enqueue_sub_backward<T>(c, p, d, B, pb, (T*)(base + gp.off_Y), (T*)(base + gp.off_dY),
                                         (const double*)(base + gp.off_cot), d.mm > 1 ? tab(d.mi) : nullptr,
                                         d.aa > 1 ? tab(d.ai) : nullptr, d.root ? nullptr : gad(d.s),
                                         d.mm > 1 ? gad(d.mi) : nullptr, d.aa > 1 ? gad(d.ai) : nullptr,
                                         yk ? (const T*)(base + yk) : nullptr);
        if (rc) return rc;
    }
    return PG_OK;
}

// First-fit allocator of byte ranges in [0, cap) of the pool (host-side
// bookkeeping; every user is ordered on the one stream, so a range freed here
// may be handed out again at once).  With dry = true it only records the
// peak.
struct PoolAlloc {
    size_t cap = 0, peak = 0;
    std::map<size_t, size_t> used;  // offset -> bytes
    static constexpr size_t kNone = ~(size_t)0;
    size_t alloc(size_t bytes)
    {
        size_t off = 0;
        for (auto& u : used) {
            if (off + bytes <= u.first) break;
            off = std::max(off, u.first + u.second);
        }
        if (off + bytes > cap) return kNone;
        used[off] = bytes;
        peak = std::max(peak, off + bytes);
        return off;
    }
    void release(size_t off)
    {
        if (off != kNone) used.erase(off);
    }
};

// The adjoint with tables recomputed: the backward step of s needs the
// tables of its children, which are rebuilt from their subtrees (forward
// steps into pool buffers, freed as soon as their parent step has run), and
// hands G_main / G_aux on to the children's steps.  Each table is rebuilt
// once per ancestor level, and the live set is one root-to-leaf path of G
// tables plus the tables of the subtree being rebuilt.
template <typename T>
struct Recompute {
    pg_ctx* c;
    const pg_plan* p;
    int B;
    bool dry;
    PoolAlloc pool;
    char* pool_base = nullptr;
    PassBufs<T> pb{};
    T* Ybuf = nullptr;
    T* dYbuf = nullptr;
    const double* cot = nullptr;
    std::vector<const DevSub*> sub;  // by subtemplate index

    T* at(size_t off) const { return off == PoolAlloc::kNone ? nullptr : (T*)(pool_base + off); }

    // rebuilds the table of non-leaf subtemplate x; *out = its pool offset
    int materialize(int x, size_t* out)
    {
        const DevSub& d = *sub[x];
        size_t m = PoolAlloc::kNone, a = PoolAlloc::kNone;
        int rc = PG_OK;
        if (d.mm > 1) rc = materialize(d.mi, &m);
        if (rc == PG_OK && d.aa > 1) rc = materialize(d.ai, &a);
        if (rc) return rc;
        *out = pool.alloc(table_bytes(c, d, B, sizeof(T)));
        if (*out == PoolAlloc::kNone) return PG_ERR_NOMEM;
        if (!dry) rc = enqueue_step<T>(c, p, d, B, pb, at(m), at(a), at(*out), nullptr);
        pool.release(m);
        pool.release(a);
        return rc;
    }

    // backward step of x with its G table at pool offset g (none at the root)
    int backward(int x, size_t g)
    {
        const DevSub& d = *sub[x];
        size_t tm = PoolAlloc::kNone, ta = PoolAlloc::kNone, gm = PoolAlloc::kNone, ga = PoolAlloc::kNone;
        int rc = PG_OK;
        if (d.mm > 1) rc = materialize(d.mi, &tm);
        if (rc == PG_OK && d.aa > 1) rc = materialize(d.ai, &ta);
        if (rc) return rc;
        if (d.mm > 1 && (gm = pool.alloc(table_bytes(c, *sub[d.mi], B, sizeof(T)))) == PoolAlloc::kNone)
            return PG_ERR_NOMEM;
        if (d.aa > 1 && (ga = pool.alloc(table_bytes(c, *sub[d.ai], B, sizeof(T)))) == PoolAlloc::kNone)
            return PG_ERR_NOMEM;
        if (!dry) rc = enqueue_sub_backward<T>(c, p, d, B, pb, Ybuf, dYbuf, cot, at(tm), at(ta), at(g), at(gm), at(ga));
        if (rc) return rc;
        pool.release(tm);
        pool.release(ta);
        pool.release(g);
        if (d.mm > 1 && (rc = backward(d.mi, gm))) return rc;
        if (d.aa > 1 && (rc = backward(d.ai, ga))) return rc;
        return PG_OK;
    }
};

// Memory plan of the recomputing adjoint: the pass-long buffers of the
// forward plan, then Y, dY and the cotangents, then the pool, whose peak a
// dry run of the sweep measures.  The forward pass's own tables overlap Y,
// dY and the pool; they are dead once the forward pass has run.
template <typename T>
GradPlan plan_memory_recompute(pg_ctx* c, const pg_plan* p, int B, size_t budget)
{
    GradPlan g;
    g.keep_all = false;
    g.fwd = plan_memory(c, p, B, sizeof(T));
    // plan_memory places the pass-long buffers first, below every table, so
    // the lowest table offset is where they end
    size_t tab_lo = ~(size_t)0;
    for (auto& d : p->subs)
        if (!d->root) tab_lo = std::min(tab_lo, g.fwd.off_tab[d->s]);
    int maxY = 1;
    for (auto& d : p->subs) maxY = std::max(maxY, d->Wy);
    size_t off = tab_lo == ~(size_t)0 ? g.fwd.total : tab_lo;
    g.off_Y = off;
    off += align_up((size_t)c->n_act * B * maxY * sizeof(T), 256);
    g.off_dY = off;
    off += align_up((size_t)c->n_act * B * maxY * sizeof(T), 256);
    g.off_pool = off;
    const size_t cot_bytes = align_up((size_t)B * sizeof(double), 256);
    Recompute<T> r{c, p, B, true};
    r.pool.cap = budget > off + cot_bytes ? budget - off - cot_bytes : 0;
    r.sub.assign(p->nsub, nullptr);
    for (auto& d : p->subs) r.sub[d->s] = d.get();
    g.pool_bytes = r.backward(0, PoolAlloc::kNone) == PG_OK ? r.pool.peak : ~(size_t)0;
    if (g.pool_bytes == ~(size_t)0) {
        g.total = ~(size_t)0;
        return g;
    }
    // the cotangents above everything else: they are uploaded before the
    // pass, while the forward pass's tables overlap Y, dY and the pool
    g.off_cot = align_up(std::max(g.fwd.total, off + g.pool_bytes), 256);
    g.total = g.off_cot + cot_bytes;
    return g;
}

template <typename T>
int count_grad_impl(pg_ctx* c, const pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed, uint32_t flags,
                    const double* cotangent, double* raw_out, void* grad, int grad_on_device)
{
    int rc = build_buckets(c);
    if (rc) return rc;
    rc = build_transpose(c);
    if (rc) return rc;
    if (!c->symmetric)
        return set_err(c, PG_ERR_INVALID, "gradients need a symmetric pattern (every entry's transpose present)");
    const size_t tsz = sizeof(T);
    size_t budget;
    if (c->mem_budget_mb > 0) {
        budget = (size_t)c->mem_budget_mb << 20;
    } else {
        size_t fr = 0, tot = 0;
        HIP_TRY(c, hipMemGetInfo(&fr, &tot));
        budget = (size_t)(0.8 * (double)(fr + c->d_arena.bytes));
    }
    // the largest batch tile that fits, with every table kept and with tables
    // recomputed; option "adjoint_recompute": 0 keep, 1 recompute, -1 keep
    // unless recomputing fits kRecomputeGain times as many colorings per pass
    // (or keeping fits none).  Rebuilding a table repeats its subtree's
    // forward sweep at every level that needs it (measured: 1.9x the time per
    // coloring for the 12-vertex tree at R-MAT scale 22, same tile), while a
    // wider pass gains less than that per coloring beyond a few colorings.
    constexpr int kRecomputeGain = 4;
    const int cap = c->batch_tile > 0 ? (int)std::min<int64_t>(c->batch_tile, ncol) : std::min(ncol, 64);
    auto lds_ok = [&](int B) {
        size_t lds = 0;
        for (auto& d : p->subs) lds = std::max(lds, sub_lds_max(c, p, *d, B, tsz));
        return lds <= kLdsLimit;
    };
    int tile_keep = 0, tile_rec = 0;
    GradPlan g_keep, g_rec;
    if (c->adjoint_recompute != 1)
        for (int B = cap; B >= 1; --B) {
            if (!lds_ok(B)) continue;
            GradPlan g = plan_memory_grad(c, p, B, tsz);
            if (g.total > budget) continue;
            tile_keep = B;
            g_keep = g;
            break;
        }
    const int rec_lo = c->adjoint_recompute == 1 ? 1 : tile_keep == 0 ? 1 : tile_keep * kRecomputeGain;
    if (c->adjoint_recompute != 0)
        for (int B = cap; B >= rec_lo; --B) {
            if (!lds_ok(B)) continue;
            GradPlan g = plan_memory_recompute<T>(c, p, B, budget);
            if (g.total > budget) continue;
            tile_rec = B;
            g_rec = g;
            break;
        }
    const bool use_rec = tile_rec > 0;
    int tile = use_rec ? tile_rec : tile_keep;
    GradPlan gp = use_rec ? g_rec : g_keep;
    if (tile == 0) return set_err(c, PG_ERR_NOMEM, "adjoint tables of one coloring do not fit the device");
    c->last_tile = tile;
    c->last_recompute = use_rec ? 1 : 0;
    if (c->autotune) {
        // the kernel variants, degree buckets and column tile tuned on
        // forward passes of this tile (the adjoint's gathers have the same
        // shapes and share the buckets): a smaller tuned tile is taken over,
        // as wide templates gather faster per coloring at a few colorings per
        // pass in both sweeps (measured for the 12-vertex tree at R-MAT scale
        // 22: 2.26 against 2.04 colorings/s at 3 and 5 per pass).  The plan
        // is made again with the tuned buckets, which size the hub partial
        // rows.
        const size_t fwd_total = plan_memory(c, p, tile, tsz).total;
        if (fwd_total > c->d_arena.bytes) {
            drop_graphs(c);
            HIP_TRY(c, c->d_arena.alloc(fwd_total));
        }
        int fwd_tile = tile;
        const int rt = autotune<T>(c, p, tile, &fwd_tile);
        if (rt) return rt;
        if (fwd_tile >= 1 && fwd_tile < tile) tile = fwd_tile;
        c->last_tile = tile;
        gp = use_rec ? plan_memory_recompute<T>(c, p, tile, budget) : plan_memory_grad(c, p, tile, tsz);
    }
    if (gp.total > c->d_arena.bytes) {
        drop_graphs(c);
        HIP_TRY(c, c->d_arena.alloc(gp.total));
    }
    if (c->weighted) {
        HIP_TRY(c, c->d_wT32.alloc(c->nnz * 4));
        HIP_TRY(c, c->d_wT64.alloc(c->nnz * 8));
        HIP_TRY(c, launch_gather_weights_f32(c->d_wT32.as<float>(), c->d_w32.as<float>(), c->d_tr.as<int32_t>(),
                                             c->nnz, c->stream));
        HIP_TRY(c, launch_gather_weights_f64(c->d_wT64.as<double>(), c->d_w32.as<float>(), c->d_tr.as<int32_t>(),
                                             c->nnz, c->stream));
    }
    HIP_TRY(c, c->d_grad.alloc(std::max<int64_t>(c->nnz, 1) * 8));
    HIP_TRY(c, launch_zero_f64(c->d_grad.as<double>(), c->nnz, c->stream));
    const uint8_t* d_in = nullptr;
    if (colors) {
        if ((size_t)ncol * c->n > c->d_colors_in.bytes) {
            drop_graphs(c);
            HIP_TRY(c, c->d_colors_in.alloc((size_t)ncol * c->n));
        }
        HIP_TRY(c, hipMemcpyAsync(c->d_colors_in.p, colors, (size_t)ncol * c->n, hipMemcpyHostToDevice, c->stream));
        d_in = c->d_colors_in.as<uint8_t>();
    }
    c->agg.clear();
    c->agg_sub.clear();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(c, hipEventCreate(&t0));
    HIP_TRY(c, hipEventCreate(&t1));
    HIP_TRY(c, hipEventRecord(t0, c->stream));
    char* base = c->d_arena.as<char>();
    std::vector<double> raw;
    for (int first = 0; first < ncol && rc == PG_OK; first += tile) {
        const int B = std::min(tile, ncol - first);
        const GradPlan g = B == tile ? gp : use_rec ? plan_memory_recompute<T>(c, p, B, gp.total)
                                                    : plan_memory_grad(c, p, B, tsz);
        if (g.total > gp.total) {
            rc = set_err(c, PG_ERR_NOMEM, "adjoint plan of the last pass exceeds the first");
            break;
        }
        // pass parameters and cotangents (slots of their own), then the
        // forward pass and the reverse sweep: replayed from a captured graph
        // per shape, or launched eagerly while timing or with graphs off
        const PassParams hp{seed, first};
        HIP_TRY(c, hipMemcpyAsync(base + g.fwd.off_params, &hp, sizeof hp, hipMemcpyHostToDevice, c->stream));
        std::vector<double> cot(B, 1.0);
        if (cotangent) std::copy(cotangent + first, cotangent + first + B, cot.begin());
        HIP_TRY(c, hipMemcpyAsync(base + g.off_cot, cot.data(), sizeof(double) * B, hipMemcpyHostToDevice, c->stream));
        auto enqueue_grad_pass = [&]() -> int {
            int r0 = enqueue_pass<T>(c, p, B, d_in, g.fwd);
            if (r0 == PG_OK && g.keep_all) r0 = enqueue_backward_kept<T>(c, p, B, g);
            if (r0 == PG_OK && !g.keep_all) {
                Recompute<T> r{c, p, B, false};
                r.pool.cap = g.pool_bytes;
                r.pool_base = base + g.off_pool;
                r.pb = pass_bufs<T>(c, g.fwd);
                r.Ybuf = (T*)(base + g.off_Y);
                r.dYbuf = (T*)(base + g.off_dY);
                r.cot = (const double*)(base + g.off_cot);
                r.sub.assign(p->nsub, nullptr);
                for (auto& d : p->subs) r.sub[d->s] = d.get();
                r0 = r.backward(0, PoolAlloc::kNone);
                if (r0 == PG_ERR_NOMEM) set_err(c, r0, "adjoint recompute pool exhausted");
            }
            return r0;
        };
        std::vector<LaunchRec> recs;  // filled when option "timing" is on
        if (!c->use_graphs || c->timing) {
            c->cur_recs = &recs;
            rc = enqueue_grad_pass();
            c->cur_recs = nullptr;
        } else {
            GraphKey key{p, B, (int)tsz, d_in, c->d_arena.p, c->generation};
            key.kind = use_rec ? 2 : 1;
            key.grad = c->d_grad.p;
            auto it = c->graphs.find(key);
            if (it == c->graphs.end()) {
                HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                rc = enqueue_grad_pass();
                hipGraph_t graph = nullptr;
                const hipError_t e = hipStreamEndCapture(c->stream, &graph);
                if (rc == PG_OK && e != hipSuccess)
                    rc = set_err(c, PG_ERR_HIP, "gradient graph capture failed: %s", hipGetErrorString(e));
                if (rc == PG_OK) {
                    PassGraph pg;
                    const hipError_t ei = hipGraphInstantiate(&pg.exec, graph, nullptr, nullptr, 0);
                    if (ei != hipSuccess)
                        rc = set_err(c, PG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
                    else
                        it = c->graphs.emplace(key, std::move(pg)).first;
                }
                if (graph) (void)hipGraphDestroy(graph);
            }
            if (rc == PG_OK) HIP_TRY(c, hipGraphLaunch(it->second.exec, c->stream));
        }
        if (rc) {
            destroy_recs(recs);
            break;
        }
        raw.assign(B, 0.0);
        HIP_TRY(c, hipMemcpyAsync(raw.data(), base + g.fwd.off_rawdev, sizeof(double) * B, hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        accumulate_timing(c, recs);
        destroy_recs(recs);
        for (int b = 0; b < B; ++b) raw_out[first + b] = finish_raw(p, sizeof(T) == 4, flags, raw[b]);
    }
    if (rc == PG_OK && grad) {
        const double factor = finish_raw(p, sizeof(T) == 4, flags, 1.0);
        if (grad_on_device) {
            if (sizeof(T) == 4)
                HIP_TRY(c, launch_scatter_grad_f32((float*)grad, c->d_grad.as<double>(), c->d_ent.as<int32_t>(),
                                                   c->nnz, factor, c->stream));
            else
                HIP_TRY(c, launch_scatter_grad_f64((double*)grad, c->d_grad.as<double>(), c->d_ent.as<int32_t>(),
                                                   c->nnz, factor, c->stream));
        } else {
            HIP_TRY(c, c->d_grad_out.alloc(std::max<int64_t>(c->nnz, 1) * 8));
            HIP_TRY(c, launch_scatter_grad_f64(c->d_grad_out.as<double>(), c->d_grad.as<double>(),
                                               c->d_ent.as<int32_t>(), c->nnz, factor, c->stream));
            HIP_TRY(c, hipMemcpyAsync(grad, c->d_grad_out.p, c->nnz * 8, hipMemcpyDeviceToHost, c->stream));
        }
    }
    (void)hipEventRecord(t1, c->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    c->last_total_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    finish_timing(c);
    return rc;
}

// colex rank of every K-bit mask among the masks of its popcount: colex
// order of sets is the numeric order of their bit masks
std::vector<int16_t> colex_rank_table(int K)
{
    std::vector<int16_t> rank((size_t)1 << K);
    std::vector<int> next(K + 1, 0);
    for (uint32_t m = 0; m < (1u << K); ++m) rank[m] = (int16_t)next[__builtin_popcount(m)]++;
    return rank;
}

// SubArgs::gidx: for a vertex of color cv, Y value i (the K-bit set ymask[i]
// relative to cv) takes from a neighbour of color cu its row value at the
// rank of (set minus cu) relative to cu, when cu is in the set
std::vector<int16_t> gather_index_table(int k, const std::vector<uint16_t>& ymask, const std::vector<int16_t>& rank,
                                        int Wy4)
{
    const int Wy = (int)ymask.size();
    std::vector<int16_t> g((size_t)(k + 1) * k * Wy4, (int16_t)-1);
    auto expand = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u)); };
    auto squeeze = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c); };
    for (int cv = 0; cv < k; ++cv)
        for (int cu = 0; cu < k; ++cu) {
            if (cu == cv) continue;
            for (int i = 0; i < Wy; ++i) {
                const uint32_t A = expand(ymask[i], (uint32_t)cv);
                if (!((A >> cu) & 1u)) continue;
                g[((size_t)cv * k + cu) * Wy4 + i] = rank[squeeze(A & ~(1u << cu), (uint32_t)cu)];
            }
        }
    return g;
}

// SubArgs::gidx of the adjoint's transposed gather: the vertex's row value i
// (the set amask[i] relative to cv, plus cv) takes from a neighbour of color
// cu its dY value at the rank of that set relative to cu, when cu is outside it
std::vector<int16_t> gather_index_table_t(int k, const std::vector<uint16_t>& amask, const std::vector<int16_t>& rank,
                                          int Wa4)
{
    const int Wa = (int)amask.size();
    std::vector<int16_t> g((size_t)(k + 1) * k * Wa4, (int16_t)-1);
    auto expand = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u)); };
    auto squeeze = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c); };
    for (int cv = 0; cv < k; ++cv)
        for (int cu = 0; cu < k; ++cu)
            for (int i = 0; i < Wa; ++i) {
                const uint32_t A = expand(amask[i], (uint32_t)cv) | (1u << cv);
                if ((A >> cu) & 1u) continue;  // also cu == cv
                g[((size_t)cv * k + cu) * Wa4 + i] = rank[squeeze(A, (uint32_t)cu)];
            }
    return g;
}

// SubArgs::xidx: value j of the row of a neighbour of color cu (the set
// amask[j] relative to cu, plus cu) adds, at a vertex of color cv, to the Y
// value of that set relative to cv, when cv is outside it
// SubArgs::ttab of the fused transposed gather: for a vertex of colour cv
// and a neighbour of colour c, position j = lane + 64 t of the neighbour's dY
// row (the set ymask[j] relative to c, plus c) adds to the vertex's row entry
// of that set minus cv (relative to cv) when the set holds cv, else -1; one
// row of rv values per lane, so a lane reads its rv indices with one load
std::vector<int16_t> runs_tg_table(int k, const std::vector<uint16_t>& ymask, const std::vector<int16_t>& rank, int rv)
{
    const int Wy = (int)ymask.size();
    std::vector<int16_t> g((size_t)k * k * 64 * rv, (int16_t)-1);
    auto expand = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u)); };
    auto squeeze = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c); };
    for (int cv = 0; cv < k; ++cv)
        for (int c = 0; c < k; ++c) {
            if (c == cv) continue;  // entries of the vertex's own colour add nothing
            for (int lane = 0; lane < 64; ++lane)
                for (int t = 0; t < rv; ++t) {
                    const int j = lane + 64 * t;
                    if (j >= Wy) continue;
                    const uint32_t A = expand(ymask[j], (uint32_t)c);
                    if (!((A >> cv) & 1u)) continue;
                    g[(((size_t)cv * k + c) * 64 + lane) * rv + t] = rank[squeeze(A & ~(1u << cv), (uint32_t)cv)];
                }
        }
    return g;
}

std::vector<int16_t> rows_index_table(int k, const std::vector<uint16_t>& amask, const std::vector<int16_t>& rank,
                                      int Wa4)
{
    const int Wa = (int)amask.size();
    std::vector<int16_t> x((size_t)k * k * Wa4, (int16_t)-1);
    auto expand = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> c) << (c + 1u)); };
    auto squeeze = [](uint32_t m, uint32_t c) { return (m & ((1u << c) - 1u)) | ((m >> (c + 1u)) << c); };
    for (int cv = 0; cv < k; ++cv)
        for (int cu = 0; cu < k; ++cu)
            for (int j = 0; j < Wa; ++j) {
                const uint32_t S = expand(amask[j], (uint32_t)cu) | (1u << cu);
                if ((S >> cv) & 1u) continue;
                x[((size_t)cv * k + cu) * Wa4 + j] = rank[squeeze(S, (uint32_t)cv)];
            }
    return x;
}

std::vector<uint16_t> masks_of_size(int K, int size)
{
    std::vector<uint16_t> out;
    for (uint32_t m = 0; m < (1u << K); ++m)
        if (__builtin_popcount(m) == size) out.push_back((uint16_t)m);
    return out;
}

}  // namespace

// =========================================================================
// C ABI
// =========================================================================

// ------------------------------------------------------------ static scratch
//
// pg_count_noalloc places a pass's buffers (colors, neighbour colors, the
// histogram, count tables, hub partial rows, root values, results) in this
// array instead of the context's arena.  Each device holds its own copy for
// the life of the process.  Host code must not pass `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current (scratch_setup, from pg_create) and kept per device.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[kMaxDevices] = {};

// with c->device current
hipError_t scratch_setup(pg_ctx* c)
{
    if (c->device < 0 || c->device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[c->device] = static_cast<char*>(ptr);
    }
    c->scratch = g_scratch_dev[c->device];
    return hipSuccess;
}
}  // namespace

extern "C" int pg_create(pg_ctx** out, int device)
{
    if (!out) return PG_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PG_ERR_HIP;
    if (device < 0 || device >= ndev) return PG_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return PG_ERR_HIP;
    pg_ctx* c = new (std::nothrow) pg_ctx();
    if (!c) return PG_ERR_NOMEM;
    c->device = device;
    if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        c->num_cus <= 0)
        c->num_cus = 256;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PG_ERR_HIP;
    }
    if (scratch_setup(c) != hipSuccess || hipStreamCreateWithFlags(&c->hub_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_hub_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_hub_join, hipEventDisableTiming) != hipSuccess) {
        if (c->ev_hub_fork) (void)hipEventDestroy(c->ev_hub_fork);
        if (c->hub_stream) (void)hipStreamDestroy(c->hub_stream);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PG_ERR_HIP;
    }
    *out = c;
    return PG_OK;
}

extern "C" void pg_destroy(pg_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (hipStream_t s : c->side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : c->ev_join) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_hub_fork) (void)hipEventDestroy(c->ev_hub_fork);
    if (c->ev_hub_join) (void)hipEventDestroy(c->ev_hub_join);
    if (c->hub_stream) (void)hipStreamDestroy(c->hub_stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* pg_last_error(const pg_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int pg_set_option(pg_ctx* c, const char* key, int64_t v)
{
    if (!c || !key) return PG_ERR_INVALID;
    std::string k(key);
    if (k == "hub_degree") {
        if (v < 16) return set_err(c, PG_ERR_INVALID, "hub_degree must be >= 16");
        c->hub_degree = v;
    } else if (k == "hub_chunk") {
        if (v < 16 || v > (1 << 24)) return set_err(c, PG_ERR_INVALID, "hub_chunk must be in [16, 2^24]");
        c->hub_chunk = v;
    } else if (k == "bucket_high") {
        if (v < 1) return set_err(c, PG_ERR_INVALID, "bucket_high must be >= 1");
        c->bucket_high = v;
    } else if (k == "bucket_mid") {
        if (v < 1) return set_err(c, PG_ERR_INVALID, "bucket_mid must be >= 1");
        c->bucket_mid = v;
    } else if (k == "batch_tile") {
        if (v < 0 || v > 4096) return set_err(c, PG_ERR_INVALID, "batch_tile must be in [0, 4096]");
        c->batch_tile = v;
    } else if (k == "mem_budget_mb") {
        if (v < 0) return set_err(c, PG_ERR_INVALID, "mem_budget_mb must be >= 0");
        c->mem_budget_mb = v;
    } else if (k == "timing") {
        c->timing = v ? 1 : 0;
    } else if (k == "graphs") {
        c->use_graphs = v != 0;
    } else if (k == "autotune") {
        c->autotune = v ? 1 : 0;
    } else if (k == "aux_rows") {
        if (v < -1 || v > 1) return set_err(c, PG_ERR_INVALID, "aux_rows must be -1, 0 or 1");
        if (v != c->aux_rows) {
            drop_graphs(c);
            c->aux_rows = v;
        }
    } else if (k == "adjoint_recompute") {
        if (v < -1 || v > 1) return set_err(c, PG_ERR_INVALID, "adjoint_recompute must be -1, 0 or 1");
        c->adjoint_recompute = v;
    } else if (k == "vertex_order") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "vertex_order must be 0 or 1");
        c->vertex_order = v;  // takes effect at the next graph upload
    } else if (k == "reg_cols") {
        if (v != 0 && v != 1 && v != kRegCols)
            return set_err(c, PG_ERR_INVALID, "reg_cols must be 0, 1 or %d", kRegCols);
        if (v != c->reg_cols) {
            drop_graphs(c);
            c->reg_cols = v;
        }
    } else if (k == "stage_rows") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "stage_rows must be 0 or 1");
        if (v != c->stage_rows) {
            drop_graphs(c);
            c->stage_rows = v;
        }
    } else if (k == "exact_order") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "exact_order must be 0 or 1");
        if (v != c->exact_order) {
            drop_graphs(c);
            c->exact_order = v;
        }
    } else if (k == "fuse_min_width") {
        if (v < 1 || v > 4096) return set_err(c, PG_ERR_INVALID, "fuse_min_width must be in 1..4096");
        if (v != c->fuse_min_width) {
            drop_graphs(c);
            c->fuse_min_width = v;
        }
    } else if (k == "seq_max_width") {
        if (v < 1 || v > 4096) return set_err(c, PG_ERR_INVALID, "seq_max_width must be in 1..4096");
        if (v != c->seq_max_width) {
            drop_graphs(c);
            c->seq_max_width = v;
        }
    } else if (k == "runs_g_min") {
        if (v < 1 || v > 4096) return set_err(c, PG_ERR_INVALID, "runs_g_min must be in 1..4096");
        if (v != c->runs_g_min) {
            drop_graphs(c);
            c->runs_g_min = v;
        }
    } else if (k == "node_rows") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "node_rows must be 0 or 1");
        if (v != c->node_rows) {
            drop_graphs(c);
            c->node_rows = v;
        }
    } else if (k == "node_fp32") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "node_fp32 must be 0 or 1");
        if (v != c->node_fp32) {
            drop_graphs(c);
            c->node_fp32 = v;
        }
    } else if (k == "fuse_grad") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "fuse_grad must be 0 or 1");
        if (v != c->fuse_grad) {
            drop_graphs(c);
            c->fuse_grad = v;
        }
    } else if (k == "hub_concurrent") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "hub_concurrent must be 0 or 1");
        if (v != c->hub_concurrent) {
            drop_graphs(c);
            c->hub_concurrent = v;
        }
    } else if (k == "color_runs") {
        if (v < -1 || v > 1) return set_err(c, PG_ERR_INVALID, "color_runs must be -1, 0 or 1");
        if (v != c->color_runs) {
            drop_graphs(c);
            c->color_runs = v;
        }
    } else if (k == "split_lds") {
        if (v != 0 && v != 1) return set_err(c, PG_ERR_INVALID, "split_lds must be 0 or 1");
        if (v != c->split_lds) {
            drop_graphs(c);
            c->split_lds = v;
        }
    } else if (k == "blocks_per_cu") {
        if (v < 0 || v > 64) return set_err(c, PG_ERR_INVALID, "blocks_per_cu must be in [0, 64]");
        if (v != c->blocks_per_cu) {
            drop_graphs(c);  // captured launches hold the old grid sizes
            c->blocks_per_cu = v;
        }
    } else {
        return set_err(c, PG_ERR_INVALID, "unknown option %s", key);
    }
    return PG_OK;
}

extern "C" int pg_get_option(const pg_ctx* c, const char* key, int64_t* v)
{
    if (!c || !key || !v) return PG_ERR_INVALID;
    std::string k(key);
    if (k == "hub_degree") *v = c->hub_degree;
    else if (k == "hub_chunk") *v = c->hub_chunk;
    else if (k == "bucket_high") *v = c->bucket_high;
    else if (k == "bucket_mid") *v = c->bucket_mid;
    else if (k == "batch_tile") *v = c->batch_tile;
    else if (k == "mem_budget_mb") *v = c->mem_budget_mb;
    else if (k == "timing") *v = c->timing;
    else if (k == "autotune") *v = c->autotune;
    else if (k == "blocks_per_cu") *v = c->blocks_per_cu;
    else if (k == "aux_rows") *v = c->aux_rows;
    else if (k == "stage_rows") *v = c->stage_rows;
    else if (k == "split_lds") *v = c->split_lds;
    else if (k == "color_runs") *v = c->color_runs;
    else if (k == "exact_order") *v = c->exact_order;
    else if (k == "hub_concurrent") *v = c->hub_concurrent;
    else if (k == "fuse_grad") *v = c->fuse_grad;
    else if (k == "node_fp32") *v = c->node_fp32;
    else if (k == "node_rows") *v = c->node_rows;
    else if (k == "runs_g_min") *v = c->runs_g_min;
    else if (k == "fuse_min_width") *v = c->fuse_min_width;
    else if (k == "seq_max_width") *v = c->seq_max_width;
    else if (k == "reg_cols") *v = c->reg_cols;
    else if (k == "adjoint_recompute") *v = c->adjoint_recompute;
    else if (k == "vertex_order") *v = c->vertex_order;
    else if (k == "last_tile") *v = c->last_tile;
    else if (k == "last_recompute") *v = c->last_recompute;
    else if (k == "graphs") *v = c->use_graphs ? 1 : 0;
    else return PG_ERR_INVALID;
    return PG_OK;
}

// Internal order: vertices that appear in any entry, by descending row
// degree (ties by caller id), so hubs form a prefix, degree buckets are
// ranges, and isolated vertices -- whose tables are all zero above the
// leaves -- are dropped.  Row v lists the entries (row v, column u) in
// ascending u, the order in which the reference accumulates row v
// (CSCGraph.hpp:199-213, 254-307).
extern "C" int pg_graph_upload(pg_ctx* c, int32_t n, int64_t nnz, const int32_t* colptr, const int32_t* rowidx,
                               const float* vals)
{
    if (!c) return PG_ERR_INVALID;
    if (n <= 0 || nnz < 0 || nnz > INT32_MAX || !colptr || (nnz > 0 && !rowidx))
        return set_err(c, PG_ERR_INVALID, "invalid graph arguments (n=%d nnz=%lld)", n, (long long)nnz);
    if (colptr[0] != 0 || colptr[n] != nnz) return set_err(c, PG_ERR_INVALID, "colptr must run from 0 to nnz");
    for (int32_t u = 0; u < n; ++u)
        if (colptr[u + 1] < colptr[u]) return set_err(c, PG_ERR_INVALID, "colptr not monotone at %d", u);
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<int32_t> rdeg(n, 0);
    std::vector<char> seen(n, 0);
    for (int32_t u = 0; u < n; ++u) {
        if (colptr[u + 1] > colptr[u]) seen[u] = 1;
        for (int64_t e = colptr[u]; e < colptr[u + 1]; ++e) {
            int32_t v = rowidx[e];
            if (v < 0 || v >= n)
                return set_err(c, PG_ERR_INVALID, "row index %d out of range at entry %lld", v, (long long)e);
            rdeg[v]++;
            seen[v] = 1;
        }
    }
    // row-major view in caller ids (ascending column within each row)
    std::vector<int64_t> rptr((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v) rptr[v + 1] = rptr[v] + rdeg[v];
    std::vector<int32_t> rcol((size_t)nnz), rent((size_t)nnz);
    {
        std::vector<int64_t> pos(rptr.begin(), rptr.end() - 1);
        for (int32_t u = 0; u < n; ++u)
            for (int64_t e = colptr[u]; e < colptr[u + 1]; ++e) {
                int32_t v = rowidx[e];
                rcol[pos[v]] = u;
                rent[pos[v]] = (int32_t)e;
                pos[v]++;
            }
    }
    // Internal order: descending degree (hubs first, each degree bucket an
    // index range).  Option vertex_order = 1 instead takes degree classes,
    // highest first, whose boundaries are every bucket threshold the
    // autotuner tries (so each tuned bucket is still an index range), and
    // within a class breadth-first order from the largest hubs, so that
    // vertices processed side by side share neighbour rows in L2.
    std::vector<int32_t> perm;
    perm.reserve(n);
    for (int32_t v = 0; v < n; ++v)
        if (seen[v]) perm.push_back(v);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return rdeg[a] > rdeg[b]; });
    const int32_t na = (int32_t)perm.size();
    if (c->vertex_order == 1) {
        std::vector<int32_t> bfs(n, -1);
        std::vector<int32_t> queue;
        queue.reserve(na);
        int32_t next = 0;
        for (int32_t root : perm) {
            if (bfs[root] >= 0) continue;
            bfs[root] = next++;
            queue.assign(1, root);
            for (size_t h = 0; h < queue.size(); ++h) {
                const int32_t v = queue[h];
                for (int64_t e = rptr[v]; e < rptr[v + 1]; ++e) {
                    const int32_t u = rcol[e];
                    if (bfs[u] < 0) {
                        bfs[u] = next++;
                        queue.push_back(u);
                    }
                }
            }
        }
        // degree >= bound: hubs above 128, 256, 512, 1024; team buckets from 4 .. 512
        static const int64_t bounds[] = {1025, 513, 512, 257, 256, 129, 128, 64, 32, 16, 8, 4};
        auto cls = [&](int32_t v) {
            int k = 0;
            for (int64_t b : bounds) {
                if (rdeg[v] >= b) break;
                ++k;
            }
            return k;
        };
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
            const int ca = cls(a), cb = cls(b);
            return ca != cb ? ca < cb : bfs[a] < bfs[b];
        });
    }
    std::vector<int32_t> inv(n, -1);
    for (int32_t i = 0; i < na; ++i) inv[perm[i]] = i;
    std::vector<int32_t> hrow((size_t)na + 1, 0), hcol((size_t)nnz), hent((size_t)nnz);
    for (int32_t i = 0; i < na; ++i) hrow[i + 1] = hrow[i] + rdeg[perm[i]];
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t i = 0; i < na; ++i) {
        int32_t v = perm[i];
        int64_t o = hrow[i];
        for (int64_t e = rptr[v]; e < rptr[v + 1]; ++e, ++o) {
            hcol[o] = inv[rcol[e]];
            hent[o] = rent[e];
        }
    }
    c->has_graph = false;
    c->fascia_acc_ready = false;
    c->n = n;
    c->nnz = nnz;
    c->n_act = na;
    c->max_deg = 0;
    for (int32_t v : perm) c->max_deg = std::max(c->max_deg, rdeg[v]);
    c->h_rowptr = hrow;
    c->h_col = hcol;
    c->h_perm = perm;
    c->has_transpose = false;
    HIP_TRY(c, c->d_rowptr.alloc(hrow.size() * 4));
    HIP_TRY(c, c->d_col.alloc(std::max<size_t>(hcol.size(), 1) * 4));
    HIP_TRY(c, c->d_ent.alloc(std::max<size_t>(hent.size(), 1) * 4));
    HIP_TRY(c, c->d_perm.alloc(std::max<size_t>(perm.size(), 1) * 4));
    HIP_TRY(c, hipMemcpy(c->d_rowptr.p, hrow.data(), hrow.size() * 4, hipMemcpyHostToDevice));
    if (nnz > 0) {
        HIP_TRY(c, hipMemcpy(c->d_col.p, hcol.data(), hcol.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->d_ent.p, hent.data(), hent.size() * 4, hipMemcpyHostToDevice));
    }
    if (na > 0) HIP_TRY(c, hipMemcpy(c->d_perm.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice));
    c->built_deg = -1;
    drop_graphs(c);
    c->tuned.clear();
    c->generation++;
    c->weighted = false;
    c->has_graph = true;
    int rc = build_buckets(c);
    if (rc) return rc;
    if (vals) return pg_graph_set_weights(c, vals, 0);
    return PG_OK;
}

extern "C" int pg_graph_set_weights(pg_ctx* c, const float* vals, int on_device)
{
    if (!c) return PG_ERR_INVALID;
    if (!c->has_graph) return set_err(c, PG_ERR_STATE, "no graph uploaded");
    HIP_TRY(c, hipSetDevice(c->device));
    // captured passes read the weights through d_w32 / d_w64: new values need
    // no new capture, a change of those buffers or of weighted-ness does
    if (!vals) {
        if (c->weighted) c->generation++;
        c->weighted = false;
        return PG_OK;
    }
    const void* old32 = c->d_w32.p;
    const void* old64 = c->d_w64.p;
    const bool was_weighted = c->weighted;
    const float* src = vals;
    if (!on_device) {
        HIP_TRY(c, c->d_wsrc.alloc(std::max<int64_t>(c->nnz, 1) * 4));
        HIP_TRY(c, hipMemcpyAsync(c->d_wsrc.p, vals, c->nnz * 4, hipMemcpyHostToDevice, c->stream));
        src = c->d_wsrc.as<float>();
    }
    HIP_TRY(c, c->d_w32.alloc(std::max<int64_t>(c->nnz, 1) * 4));
    HIP_TRY(c, c->d_w64.alloc(std::max<int64_t>(c->nnz, 1) * 8));
    HIP_TRY(c, launch_gather_weights_f32(c->d_w32.as<float>(), src, c->d_ent.as<int32_t>(), c->nnz, c->stream));
    HIP_TRY(c, launch_gather_weights_f64(c->d_w64.as<double>(), src, c->d_ent.as<int32_t>(), c->nnz, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->weighted = true;
    if (!was_weighted || old32 != c->d_w32.p || old64 != c->d_w64.p) c->generation++;
    return PG_OK;
}

extern "C" int pg_graph_stats(const pg_ctx* c, int64_t* n, int64_t* nnz, int64_t* n_active, int64_t* max_degree)
{
    if (!c) return PG_ERR_INVALID;
    if (!c->has_graph) return PG_ERR_STATE;
    if (n) *n = c->n;
    if (nnz) *nnz = c->nnz;
    if (n_active) *n_active = c->n_act;
    if (max_degree) *max_degree = c->max_deg;
    return PG_OK;
}

// The reference tables (IndexSys) fix the column of every color set (colex
// rank) and the order of the splits (lexicographic position subsets, aux =
// complement).  The kernels rely on exactly that convention, re-expressed
// per vertex color: the plan checks the caller's tables against it, then
// builds the color-relative tables.  Restricting the reference's split list
// of a set S to the splits whose main part holds col(v), and dropping col(v),
// leaves the lexicographic position-subset order of S \ {col(v)}: removing a
// common element preserves lexicographic order.
extern "C" int pg_plan_create(pg_ctx* c, int k, int nsub, const int32_t* sub_size, const int32_t* main_idx,
                              const int32_t* aux_idx, const int32_t* comb_to_count, const int32_t* split_main,
                              const int32_t* split_aux, pg_plan** out)
{
    if (!c || !out) return PG_ERR_INVALID;
    *out = nullptr;
    if (k < 2 || k > kMaxK || nsub != 2 * k - 1 || !sub_size || !main_idx || !aux_idx || !comb_to_count ||
        !split_main || !split_aux)
        return set_err(c, PG_ERR_INVALID, "plan needs 2 <= k <= %d and nsub = 2k-1", kMaxK);
    if (sub_size[0] != k) return set_err(c, PG_ERR_INVALID, "subtemplate 0 must be the whole template");
    int nonleaf = 0;
    for (int s = 0; s < nsub; ++s) {
        const int sz = sub_size[s];
        if (sz < 1 || sz > k) return set_err(c, PG_ERR_INVALID, "subtemplate %d has size %d", s, sz);
        if (sz == 1) continue;
        ++nonleaf;
        const int mi = main_idx[s], ai = aux_idx[s];
        if (mi <= s || ai <= s || mi >= nsub || ai >= nsub || mi == ai)
            return set_err(c, PG_ERR_INVALID, "subtemplate %d children must follow it", s);
        if (sub_size[mi] + sub_size[ai] != sz)
            return set_err(c, PG_ERR_INVALID, "subtemplate %d: child sizes do not add up", s);
    }
    // the caller's index tables must be the IndexSys convention
    {
        std::vector<int64_t> c2c_off(nsub + 1, 0), split_off(nsub + 1, 0);
        for (int s = 0; s < nsub; ++s) {
            c2c_off[s + 1] = c2c_off[s] + IndexSys::comb_calc(k, sub_size[s]);
            int64_t ns = 0;
            if (sub_size[s] > 1)
                ns = (int64_t)IndexSys::comb_calc(k, sub_size[s]) *
                     IndexSys::comb_calc(sub_size[s], sub_size[main_idx[s]]);
            split_off[s + 1] = split_off[s] + ns;
        }
        for (int s = 0; s < nsub; ++s) {
            const int sz = sub_size[s];
            const int nc = IndexSys::comb_calc(k, sz);
            // i-th lexicographic set of size sz
            std::vector<int> set(sz);
            std::iota(set.begin(), set.end(), 0);
            const int mm = sz > 1 ? sub_size[main_idx[s]] : 0;
            const int P = sz > 1 ? IndexSys::comb_calc(sz, mm) : 0;
            for (int i = 0; i < nc; ++i) {
                uint32_t mask = 0;
                for (int x : set) mask |= 1u << x;
                int rank = 0;
                for (int q = 0; q < sz; ++q) rank += IndexSys::comb_calc(set[q], q + 1);
                if (comb_to_count[c2c_off[s] + i] != rank)
                    return set_err(c, PG_ERR_INVALID, "comb_to_count of subtemplate %d is not IndexSys's", s);
                if (sz > 1) {
                    std::vector<int> pos(mm);
                    std::iota(pos.begin(), pos.end(), 0);
                    for (int j = 0; j < P; ++j) {
                        uint32_t mmask = 0;
                        for (int q : pos) mmask |= 1u << set[q];
                        const uint32_t amask = mask & ~mmask;
                        int rm = 0, ra = 0, qm = 0, qa = 0;
                        for (int bit = 0; bit < k; ++bit) {
                            if (mmask >> bit & 1u) rm += IndexSys::comb_calc(bit, ++qm);
                            if (amask >> bit & 1u) ra += IndexSys::comb_calc(bit, ++qa);
                        }
                        const int64_t at = split_off[s] + (int64_t)i * P + j;
                        if (split_main[at] != rm || split_aux[at] != ra)
                            return set_err(c, PG_ERR_INVALID, "split tables of subtemplate %d are not IndexSys's", s);
                        int q = mm - 1;
                        while (q >= 0 && pos[q] == sz - mm + q) --q;
                        if (q < 0) break;
                        ++pos[q];
                        for (int r = q + 1; r < mm; ++r) pos[r] = pos[r - 1] + 1;
                    }
                }
                int q = sz - 1;
                while (q >= 0 && set[q] == k - sz + q) --q;
                if (q < 0) break;
                ++set[q];
                for (int r = q + 1; r < sz; ++r) set[r] = set[r - 1] + 1;
            }
        }
    }
    HIP_TRY(c, hipSetDevice(c->device));
    auto plan = std::make_unique<pg_plan>();
    plan->ctx = c;
    plan->device = c->device;
    plan->k = k;
    plan->K = k - 1;
    plan->nsub = nsub;
    plan->nonleaf = nonleaf;
    plan->sizes.assign(sub_size, sub_size + nsub);
    plan->mains.assign(main_idx, main_idx + nsub);
    plan->auxs.assign(aux_idx, aux_idx + nsub);
    plan->h_c2c.resize(nsub);
    plan->h_fsplit.resize(nsub);
    plan->d_c2c.resize(nsub);
    plan->d_fsplit.resize(nsub);
    {
        int64_t co = 0, so = 0;
        for (int s = 0; s < nsub; ++s) {
            const int nc = IndexSys::comb_calc(k, sub_size[s]);
            if (sub_size[s] > 1) {
                const int P = IndexSys::comb_calc(sub_size[s], sub_size[main_idx[s]]);
                plan->h_c2c[s].assign(comb_to_count + co, comb_to_count + co + nc);
                for (int64_t q = 0; q < (int64_t)nc * P; ++q)
                    plan->h_fsplit[s].push_back(((uint32_t)split_main[so + q] << 16) | (uint32_t)split_aux[so + q]);
                so += (int64_t)nc * P;
            }
            co += nc;
        }
    }
    const int K = k - 1;
    const std::vector<int16_t> rank = colex_rank_table(K);
    HIP_TRY(c, plan->rank_tab.alloc(rank.size() * 2));
    HIP_TRY(c, hipMemcpy(plan->rank_tab.p, rank.data(), rank.size() * 2, hipMemcpyHostToDevice));
    // the reference scales the first non-leaf it processes (highest index)
    // unless that is the root (CountMat.cpp:783-790, 811)
    int first_nonleaf = -1;
    for (int s = nsub - 1; s >= 0; --s)
        if (sub_size[s] > 1) { first_nonleaf = s; break; }
    plan->scaled_sub = first_nonleaf > 0 ? first_nonleaf : -1;
    for (int s = nsub - 1; s >= 0; --s) {
        if (sub_size[s] == 1) continue;
        auto d = std::make_unique<DevSub>();
        d->s = s;
        d->size = sub_size[s];
        d->mi = main_idx[s];
        d->ai = aux_idx[s];
        d->mm = sub_size[d->mi];
        d->aa = sub_size[d->ai];
        d->Wo = IndexSys::comb_calc(K, d->size - 1);
        d->Wy = IndexSys::comb_calc(K, d->aa);
        d->Wa = d->aa > 1 ? IndexSys::comb_calc(K, d->aa - 1) : k;
        d->Wm = d->mm > 1 ? IndexSys::comb_calc(K, d->mm - 1) : 0;
        d->P = IndexSys::comb_calc(d->size - 1, d->mm - 1);
        d->Cs = IndexSys::comb_calc(k, d->size);
        d->Ca = IndexSys::comb_calc(k, d->aa);
        d->Cm = IndexSys::comb_calc(k, d->mm);
        d->root = s == 0;
        d->scaled = s == plan->scaled_sub;
        // split table over output sets U (size-1 colors of the K others):
        // j-th lexicographic (mm-1)-subset of U's positions is the main part
        if (d->mm > 1) {
            const std::vector<uint16_t> outs = masks_of_size(K, d->size - 1);
            std::vector<uint32_t> tab;
            tab.reserve((size_t)d->Wo * d->P);
            const int nm = d->mm - 1, n1 = d->size - 1;
            for (uint16_t U : outs) {
                int el[kMaxK], ne = 0;
                for (int bit = 0; bit < K; ++bit)
                    if (U >> bit & 1) el[ne++] = bit;
                std::vector<int> pos(nm);
                std::iota(pos.begin(), pos.end(), 0);
                for (int j = 0; j < d->P; ++j) {
                    uint32_t mmask = 0;
                    for (int q : pos) mmask |= 1u << el[q];
                    const uint32_t amask = U & ~mmask;
                    tab.push_back(((uint32_t)rank[mmask] << 16) | (uint32_t)rank[amask]);
                    int q = nm - 1;
                    while (q >= 0 && pos[q] == n1 - nm + q) --q;
                    if (q < 0) break;
                    ++pos[q];
                    for (int r = q + 1; r < nm; ++r) pos[r] = pos[r - 1] + 1;
                }
            }
            HIP_TRY(c, d->split.alloc(tab.size() * 4));
            HIP_TRY(c, hipMemcpy(d->split.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
            // adjoint lists: every (output, split) pair filed under its aux
            // set and under its main set, in output order (equal lengths);
            // set-major for the flat adjoint kernel, and split-major (threads
            // on consecutive sets read consecutive words) for the staged one
            std::vector<std::vector<uint32_t>> by_a(d->Wy), by_m(d->Wm);
            for (int o = 0; o < d->Wo; ++o)
                for (int j = 0; j < d->P; ++j) {
                    const uint32_t e = tab[(size_t)o * d->P + j];
                    by_a[e & 0xffffu].push_back(((uint32_t)o << 16) | (e >> 16));
                    by_m[e >> 16].push_back(((uint32_t)o << 16) | (e & 0xffffu));
                }
            d->Pa = (int)by_a[0].size();
            d->Pm = (int)by_m[0].size();
            for (auto& l : by_a)
                if ((int)l.size() != d->Pa) return set_err(c, PG_ERR_INVALID, "uneven adjoint split lists");
            for (auto& l : by_m)
                if ((int)l.size() != d->Pm) return set_err(c, PG_ERR_INVALID, "uneven adjoint split lists");
            std::vector<uint32_t> fa, fm, fat((size_t)d->Pa * d->Wy), fmt((size_t)d->Pm * d->Wm);
            for (int x = 0; x < d->Wy; ++x) {
                fa.insert(fa.end(), by_a[x].begin(), by_a[x].end());
                for (int j = 0; j < d->Pa; ++j) fat[(size_t)j * d->Wy + x] = by_a[x][j];
            }
            for (int x = 0; x < d->Wm; ++x) {
                fm.insert(fm.end(), by_m[x].begin(), by_m[x].end());
                for (int j = 0; j < d->Pm; ++j) fmt[(size_t)j * d->Wm + x] = by_m[x][j];
            }
            const std::pair<DevBuf*, const std::vector<uint32_t>*> lists[] = {
                {&d->by_aux, &fa}, {&d->by_main, &fm}, {&d->by_auxT, &fat}, {&d->by_mainT, &fmt}};
            for (const auto& l : lists) {
                HIP_TRY(c, l.first->alloc(l.second->size() * 4));
                HIP_TRY(c, hipMemcpy(l.first->p, l.second->data(), l.second->size() * 4, hipMemcpyHostToDevice));
            }
        }
        if (d->aa > 1) {
            const std::vector<uint16_t> ym = masks_of_size(K, d->aa);
            HIP_TRY(c, d->ymask.alloc(ym.size() * 2));
            HIP_TRY(c, hipMemcpy(d->ymask.p, ym.data(), ym.size() * 2, hipMemcpyHostToDevice));
            const std::vector<uint16_t> am = masks_of_size(K, d->aa - 1);
            HIP_TRY(c, d->amask.alloc(am.size() * 2));
            HIP_TRY(c, hipMemcpy(d->amask.p, am.data(), am.size() * 2, hipMemcpyHostToDevice));
            d->Wy4 = (d->Wy + kRegCols - 1) / kRegCols * kRegCols;
            const size_t gbytes = (size_t)2 * (k + 1) * k * d->Wy4;
            if (gbytes <= kGidxMaxBytes) {
                const std::vector<int16_t> g = gather_index_table(k, ym, rank, d->Wy4);
                HIP_TRY(c, d->gidx.alloc(gbytes));
                HIP_TRY(c, hipMemcpy(d->gidx.p, g.data(), gbytes, hipMemcpyHostToDevice));
            }
            d->Wa4 = (d->Wa + 3) / 4 * 4;
            const size_t xbytes = (size_t)2 * k * k * d->Wa4;
            if (xbytes <= kGidxMaxBytes || runs_values_per_lane(d->Wa) > 0) {
                const std::vector<int16_t> x = rows_index_table(k, am, rank, d->Wa4);
                if (xbytes <= kGidxMaxBytes) {
                    HIP_TRY(c, d->xidx.alloc(xbytes));
                    HIP_TRY(c, hipMemcpy(d->xidx.p, x.data(), xbytes, hipMemcpyHostToDevice));
                }
                if (const int rv = runs_values_per_lane(d->Wa)) {
                    // per lane: value lane + 64 t of the row at [lane][t]
                    std::vector<int16_t> xl((size_t)k * k * 64 * rv, (int16_t)-1);
                    for (int cc = 0; cc < k * k; ++cc)
                        for (int lane = 0; lane < 64; ++lane)
                            for (int t = 0; t < rv; ++t) {
                                const int j = lane + 64 * t;
                                if (j < d->Wa) xl[((size_t)cc * 64 + lane) * rv + t] = x[(size_t)cc * d->Wa4 + j];
                            }
                    HIP_TRY(c, d->xidx_g.alloc(xl.size() * 2));
                    HIP_TRY(c, hipMemcpy(d->xidx_g.p, xl.data(), xl.size() * 2, hipMemcpyHostToDevice));
                }
            }
            const size_t tbytes = (size_t)2 * (k + 1) * k * d->Wa4;
            if (tbytes <= kGidxMaxBytes) {
                const std::vector<int16_t> g = gather_index_table_t(k, am, rank, d->Wa4);
                HIP_TRY(c, d->gidxT.alloc(tbytes));
                HIP_TRY(c, hipMemcpy(d->gidxT.p, g.data(), tbytes, hipMemcpyHostToDevice));
            }
            if (const int rv = runs_values_per_lane(d->Wy)) {
                const std::vector<int16_t> g = runs_tg_table(k, ym, rank, rv);
                HIP_TRY(c, d->ttab.alloc(g.size() * 2));
                HIP_TRY(c, hipMemcpy(d->ttab.p, g.data(), g.size() * 2, hipMemcpyHostToDevice));
            }
        }
        plan->subs.push_back(std::move(d));
    }
    *out = plan.release();
    return PG_OK;
}

extern "C" int pg_plan_from_edges(pg_ctx* c, int nv, int ne, const int32_t* src, const int32_t* dst, pg_plan** out)
{
    if (!c || !out || !src || !dst) return PG_ERR_INVALID;
    if (nv < 2 || nv > kMaxK || ne != nv - 1)
        return set_err(c, PG_ERR_INVALID, "template must be a tree with 2..%d vertices", kMaxK);
    for (int j = 0; j < ne; ++j)
        if (src[j] < 0 || src[j] >= nv || dst[j] < 0 || dst[j] >= nv)
            return set_err(c, PG_ERR_INVALID, "template vertex id out of range");
    Graph g;
    std::vector<int> s(src, src + ne), d(dst, dst + ne);
    g.build_graph(nv, ne, s.data(), d.data());
    FlatPlan fp;
    std::string err;
    int rc = flatten_plan(g, &fp, &err);
    if (rc) return set_err(c, rc, "%s", err.c_str());
    return pg_plan_create(c, fp.k, fp.nsub, fp.sub_size.data(), fp.main_idx.data(), fp.aux_idx.data(),
                          fp.comb_to_count.data(), fp.split_main.data(), fp.split_aux.data(), out);
}

extern "C" void pg_plan_destroy(pg_plan* p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->ctx) {
        drop_graphs(p->ctx, p);
        for (auto it = p->ctx->tuned.begin(); it != p->ctx->tuned.end();)
            it = std::get<0>(it->first) == p ? p->ctx->tuned.erase(it) : std::next(it);
    }
    delete p;
}

extern "C" int pg_plan_info(const pg_plan* p, int* k, int* nsub, int* n_nonleaf, int* scaled_sub)
{
    if (!p) return PG_ERR_INVALID;
    if (k) *k = p->k;
    if (nsub) *nsub = p->nsub;
    if (n_nonleaf) *n_nonleaf = p->nonleaf;
    if (scaled_sub) *scaled_sub = p->scaled_sub;
    return PG_OK;
}

// Device side of CountMat::estimatePeakMemUsage (CountMat.cpp:1067-1115):
// what a forward pass of B colorings holds, and the byte model of the
// per-kernel timing records (enqueue_pass / enqueue_step) per coloring.
extern "C" int pg_plan_footprint(pg_ctx* c, const pg_plan* p, int B, int precision, int64_t* device_bytes,
                                 double* algorithmic_bytes_per_coloring)
{
    if (!c) return PG_ERR_INVALID;
    if (!p || p->ctx != c || B <= 0) return set_err(c, PG_ERR_INVALID, "invalid footprint arguments");
    if (precision != PG_FP32 && precision != PG_FP64) return set_err(c, PG_ERR_INVALID, "precision must be PG_FP32 or PG_FP64");
    if (!c->has_graph) return set_err(c, PG_ERR_STATE, "no graph uploaded");
    int rc = build_buckets(c);
    if (rc) return rc;
    const size_t tsz = precision == PG_FP64 ? 8 : 4;
    if (device_bytes) {
        size_t b = plan_memory(c, p, B, tsz).total;
        const DevBuf* graph_bufs[] = {&c->d_rowptr, &c->d_col, &c->d_ent, &c->d_perm, &c->d_w32, &c->d_w64,
                                      &c->d_wsrc, &c->d_hub_items, &c->d_slot_begin, &p->rank_tab};
        for (const DevBuf* g : graph_bufs)
            b += g->bytes;
        for (auto& d : p->subs)
            for (const DevBuf* t : {&d->split, &d->ymask, &d->by_aux, &d->by_main, &d->by_auxT, &d->by_mainT,
                                    &d->amask, &d->gidx, &d->xidx, &d->gidxT})
                b += t->bytes;
        *device_bytes = (int64_t)b;
    }
    if (algorithmic_bytes_per_coloring) {
        const double na = (double)c->n_act;
        double bytes = 0;
        bool need_H = false;
        for (auto& d : p->subs) {
            need_H |= d->aa == 1;
            bytes += 4.0 * na * B * ((d->aa > 1 ? d->Ca : 0) + (d->mm > 1 ? d->Cm : 0) + (d->root ? 0 : d->Cs));
            if (d->aa > 1) bytes += csc_bytes(c);
        }
        if (need_H) bytes += csc_bytes(c) + na * B;
        *algorithmic_bytes_per_coloring = bytes / B;
    }
    return PG_OK;
}

// ------------------------------------------------------- unpruned mode
//
// PG_FLAG_UNPRUNED: the reference's isPruned = 0 path (countNonBottomeOriginal,
// CountMat.cpp:851-977) with its numerics, colorings one after another (each
// continues the accumulator the previous one left, as the reference's
// _bufVec does).  Absolute-set column-major FP32 tables of every non-leaf
// subtemplate live for the whole coloring; launches run eagerly.
int count_unpruned(pg_ctx* c, pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed, double* raw_out)
{
    const hipStream_t st = c->stream;
    const size_t na = (size_t)c->n_act;
    if (!c->fascia_acc_ready) {
        HIP_TRY(c, c->d_fascia_acc.alloc(na * sizeof(float)));
        HIP_TRY(c, hipMemsetAsync(c->d_fascia_acc.p, 0, na * sizeof(float), st));
        c->fascia_acc_ready = true;
    }
    for (auto& d : p->subs) {
        const int s = d->s;
        if (p->d_c2c[s]) continue;
        auto c2c = std::make_unique<DevBuf>(), sp = std::make_unique<DevBuf>();
        HIP_TRY(c, c2c->alloc(p->h_c2c[s].size() * 4));
        HIP_TRY(c, sp->alloc(p->h_fsplit[s].size() * 4));
        HIP_TRY(c, hipMemcpy(c2c->p, p->h_c2c[s].data(), p->h_c2c[s].size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(sp->p, p->h_fsplit[s].data(), p->h_fsplit[s].size() * 4, hipMemcpyHostToDevice));
        p->d_c2c[s] = std::move(c2c);
        p->d_fsplit[s] = std::move(sp);
    }
    // layout: tables by subtemplate index, root buffer, caller-order root
    // values, colors, parameters, the FP64 result
    std::vector<size_t> off_tab(p->nsub, 0);
    size_t total = 0;
    for (auto& d : p->subs) {
        off_tab[d->s] = total;
        total += align_up((d->root ? 1 : (size_t)d->Cs) * na * sizeof(float), 256);
    }
    const size_t off_rootc = total;
    total += align_up((size_t)c->n * sizeof(float), 256);
    const size_t off_colors = total;
    total += align_up(na, 256);
    const size_t off_params = total;
    total += align_up(sizeof(PassParams), 256);
    const size_t off_out = total;
    total += 256;
    HIP_TRY(c, c->d_fascia.alloc(total));
    char* base = c->d_fascia.as<char>();
    const uint8_t* d_in = nullptr;
    if (colors) {
        if ((size_t)ncol * c->n > c->d_colors_in.bytes) {
            drop_graphs(c);  // captured passes hold the former colors buffer
            HIP_TRY(c, c->d_colors_in.alloc((size_t)ncol * c->n));
        }
        HIP_TRY(c, hipMemcpyAsync(c->d_colors_in.p, colors, (size_t)ncol * c->n, hipMemcpyHostToDevice, st));
        d_in = c->d_colors_in.as<uint8_t>();
    }
    uint8_t* d_colors = (uint8_t*)(base + off_colors);
    const PassParams* pp = (const PassParams*)(base + off_params);
    c->agg.clear();
    c->agg_sub.clear();
    std::vector<LaunchRec> recs;
    c->cur_recs = &recs;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(c, hipEventCreate(&t0));
    HIP_TRY(c, hipEventCreate(&t1));
    HIP_TRY(c, hipEventRecord(t0, st));
    int rc = PG_OK;
    for (int i = 0; i < ncol && rc == PG_OK; ++i) {
        const PassParams hp{seed, i};
        HIP_TRY(c, hipMemcpyAsync(base + off_params, &hp, sizeof hp, hipMemcpyHostToDevice, st));
        record(c, "k_colors", 0, true);
        HIP_TRY(c, d_in ? launch_colors_gather(d_colors, d_in, c->d_perm.as<int32_t>(), c->n_act, 1, c->n, pp, st)
                        : launch_colors_rng(d_colors, c->d_perm.as<int32_t>(), c->n_act, 1, pp, p->k, st));
        record(c, "", 0, false);
        for (auto& dp : p->subs) {
            const DevSub& d = *dp;
            const SubTag tag(c, d.s);
            FasciaArgs a{};
            a.n_act = c->n_act;
            a.rowptr = c->d_rowptr.as<int32_t>();
            a.col = c->d_col.as<int32_t>();
            a.colors = d_colors;
            a.acc = c->d_fascia_acc.as<float>();
            a.Cs = d.Cs;
            a.P = IndexSys::comb_calc(d.size, d.mm);
            a.c2c = p->d_c2c[d.s]->as<int32_t>();
            a.split = p->d_fsplit[d.s]->as<uint32_t>();
            a.main_leaf = d.mm == 1;
            a.aux_leaf = d.aa == 1;
            a.tm = d.mm > 1 ? (const float*)(base + off_tab[d.mi]) : nullptr;
            a.ta = d.aa > 1 ? (const float*)(base + off_tab[d.ai]) : nullptr;
            a.out = (float*)(base + off_tab[d.s]);
            // the reference's Fascia byte model per pair: the aux column, the
            // main column and the output (estimateMemCommFascia)
            record(c, "k_fascia_step", 4.0 * (double)a.Cs * a.P * (double)(c->nnz + 3 * na), true);
            HIP_TRY(c, launch_fascia_step(a, c->num_cus, st));
            record(c, "", 0, false);
            if (d.root) {
                record(c, "k_fascia_sum", 4.0 * (double)(na + c->n), true);
                HIP_TRY(c, launch_fascia_root_sum(a.out, c->d_perm.as<int32_t>(), c->n_act,
                                                  (float*)(base + off_rootc), c->n, (double*)(base + off_out), st));
                record(c, "", 0, false);
            }
        }
        // CountMat.cpp:952-955: the FP64 sum is the raw count (no 1e12 here)
        HIP_TRY(c, hipMemcpyAsync(&raw_out[i], base + off_out, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
    }
    c->cur_recs = nullptr;
    (void)hipEventRecord(t1, st);
    (void)hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    c->last_total_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    accumulate_timing(c, recs);
    destroy_recs(recs);
    finish_timing(c);
    c->last_tile = 1;
    return rc;
}

extern "C" int pg_count(pg_ctx* c, const pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed, int precision,
                        uint32_t flags, double* raw_out)
{
    if (!c) return PG_ERR_INVALID;
    if (!p || p->ctx != c || ncol <= 0 || !raw_out) return set_err(c, PG_ERR_INVALID, "invalid count arguments");
    if (precision != PG_FP32 && precision != PG_FP64)
        return set_err(c, PG_ERR_INVALID, "precision must be PG_FP32 or PG_FP64");
    if (!c->has_graph) return set_err(c, PG_ERR_STATE, "no graph uploaded");
    if (colors) {
        const size_t total = (size_t)ncol * c->n;
        for (size_t i = 0; i < total; ++i)
            if (colors[i] >= p->k) return set_err(c, PG_ERR_INVALID, "color %d >= k at %zu", colors[i], i);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->n_act == 0) {
        for (int i = 0; i < ncol; ++i) raw_out[i] = 0.0;
        c->last_total_ms = 0;
        return PG_OK;
    }
    if (flags & PG_FLAG_UNPRUNED) {
        if (precision != PG_FP32)
            return set_err(c, PG_ERR_INVALID, "the unpruned mode has the reference's FP32 numerics only");
        return count_unpruned(c, const_cast<pg_plan*>(p), ncol, colors, seed, raw_out);
    }
    return precision == PG_FP32 ? count_impl<float>(c, p, ncol, colors, seed, flags, raw_out)
                                : count_impl<double>(c, p, ncol, colors, seed, flags, raw_out);
}

// pg_count's passes with every buffer in the static scratch: eager launches
// (no captured graphs, no events), no device or pinned allocation.
template <typename T>
int count_noalloc_impl(pg_ctx* c, const pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed, uint32_t flags,
                       double* raw_out)
{
    if (c->built_deg != c->hub_degree || c->built_chunk != c->hub_chunk || c->built_high != c->bucket_high ||
        c->built_mid != c->bucket_mid)
        return set_err(c, PG_ERR_STATE,
                       "degree-bucket options changed since the graph upload; their work lists are built by the next "
                       "allocating call (pg_count), not by pg_count_noalloc");
    const size_t colors_bytes = colors ? align_up((size_t)ncol * c->n, 256) : 0;
    const int cap = std::min(64, c->batch_tile > 0 ? (int)std::min<int64_t>(c->batch_tile, ncol) : ncol);
    int tile = 0;
    MemPlan mp;
    size_t need1 = 0;
    for (int B = cap; B >= 1; --B) {
        size_t lds = 0;
        for (auto& d : p->subs) lds = std::max(lds, sub_lds_max(c, p, *d, B, sizeof(T)));
        if (lds > kLdsLimit) continue;
        const MemPlan m = plan_memory(c, p, B, sizeof(T));
        if (B == 1) need1 = m.total + colors_bytes;
        if (m.total + colors_bytes <= kScratchBytes) {
            tile = B;
            mp = m;
            break;
        }
    }
    if (tile == 0)
        return set_err(c, PG_ERR_RANGE,
                       "pg_count_noalloc: one coloring of this graph and template needs %zu bytes of device "
                       "scratch (with the caller's colors), beyond the static %zu",
                       need1, kScratchBytes);
    const hipStream_t st = c->stream;
    const uint8_t* d_in = nullptr;
    if (colors) {
        // pageable source: the copy is complete when the call returns
        HIP_TRY(c, hipMemcpyAsync(c->scratch + mp.total, colors, (size_t)ncol * c->n, hipMemcpyHostToDevice, st));
        d_in = reinterpret_cast<const uint8_t*>(c->scratch + mp.total);
    }
    struct Override {
        pg_ctx* c;
        int64_t timing;
        Override(pg_ctx* ctx) : c(ctx), timing(ctx->timing)
        {
            c->arena_override = c->scratch;
            c->timing = 0;
        }
        ~Override()
        {
            c->arena_override = nullptr;
            c->timing = timing;
        }
    } guard(c);
    c->last_tile = tile;
    double raw[64];
    for (int first = 0; first < ncol; first += tile) {
        const int B = std::min(tile, ncol - first);
        const MemPlan m = B == tile ? mp : plan_memory(c, p, B, sizeof(T));
        const PassParams hp{seed, first};
        HIP_TRY(c, hipMemcpyAsync(c->scratch + m.off_params, &hp, sizeof hp, hipMemcpyHostToDevice, st));
        int rc = enqueue_pass<T>(c, p, B, d_in, m);
        if (rc) return rc;
        HIP_TRY(c, hipMemcpyAsync(raw, c->scratch + m.off_rawdev, sizeof(double) * B, hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
        for (int b = 0; b < B; ++b) raw_out[first + b] = finish_raw(p, sizeof(T) == 4, flags, raw[b]);
    }
    return PG_OK;
}

extern "C" int pg_count_noalloc(pg_ctx* c, const pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed,
                                int precision, uint32_t flags, double* raw_out)
{
    if (!c) return PG_ERR_INVALID;
    if (!p || p->ctx != c || ncol <= 0 || !raw_out) return set_err(c, PG_ERR_INVALID, "invalid count arguments");
    if (precision != PG_FP32 && precision != PG_FP64)
        return set_err(c, PG_ERR_INVALID, "precision must be PG_FP32 or PG_FP64");
    if (flags & ~PG_FLAG_FIX_2VERTEX) return set_err(c, PG_ERR_INVALID, "pg_count_noalloc takes PG_FLAG_FIX_2VERTEX only");
    if (!c->has_graph) return set_err(c, PG_ERR_STATE, "no graph uploaded");
    if (!c->scratch) return set_err(c, PG_ERR_STATE, "no static scratch on this device");
    if (colors) {
        const size_t total = (size_t)ncol * c->n;
        for (size_t i = 0; i < total; ++i)
            if (colors[i] >= p->k) return set_err(c, PG_ERR_INVALID, "color %d >= k at %zu", colors[i], i);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->n_act == 0) {
        for (int i = 0; i < ncol; ++i) raw_out[i] = 0.0;
        return PG_OK;
    }
    return precision == PG_FP32 ? count_noalloc_impl<float>(c, p, ncol, colors, seed, flags, raw_out)
                                : count_noalloc_impl<double>(c, p, ncol, colors, seed, flags, raw_out);
}

extern "C" int pg_count_multi(pg_ctx* c, int nplans, const pg_plan* const* plans, const int32_t* ncolorings,
                              const uint64_t* seeds, int precision, uint32_t flags, double* raw_out)
{
    if (!c) return PG_ERR_INVALID;
    if (nplans <= 0 || !plans || !ncolorings || !seeds || !raw_out)
        return set_err(c, PG_ERR_INVALID, "invalid multi-count arguments");
    for (int i = 0; i < nplans; ++i) {
        if (!plans[i] || plans[i]->ctx != c || ncolorings[i] <= 0)
            return set_err(c, PG_ERR_INVALID, "invalid plan or coloring count at %d", i);
    }
    if (precision != PG_FP32 && precision != PG_FP64)
        return set_err(c, PG_ERR_INVALID, "precision must be PG_FP32 or PG_FP64");
    if (!c->has_graph) return set_err(c, PG_ERR_STATE, "no graph uploaded");
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->n_act > 0) {
        bool done = false;
        int rc = precision == PG_FP32
                     ? count_multi_impl<float>(c, nplans, plans, ncolorings, seeds, flags, raw_out, &done)
                     : count_multi_impl<double>(c, nplans, plans, ncolorings, seeds, flags, raw_out, &done);
        if (rc) return rc;
        if (done) return PG_OK;
    }
    int64_t off = 0;
    double total = 0;
    for (int i = 0; i < nplans; ++i) {
        int rc = pg_count(c, plans[i], ncolorings[i], nullptr, seeds[i], precision, flags, raw_out + off);
        if (rc) return rc;
        total += c->last_total_ms;
        off += ncolorings[i];
    }
    c->last_total_ms = total;
    return PG_OK;
}

extern "C" int pg_count_grad(pg_ctx* c, const pg_plan* p, int ncol, const uint8_t* colors, uint64_t seed,
                             int precision, uint32_t flags, const double* cotangent, double* raw_out, void* grad,
                             int grad_on_device)
{
    if (!c) return PG_ERR_INVALID;
    if (!p || p->ctx != c || ncol <= 0 || !raw_out)
        return set_err(c, PG_ERR_INVALID, "invalid gradient arguments");
    if (precision != PG_FP32 && precision != PG_FP64)
        return set_err(c, PG_ERR_INVALID, "precision must be PG_FP32 or PG_FP64");
    if (!c->has_graph) return set_err(c, PG_ERR_STATE, "no graph uploaded");
    if (colors) {
        const size_t total = (size_t)ncol * c->n;
        for (size_t i = 0; i < total; ++i)
            if (colors[i] >= p->k) return set_err(c, PG_ERR_INVALID, "color %d >= k at %zu", colors[i], i);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->n_act == 0) {
        for (int i = 0; i < ncol; ++i) raw_out[i] = 0.0;
        if (grad && !grad_on_device)
            for (int64_t e = 0; e < c->nnz; ++e) ((double*)grad)[e] = 0.0;
        return PG_OK;
    }
    return precision == PG_FP32
               ? count_grad_impl<float>(c, p, ncol, colors, seed, flags, cotangent, raw_out, grad, grad_on_device)
               : count_grad_impl<double>(c, p, ncol, colors, seed, flags, cotangent, raw_out, grad, grad_on_device);
}

extern "C" int pg_last_timing(const pg_ctx* c, double* total_ms, double* spmm_ms, int64_t* spmm_launches,
                              double* spmm_bytes)
{
    if (!c) return PG_ERR_INVALID;
    if (total_ms) *total_ms = c->last_total_ms;
    if (spmm_ms) *spmm_ms = c->dom_ms;
    if (spmm_launches) *spmm_launches = c->dom_launches;
    if (spmm_bytes) *spmm_bytes = c->dom_bytes;
    return PG_OK;
}

extern "C" int pg_last_timing_kernel(const pg_ctx* c, char* buf, int len)
{
    if (!c || !buf || len <= 0) return PG_ERR_INVALID;
    snprintf(buf, (size_t)len, "%s", c->dom_name.c_str());
    return PG_OK;
}

extern "C" int pg_last_timing_sub(const pg_ctx* c, int idx, int* sub, char* name, int len, double* ms,
                                  int64_t* launches, double* bytes)
{
    if (!c || idx < 0 || (name && len <= 0)) return PG_ERR_INVALID;
    if ((size_t)idx >= c->agg_sub.size()) return PG_ERR_RANGE;
    auto it = c->agg_sub.begin();
    std::advance(it, idx);
    if (sub) *sub = it->first.first;
    if (name) snprintf(name, (size_t)len, "%s", it->first.second.c_str());
    if (ms) *ms = it->second.ms;
    if (launches) *launches = it->second.launches;
    if (bytes) *bytes = it->second.bytes;
    return (int)c->agg_sub.size();
}

extern "C" int pg_last_timing_entry(const pg_ctx* c, int idx, char* name, int len, double* ms,
                                    int64_t* launches, double* bytes)
{
    if (!c || idx < 0 || (name && len <= 0)) return PG_ERR_INVALID;
    if ((size_t)idx >= c->agg.size()) return PG_ERR_RANGE;
    auto it = c->agg.begin();
    std::advance(it, idx);
    if (name) snprintf(name, (size_t)len, "%s", it->first.c_str());
    if (ms) *ms = it->second.ms;
    if (launches) *launches = it->second.launches;
    if (bytes) *bytes = it->second.bytes;
    return (int)c->agg.size();
}
