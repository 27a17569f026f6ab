// Template graph, subtemplate split and color-set indexing (host side).
// Reference: Graph.cpp:9-129, DivideTemplates.cpp:6-327, IndexSys.cpp:9-436.
#include "pg_host.hpp"
#include "pgbsc.h"

#include <algorithm>
#include <fstream>
#include <functional>
#include <numeric>

namespace pgbsc {

static int fail(std::string* err, const std::string& msg, int code)
{
    if (err) *err = msg;
    return code;
}

// ---------------------------------------------------------------- Graph ---

void Graph::build_graph(int verts, int edges, const int* src, const int* dst)
{
    nv_ = verts;
    src_.assign(src, src + edges);
    dst_.assign(dst, dst + edges);
    off_.assign(verts + 1, 0);
    for (int j = 0; j < edges; ++j) { off_[src[j] + 1]++; off_[dst[j] + 1]++; }
    std::partial_sum(off_.begin(), off_.end(), off_.begin());
    adj_.assign(2 * (size_t)edges, 0);
    std::vector<int> pos(off_.begin(), off_.end() - 1);
    // Graph.cpp:111-115: each edge appends dst to src's list, then src to dst's
    for (int j = 0; j < edges; ++j) {
        adj_[pos[src[j]]++] = dst[j];
        adj_[pos[dst[j]]++] = src[j];
    }
}

int Graph::read_enlist(const std::string& path, std::string* err)
{
    std::ifstream in(path);
    if (!in) return fail(err, "cannot open template file " + path, PG_ERR_IO);
    std::string line;
    if (!std::getline(in, line)) return fail(err, "template file has no vertex count", PG_ERR_INVALID);
    int verts = atoi(line.c_str());
    if (!std::getline(in, line)) return fail(err, "template file has no edge count", PG_ERR_INVALID);
    int edges = atoi(line.c_str());
    if (verts < 2 || verts > kMaxTemplate || edges != verts - 1)
        return fail(err, "template must be a tree with 2.." + std::to_string(kMaxTemplate) +
                             " vertices and n-1 edges", PG_ERR_INVALID);
    std::vector<int> s(edges), d(edges);
    int max_id = 0;
    for (int j = 0; j < edges; ++j) {
        // same tokenisation as Graph.cpp:26-36: up to a space, then to end of line
        if (!std::getline(in, line, ' ')) return fail(err, "template edge list truncated", PG_ERR_INVALID);
        s[j] = atoi(line.c_str());
        if (!std::getline(in, line)) return fail(err, "template edge list truncated", PG_ERR_INVALID);
        d[j] = atoi(line.c_str());
        if (s[j] < 0 || d[j] < 0) return fail(err, "negative template vertex id", PG_ERR_INVALID);
        max_id = std::max(max_id, std::max(s[j], d[j]));
    }
    if (max_id != verts - 1) {
        // renumber the ids that occur, ascending; n stays as declared
        std::vector<int> id(max_id + 1, -1);
        for (int j = 0; j < edges; ++j) id[s[j]] = id[d[j]] = 0;
        int next = 0;
        for (auto& x : id) if (x == 0) x = next++;
        for (int j = 0; j < edges; ++j) { s[j] = id[s[j]]; d[j] = id[d[j]]; }
        if (next > verts) return fail(err, "template ids exceed the vertex count", PG_ERR_INVALID);
    }
    build_graph(verts, edges, s.data(), d.data());
    return PG_OK;
}

// ------------------------------------------------------ DivideTemplates ---

static bool is_tree(const Graph& g)
{
    int n = g.get_vert_num();
    if (n < 1 || g.get_edge_num() != 2 * (n - 1)) return false;
    std::vector<char> seen(n, 0);
    std::vector<int> q{0};
    seen[0] = 1;
    for (size_t i = 0; i < q.size(); ++i)
        for (int j = 0; j < g.get_out_deg(q[i]); ++j) {
            int v = g.get_adj_list(q[i])[j];
            if (v < 0 || v >= n) return false;
            if (!seen[v]) { seen[v] = 1; q.push_back(v); }
        }
    return (int)q.size() == n;
}

// One side of a cut: breadth-first from `root` in adjacency order, never
// entering `excluded`; vertices are renumbered in discovery order (root = 0)
// and edges kept in discovery order (DivideTemplates.cpp:153-246).
int DivideTemplates::cut(int sub, int root, int excluded)
{
    const Graph g = subs_[sub];  // copy: subs_ grows below
    int n = g.get_vert_num();
    std::vector<int> label(n, -1), order{root}, es, ed;
    label[root] = 0;
    for (size_t i = 0; i < order.size(); ++i) {
        int a = order[i];
        for (int j = 0; j < g.get_out_deg(a); ++j) {
            int v = g.get_adj_list(a)[j];
            if (v == excluded || label[v] >= 0) continue;
            label[v] = (int)order.size();
            order.push_back(v);
            es.push_back(label[a]);
            ed.push_back(label[v]);
        }
    }
    Graph part;
    part.build_graph((int)order.size(), (int)es.size(), es.data(), ed.data());
    subs_.push_back(part);
    main_.push_back(kDummy);
    aux_.push_back(kDummy);
    parent_.push_back(kDummy);
    return 0;  // the part's root is always relabelled 0
}

void DivideTemplates::divide(int sub, int root)
{
    int nb = subs_[sub].get_adj_list(root)[0];
    int root_main = cut(sub, root, nb);
    int root_aux = cut(sub, nb, root);
    int mi = (int)subs_.size() - 2, ai = (int)subs_.size() - 1;
    main_[sub] = mi;
    aux_[sub] = ai;
    parent_[mi] = parent_[ai] = sub;
    if (subs_[mi].get_vert_num() > 1) divide(mi, root_main);
    if (subs_[ai].get_vert_num() > 1) divide(ai, root_aux);
}

int DivideTemplates::DivideTp(const Graph& tp, std::string* err)
{
    if (tp.get_vert_num() < 2 || tp.get_vert_num() > kMaxTemplate || !is_tree(tp))
        return fail(err, "template must be a tree with 2.." + std::to_string(kMaxTemplate) + " vertices",
                    PG_ERR_INVALID);
    subs_.assign(1, tp);
    main_.assign(1, kDummy);
    aux_.assign(1, kDummy);
    parent_.assign(1, kDummy);
    divide(0, 0);
    return PG_OK;
}

// Adjacent swaps while a subtemplate's parent index is smaller than its left
// neighbour's (positions 1.., the root stays first); every link to the two
// swapped positions is renamed (DivideTemplates.cpp:264-327).
void DivideTemplates::sort_tps()
{
    const int S = (int)subs_.size();
    auto rename = [](int x, int a, int b) { return x == a ? b : (x == b ? a : x); };
    for (bool swapped = true; swapped;) {
        swapped = false;
        for (int i = 2; i < S; ++i) {
            if (parent_[i] >= parent_[i - 1]) continue;
            const int a = i - 1, b = i;
            for (int x = 0; x < S; ++x) {
                if (main_[x] != kDummy) main_[x] = rename(main_[x], a, b);
                if (aux_[x] != kDummy) aux_[x] = rename(aux_[x], a, b);
                if (parent_[x] != kDummy) parent_[x] = rename(parent_[x], a, b);
            }
            std::swap(subs_[a], subs_[b]);
            std::swap(main_[a], main_[b]);
            std::swap(aux_[a], aux_[b]);
            std::swap(parent_[a], parent_[b]);
            swapped = true;
        }
    }
}

// ------------------------------------------------------------- IndexSys ---

int IndexSys::comb_calc(int n, int k)
{
    if (k < 0 || k > n) return 0;
    if (2 * k > n) k = n - k;
    long long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return (int)r;
}

// all `len`-subsets of {0..val-1} in lexicographic order, row-major
static std::vector<int> lex_subsets(int val, int len)
{
    std::vector<int> out, cur(len);
    std::iota(cur.begin(), cur.end(), 0);
    int cnt = IndexSys::comb_calc(val, len);
    out.reserve((size_t)cnt * len);
    for (int r = 0; r < cnt; ++r) {
        out.insert(out.end(), cur.begin(), cur.end());
        int i = len - 1;
        while (i >= 0 && cur[i] == val - len + i) --i;
        if (i < 0) break;
        ++cur[i];
        for (int j = i + 1; j < len; ++j) cur[j] = cur[j - 1] + 1;
    }
    return out;
}

static int colex_rank(const int* set, int len)
{
    int r = 0;
    for (int i = 0; i < len; ++i) r += IndexSys::comb_calc(set[i], i + 1);
    return r;
}

void IndexSys::initialization(int color_num, const DivideTemplates& div)
{
    k_ = color_num;
    const int S = div.get_subtps_num();
    sizes_.assign(S, 0);
    nsplit_.assign(S, 0);
    c2c_.assign(S, {});
    split_main_.assign(S, {});
    split_aux_.assign(S, {});
    for (int s = 0; s < S; ++s) sizes_[s] = div.get_subtp(s).get_vert_num();
    for (int s = 0; s < S; ++s) {
        const int sz = sizes_[s];
        const std::vector<int> sets = lex_subsets(k_, sz);
        const int nc = comb_calc(k_, sz);
        c2c_[s].resize(nc);
        for (int i = 0; i < nc; ++i) c2c_[s][i] = colex_rank(&sets[(size_t)i * sz], sz);
        if (sz == 1) continue;
        const int mm = div.get_main_node_vert_num(s), aa = div.get_aux_node_vert_num(s);
        const int P = comb_calc(sz, mm);
        nsplit_[s] = P;
        const std::vector<int> pm = lex_subsets(sz, mm), pa = lex_subsets(sz, aa);
        split_main_[s].resize((size_t)nc * P);
        split_aux_[s].resize((size_t)nc * P);
        int mset[kMaxTemplate], aset[kMaxTemplate];
        for (int i = 0; i < nc; ++i) {
            const int* set = &sets[(size_t)i * sz];
            for (int j = 0; j < P; ++j) {
                for (int q = 0; q < mm; ++q) mset[q] = set[pm[(size_t)j * mm + q]];
                for (int q = 0; q < aa; ++q) aset[q] = set[pa[(size_t)(P - 1 - j) * aa + q]];
                split_main_[s][(size_t)i * P + j] = colex_rank(mset, mm);
                split_aux_[s][(size_t)i * P + j] = colex_rank(aset, aa);
            }
        }
    }
}

int flatten_plan(const Graph& tmpl, FlatPlan* out, std::string* err)
{
    DivideTemplates div;
    int rc = div.DivideTp(tmpl, err);
    if (rc) return rc;
    div.sort_tps();
    IndexSys idx;
    idx.initialization(tmpl.get_vert_num(), div);
    FlatPlan p;
    p.k = tmpl.get_vert_num();
    p.nsub = div.get_subtps_num();
    for (int s = 0; s < p.nsub; ++s) {
        int sz = idx.getSubsSize(s);
        p.sub_size.push_back(sz);
        p.main_idx.push_back(sz > 1 ? div.get_main_node_idx(s) : -1);
        p.aux_idx.push_back(sz > 1 ? div.get_aux_node_idx(s) : -1);
        const auto& c = idx.getCombToCount(s);
        p.comb_to_count.insert(p.comb_to_count.end(), c.begin(), c.end());
        if (sz > 1) {
            const auto& m = idx.getSplitMain(s);
            const auto& a = idx.getSplitAux(s);
            p.split_main.insert(p.split_main.end(), m.begin(), m.end());
            p.split_aux.insert(p.split_aux.end(), a.begin(), a.end());
        }
    }
    *out = std::move(p);
    return PG_OK;
}

}  // namespace pgbsc

static int template_plan(int nv, int ne, const int32_t* src, const int32_t* dst, pgbsc::FlatPlan* fp)
{
    if (nv < 2 || nv > pgbsc::kMaxTemplate || ne != nv - 1 || !src || !dst) return PG_ERR_INVALID;
    for (int j = 0; j < ne; ++j)
        if (src[j] < 0 || src[j] >= nv || dst[j] < 0 || dst[j] >= nv) return PG_ERR_INVALID;
    pgbsc::Graph g;
    std::vector<int> s(src, src + ne), d(dst, dst + ne);
    g.build_graph(nv, ne, s.data(), d.data());
    return pgbsc::flatten_plan(g, fp);
}

extern "C" int pg_template_sizes(int nv, int ne, const int32_t* src, const int32_t* dst, int32_t* nsub,
                                 int64_t* n_comb, int64_t* n_split)
{
    pgbsc::FlatPlan fp;
    int rc = template_plan(nv, ne, src, dst, &fp);
    if (rc) return rc;
    if (nsub) *nsub = fp.nsub;
    if (n_comb) *n_comb = (int64_t)fp.comb_to_count.size();
    if (n_split) *n_split = (int64_t)fp.split_main.size();
    return PG_OK;
}

extern "C" int pg_template_tables(int nv, int ne, const int32_t* src, const int32_t* dst, int32_t* sub_size,
                                  int32_t* main_idx, int32_t* aux_idx, int32_t* comb_to_count,
                                  int32_t* split_main, int32_t* split_aux)
{
    pgbsc::FlatPlan fp;
    int rc = template_plan(nv, ne, src, dst, &fp);
    if (rc) return rc;
    if (!sub_size || !main_idx || !aux_idx || !comb_to_count || !split_main || !split_aux) return PG_ERR_INVALID;
    std::copy(fp.sub_size.begin(), fp.sub_size.end(), sub_size);
    std::copy(fp.main_idx.begin(), fp.main_idx.end(), main_idx);
    std::copy(fp.aux_idx.begin(), fp.aux_idx.end(), aux_idx);
    std::copy(fp.comb_to_count.begin(), fp.comb_to_count.end(), comb_to_count);
    std::copy(fp.split_main.begin(), fp.split_main.end(), split_main);
    std::copy(fp.split_aux.begin(), fp.split_aux.end(), split_aux);
    return PG_OK;
}

// CountMat::automorphismNum (CountMat.cpp:1498-1572).  The reference tries
// all k! permutations s (calcAutomorphismRecursive builds mappingID, s(i) =
// mappingID[i]) and accepts s when, for every vertex i, deg(i) = deg(s(i))
// and every neighbour u of i equals s(w) for some neighbour w of s(i)
// (calcAutomorphismZero, :1510-1544): N(i) is a subset of s(N(s(i))).  That
// is not always the automorphism group (on a 3-leg spider it accepts some
// non-automorphisms and rejects others), so the same predicate is counted.
// Backtracking assigns s(0), s(1), ... and tests vertex i as soon as s is
// known on i and on N(s(i)); a branch is cut only when a fully determined
// test fails, so the count is the reference's over all k! permutations.
extern "C" int64_t pg_template_automorphisms(int nv, int ne, const int32_t* src, const int32_t* dst)
{
    if (nv < 1 || nv > pgbsc::kMaxTemplate || ne < 0 || (ne > 0 && (!src || !dst))) return PG_ERR_INVALID;
    for (int j = 0; j < ne; ++j)
        if (src[j] < 0 || src[j] >= nv || dst[j] < 0 || dst[j] >= nv) return PG_ERR_INVALID;
    pgbsc::Graph g;
    std::vector<int> sv(src, src + ne), dv(dst, dst + ne);
    g.build_graph(nv, ne, sv.data(), dv.data());
    const int n = nv;
    std::vector<int> map(n, -1), reach(n, 0);
    std::vector<char> used(n, 0), done(n, 0);
    // reach[c]: the largest neighbour of c, so s is known on N(c) once s(0..reach[c]) are
    for (int c = 0; c < n; ++c)
        for (int j = 0; j < g.get_out_deg(c); ++j) reach[c] = std::max(reach[c], g.get_adj_list(c)[j]);
    auto holds = [&](int i) {
        const int c = map[i];
        for (int j = 0; j < g.get_out_deg(i); ++j) {
            const int u = g.get_adj_list(i)[j];
            bool hit = false;
            for (int q = 0; q < g.get_out_deg(c) && !hit; ++q) hit = map[g.get_adj_list(c)[q]] == u;
            if (!hit) return false;
        }
        return true;
    };
    int64_t count = 0;
    std::function<void(int)> rec = [&](int d) {
        if (d == n) { ++count; return; }
        for (int c = 0; c < n; ++c) {
            if (used[c] || g.get_out_deg(c) != g.get_out_deg(d)) continue;
            used[c] = 1;
            map[d] = c;
            std::vector<int> now;  // tests made at this depth
            bool ok = true;
            for (int i = 0; i <= d && ok; ++i)
                if (!done[i] && reach[map[i]] <= d) {
                    ok = holds(i);
                    if (ok) { done[i] = 1; now.push_back(i); }
                }
            if (ok) rec(d + 1);
            for (int i : now) done[i] = 0;
            map[d] = -1;
            used[c] = 0;
        }
    };
    rec(0);
    return count;
}
