// CountMat driver over the libpgbsc C ABI (reference: CountMat.cpp:21-266,
// 979-999, 1498-1572).
#include "countmat.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>

namespace pgbsc {

CountMat::~CountMat()
{
    if (ctx_) pg_destroy(ctx_);
}

int CountMat::initialization(const CSCGraph* g, int thd_num, int itr_num, int isPruned, int useSPMM, int,
                             bool calculate_automorphisms, int device, int precision, uint64_t seed)
{
    if (!g || itr_num <= 0) {
        err_ = "initialization needs a graph and itr_num > 0";
        return PG_ERR_INVALID;
    }
    (void)isPruned;
    (void)useSPMM;
    itr_ = itr_num;
    thd_ = thd_num;
    autom_ = calculate_automorphisms;
    precision_ = precision;
    seed_ = seed;
    if (ctx_) pg_destroy(ctx_);
    ctx_ = nullptr;
    int rc = pg_create(&ctx_, device);
    if (rc) {
        err_ = "no usable HIP device";
        return rc;
    }
    rc = pg_graph_upload(ctx_, g->getNumVertices(), g->getNNZ(), g->getIndexCol(), g->getIndexRow(), nullptr);
    if (rc) err_ = pg_last_error(ctx_);
    return rc;
}

static int factorial(int n) { return n <= 1 ? 1 : n * factorial(n - 1); }

double CountMat::compute(const Graph& t, bool isEstimate)
{
    tmpl_ = &t;
    if (!ctx_) {
        err_ = "compute before initialization";
        return PG_ERR_STATE;
    }
    FlatPlan fp;
    int rc = flatten_plan(t, &fp, &err_);
    if (rc) return rc;
    pg_plan* plan = nullptr;
    rc = pg_plan_create(ctx_, fp.k, fp.nsub, fp.sub_size.data(), fp.main_idx.data(), fp.aux_idx.data(),
                        fp.comb_to_count.data(), fp.split_main.data(), fp.split_aux.data(), &plan);
    if (rc) {
        err_ = pg_last_error(ctx_);
        return rc;
    }
    if (isEstimate) {
        pg_plan_destroy(plan);
        return 0.0;
    }
    raw_.assign(itr_, 0.0);
    auto t0 = std::chrono::steady_clock::now();
    rc = pg_count(ctx_, plan, itr_, nullptr, seed_, precision_, 0u, raw_.data());
    auto t1 = std::chrono::steady_clock::now();
    pg_plan_destroy(plan);
    if (rc) {
        err_ = pg_last_error(ctx_);
        return rc;
    }
    sec_per_col_ = std::chrono::duration<double>(t1 - t0).count() / itr_;
    printf("\nTime for count per iter: %9.6lf seconds\n", sec_per_col_);
    double mean = 0;
    for (double r : raw_) mean += r;
    mean /= itr_;
    const int k = t.get_vert_num();
    // CountMat.cpp:248-259: colorful probability k!/k^k, times |Aut| if asked
    const double prob = factorial(k) / (factorial(0) * std::pow((double)k, k));
    printf("Final raw count is %e\n", mean);
    printf("Prob is %f\n", prob);
    const int aut = autom_ ? automorphismNum() : 1;
    const double fin = std::floor(mean / (prob * (double)aut) + 0.5);
    printf("Final count is %e\n", fin);
    fflush(stdout);
    return fin;
}

// Permutations of the template's vertices that map edges to edges, found by
// backtracking with a degree filter (the reference enumerates all k!
// permutations and then checks them, CountMat.cpp:1498-1572).
int CountMat::automorphismNum() const
{
    if (!tmpl_) return 0;
    const Graph& g = *tmpl_;
    const int n = g.get_vert_num();
    std::vector<std::vector<char>> adj(n, std::vector<char>(n, 0));
    for (int v = 0; v < n; ++v)
        for (int j = 0; j < g.get_out_deg(v); ++j) adj[v][g.get_adj_list(v)[j]] = 1;
    std::vector<int> map(n, -1);
    std::vector<char> used(n, 0);
    int count = 0;
    std::function<void(int)> rec = [&](int i) {
        if (i == n) { ++count; return; }
        for (int c = 0; c < n; ++c) {
            if (used[c] || g.get_out_deg(c) != g.get_out_deg(i)) continue;
            bool ok = true;
            for (int j = 0; j < i && ok; ++j) ok = adj[i][j] == adj[c][map[j]];
            if (!ok) continue;
            used[c] = 1;
            map[i] = c;
            rec(i + 1);
            used[c] = 0;
        }
    };
    rec(0);
    return count;
}

}  // namespace pgbsc
