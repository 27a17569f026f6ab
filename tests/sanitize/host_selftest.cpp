// Host-side self-test run under AddressSanitizer + UndefinedBehaviorSanitizer
// (tests/sanitize/Makefile, driven by tests/test_sanitize.py).  It exercises,
// on random inputs, the CPU restatement of the DP (oracle/oracle.c) and every
// host-only part of the product: the template split and index tables, the
// cost models, the R-MAT source, the CSC build, the hole rule, the color
// stream, and the text / binary graph loaders (including malformed files).
// Results are checked against each other and against brute force, so the run
// also fails on a wrong answer.  Exit status 0 = every check passed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "oracle.h"
#include "pg_host.hpp"
#include "pgbsc.h"

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            ++failures;                                   \
            fprintf(stderr, "CHECK failed %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                 \
            fprintf(stderr, "\n");                        \
        }                                                 \
    } while (0)

static void random_tree(std::mt19937_64& rng, int k, std::vector<int>& s, std::vector<int>& d)
{
    s.clear();
    d.clear();
    std::vector<int> perm(k);
    for (int i = 0; i < k; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    for (int j = 1; j < k; ++j) {
        s.push_back(perm[std::uniform_int_distribution<int>(0, j - 1)(rng)]);
        d.push_back(perm[j]);
    }
}

// template split: product tables against the oracle's
static void check_templates(std::mt19937_64& rng)
{
    std::vector<int> s, d;
    for (int trial = 0; trial < 40; ++trial) {
        const int k = 2 + trial % 11;
        random_tree(rng, k, s, d);
        int32_t nsub = 0;
        int64_t ncomb = 0, nsplit = 0;
        CHECK(pg_template_sizes(k, k - 1, s.data(), d.data(), &nsub, &ncomb, &nsplit) == PG_OK, "sizes k=%d", k);
        std::vector<int32_t> sz(nsub), mi(nsub), ai(nsub), c2c(ncomb), sm(std::max<int64_t>(nsplit, 1)),
            sa(std::max<int64_t>(nsplit, 1));
        CHECK(pg_template_tables(k, k - 1, s.data(), d.data(), sz.data(), mi.data(), ai.data(), c2c.data(),
                                 sm.data(), sa.data()) == PG_OK,
              "tables k=%d", k);
        or_plan* op = or_plan_create(k, k - 1, s.data(), d.data());
        CHECK(op && or_plan_nsub(op) == nsub, "oracle plan k=%d", k);
        for (int t = 0; op && t < nsub; ++t) {
            int size, m, a, parent, nc, ns;
            or_plan_sub(op, t, &size, &m, &a, &parent, &nc, &ns);
            CHECK(size == sz[t] && m == mi[t] && a == ai[t], "sub %d of k=%d", t, k);
        }
        pg_estimate e;
        CHECK(pg_estimate_model(1000, 8000, k, nsub, sz.data(), mi.data(), ai.data(), &e) == PG_OK, "model");
        CHECK(e.peak_mem_gb > 0 && std::isfinite(e.fascia_gbytes), "model values");
        if (op) or_plan_free(op);
    }
    // malformed trees are rejected
    int bs[] = {0, 1, 2}, bd[] = {1, 0, 2};
    int32_t nsub = 0;
    int64_t nc = 0, ns = 0;
    CHECK(pg_template_sizes(3, 3, bs, bd, &nsub, &nc, &ns) != PG_OK, "cycle accepted");
    CHECK(pg_template_sizes(0, 0, nullptr, nullptr, &nsub, &nc, &ns) != PG_OK, "empty accepted");
}

// CSC build, hole rule, color stream, R-MAT, and the DP against brute force
static void check_graphs(std::mt19937_64& rng)
{
    for (int trial = 0; trial < 12; ++trial) {
        const int n = 6 + (int)(rng() % 40);
        const int m = n + (int)(rng() % (3 * n));
        std::vector<int32_t> s(m), d(m);
        for (int i = 0; i < m; ++i) {
            s[i] = (int32_t)(rng() % n);
            d[i] = (int32_t)(rng() % n);
        }
        s[0] = 0;
        d[0] = n - 1;
        std::vector<int32_t> colptr(n + 1), rowidx(2 * (size_t)m);
        CHECK(pg_csc_from_edges(n, m, s.data(), d.data(), colptr.data(), rowidx.data()) == PG_OK, "csc");
        std::vector<int> oc(n + 1), orow(2 * (size_t)m);
        or_csc_build(n, m, s.data(), d.data(), oc.data(), orow.data());
        CHECK(std::equal(oc.begin(), oc.end(), colptr.begin()) && std::equal(orow.begin(), orow.end(), rowidx.begin()),
              "csc differs from the oracle, trial %d", trial);
        std::vector<int32_t> hs = s, hd = d;
        const int64_t nn = pg_edges_remove_holes(n + 5, m, hs.data(), hd.data());
        CHECK(nn >= 2 && nn <= n, "holes: %lld", (long long)nn);
        for (int k = 2; k <= 5; ++k) {
            std::vector<int> ts, td;
            random_tree(rng, k, ts, td);
            or_plan* op = or_plan_create(k, k - 1, ts.data(), td.data());
            std::vector<uint8_t> colors(n);
            pg_colors_fill(7 + trial, k, 1, n, k, colors.data());
            for (int v = 0; v < n; ++v) CHECK(colors[v] < k, "color out of range");
            const double f64 = or_count_f64(op, n, oc.data(), orow.data(), nullptr, colors.data(), 0);
            const double bf = or_brute_force(k, k - 1, ts.data(), td.data(), n, oc.data(), orow.data(), nullptr,
                                             colors.data(), nullptr);
            if (k > 2) CHECK(f64 == bf, "DP %g != brute force %g (k=%d)", f64, bf, k);
            const double f32 = or_count_f32(op, n, oc.data(), orow.data(), colors.data());
            if (k > 2 && bf > 0) CHECK(std::fabs(f32 - bf) <= 1e-5 * bf, "FP32 %g vs %g", f32, bf);
            or_plan_free(op);
        }
    }
    int64_t m = 0;
    CHECK(pg_rmat_edges(8, 8, 0.57, 0.19, 0.19, 3, 0, nullptr, nullptr, &m) == PG_OK && m > 0, "rmat size");
    std::vector<int32_t> rs(m), rd(m);
    CHECK(pg_rmat_edges(8, 8, 0.57, 0.19, 0.19, 3, m, rs.data(), rd.data(), &m) == PG_OK, "rmat");
    for (int64_t i = 0; i < m; ++i) CHECK(rs[i] < rd[i] && rd[i] < 256, "rmat edge %lld", (long long)i);
}

// text and binary loaders, good and malformed files
static void check_files()
{
    const std::string dir = getenv("PGBSC_SANITIZE_TMP") ? getenv("PGBSC_SANITIZE_TMP") : "/tmp";
    const std::string el = dir + "/sel.el", bin = dir + "/sel.data", bad = dir + "/sel_bad.el";
    FILE* f = fopen(el.c_str(), "w");
    fprintf(f, "9\n5\n0 2\n2 5\n5 7\n5 2\n7 0\n");
    fclose(f);
    pgbsc::EdgeList e;
    std::string err;
    CHECK(e.read(el, &err) == PG_OK && e.getNumVertices() == 4 && e.getNumEdges() == 5, "edge list: %s", err.c_str());
    pgbsc::CSCGraph g;
    g.createFromEdgeListFile(e.getNumVertices(), e.getNumEdges(), e.getSrcList(), e.getDstList());
    CHECK(g.serialize(bin, &err) == PG_OK, "serialize: %s", err.c_str());
    pgbsc::CSCGraph h;
    CHECK(h.deserialize(bin, &err) == PG_OK && h.getNNZ() == g.getNNZ() && h.getNumVertices() == g.getNumVertices(),
          "deserialize: %s", err.c_str());
    // truncated binary: an error, not a crash
    {
        FILE* r = fopen(bin.c_str(), "rb");
        std::vector<char> buf(4096);
        const size_t got = fread(buf.data(), 1, buf.size(), r);
        fclose(r);
        FILE* w = fopen(bin.c_str(), "wb");
        fwrite(buf.data(), 1, got / 2, w);
        fclose(w);
        pgbsc::CSCGraph t;
        CHECK(t.deserialize(bin, &err) != PG_OK, "truncated binary accepted");
    }
    pgbsc::CSCGraph missing;
    CHECK(missing.deserialize(dir + "/no_such_graph.data", &err) != PG_OK, "missing binary accepted");
    f = fopen(bad.c_str(), "w");
    fprintf(f, "4\n3\n0 1\n1 x\n");
    fclose(f);
    pgbsc::EdgeList b;
    CHECK(b.read(bad, &err) != PG_OK, "malformed edge list accepted");
    pgbsc::Graph t;
    f = fopen(bad.c_str(), "w");
    fprintf(f, "3\n5\n0 1\n1 2\n");
    fclose(f);
    CHECK(t.read_enlist(bad, &err) != PG_OK, "short template accepted");
    remove(el.c_str());
    remove(bin.c_str());
    remove(bad.c_str());
}

int main()
{
    std::mt19937_64 rng(20261202);
    check_templates(rng);
    check_graphs(rng);
    check_files();
    printf("host self-test: %s (%d failures)\n", failures ? "FAILED" : "passed", failures);
    return failures ? 1 : 0;
}
