// Graph loaders (text edge list, binary CSC) and the synthetic R-MAT source.
// Reference: EdgeList.cpp:7-114, CSCGraph.hpp:96-178 and 428-463.
#include "pg_host.hpp"
#include "pgbsc.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <omp.h>

namespace pgbsc {

static int fail(std::string* err, const std::string& msg, int code)
{
    if (err) *err = msg;
    return code;
}

// ------------------------------------------------------------- EdgeList ---

int64_t EdgeList::remove_holes(int64_t n, int64_t m, int32_t* src, int32_t* dst)
{
    int32_t max_id = 0;
    for (int64_t i = 0; i < m; ++i) max_id = std::max(max_id, std::max(src[i], dst[i]));
    if (max_id == n - 1) return n;
    std::vector<int32_t> id((size_t)max_id + 1, -1);
    for (int64_t i = 0; i < m; ++i) id[src[i]] = id[dst[i]] = 0;
    int32_t next = 0;
    for (auto& x : id) if (x == 0) x = next++;
#pragma omp parallel for
    for (int64_t i = 0; i < m; ++i) { src[i] = id[src[i]]; dst[i] = id[dst[i]]; }
    return next;
}

// Parses "n\nm\n" then m lines "u v" (EdgeList.cpp:42-66).  A buffered
// whole-file parse: the reference reads line by line with getline/atoi,
// which this reproduces for well-formed files at a fraction of the cost.
int EdgeList::read(const std::string& path, std::string* err)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return fail(err, "cannot open graph file " + path, PG_ERR_IO);
    std::vector<char> buf;
    {
        fseek(f, 0, SEEK_END);
        long sz = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (sz < 0) { fclose(f); return fail(err, "cannot size graph file " + path, PG_ERR_IO); }
        buf.resize((size_t)sz + 1);
        size_t got = fread(buf.data(), 1, (size_t)sz, f);
        buf[got] = 0;
    }
    fclose(f);
    const char* p = buf.data();
    const char* end = buf.data() + buf.size() - 1;
    auto next_line = [&](long long* v) -> bool {
        if (p >= end) return false;
        *v = atoll(p);
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
        return true;
    };
    long long n = 0, m = 0;
    if (!next_line(&n) || !next_line(&m)) return fail(err, "graph header truncated", PG_ERR_INVALID);
    if (n <= 0 || n > INT32_MAX || m < 0 || 2 * m > INT32_MAX)
        return fail(err, "graph size outside the int32 CSC range", PG_ERR_RANGE);
    src_.resize((size_t)m);
    dst_.resize((size_t)m);
    for (long long j = 0; j < m; ++j) {
        if (p >= end) return fail(err, "graph edge list truncated", PG_ERR_INVALID);
        long long u = strtoll(p, (char**)&p, 10);
        while (p < end && *p == ' ') ++p;
        long long v = strtoll(p, (char**)&p, 10);
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
        if (u < 0 || v < 0 || u > INT32_MAX || v > INT32_MAX)
            return fail(err, "graph vertex id out of range", PG_ERR_INVALID);
        src_[j] = (int32_t)u;
        dst_[j] = (int32_t)v;
    }
    n_ = (int32_t)remove_holes(n, m, src_.data(), dst_.data());
    for (long long j = 0; j < m; ++j)
        if (src_[j] >= n_ || dst_[j] >= n_) return fail(err, "graph vertex id >= n", PG_ERR_INVALID);
    return PG_OK;
}

// ------------------------------------------------------------- CSCGraph ---

void CSCGraph::createFromEdgeListFile(int32_t n, int64_t m, const int32_t* src, const int32_t* dst)
{
    n_ = n;
    num_edges_ = (int32_t)m;
    deg_.assign(n, 0);
    for (int64_t i = 0; i < m; ++i) { deg_[dst[i]]++; deg_[src[i]]++; }
    colptr_.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) colptr_[i + 1] = colptr_[i] + deg_[i];
    rowidx_.assign((size_t)colptr_[n], 0);
    std::vector<int32_t> pos(colptr_.begin(), colptr_.end() - 1);
    for (int64_t i = 0; i < m; ++i) {
        rowidx_[pos[dst[i]]++] = src[i];
        rowidx_[pos[src[i]]++] = dst[i];
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) std::sort(rowidx_.begin() + colptr_[i], rowidx_.begin() + colptr_[i + 1]);
    val_.assign(rowidx_.size(), 1.0f);
}

int CSCGraph::serialize(const std::string& path, std::string* err) const
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return fail(err, "cannot create " + path, PG_ERR_IO);
    int32_t nnz = colptr_.empty() ? 0 : colptr_[n_];
    bool ok = fwrite(&num_edges_, 4, 1, f) == 1 && fwrite(&n_, 4, 1, f) == 1 &&
              fwrite(deg_.data(), 4, (size_t)n_, f) == (size_t)n_ &&
              fwrite(colptr_.data(), 4, (size_t)n_ + 1, f) == (size_t)n_ + 1 &&
              fwrite(rowidx_.data(), 4, (size_t)nnz, f) == (size_t)nnz &&
              fwrite(val_.data(), 4, (size_t)nnz, f) == (size_t)nnz;
    ok = (fclose(f) == 0) && ok;
    return ok ? PG_OK : fail(err, "short write to " + path, PG_ERR_IO);
}

int CSCGraph::deserialize(const std::string& path, std::string* err)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return fail(err, "cannot open binary graph " + path, PG_ERR_IO);
    int32_t ne = 0, n = 0;
    auto rd = [&](void* dst, size_t cnt) { return fread(dst, 4, cnt, f) == cnt; };
    if (!rd(&ne, 1) || !rd(&n, 1) || n < 0) { fclose(f); return fail(err, "binary graph header truncated", PG_ERR_IO); }
    std::vector<int32_t> deg(n), colptr((size_t)n + 1);
    if (!rd(deg.data(), (size_t)n) || !rd(colptr.data(), (size_t)n + 1)) {
        fclose(f);
        return fail(err, "binary graph truncated", PG_ERR_IO);
    }
    int32_t nnz = colptr[n];
    bool mono = colptr[0] == 0 && nnz >= 0;
    for (int32_t i = 0; mono && i < n; ++i) mono = colptr[i + 1] >= colptr[i];
    if (!mono) { fclose(f); return fail(err, "binary graph has a malformed column index", PG_ERR_INVALID); }
    std::vector<int32_t> rowidx((size_t)nnz);
    std::vector<float> val((size_t)nnz);
    if (!rd(rowidx.data(), (size_t)nnz) || !rd(val.data(), (size_t)nnz)) {
        fclose(f);
        return fail(err, "binary graph truncated", PG_ERR_IO);
    }
    fclose(f);
    for (int32_t r : rowidx)
        if (r < 0 || r >= n) return fail(err, "binary graph row index out of range", PG_ERR_INVALID);
    n_ = n;
    num_edges_ = ne;
    deg_.swap(deg);
    colptr_.swap(colptr);
    rowidx_.swap(rowidx);
    val_.swap(val);
    return PG_OK;
}

}  // namespace pgbsc

// ------------------------------------------------------- C ABI helpers ---

extern "C" int pg_csc_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                                 int32_t* colptr, int32_t* rowidx)
{
    if (n < 0 || m < 0 || 2 * m > INT32_MAX || (m > 0 && (!src || !dst)) || !colptr || (m > 0 && !rowidx))
        return PG_ERR_INVALID;
    for (int64_t i = 0; i < m; ++i)
        if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) return PG_ERR_INVALID;
    pgbsc::CSCGraph g;
    g.createFromEdgeListFile(n, m, src, dst);
    memcpy(colptr, g.getIndexCol(), sizeof(int32_t) * ((size_t)n + 1));
    if (m > 0) memcpy(rowidx, g.getIndexRow(), sizeof(int32_t) * (size_t)g.getNNZ());
    return PG_OK;
}

extern "C" void pg_colors_fill(uint64_t seed, int64_t first_coloring, int ncolorings, int32_t n, int k,
                               uint8_t* out)
{
    if (!out || n <= 0 || ncolorings <= 0 || k <= 0) return;
#pragma omp parallel for collapse(2)
    for (int c = 0; c < ncolorings; ++c)
        for (int32_t v = 0; v < n; ++v)
            out[(size_t)c * n + v] = pgbsc::color_of(seed, (uint64_t)(first_coloring + c), (uint64_t)v, (uint32_t)k);
}

// R-MAT: each of edge_factor * 2^scale draws descends `scale` levels, picking
// quadrant a (0,0), b (0,1), c (1,0) or d (1,1) from a counter-based uniform
// per (draw, level); self loops are dropped, pairs made (min,max), labels
// shuffled with a seeded permutation, and duplicates removed.  Output is
// sorted by (u, v) so it does not depend on the thread count.
extern "C" int pg_rmat_edges(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                             int64_t cap, int32_t* src, int32_t* dst, int64_t* m_out)
{
    if (scale < 1 || scale > 30 || edge_factor < 1 || !m_out || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return PG_ERR_INVALID;
    const int64_t n = (int64_t)1 << scale;
    const int64_t draws = n * edge_factor;
    const double ab = a + b, abc = a + b + c;
    std::vector<uint32_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0u);
    // Fisher-Yates with a counter-based stream: permutation fixed by seed
    for (int64_t i = n - 1; i > 0; --i) {
        uint64_t h = pgbsc::splitmix64(seed ^ pgbsc::splitmix64(0xA5A5A5A5ull + (uint64_t)i));
        int64_t j = (int64_t)(((h >> 11) * (uint64_t)(i + 1)) >> 53);
        std::swap(perm[i], perm[j]);
    }
    std::vector<uint64_t> keys((size_t)draws);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < draws; ++e) {
        uint64_t u = 0, v = 0;
        for (int l = 0; l < scale; ++l) {
            uint64_t h = pgbsc::splitmix64(seed + pgbsc::splitmix64((uint64_t)e * 64 + (uint64_t)l));
            double r = (double)(h >> 11) * (1.0 / 9007199254740992.0);
            int bu = r >= ab, bv = (r >= a && r < ab) || r >= abc;
            u = (u << 1) | (uint64_t)bu;
            v = (v << 1) | (uint64_t)bv;
        }
        if (u == v) { keys[e] = ~0ull; continue; }
        uint64_t pu = perm[u], pv = perm[v];
        if (pu > pv) std::swap(pu, pv);
        keys[e] = (pu << 32) | pv;
    }
    // parallel sort: bucket by the top bits of u, sort buckets independently
    const int nb = 256;
    const int shift = 32 + scale - 8 > 32 ? 32 + scale - 8 : 32;
    std::vector<int64_t> bcnt((size_t)nb + 1, 0);
    for (uint64_t k : keys)
        if (k != ~0ull) bcnt[(k >> shift) + 1]++;
    std::partial_sum(bcnt.begin(), bcnt.end(), bcnt.begin());
    std::vector<uint64_t> sorted((size_t)bcnt[nb]);
    {
        std::vector<int64_t> pos(bcnt.begin(), bcnt.end() - 1);
        for (uint64_t k : keys)
            if (k != ~0ull) sorted[pos[k >> shift]++] = k;
    }
    std::vector<uint64_t>().swap(keys);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nb; ++i) std::sort(sorted.begin() + bcnt[i], sorted.begin() + bcnt[i + 1]);
    auto last = std::unique(sorted.begin(), sorted.end());
    int64_t m = last - sorted.begin();
    *m_out = m;
    if (cap > 0) {
        if (!src || !dst) return PG_ERR_INVALID;
        int64_t w = std::min(cap, m);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < w; ++i) {
            src[i] = (int32_t)(sorted[i] >> 32);
            dst[i] = (int32_t)(sorted[i] & 0xffffffffu);
        }
    }
    return PG_OK;
}

extern "C" int64_t pg_edges_remove_holes(int64_t n, int64_t m, int32_t* src, int32_t* dst)
{
    if (n <= 0 || m < 0 || (m > 0 && (!src || !dst))) return PG_ERR_INVALID;
    return pgbsc::EdgeList::remove_holes(n, m, src, dst);
}
