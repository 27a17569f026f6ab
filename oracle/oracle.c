/*
 * CPU restatement of the reference's color-coding DP -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker: the product path never calls
 * it.  See oracle.h for what each function restates and how it is pinned.
 * Build: make -C oracle oracle  (gcc -std=c11 -O2 -fopenmp -ffp-contract=off).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* ------------------------------------------------------------------------ */
/* small insertion-ordered graph: Graph::build_graph (Graph.cpp:82-129)     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nv, ne;              /* ne = undirected edges */
    int off[OR_MAXK + 1];
    int adj[2 * OR_MAXK];
} tgraph;

static void tg_build(tgraph* g, int nv, int ne, const int* src, const int* dst)
{
    int cnt[OR_MAXK] = {0}, pos[OR_MAXK];
    g->nv = nv;
    g->ne = ne;
    for (int j = 0; j < ne; ++j) { cnt[src[j]]++; cnt[dst[j]]++; }
    g->off[0] = 0;
    for (int j = 0; j < nv; ++j) g->off[j + 1] = g->off[j] + cnt[j];
    memcpy(pos, g->off, nv * sizeof(int));
    for (int j = 0; j < ne; ++j) {
        g->adj[pos[src[j]]++] = dst[j];
        g->adj[pos[dst[j]]++] = src[j];
    }
}

/* ------------------------------------------------------------------------ */
/* IndexSys::comb_calc (IndexSys.cpp:206-220)                               */
/* ------------------------------------------------------------------------ */
int or_comb(int n, int k)
{
    if (k > n) return 0;
    if (k * 2 > n) k = n - k;
    if (k == 0) return 1;
    int r = n;
    for (int i = 2; i <= k; ++i) { r *= (n - i + 1); r /= i; }
    return r;
}

/* ------------------------------------------------------------------------ */
/* DivideTemplates (DivideTemplates.cpp:6-327)                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    tgraph sub[OR_MAXSUB];
    int node_main[OR_MAXSUB], node_aux[OR_MAXSUB], parents[OR_MAXSUB];
    int cur_index;
} divider;

/* set_nodes (DivideTemplates.cpp:87-93): unset slots hold the dummy value */
static void set_node(int* list, int idx, int val) { list[idx] = val; }

/* split_nodes (DivideTemplates.cpp:153-217): BFS from root that never enters
 * `rest`, edges kept in discovery order, relabelled by check_val (root -> 0,
 * then first appearance in edge order, DivideTemplates.cpp:219-246). */
static int split_nodes(divider* d, int sub, int root, int rest)
{
    const tgraph* g = &d->sub[sub];
    int src[OR_MAXK], dst[OR_MAXK], ns = 0;
    int added[OR_MAXK * 2], na = 0;
    added[na++] = root;
    for (int i = 0; i < na; ++i) {
        int a = added[i];
        for (int j = g->off[a]; j < g->off[a + 1]; ++j) {
            int v = g->adj[j];
            int add = 1;
            for (int q = 0; q < ns; ++q)
                if (src[q] == v && dst[q] == a) { add = 0; break; }
            if (add && v != rest) {
                src[ns] = a; dst[ns] = v; ns++;
                added[na++] = v;
            }
        }
    }
    int nv, ne;
    if (ns > 0) {
        ne = ns;
        nv = ns + 1;
        int maxv = 0;
        for (int i = 0; i < ns; ++i) {
            if (src[i] > maxv) maxv = src[i];
            if (dst[i] > maxv) maxv = dst[i];
        }
        int map[OR_MAXK * 2];
        for (int i = 0; i <= maxv; ++i) map[i] = -1;
        map[root] = 0;
        int next = 1;
        for (int i = 0; i < ns; ++i) {
            if (map[src[i]] == -1) map[src[i]] = next++;
            if (map[dst[i]] == -1) map[dst[i]] = next++;
        }
        for (int i = 0; i < ns; ++i) { src[i] = map[src[i]]; dst[i] = map[dst[i]]; }
    } else {
        ne = 0;
        nv = 1;
        src[0] = 0;
    }
    int ret = src[0];
    tg_build(&d->sub[d->cur_index], nv, ne, src, dst);
    d->cur_index++;
    return ret;
}

/* divide + split (DivideTemplates.cpp:45-85, 128-141): the root's FIRST
 * adjacency entry becomes the aux root; main = side with the root. */
static void divide(divider* d, int sub, int root)
{
    int new_node = d->sub[sub].adj[d->sub[sub].off[root]];
    int root_main = split_nodes(d, sub, root, new_node);
    int root_aux = split_nodes(d, sub, new_node, root);
    int main_index = d->cur_index - 2;
    int aux_index = d->cur_index - 1;
    set_node(d->node_main, sub, main_index);
    set_node(d->node_aux, sub, aux_index);
    set_node(d->parents, main_index, sub);
    set_node(d->parents, aux_index, sub);
    if (d->sub[main_index].nv > 1) divide(d, main_index, root_main);
    else { set_node(d->node_main, main_index, OR_DUMMY); set_node(d->node_aux, main_index, OR_DUMMY); }
    if (d->sub[aux_index].nv > 1) divide(d, aux_index, root_aux);
    else { set_node(d->node_main, aux_index, OR_DUMMY); set_node(d->node_aux, aux_index, OR_DUMMY); }
}

/* sort_tps (DivideTemplates.cpp:264-327): bubble sort by parent index with
 * the reference's child/parent link fix-ups, in the same statement order. */
static void sort_tps(divider* d, int total)
{
    int* pm = d->node_main; int* pa = d->node_aux; int* pp = d->parents;
    int running = 1;
    while (running) {
        int swapped = 0;
        for (int i = 2; i < total; ++i) {
            if (pp[i] < pp[i - 1]) {
                tgraph cur = d->sub[i];
                int cp = pp[i], cm = pm[i], ca = pa[i];
                if (pm[i] != OR_DUMMY) pp[pm[i]] = i - 1;
                if (pm[i - 1] != OR_DUMMY) pp[pm[i - 1]] = i;
                if (pa[i] != OR_DUMMY) pp[pa[i]] = i - 1;
                if (pa[i - 1] != OR_DUMMY) pp[pa[i - 1]] = i;
                if (pm[pp[i]] == i) pm[pp[i]] = i - 1;
                else if (pa[pp[i]] == i) pa[pp[i]] = i - 1;
                if (pm[pp[i - 1]] == i - 1) pm[pp[i - 1]] = i;
                else if (pa[pp[i - 1]] == i - 1) pa[pp[i - 1]] = i;
                d->sub[i] = d->sub[i - 1];
                pp[i] = pp[i - 1]; pm[i] = pm[i - 1]; pa[i] = pa[i - 1];
                d->sub[i - 1] = cur;
                pp[i - 1] = cp; pm[i - 1] = cm; pa[i - 1] = ca;
                swapped = 1;
            }
        }
        if (!swapped) running = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* IndexSys (IndexSys.cpp:169-436)                                          */
/* ------------------------------------------------------------------------ */
struct or_plan {
    int k, nsub;
    int size[OR_MAXSUB], main_idx[OR_MAXSUB], aux_idx[OR_MAXSUB], parent[OR_MAXSUB];
    int ncomb[OR_MAXSUB], nsplit[OR_MAXSUB];
    int* c2c[OR_MAXSUB];
    int* smain[OR_MAXSUB];
    int* saux[OR_MAXSUB];
};

/* perm_set_next (IndexSys.cpp:228-243): next lexicographic subset (1-based) */
static void perm_next(int* s, int size, int colors)
{
    for (int i = size - 1; i >= 0; i--) {
        if (s[i] < colors - (size - i - 1)) {
            s[i]++;
            for (int j = i + 1; j < size; j++) s[j] = s[j - 1] + 1;
            break;
        }
    }
}

/* get_color_hash (IndexSys.cpp:303-313): colex rank sum_i C(c_i - 1, i + 1) */
static int color_hash(const int* s, int size)
{
    int c = 0;
    for (int i = 0; i < size; i++) c += or_comb(s[i] - 1, i + 1);
    return c;
}

/* gen_index (IndexSys.cpp:169-203): all lex `len`-subsets of positions
 * 0..val-1, row-major [C(val,len)][len]. */
static int* lex_subsets(int val, int len)
{
    int cnt = or_comb(val, len);
    int* out = (int*)malloc(sizeof(int) * (size_t)(cnt > 0 ? cnt : 1) * (len > 0 ? len : 1));
    int s[OR_MAXK];
    for (int i = 0; i < len; i++) s[i] = i + 1;
    for (int r = 0; r < cnt; r++) {
        for (int q = 0; q < len; q++) out[r * len + q] = s[q] - 1;
        perm_next(s, len, val);
    }
    return out;
}

or_plan* or_plan_create(int nv, int ne, const int* src, const int* dst)
{
    if (nv < 2 || nv > OR_MAXK || ne != nv - 1) return NULL;
    divider* d = (divider*)calloc(1, sizeof(divider));
    for (int i = 0; i < OR_MAXSUB; ++i) d->node_main[i] = d->node_aux[i] = d->parents[i] = OR_DUMMY;
    tg_build(&d->sub[0], nv, ne, src, dst);
    d->cur_index = 1;
    if (nv > 1) divide(d, 0, 0);
    else { d->node_main[0] = d->node_aux[0] = OR_DUMMY; }
    int total = d->cur_index;
    sort_tps(d, total);

    or_plan* p = (or_plan*)calloc(1, sizeof(or_plan));
    p->k = nv;
    p->nsub = total;
    for (int s = 0; s < total; ++s) {
        p->size[s] = d->sub[s].nv;
        p->main_idx[s] = d->node_main[s] == OR_DUMMY ? -1 : d->node_main[s];
        p->aux_idx[s] = d->node_aux[s] == OR_DUMMY ? -1 : d->node_aux[s];
        p->parent[s] = d->parents[s] == OR_DUMMY ? -1 : d->parents[s];
    }
    int k = nv;
    for (int s = 0; s < total; ++s) {
        int sz = p->size[s];
        int nc = or_comb(k, sz);
        p->ncomb[s] = nc;
        p->c2c[s] = (int*)malloc(sizeof(int) * nc);
        int mm = 0, aa = 0, ns = 0;
        int *pos_main = NULL, *pos_aux = NULL;
        if (sz > 1) {
            mm = p->size[p->main_idx[s]];
            aa = p->size[p->aux_idx[s]];
            ns = or_comb(sz, mm);
            p->smain[s] = (int*)malloc(sizeof(int) * (size_t)nc * ns);
            p->saux[s] = (int*)malloc(sizeof(int) * (size_t)nc * ns);
            /* gen_colors (IndexSys.cpp:245-301) uses _index_tmp[sz-2][mm-1]
             * for the main part and _index_tmp[sz-2][aa-1] for the aux part */
            pos_main = lex_subsets(sz, mm);
            pos_aux = lex_subsets(sz, aa);
        }
        p->nsplit[s] = ns;
        int perm[OR_MAXK];
        for (int i = 0; i < sz; i++) perm[i] = i + 1;
        for (int n = 0; n < nc; n++) {
            p->c2c[s][n] = color_hash(perm, sz);
            if (sz > 1) {
                /* gen_comb_hash_table (IndexSys.cpp:377-414): main from the
                 * j-th lex position subset, aux from the (C-1-j)-th */
                int aux_itr = ns - 1;
                for (int j = 0; j < ns; ++j, --aux_itr) {
                    int cm[OR_MAXK], ca[OR_MAXK];
                    for (int q = 0; q < mm; q++) cm[q] = perm[pos_main[j * mm + q]];
                    for (int q = 0; q < aa; q++) ca[q] = perm[pos_aux[aux_itr * aa + q]];
                    p->smain[s][n * ns + j] = color_hash(cm, mm);
                    p->saux[s][n * ns + j] = color_hash(ca, aa);
                }
            }
            perm_next(perm, sz, k);
        }
        free(pos_main);
        free(pos_aux);
    }
    free(d);
    return p;
}

void or_plan_free(or_plan* p)
{
    if (!p) return;
    for (int s = 0; s < p->nsub; ++s) { free(p->c2c[s]); free(p->smain[s]); free(p->saux[s]); }
    free(p);
}

int or_plan_k(const or_plan* p) { return p->k; }
int or_plan_nsub(const or_plan* p) { return p->nsub; }

void or_plan_sub(const or_plan* p, int s, int* size, int* main_idx, int* aux_idx,
                 int* parent, int* ncomb, int* nsplit)
{
    *size = p->size[s]; *main_idx = p->main_idx[s]; *aux_idx = p->aux_idx[s];
    *parent = p->parent[s]; *ncomb = p->ncomb[s]; *nsplit = p->nsplit[s];
}

void or_plan_tables(const or_plan* p, int s, int* comb_to_count, int* split_main, int* split_aux)
{
    memcpy(comb_to_count, p->c2c[s], sizeof(int) * p->ncomb[s]);
    if (p->nsplit[s] > 0) {
        memcpy(split_main, p->smain[s], sizeof(int) * (size_t)p->ncomb[s] * p->nsplit[s]);
        memcpy(split_aux, p->saux[s], sizeof(int) * (size_t)p->ncomb[s] * p->nsplit[s]);
    }
}

/* ------------------------------------------------------------------------ */
/* text readers (EdgeList.cpp:34-114, Graph.cpp:9-80)                       */
/* ------------------------------------------------------------------------ */
/* getline(..., ' ') then getline(...) then atoi, as the reference parses */
static int read_pairs(const char* path, int* n, int* m, int** src, int** dst, int* max_id)
{
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char* line = NULL;
    size_t cap = 0;
    if (getline(&line, &cap, f) < 0) { fclose(f); free(line); return -2; }
    *n = atoi(line);
    if (getline(&line, &cap, f) < 0) { fclose(f); free(line); return -2; }
    *m = atoi(line);
    if (*m < 0) { fclose(f); free(line); return -2; }
    int* s = (int*)malloc(sizeof(int) * (size_t)(*m > 0 ? *m : 1));
    int* d = (int*)malloc(sizeof(int) * (size_t)(*m > 0 ? *m : 1));
    int mx = 0;
    for (int j = 0; j < *m; ++j) {
        if (getdelim(&line, &cap, ' ', f) < 0) line[0] = 0;
        s[j] = atoi(line);
        if (s[j] > mx) mx = s[j];
        if (getline(&line, &cap, f) < 0) line[0] = 0;
        d[j] = atoi(line);
        if (d[j] > mx) mx = d[j];
    }
    fclose(f);
    free(line);
    *src = s; *dst = d; *max_id = mx;
    return 0;
}

/* hole removal: ids that appear get consecutive new ids in ascending order */
static int remove_holes(int m, int* s, int* d, int max_id)
{
    int* vid = (int*)calloc((size_t)max_id + 1, sizeof(int));
    for (int i = 0; i < m; i++) { vid[s[i]] = 1; vid[d[i]] = 1; }
    int itr = 0;
    for (int i = 0; i <= max_id; i++) if (vid[i] == 1) vid[i] = itr++;
    for (int i = 0; i < m; i++) { s[i] = vid[s[i]]; d[i] = vid[d[i]]; }
    free(vid);
    return itr;
}

int or_read_edgelist(const char* path, int* n, int* m, int** src, int** dst)
{
    int mx;
    int rc = read_pairs(path, n, m, src, dst, &mx);
    if (rc) return rc;
    if (mx != *n - 1) *n = remove_holes(*m, *src, *dst, mx); /* EdgeList.cpp:70-112 */
    return 0;
}

int or_read_template(const char* path, int* n, int* m, int** src, int** dst)
{
    int mx;
    int rc = read_pairs(path, n, m, src, dst, &mx);
    if (rc) return rc;
    if (mx != *n - 1) remove_holes(*m, *src, *dst, mx); /* Graph.cpp:40-74: n kept */
    return 0;
}

void or_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* CSC build (CSCGraph.hpp:96-178)                                          */
/* ------------------------------------------------------------------------ */
static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

void or_csc_build(int n, int m, const int* src, const int* dst, int* colptr, int* rowidx)
{
    int* deg = (int*)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < m; i++) { deg[dst[i]]++; deg[src[i]]++; }
    colptr[0] = 0;
    for (int i = 1; i <= n; i++) colptr[i] = colptr[i - 1] + deg[i - 1];
    int* pos = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(pos, colptr, sizeof(int) * (size_t)n);
    for (int i = 0; i < m; i++) {
        rowidx[pos[dst[i]]++] = src[i];
        rowidx[pos[src[i]]++] = dst[i];
    }
    for (int i = 0; i < n; i++)
        qsort(rowidx + colptr[i], (size_t)(colptr[i + 1] - colptr[i]), sizeof(int), cmp_int);
    free(pos);
    free(deg);
}

/* ------------------------------------------------------------------------ */
/* the DP                                                                   */
/* ------------------------------------------------------------------------ */
/* Row v of the SpMM output in the reference is accumulated over the row-owner
 * split lists, which hold the entries of row v in ascending column order
 * (CSCGraph.hpp:199-213, 254-307).  For a symmetric pattern that is column v's
 * rows in order; for generality build the transpose (rows -> columns). */
typedef struct { int* ptr; int* idx; int* ent; } rowview;

static rowview row_view(int n, const int* colptr, const int* rowidx)
{
    rowview r;
    int nnz = colptr[n];
    r.ptr = (int*)calloc((size_t)n + 1, sizeof(int));
    r.idx = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    r.ent = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    for (int e = 0; e < nnz; e++) r.ptr[rowidx[e] + 1]++;
    for (int i = 0; i < n; i++) r.ptr[i + 1] += r.ptr[i];
    int* pos = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(pos, r.ptr, sizeof(int) * (size_t)n);
    for (int u = 0; u < n; u++)          /* columns ascending */
        for (int e = colptr[u]; e < colptr[u + 1]; e++) {
            int v = rowidx[e];
            r.idx[pos[v]] = u;
            r.ent[pos[v]] = e;
            pos[v]++;
        }
    free(pos);
    return r;
}

static void row_view_free(rowview* r) { free(r->ptr); free(r->idx); free(r->ent); }

double or_count_f32(const or_plan* p, int n, const int* colptr, const int* rowidx,
                    const uint8_t* colors)
{
    const int k = p->k, S = p->nsub;
    const size_t N = (size_t)n;
    rowview rv = row_view(n, colptr, rowidx);
    float* tab[OR_MAXSUB] = {0};
    /* leaf table: one-hot, column c = colex rank of {c} = c
     * (DataTableColMajor.cpp:553-567; all size-1 subtemplates alias it,
     * DataTableColMajor.cpp:134-147) */
    float* leaf = (float*)calloc(N * k, sizeof(float));
    for (size_t v = 0; v < N; v++) leaf[(size_t)colors[v] * N + v] = 1.0f;
    int is_scaled = 0;
    double count = 0.0;
    const float scale = 1.0e-12f; /* CountMat.cpp:785 passes 1e-12 as float */
    for (int s = S - 1; s >= 0; --s) {
        if (p->size[s] == 1) continue;
        int mi = p->main_idx[s], ai = p->aux_idx[s];
        int mm = p->size[mi], aa = p->size[ai];
        int Cs = p->ncomb[s], P = p->nsplit[s], Ca = or_comb(k, aa);
        const float* Ta = aa > 1 ? tab[ai] : leaf;
        const float* Tm = mm > 1 ? tab[mi] : leaf;
        /* SpMM: Y[c][v] = sum over row v's entries in ascending column */
        float* Y = (float*)malloc(sizeof(float) * N * Ca);
        #pragma omp parallel for schedule(static)
        for (long v = 0; v < (long)N; v++) {
            for (int c = 0; c < Ca; c++) {
                float acc = 0.0f;
                const float* x = Ta + (size_t)c * N;
                for (int e = rv.ptr[v]; e < rv.ptr[v + 1]; e++) acc = acc + x[rv.idx[e]];
                Y[(size_t)c * N + v] = acc;
            }
        }
        if (s > 0) {
            float* out = (float*)calloc(N * Cs, sizeof(float));
            const int* c2c = p->c2c[s];
            const int* sm = p->smain[s];
            const int* sa = p->saux[s];
            #pragma omp parallel for schedule(static)
            for (long v = 0; v < (long)N; v++) {
                for (int i = 0; i < Cs; i++) {
                    float* dst = out + (size_t)c2c[i] * N + v;
                    for (int j = 0; j < P; j++) {
                        float a = Y[(size_t)sa[i * P + j] * N + v];
                        float b = Tm[(size_t)sm[i * P + j] * N + v];
                        if (!is_scaled) {
                            /* DataTableColMajor.cpp:367 */
                            double t = ((double)a * (double)b) * (double)scale;
                            *dst = (float)((double)*dst + t);
                        } else {
                            float prod = a * b; /* DataTableColMajor.cpp:489-494 */
                            *dst = *dst + prod;
                        }
                    }
                }
            }
            tab[s] = out;
        } else {
            double* last = (double*)calloc(N, sizeof(double));
            const int* sm = p->smain[s];
            const int* sa = p->saux[s];
            #pragma omp parallel for schedule(static)
            for (long v = 0; v < (long)N; v++) {
                for (int i = 0; i < Cs; i++)
                    for (int j = 0; j < P; j++) {
                        float a = Y[(size_t)sa[i * P + j] * N + v];
                        float b = Tm[(size_t)sm[i * P + j] * N + v];
                        float prod = a * b; /* DataTableColMajor.cpp:396: f32 product */
                        last[v] = last[v] + (double)prod;
                    }
            }
            is_scaled = 1; /* CountMat.cpp:811 precedes the root check */
            for (size_t v = 0; v < N; v++) count += last[v];
            if (is_scaled == 1) count *= 1.0e12;
            free(last);
        }
        is_scaled = 1;
        free(Y);
        if (mm > 1) { free(tab[mi]); tab[mi] = NULL; }
        if (aa > 1) { free(tab[ai]); tab[ai] = NULL; }
    }
    for (int s = 0; s < S; s++) free(tab[s]);
    free(leaf);
    row_view_free(&rv);
    return count;
}

double or_count_f64(const or_plan* p, int n, const int* colptr, const int* rowidx,
                    const double* vals, const uint8_t* colors, int compat_2vertex)
{
    const int k = p->k, S = p->nsub;
    const size_t N = (size_t)n;
    rowview rv = row_view(n, colptr, rowidx);
    double* tab[OR_MAXSUB] = {0};
    double* leaf = (double*)calloc(N * k, sizeof(double));
    for (size_t v = 0; v < N; v++) leaf[(size_t)colors[v] * N + v] = 1.0;
    double count = 0.0;
    int nonleaf = 0;
    for (int s = 0; s < S; s++) nonleaf += p->size[s] > 1;
    for (int s = S - 1; s >= 0; --s) {
        if (p->size[s] == 1) continue;
        int mi = p->main_idx[s], ai = p->aux_idx[s];
        int mm = p->size[mi], aa = p->size[ai];
        int Cs = p->ncomb[s], P = p->nsplit[s], Ca = or_comb(k, aa);
        const double* Ta = aa > 1 ? tab[ai] : leaf;
        const double* Tm = mm > 1 ? tab[mi] : leaf;
        double* Y = (double*)malloc(sizeof(double) * N * Ca);
        #pragma omp parallel for schedule(static)
        for (long v = 0; v < (long)N; v++) {
            for (int c = 0; c < Ca; c++) {
                double acc = 0.0;
                const double* x = Ta + (size_t)c * N;
                for (int e = rv.ptr[v]; e < rv.ptr[v + 1]; e++)
                    acc += (vals ? vals[rv.ent[e]] : 1.0) * x[rv.idx[e]];
                Y[(size_t)c * N + v] = acc;
            }
        }
        const int* c2c = p->c2c[s];
        const int* sm = p->smain[s];
        const int* sa = p->saux[s];
        if (s > 0) {
            double* out = (double*)calloc(N * Cs, sizeof(double));
            #pragma omp parallel for schedule(static)
            for (long v = 0; v < (long)N; v++)
                for (int i = 0; i < Cs; i++) {
                    double acc = 0.0;
                    for (int j = 0; j < P; j++)
                        acc += Y[(size_t)sa[i * P + j] * N + v] * Tm[(size_t)sm[i * P + j] * N + v];
                    out[(size_t)c2c[i] * N + v] = acc;
                }
            tab[s] = out;
        } else {
            double tot = 0.0;
            #pragma omp parallel for schedule(static) reduction(+:tot)
            for (long v = 0; v < (long)N; v++)
                for (int j = 0; j < P; j++)
                    tot += Y[(size_t)sa[j] * N + v] * Tm[(size_t)sm[j] * N + v];
            count = tot;
            if (compat_2vertex && nonleaf == 1) count *= 1.0e12;
        }
        free(Y);
        if (mm > 1) { free(tab[mi]); tab[mi] = NULL; }
        if (aa > 1) { free(tab[ai]); tab[ai] = NULL; }
    }
    for (int s = 0; s < S; s++) free(tab[s]);
    free(leaf);
    row_view_free(&rv);
    return count;
}

/* ------------------------------------------------------------------------ */
/* brute force                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int tnv;
    int order[OR_MAXK], par[OR_MAXK];   /* BFS order from template vertex 0 */
    int n;
    const int* colptr; const int* rowidx; const double* vals;
    const uint8_t* colors;
    int phi[OR_MAXK];
    int ent[OR_MAXK];                    /* CSC entry used by order[i]'s edge */
    double* edge_count;
    double total;
} bf_state;

static int find_entry(const bf_state* st, int row, int col)
{
    int lo = st->colptr[col], hi = st->colptr[col + 1];
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (st->rowidx[mid] < row) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void bf_rec(bf_state* st, int depth, unsigned used, double w)
{
    if (depth == st->tnv) {
        st->total += w;
        if (st->edge_count) {
            for (int i = 1; i < st->tnv; i++) {
                double d = 1.0;
                for (int q = 1; q < st->tnv; q++)
                    if (q != i) d *= st->vals ? st->vals[st->ent[q]] : 1.0;
                st->edge_count[st->ent[i]] += d;
            }
        }
        return;
    }
    int t = st->order[depth];
    int pv = st->phi[st->par[t]];
    /* neighbours u of pv: column pv's rows (symmetric pattern); weight is
     * the entry (row = pv, col = u), matching Y[v] += A[v][u] * T[u] */
    for (int e = st->colptr[pv]; e < st->colptr[pv + 1]; e++) {
        int u = st->rowidx[e];
        unsigned bit = 1u << st->colors[u];
        if (used & bit) continue;
        int ent = find_entry(st, pv, u);
        double we = st->vals ? st->vals[ent] : 1.0;
        st->phi[t] = u;
        st->ent[depth] = ent;
        bf_rec(st, depth + 1, used | bit, w * we);
    }
}

double or_brute_force(int tnv, int tne, const int* tsrc, const int* tdst,
                      int n, const int* colptr, const int* rowidx, const double* vals,
                      const uint8_t* colors, double* edge_count)
{
    if (tnv < 1 || tnv > OR_MAXK || tne != tnv - 1) return -1.0;
    tgraph tg;
    tg_build(&tg, tnv, tne, tsrc, tdst);
    bf_state st;
    memset(&st, 0, sizeof(st));
    st.tnv = tnv;
    int seen[OR_MAXK] = {0}, no = 0;
    st.order[no++] = 0; seen[0] = 1; st.par[0] = -1;
    for (int i = 0; i < no; i++) {
        int a = st.order[i];
        for (int j = tg.off[a]; j < tg.off[a + 1]; j++) {
            int b = tg.adj[j];
            if (!seen[b]) { seen[b] = 1; st.par[b] = a; st.order[no++] = b; }
        }
    }
    st.n = n; st.colptr = colptr; st.rowidx = rowidx; st.vals = vals;
    st.colors = colors; st.edge_count = edge_count;
    /* the edge of order[i] (i >= 1) is stored at ent[i]; bf_rec writes ent[depth]
     * where depth == i for vertex order[i] */
    for (int v = 0; v < n; v++) {
        st.phi[0] = v;
        bf_rec(&st, 1, 1u << colors[v], 1.0);
    }
    return st.total;
}
