/*
 * CPU restatement of the reference's color-coding DP (dsc-nec/subgraph),
 * used ONLY as test infrastructure: the parity checker for the HIP path, and
 * the CPU baseline leg of bench.py.  The product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  What pins each
 * part (golden vectors from the compiled reference, closed forms, brute
 * force) is listed per function; nothing here is "parity unpinned".
 */
#ifndef PGBSC_ORACLE_H
#define PGBSC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXK 16
#define OR_MAXSUB (2 * OR_MAXK)
#define OR_DUMMY 93620 /* Helper.hpp:7 DUMMY_VAL */

typedef struct or_plan or_plan;

/* Template edge list -> decomposition + index tables.
 * DivideTemplates.cpp:6-327 (DivideTp, divide, split, split_nodes,
 * check_val, sort_tps) and IndexSys.cpp:9-57, 169-436 (comb table, gen_index,
 * gen_colors, gen_comb_hash_table).  Vertices must already be 0..nv-1
 * (Graph::read_enlist's hole removal is or_read_template's job). */
or_plan* or_plan_create(int nv, int ne, const int* src, const int* dst);
void or_plan_free(or_plan* p);
int or_plan_k(const or_plan* p);
int or_plan_nsub(const or_plan* p);
/* per subtemplate: size, main/aux child index (-1 for a leaf), parent (-1 at
 * the root), ncomb = C(k,size), nsplit = C(size, main size) (0 for a leaf) */
void or_plan_sub(const or_plan* p, int s, int* size, int* main_idx, int* aux_idx,
                 int* parent, int* ncomb, int* nsplit);
/* comb_to_count[ncomb]; split_main/aux[ncomb*nsplit] (row i = i-th lex set) */
void or_plan_tables(const or_plan* p, int s, int* comb_to_count, int* split_main, int* split_aux);

/* IndexSys::comb_calc (IndexSys.cpp:206-220) */
int or_comb(int n, int k);

/* Fascia/edge-list text readers: header n, m then "u v" lines.
 * EdgeList.cpp:34-114 (graph: hole removal renumbers and shrinks n) and
 * Graph.cpp:9-80 (template: hole removal renumbers, keeps n).
 * Return 0 on success; src and dst are malloc-ed (free with or_free). */
int or_read_edgelist(const char* path, int* n, int* m, int** src, int** dst);
int or_read_template(const char* path, int* n, int* m, int** src, int** dst);
void or_free(void* p);

/* CSCGraph::createFromEdgeListFile (CSCGraph.hpp:96-178): every listed edge
 * stored in both columns, rows sorted per column (duplicates and self loops
 * kept).  colptr[n+1], rowidx[2m]. */
void or_csc_build(int n, int m, const int* src, const int* dst, int* colptr, int* rowidx);

/* One coloring through the pruned SpMM DP with the reference's FP32
 * numerics: CountMat.cpp:268-322 (subtemplate loop), 556-849 (SpMM in
 * ascending-neighbour order, eMA in lexicographic-set / split order,
 * float(1e-12) on the first processed non-leaf non-root subtemplate, FP64 root
 * and serial FP64 vertex sum times 1e12), DataTableColMajor.cpp:346-400,
 * 489-494 (unfused dst + a*b), 553-567 (one-hot leaf).
 * Pinned by tests/golden (compiled reference), closed forms and brute force. */
double or_count_f32(const or_plan* p, int n, const int* colptr, const int* rowidx,
                    const uint8_t* colors);

/* Same DP in FP64 without the 1e-12 scale: exact integer counts below 2^53.
 * vals (per CSC entry, may be NULL for 1.0) weight the SpMM: Y[v] += A[v][u]*T[u].
 * compat_2vertex reproduces the reference's x1e12 on templates whose only
 * non-leaf subtemplate is the root (CountMat.cpp:811-824).
 * Pinned by brute force, closed forms and the FP32 golden vectors (within 1e-5). */
double or_count_f64(const or_plan* p, int n, const int* colptr, const int* rowidx,
                    const double* vals, const uint8_t* colors, int compat_2vertex);

/* Brute-force enumeration of colorful template maps (the quantity the DP
 * counts: maps phi from template vertices to graph vertices with every
 * template edge on a graph edge and all colors distinct).  Returns the
 * weighted count sum_phi prod_e A[phi(x)][phi(y)] (vals NULL = 1.0) and, if
 * edge_count != NULL, adds for every map and template edge (x,y), x nearer to
 * template vertex 0, the derivative of the map's weight with respect to that
 * entry (the product of the other edges' weights) to
 * edge_count[entry(row=phi(x), col=phi(y))] (entry = CSC index).
 * Independent of the DP: pins it on small graphs. */
double or_brute_force(int tnv, int tne, const int* tsrc, const int* tdst,
                      int n, const int* colptr, const int* rowidx, const double* vals,
                      const uint8_t* colors, double* edge_count);

#ifdef __cplusplus
}
#endif
#endif
