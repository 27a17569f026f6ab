// Host-side surface of the reference (dsc-nec/subgraph), rebuilt for the
// MI355X library: the text/binary graph loaders, the Fascia template graph,
// the subtemplate split and the color-set index system.  Class and method
// names follow the reference so that code written against it ports
// unchanged; behaviour (including the reference's numbering rules) is kept
// where it fixes results, while errors are reported instead of crashing.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pgbsc {

constexpr int kDummy = 93620;  // Helper.hpp:7 DUMMY_VAL: "no child"
constexpr int kMaxTemplate = 16;

// --------------------------------------------------------------------------
// Template graph: Graph.hpp:11-64 / Graph.cpp:9-129.  Adjacency entries are
// kept in edge-insertion order, which fixes every later "first neighbour".
// --------------------------------------------------------------------------
class Graph {
public:
    // Fascia text: "n\nm\n" then m lines "u v".  Ids with a gap are renumbered
    // consecutively in ascending order; n is kept (Graph.cpp:40-74).
    // Returns 0, or a negative PG_ERR_* code with `err` set.
    int read_enlist(const std::string& path, std::string* err = nullptr);
    void build_graph(int verts, int edges, const int* src, const int* dst);

    int get_vert_num() const { return nv_; }
    int get_edge_num() const { return 2 * (int)src_.size(); }
    int get_out_deg(int v) const { return off_[v + 1] - off_[v]; }
    const int* get_adj_list(int v) const { return adj_.data() + off_[v]; }
    const std::vector<int>& src() const { return src_; }
    const std::vector<int>& dst() const { return dst_; }

private:
    int nv_ = 0;
    std::vector<int> off_, adj_, src_, dst_;
};

// --------------------------------------------------------------------------
// Text edge list: EdgeList.hpp:19-68 / EdgeList.cpp:34-114.
// --------------------------------------------------------------------------
class EdgeList {
public:
    EdgeList() = default;
    int read(const std::string& path, std::string* err = nullptr);
    // Apply the reference's "hole" rule to an in-memory list: when the largest
    // id is not n-1, ids that occur are renumbered 0.. in ascending order and
    // n becomes their count (EdgeList.cpp:70-112).
    static int64_t remove_holes(int64_t n, int64_t m, int32_t* src, int32_t* dst);

    int32_t getNumVertices() const { return n_; }
    int64_t getNumEdges() const { return (int64_t)src_.size(); }
    const int32_t* getSrcList() const { return src_.data(); }
    const int32_t* getDstList() const { return dst_.data(); }

private:
    int32_t n_ = 0;
    std::vector<int32_t> src_, dst_;
};

// --------------------------------------------------------------------------
// CSC adjacency: CSCGraph.hpp:22-178, 428-463.
// --------------------------------------------------------------------------
class CSCGraph {
public:
    // Every listed edge lands in both columns; rows are sorted per column;
    // duplicates and self loops are kept; values are 1.0.
    void createFromEdgeListFile(int32_t n, int64_t m, const int32_t* src, const int32_t* dst);
    // Binary "graph.data" layout (CSCGraph.hpp:428-436): int32 numEdges,
    // int32 n, int32 deg[n], int32 colptr[n+1], int32 rowidx[nnz], float val[nnz].
    int serialize(const std::string& path, std::string* err = nullptr) const;
    // Opens read-only (the reference opens with O_CREAT|O_RDWR and silently
    // creates a missing file, sc-main.cpp:1669) and checks every length.
    int deserialize(const std::string& path, std::string* err = nullptr);

    int32_t getNumVertices() const { return n_; }
    int64_t getNNZ() const { return (int64_t)rowidx_.size(); }
    int32_t getNumEdges() const { return num_edges_; }
    const int32_t* getIndexCol() const { return colptr_.data(); }
    const int32_t* getIndexRow() const { return rowidx_.data(); }
    const float* getNNZVal() const { return val_.data(); }
    const int32_t* getDegList() const { return deg_.data(); }

private:
    int32_t n_ = 0, num_edges_ = 0;
    std::vector<int32_t> deg_, colptr_, rowidx_;
    std::vector<float> val_;
};

// --------------------------------------------------------------------------
// Subtemplate chain: DivideTemplates.hpp:14-65 / DivideTemplates.cpp:6-327.
// Subtemplate 0 is the template rooted at its vertex 0.  Each non-leaf
// subtemplate is cut at the edge to its root's FIRST neighbour: "main" keeps
// the root, "aux" is rooted at that neighbour.  Subtemplates are then ordered
// by parent index with the reference's adjacent-swap pass, so children always
// follow their parent and the DP runs from the last index down to 0.
// --------------------------------------------------------------------------
class DivideTemplates {
public:
    int DivideTp(const Graph& tp, std::string* err = nullptr);
    void sort_tps();

    int get_subtps_num() const { return (int)subs_.size(); }
    const Graph& get_subtp(int s) const { return subs_[s]; }
    int get_main_node_idx(int s) const { return main_[s]; }
    int get_aux_node_idx(int s) const { return aux_[s]; }
    int get_parent_idx(int s) const { return parent_[s]; }
    int get_main_node_vert_num(int s) const { return subs_[main_[s]].get_vert_num(); }
    int get_aux_node_vert_num(int s) const { return subs_[aux_[s]].get_vert_num(); }

private:
    void divide(int sub, int root);
    int cut(int sub, int root, int excluded);
    std::vector<Graph> subs_;
    std::vector<int> main_, aux_, parent_;
};

// --------------------------------------------------------------------------
// Color-set index system: IndexSys.hpp:10-90 / IndexSys.cpp.
// The table column of a color set {c_0 < c_1 < ...} (0-based) is its colex
// rank sum_i C(c_i, i+1).  For every subtemplate the reference walks the
// color sets in lexicographic order; getCombToCount()[s][i] is the column of
// the i-th set and the split tables give, for the j-th split of that set, the
// main-child column (j-th lexicographic subset of the set's positions) and
// the aux-child column (the complementary, (C-1-j)-th, subset).
// --------------------------------------------------------------------------
class IndexSys {
public:
    void initialization(int color_num, const DivideTemplates& div);
    static int comb_calc(int n, int k);

    int getSubsSize(int s) const { return sizes_[s]; }
    const std::vector<int>& getCombToCount(int s) const { return c2c_[s]; }
    // row-major [C(k,size)][C(size, main size)]
    const std::vector<int>& getSplitMain(int s) const { return split_main_[s]; }
    const std::vector<int>& getSplitAux(int s) const { return split_aux_[s]; }
    int getSplitNum(int s) const { return nsplit_[s]; }

private:
    int k_ = 0;
    std::vector<int> sizes_, nsplit_;
    std::vector<std::vector<int>> c2c_, split_main_, split_aux_;
};

// Flattened plan arrays in the layout pg_plan_create takes.
struct FlatPlan {
    int k = 0, nsub = 0;
    std::vector<int32_t> sub_size, main_idx, aux_idx, comb_to_count, split_main, split_aux;
};
int flatten_plan(const Graph& tmpl, FlatPlan* out, std::string* err = nullptr);

// splitmix64-based color stream shared by host and device.
inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint8_t color_of(uint64_t seed, uint64_t coloring, uint64_t v, uint32_t k)
{
    uint64_t h = splitmix64(seed ^ splitmix64(coloring * 0x9E3779B97F4A7C15ull + v));
    return (uint8_t)(((h >> 32) * (uint64_t)k) >> 32);
}

}  // namespace pgbsc
