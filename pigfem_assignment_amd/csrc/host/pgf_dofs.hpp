// Degree-of-freedom numbering, CSR sparsity pattern, element scatter offsets
// and element colouring for the assembly kernels.
//
// Numbering follows DofObject::distribute_dofs (src/DofObject.cpp:241-278,
// 336-390, 498-538): free dofs 0..n_free-1 in node order (regular nodes, then
// enrichment nodes), node-major with the component index minor; prescribed
// dofs n_free.. in the same traversal.  The pattern is the one
// Solver::preallocate_matrix / populate_local_neighbors build
// (src/Solver.cpp:365-441,723-829) and set_matrix fills
// (src/Assembler.cpp:811-831): row dof r of node a holds every dof of every
// node sharing an element with a, zero entries included.  It is stored as one
// CSR over all rows (free rows first, then prescribed rows).  Because all free
// dof ids precede all prescribed ids, each row is [K_ff | K_fp] (free rows) or
// [K_pf | K_pp] (prescribed rows); rowsplit marks the boundary.
#pragma once

#include <cstdint>
#include <vector>

#include "pgf_mesh.hpp"

namespace pgf {

struct Dirichlet {
    int32_t node;
    int32_t dof;
    double value;
};

struct DofMap {
    int dim = 0;
    int64_t n_all = 0, n_free = 0, n_pres = 0;
    std::vector<int32_t> dof;        // [dim][n_all] global dof id
    std::vector<double> pres_value;  // [n_pres] full-load prescribed values
    void build(const Mesh& mesh, const std::vector<Dirichlet>& bcs);
};

// Plain-array view of the element topology the pattern needs.
struct MeshView {
    int dim = 0;
    int64_t n_nodes = 0, n_all = 0, n_elem = 0, n_cut = 0;
    const int32_t* conn = nullptr;          // [n_elem][4]
    const int32_t* cut_elem = nullptr;      // [n_cut]
    const int32_t* cut_n_enrich = nullptr;  // [n_cut]
    const int32_t* cut_enrich = nullptr;    // [n_cut][16]
    std::vector<int32_t> cut_index;         // per element, -1 if uncut
    void finalize();
    int n_total(int64_t e) const { return cut_index[e] < 0 ? 4 : 4 + cut_n_enrich[cut_index[e]]; }
    int32_t node(int64_t e, int s) const {
        return s < 4 ? conn[e * 4 + s] : (int32_t)(n_nodes + cut_enrich[(int64_t)cut_index[e] * 16 + s - 4]);
    }
    static MeshView of(const Mesh& m, std::vector<int32_t>& cut_elem, std::vector<int32_t>& cut_ne,
                       std::vector<int32_t>& cut_en);
};

struct Pattern {
    int64_t n_rows = 0, nnz = 0;
    std::vector<int64_t> rowptr;     // [n_rows+1]
    std::vector<int32_t> rowsplit;   // [n_rows] offset of the first prescribed column inside the row
    std::vector<int32_t> colidx;     // [nnz]
    // scatter offsets: for element slot pair (i, j), offset (from row start) of
    // the first free and the first prescribed column of slot j's node inside
    // any row of slot i's node.
    std::vector<uint8_t> elem_off;   // uncut layout: [n_elem][4][4][2]
    std::vector<int64_t> cut_off_ptr;  // [n_cut+1]
    std::vector<uint8_t> cut_off;    // per cut element [nt][nt][2]
    // greedy colouring: no two elements of one colour share a node
    int n_colors = 0;
    std::vector<int32_t> color;      // per element
    // row-gather plan: for every node its incident elements (element order),
    // the node's slot in each, and for uncut elements the packed (free, pres)
    // column offsets of the element's 4 nodes inside the node's rows
    std::vector<int64_t> inc_ptr;    // [n_all+1]
    std::vector<int32_t> inc_elem;   // [n_inc]
    std::vector<uint8_t> inc_slot;   // [n_inc]
    std::vector<uint64_t> inc_off;   // [n_inc] uncut elements only
    // nodes that share dofs (periodic images): only the first node holding a
    // dof owns its rows and incidence list; the others have empty lists
    bool shared_dofs = false;
    void build(const MeshView& mesh, const int32_t* dof, int64_t n_free, int64_t n_pres);
};

}  // namespace pgf
