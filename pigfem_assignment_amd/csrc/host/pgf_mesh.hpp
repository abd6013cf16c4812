// Host mesh for the PIGFEM MI355X path: structured background mesh, circle /
// sphere inclusions, IGFEM node and element detection, enrichment nodes and the
// per-quadrature-point tables of cut elements.
//
// Mirrors the reference's Mesh surface (src/Mesh.cpp): generate_mesh
// (:191-583), add_inclusion (:2592), analyze_nodes / detect_node /
// modify_node_location (:2676-2918), analyze_elements (:2922-3027),
// add_enrichments serial branch (:3064-3202), add_enrich_node_to_elem
// (:3600-3656) and store_shape_functions (:1443-1605) for cut elements.
// Storage is structure-of-arrays so it can be shipped to HBM unchanged.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "pgf_common.hpp"

namespace pgf {

struct Inclusion {
    int type = 0;         // 0 circle (2D), 1 sphere (3D)
    double c[3] = {0, 0, 0};
    double r = 0.0;
    int mat = 1;          // volumetric material of the inclusion
    int coh_law = -1;     // cohesive law on its interface (-1: mesh default)
};

struct ElasticMaterial {
    double E = 0.0, nu = 0.0;
};

// Cohesive traction-separation law parameters.
//   law 0: Ortiz-Pandolfi exponential with linear unloading and viscous term,
//          the active law of src/Material_OPCohesive.cpp:146-166,248-253.
//   law 1: bilinear (linear rise to sigma_c at delta_c, linear softening to
//          zero at delta_f), the parameterised form of the commented law at
//          src/Material_OPCohesive.cpp:175-199,256-271.
struct CohesiveLaw {
    int law = 0;
    double sigma_c = 0.0, delta_c = 0.0, delta_f = 0.0, beta = 1.0, visc = 0.0;
    double t_floor = 0.0;  // residual traction floor of the bilinear law (0 = off)
};

// One cut (intersected) background element.  Node slots 0..nn-1 are the
// element's regular nodes, slots nn.. its enrichment nodes in cut-edge order,
// followed (cohesive meshes) by the mirror nodes.
struct CutElem {
    int32_t elem = -1;
    uint8_t n_cut = 0;
    uint8_t cut_edge[8] = {0};
    int32_t enrich_incl[8] = {0};
    uint8_t n_sub = 0;
    uint8_t sub_nn[MAX_SUBELEM] = {0};
    uint8_t sub_nodes[MAX_SUBELEM][8] = {{0}};
    int32_t sub_mat[MAX_SUBELEM] = {0};
    uint8_t n_coh = 0;
    uint8_t coh_nn[MAX_COHELEM] = {0};
    uint8_t coh_nodes[MAX_COHELEM][8] = {{0}};
    int32_t coh_law[MAX_COHELEM] = {0};
    uint8_t n_enrich = 0;        // n_cut, or 2*n_cut with mirror nodes
    int32_t enrich[16] = {0};    // enrichment-node index per enrichment slot
};

// the device set-up writes these records with the same layout
static_assert(sizeof(CutElem) == 224, "CutElem layout");

struct MeshOptions {
    bool cohesive = false;
    bool plane_strain = true;
    // Reproduce the reference's sphere bounding box, whose z upper bound is
    // c_y + r instead of c_z + r (src/InclusionSphere.cpp:158).  Off by default;
    // with it on, spheres with c_z > c_y lose nodes or fail root bracketing
    // exactly as the reference does.
    bool sphere_zlim_as_reference = false;
    // Scale applied to every volumetric tet weight.  1.0 keeps the reference's
    // 5-point rule whose weights sum to 1 with J = 6*volume (src/tet4.cpp:114-121);
    // 1/6 gives true volumes.
    double tet_weight_scale = 1.0;
    int default_coh_law = 0;
};

class Mesh;
// IGFEM preprocessing on the current HIP device (kernels/pgf_preprocess.hip):
// fills the members analyze_nodes, analyze_elements, add_enrichments and
// store_shape_functions fill, with the same bits; phase_ms[4] receives the
// wall time of the node, element, enrichment and table phases (transfers
// included).
void device_preprocess(Mesh& m, double* phase_ms);

class Mesh {
   public:
    // ---- construction (src/Mesh.cpp:191-583) ----
    void generate_mesh(ElemKind kind, const double* mins, const double* maxs, const int* N);
    void add_inclusion(const Inclusion& inc) { _incl.push_back(inc); }
    void set_options(const MeshOptions& o) { _opt = o; }

    // ---- IGFEM preprocessing ----
    void analyze_nodes();          // node detection + surface nudges
    void analyze_elements();       // element detection tables
    void add_enrichments();        // enrichment + mirror nodes, element slots
    void store_shape_functions();  // cut-element integration-point tables

    // ---- accessors ----
    int dim() const { return _dim; }
    ElemKind kind() const { return _kind; }
    int nodes_per_elem() const { return _nn; }
    int64_t n_nodes() const { return _n_nodes; }
    int64_t n_enrich_nodes() const { return (int64_t)_enrich_edge.size(); }
    int64_t n_all_nodes() const { return _n_nodes + n_enrich_nodes(); }
    int64_t n_elem() const { return _n_elem; }
    const int* N() const { return _N; }
    const std::vector<double>& X(int d) const { return _X[d]; }
    std::vector<double>& X(int d) { return _X[d]; }
    const std::vector<int32_t>& conn() const { return _conn; }
    const std::vector<int32_t>& elem_mat() const { return _elem_mat; }
    std::vector<int32_t>& elem_mat() { return _elem_mat; }
    const std::vector<int32_t>& node_detect() const { return _node_detect; }
    const std::vector<CutElem>& cut() const { return _cut; }
    const std::vector<int32_t>& cut_index() const { return _cut_index; }
    const std::vector<std::array<int32_t, 3>>& enrich_edge() const { return _enrich_edge; }  // (n0, n1, inclusion)
    const std::vector<Inclusion>& inclusions() const { return _incl; }
    const MeshOptions& options() const { return _opt; }
    int64_t n_moved_nodes() const { return _n_moved; }
    bool is_cohesive() const { return _opt.cohesive; }

    // nodesets of the generated mesh (upper-case names as in the reference)
    std::vector<int32_t> nodeset(const std::string& name) const;

    // ---- periodic RVE (src/Mesh.cpp:603-741) ----
    // Periodic images of a generated grid: every node with a coordinate index
    // at the upper bound N is an image of the node with that index 0 (corners,
    // edges and faces alike), so each periodic set's primary node -- the
    // lowest id of the set, the reference's set[0] -- is the node whose grid
    // indices are (x mod Nx, y mod Ny, z mod Nz).
    void build_periodic_nodesets();
    bool periodic() const { return !_periodic_primary.empty(); }
    // primary node of every regular node's periodic set (itself if none)
    const std::vector<int32_t>& periodic_primary() const { return _periodic_primary; }

    // element-local node slot -> all-node id (regular i -> i, enrichment k -> n_nodes + k)
    int32_t elem_node(int64_t e, int slot) const;
    int elem_n_total(int64_t e) const;

    // ---- cut-element integration tables (flat, CSR-like by cut element) ----
    // volumetric points: weight J*W, material, gradients of all n_total slots
    std::vector<int64_t> vpt_ptr;        // [n_cut+1]
    std::vector<double> vpt_w;
    std::vector<int32_t> vpt_mat;
    std::vector<int64_t> vpt_grad_off;   // offset of the point's grad block
    std::vector<double> vpt_grad;        // n_total*dim per point, node-major
    // cohesive points
    std::vector<int64_t> cpt_ptr;        // [n_cut+1]
    std::vector<double> cpt_w;           // dA*W
    std::vector<double> cpt_N;           // 4 per point (padded)
    std::vector<double> cpt_R;           // 9 per point (dim x dim, row-major, padded)
    std::vector<uint8_t> cpt_nodes;      // 8 per point: element slots, lower half then upper half
    std::vector<uint8_t> cpt_nh;         // nodes per surface (2, 3 or 4)
    std::vector<int32_t> cpt_law;
    int64_t n_coh_points() const { return (int64_t)cpt_w.size(); }

   private:
    int detect_node(int64_t n, int& times_moved);
    void move_node(int64_t n, const Inclusion& inc);
    void detect_element(int64_t e, const int* nd, CutElem& ce, bool& cut);
    // the device set-up (kernels/pgf_preprocess.hip) fills the same members
    friend struct DevicePreprocess;
    void edge_nodes(int edge, int& a, int& b) const;

    int _dim = 0, _nn = 0;
    ElemKind _kind = QUAD4;
    int _N[3] = {0, 0, 0};
    double _min[3] = {0, 0, 0}, _max[3] = {0, 0, 0};
    int64_t _n_nodes = 0, _n_elem = 0, _n_moved = 0;
    std::vector<double> _X[3];
    std::vector<int32_t> _conn;
    std::vector<int32_t> _elem_mat;
    std::vector<int32_t> _node_detect;
    std::vector<int32_t> _first_elem;  // lowest element id holding each node
    std::vector<Inclusion> _incl;
    std::vector<CutElem> _cut;
    std::vector<int32_t> _cut_index;   // per element: index into _cut or -1
    std::vector<std::array<int32_t, 3>> _enrich_edge;
    std::vector<int32_t> _periodic_primary;
    MeshOptions _opt;
    bool _nodes_detected = false, _elements_detected = false, _enriched = false;
};

}  // namespace pgf
