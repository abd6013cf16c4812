// Legacy ASCII VTK output of an IGFEM mesh and its displacement field.
//
// Mirrors Writer_ASCII_VTK / Writer_VTK (src/Writer_ASCII_VTK.cpp:43-282,
// src/Writer_VTK.cpp:221-335, 639-730):
//   * POINTS: every regular node, then every enrichment node (cohesive
//     mirrors included), in id order;
//   * CELLS: an uncut element as itself, a cut element as its integration
//     sub-elements (VTK triangle 5, polygon 7, quad 9, tetra 10);
//   * POINT_DATA displacement: the nodal solution at regular nodes; at an
//     enrichment node the physical displacement, i.e. the regular field
//     interpolated along the parent edge plus the enrichment amplitude
//     (Problem::interpolate_nodal_enrichment); for a periodic RVE the macro
//     displacement eps_bar x is added to every node
//     (SubscaleModel::compute_macro_displacement);
//   * CELL_DATA material of every cell.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "pgf_mesh.hpp"

namespace pgf {

void write_vtk(const Mesh& m, const std::string& path, const double* u, const double* eps_bar, double t) {
    std::ofstream f(path.c_str());
    if (!f.good()) throw Error(PGF_ERR_ARG, "cannot open " + path + " for writing");
    const int dim = m.dim();
    const int64_t nn = m.n_nodes(), na = m.n_all_nodes();
    f << "# vtk DataFile Version 3.0\n";
    f << (m.cut().empty() ? "FEM " : "IGFEM ") << dim << "D simulation: t=" << t << "\n";
    f << "ASCII\n\nDATASET UNSTRUCTURED_GRID\n";
    f << "POINTS " << na << " float\n";
    f << std::setprecision(9);
    for (int64_t n = 0; n < na; ++n) f << (float)m.X(0)[n] << " " << (float)m.X(1)[n] << " " << (float)m.X(2)[n] << "\n";
    // cells
    std::vector<int64_t> eptr{0};
    std::vector<int32_t> eind, emat;
    std::vector<uint8_t> etype;
    const auto& cidx = m.cut_index();
    for (int64_t e = 0; e < m.n_elem(); ++e) {
        const int ci = cidx.empty() ? -1 : cidx[e];
        if (ci < 0) {
            for (int k = 0; k < 4; ++k) eind.push_back(m.conn()[e * 4 + k]);
            eptr.push_back((int64_t)eind.size());
            etype.push_back(dim == 2 ? 9 : 10);
            emat.push_back(m.elem_mat()[e]);
            continue;
        }
        const CutElem& ce = m.cut()[ci];
        for (int ie = 0; ie < ce.n_sub; ++ie) {
            const int cn = ce.sub_nn[ie];
            for (int k = 0; k < cn; ++k) eind.push_back(m.elem_node(e, ce.sub_nodes[ie][k]));
            eptr.push_back((int64_t)eind.size());
            etype.push_back(dim == 3 ? 10 : (cn == 3 ? 5 : (cn == 4 ? 9 : 7)));
            emat.push_back(ce.sub_mat[ie]);
        }
    }
    const int64_t ncell = (int64_t)etype.size();
    f << "\nCELLS " << ncell << " " << (ncell + (int64_t)eind.size()) << "\n";
    for (int64_t c = 0; c < ncell; ++c) {
        f << (eptr[c + 1] - eptr[c]);
        for (int64_t k = eptr[c]; k < eptr[c + 1]; ++k) f << " " << eind[k];
        f << "\n";
    }
    f << "\nCELL_TYPES " << ncell << "\n";
    for (uint8_t tp : etype) f << (int)tp << "\n";
    if (u) {
        f << "\nPOINT_DATA " << na << "\nVECTORS displacement float\n";
        const auto& edges = m.enrich_edge();
        for (int64_t n = 0; n < na; ++n) {
            double v[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) v[d] = u[(size_t)d * na + n];
            if (n >= nn) {
                // regular field along the parent edge at the enrichment point
                const auto& ed = edges[n - nn];
                double l2 = 0.0, s = 0.0;
                for (int d = 0; d < dim; ++d) {
                    const double dx = m.X(d)[ed[1]] - m.X(d)[ed[0]];
                    l2 += dx * dx;
                    s += dx * (m.X(d)[n] - m.X(d)[ed[0]]);
                }
                const double a = s / l2;
                for (int d = 0; d < dim; ++d)
                    v[d] += (1.0 - a) * u[(size_t)d * na + ed[0]] + a * u[(size_t)d * na + ed[1]];
            }
            if (eps_bar) {
                // macro displacement eps_bar x (Voigt, engineering shear)
                const double x = m.X(0)[n], y = m.X(1)[n], z = m.X(2)[n];
                if (dim == 2) {
                    v[0] += eps_bar[0] * x + 0.5 * eps_bar[2] * y;
                    v[1] += 0.5 * eps_bar[2] * x + eps_bar[1] * y;
                } else {
                    v[0] += eps_bar[0] * x + 0.5 * eps_bar[5] * y + 0.5 * eps_bar[4] * z;
                    v[1] += 0.5 * eps_bar[5] * x + eps_bar[1] * y + 0.5 * eps_bar[3] * z;
                    v[2] += 0.5 * eps_bar[4] * x + 0.5 * eps_bar[3] * y + eps_bar[2] * z;
                }
            }
            f << (float)v[0] << " " << (float)v[1] << " " << (float)v[2] << "\n";
        }
    }
    f << "\nCELL_DATA " << ncell << "\nSCALARS material int 1\nLOOKUP_TABLE default\n";
    for (int32_t mt : emat) f << mt << "\n";
    if (!f.good()) throw Error(PGF_ERR_ARG, "error writing " + path);
}

}  // namespace pgf
