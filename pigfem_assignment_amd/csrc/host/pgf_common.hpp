// Common host-side types for the PIGFEM MI355X path.
//
// Error convention: every fallible host routine returns a pgf::Status.  The
// reference aborts the process through err_message() -> exit(0)
// (include/common.h:292-297); here failures travel back to the caller as a
// negative status code plus a message, and nothing ever exits.
#pragma once

#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>

// Functions marked PGF_HD are compiled for the host (g++) and, from the HIP
// set-up kernels, for the device; that code is built without contraction of
// multiply-adds so both sides round every operation the same way.
#if defined(__HIPCC__)
#define PGF_HD __host__ __device__
#else
#define PGF_HD
#endif

namespace pgf {

enum Status : int {
    ST_OK = 0,
    PGF_ERR_ARG = -1,          // invalid argument / shape
    PGF_ERR_UNSUPPORTED = -2,  // geometric case the reference also rejects
    PGF_ERR_GEOMETRY = -3,     // negative Jacobian, bad inverse map, no root
    PGF_ERR_DEVICE = -4,       // HIP runtime failure
    PGF_ERR_NOT_CONVERGED = -5,
    PGF_ERR_STATE = -6,        // call sequence violated
    PGF_ERR_LIMIT = -7,        // a table limit (uint8 offsets, node counts) exceeded
};

// Thrown internally by host preprocessing; converted to a Status at the C-ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum ElemKind : int { QUAD4 = 0, TET4 = 1 };

// Largest element-local node count the tables support: a tet cut 2-vs-2 with
// cohesive mirrors has 4 regular + 4 enrichment + 4 mirror nodes.
constexpr int MAX_EL_NODES = 12;
constexpr int MAX_SUBELEM = 6;
constexpr int MAX_COHELEM = 2;

// Isotropic linear-elastic constitutive matrix in engineering-shear Voigt
// notation.  2D order [xx, yy, xy], 3D order [xx, yy, zz, yz, xz, xy].
// Follows LinearElasticIsotropicMaterial::computeLinearDmat
// (src/material_lei.cpp:52-97): plane strain and 3D use E/((1+nu)(1-2nu)),
// plane stress uses E/(1-nu^2).
inline void isotropic_D(int dim, bool plane_strain, double E, double nu, double* D) {
    const int s = dim == 2 ? 3 : 6;
    for (int i = 0; i < s * s; ++i) D[i] = 0.0;
    if (dim == 2) {
        if (plane_strain) {
            double coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
            D[0] = coef * (1.0 - nu); D[1] = coef * nu;
            D[3] = coef * nu;         D[4] = coef * (1.0 - nu);
            D[8] = coef * (1.0 - 2.0 * nu) / 2.0;
        } else {
            double coef = E / (1.0 - nu * nu);
            D[0] = coef;      D[1] = coef * nu;
            D[3] = coef * nu; D[4] = coef;
            D[8] = coef * (1.0 - nu) / 2.0;
        }
    } else {
        double coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[i * 6 + j] = coef * (i == j ? (1.0 - nu) : nu);
        for (int i = 3; i < 6; ++i) D[i * 6 + i] = coef * (1.0 - 2.0 * nu) / 2;
    }
}

}  // namespace pgf
