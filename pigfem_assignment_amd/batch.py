"""Batches of 2D RVEs solved together on one device context.

Two homogenisation set-ups are provided:

* periodic (the reference's linear subscale model, SubscaleModelLinear,
  src/SubscaleModelLinear.cpp): a periodic mesh whose upper-face nodes share
  the dofs of their lower-face images (src/Mesh.cpp:603-741), node 0 fixed
  (src/SubscaleModel.cpp:232-243), the macro strain applied as the body load
  -int B^T D eps_bar (src/BodyLoadMacroStrainLinear.cpp:10-30), and the
  effective stiffness from the volume-averaged stress
  (src/SubscaleModel.cpp:42-196);
* affine boundary displacements, described below.

Each RVE is an N x N QUAD4 unit square with seeded circular inclusions, built
by the host IGFEM set-up exactly like a single problem.  The RVE meshes are
concatenated into one mesh description (regular nodes of all RVEs, then their
enrichment nodes; free dofs numbered RVE by RVE so K_ff is block diagonal with
contiguous blocks; prescribed dofs after all free dofs), so one assembly
launch covers every element of the batch and pgf_batch_pcg solves every RVE
for all of its load cases at once.

Homogenisation load cases are affine displacements u = E x on the RVE boundary
(E = e_xx, e_yy, gamma_xy), the kinematic counterpart of the reference's
macro-strain RVE problem (src/SubscaleModel.cpp:42-196,
src/BodyLoadMacroStrainLinear.cpp:10-30); the effective stiffness follows from
the boundary reactions, sigma_ij = (1/A) sum_b R_i x_j.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from . import configs
from .api import HostProblem

# Voigt strains of the three load cases: e_xx, e_yy, gamma_xy
CASES = [np.array([[1.0, 0.0], [0.0, 0.0]]), np.array([[0.0, 0.0], [0.0, 1.0]]),
         np.array([[0.0, 0.5], [0.5, 0.0]])]


def rve_inclusions(N, seed, vf_lo=0.2, vf_hi=0.4, n_lo=6, n_hi=20):
    """Seeded RSA circles with a target area fraction U[vf_lo, vf_hi]."""
    rng = np.random.RandomState(seed)
    n = int(rng.randint(n_lo, n_hi + 1))
    vf = rng.uniform(vf_lo, vf_hi)
    gap = 2.0 / N  # keeps every element cut by at most one interface
    while True:
        r = np.sqrt(vf / (n * np.pi))
        try:
            return configs.rsa(n, 0.85 * r, 1.15 * r, 2, gap=gap, seed=seed, max_tries=20_000)
        except RuntimeError:
            # too dense for this mesh: fewer inclusions, then a lower fraction
            if n > 2:
                n -= 1
            else:
                vf *= 0.9
                if vf < 0.02:
                    raise


def rve_problem(N, seed, materials=(configs.MATRIX, configs.INCLUSION), inclusions=None):
    incl = rve_inclusions(N, seed) if inclusions is None else inclusions
    bcs = []
    for y in range(N + 1):
        for x in range(N + 1):
            if x in (0, N) or y in (0, N):
                node = y * (N + 1) + x
                bcs.append((L.SET_NODE, node, 0, 0.0))
                bcs.append((L.SET_NODE, node, 1, 0.0))
    return HostProblem("quad4", (N, N), (0, 0), (1, 1), inclusions=incl, materials=materials, bcs=bcs), incl


def periodic_rve_problem(N, seed, materials=(configs.MATRIX, configs.INCLUSION), inclusions=None,
                         device_preprocess=False):
    """Periodic N x N QUAD4 RVE with node 0 (all corners) fixed."""
    incl = rve_inclusions(N, seed) if inclusions is None else inclusions
    bcs = [(L.SET_NODE, 0, 0, 0.0), (L.SET_NODE, 0, 1, 0.0)]
    return HostProblem("quad4", (N, N), (0, 0), (1, 1), inclusions=incl, materials=materials, bcs=bcs,
                       periodic=True, device_preprocess=device_preprocess), incl


class BatchProblem:
    """Concatenation of independent HostProblems (same dimension, materials,
    element kind) built by the library (pgfh_batch_create), exposing what
    DeviceContext needs: desc() and dofs()."""

    def __init__(self, problems):
        self.parts = problems
        lib = L.lib()
        lib.pgfh_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
        lib.pgfh_batch_destroy.argtypes = [C.c_void_p]
        lib.pgfh_batch_destroy.restype = None
        lib.pgfh_batch_desc.argtypes = [C.c_void_p, C.POINTER(L.MeshDesc)]
        lib.pgfh_batch_offsets.argtypes = [C.c_void_p] + [L.i64p] * 5
        lib.pgfh_batch_get_dofs.argtypes = [C.c_void_p, L.i32p, L.dp]
        arr = (C.c_void_p * len(problems))(*[p.h for p in problems])
        h = C.c_void_p()
        L.check(lib.pgfh_batch_create(arr, len(problems), C.byref(h)), host=True)
        self.h = h
        D = L.MeshDesc()
        L.check(lib.pgfh_batch_desc(h, C.byref(D)), host=True)
        self._desc = D
        self.dim = D.dim
        self.n_nodes, self.n_all_nodes, self.n_elem = D.n_nodes, D.n_all_nodes, D.n_elem
        self.n_free, self.n_pres, self.n_cut = D.n_free, D.n_pres, D.n_cut
        offs = [np.zeros(len(problems) + 1, np.int64) for _ in range(5)]
        L.check(lib.pgfh_batch_offsets(h, *[L.ptr(o, C.c_int64) for o in offs]), host=True)
        self.sys_ptr, self.pres_ptr, self.elem_ptr, self.reg_ptr, self.enr_ptr = offs
        self.dof_full = np.zeros((self.dim, self.n_all_nodes), np.int32)
        self._pres = np.zeros(max(self.n_pres, 1))
        L.check(lib.pgfh_batch_get_dofs(h, L.ptr(self.dof_full, C.c_int32), L.ptr(self._pres, C.c_double)),
                host=True)
        self._pres = self._pres[: self.n_pres]
        self.coords_full = np.ctypeslib.as_array(D.coords, shape=(3 * self.n_all_nodes,)).reshape(
            3, self.n_all_nodes).copy()

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            L.lib().pgfh_batch_destroy(h)
            self.h = None

    def desc(self):
        return self._desc

    def dofs(self):
        return self.dof_full, self._pres

    def case_values(self):
        """Prescribed values of the three load cases, [3][n_pres], from the
        original grid coordinates of the prescribed nodes."""
        UP = np.zeros((len(CASES), self.n_pres))
        for d in range(self.dim):
            ids = self.dof_full[d]
            sel = np.nonzero(ids >= self.n_free)[0]
            x = self.coords_full[: self.dim, sel]
            for c, E in enumerate(CASES):
                UP[c, ids[sel] - self.n_free] = E[d] @ x
        return UP

    def pres_coords(self):
        """(component, x, y) of every prescribed dof."""
        out = np.zeros((self.n_pres, 3))
        for d in range(self.dim):
            ids = self.dof_full[d]
            sel = np.nonzero(ids >= self.n_free)[0]
            out[ids[sel] - self.n_free, 0] = d
            out[ids[sel] - self.n_free, 1:] = self.coords_full[:2, sel].T
        return out


def effective_stiffness(batch, R):
    """C_eff [n_rve][3][3] (Voigt, gamma_xy) from reactions R [3][n_pres] of the
    three cases: sigma_ij = sum_b R_i x_j over the RVE's boundary (area 1)."""
    pc = batch.pres_coords()
    n_rve = len(batch.parts)
    Ce = np.zeros((n_rve, 3, 3))
    for k in range(n_rve):
        sl = slice(batch.pres_ptr[k], batch.pres_ptr[k + 1])
        comp, x, y = pc[sl, 0].astype(int), pc[sl, 1], pc[sl, 2]
        for c in range(3):
            Rk = R[c, sl]
            sxx = np.sum(Rk[comp == 0] * x[comp == 0])
            syy = np.sum(Rk[comp == 1] * y[comp == 1])
            sxy = 0.5 * (np.sum(Rk[comp == 0] * y[comp == 0]) + np.sum(Rk[comp == 1] * x[comp == 1]))
            Ce[k, :, c] = (sxx, syy, sxy)
    return Ce


# Voigt macro strains of the three load cases (engineering shear)
VOIGT_CASES = [np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])]


def affine_field(prob, E):
    """nodal field E x on the regular nodes, zero enrichment amplitudes"""
    X = prob.coords_full if hasattr(prob, "coords_full") else prob.coords()
    nr = prob.n_nodes if hasattr(prob, "n_nodes") else prob.info.n_nodes
    U = np.zeros((2, X.shape[1]))
    U[:, :nr] = E @ X[:2, :nr]
    return U


def macro_strain_rhs(ctx, prob, cases=CASES):
    """Right-hand sides of the macro-strain load cases, [n_case][n_free]: the
    body load -int B^T D eps_bar assembled on the device as minus the internal
    force of the affine field eps_bar x (its strain is eps_bar at every
    integration point, cut elements included), periodic images combined."""
    nf = ctx.csr.n_free
    B = np.zeros((len(cases), nf))
    for k, E in enumerate(cases):
        ctx.set_solution(affine_field(prob, E))
        ctx.assemble(L.PGF_ASM_FUNC)
        B[k] = -ctx.internal_force()[:nf]
    return B


def homogenized_stiffness(ctx, prob, X, seg_elem, cases=VOIGT_CASES):
    """C_eff [n_seg][3][3] (Voigt) from the volume-averaged stress of each load
    case's fluctuation solution X[k] (free dofs; prescribed dofs zero)."""
    n_seg = len(seg_elem) - 1
    C_eff = np.zeros((n_seg, 3, 3))
    for k, e in enumerate(cases):
        ctx.set_solution_dofs(X[k])
        sig, _, _ = ctx.volume_average(e, seg_elem)
        C_eff[:, :, k] = sig
    return C_eff
