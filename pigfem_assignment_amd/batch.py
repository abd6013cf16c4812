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


def periodic_rve_problem(N, seed, materials=(configs.MATRIX, configs.INCLUSION), inclusions=None):
    """Periodic N x N QUAD4 RVE with node 0 (all corners) fixed."""
    incl = rve_inclusions(N, seed) if inclusions is None else inclusions
    bcs = [(L.SET_NODE, 0, 0, 0.0), (L.SET_NODE, 0, 1, 0.0)]
    return HostProblem("quad4", (N, N), (0, 0), (1, 1), inclusions=incl, materials=materials, bcs=bcs,
                       periodic=True), incl


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(dtype, copy=True)


class BatchProblem:
    """Concatenation of independent HostProblems (same dimension, materials,
    element kind), exposing what DeviceContext needs: desc() and dofs()."""

    def __init__(self, problems):
        self.parts = problems
        descs = [p.desc() for p in problems]
        d0 = descs[0]
        dim = d0.dim
        self.dim = dim
        n_reg = [d.n_nodes for d in descs]
        n_enr = [d.n_all_nodes - d.n_nodes for d in descs]
        NR, NE = sum(n_reg), sum(n_enr)
        n_all = NR + NE
        R_off = np.concatenate([[0], np.cumsum(n_reg)])
        E_off = np.concatenate([[0], np.cumsum(n_enr)])
        nf = [d.n_free for d in descs]
        npr = [d.n_pres for d in descs]
        F_off = np.concatenate([[0], np.cumsum(nf)])
        P_off = np.concatenate([[0], np.cumsum(npr)])
        NF, NP = int(F_off[-1]), int(P_off[-1])
        self.sys_ptr = F_off.astype(np.int64)
        self.pres_ptr = P_off.astype(np.int64)
        coords = np.zeros((3, n_all))
        dof = np.zeros((dim, n_all), np.int32)
        conn, emat, cut_elem, cut_ne, cut_en = [], [], [], [], []
        vpt_ptr, vpt_w, vpt_mat, vpt_goff, vpt_grad = [np.zeros(1, np.int64)], [], [], [], []
        e_off = v_off = g_off = 0
        pres = []
        for k, (p, d) in enumerate(zip(problems, descs)):
            na = d.n_all_nodes
            X = _arr(d.coords, 3 * na, np.float64).reshape(3, na)
            reg = slice(R_off[k], R_off[k + 1])
            enr = slice(NR + E_off[k], NR + E_off[k + 1])
            coords[:, reg] = X[:, : d.n_nodes]
            coords[:, enr] = X[:, d.n_nodes:]
            dk = _arr(d.dof, dim * na, np.int32).reshape(dim, na)
            newd = np.where(dk < d.n_free, dk + F_off[k], NF + P_off[k] + (dk - d.n_free)).astype(np.int32)
            dof[:, reg] = newd[:, : d.n_nodes]
            dof[:, enr] = newd[:, d.n_nodes:]
            conn.append((_arr(d.conn, 4 * d.n_elem, np.int32) + R_off[k]).astype(np.int32))
            emat.append(_arr(d.elem_mat, d.n_elem, np.int32))
            nc = d.n_cut
            cut_elem.append((_arr(d.cut_elem, nc, np.int32) + e_off).astype(np.int32))
            cut_ne.append(_arr(d.cut_n_enrich, nc, np.int32))
            ce = _arr(d.cut_enrich, 16 * nc, np.int32).reshape(max(nc, 0), 16)
            ne = cut_ne[-1]
            mask = np.arange(16)[None, :] < ne[:, None]
            ce = np.where(mask, ce + E_off[k], ce).astype(np.int32)
            cut_en.append(ce.ravel())
            vp = _arr(d.vpt_ptr, nc + 1, np.int64) if nc else np.zeros(1, np.int64)
            nv = int(vp[-1])
            vpt_ptr.append((vp[1:] + v_off).astype(np.int64))
            vpt_w.append(_arr(d.vpt_w, nv, np.float64))
            vpt_mat.append(_arr(d.vpt_mat, nv, np.int32))
            go = _arr(d.vpt_grad_off, nv, np.int64)
            ng = int(p.info.n_vgrad)
            vpt_goff.append((go + g_off).astype(np.int64))
            vpt_grad.append(_arr(d.vpt_grad, ng, np.float64))
            if p.info.n_cpt > 0:
                raise ValueError("cohesive RVEs are not batched")
            e_off += d.n_elem
            v_off += nv
            g_off += ng
            pres.append(p.dofs()[1])
        self.n_nodes, self.n_all_nodes, self.n_elem = NR, n_all, e_off
        self.elem_ptr = np.concatenate([[0], np.cumsum([d.n_elem for d in descs])]).astype(np.int64)
        self.reg_ptr = R_off.astype(np.int64)
        self.n_free, self.n_pres = NF, NP
        keep = dict(
            coords=np.ascontiguousarray(coords.ravel()), dof=np.ascontiguousarray(dof.ravel()),
            conn=np.ascontiguousarray(np.concatenate(conn)), emat=np.ascontiguousarray(np.concatenate(emat)),
            cut_elem=np.ascontiguousarray(np.concatenate(cut_elem)), cut_ne=np.ascontiguousarray(np.concatenate(cut_ne)),
            cut_en=np.ascontiguousarray(np.concatenate(cut_en)), vpt_ptr=np.ascontiguousarray(np.concatenate(vpt_ptr)),
            vpt_w=np.ascontiguousarray(np.concatenate(vpt_w)), vpt_mat=np.ascontiguousarray(np.concatenate(vpt_mat)),
            vpt_goff=np.ascontiguousarray(np.concatenate(vpt_goff)),
            vpt_grad=np.ascontiguousarray(np.concatenate(vpt_grad)),
            cpt_ptr=np.zeros(len(np.concatenate(cut_elem)) + 1, np.int64),
            mat_E=np.ascontiguousarray([m[0] for m in problems[0].materials], np.float64),
            mat_nu=np.ascontiguousarray([m[1] for m in problems[0].materials], np.float64),
            pres=np.ascontiguousarray(np.concatenate(pres)))
        self._keep = keep
        n_cut = len(keep["cut_elem"])
        self.n_cut = n_cut
        D = L.MeshDesc()
        D.dim, D.elem_kind, D.plane_strain = dim, d0.elem_kind, d0.plane_strain
        D.n_nodes, D.n_all_nodes = NR, n_all
        D.coords = L.ptr(keep["coords"], C.c_double)
        D.n_elem = e_off
        D.conn = L.ptr(keep["conn"], C.c_int32)
        D.elem_mat = L.ptr(keep["emat"], C.c_int32)
        D.n_mat = len(keep["mat_E"])
        D.n_coh_law = 0
        D.mat_E = L.ptr(keep["mat_E"], C.c_double)
        D.mat_nu = L.ptr(keep["mat_nu"], C.c_double)
        D.n_cut = n_cut
        D.cut_elem = L.ptr(keep["cut_elem"], C.c_int32)
        D.cut_n_enrich = L.ptr(keep["cut_ne"], C.c_int32)
        D.cut_enrich = L.ptr(keep["cut_en"], C.c_int32)
        D.vpt_ptr = L.ptr(keep["vpt_ptr"], C.c_int64)
        D.vpt_w = L.ptr(keep["vpt_w"], C.c_double)
        D.vpt_mat = L.ptr(keep["vpt_mat"], C.c_int32)
        D.vpt_grad_off = L.ptr(keep["vpt_goff"], C.c_int64)
        D.vpt_grad = L.ptr(keep["vpt_grad"], C.c_double)
        D.cpt_ptr = L.ptr(keep["cpt_ptr"], C.c_int64)
        D.n_free, D.n_pres = NF, NP
        D.dof = L.ptr(keep["dof"], C.c_int32)
        self._desc = D
        self.coords_full = coords
        self.dof_full = dof

    def desc(self):
        return self._desc

    def dofs(self):
        return self.dof_full, self._keep["pres"]

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
