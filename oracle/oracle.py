"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                 C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                 C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_sizes.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_dofs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_pattern.restype = C.c_int64
        L.orc_pattern.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_solution.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get_solution.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_assemble.restype = C.c_double
        L.orc_assemble.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_isv.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_assemble_abs.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        L.orc_isv_commit.argtypes = [C.c_void_p]
        L.orc_pcg_last.restype = C.c_int
        L.orc_pcg_last.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int]
        L.orc_newton_step.restype = C.c_int
        L.orc_newton_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                      C.c_int, C.c_void_p]
        L.orc_detect_nodes.argtypes = [C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_law.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        L.orc_sample_range.argtypes = [C.c_int64, C.c_int64]
        L.orc_shape.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_periodic_grid.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_set_fext.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_macro_strain_load.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_homogenize.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_pcg_csr_ms_per_iter.restype = C.c_double
        L.orc_pcg_csr_ms_per_iter.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_int, C.c_int]
        L.orc_edge_root.restype = C.c_double
        L.orc_edge_root.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """Reference-structured CPU model built from a HostProblem's geometry and
    cut topology.  `bc` is a list of (node, dof, value)."""

    def __init__(self, prob, bcs, sample=None):
        """sample=(e0, e1) builds and assembles only elements [e0, e1) (timing)."""
        L = lib()
        L.orc_sample_range(*(sample if sample else (-1, -1)))
        L.orc_periodic_grid(*(prob.grid if getattr(prob, "periodic", False) else (0, 0, 0)))
        self.prob = prob
        i = prob.info
        self.dim = i.dim
        self.n_all = i.n_all_nodes
        X = np.ascontiguousarray(prob.coords())
        conn = np.ascontiguousarray(prob.conn())
        em = np.ascontiguousarray(prob.elem_mat())
        E = np.ascontiguousarray(prob.materials[:, 0])
        nu = np.ascontiguousarray(prob.materials[:, 1])
        laws = np.ascontiguousarray(np.array([list(lw) + [0.0] * (7 - len(lw)) for lw in prob.coh_laws]
                                             if prob.coh_laws else np.zeros((0, 7)), np.float64))
        cut = np.ascontiguousarray(prob.cut())
        bn = np.ascontiguousarray(np.array([b[0] for b in bcs], np.int32))
        bd = np.ascontiguousarray(np.array([b[1] for b in bcs], np.int32))
        bv = np.ascontiguousarray(np.array([b[2] for b in bcs], np.float64))
        self._keep = [X, conn, em, E, nu, laws, cut, bn, bd, bv]
        self.h = L.orc_create(self.dim, int(prob.plane_strain), 1.0 if prob.kind != "tet4" else prob_tet_scale(prob),
                              i.n_nodes, i.n_all_nodes, _p(X), i.n_elem, _p(conn), _p(em), len(E), _p(E), _p(nu),
                              laws.shape[0], _p(laws), i.n_cut, _p(cut), len(bcs), _p(bn), _p(bd), _p(bv))
        L.orc_sample_range(-1, -1)
        L.orc_periodic_grid(0, 0, 0)
        sz = np.zeros(4, np.int64)
        L.orc_sizes(self.h, _p(sz))
        self.n_free, self.n_pres, self.n_isv, self.n_qp = [int(v) for v in sz]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def dofs(self):
        dof = np.zeros((self.dim, self.n_all), np.int32)
        pres = np.zeros(max(self.n_pres, 1))
        lib().orc_dofs(self.h, _p(dof), _p(pres))
        return dof, pres[: self.n_pres]

    def pattern(self):
        n = self.n_free + self.n_pres
        nnz = lib().orc_pattern(self.h, None, None)
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(nnz, np.int32)
        lib().orc_pattern(self.h, _p(rp), _p(ci))
        return rp, ci

    def set_solution(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        lib().orc_set_solution(self.h, _p(U))

    def solution(self):
        U = np.zeros((self.dim, self.n_all))
        lib().orc_get_solution(self.h, _p(U))
        return U

    def assemble(self, dt=1.0, bvariant=False):
        """returns (vals in orc_pattern CSR order, P over all rows, ms)"""
        nnz = lib().orc_pattern(self.h, None, None)
        vals = np.zeros(nnz)
        P = np.zeros(self.n_free + self.n_pres)
        ms = lib().orc_assemble(self.h, dt, int(bvariant), _p(vals), _p(P))
        return vals, P, ms

    def assemble_abs(self, dt=1.0):
        """per K entry (orc_pattern order) the sum of the absolute values of the
        products that enter it: the scale of its rounding error"""
        nnz = lib().orc_pattern(self.h, None, None)
        vals = np.zeros(nnz)
        lib().orc_assemble_abs(self.h, dt, _p(vals))
        return vals

    def set_fext(self, f_free):
        """external load of the Newton residual (free rows)"""
        lib().orc_set_fext(self.h, _p(np.ascontiguousarray(f_free, np.float64)))

    def macro_strain_load(self, eps_bar):
        """-int B^T D eps_bar over all rows (BodyLoadMacroStrainLinear)"""
        e = np.ascontiguousarray(eps_bar, np.float64)
        f = np.zeros(self.n_free + self.n_pres)
        lib().orc_macro_strain_load(self.h, _p(e), _p(f))
        return f

    def homogenize(self, eps_bar, energy=False):
        """(average stress, average fluctuation strain, volume[, average
        sigma : eps]) at the stored solution"""
        S = 3 if self.dim == 2 else 6
        e = np.ascontiguousarray(eps_bar, np.float64)
        sig, eps, vol, en = np.zeros(S), np.zeros(S), np.zeros(1), np.zeros(1)
        lib().orc_homogenize(self.h, _p(e), _p(sig), _p(eps), _p(vol), _p(en))
        if energy:
            return sig, eps, float(vol[0]), float(en[0])
        return sig, eps, float(vol[0])

    def isv(self):
        dt, dc = np.zeros(max(self.n_isv, 1)), np.zeros(max(self.n_isv, 1))
        fl = np.zeros(max(self.n_isv, 1), np.int32)
        lib().orc_isv(self.h, _p(dt), _p(dc), _p(fl))
        return dt[: self.n_isv], dc[: self.n_isv], fl[: self.n_isv]

    def isv_commit(self):
        lib().orc_isv_commit(self.h)

    def pcg_last(self, b, rtol=1e-14, maxit=100000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n_free)
        it = lib().orc_pcg_last(self.h, _p(b), _p(x), rtol, maxit)
        return x, it

    def newton_step(self, frac, dt=1.0, rel_tol=1e-8, abs_tol=0.0, max_iter=40, lin_rtol=1e-13, bvariant=False):
        info = np.zeros(3)
        it = lib().orc_newton_step(self.h, frac, dt, rel_tol, abs_tol, max_iter, lin_rtol, int(bvariant), _p(info))
        return it, info


def prob_tet_scale(prob):
    return getattr(prob, "tet_weight_scale", 1.0)


def detect_nodes(coords_regular, inclusions, conn):
    """Independent node detection (with surface nudges) on the original coords.
    Returns (detection, moved coords)."""
    X = np.ascontiguousarray(coords_regular.copy())
    n = X.shape[1]
    # "other" node of the nudge rule: node 0 (or 1) of the lowest element holding the node
    first = np.full(n, -1, np.int64)
    for e in range(conn.shape[0] - 1, -1, -1):
        first[conn[e]] = e
    other = np.where(conn[first, 0] == np.arange(n), conn[first, 1], conn[first, 0]).astype(np.int32)
    other = np.ascontiguousarray(other)
    inc = np.ascontiguousarray(np.asarray(inclusions, np.float64).reshape(-1, 5))
    det = np.zeros(n, np.int32)
    lib().orc_detect_nodes(n, _p(X), inc.shape[0], _p(inc), _p(other), _p(det))
    return det, X


def law(params, delta, dmax=0.0, dt=1.0):
    """traction, tangent, new delta_max, flag for one opening (ttn frame)."""
    p = np.ascontiguousarray(list(params) + [0.0] * (7 - len(params)), np.float64)
    d = np.ascontiguousarray(delta, np.float64)
    dim = len(d)
    out = np.zeros(dim + dim * dim + 2)
    lib().orc_law(_p(p), dim, _p(d), dmax, dt, _p(out))
    return out[:dim], out[dim:dim + dim * dim].reshape(dim, dim), out[-2], int(out[-1])


def edge_root(p0, p1, incl):
    p0 = np.ascontiguousarray(p0, np.float64)
    p1 = np.ascontiguousarray(p1, np.float64)
    incl = np.ascontiguousarray(incl, np.float64)
    return lib().orc_edge_root(_p(p0), _p(p1), _p(incl))


def shape(dim, nn, xi):
    """Parent shape values [nn] and gradients [nn][dim] at parent point xi;
    for the Wachspress n-gon (dim 2, nn >= 5) also its reference vertices
    [nn][2] and fan rule [4 nn][3] (xi, eta, weight)."""
    x = np.ascontiguousarray(xi, np.float64)
    N = np.zeros(nn)
    dN = np.zeros((nn, dim))
    vtx = np.zeros((nn, 2))
    rule = np.zeros((4 * nn, 3))
    poly = dim == 2 and nn >= 5
    lib().orc_shape(dim, nn, _p(x), _p(N), _p(dN), _p(vtx) if poly else None, _p(rule) if poly else None)
    return (N, dN, vtx, rule) if poly else (N, dN)


def set_threads(n):
    """threads of the oracle's element loop (1 = the serial reference loop)"""
    lib().orc_set_threads(int(n))


def pcg_ms_per_iter(rowptr, rowsplit, colidx, vals, b, iters, threads):
    """wall ms per Jacobi-PCG iteration on a given CSR K_ff (CPU baseline)"""
    rp = np.ascontiguousarray(rowptr, np.int64)
    rs = np.ascontiguousarray(rowsplit, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    return lib().orc_pcg_csr_ms_per_iter(len(bb), _p(rp), _p(rs), _p(ci), _p(v), _p(bb), int(iters), int(threads))
