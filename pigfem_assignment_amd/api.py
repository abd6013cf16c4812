"""Python handles over the C-ABI: HostProblem (pgfh_*) and DeviceContext (pgf_*).

These mirror the reference's driver flow (Tests/Main.cpp): build a mesh with
inclusions and materials, set Dirichlet conditions, Problem::init, then solve.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class HostProblem:
    """Mesh + IGFEM preprocessing + dof numbering (pgfh_problem_create)."""

    def __init__(self, kind, N, mins, maxs, inclusions=(), materials=((3.2e9, 0.35),), incl_mat=None,
                 plane_strain=True, cohesive=False, coh_laws=(), incl_coh_law=None, default_coh_law=0,
                 bcs=(), tet_weight_scale=1.0, sphere_zlim_as_reference=False, periodic=False,
                 device_preprocess=False):
        lib = L.lib()
        self._keep = []
        cfg = L.HostConfig()
        cfg.elem_kind = L.PGF_TET4 if kind == "tet4" else L.PGF_QUAD4
        for d in range(3):
            cfg.N[d] = int(N[d]) if d < len(N) else 0
            cfg.mins[d] = float(mins[d]) if d < len(mins) else 0.0
            cfg.maxs[d] = float(maxs[d]) if d < len(maxs) else 0.0
        inc = np.ascontiguousarray(np.asarray(inclusions, dtype=np.float64).reshape(-1, 5))
        cfg.n_incl = inc.shape[0]
        cfg.incl = L.ptr(inc, C.c_double)
        im = np.ascontiguousarray(np.full(inc.shape[0], 1, np.int32) if incl_mat is None
                                  else np.asarray(incl_mat, np.int32))
        cfg.incl_mat = L.ptr(im, C.c_int32)
        icl = None if incl_coh_law is None else np.ascontiguousarray(np.asarray(incl_coh_law, np.int32))
        cfg.incl_coh_law = L.ptr(icl, C.c_int32)
        mats = np.asarray(materials, dtype=np.float64).reshape(-1, 2)
        E = np.ascontiguousarray(mats[:, 0])
        nu = np.ascontiguousarray(mats[:, 1])
        cfg.n_mat = len(E)
        cfg.mat_E = L.ptr(E, C.c_double)
        cfg.mat_nu = L.ptr(nu, C.c_double)
        cfg.plane_strain = int(bool(plane_strain))
        cfg.cohesive = int(bool(cohesive))
        laws = list(coh_laws)
        ct = np.ascontiguousarray(np.array([lw[0] for lw in laws], np.int32)) if laws else None
        cp = np.ascontiguousarray(np.array([list(lw[1:]) + [0.0] * (7 - len(lw)) for lw in laws],
                                           np.float64)[:, :6]) if laws else None
        cfg.n_coh_law = len(laws)
        cfg.coh_type = L.ptr(ct, C.c_int32)
        cfg.coh_params = L.ptr(cp, C.c_double)
        cfg.default_coh_law = default_coh_law
        cfg.sphere_zlim_as_reference = int(bool(sphere_zlim_as_reference))
        cfg.tet_weight_scale = tet_weight_scale
        bset = np.ascontiguousarray(np.array([b[0] for b in bcs], np.int32))
        bnode = np.ascontiguousarray(np.array([b[1] for b in bcs], np.int32))
        bdof = np.ascontiguousarray(np.array([b[2] for b in bcs], np.int32))
        bval = np.ascontiguousarray(np.array([b[3] for b in bcs], np.float64))
        cfg.n_bc = len(bcs)
        cfg.bc_set, cfg.bc_node = L.ptr(bset, C.c_int32), L.ptr(bnode, C.c_int32)
        cfg.bc_dof, cfg.bc_val = L.ptr(bdof, C.c_int32), L.ptr(bval, C.c_double)
        cfg.periodic = int(bool(periodic))
        cfg.device_preprocess = int(bool(device_preprocess))
        self._keep += [inc, im, icl, E, nu, ct, cp, bset, bnode, bdof, bval]
        h = C.c_void_p()
        L.check(lib.pgfh_problem_create(C.byref(cfg), C.byref(h)), host=True)
        self.h = h
        self.info = L.HostInfo()
        L.check(lib.pgfh_problem_info(self.h, C.byref(self.info)), host=True)
        self.materials = mats
        self.coh_laws = laws
        self.plane_strain = plane_strain
        self.kind = kind
        self.tet_weight_scale = tet_weight_scale
        self.inclusions = inc
        self.periodic = bool(periodic)
        self.grid = tuple(int(N[d]) if d < len(N) else 0 for d in range(3))

    def bc_list(self):
        """Dirichlet conditions as (node, dof, full-load value) over regular
        nodes (for a periodic mesh only the nodes that own their dofs)."""
        dof, pres = self.dofs()
        nf = self.info.n_free
        out = []
        seen = set()
        for d in range(self.dim):
            for n in np.nonzero(dof[d, : self.info.n_nodes] >= nf)[0]:
                if (d, int(dof[d, n])) in seen:
                    continue
                seen.add((d, int(dof[d, n])))
                out.append((int(n), d, float(pres[dof[d, n] - nf])))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            L.lib().pgfh_problem_destroy(self.h)
            self.h = None

    # ---- exports ----
    @property
    def dim(self):
        return self.info.dim

    def coords(self):
        a = np.zeros((3, self.info.n_all_nodes))
        L.lib().pgfh_get_coords(self.h, L.ptr(a, C.c_double))
        return a

    def conn(self):
        a = np.zeros((self.info.n_elem, 4), np.int32)
        L.lib().pgfh_get_conn(self.h, L.ptr(a, C.c_int32))
        return a

    def elem_mat(self):
        a = np.zeros(self.info.n_elem, np.int32)
        L.lib().pgfh_get_elem_mat(self.h, L.ptr(a, C.c_int32))
        return a

    def node_detect(self):
        a = np.zeros(self.info.n_nodes, np.int32)
        L.lib().pgfh_get_node_detect(self.h, L.ptr(a, C.c_int32))
        return a

    def dofs(self):
        dof = np.zeros((self.dim, self.info.n_all_nodes), np.int32)
        pres = np.zeros(max(self.info.n_pres, 1))
        L.lib().pgfh_get_dofs(self.h, L.ptr(dof, C.c_int32), L.ptr(pres, C.c_double))
        return dof, pres[: self.info.n_pres]

    def enrich_edges(self):
        n = self.info.n_all_nodes - self.info.n_nodes
        a = np.zeros((max(n, 1), 3), np.int32)
        L.lib().pgfh_get_enrich_edges(self.h, L.ptr(a, C.c_int32))
        return a[:n]

    def cut(self):
        a = np.zeros((max(self.info.n_cut, 1), 128), np.int32)
        L.lib().pgfh_get_cut(self.h, L.ptr(a, C.c_int32))
        return a[: self.info.n_cut]

    def pattern(self):
        """(rowptr, rowsplit, colidx, color) built on the host exactly as pgf_create does."""
        lib = L.lib()
        n, nnz, nc = C.c_int64(), C.c_int64(), C.c_int32()
        L.check(lib.pgfh_pattern_info(self.h, C.byref(n), C.byref(nnz), C.byref(nc)), host=True)
        rp = np.zeros(n.value + 1, np.int64)
        rs = np.zeros(n.value, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        col = np.zeros(self.info.n_elem, np.int32)
        L.check(lib.pgfh_get_pattern(self.h, L.ptr(rp, C.c_int64), L.ptr(rs, C.c_int32), L.ptr(ci, C.c_int32),
                                     L.ptr(col, C.c_int32)), host=True)
        return rp, rs, ci, col

    def desc(self):
        d = L.MeshDesc()
        L.check(L.lib().pgfh_mesh_desc(self.h, C.byref(d)), host=True)
        return d

    def tables(self):
        """Copies of the cut elements' integration tables (store_shape_functions)
        as a dict of arrays."""
        d = self.desc()
        nc = d.n_cut

        def arr(ptr, n, ct):
            if n <= 0:
                return np.zeros(0, np.dtype(ct))
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()

        vptr = arr(d.vpt_ptr, nc + 1, C.c_int64)
        cptr = arr(d.cpt_ptr, nc + 1, C.c_int64)
        nv, ncp = int(vptr[-1]), int(cptr[-1])
        return {"cut_elem": arr(d.cut_elem, nc, C.c_int32), "cut_n_enrich": arr(d.cut_n_enrich, nc, C.c_int32),
                "cut_enrich": arr(d.cut_enrich, 16 * nc, C.c_int32), "vpt_ptr": vptr,
                "vpt_w": arr(d.vpt_w, nv, C.c_double), "vpt_mat": arr(d.vpt_mat, nv, C.c_int32),
                "vpt_grad_off": arr(d.vpt_grad_off, nv, C.c_int64),
                "vpt_grad": arr(d.vpt_grad, int(self.info.n_vgrad), C.c_double), "cpt_ptr": cptr,
                "cpt_w": arr(d.cpt_w, ncp, C.c_double), "cpt_N": arr(d.cpt_N, 4 * ncp, C.c_double),
                "cpt_R": arr(d.cpt_R, 9 * ncp, C.c_double), "cpt_nodes": arr(d.cpt_nodes, 8 * ncp, C.c_uint8),
                "cpt_nh": arr(d.cpt_nh, ncp, C.c_uint8), "cpt_law": arr(d.cpt_law, ncp, C.c_int32)}

    def write_vtk(self, path, U=None, eps_bar=None, t=0.0):
        """legacy ASCII VTK file: mesh (cut elements as their sub-elements),
        displacement (physical at enrichment nodes; plus eps_bar x if given),
        material per cell (pgfh_write_vtk)"""
        lib = L.lib()
        lib.pgfh_write_vtk.argtypes = [C.c_void_p, C.c_char_p, L.dp, L.dp, C.c_double]
        u = None if U is None else np.ascontiguousarray(U, np.float64)
        e = None if eps_bar is None else np.ascontiguousarray(eps_bar, np.float64)
        L.check(lib.pgfh_write_vtk(self.h, str(path).encode(), L.ptr(u, C.c_double), L.ptr(e, C.c_double), float(t)),
                host=True)


class DeviceContext:
    """One mesh resident on the GPU (pgf_create)."""

    def __init__(self, prob):
        self.prob = prob
        self.lib = L.lib()
        d = prob.desc()
        h = C.c_void_p()
        L.check(self.lib.pgf_create(C.byref(d), C.byref(h)), None)
        self.h = h
        self.csr = L.CsrInfo()
        self.lib.pgf_csr_info_get(self.h, C.byref(self.csr))
        dof, pres = prob.dofs()
        self.pres_full = np.ascontiguousarray(pres)
        L.check(self.lib.pgf_set_load(self.h, L.ptr(self.pres_full, C.c_double), None), self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pgf_destroy(self.h)
            self.h = None

    def _chk(self, rc):
        L.check(rc, self.h)

    def pattern(self):
        n, nnz = self.csr.n_rows, self.csr.nnz
        rp = np.zeros(n + 1, np.int64)
        rs = np.zeros(n, np.int32)
        ci = np.zeros(nnz, np.int32)
        self._chk(self.lib.pgf_get_csr(self.h, L.ptr(rp, C.c_int64), L.ptr(rs, C.c_int32), L.ptr(ci, C.c_int32)))
        return rp, rs, ci

    def colors(self):
        a = np.zeros(self.prob.info.n_elem, np.int32)
        self._chk(self.lib.pgf_get_colors(self.h, L.ptr(a, C.c_int32)))
        return a

    def set_solution(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._chk(self.lib.pgf_set_solution(self.h, L.ptr(u, C.c_double)))

    def solution(self):
        u = np.zeros((self.prob.dim, self.prob.info.n_all_nodes))
        self._chk(self.lib.pgf_get_solution(self.h, L.ptr(u, C.c_double)))
        return u

    def set_solution_dofs(self, x_free, u_pres=None):
        """nodal field from a dof vector (periodic images take their dof's value)"""
        x = np.ascontiguousarray(x_free, np.float64)
        up = None if u_pres is None else np.ascontiguousarray(u_pres, np.float64)
        self._chk(self.lib.pgf_set_solution_dofs(self.h, L.ptr(x, C.c_double), L.ptr(up, C.c_double)))

    def volume_average(self, eps_bar, seg_elem=None):
        """per element segment: (average stress [S], average fluctuation strain
        [S], volume) at the current nodal solution and macro strain eps_bar"""
        e = np.ascontiguousarray(eps_bar, np.float64)
        n_elem = self.prob.n_elem if hasattr(self.prob, "n_elem") else self.prob.info.n_elem
        seg = np.ascontiguousarray([0, n_elem] if seg_elem is None else seg_elem, np.int64)
        S = len(e)
        n = len(seg) - 1
        out = np.zeros((n, 2 * S + 1))
        self._chk(self.lib.pgf_volume_average(self.h, L.ptr(e, C.c_double), n, seg.ctypes.data_as(C.POINTER(C.c_int64)),
                                              L.ptr(out, C.c_double)))
        return out[:, :S], out[:, S:2 * S], out[:, 2 * S]

    def set_load(self, u_pres=None, fext=None):
        """full-load prescribed values [n_pres] (default: the problem's) and
        external free-dof force [n_free] (default: none)"""
        up = self.pres_full if u_pres is None else np.ascontiguousarray(u_pres, np.float64)
        f = None if fext is None else np.ascontiguousarray(fext, np.float64)
        self._chk(self.lib.pgf_set_load(self.h, L.ptr(up, C.c_double), L.ptr(f, C.c_double)))

    def apply_load_fraction(self, frac):
        self._chk(self.lib.pgf_apply_load_fraction(self.h, frac))

    def assemble(self, flags=L.PGF_ASM_JAC | L.PGF_ASM_FUNC, dt=1.0):
        self._chk(self.lib.pgf_assemble(self.h, dt, flags))

    def matrix(self):
        v = np.zeros(self.csr.nnz)
        self._chk(self.lib.pgf_get_matrix(self.h, L.ptr(v, C.c_double)))
        return v

    def matrix32(self):
        v = np.zeros(self.csr.nnz, np.float32)
        self._chk(self.lib.pgf_get_matrix32(self.h, L.ptr(v, C.c_float)))
        return v

    def internal_force(self):
        v = np.zeros(self.csr.n_rows)
        self._chk(self.lib.pgf_get_internal_force(self.h, L.ptr(v, C.c_double)))
        return v

    def isv(self):
        n = C.c_int64()
        self.lib.pgf_n_cohesive_points(self.h, C.byref(n))
        n = n.value
        dt, dc, fl = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1), np.uint8)
        self._chk(self.lib.pgf_get_isv(self.h, L.ptr(dt, C.c_double), L.ptr(dc, C.c_double), L.ptr(fl, C.c_uint8)))
        return dt[:n], dc[:n], fl[:n]

    def isv_commit(self):
        self._chk(self.lib.pgf_isv_commit(self.h))

    def isv_rollback(self):
        """trial delta_max back to the last committed state (pgf_isv_rollback)"""
        self._chk(self.lib.pgf_isv_rollback(self.h))

    def set_isv(self, dmax_conv):
        d = np.ascontiguousarray(dmax_conv, np.float64)
        self._chk(self.lib.pgf_set_isv(self.h, L.ptr(d, C.c_double)))

    def reset(self):
        """Back to the unloaded state: zero displacement and damage."""
        self.set_solution(np.zeros((self.prob.dim, self.prob.info.n_all_nodes)))
        n = C.c_int64()
        self.lib.pgf_n_cohesive_points(self.h, C.byref(n))
        if n.value:
            self.set_isv(np.zeros(n.value))

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.csr.n_free)
        self._chk(self.lib.pgf_spmv(self.h, L.ptr(x, C.c_double), L.ptr(y, C.c_double)))
        return y

    def pcg(self, b, rtol=1e-10, maxit=20000, precision=64, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.csr.n_free)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        info = L.SolveInfo()
        self._chk(self.lib.pgf_pcg(self.h, L.ptr(b, C.c_double), L.ptr(x, C.c_double), L.ptr(x0, C.c_double),
                                   rtol, maxit, precision, C.byref(info)))
        return x, info

    def newton_step(self, frac, rel_tol=1e-8, abs_tol=0.0, max_iter=40, precision=64, lin_rtol=1e-10,
                    lin_maxit=50000, dt=1.0, use_guess=False, adaptive_lin=False, fp32_maxit=0):
        o = L.NewtonOpts(rel_tol, abs_tol, max_iter, precision, lin_rtol, lin_maxit, int(use_guess), dt,
                         int(adaptive_lin), int(fp32_maxit))
        si = L.StepInfo()
        rc = self.lib.pgf_newton_step(self.h, frac, C.byref(o), C.byref(si))
        if rc not in (0, -5):
            self._chk(rc)
        return si

    def time_assemble(self, flags, reps=10):
        ms = C.c_double()
        self._chk(self.lib.pgf_time_assemble(self.h, flags, reps, C.byref(ms)))
        return ms.value

    def time_assemble_split(self, flags, reps=10):
        """(ms per call, ms cut-element pass, ms row-gather pass)"""
        out = np.zeros(3)
        self._chk(self.lib.pgf_time_assemble_split(self.h, flags, reps, L.ptr(out, C.c_double)))
        return tuple(out)

    @property
    def gather_layout(self):
        """(slots, node_rows) of the row-gather assembly pass: threads sharing
        one unit's incident elements, and whether a unit is a matrix row (0)
        or all rows of a node (1)."""
        sl, nr = C.c_int32(), C.c_int32()
        self._chk(self.lib.pgf_get_gather_layout(self.h, C.byref(sl), C.byref(nr)))
        return sl.value, nr.value

    @gather_layout.setter
    def gather_layout(self, layout):
        slots, node_rows = layout
        self._chk(self.lib.pgf_set_gather_layout(self.h, int(slots), int(node_rows)))

    def time_spmv(self, precision=64, reps=20):
        ms = C.c_double()
        self._chk(self.lib.pgf_time_spmv(self.h, precision, reps, C.byref(ms)))
        return ms.value

    # ---- batched solves (pgf_batch_pcg) ----
    def batch_pcg(self, sys_ptr, B, rtol=1e-10, maxit=100000, tune=(0, 0), persistent=False):
        """Solve K_ff X = B per system (free-row ranges sys_ptr) for the rows of
        B [nrhs][n_free]; returns X, iterations [n_sys][nrhs], relres, ms, tune.
        persistent: the whole solve in one cooperative launch
        (pgf_batch_pcg_persistent) instead of graph replays."""
        sp = np.ascontiguousarray(sys_ptr, np.int64)
        B = np.ascontiguousarray(B, np.float64)
        nrhs = B.shape[0]
        X = np.zeros_like(B)
        n_sys = len(sp) - 1
        its = np.zeros((n_sys, nrhs), np.int32)
        rel = np.zeros((n_sys, nrhs))
        ms = C.c_double()
        tn = (C.c_int32 * 2)(int(tune[0]), int(tune[1]))
        fn = self.lib.pgf_batch_pcg_persistent if persistent else self.lib.pgf_batch_pcg
        rc = fn(self.h, n_sys, sp.ctypes.data_as(C.POINTER(C.c_int64)), nrhs, L.ptr(B, C.c_double),
                                    L.ptr(X, C.c_double), rtol, maxit, its.ctypes.data_as(C.POINTER(C.c_int32)),
                                    L.ptr(rel, C.c_double), C.byref(ms), tn)
        if rc not in (0, -5):
            self._chk(rc)
        return X, its, rel, ms.value, (tn[0], tn[1])

    def rhs_prescribed(self, UP):
        UP = np.ascontiguousarray(UP, np.float64)
        B = np.zeros((UP.shape[0], self.csr.n_free))
        self._chk(self.lib.pgf_rhs_prescribed(self.h, UP.shape[0], L.ptr(UP, C.c_double), L.ptr(B, C.c_double)))
        return B

    def reactions(self, UF, UP):
        UF = np.ascontiguousarray(UF, np.float64)
        UP = np.ascontiguousarray(UP, np.float64)
        R = np.zeros_like(UP)
        self._chk(self.lib.pgf_reactions(self.h, UF.shape[0], L.ptr(UF, C.c_double), L.ptr(UP, C.c_double),
                                         L.ptr(R, C.c_double)))
        return R

    # ---- learned initial guess (include/pgfem_learn.h) ----
    def learn_attach(self, net=None, box=(0, 0, 0, 1, 1, 1), indicator=None):
        b = np.ascontiguousarray(box, np.float64)
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float32).ravel()
        self._learn_net = net  # keep the network alive while attached
        self._chk(self.lib.pgf_learn_attach(self.h, net.h if net is not None else None, L.ptr(b, C.c_double),
                                            L.ptr(ind, C.c_float)))

    def learn_commit(self):
        self._chk(self.lib.pgf_learn_commit(self.h))

    def learn_record(self):
        self._chk(self.lib.pgf_learn_record(self.h))

    def learn_last(self):
        out = np.zeros(6)
        self._chk(self.lib.pgf_learn_last(self.h, L.ptr(out, C.c_double)))
        return dict(alpha=out[:4].copy(), cols=int(out[4]), ms=out[5])

    def learn_features(self, grid=64):
        f = np.zeros((grid, grid, grid, 8), np.float32)
        self._chk(self.lib.pgf_learn_features(self.h, L.ptr(f, C.c_float)))
        return f

    def solve_load_steps(self, T_final=1.0, max_time_step=0.05, min_time_step=1e-6, max_repeats=20, max_steps=0,
                         callback=None, **newton):
        nw = dict(rel_tol=1e-8, abs_tol=0.0, max_iter=40, precision=64, lin_rtol=1e-10, lin_maxit=50000,
                  use_guess=0, dt=max_time_step)
        nw.update(newton)
        o = L.LoadOpts(T_final, max_time_step, min_time_step, max_repeats, max_steps,
                       L.NewtonOpts(nw["rel_tol"], nw["abs_tol"], nw["max_iter"], nw["precision"], nw["lin_rtol"],
                                    nw["lin_maxit"], int(nw["use_guess"]), nw["dt"]))
        records = []

        def _cb(user, step, t, dt, si):
            s = si.contents
            records.append(dict(step=step, t=t, dt=dt, newton=s.newton_iterations, pcg=s.pcg_iterations,
                                res0=s.res0, res=s.res, ms=s.ms_total, ms_asm=s.ms_assemble, ms_solve=s.ms_solve))
            if callback:
                callback(records[-1])

        cb = L.STEP_CB(_cb)
        info = L.LoadInfo()
        rc = self.lib.pgfh_solve_load_steps(self.prob.h, self.h, C.byref(o), cb, None, C.byref(info))
        if rc != 0:
            raise RuntimeError("load stepping failed (%d): %s" % (rc, self.lib.pgfh_last_error().decode()))
        return info, records


class GuessNet:
    """The learned-guess network (include/pgfem_learn.h) on the device."""

    GRID = 64
    IN_CH = 8
    OUT_CH = 3

    def __init__(self, seed=1):
        self.lib = L.lib()
        h = C.c_void_p()
        L.check(self.lib.pgf_net_create(C.c_uint64(seed), C.byref(h)), net=None)
        self.h = h
        n = C.c_int64()
        self._chk(self.lib.pgf_net_num_params(self.h, C.byref(n)))
        self.n_params = n.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pgf_net_destroy(self.h)
            self.h = None

    def _chk(self, rc):
        L.check(rc, net=self.h)

    def params(self):
        w = np.zeros(self.n_params, np.float32)
        self._chk(self.lib.pgf_net_get_params(self.h, L.ptr(w, C.c_float)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        self._chk(self.lib.pgf_net_set_params(self.h, L.ptr(w, C.c_float)))

    def _x(self, x):
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.GRID,) * 3 + (self.IN_CH,)
        return x

    def forward(self, x):
        x = self._x(x)
        y = np.zeros((self.GRID,) * 3 + (self.OUT_CH,), np.float32)
        self._chk(self.lib.pgf_net_forward(self.h, L.ptr(x, C.c_float), L.ptr(y, C.c_float)))
        return y

    def backward(self, x, dy):
        x = self._x(x)
        dy = np.ascontiguousarray(dy, np.float32)
        assert dy.shape == (self.GRID,) * 3 + (self.OUT_CH,)
        g = np.zeros(self.n_params, np.float32)
        self._chk(self.lib.pgf_net_backward(self.h, L.ptr(x, C.c_float), L.ptr(dy, C.c_float), L.ptr(g, C.c_float)))
        return g

    def add_sample(self, x, target):
        x = self._x(x)
        t = np.ascontiguousarray(target, np.float32)
        assert t.shape == (self.GRID,) * 3 + (self.OUT_CH,)
        self._chk(self.lib.pgf_net_add_sample(self.h, L.ptr(x, C.c_float), L.ptr(t, C.c_float)))

    def num_samples(self):
        n = C.c_int32()
        self._chk(self.lib.pgf_net_num_samples(self.h, C.byref(n)))
        return n.value

    def train(self, epochs, lr=1e-3):
        loss = (C.c_float * max(epochs, 1))()
        self._chk(self.lib.pgf_net_train(self.h, int(epochs), float(lr), loss))
        return np.array(loss[:epochs], np.float64)
