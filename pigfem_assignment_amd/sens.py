"""Direct shape sensitivities of the effective stiffness with respect to the
inclusion radii (C3's sensitivity right-hand sides).

The reference's direct method (SensitivityDirect::solveProblem,
src/SensitivityDirect.cpp:25-188) for one shape parameter d: assemble the
pseudo-load dR/dd at the converged solution (SensitivityShapeParameter /
SensitivityShapeParameterStructural, src/SensitivityShapeParameter.cpp:49-182,
src/SensitivityShapeParameterStructural.cpp:18-61), solve K du/dd = -dR/dd
with the same K, and differentiate the output.  Here the output is the
homogenised stress of each load case (src/SubscaleModel.cpp:42-196), so

    d sigma_bar / dr = ( int D B du/dr + int (D dB u + sigma div v) ) / V,

the first integral from the volume average of du/dr, the second (explicit)
one from pgf_shape_pseudo_load.  The velocity tables come from the host
(pgfh_shape_sensitivity); every RVE's parameter slot j is its j-th inclusion.
"""
import ctypes as C

import numpy as np

from . import _lib as L

_bound = False


def _lib():
    global _bound
    lib = L.lib()
    if not _bound:
        lib.pgfh_shape_sensitivity.argtypes = [C.c_void_p, C.c_int32, L.dp, L.dp, L.i32p, L.dp]
        lib.pgf_set_shape_sensitivity.argtypes = [C.c_void_p, L.dp, L.dp, L.i32p]
        lib.pgf_shape_pseudo_load.argtypes = [C.c_void_p, L.dp, C.c_int32, L.dp, C.c_int32, L.i64p, L.dp]
        _bound = True
    return lib


def host_tables(prob, as_reference=False):
    """(dgrad, divv, param, vel) of a HostProblem: derivatives of every
    volumetric point's slot gradients, velocity divergence per point, the
    inclusion of each cut element, enrichment-node velocities."""
    lib = _lib()
    i = prob.info
    dgrad = np.zeros(max(i.n_vgrad, 1))
    divv = np.zeros(max(i.n_vpt, 1))
    param = np.zeros(max(i.n_cut, 1), np.int32)
    vel = np.zeros((max(i.n_all_nodes - i.n_nodes, 1), 3))
    L.check(lib.pgfh_shape_sensitivity(prob.h, int(bool(as_reference)), L.ptr(dgrad, C.c_double),
                                       L.ptr(divv, C.c_double), L.ptr(param, C.c_int32), L.ptr(vel, C.c_double)),
            host=True)
    return dgrad[: i.n_vgrad], divv[: i.n_vpt], param[: i.n_cut], vel[: i.n_all_nodes - i.n_nodes]


def batch_tables(parts, as_reference=False):
    """Concatenated tables of the RVEs of a BatchProblem (in its order); the
    parameter of a cut element is its inclusion's index inside its own RVE."""
    tabs = [host_tables(p, as_reference) for p in parts]
    return (np.concatenate([t[0] for t in tabs]), np.concatenate([t[1] for t in tabs]),
            np.concatenate([t[2] for t in tabs]).astype(np.int32))


def attach(ctx, dgrad, divv, param):
    lib = _lib()
    dg = np.ascontiguousarray(dgrad, np.float64)
    dv = np.ascontiguousarray(divv, np.float64)
    pr = np.ascontiguousarray(param, np.int32)
    ctx._sens_keep = (dg, dv, pr)
    ctx._chk(lib.pgf_set_shape_sensitivity(ctx.h, L.ptr(dg, C.c_double), L.ptr(dv, C.c_double),
                                           L.ptr(pr, C.c_int32)))


def pseudo_load(ctx, eps_bar, slot, seg_elem=None):
    """(rhs [n_free] = -dR/dr_slot, explicit stress-derivative integrals
    [n_seg][S] with the segment volume's derivative as column S) at the
    context's current nodal solution"""
    lib = _lib()
    e = np.ascontiguousarray(eps_bar, np.float64)
    rhs = np.zeros(ctx.csr.n_free)
    seg = None if seg_elem is None else np.ascontiguousarray(seg_elem, np.int64)
    n_seg = 0 if seg is None else len(seg) - 1
    dsig = np.zeros((max(n_seg, 1), len(e) + 1))
    ctx._chk(lib.pgf_shape_pseudo_load(ctx.h, L.ptr(e, C.c_double), int(slot), L.ptr(rhs, C.c_double), n_seg,
                                       None if seg is None else seg.ctypes.data_as(C.POINTER(C.c_int64)),
                                       L.ptr(dsig, C.c_double)))
    return rhs, dsig[:n_seg]


def stiffness_sensitivities(ctx, X, seg_elem, sys_ptr, n_slots, cases, rtol=1e-12, log=None):
    """dC_eff/dr [n_seg][n_slots][3][3] by the direct method: for every
    parameter slot, the pseudo-loads of the load cases (fluctuations X[k]) are
    solved together in one segmented batched PCG, and the homogenised-stress
    derivative is the volume average of D B du/dr plus the explicit part,
    less sigma_bar times the relative volume derivative (the discrete volume,
    a sum of J*W, moves with r where a quadrature is not exact for the area)."""
    n_seg = len(seg_elem) - 1
    S = len(cases[0])
    dC = np.zeros((n_seg, n_slots, S, len(cases)))
    vol = None
    info = []
    sig_bar = []
    for k, e in enumerate(cases):
        ctx.set_solution_dofs(X[k])
        sig_bar.append(ctx.volume_average(np.asarray(e, np.float64), seg_elem)[0])
    for j in range(n_slots):
        rhs, expl = [], []
        for k, e in enumerate(cases):
            ctx.set_solution_dofs(X[k])
            r, ds = pseudo_load(ctx, e, j, seg_elem)
            rhs.append(r)
            expl.append(ds)
        dX, its, rel, ms, _ = ctx.batch_pcg(sys_ptr, np.array(rhs), rtol=rtol)
        info.append((its, rel, ms))
        for k in range(len(cases)):
            ctx.set_solution_dofs(dX[k])
            sig, _, v = ctx.volume_average(np.zeros(S), seg_elem)
            vol = v
            dC[:, j, :, k] = sig + (expl[k][:, :S] - sig_bar[k] * expl[k][:, S:]) / v[:, None]
        if log:
            log("# slot %d: batched sensitivity solves %.1f ms, iterations %d..%d" % (j, ms, its.min(), its.max()))
    return dC, info
