"""Training data, training and evaluation of the learned initial guess.

A load-step sequence of a cohesive inclusion problem is solved with plain
Jacobi-PCG inside the device Newton step; after every accepted step the
step-start features (inclusion indicator, damage, step residual, previous
increment on the 64^3 voxel grid) and the normalised converged increment are
added to the network's device-resident training set.  The network is then
trained on the device (bf16 MFMA forward/backward, Adam) and used as the
initial guess of the first PCG solve of every Newton step (use_guess=2),
Galerkin-combined with the previous increment so the start is never worse in
the energy norm than a zero start.  Plain PCG stays the solver: every step
still has to reach the same relative residual.
"""
import time

import numpy as np

from . import configs
from .api import DeviceContext, GuessNet

GRID = 64


def indicator(inclusions, grid=GRID, box=(0, 0, 0, 1, 1, 1)):
    """1 at voxel centres inside any inclusion (spheres or circles), else 0."""
    lo = np.asarray(box[:3], np.float64)
    ext = np.asarray(box[3:], np.float64) - lo
    c = (np.arange(grid) + 0.5) / grid
    z, y, x = np.meshgrid(lo[2] + c * ext[2], lo[1] + c * ext[1], lo[0] + c * ext[0], indexing="ij")
    ind = np.zeros((grid, grid, grid), np.float32)
    for inc in np.asarray(inclusions, np.float64).reshape(-1, 5):
        d2 = (x - inc[1]) ** 2 + (y - inc[2]) ** 2 + (z - inc[3]) ** 2
        ind[d2 <= inc[4] ** 2] = 1.0
    return ind


def cohesive_problem(n, n_incl, seed, law=configs.BILINEAR_C2, r_lo=0.06, r_hi=0.10, strain=0.02):
    """Unit cube, n^3 x 6 TET4, n_incl RSA spheres with cohesive interfaces."""
    prob = configs.c2_problem(n=n, n_incl=n_incl, seed=seed, law=law, r_lo=r_lo, r_hi=r_hi, strain=strain)
    incl = configs.c2_inclusions(n_incl, n, seed, r_lo, r_hi)
    return prob, incl


def run_steps(prob, incl, n_steps, frac_step, use_guess=0, net=None, record=False, lin_rtol=1e-10, dt=None,
              precision=64, callback=None, ctx=None, adaptive_lin=False, fp32_maxit=0, keep_solutions=None):
    """Solve load steps frac = k * frac_step, k = 1..n_steps from the unloaded
    state; returns per-step records, the damage flags after every step and the
    context (pass it back in to reuse its device set-up).  keep_solutions, a
    list, receives the nodal displacement after every step."""
    if ctx is None:
        ctx = DeviceContext(prob)
    else:
        ctx.reset()
    ctx.learn_attach(net, indicator=indicator(incl))
    out, flags = [], []
    for k in range(1, n_steps + 1):
        t0 = time.time()
        si = ctx.newton_step(k * frac_step, dt=dt if dt is not None else frac_step, use_guess=use_guess,
                             lin_rtol=lin_rtol, precision=precision, adaptive_lin=adaptive_lin,
                             fp32_maxit=fp32_maxit)
        wall = time.time() - t0
        if not si.converged:
            raise RuntimeError("load step %d did not converge (mode %d)" % (k, use_guess))
        ctx.isv_commit()
        if record:
            ctx.learn_record()
        ctx.learn_commit()
        g = ctx.learn_last() if use_guess in (2, 3) else None
        rec = dict(step=k, newton=si.newton_iterations, pcg=si.pcg_iterations, ms=si.ms_total,
                   ms_solve=si.ms_solve, ms_asm=si.ms_assemble, res0=si.res0, res=si.res, wall_s=wall,
                   guess=g, fp64_fallback=si.fp64_fallback)
        out.append(rec)
        flags.append(ctx.isv()[2].copy())
        if keep_solutions is not None:
            keep_solutions.append(ctx.solution())
        if callback:
            callback(rec)
    return out, flags, ctx


def train_net(net, problems, n_steps, frac_step, epochs, lr=1e-3, lin_rtol=1e-10, log=None):
    """Generate samples from plain solves of `problems` [(prob, incl)], then train."""
    t0 = time.time()
    for prob, incl in problems:
        run_steps(prob, incl, n_steps, frac_step, use_guess=0, net=net, record=True, lin_rtol=lin_rtol)
        if log:
            log("# samples after problem: %d (%.1f s)" % (net.num_samples(), time.time() - t0))
    t1 = time.time()
    loss = net.train(epochs, lr=lr)
    if log:
        log("# trained %d epochs on %d samples in %.1f s: loss %.4f -> %.4f" % (
            epochs, net.num_samples(), time.time() - t1, loss[0], loss[-1]))
    return loss
