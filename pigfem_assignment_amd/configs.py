"""Synthetic inclusion workloads (seeded), mirroring the reference's driver set-up.

Materials are SI: matrix E=3.2 GPa nu=0.35, inclusion E=70 GPa nu=0.22.
Inclusion centres are placed by random sequential addition (RSA): uniform
centres, radii uniform in [r_lo, r_hi], rejected if closer than `gap` to a
placed inclusion or to the domain boundary, so that no element is cut by two
interfaces (the reference rejects such elements for tets, src/tet4.cpp:259-263).
"""
import math

import numpy as np

from . import _lib as L
from .api import HostProblem

MATRIX = (3.2e9, 0.35)
INCLUSION = (70e9, 0.22)
SEED = 20261202


def rsa(n, r_lo, r_hi, dim, gap, seed=SEED, max_tries=2_000_000, lo=0.0, hi=1.0):
    rng = np.random.RandomState(seed)
    cs, rs = [], []
    tries = 0
    while len(cs) < n and tries < max_tries:
        tries += 1
        r = rng.uniform(r_lo, r_hi)
        c = rng.uniform(lo + r + gap, hi - r - gap, size=dim)
        ok = True
        for c2, r2 in zip(cs, rs):
            if np.sum((c - c2) ** 2) < (r + r2 + gap) ** 2:
                ok = False
                break
        if ok:
            cs.append(c)
            rs.append(r)
    if len(cs) < n:
        raise RuntimeError("RSA placed only %d of %d inclusions" % (len(cs), n))
    incl = np.zeros((n, 5))
    incl[:, 0] = 0 if dim == 2 else 1
    incl[:, 1:1 + dim] = np.array(cs)
    incl[:, 4] = np.array(rs)
    return incl


def volume_fraction(incl, dim):
    if dim == 2:
        return float(np.sum(math.pi * incl[:, 4] ** 2))
    return float(np.sum(4.0 / 3.0 * math.pi * incl[:, 4] ** 3))


def c0_problem(n=100, **kw):
    """Unit square, n x n QUAD4, one centred circle r=0.2, perfectly bonded,
    left edge fixed, right edge u_x = 0.01 (linear FP64 solve)."""
    incl = np.array([[0, 0.5, 0.5, 0.0, 0.2]])
    bcs = [(L.SET_LEFT, 0, 0, 0.0), (L.SET_LEFT, 0, 1, 0.0), (L.SET_RIGHT, 0, 0, 0.01)]
    return HostProblem("quad4", (n, n), (0, 0), (1, 1), inclusions=incl, materials=(MATRIX, INCLUSION), bcs=bcs, **kw)


def c1_inclusions(n_incl=400, n=1000, seed=SEED):
    return rsa(n_incl, 0.005, 0.02, 2, gap=2.0 / n, seed=seed)


def c1_problem(n=1000, n_incl=400, seed=SEED, **kw):
    """Unit square, n x n QUAD4, RSA circles r~U[0.005,0.02], linear elastic."""
    incl = c1_inclusions(n_incl, n, seed)
    bcs = [(L.SET_LEFT, 0, 0, 0.0), (L.SET_LEFT, 0, 1, 0.0), (L.SET_RIGHT, 0, 0, 0.01)]
    return HostProblem("quad4", (n, n), (0, 0), (1, 1), inclusions=incl, materials=(MATRIX, INCLUSION), bcs=bcs, **kw)


BILINEAR_C2 = (L.PGF_LAW_BILINEAR, 30e6, 1e-3, 1e-2, 1.0, 0.0, 0.0)
OP_EXP = (L.PGF_LAW_OP_EXPONENTIAL, 30e6, 1e-3, 0.0, 1.0, 0.0, 0.0)


def c2_inclusions(n_incl=200, n=128, seed=SEED, r_lo=0.03, r_hi=0.06):
    return rsa(n_incl, r_lo, r_hi, 3, gap=2.0 / n, seed=seed)


def c2_problem(n=128, n_incl=200, seed=SEED, law=BILINEAR_C2, r_lo=0.03, r_hi=0.06, strain=0.02, **kw):
    """Unit cube, n^3 cubes x 6 TET4, RSA spheres with cohesive interfaces,
    left face fixed, right face u_x = strain (applied in load increments)."""
    incl = c2_inclusions(n_incl, n, seed, r_lo, r_hi)
    bcs = [(L.SET_LEFT, 0, d, 0.0) for d in range(3)] + [(L.SET_RIGHT, 0, 0, strain)]
    return HostProblem("tet4", (n, n, n), (0, 0, 0), (1, 1, 1), inclusions=incl, materials=(MATRIX, INCLUSION),
                       cohesive=True, coh_laws=[law], bcs=bcs, **kw)
