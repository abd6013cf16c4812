"""Parity bar of assembled stiffness entries, shared by the GPU parity tests.

|K_gpu - K_oracle| <= 1e-12 * |K|_abs + 1e-14 * rowmax for every entry, where
|K|_abs is the oracle's sum of the absolute values of every product entering
the entry (the scale of any rounding difference of the sums) and rowmax the
largest entry of the row (the scale of the rounding of the point gradients
themselves, which carry absolute errors of a few units in the last place of
the largest gradient: an entry whose exact products all vanish comes out as a
few ulps of the row).  Entries whose terms do not cancel (0 < |K|_abs <= 100
|K|) of rows that touch no cut element must also be within 1e-12 of their own
magnitude.
"""
import numpy as np


def k_entries_ok(K, o, rp, label, dt=1.0, vals=None):
    """The K bar above; o is the oracle at the same solution (assembled)."""
    K = np.asarray(K, np.float64)
    if vals is None:
        vals = o.assemble(dt=dt)[0]
    kabs = o.assemble_abs(dt=dt)
    d = np.abs(K - vals)
    rowmax = np.repeat(np.maximum.reduceat(np.abs(vals), rp[:-1]), np.diff(rp))
    tol = 1e-12 * kabs + 1e-14 * rowmax
    ok_abs = d <= tol
    p = o.prob
    cut_nodes = np.zeros(p.info.n_all_nodes, bool)
    conn = p.conn()
    for r in p.cut():
        cut_nodes[conn[r[0]]] = True
    cut_nodes[p.info.n_nodes:] = True
    dof, _ = o.dofs()
    row_cut = np.zeros(len(rp) - 1, bool)
    for k in range(dof.shape[0]):
        row_cut[dof[k][cut_nodes]] = True
    entry_cut = np.repeat(row_cut, np.diff(rp))
    wc = (kabs > 0) & (kabs <= 100 * np.abs(vals))
    plain = wc & ~entry_cut
    own = d / np.maximum(np.abs(vals), 1e-300)
    ratio = d / np.maximum(tol, 1e-300)
    w = int(np.argmax(ratio))
    row = int(np.searchsorted(rp, w, side="right") - 1)
    print("%s: %d entries; all within the bar: %s (max ratio %.3f at row %d%s: gpu %.17e oracle %.17e |K|_abs %.3e "
          "rowmax %.3e); non-cancelling entries of uncut rows %d, max own rel err %.2e; non-cancelling entries of "
          "cut rows %d, within 1e-12 own %d, max own rel %.2e; cancelling or structurally zero entries %d" % (
              label, len(vals), ok_abs.all(), float(ratio[w]), row, " (cut row)" if row_cut[row] else "", K[w],
              vals[w], kabs[w], rowmax[w], plain.sum(), float(own[plain].max()) if plain.any() else 0.0,
              (wc & entry_cut).sum(), (own[wc & entry_cut] <= 1e-12).sum(),
              float(own[wc & entry_cut].max()) if (wc & entry_cut).any() else 0, (~wc).sum()))
    return bool(ok_abs.all() and (own[plain] <= 1e-12).all())
