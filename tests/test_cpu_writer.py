"""VTK output (pgfh_write_vtk) read back: point and cell counts, cell types,
the cells tile the domain (their areas sum to the square's), and the
displacement written at an enrichment node is the physical one (for an
affine field on the regular nodes with zero enrichment amplitudes it is the
affine field at the enrichment point; an enrichment amplitude adds to it)."""
import numpy as np

from pigfem_assignment_amd import configs


def _read(path):
    lines = open(path).read().split("\n")
    out = {}
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if tok and tok[0] == "POINTS":
            n = int(tok[1])
            out["points"] = np.array([list(map(float, lines[i + 1 + k].split())) for k in range(n)])
            i += n
        elif tok and tok[0] == "CELLS":
            n = int(tok[1])
            out["cells"] = [list(map(int, lines[i + 1 + k].split()))[1:] for k in range(n)]
            i += n
        elif tok and tok[0] == "CELL_TYPES":
            n = int(tok[1])
            out["types"] = np.array([int(lines[i + 1 + k]) for k in range(n)])
            i += n
        elif tok and tok[0] == "VECTORS":
            n = len(out["points"])
            out["disp"] = np.array([list(map(float, lines[i + 1 + k].split())) for k in range(n)])
            i += n
        elif tok and tok[:2] == ["SCALARS", "material"]:
            n = len(out["cells"])
            out["mat"] = np.array([int(lines[i + 2 + k]) for k in range(n)])
            i += n + 1
        i += 1
    return out


def _area(P):
    x, y = P[:, 0], P[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_vtk_round_trip(tmp_path):
    p = configs.c0_problem(n=20)
    nn, na = p.info.n_nodes, p.info.n_all_nodes
    X = p.coords()
    U = np.zeros((2, na))
    U[0, :nn] = 1e-3 * X[0, :nn] + 2e-4 * X[1, :nn]
    U[1, :nn] = -5e-4 * X[1, :nn]
    path = tmp_path / "c0.vtk"
    p.write_vtk(path, U)
    v = _read(path)
    assert v["points"].shape == (na, 3)
    n_sub = sum(r[18] for r in p.cut())
    assert len(v["cells"]) == p.info.n_elem - p.info.n_cut + n_sub
    assert set(np.unique(v["types"])) <= {5, 7, 9}
    assert abs(sum(_area(v["points"][c, :2]) for c in v["cells"]) - 1.0) < 1e-6
    # enrichment points carry the affine field (float output)
    P = v["points"][nn:, :2]
    want = np.stack([1e-3 * P[:, 0] + 2e-4 * P[:, 1], -5e-4 * P[:, 1]], 1)
    assert np.allclose(v["disp"][nn:, :2], want, rtol=1e-6, atol=1e-9)
    assert np.allclose(v["disp"][:nn, :2], U[:, :nn].T, rtol=1e-6, atol=1e-9)
    # an enrichment amplitude adds to the interpolated field
    U[0, nn] = 1e-3
    p.write_vtk(path, U)
    v2 = _read(path)
    assert abs(v2["disp"][nn, 0] - (v["disp"][nn, 0] + 1e-3)) < 1e-8
    # materials: matrix cells and inclusion cells both present
    assert set(np.unique(v["mat"])) == {0, 1}
