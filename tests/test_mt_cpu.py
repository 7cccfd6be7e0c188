"""The oracle's marching tetrahedra (dax/worklet/MarchingTetrahedra.h:29-143
through the Serial adapter): its tables against the reference's bytes
(SHA-256) and per-case invariants, the reference's unit test
(dax/worklet/testing/UnitTestWorkletMarchingTetrahedra.cxx:41-135) and a
brute-force count of the mesh's sign-change edges."""
import hashlib

import numpy as np

from oracle import oracle as orc

EDGES = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 1), (3, 2)]


def test_tet_tables_match_reference_bytes():
    nf, tri = orc.tet_tables()
    assert hashlib.sha256(nf.tobytes()).hexdigest() == (
        "74e246ffba90913c1fb68de7910c936cafd863b7e0609a95901b6c16a669b9da")
    assert hashlib.sha256(tri.tobytes()).hexdigest() == (
        "aa1d16e253be25134238c3abaa96b7d32f3ac07d4dafa6238de6097705c32b2f")


def test_tet_tables_cut_exactly_the_sign_change_edges():
    nf, tri = orc.tet_tables()
    for c in range(16):
        inside = [(c >> v) & 1 for v in range(4)]
        change = {e for e, (a, b) in enumerate(EDGES) if inside[a] != inside[b]}
        used = set(int(e) for e in tri[c][:3 * nf[c]])
        assert used == change
        assert all(e == 255 for e in tri[c][3 * nf[c]:])
        # one triangle for 3 cut edges, two (a quad) for 4
        assert nf[c] == {0: 0, 3: 1, 4: 2}[len(change)]


def _field_xyz(pts):
    return (pts[:, 0] + pts[:, 1] + pts[:, 2]).astype(np.float32)


def test_reference_unit_test_26():
    """26^3 Freudenthal tetrahedra, f = x+y+z, iso 70: 3T points unmerged,
    151 merged."""
    tets, pts = orc.freudenthal_tets((26, 26, 26))
    f = _field_xyz(pts)
    r0 = orc.mt_generate(tets, f, pts, 70.0, dedup=False)
    r1 = orc.mt_generate(tets, f, pts, 70.0, dedup=True)
    assert r0["n_pts"] == 3 * r0["n_tris"] and r0["n_tris"] == r1["n_tris"]
    assert r1["n_pts"] == 151


def test_dedup_points_are_the_sign_change_edges():
    rng = np.random.default_rng(3)
    tets, pts = orc.freudenthal_tets((9, 7, 6))
    f = rng.standard_normal(pts.shape[0]).astype(np.float32)
    r = orc.mt_generate(tets, f, pts, 0.2, dedup=True)
    e = np.concatenate([tets[:, [a, b]] for a, b in EDGES])
    e = np.sort(e, axis=1)
    e = np.unique(e, axis=0)
    cut = e[(f[e[:, 0]] > 0.2) != (f[e[:, 1]] > 0.2)]
    assert r["n_pts"] == cut.shape[0]
    assert np.array_equal(np.stack([r["first"], r["second"]], axis=1), cut)
