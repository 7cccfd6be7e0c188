"""The case tables: both encodings expand to the reference's bytes
(SHA-256 of NumFaces[256] and TriTable[256][16] from
dax/worklet/internal/MarchingCubesTable.h:28-306) and satisfy its invariants."""
import hashlib
import re
from pathlib import Path

import numpy as np

from oracle import oracle as orc

ROOT = Path(__file__).resolve().parent.parent
SHA_NUMFACES = "fc9278a1778172f129c18226368e8413cb551e0a7ab5646152e7dfa4532a246f"
SHA_TRITABLE = "19bf7699e214903d72c94c296546f2e31337d637a1e4b118c3108a0f428e809b"

# MarchingCubes.h:131-135
EDGE_VERTS = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5),
              (2, 6), (3, 7)]


def library_table():
    text = (ROOT / "daxtoolkit_amd" / "csrc" / "mc_case_table.h").read_text()
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-f]{16})ULL", text)]
    assert len(vals) == 256
    nf = np.array([v >> 60 for v in vals], np.uint8)
    tri = np.full((256, 16), 255, np.uint8)
    for c, v in enumerate(vals):
        for k in range(3 * nf[c]):
            tri[c, k] = (v >> (4 * k)) & 0xF
    return nf, tri


def test_oracle_table_sha():
    nf, tri = orc.tables()
    assert hashlib.sha256(nf.tobytes()).hexdigest() == SHA_NUMFACES
    assert hashlib.sha256(tri.tobytes()).hexdigest() == SHA_TRITABLE


def test_library_table_sha():
    nf, tri = library_table()
    assert hashlib.sha256(nf.tobytes()).hexdigest() == SHA_NUMFACES
    assert hashlib.sha256(tri.tobytes()).hexdigest() == SHA_TRITABLE


def test_table_invariants():
    nf, tri = orc.tables()
    assert int(nf.sum()) == 820
    hist = np.bincount(nf, minlength=6)
    assert list(hist) == [2, 16, 50, 80, 76, 32]
    for c in range(256):
        n = 3 * nf[c]
        assert (tri[c, :n] < 12).all() and (tri[c, n:] == 255).all()
        inside = [(c >> v) & 1 for v in range(8)]
        sign_change = {e for e, (a, b) in enumerate(EDGE_VERTS) if inside[a] != inside[b]}
        assert set(tri[c, :n].tolist()) == sign_change


# Cube faces and body diagonals in the corner order of TopologyUniform.h:139-159.
FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (0, 3, 7, 4), (1, 2, 6, 5)]
DIAGONALS = [(0, 6), (1, 7), (2, 4), (3, 5)]


def test_numfaces_closed_form():
    """The count kernel sums NumFaces over 64 cells with bit-plane popcounts
    of NumFaces = V - 2F - 4H (V inside corners, F faces with four inside
    corners, H = 1 iff the two outside corners end a body diagonal).  It
    must equal the table for every case with mixed corners, and the four
    H cases must be exactly those the plain V - 2F overcounts."""
    nf, _ = orc.tables()
    h_cases = []
    for c in range(1, 255):
        inside = [(c >> v) & 1 for v in range(8)]
        V = sum(inside)
        F = sum(all(inside[v] for v in f) for f in FACES)
        H = int(V == 6 and any(not inside[p] and not inside[q] for p, q in DIAGONALS))
        assert V - 2 * F - 4 * H == nf[c], c
        if H:
            h_cases.append(c)
    assert h_cases == [125, 190, 215, 235]


def test_numfaces_word_sum_bitplanes():
    """The same sum formed the way the kernel forms it: eight 64-bit corner
    planes, AND / AND-NOT of planes and popcounts (word_triangles in
    mc_common.h), against the per-cell table lookup on random words."""
    nf, _ = orc.tables()
    rng = np.random.default_rng(5)
    mask = (1 << 64) - 1
    pc = lambda x: bin(x & mask).count("1")
    for _ in range(300):
        c = [int(rng.integers(0, 1 << 63)) << 1 | int(rng.integers(0, 2)) for _ in range(8)]
        if rng.random() < 0.5:  # mostly-inside words reach the 6- and 7-corner cases
            c = [x | int(rng.integers(0, 1 << 63)) for x in c]
        all_, any_ = mask, 0
        for x in c:
            all_ &= x
            any_ |= x
        a = any_ & ~all_ & mask
        v = sum(pc(x & a) for x in c)
        p01, p23, p45, p67 = c[0] & c[1], c[2] & c[3], c[4] & c[5], c[6] & c[7]
        f = (pc(p01 & p23 & a) + pc(p45 & p67 & a) + pc(p01 & p45 & a) + pc(p23 & p67 & a) +
             pc(c[0] & c[3] & c[4] & c[7] & a) + pc(c[1] & c[2] & c[5] & c[6] & a))
        h = ((p23 & p45 & ((c[1] & ~c[0] & c[7] & ~c[6]) | (c[0] & ~c[1] & c[6] & ~c[7]))) |
             (p01 & p67 & ((c[3] & ~c[2] & c[5] & ~c[4]) | (c[2] & ~c[3] & c[4] & ~c[5]))))
        got = v - 2 * f - 4 * pc(h & a)
        want = sum(int(nf[sum(((c[k] >> i) & 1) << k for k in range(8))]) for i in range(64)
                   if (a >> i) & 1)
        assert got == want
