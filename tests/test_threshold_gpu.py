"""GPU Threshold (ThresholdCount / ThresholdTopology through DispatcherMapCell
and DispatcherGenerateTopology, dax/worklet/Threshold.h:30-110,
dax/cont/DispatcherGenerateTopology.h:146-256): bit-exact parity with the
oracle (counts, connectivity, used-point coordinates, point mask,
CompactPointField) on the reference's unit test, random uniform grids with
origin and spacing, shuffled tetrahedral grids and hexahedral grids, both
duplicate-removal modes; cells repeated by counts above 1; and at 256^3 on
the reference benchmark's field (point magnitude, [0, 100]) the output
against a torch restatement of the definition."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import _lib, topology
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _check(out, ref, dedup):
    assert out.num_cells == ref["n_cells"]
    assert np.array_equal(out.counts.cpu().numpy(), ref["counts"])
    assert np.array_equal(out.connectivity.cpu().numpy(), ref["conn"])
    if dedup:
        assert out.num_points == ref["n_points"]
        assert np.array_equal(out.points.cpu().numpy().view(np.uint32),
                              ref["coords"].view(np.uint32))
        assert np.array_equal(out.point_mask.cpu().numpy(), ref["mask"])


def test_reference_unit_test_26():
    n = 26
    f = np.add.outer(np.add.outer(np.arange(n), np.arange(n)), np.arange(n)).astype(np.float32)
    print("[case] threshold unit26", flush=True)
    ft = torch.from_numpy(f).cuda()
    out = topology.threshold(ft, 70, 82)
    ref = orc.threshold(f, 70, 82)
    _check(out, ref, True)
    vals = out.compact_point_field(ft.reshape(-1)).cpu().numpy()
    assert np.array_equal(vals, ref["field"])
    assert out.num_cells == 10 and np.all((vals >= 70) & (vals <= 82))


@pytest.mark.parametrize("dedup", [True, False])
@pytest.mark.parametrize("dims", [(9, 7, 6), (70, 33, 20), (130, 5, 40), (2, 2, 2), (68, 9, 7),
                                  (64, 3, 2), (260, 6, 19)])
def test_uniform_matches_oracle(dims, dedup):
    rng = np.random.default_rng(sum(dims))
    nx, ny, nz = dims
    f = rng.random((nz, ny, nx)).astype(np.float32)
    origin, spacing = (0.5, -1.0, 2.0), (0.25, 1.5, 1.0)
    print(f"[case] threshold uniform {dims} dedup={dedup}", flush=True)
    out = topology.threshold(torch.from_numpy(f).cuda(), 0.03, 0.97, origin=origin,
                             spacing=spacing, dedup=dedup)
    ref = orc.threshold(f, 0.03, 0.97, dedup=dedup, origin=origin, spacing=spacing)
    _check(out, ref, dedup)


@pytest.mark.parametrize("dedup", [True, False])
def test_shuffled_tetrahedra_and_hexahedra_match_oracle(dedup):
    rng = np.random.default_rng(4)
    for name, builder, dims in (("tets", orc.freudenthal_tets, (12, 9, 8)),
                                ("hexes", orc.voxel_hexes, (20, 13, 9))):
        cells, pts = builder(dims)
        perm = rng.permutation(pts.shape[0]).astype(np.int32)
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size, dtype=np.int32)
        cells, pts = inv[cells], pts[perm]
        f = rng.random(pts.shape[0]).astype(np.float32)
        print(f"[case] threshold {name} {dims} dedup={dedup}", flush=True)
        ft = torch.from_numpy(f).cuda()
        out = topology.threshold(ft, 0.1, 0.95, cells=torch.from_numpy(cells).cuda(),
                                 points=torch.from_numpy(pts).cuda(), dedup=dedup)
        ref = orc.threshold(f, 0.1, 0.95, dedup=dedup, cells=cells, points=pts)
        _check(out, ref, dedup)
        if dedup:
            assert np.array_equal(out.compact_point_field(ft).cpu().numpy(), ref["field"])


def test_counts_above_one_repeat_cells():
    """DispatcherGenerateTopology with counts 0..3: cell c appears counts[c]
    times, in cell order."""
    dims = (7, 6, 5)
    C = 6 * 5 * 4
    rng = np.random.default_rng(2)
    counts = rng.integers(0, 4, C).astype(np.int32)
    hexes, pts = orc.voxel_hexes(dims)
    print("[case] topology counts 0..3", flush=True)
    out = topology.generate_topology(torch.from_numpy(counts).cuda(), dims=dims, dedup=False)
    assert np.array_equal(out.connectivity.cpu().numpy(), np.repeat(hexes, counts, axis=0))
    out2 = topology.generate_topology(torch.from_numpy(counts).cuda(), dims=dims, dedup=True)
    used = np.unique(np.repeat(hexes, counts, axis=0))
    assert np.array_equal(out2.connectivity.cpu().numpy(),
                          np.searchsorted(used, np.repeat(hexes, counts, axis=0)))
    assert np.array_equal(out2.points.cpu().numpy(), pts[used])


def test_magnitude_256_against_torch_definition():
    """The reference benchmark's pipeline (Benchmarks/Threshold/Pipeline.h:
    163-190) at 256^3: field = |p| (origin 0, spacing 1), [0, 100]."""
    n = 256
    print("[case] threshold magnitude 256^3", flush=True)
    ax = torch.arange(n, device="cuda", dtype=torch.float32)
    z, y, x = torch.meshgrid(ax, ax, ax, indexing="ij")
    f = torch.sqrt(x * x + y * y + z * z).contiguous()
    out = topology.threshold(f, 0.0, 100.0)
    inr = (f >= 0) & (f <= 100)
    keep = torch.ones((n - 1, n - 1, n - 1), dtype=torch.bool, device="cuda")
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                keep &= inr[dz:dz + n - 1, dy:dy + n - 1, dx:dx + n - 1]
    assert torch.equal(out.counts.view(n - 1, n - 1, n - 1), keep.to(torch.int32))
    T = int(keep.sum())
    assert out.num_cells == T > 400_000
    used = torch.zeros((n, n, n), dtype=torch.bool, device="cuda")
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                used[dz:dz + n - 1, dy:dy + n - 1, dx:dx + n - 1] |= keep
    assert torch.equal(out.point_mask.view(n, n, n), used.to(torch.int32))
    assert out.num_points == int(used.sum())
    # every output cell's points are its voxel's corners, in order
    idx = torch.nonzero(keep.reshape(-1)).reshape(-1)
    i, j, k = idx % (n - 1), (idx // (n - 1)) % (n - 1), idx // ((n - 1) * (n - 1))
    corner = torch.tensor([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1],
                           [1, 1, 1], [0, 1, 1]], device="cuda")
    expect = torch.stack([i, j, k], 1)[:, None, :] + corner[None]
    got = out.points[out.connectivity.long()]
    assert torch.equal(got, expect.to(torch.float32))
    vals = out.compact_point_field(f.reshape(-1))
    assert torch.equal(vals, f.reshape(-1)[used.reshape(-1)])


def test_counts_whose_sum_wraps_int32_are_rejected():
    """DispatcherGenerateTopology scans the counts in int32: three counts of
    2^31-1 wrap to a positive total, and a negative count shifts every later
    offset.  Both are rejected from the exact 64-bit sum before any cell is
    emitted, on a uniform and on an unstructured grid."""
    dims = (4, 2, 2)  # three voxels
    hexes, pts = orc.voxel_hexes(dims)
    cells = torch.from_numpy(np.ascontiguousarray(hexes, dtype=np.int32)).cuda()
    points = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float32)).cuda()
    for counts in ([2**31 - 1] * 3, [1, -1, 1], [2**28, 0, 0]):
        c = torch.tensor(counts, dtype=torch.int32, device="cuda")
        for dedup in (False, True):
            print(f"[case] topology counts {counts} dedup={dedup}", flush=True)
            with pytest.raises(_lib.ErrorControlBadValue):
                topology.generate_topology(c, dims=dims, dedup=dedup)
            with pytest.raises(_lib.ErrorControlBadValue):
                topology.generate_topology(c, cells=cells, points=points, dedup=dedup)


def test_bad_ids_are_rejected():
    pts = torch.zeros((4, 3), device="cuda")
    f = torch.zeros(4, device="cuda")
    tets = torch.tensor([[0, 1, 2, 7]], dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.ErrorControlBadValue):
        topology.threshold(f, -1.0, 1.0, cells=tets, points=pts)


def test_dedup_generate_after_plain_prepare_is_rejected():
    """generate / CompactPointField with duplicate removal need a prepare that
    ranked the points: otherwise an error, not a read of unset ranks."""
    import ctypes
    dims = (40, 30, 20)
    rng = np.random.default_rng(6)
    f = torch.from_numpy(rng.random(dims[::-1]).astype(np.float32)).cuda()
    counts = topology.threshold_count(f, 0.1, 0.9)
    C, P = counts.numel(), f.numel()
    lib = _lib.load()
    ws = torch.empty(int(lib.dax_hip_topology_workspace_bytes(C, P)), dtype=torch.uint8,
                     device="cuda")
    mask = torch.empty(P, dtype=torch.int32, device="cuda")
    t, u = ctypes.c_int64(), ctypes.c_int64()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("[case] topology flags mismatch", flush=True)
    _lib.call("dax_hip_topology_prepare_uniform", *dims, counts.data_ptr(), 0, ws.data_ptr(),
              ws.numel(), None, ctypes.byref(t), ctypes.byref(u), s)
    conn = torch.empty((max(t.value, 1), 8), dtype=torch.int32, device="cuda")
    coords = torch.empty((P, 3), dtype=torch.float32, device="cuda")
    o3 = (ctypes.c_float * 3)(0, 0, 0)
    h3 = (ctypes.c_float * 3)(1, 1, 1)
    with pytest.raises(_lib.ErrorControlBadValue):
        _lib.call("dax_hip_topology_generate_uniform", *dims, o3, h3, counts.data_ptr(), 1,
                  ws.data_ptr(), ws.numel(), mask.data_ptr(), conn.data_ptr(), coords.data_ptr(),
                  s)
    out = torch.empty(P, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.ErrorControlBadValue):
        _lib.call("dax_hip_topology_compact_point_field", f.data_ptr(), 4, C, P, mask.data_ptr(),
                  ws.data_ptr(), ws.numel(), out.data_ptr(), s)


@pytest.mark.parametrize("dims", [(68, 20, 9), (256, 7, 5), (70, 33, 20)])
def test_fused_prepare_without_mask_matches_oracle(dims):
    """ThresholdCount and prepare in one call (the count kernel forming the
    cell words when nx % 4 == 0), the point mask left out: the same counts,
    connectivity, coordinates and CompactPointField as the oracle."""
    rng = np.random.default_rng(sum(dims) + 5)
    f = rng.random(dims[::-1]).astype(np.float32)
    print(f"[case] threshold fused {dims}", flush=True)
    ft = torch.from_numpy(f).cuda()
    out = topology.threshold(ft, 0.05, 0.97, point_mask=False)
    ref = orc.threshold(f, 0.05, 0.97)
    assert out.point_mask is None
    assert out.num_cells == ref["n_cells"] and out.num_points == ref["n_points"]
    assert np.array_equal(out.counts.cpu().numpy(), ref["counts"])
    assert np.array_equal(out.connectivity.cpu().numpy(), ref["conn"])
    assert np.array_equal(out.points.cpu().numpy(), ref["coords"])
    assert np.array_equal(out.compact_point_field(ft.reshape(-1)).cpu().numpy(), ref["field"])


@pytest.mark.parametrize("dims", [(64, 5, 2), (132, 6, 17), (68, 4, 18), (260, 3, 34)])
def test_count_only_matches_oracle(dims):
    """ThresholdCount alone (DispatcherMapCell, the float4-row count kernel
    when nx % 4 == 0) over z-runs of 1, 16 + 0, 16 + 1 and 2 x 16 + 1 cell
    planes, then DispatcherGenerateTopology on its counts: the same counts,
    connectivity and coordinates as the oracle."""
    rng = np.random.default_rng(sum(dims) + 11)
    f = rng.random(dims[::-1]).astype(np.float32)
    print(f"[case] threshold count only {dims}", flush=True)
    ft = torch.from_numpy(f).cuda()
    counts = topology.threshold_count(ft, 0.04, 0.96)
    ref = orc.threshold(f, 0.04, 0.96)
    C = (dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1)
    assert np.array_equal(counts[:C].cpu().numpy(), ref["counts"])
    out = topology.generate_topology(counts, dims=dims)
    assert out.num_cells == ref["n_cells"] and out.num_points == ref["n_points"]
    assert np.array_equal(out.connectivity.cpu().numpy(), ref["conn"])
    assert np.array_equal(out.points.cpu().numpy(), ref["coords"])
