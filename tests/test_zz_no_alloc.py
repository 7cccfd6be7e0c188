"""dax_hip_mc_forward_noalloc (the forward with its workspace in the
library's per-device array) against dax_hip_mc_forward with a caller
workspace, through the C-ABI: bit for bit, on two calls in a row (the second
on a smaller grid, over the first call's leftover scratch), and a grid whose
workspace exceeds the array is rejected."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import _lib, fields, mc

pytestmark = pytest.mark.gpu


def _outputs(n_pts, n_tris):
    dev = "cuda"
    return {"v": torch.full((n_pts, 3), np.nan, dtype=torch.float32, device=dev),
            "n": torch.full((n_pts, 3), np.nan, dtype=torch.float32, device=dev),
            "c": torch.full((n_tris, 3), -7, dtype=torch.int32, device=dev),
            "e": torch.full((n_pts,), -7, dtype=torch.int32, device=dev),
            "f": torch.full((n_pts,), -7, dtype=torch.int32, device=dev),
            "s": torch.full((n_pts,), -7, dtype=torch.int32, device=dev),
            "w": torch.full((n_pts,), np.nan, dtype=torch.float32, device=dev),
            "k": torch.full((2,), -1, dtype=torch.int64, device=dev)}


def _args(o):
    return (o["v"].data_ptr(), o["n"].data_ptr(), o["c"].data_ptr(), o["e"].data_ptr(),
            o["f"].data_ptr(), o["s"].data_ptr(), o["w"].data_ptr(), o["v"].shape[0],
            o["c"].shape[0], o["k"].data_ptr())


def _run_both(dims, seed, iso, origin, spacing):
    nx, ny, nz = dims
    field = fields.gaussian_mixture_torch(
        dims, fields.gaussian_mixture_params(seed, dims, n_gauss=6, sigma=(8.0, 25.0)))
    assert field.shape == (nz, ny, nx)
    out = mc.forward(field, iso, origin=origin, spacing=spacing, normals=True, edges=True)
    U, T = out.num_points, out.num_triangles
    assert U > 1000
    lib = _lib.load()
    flags = mc.MC_DEDUP | mc.MC_NORMALS
    s = torch.cuda.current_stream().cuda_stream
    org_a, sp_a = np.asarray(origin, np.float32), np.asarray(spacing, np.float32)
    ref, got = _outputs(U + 5, T + 5), _outputs(U + 5, T + 5)
    ws = torch.empty(int(lib.dax_hip_mc_workspace_bytes(nx, ny, nz)), dtype=torch.uint8,
                     device="cuda")
    _lib.call("dax_hip_mc_forward", field.data_ptr(), nx, ny, nz, org_a.ctypes.data,
              sp_a.ctypes.data, float(iso), flags, ws.data_ptr(), ws.numel(), *_args(ref), s)
    _lib.call("dax_hip_mc_forward_noalloc", field.data_ptr(), nx, ny, nz, org_a.ctypes.data,
              sp_a.ctypes.data, float(iso), flags, *_args(got), s)
    torch.cuda.synchronize()
    assert ref["k"].tolist() == [T, U] and got["k"].tolist() == [T, U]
    for k in "vncefsw":
        a, b = ref[k].cpu().numpy(), got[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b), (dims, k)
    return T, U


def test_noalloc_forward_matches_workspace_forward():
    print("[case] noalloc 200x150x170 then 96x80x70", flush=True)
    T1, U1 = _run_both((200, 150, 170), 3, 0.5, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    # the second, smaller grid runs over the first call's scratch contents
    T2, U2 = _run_both((96, 80, 70), 4, 0.45, (1.0, -2.0, 0.5), (0.5, 1.0, 2.0))
    assert (T2, U2) != (T1, U1)


def test_noalloc_forward_rejects_grids_beyond_its_scratch():
    lib = _lib.load()
    assert lib.dax_hip_mc_workspace_bytes(640, 640, 640) > (48 << 20)
    tiny = torch.zeros(8, device="cuda")
    o = _outputs(4, 4)
    s = torch.cuda.current_stream().cuda_stream
    with pytest.raises(Exception, match="needs .* bytes of scratch"):
        _lib.call("dax_hip_mc_forward_noalloc", tiny.data_ptr(), 640, 640, 640, None, None, 0.5,
                  mc.MC_DEDUP, *_args(o), s)
    assert o["k"].tolist() == [-1, -1]
