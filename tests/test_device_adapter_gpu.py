"""DeviceAdapterAlgorithm<Hip> primitives through the C-ABI, following the
reference's conformance suite (dax/cont/testing/TestingDeviceAdapter.h) and
checked against the oracle's Serial-adapter restatement / numpy."""
import ctypes

import numpy as np
import pytest
import torch

from daxtoolkit_amd import _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ARRAY_SIZE, OFFSET = 500, 1000  # TestingDeviceAdapter.h:25-26
S = lambda: torch.cuda.current_stream().cuda_stream


def stable_order(k, desc):
    """Permutation of a stable sort of k, ascending or descending."""
    if not desc:
        return np.argsort(k, kind="stable")
    _, rank = np.unique(k, return_inverse=True)
    return np.argsort(-rank.astype(np.int64), kind="stable")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_scan_golden():
    """TestingDeviceAdapter.h:940-1001."""
    for suf, dt in (("i32", np.int32), ("i64", np.int64)):
        a = dev(np.full(ARRAY_SIZE, OFFSET, dt))
        tot = (ctypes.c_int32 if dt == np.int32 else ctypes.c_int64)()
        _lib.call(f"dax_hip_scan_inclusive_{suf}", a.data_ptr(), a.data_ptr(), ARRAY_SIZE,
                  ctypes.byref(tot), S())
        assert tot.value == OFFSET * ARRAY_SIZE
        pos = np.arange(1, ARRAY_SIZE + 1)
        assert np.array_equal(np.cumsum(a.cpu().numpy().astype(np.int64)),
                              pos * (pos + 1) // 2 * OFFSET)
        b = dev(np.full(ARRAY_SIZE, OFFSET, dt))
        _lib.call(f"dax_hip_scan_exclusive_{suf}", b.data_ptr(), b.data_ptr(), ARRAY_SIZE,
                  ctypes.byref(tot), S())
        assert tot.value == OFFSET * ARRAY_SIZE
        pos = np.arange(ARRAY_SIZE)
        assert np.array_equal(np.cumsum(b.cpu().numpy().astype(np.int64)),
                              pos * (pos + 1) // 2 * OFFSET)


@pytest.mark.parametrize("n", [1, 7, 2048, 2049, 1_000_003])
def test_scan_vs_oracle(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 6, size=n).astype(np.int32)
    d = dev(a)
    out = torch.empty_like(d)
    tot = ctypes.c_int32()
    _lib.call("dax_hip_scan_inclusive_i32", d.data_ptr(), out.data_ptr(), n, ctypes.byref(tot), S())
    ref, rt = orc.scan_inclusive(a)
    assert np.array_equal(out.cpu().numpy(), ref) and tot.value == rt
    _lib.call("dax_hip_scan_exclusive_i32", d.data_ptr(), out.data_ptr(), n, ctypes.byref(tot), S())
    ref, rt = orc.scan_exclusive(a)
    assert np.array_equal(out.cpu().numpy(), ref) and tot.value == rt
    f = rng.random(n).astype(np.float32)
    fd = dev(f)
    fo = torch.empty_like(fd)
    ft = ctypes.c_float()
    _lib.call("dax_hip_scan_inclusive_f32", fd.data_ptr(), fo.data_ptr(), n, ctypes.byref(ft), S())
    np.testing.assert_allclose(fo.cpu().numpy(), np.cumsum(f.astype(np.float64)), rtol=1e-4)


def test_bounds_golden():
    """TestingDeviceAdapter.h:621-717 (sort, unique, lower/upper bounds)."""
    data = np.array([500, 955, 955, 120, 320, 955, 986, 281], np.int32)
    keys = dev(data.copy())
    _lib.call("dax_hip_sort_i32", keys.data_ptr(), 8, 0, S())
    n_new = ctypes.c_int64()
    _lib.call("dax_hip_unique_i32", keys.data_ptr(), 8, ctypes.byref(n_new), S())
    assert n_new.value == 6
    vals = dev(data)
    lo = torch.empty(8, dtype=torch.int32, device="cuda")
    hi = torch.empty(8, dtype=torch.int32, device="cuda")
    _lib.call("dax_hip_lower_bounds_i32", keys.data_ptr(), 6, vals.data_ptr(), 8, 0, lo.data_ptr(),
              None, S())
    _lib.call("dax_hip_upper_bounds_i32", keys.data_ptr(), 6, vals.data_ptr(), 8, 0, hi.data_ptr(),
              None, S())
    assert lo.cpu().tolist() == [3, 4, 4, 0, 2, 4, 5, 1]
    assert hi.cpu().tolist() == [4, 5, 5, 1, 3, 5, 6, 2]
    test = (OFFSET + np.arange(ARRAY_SIZE) % 50).astype(np.int32)
    t = dev(test.copy())
    _lib.call("dax_hip_sort_i32", t.data_ptr(), ARRAY_SIZE, 0, S())
    _lib.call("dax_hip_unique_i32", t.data_ptr(), ARRAY_SIZE, ctypes.byref(n_new), S())
    assert n_new.value == 50
    out = torch.empty(ARRAY_SIZE, dtype=torch.int64, device="cuda")
    _lib.call("dax_hip_lower_bounds_i32", t.data_ptr(), 50, dev(test).data_ptr(), ARRAY_SIZE, 0,
              None, out.data_ptr(), S())
    assert np.array_equal(out.cpu().numpy(), np.arange(ARRAY_SIZE) % 50)


@pytest.mark.parametrize("suf,dt", [("i32", np.int32), ("i64", np.int64), ("u32", np.uint32),
                                    ("f32", np.float32), ("f64", np.float64)])
@pytest.mark.parametrize("n", [3, 4096, 4097, 300_001])
def test_sort_and_sort_by_key(suf, dt, n):
    rng = np.random.default_rng(n)
    if np.issubdtype(dt, np.floating):
        k = (rng.standard_normal(n) * 100).astype(dt)
    elif dt == np.uint32:
        k = rng.integers(0, 2**32 - 1, size=n, dtype=np.uint64).astype(dt)
    else:
        k = rng.integers(-2**31, 2**31 - 1, size=n).astype(dt)
    k[: n // 3] = k[n // 3: 2 * (n // 3)]  # duplicates: stability matters
    for desc in (0, 1):
        keys = dev(k.copy())
        vals = dev(np.arange(n, dtype=np.int32))
        _lib.call(f"dax_hip_sort_by_key_{suf}", keys.data_ptr(), vals.data_ptr(), 4, n, desc, S())
        order = stable_order(k, desc)
        assert np.array_equal(keys.cpu().numpy(), k[order])
        assert np.array_equal(vals.cpu().numpy(), order.astype(np.int32))
        s = dev(k.copy())
        _lib.call(f"dax_hip_sort_{suf}", s.data_ptr(), n, desc, S())
        assert np.array_equal(s.cpu().numpy(), k[order])


def test_sort_by_key_vector_values():
    """TestingDeviceAdapter.h:770-838: keys ARRAY_SIZE-i, Vector3 values i."""
    keys = dev(np.arange(ARRAY_SIZE, 0, -1).astype(np.int32))
    vals = dev(np.repeat(np.arange(ARRAY_SIZE, dtype=np.float32)[:, None], 3, axis=1))
    _lib.call("dax_hip_sort_by_key_i32", keys.data_ptr(), vals.data_ptr(), 12, ARRAY_SIZE, 0, S())
    assert np.array_equal(keys.cpu().numpy(), np.arange(1, ARRAY_SIZE + 1))
    assert np.array_equal(vals.cpu().numpy()[:, 0], np.arange(ARRAY_SIZE - 1, -1, -1))
    _lib.call("dax_hip_sort_by_key_i32", keys.data_ptr(), vals.data_ptr(), 12, ARRAY_SIZE, 1, S())
    assert np.array_equal(keys.cpu().numpy(), np.arange(ARRAY_SIZE, 0, -1))
    assert np.array_equal(vals.cpu().numpy()[:, 0], np.arange(ARRAY_SIZE))


def test_stream_compact():
    """TestingDeviceAdapter.h:564-619."""
    stencil = dev((np.arange(ARRAY_SIZE) % 2).astype(np.int32))   # MarkOddNumbers
    idx = torch.empty(ARRAY_SIZE, dtype=torch.int32, device="cuda")
    cnt = ctypes.c_int64()
    _lib.call("dax_hip_stream_compact", None, 0, stencil.data_ptr(), 0, ARRAY_SIZE, None,
              idx.data_ptr(), None, ctypes.byref(cnt), S())
    assert cnt.value == ARRAY_SIZE // 2
    assert np.array_equal(idx.cpu().numpy()[:cnt.value], np.arange(cnt.value) * 2 + 1)
    vals = dev((OFFSET + np.arange(ARRAY_SIZE)).astype(np.int32))  # OffsetPlusIndex
    out = torch.empty_like(vals)
    _lib.call("dax_hip_stream_compact", vals.data_ptr(), 4, stencil.data_ptr(), 0, ARRAY_SIZE,
              out.data_ptr(), None, None, ctypes.byref(cnt), S())
    assert np.array_equal(out.cpu().numpy()[:cnt.value], OFFSET + np.arange(cnt.value) * 2 + 1)


def test_copy_ragged():
    for n in (1, 15, 16, 17, 1 << 20, (1 << 20) + 3):
        src = torch.randint(0, 255, (n + 1,), dtype=torch.uint8, device="cuda")
        dst = torch.zeros(n + 1, dtype=torch.uint8, device="cuda")
        _lib.call("dax_hip_copy", dst.data_ptr() + 1, src.data_ptr() + 1, n, S())
        torch.cuda.synchronize()
        assert torch.equal(dst[1:], src[1:]) and dst[0] == 0


@pytest.mark.parametrize("n,off", [(1, 0), (3, 1), (4097, 0), (4097, 3), (1_000_003, 2)])
def test_reduce_sum_i32(n, off):
    """dax_hip_reduce_sum_i32 (the count-handle total check of the generate
    dispatcher) against numpy's int64 sum, at unaligned starts and ragged
    tails, with values whose int32 sum would wrap."""
    rng = np.random.default_rng(n + off)
    a = rng.integers(-2**31, 2**31 - 1, size=n + off, dtype=np.int64).astype(np.int32)
    d = dev(a)
    total = ctypes.c_int64()
    _lib.call("dax_hip_reduce_sum_i32", d.data_ptr() + 4 * off, n, ctypes.byref(total), S())
    assert total.value == int(a[off:].astype(np.int64).sum())
