"""rumdx_reduce_f4_static (no caller workspace, no allocation) against
rumdx_reduce_f4 through the C ABI: bit-identical sums on two calls in a row,
the second with a smaller input, and the size limit rejected on the host."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from capi import DeviceBuffer, check, gpu_available, lib

pytestmark = pytest.mark.gpu

MAX_N = 8388608


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _bind():
    L = lib()
    L.rumdx_reduce_f4_static.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    L.rumdx_reduce_f4_static.restype = C.c_int
    return L


def _both(L, buf, n, mask, ws, out_ws, out_static):
    print(f"[case] reduce_f4 n={n} mask={mask:#x}", flush=True)
    check(L.rumdx_reduce_f4(buf.value, n, mask, out_ws.value, ws.value, None))
    check(L.rumdx_reduce_f4_static(buf.value, n, mask, out_static.value, None))
    return out_ws.download(np.float64, 4), out_static.download(np.float64, 4)


def test_static_reduce_matches_workspace_reduce_twice():
    L = _bind()
    rng = np.random.default_rng(11)
    n1, n2 = 1_048_577, 4_099  # second call smaller: stale partials past its blocks must not leak in
    a = rng.normal(size=(n1, 4)).astype(np.float32)
    buf = DeviceBuffer.from_array(a)
    ws = DeviceBuffer(L.rumdx_workspace_bytes(n1))
    out_ws, out_static = DeviceBuffer(32), DeviceBuffer(32)
    s_ws, s_st = _both(L, buf, n1, 15, ws, out_ws, out_static)
    assert np.array_equal(s_ws, s_st)
    ref = a.astype(np.float64).sum(axis=0)
    np.testing.assert_allclose(s_st, ref, rtol=1e-12, atol=1e-9)
    s_ws, s_st = _both(L, buf, n2, 0b1010, ws, out_ws, out_static)
    assert np.array_equal(s_ws, s_st)
    assert s_st[0] == 0.0 and s_st[2] == 0.0
    ref2 = a[:n2].astype(np.float64).sum(axis=0)
    np.testing.assert_allclose(s_st[[1, 3]], ref2[[1, 3]], rtol=1e-12, atol=1e-9)
    # integer-valued data sums exactly
    b = rng.integers(-1000, 1000, size=(n2, 4)).astype(np.float32)
    buf.upload(b)
    s_ws, s_st = _both(L, buf, n2, 15, ws, out_ws, out_static)
    assert np.array_equal(s_st, b.astype(np.int64).sum(axis=0).astype(np.float64))
    assert np.array_equal(s_ws, s_st)


def test_static_reduce_rejects_sizes_beyond_its_scratch():
    L = _bind()
    buf, out = DeviceBuffer(64), DeviceBuffer(32)
    rc = L.rumdx_reduce_f4_static(buf.value, MAX_N + 1, 15, out.value, None)
    assert rc != 0
    assert b"exceeds the static scratch" in L.rumdx_last_error()
    assert L.rumdx_reduce_f4_static(buf.value, 0, 15, out.value, None) != 0
