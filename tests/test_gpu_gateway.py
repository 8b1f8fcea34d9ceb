"""The gateway-shaped pair gm_AxB_dot3_branch / gm_AxB_dot3 (the reference's
GB_cuda_AxB_dot3_branch / GB_cuda_AxB_dot3 argument list, flipxy included,
Source/gateway/GB_cuda_gateway.h:237-256) and the caller stream
(gm_set_stream), against the oracle.

Bar as for dot3: bit-exact for bool and integer semirings, 1e-12 relative
per entry for fp64.  flipxy swaps the multiplier's operands; every built-in
multiplier here commutes (TIMES, PLUS, PAIR, and the inner index k), so both
settings must give the oracle's result, and a multiplier outside that set is
declined rather than computed unswapped.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from helpers import SR, ZT, DevCsr, assert_csr_equal, download, gm_dot3, power_law_csr, rand_csr

pytestmark = pytest.mark.gpu

FP_TOL = 1e-12
ZNAME = {8: "int64", 6: "int32", 1: "bool", 11: "fp64"}
GM_NO_VALUE = 1


def _vals(handle):
    return {"GrB_PLUS_TIMES_SEMIRING_FP64": "fp64", "GrB_MIN_PLUS_SEMIRING_INT32": "int32",
            "GrB_PLUS_TIMES_SEMIRING_INT64": "int64", "GrB_MIN_PLUS_SEMIRING_INT64": "int64",
            "GrB_MAX_PLUS_SEMIRING_INT32": "int32", "GrB_MIN_PLUS_SEMIRING_FP64": "fp64"}.get(handle)


def _inputs(handle, seed):
    rng = np.random.default_rng(seed)
    vt = _vals(handle)
    A = power_law_csr(500, 12, rng, values=vt)
    B = power_law_csr(500, 12, rng, values=vt)
    M = rand_csr(500, 500, 0.04, rng, "int64", empty_rows=0.1)
    return M, A, B


def _gateway(trip, M, A, B, flipxy):
    ztype = trip[2]
    L = capi.lib()
    dM, dA, dB = DevCsr(M, 8), DevCsr(A, ztype), DevCsr(B, ztype)
    s = capi.GmSemiring(*trip)
    out = capi.GmCsrOut()
    try:
        branch = L.gm_AxB_dot3_branch(ctypes.byref(dM.view), True, ctypes.byref(dA.view),
                                      ctypes.byref(dB.view), ctypes.byref(s), flipxy)
        rc = L.gm_AxB_dot3(ctypes.byref(out), ctypes.byref(dM.view), True, ctypes.byref(dA.view),
                           ctypes.byref(dB.view), ctypes.byref(s), flipxy)
    finally:
        for d in (dM, dA, dB):
            d.free()
    return bool(branch), rc, out


@pytest.mark.parametrize("handle", list(SR))
@pytest.mark.parametrize("flipxy", [False, True])
def test_gateway_pair_flipxy(grb, handle, flipxy):
    print(f"[gateway] {handle} flipxy={flipxy}", flush=True)
    name, trip = SR[handle]
    M, A, B = _inputs(handle, abs(hash(handle)) % (2 ** 32))
    branch, rc, out = _gateway(trip, M, A, B, flipxy)
    assert branch
    assert rc == 0, capi.last_error()
    got = download(out, ZT[ZNAME[trip[2]]])
    ref = O.dot3(name, M, A, B, mask_struct=True)
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=handle)


def test_gateway_declines_other_multipliers(grb):
    """A multiplier code outside the built-in set: the branch says no and the
    call returns GM_NO_VALUE (the reference's 'not handled here'), whatever
    flipxy asks."""
    rng = np.random.default_rng(3)
    A = rand_csr(50, 40, 0.1, rng, "int64")
    M = rand_csr(50, 50, 0.1, rng, "int64")
    for flipxy in (False, True):
        branch, rc, out = _gateway((1, 99, 8), M, A, A, flipxy)
        assert not branch
        assert rc == GM_NO_VALUE


def _hip():
    """The HIP runtime the library already loaded (same process, same copy)."""
    capi.init()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "the HIP runtime is not loaded"
    h = ctypes.CDLL(path)
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    h.hipHostFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                 ctypes.c_void_p]
    return h


def test_caller_stream(grb):
    """gm_set_stream puts the library's work on the caller's stream: a column
    array copied to the device by an asynchronous copy on that stream, not
    waited for, is read after the copy; gm_set_stream(NULL) restores the
    library's own stream."""
    print("[gateway] caller stream", flush=True)
    hip = _hip()
    L = capi.lib()
    name, trip = SR["GxB_PLUS_PAIR_INT64"]
    rng = np.random.default_rng(11)
    A = power_law_csr(3000, 16, rng)
    M = rand_csr(3000, 3000, 0.01, rng, "int64")
    ref = O.dot3(name, M, A, A, mask_struct=True)
    own = L.gm_get_stream()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    pinned = ctypes.c_void_p()
    nb = A.j.nbytes
    assert hip.hipHostMalloc(ctypes.byref(pinned), nb, 0) == 0
    dM, dA = DevCsr(M, 8), DevCsr(A, 8)
    try:
        assert L.gm_set_stream(st) == 0
        assert L.gm_get_stream() == st.value
        # A's columns arrive by a copy on the caller's stream (pinned host
        # source, so the copy is asynchronous), then the product is queued
        # behind it on the same stream with no wait in between
        ctypes.memmove(pinned, A.j.ctypes.data, nb)
        dj = L.gm_malloc(nb)
        assert dj
        assert L.gm_memset(dj, 0, nb) == 0
        view = capi.GmCsr(A.nrows, A.ncols, dA.view.p, dj, None, 0, 8)
        assert hip.hipMemcpyAsync(dj, pinned, nb, 1, st) == 0
        s = capi.GmSemiring(*trip)
        out = capi.GmCsrOut()
        rc = L.gm_dot3(ctypes.byref(out), ctypes.byref(dM.view), True, ctypes.byref(view),
                       ctypes.byref(view), ctypes.byref(s), None)
        assert rc == 0, capi.last_error()
        got = download(out, np.int64)
        assert_csr_equal(got, ref, what="dot3 on the caller's stream")
        L.gm_free(dj)
        # the same product through gm_dot3's usual inputs, still on st
        assert_csr_equal(gm_dot3("GxB_PLUS_PAIR_INT64", M, A, A), ref, what="helpers on st")
    finally:
        assert L.gm_set_stream(None) == 0
        dM.free()
        dA.free()
        assert hip.hipStreamSynchronize(st) == 0
        hip.hipHostFree(pinned)
        hip.hipStreamDestroy(st)
    assert L.gm_get_stream() == own
    # and the library's own stream computes the same product afterwards
    assert_csr_equal(gm_dot3("GxB_PLUS_PAIR_INT64", M, A, A), ref, what="own stream")
