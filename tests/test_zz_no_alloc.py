"""The allocation-free filter entry point (pach_etc_to_filter_noalloc) against
pach_etc_to_filter, through the C ABI, on two calls in a row -- the second
smaller, so what the first left in the shared scratch array must not leak
into it -- and its refusal of inputs beyond the fixed scratch size.

The two entry points run the same convolution kernels on the same envelopes
(|prms| x 10^((120 - SWL) / 20), the 16 kHz band from the 8 kHz one), so the
outputs agree to the last bit; the bar written here is 1e-12 of the peak."""
import ctypes as C

import numpy as np
import pytest

from pachydermacoustic_universal_amd import _abi as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tracer():
    t = A.Tracer(0)
    yield t
    t.close()


def envelopes(n_ch, T, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(T) / 1000.0
    decay = np.exp(-t / 0.35)
    return rng.normal(size=(n_ch, 8, T)) * decay


def hip():
    """The HIP runtime libpach.so itself uses (the caller's device buffers are
    plain device allocations; the entry point under test allocates none)."""
    A.lib()
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


H2D, D2H = 1, 2


def run_noalloc(tracer, prms, swl, fs):
    h = hip()
    n_ch, _, T = prms.shape
    src = np.ascontiguousarray(prms, dtype=np.float64)
    out = np.full((n_ch, T + A.CARRIER_LEN - 1), np.nan)
    d_in, d_out = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(d_in), src.nbytes) == 0
    assert h.hipMalloc(C.byref(d_out), out.nbytes) == 0
    try:
        assert h.hipMemcpy(d_in, src.ctypes.data, src.nbytes, H2D) == 0
        assert h.hipMemcpy(d_out, out.ctypes.data, out.nbytes, H2D) == 0  # NaN: every element must be written
        tracer.etc_to_filter_noalloc(d_in.value, n_ch, T, swl, fs, d_out.value)
        assert h.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    finally:
        h.hipFree(d_in)
        h.hipFree(d_out)
    return out


def test_noalloc_filters_equal_the_allocating_entry_point(tracer):
    swl = np.array([100.0, 104, 108, 110, 109, 105, 99, 92])
    for n_ch, T, seed in ((11, 2400, 1), (3, 900, 2)):  # the second call smaller
        print(f"[case] noalloc filters n_ch={n_ch} T={T}", flush=True)
        prms = envelopes(n_ch, T, seed)
        ref = tracer.etc_to_filter(prms, swl, 1000)
        got = run_noalloc(tracer, prms, swl, 1000)
        assert got.shape == ref.shape
        assert np.isfinite(got).all()
        scale = np.abs(ref).max()
        assert scale > 0
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale)


def test_noalloc_rejects_inputs_beyond_the_scratch(tracer):
    # 64 channels of 20,000 samples need ~330 MB of band planes: refused
    # before any device work (the pointers are never dereferenced)
    sw = np.full(8, 120.0)
    rc = A.lib().pach_etc_to_filter_noalloc(tracer.h, C.c_void_p(16), 64, 20000, A._p(sw), 1000, C.c_void_p(16))
    assert rc == -6  # PACH_ERANGE
    assert b"scratch" in A.lib().pach_last_error(tracer.h)
