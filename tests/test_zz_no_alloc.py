"""gp_psd_filter_batch_static, the PSD filter entry point that allocates nothing
(device words and pass flags in, working memory in a static device array), vs
gp_psd_filter_batch on the same words: the same pass flags, bit for bit, on
calls in a row (a large batch, then a smaller one, then other word lengths, so
the array's table and counters are reused and replaced)."""
import ctypes

import numpy as np
import pytest

import golaypair_amd as g

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself maps (device buffers for the test come
    from it: the entry point takes device pointers)."""
    g.load_library()
    path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l)
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def _static(ctx, words, n, order, paf):
    hip = _hip()
    words = np.ascontiguousarray(words)
    count, stride = words.shape
    dw, dp = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dw), max(words.nbytes, 1)) == 0
    assert hip.hipMalloc(ctypes.byref(dp), max(count, 1)) == 0
    try:
        assert hip.hipMemcpy(dw, words.ctypes.data, words.nbytes, H2D) == 0
        garbage = np.full(max(count, 1), 7, np.uint8)                   # overwritten by the call
        assert hip.hipMemcpy(dp, garbage.ctypes.data, garbage.nbytes, H2D) == 0
        st = g.GpStats()
        rc = ctx.lib.gp_psd_filter_batch_static(ctx.h, dw, stride, n, order, paf, count, dp, ctypes.byref(st))
        out = np.zeros(max(count, 1), np.uint8)
        assert hip.hipMemcpy(out.ctypes.data, dp, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(dw)
        hip.hipFree(dp)
    return rc, out[:count], st


def _words(rng, count, n, c, stride):
    w = np.zeros((count, stride), np.int8)
    w[:, :n] = -c + 2 * rng.integers(0, c + 1, size=(count, n))
    return w


@pytest.mark.parametrize("calls", [
    [(45, 1, 45, -2, 200000), (45, 1, 45, -2, 3000)],            # the second call smaller
    [(18, 5, 90, 0, 60000), (10, 5, 50, 0, 7000), (18, 5, 90, 0, 5000), (64, 3, 400, 0, 4000)],
])
def test_static_entry_matches_batch_entry(gpu_ctx, calls):
    rng = np.random.default_rng(len(calls))
    for n, c, order, paf, count in calls:
        stride = 16 * ((n + 15) // 16)
        w = _words(rng, count, n, c, stride)
        rc, got, st = _static(gpu_ctx, w, n, order, paf)
        assert rc == g.GP_OK, gpu_ctx.lib.gp_last_error(gpu_ctx.h)
        want, _, st2 = gpu_ctx.psd_filter_batch(w, np.full(count, n, np.int32), np.full(count, order, np.int32),
                                                np.full(count, paf, np.int32))
        assert 0 < want.sum() < count, "threshold should split the batch"
        assert (got == want).all(), (n, count, np.nonzero(got != want)[0][:10])
        assert st.psd_decisions == count and st.psd_pass == want.sum()
        assert st.tie_band == 0


def test_static_entry_rejects_sizes_beyond_its_scratch(gpu_ctx):
    w = np.zeros((4, 80), np.int8)
    rc, _, _ = _static(gpu_ctx, w, 65, 100, 0)             # longer than the one table it holds
    assert rc == g.GP_ERR_ARG
    assert b"outside [2, 64]" in gpu_ctx.lib.gp_last_error(gpu_ctx.h)
    w = np.zeros((4, 16), np.int8)
    rc, _, _ = _static(gpu_ctx, w, 20, 100, 0)             # stride below 16 * ceil(n / 16)
    assert rc == g.GP_ERR_ARG
