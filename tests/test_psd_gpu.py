"""Batched PSD filter on the GPU vs the FP64 oracle (fourier.cpp:25-52).

Decisions must be bit-exact (the kernels settle every candidate inside their
stated error band in FP64); the kernels' own max-PSD values must lie within the
error bound of the split-BF16 and FP32 paths, row by row:
  |PSD' - PSD| <= 2 E sqrt(2 P) + 2 E^2 + 2.0001 u P,   P = PSD', u = 2^-24,
with E the bound on the error of Re X_k and Im X_k recomputed here for each
(n, order, paf) segment the kernel runs:
  split-BF16  E = S1 (rho + 1.004 (32 nb + 1) 2^-23), rho = max |w - hi - lo|
              over the segment's cos/sin entries split into two BF16 parts;
  FP32        E = S1 (n + 2) 2^-24 1.01;
S1 = max over the segment's rows of sum |x_p|, nb = ceil(n/16).  Near the
threshold thr this is the band half-width the kernels use (P <= thr + 1), so a
row whose error exceeded it could be decided wrongly outside the FP64 band.
"""
import math

import numpy as np
import pytest

import golaypair_amd as g
from tests import _oracle

pytestmark = pytest.mark.gpu

MODES = [("split_bf16", g.GP_PSD_SPLIT_BF16), ("fp32", g.GP_PSD_FP32)]


def _random_batch(rng, count, lens_choice, c_choice):
    lens = rng.choice(lens_choice, size=count).astype(np.int32)
    cs = rng.choice(c_choice, size=count)
    words = np.zeros((count, 64), dtype=np.int8)
    for i in range(count):
        c = int(cs[i])
        idx = rng.integers(0, c + 1, size=lens[i])
        words[i, :lens[i]] = -c + 2 * idx
    return words, lens, cs


def _twiddles(n):
    """FP64 cos/sin of 2 pi m / n, symmetric in m <-> n - m, exact at m = 0 and n/2."""
    c, s = np.zeros(n), np.zeros(n)
    for m in range(n // 2 + 1):
        a = 2.0 * math.pi * m / n
        cm, sm = math.cos(a), math.sin(a)
        if m == 0:
            cm, sm = 1.0, 0.0
        if 2 * m == n:
            cm, sm = -1.0, 0.0
        c[m], s[m] = cm, sm
        if m != 0 and 2 * m != n:
            c[n - m], s[n - m] = cm, -sm
    return c, s


def _bf16(x):
    """Round-to-nearest-even of float32 values to BF16, returned as float64."""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _split_residual(n):
    """w - hi - lo for every (channel k = 1..n/2, Re/Im, entry e) of the length-n table."""
    c, s = _twiddles(n)
    K = n // 2
    m = (np.arange(1, K + 1)[:, None] * np.arange(n)[None, :]) % n
    w = np.stack([c[m], s[m]])                       # [2, K, n]
    hi = _bf16(w)
    lo = _bf16(w - hi)
    return w - hi - lo


def _row_bounds(words, lens, orders, pafs, mode, gp_max):
    """Per-row bound on |PSD' - PSD| at P = the row's PSD' maximum."""
    bound = np.zeros(len(lens))
    keys = {}
    for i, key in enumerate(zip(lens.tolist(), orders.tolist(), pafs.tolist())):
        keys.setdefault(key, []).append(i)
    for (n, _, _), rows in keys.items():
        rows = np.array(rows)
        s1 = np.abs(words[rows, :n].astype(np.int64)).sum(axis=1).max()
        if mode == g.GP_PSD_SPLIT_BF16:
            rho = float(np.float32(np.abs(_split_residual(n)).max() * 1.0000001)) if n >= 2 else 0.0
            E = s1 * (rho + 1.004 * (32.0 * ((n + 15) // 16) + 1.0) * 2.0 ** -23)
        else:
            E = s1 * (n + 2) * 2.0 ** -24 * 1.01
        P = gp_max[rows].astype(np.float64)
        bound[rows] = 2 * E * np.sqrt(2 * P) * 1.0000001 + 2 * E * E + 2.0001 * 2.0 ** -24 * P
    return bound


def _adversarial_rows(c):
    """Words x_e = c sign(w - hi - lo) of channel Re X_1: the BF16 split residuals
    of that channel all add up."""
    rows, lens = [], []
    for n in (9, 10, 15, 17, 18, 25, 45, 50, 64):
        res = _split_residual(n)[0, 0]
        w = np.zeros(64, np.int8)
        w[:n] = np.where(res >= 0, c, -c)
        rows.append(w)
        lens.append(n)
    return rows, lens


@pytest.mark.parametrize("name,mode", MODES)
def test_psd_batch_matches_oracle(gpu_ctx, name, mode):
    rng = np.random.default_rng(1234)
    words, lens, cs = _random_batch(rng, 20000, [2, 3, 8, 9, 10, 15, 16, 17, 18, 25, 32, 33, 45, 50, 64], [1, 2, 3, 5])
    adv, adv_lens = _adversarial_rows(5)
    words = np.concatenate([words, np.stack(adv)])
    lens = np.concatenate([lens, np.array(adv_lens, np.int32)])
    cs = np.concatenate([cs, np.full(len(adv), 5)])
    # orders chosen so that the threshold cuts through the PSD distribution
    orders = np.maximum(1, (lens * (cs ** 2 + 1)) // 2).astype(np.int32)
    pafs = np.zeros_like(orders)
    pafs[::3] = -2
    gpu_ctx.set_mode(mode)
    try:
        gp_pass, gp_max, st = gpu_ctx.psd_filter_batch(words, lens, orders, pafs, want_max=True)
    finally:
        gpu_ctx.set_mode(g.GP_PSD_SPLIT_BF16)
    ref_pass, ref_max = _oracle.psd_filter_batch(words, lens, orders, pafs)
    assert 0.05 < ref_pass.mean() < 0.95, "threshold should split the batch"
    mism = np.nonzero(gp_pass != ref_pass)[0]
    assert len(mism) == 0, "decision mismatch at rows %s" % mism[:10]
    # the kernel's own max PSD is within its stated bound of the FP64 value, row by row
    err = np.abs(gp_max.astype(np.float64) - ref_max)
    bound = _row_bounds(words, lens, orders, pafs, mode, gp_max)
    worst = np.argmax(err / np.maximum(bound, 1e-300))
    assert (err <= bound).all(), "row %d: err %.3g > bound %.3g (n=%d)" % (worst, err[worst], bound[worst],
                                                                          lens[worst])
    assert st.psd_decisions == len(words)
    assert st.tie_band == 0


@pytest.mark.parametrize("name,mode", MODES)
def test_psd_edge_cases(gpu_ctx, name, mode):
    # all-zero words (PSD 0 passes), a single spike (flat PSD = x^2), an
    # alternating word (all energy at Nyquist), maximal letters.
    rows = []
    lens = []
    for n in (2, 7, 16, 31, 64):
        z = np.zeros(64, np.int8)
        rows.append(z.copy()); lens.append(n)
        s = z.copy(); s[0] = 10; rows.append(s); lens.append(n)       # PSD_k = 100 for all k
        a = z.copy(); a[:n] = [(-1) ** j for j in range(n)]; rows.append(a); lens.append(n)
        m = z.copy(); m[:n] = 127; rows.append(m); lens.append(n)       # PSD_k = 0 for k != 0
    words = np.stack(rows)
    lens = np.array(lens, np.int32)
    # threshold exactly at the spike's flat PSD: 2*order - paf + 0.001 = 100.001 passes
    orders = np.full(len(lens), 50, np.int32)
    pafs = np.zeros(len(lens), np.int32)
    gpu_ctx.set_mode(mode)
    try:
        gp_pass, _, _ = gpu_ctx.psd_filter_batch(words, lens, orders, pafs)
    finally:
        gpu_ctx.set_mode(g.GP_PSD_SPLIT_BF16)
    ref_pass, _ = _oracle.psd_filter_batch(words, lens, orders, pafs)
    assert (gp_pass == ref_pass).all()


MIX_SMALL = [  # n, c, order, paf, share  (the microbenchmark mix at 1/64 scale)
    [9, 10, 90, 0, 1 << 19],
    [10, 5, 50, 0, 1 << 18],
    [15, 3, 45, -2, 1 << 17],
    [17, 3, 51, -2, 1 << 16],
    [18, 5, 90, 0, 1 << 15],
    [25, 2, 50, 0, 1 << 15],
]


@pytest.mark.parametrize("name,mode", MODES)
def test_microbench_pass_set_equals_oracle(gpu_ctx, name, mode):
    seed = 0x601A7E5EED
    mix = np.array(MIX_SMALL, np.int32)
    count = int(mix[:, 4].sum())
    gpu_ctx.set_mode(mode)
    try:
        acc, digest, pas, st = gpu_ctx.psd_microbench(seed, mix, pass_cap=count)
    finally:
        gpu_ctx.set_mode(g.GP_PSD_SPLIT_BF16)
    # regenerate every word on the host and run the oracle
    kinds, W = g.microbench_words(seed, mix, 0, count, stride=32)
    lens = mix[kinds, 0]
    ref_pass, _ = _oracle.psd_filter_batch(W, lens, mix[kinds, 2], mix[kinds, 3])
    assert (pas == ref_pass).all(), np.nonzero(pas != ref_pass)[0][:10]
    for k in range(len(mix)):
        assert acc[k] == ref_pass[kinds == k].sum()
        assert (kinds == k).sum() == mix[k, 4]
    assert st.tie_band == 0
