"""Batched PSD filter on the GPU vs the FP64 oracle (fourier.cpp:25-52).

Decisions must be bit-exact (the kernels settle every candidate inside their
stated error band in FP64); the kernels' own max-PSD values must lie within the
error bound of the split-BF16 and FP32 paths (|PSD' - PSD| <= eps(n, c), with
the bound recomputed here from the table constants).
"""
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


@pytest.mark.parametrize("name,mode", MODES)
def test_psd_batch_matches_oracle(gpu_ctx, name, mode):
    rng = np.random.default_rng(1234)
    words, lens, cs = _random_batch(rng, 20000, [2, 3, 8, 9, 10, 15, 16, 17, 18, 25, 32, 33, 45, 50, 64], [1, 2, 3, 5])
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
    # the kernel's own max PSD is within its bound of the FP64 value
    err = np.abs(gp_max.astype(np.float64) - ref_max)
    rel = err / np.maximum(1.0, ref_max)
    assert rel.max() < 1e-3, rel.max()
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
