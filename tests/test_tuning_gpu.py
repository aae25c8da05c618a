"""The tuning table (golaypair_amd/csrc/tuning.hip, gp_set_tuning / gp_get_tuning):
every setting gives the same output, only the time changes.  Non-default
settings (one wave per SIMD, small batches, extreme ranks per lane, filter
work items by wave index but for the last rounds, one, three or five work
counters) against the built-in defaults, on gen and on a slice of an
uncompression level."""
import ctypes

import numpy as np
import pytest

import golaypair_amd as g

pytestmark = pytest.mark.gpu


def _set(ctx, key, value):
    assert ctx.lib.gp_set_tuning(ctx.h, key.encode(), value) == g.GP_OK


def _get(ctx, key):
    v = ctypes.c_int64()
    rc = ctx.lib.gp_get_tuning(ctx.h, key.encode(), ctypes.byref(v))
    return v.value if rc == g.GP_OK else None


def test_tuning_set_get_roundtrip(gpu_ctx):
    key = "unc_waves/9999"
    _set(gpu_ctx, key, 3)
    assert _get(gpu_ctx, key) == 3
    _set(gpu_ctx, key, -1)
    assert _get(gpu_ctx, key) is None


@pytest.mark.parametrize("order,c0,c1,paf,pairs,settings", [
    (45, 3, 1, -2, (0, 300), {"unc_waves/45": 1, "unc_leaf_budget/45": 1 << 20, "gen_ranks_per_lane/15/4": 4,
                              "unc_counter_rounds/45": 0, "unc_counter_streams/45": 1}),
    # matched pairs 46..48 of v=50 emit c=1 pairs (tests/golden/slice_v50_c5_1.json)
    (50, 5, 1, 0, (46, 49), {"unc_waves/50": 2, "unc_leaf_budget/50": 1 << 22, "gen_ranks_per_lane/10/6": 1024,
                             "unc_counter_rounds/50": 2, "unc_counter_streams/50": 3}),
    (90, 10, 5, 0, (0, 40), {"unc_waves/18": 1, "unc_leaf_budget/18": 1 << 18, "gen_ranks_per_lane/9/11": 7,
                             "unc_counter_rounds/18": 1, "unc_counter_streams/18": 5}),
])
def test_tuned_settings_give_default_output(gpu_ctx, order, c0, c1, paf, pairs, settings):
    saved = {k: _get(gpu_ctx, k) for k in settings}
    try:
        results = []
        for use in (False, True):
            for k, v in settings.items():
                _set(gpu_ctx, k, v if use else -1)
            a, b, _ = gpu_ctx.generate(order, c0, paf)
            rc, pa, pb, _ = gpu_ctx.match(order, c0, paf, gpu_ctx.sort_candidates(a), gpu_ctx.sort_candidates(b))
            oa, ob, _, _, st = gpu_ctx.uncompress(order, c0, c1, paf, pa[pairs[0]:pairs[1]], pb[pairs[0]:pairs[1]])
            results.append((a.words.tobytes(), a.keys.tobytes(), b.words.tobytes(), b.keys.tobytes(),
                            oa.tobytes(), ob.tobytes(), st.psd_decisions, st.psd_pass))
        assert results[0] == results[1]
        assert len(results[0][4]) > 0
    finally:
        for k, v in saved.items():
            _set(gpu_ctx, k, -1 if v is None else v)
