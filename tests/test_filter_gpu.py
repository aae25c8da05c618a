"""GPU equivalence filter vs the oracle (filter.cpp:18-163)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golaypair_amd as g
from tests import _oracle
from tests.test_oracle import apply, param_group

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _pairs_text(a, b):
    return "".join(g.pair_line(x, y) + "\n" for x, y in zip(a, b))


def _read(path, n):
    toks = np.array(open(path).read().split(), np.int64).reshape(-1, 2 * n)
    return toks[:, :n].astype(np.int8), toks[:, n:].astype(np.int8)


def _gpu_filter(ctx, order, compress, a, b):
    a = np.ascontiguousarray(a, np.int8)
    b = np.ascontiguousarray(b, np.int8)
    n = order // compress
    oa = np.zeros((max(len(a), 1), n), np.int8)
    ob = np.zeros((max(len(a), 1), n), np.int8)
    inp = g.GpPairs(n, len(a), len(a), g.ptr(a, g.ctypes.c_int8), g.ptr(b, g.ctypes.c_int8))
    out = g.GpPairs(n, 0, len(oa), g.ptr(oa, g.ctypes.c_int8), g.ptr(ob, g.ctypes.c_int8))
    st = g.GpStats()
    ctx.check(ctx.lib.gp_filter(ctx.h, order, compress, g.ctypes.byref(inp), g.ctypes.byref(out),
                                g.ctypes.byref(st)), "gp_filter")
    return oa[:out.count], ob[:out.count]


def test_filter_v34_known_answer(gpu_ctx):
    a, b = _read(os.path.join(GOLDEN, "v34", "34-pairs-1"), 34)
    oa, ob = _gpu_filter(gpu_ctx, 34, 1, a, b)
    assert len(oa) == 373
    assert _pairs_text(oa, ob) == open(os.path.join(GOLDEN, "v34", "34-unique")).read()


def test_filter_c1_golden(gpu_ctx):
    a, b = _read(os.path.join(GOLDEN, "v16_c2_1", "16-pairs-1"), 16)
    oa, ob = _gpu_filter(gpu_ctx, 16, 1, a, b)
    assert _pairs_text(oa, ob) == open(os.path.join(GOLDEN, "v16_c2_1", "16-unique")).read()


@pytest.mark.parametrize("n,compress", [(6, 1), (7, 1), (8, 1), (5, 3), (6, 2), (10, 1), (9, 1)])
def test_filter_random_orbits_vs_oracle(gpu_ctx, n, compress):
    rng = np.random.default_rng(100 + n)
    alt = compress == 1 and n % 2 == 0
    maps = param_group(n, alt)
    pairs = []
    for _ in range(8):
        p = (tuple(rng.choice([-1, 1], n)), tuple(rng.choice([-1, 1], n)))
        for _ in range(5):
            pairs.append(apply(maps[int(rng.integers(len(maps)))], *p))
    a = np.array([p[0] for p in pairs], np.int8)
    b = np.array([p[1] for p in pairs], np.int8)
    order = n * compress
    with tempfile.TemporaryDirectory() as d:
        wd = os.path.join(d, "order-%d" % order)
        os.makedirs(wd)
        open(os.path.join(wd, "%d-pairs-%d" % (order, compress)), "w").write(_pairs_text(a, b))
        subprocess.run([_oracle.CLI, str(order), "-c", str(compress), "-d", d, "filter"], check=True,
                       capture_output=True)
        ref = open(os.path.join(wd, "%d-unique" % order)).read()
    oa, ob = _gpu_filter(gpu_ctx, order, compress, a, b)
    assert _pairs_text(oa, ob) == ref
