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


def _random_map(rng, n, alt):
    units = [u for u in range(1, max(n, 2)) if np.gcd(u, n) == 1] or [0]
    return (int(rng.integers(2)), int(rng.choice(units)), int(rng.choice([1, -1])),
            int(rng.integers(2)) if alt else 0, int(rng.choice([1, -1])) if alt else 1,
            int(rng.choice([1, -1])) if alt else 1, int(rng.integers(n)), int(rng.integers(n)))


# (n, compress, letters): odd and even alphabets, packed keys of 1..8 words, and
# symmetric pairs whose a-part minimum is reached by many transforms
@pytest.mark.parametrize("n,compress,letters", [
    (45, 1, (-1, 1)), (20, 1, (-1, 1)), (15, 3, (-3, -1, 1, 3)), (18, 5, (-5, -3, -1, 1, 3, 5)),
    (12, 2, (-2, 0, 2)), (20, 10, tuple(range(-10, 11, 2))), (40, 4, (-4, -2, 0, 2, 4)),
    (60, 20, tuple(range(-20, 21, 2)))])
def test_filter_alphabets_vs_oracle(gpu_ctx, n, compress, letters):
    rng = np.random.default_rng(7 * n + compress)
    alt = compress == 1 and n % 2 == 0
    pairs = []
    for k in range(6):
        if k == 0:
            p = (tuple([letters[-1]] * n), tuple([letters[0]] * n))
        elif k == 1:
            p = (tuple(letters[i % 2] for i in range(n)), tuple(letters[(i // 2) % len(letters)] for i in range(n)))
        else:
            p = (tuple(int(x) for x in rng.choice(letters, n)), tuple(int(x) for x in rng.choice(letters, n)))
        for _ in range(4):
            pairs.append(apply(_random_map(rng, n, alt), *p))
    a = np.array([p[0] for p in pairs], np.int8)
    b = np.array([p[1] for p in pairs], np.int8)
    order = n * compress
    with tempfile.TemporaryDirectory() as d:
        wd = os.path.join(d, "order-%d" % order)
        os.makedirs(wd)
        open(os.path.join(wd, "%d-pairs-%d" % (order, compress)), "w").write(_pairs_text(a, b))
        subprocess.run([_oracle.CLI, str(order), "-c", str(compress), "-d", d, "filter"], check=True,
                       capture_output=True, timeout=600)
        ref = open(os.path.join(wd, "%d-unique" % order)).read()
    oa, ob = _gpu_filter(gpu_ctx, order, compress, a, b)
    assert _pairs_text(oa, ob) == ref


def _lexmin_rows(m):
    """Lexicographically smallest row of a 2-D integer array."""
    order = np.lexsort(m.T[::-1])
    return m[order[0]]


def orbit_min_brute(a, b, alt):
    """Minimum of the pair's orbit under the signed affine family, by brute force
    over every group element (swap, u, e, kappa, sa, sb, ta, tb): for a fixed
    (swap, u, e, kappa, sa, sb) the image's a-part depends only on ta and its b-part
    only on tb, so the lexicographic minimum over (ta, tb) is (min over ta, min over tb)."""
    n = len(a)
    i = np.arange(n)
    shifts = np.arange(n)[:, None]
    best = None
    for sw in (0, 1):
        X, Y = (np.asarray(b), np.asarray(a)) if sw else (np.asarray(a), np.asarray(b))
        for u in [u for u in range(1, n) if np.gcd(u, n) == 1]:
            for e in (1, -1):
                for kappa in ((0, 1) if alt else (0,)):
                    sgn = np.where((i % 2 == 1) & (kappa == 1), -1, 1)
                    imgs_a = X[(u * i + shifts) % n] * sgn          # row ta: the image's a-part
                    imgs_b = Y[(e * u * i + shifts) % n] * sgn      # row tb: the image's b-part
                    for sa in ((1, -1) if alt else (1,)):
                        ma = tuple(_lexmin_rows(imgs_a * sa).tolist())
                        for sb in ((1, -1) if alt else (1,)):
                            cand = ma + tuple(_lexmin_rows(imgs_b * sb).tolist())
                            if best is None or cand < best:
                                best = cand
    return best


def brute_unique(a, b, alt):
    """The lexicographically smallest input pair of each orbit, in (a, b) order."""
    keep = {}
    for x, y in zip(a, b):
        key = orbit_min_brute(x, y, alt)
        p = (tuple(int(v) for v in x), tuple(int(v) for v in y))
        if key not in keep or p < keep[key]:
            keep[key] = p
    return sorted(keep.values())


def test_filter_length50_altnegative_vs_brute_force(gpu_ctx):
    """Length 50 with altnegative (order 50, compress 1, filter.cpp:83-87): the group
    has 2 * 20 * 2 * 2 * 4 * 50^2 = 1.6e6 elements, beyond the oracle's closure.
    Planted orbits (random +-1 pairs, five random images each, plus a symmetric
    pair) against the brute-force orbit minimum over every group element."""
    n = 50
    rng = np.random.default_rng(50)
    pairs = []
    bases = [(tuple(int(x) for x in rng.choice([-1, 1], n)), tuple(int(x) for x in rng.choice([-1, 1], n)))
             for _ in range(5)]
    bases.append((tuple((-1) ** (i // 2) for i in range(n)), tuple((-1) ** (i // 5) for i in range(n))))
    for p in bases:
        for _ in range(5):
            pairs.append(apply(_random_map(rng, n, True), *p))
    rng.shuffle(pairs)
    a = np.array([p[0] for p in pairs], np.int8)
    b = np.array([p[1] for p in pairs], np.int8)
    want = brute_unique(a, b, True)
    assert len(want) == len(bases)
    oa, ob = _gpu_filter(gpu_ctx, n, 1, a, b)
    assert [(tuple(int(v) for v in x), tuple(int(v) for v in y)) for x, y in zip(oa, ob)] == want
