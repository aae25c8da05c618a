"""Device sort + uniq of candidate records (gp_sort_candidates, the GNU_sort seam:
src/core/sort/sort.cpp:7-23 run as `LC_ALL=C sort | uniq` on the lines of
src/utils/io.cpp:14-36).  The sorted records, written as text, must be the
byte-ordered unique lines: against Python's code-point order on ASCII text (the
C locale's byte order) for random record sets, and against coreutils itself on
the gen files."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golaypair_amd as g
from tests import _oracle

pytestmark = pytest.mark.gpu


def _random_records(rng, m, n, letters, key_values):
    words = rng.choice(letters, size=(m, n)).astype(np.int8)
    keys = rng.choice(key_values, size=(m, n // 2)).astype(np.uint8)
    if m > 4:  # duplicates: copies of earlier rows, and rows equal but for one key digit
        dup = rng.integers(0, m, size=m // 4)
        words[rng.integers(0, m, size=m // 4)] = words[dup]
        keys[rng.integers(0, m, size=m // 4)] = keys[dup]
    return g.Cands(words, keys)


@pytest.mark.parametrize("n,m,letters,key_values", [
    (2, 50, [-1, 1], [0, 1, 2, 10]),
    (3, 1, [-1, 1], [5]),
    (8, 3000, [-2, 0, 2], list(range(0, 40))),
    (15, 60000, [-3, -1, 1, 3], list(range(0, 93))),                    # 15 radix blocks of 4096
    (18, 20000, list(range(-10, 11, 2)), [0, 1, 9, 10, 11, 99, 100, 101, 180, 255]),
    (25, 5000, list(range(-20, 21, 2)), list(range(0, 256))),           # 21 letters: 5-bit symbols
    (45, 8000, [-1, 1], [1, 12, 2, 23, 3, 123, 31, 231]),               # key texts that collide as prefixes
    (50, 30000, [-1, 1], list(range(0, 101))),
    (64, 2000, [-128, -100, -7, 0, 7, 100, 127], [0, 255, 100, 99, 9, 10]),
])
def test_sort_random_records_byte_order(gpu_ctx, n, m, letters, key_values):
    rng = np.random.default_rng(n * 1000 + m)
    c = _random_records(rng, m, n, letters, key_values)
    want = sorted(set(c.lines()))
    got = gpu_ctx.sort_candidates(c)
    assert got.lines() == want


def test_sort_empty_and_all_equal(gpu_ctx):
    e = g.Cands(np.zeros((0, 10), np.int8), np.zeros((0, 5), np.uint8))
    assert len(gpu_ctx.sort_candidates(e)) == 0
    w = np.tile(np.array([[1, -1, 1, 1, -1, 1, -1, -1, 1, 1]], np.int8), (9000, 1))
    k = np.tile(np.array([[3, 14, 15, 92, 65]], np.uint8), (9000, 1))
    got = gpu_ctx.sort_candidates(g.Cands(w, k))
    assert len(got) == 1 and got.lines() == [g.cand_line(w[0], k[0])]


@pytest.mark.parametrize("order,compress,paf", [(26, 2, 0), (34, 2, 0), (45, 3, -2), (50, 5, 0)])
def test_sort_gen_files_equal_coreutils(gpu_ctx, order, compress, paf):
    """The GPU's gen candidates, shuffled and duplicated, sorted on the GPU: the same
    lines as coreutils `sort | uniq` with LC_ALL=C on the oracle's gen files."""
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([_oracle.CLI, "-d", d, str(order), "-c", str(compress), "--paf", str(paf), "gen"], check=True,
                       capture_output=True)
        wd = os.path.join(d, "order-%d" % order)
        a, b, _ = gpu_ctx.generate(order, compress, paf)
        rng = np.random.default_rng(order)
        for side, cands in (("a", a), ("b", b)):
            path = os.path.join(wd, "%d-filtered-%s#0" % (order, side))
            env = dict(os.environ, LC_ALL="C")
            ref = subprocess.run("sort %s | uniq" % path, shell=True, check=True, capture_output=True, env=env,
                                 text=True).stdout
            perm = rng.permutation(np.concatenate([np.arange(len(cands)), rng.integers(0, len(cands), 100)]))
            mixed = g.Cands(cands.words[perm], cands.keys[perm])
            got = gpu_ctx.sort_candidates(mixed)
            assert "".join(l + "\n" for l in got.lines()) == ref, side
