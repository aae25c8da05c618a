"""GPU gen / sort / match vs the oracle, byte-exact on the file text.

gen: generate_hybrid.cpp:107-163 (A/B candidate files in FKM order)
sort: sort.cpp:7-23 under LC_ALL=C (+ uniq)
match: match_pairs.cpp:15-121 (pair lines in emission order)
"""
import hashlib
import json
import os

import numpy as np
import pytest

import golaypair_amd as g
from tests import _oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _text(lines):
    return "".join(l + "\n" for l in lines)


SMALL = [  # (order, compress, paf, job, njobs): oracle runs in < ~10 s
    (8, 1, 0, 0, 1),
    (10, 1, 0, 0, 1),
    (16, 2, 0, 0, 1),
    (16, 1, 0, 0, 1),
    (26, 2, 0, 0, 1),
    (26, 2, 0, 1, 3),
    (26, 2, 0, 2, 3),
    (20, 2, 0, 0, 1),
    (18, 3, -2, 0, 1),
    (50, 5, 0, 0, 1),
    (34, 2, 0, 0, 1),
]


@pytest.mark.parametrize("order,compress,paf,job,njobs", SMALL)
def test_gen_matches_oracle(gpu_ctx, order, compress, paf, job, njobs):
    rc, ta, tb, _ = _oracle.generate(order, compress, paf, job, njobs)
    a, b, st = gpu_ctx.generate(order, compress, paf, job, njobs)
    if rc == 1:
        assert a is None
        return
    assert _text(a.lines()) == ta
    assert _text(b.lines()) == tb
    assert st.tie_band == 0


def test_gen_no_decomposition(gpu_ctx):
    # 2*7 = 14 is not a sum of two squares: the reference prints and returns 1
    a, b, st = gpu_ctx.generate(7, 1, 0)
    assert a is None and b is None


@pytest.mark.parametrize("order,compress,paf", [(16, 2, 0), (26, 2, 0), (50, 5, 0), (34, 2, 0)])
def test_sort_and_match_match_oracle(gpu_ctx, order, compress, paf):
    rc, ta, tb, _ = _oracle.generate(order, compress, paf)
    a, b, _ = gpu_ctx.generate(order, compress, paf)
    sa = gpu_ctx.sort_candidates(a)
    sb = gpu_ctx.sort_candidates(b)
    ref_sa = sorted(set(ta.splitlines()))
    ref_sb = sorted(set(tb.splitlines()))
    assert sa.lines() == ref_sa
    assert sb.lines() == ref_sb
    rc, pa, pb, st = gpu_ctx.match(order, compress, paf, sa, sb)
    got = _text(g.pair_line(x, y) for x, y in zip(pa, pb))
    ref = _oracle_match(order, compress, paf, _text(ref_sa), _text(ref_sb))
    assert got == ref
    assert st.pairs == len(pa)


def _oracle_match(order, compress, paf, ta, tb):
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        wd = os.path.join(d, "order-%d" % order)
        os.makedirs(wd)
        open(os.path.join(wd, "%d-filtered_sorted-a" % order), "w").write(ta)
        open(os.path.join(wd, "%d-filtered_sorted-b" % order), "w").write(tb)
        subprocess.run([_oracle.CLI, str(order), "-c", str(compress), "--paf", str(paf), "-d", d, "match"],
                       check=True, capture_output=True)
        return open(os.path.join(wd, "%d-pairs-%d" % (order, compress))).read()


def test_match_empty_input_returns_1(gpu_ctx):
    a, b, _ = gpu_ctx.generate(16, 2, 0)
    empty = g.Cands(np.zeros((0, 8), np.int8), np.zeros((0, 4), np.uint8))
    rc, pa, pb, st = gpu_ctx.match(16, 2, 0, empty, b)
    assert rc == g.GP_NO_INPUT and len(pa) == 0


def test_psd_keys_match_oracle_lines(gpu_ctx):
    rc, ta, tb, _ = _oracle.generate(26, 2, 0)
    for text, inv in ((ta, False), (tb, True)):
        lines = text.splitlines()
        words = np.array([[int(t) for t in l.split()[1:]] for l in lines], np.int8)
        keys = gpu_ctx.psd_keys(words, 52, inv)
        assert [g.cand_line(w, k) for w, k in zip(words, keys)] == lines


def _golden(name):
    p = os.path.join(GOLDEN, name)
    if not os.path.exists(p):
        pytest.skip("golden fixture %s not generated" % name)
    return json.load(open(p))


@pytest.mark.parametrize("fixture", ["gen_v45_c3_paf-2.json", "gen_v90_c10.json"])
def test_gen_large_matches_golden_digest(gpu_ctx, fixture):
    fx = _golden(fixture)
    a, b, st = gpu_ctx.generate(fx["order"], fx["compress"], fx["paf"])
    for side, c in (("a", a), ("b", b)):
        text = _text(c.lines())
        assert len(c) == fx[side]["lines"]
        assert hashlib.sha256(text.encode()).hexdigest() == fx[side]["sha256"]
    assert st.tie_band == 0
