"""GPU uncompression vs the oracle (uncompression.cpp:32-210).

Per sequence: the candidate lines in DFS order (the reference's
uncompress_tmp_{a,b} content) must be byte-identical; per pair file: the pairs
at the finer compression must be byte-identical, in order.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golaypair_amd as g
from tests import _oracle

pytestmark = pytest.mark.gpu


def _oracle_dir(order, compress, paf, stages):
    d = tempfile.mkdtemp()
    subprocess.run([_oracle.CLI, str(order), "-c"] + [str(c) for c in compress] +
                   ["--paf", str(paf), "-d", d] + stages, check=True, capture_output=True)
    return os.path.join(d, "order-%d" % order)


def _read_pairs(path, L):
    rows = [list(map(int, l.split())) for l in open(path).read().splitlines() if l.strip()]
    a = np.array([r[:L] for r in rows], np.int8).reshape(-1, L)
    b = np.array([r[L:] for r in rows], np.int8).reshape(-1, L)
    return a, b


CASES = [  # order, compress pair, paf: r = 2, 3, 5 and odd/even finer alphabets
    (16, [2, 1], 0),
    (26, [2, 1], 0),
    (21, [3, 1], -2),
    (27, [3, 1], -2),
    (20, [5, 1], 0),
    (16, [4, 2], 0),
    (20, [10, 5], 0),
]


@pytest.mark.parametrize("order,compress,paf", CASES)
def test_uncompress_side_matches_oracle(gpu_ctx, order, compress, paf):
    wd = _oracle_dir(order, compress, paf, ["gen", "sort", "match"])
    c, c2 = compress
    L = order // c
    pa, pb = _read_pairs(os.path.join(wd, "%d-pairs-%d" % (order, c)), L)
    assert len(pa) > 0
    for q in range(min(len(pa), 6)):
        for seqflag, seq in ((0, pa[q]), (1, pb[q])):
            ref_text, ref_leaves = _oracle.uncompress_side(seq, c, c2, paf, seqflag)
            cands, rot, st = gpu_ctx.uncompress_side(seq, c, c2, paf, seqflag)
            got = "".join(l + "\n" for l in cands.lines())
            assert got == ref_text, (q, seqflag)
            assert st.psd_decisions == ref_leaves


@pytest.mark.parametrize("mode", [g.GP_PSD_SPLIT_BF16, g.GP_PSD_FP32])
@pytest.mark.parametrize("order,compress,paf", CASES)
def test_uncompress_pipeline_matches_oracle(gpu_ctx, order, compress, paf, mode):
    wd = _oracle_dir(order, compress, paf, ["gen", "sort", "match", "uncompress"])
    c, c2 = compress
    pa, pb = _read_pairs(os.path.join(wd, "%d-pairs-%d" % (order, c)), order // c)
    gpu_ctx.set_mode(mode)
    try:
        oa, ob, ta, tb, st = gpu_ctx.uncompress(order, c, c2, paf, pa, pb, want_tmp=True)
    finally:
        gpu_ctx.set_mode(g.GP_PSD_SPLIT_BF16)
    got = "".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob))
    assert got == open(os.path.join(wd, "%d-pairs-%d" % (order, c2))).read()
    # the last pair's temporary candidate files
    assert "".join(l + "\n" for l in ta.lines()) == open(os.path.join(wd, "uncompress_tmp_a")).read()
    assert "".join(l + "\n" for l in tb.lines()) == open(os.path.join(wd, "uncompress_tmp_b")).read()
    assert st.tie_band == 0


@pytest.mark.parametrize("batch_leaves,emit_cap", [(64, None), (1 << 12, 1), (1 << 16, 3)])
def test_uncompress_batches_and_overflow(gpu_ctx, monkeypatch, batch_leaves, emit_cap):
    """Many small device batches (one input pair each at 64 leaves, so batch
    i + 1 is queued while batch i is collected) and emitted-pair lists too small
    for a batch (which is joined again with exact sizes): same pairs, same order."""
    order, compress, paf = 26, [2, 1], 0
    wd = _oracle_dir(order, compress, paf, ["gen", "sort", "match", "uncompress"])
    c, c2 = compress
    pa, pb = _read_pairs(os.path.join(wd, "%d-pairs-%d" % (order, c)), order // c)
    assert len(pa) > 4
    monkeypatch.setenv("GOLAY_UNC_BATCH_LEAVES", str(batch_leaves))
    if emit_cap is not None:
        monkeypatch.setenv("GOLAY_UNC_EMIT_CAP", str(emit_cap))
    oa, ob, ta, tb, st = gpu_ctx.uncompress(order, c, c2, paf, pa, pb, want_tmp=True)
    got = "".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob))
    assert got == open(os.path.join(wd, "%d-pairs-%d" % (order, c2))).read()
    assert "".join(l + "\n" for l in ta.lines()) == open(os.path.join(wd, "uncompress_tmp_a")).read()
    assert "".join(l + "\n" for l in tb.lines()) == open(os.path.join(wd, "uncompress_tmp_b")).read()
    assert st.tie_band == 0


def test_uncompress_empty_input(gpu_ctx):
    z = np.zeros((0, 8), np.int8)
    oa, ob, ta, tb, st = gpu_ctx.uncompress(16, 2, 1, 0, z, z, want_tmp=True)
    assert len(oa) == 0 and len(ta) == 0


def test_uncompress_compress1_is_noop(gpu_ctx):
    a = np.ones((2, 16), np.int8)
    oa, ob, _, _, _ = gpu_ctx.uncompress(16, 1, 1, 0, a, a)
    assert len(oa) == 0


@pytest.mark.parametrize("order,compress,paf,npairs,max_leaves,with_filter", [
    (90, [10, 5], 0, 3, None, True), (45, [3, 1], -2, 2, None, True)])
def test_uncompress_slice_at_full_shape(gpu_ctx, order, compress, paf, npairs, max_leaves, with_filter):
    """Matched pairs of a full-size level (v=90 c=10 -> 5, filtered at length 18;
    v=45 c=3 -> 1, length 45): GPU gen/sort/match, the
    first `npairs` input pairs that emit pairs at the finer level (found one pair
    at a time on the GPU, at most `max_leaves` leaves each) plus an input pair
    that emits nothing, then uncompress (and the equivalence filter) on the GPU
    and in the oracle over that slice."""
    c, c2 = compress
    a, b, _ = gpu_ctx.generate(order, c, paf)
    rc, pa, pb, _ = gpu_ctx.match(order, c, paf, gpu_ctx.sort_candidates(a), gpu_ctx.sort_candidates(b))
    assert rc == 0
    pick = []
    for i in range(min(len(pa), 400 if max_leaves is None else len(pa))):
        oa, _, _, _, st = gpu_ctx.uncompress(order, c, c2, paf, pa[i:i + 1], pb[i:i + 1])
        if max_leaves is not None and st.psd_decisions > max_leaves:
            continue  # keep the oracle's share of the test to seconds
        if len(oa):
            if not pick and i > 0:
                pick.append(i - 1)  # an input pair that emits nothing
            pick.append(i)
            if len(pick) > npairs:
                break
    assert len(pick) >= npairs
    pa, pb = pa[pick], pb[pick]
    d = tempfile.mkdtemp()
    wd = os.path.join(d, "order-%d" % order)
    os.makedirs(wd)
    open(os.path.join(wd, "%d-pairs-%d" % (order, c)), "w").write(
        "".join(g.pair_line(x, y) + "\n" for x, y in zip(pa, pb)))
    subprocess.run([_oracle.CLI, str(order), "-c", str(c), str(c2), "--paf", str(paf), "-d", d, "uncompress"] +
                   (["filter"] if with_filter else []), check=True, capture_output=True, timeout=600)
    oa, ob, _, _, st = gpu_ctx.uncompress(order, c, c2, paf, pa, pb)
    assert st.tie_band == 0 and len(oa) > 0
    assert "".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob)) == \
        open(os.path.join(wd, "%d-pairs-%d" % (order, c2))).read()
    if not with_filter:
        return
    fa, fb, _ = gpu_ctx.filter(order, c2, oa, ob)
    assert "".join(g.pair_line(x, y) + "\n" for x, y in zip(fa, fb)) == \
        open(os.path.join(wd, "%d-unique" % order)).read()


@pytest.mark.parametrize("seqflag", [0, 1])
def test_uncompress_side_mid_rows_beyond_one_chunk(gpu_ctx, seqflag):
    """A length-35 side expanded 5 -> 1 whose active positions all have 10
    splits (position 0: 2 after the dihedral reduction): suffix 3 positions
    (1000 leaves), mid 2 positions (100 rows), outer position 0.  The filter's
    one-block one-tile class stages mid-row seeds 32 rows at a time, so this side
    runs four chunks (32, 32, 32, 4) per work item.  DFS candidate lines vs the
    oracle."""
    seq = np.array([1, -1, 1, 5, 1, -1, -1], np.int8)
    ref_text, ref_leaves = _oracle.uncompress_side(seq, 5, 1, -2, seqflag)
    cands, rot, st = gpu_ctx.uncompress_side(seq, 5, 1, -2, seqflag)
    assert ref_leaves == 200000 and st.psd_decisions == ref_leaves
    assert len(ref_text) > 0
    assert "".join(l + "\n" for l in cands.lines()) == ref_text


@pytest.mark.parametrize("seq", [
    [-3, -1, 1, -3, -1, -3, -3, -1, 1, 1, -1, -3, 1, -3, 1, -1, 1],              # N = 51: 17 channels
    [-3, -1, 1, 1, -3, -1, -3, -3, 3, 1, -3, -1, -1, -1, -3, -1, -1, 1],         # N = 54: 18
    [3, -1, -1, -1, -3, -1, -3, 1, -3, -3, -1, -3, -1, 1, -3, -3, -1, 1, -1],    # N = 57: 19
    [-1, -1, 1, -1, -1, -3, -3, -3, 1, -3, -3, -3, -3, -1, 1, -1, -3, 1, -3, 1],  # N = 60: 20
])
@pytest.mark.parametrize("seqflag", [0, 1])
def test_uncompress_side_small_last_tile(gpu_ctx, seq, seqflag):
    """Sides expanded 3 -> 1 to lengths 51, 54, 57 and 60: 17 to 20 filter
    channels (the k = 3m are fixed), so the second M-tile holds 1 to 4 of them
    and runs on the 16x16x32 MFMAs of the small-tile class, padding slots
    included (the v=50 level has 4).  Words of 11 letters +-1 (3 splits each)
    and +-3 elsewhere, whose fixed channels pass the bound (found by a seeded
    search); DFS candidate lines vs the oracle."""
    seq = np.array(seq, np.int8)
    ref_text, ref_leaves = _oracle.uncompress_side(seq, 3, 1, -2, seqflag)
    cands, rot, st = gpu_ctx.uncompress_side(seq, 3, 1, -2, seqflag)
    assert ref_leaves > 0 and st.psd_decisions == ref_leaves
    assert len(ref_text) > 0
    assert "".join(l + "\n" for l in cands.lines()) == ref_text


def test_uncompress_side_v50_full_shape(gpu_ctx):
    """Single sequences of the v=50 5-compression level expanded 5 -> 1 (length 50:
    two M-tiles, fixed channels k = 5m dropped, single-split positions folded):
    gen candidates with the fewest leaves, DFS candidate lines vs the oracle."""
    split = {1: 10, -1: 10, 3: 5, -3: 5, 5: 1, -5: 1}
    a, b, _ = gpu_ctx.generate(50, 5, 0)
    picked = []
    for cands, seqflag in ((a, 0), (b, 1)):
        est = [(int(np.prod([split[int(x)] for x in w[1:]])), i) for i, w in enumerate(cands.words)]
        picked += [(cands.words[i], seqflag) for _, i in sorted(est)[:1]]  # about 2.5e6 leaves, ~30 s in the oracle
    total = 0
    for seq, seqflag in picked:
        ref_text, ref_leaves = _oracle.uncompress_side(seq, 5, 1, 0, seqflag)
        cands, rot, st = gpu_ctx.uncompress_side(seq, 5, 1, 0, seqflag)
        assert "".join(l + "\n" for l in cands.lines()) == ref_text, (list(seq), seqflag)
        assert st.psd_decisions == ref_leaves
        total += len(cands)
    assert total > 0  # some expansions pass the filter


def test_uncompress_v50_pair_slice_vs_oracle_fixture(gpu_ctx):
    """Config 4's own level (v=50, 5 -> 1, length 50, two M-tiles): the matched c=5
    pairs with the fewest expansion leaves (about 10^8 leaves each, written with the
    oracle's output by tests/golden/make_golden.py slice50).  The GPU's 50-pairs-1
    lines for those input pairs must be the oracle's, in order; the input lines must
    be the GPU's own match output at those positions; and the equivalence filter at
    length 50 with altnegative must keep the brute-force survivors."""
    import json
    from tests.test_filter_gpu import brute_unique
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "slice_v50_c5_1.json")))
    order, (c, c2), paf = fx["order"], fx["compress"], fx["paf"]
    L = order // c
    rows = [list(map(int, l.split())) for l in fx["pairs_in"].splitlines()]
    pa = np.array([r[:L] for r in rows], np.int8)
    pb = np.array([r[L:] for r in rows], np.int8)
    a, b, _ = gpu_ctx.generate(order, c, paf)
    rc, ma, mb, _ = gpu_ctx.match(order, c, paf, gpu_ctx.sort_candidates(a), gpu_ctx.sort_candidates(b))
    assert rc == 0
    for k, i in enumerate(fx["input_lines"]):
        assert g.pair_line(ma[i], mb[i]) == fx["pairs_in"].splitlines()[k]
    oa, ob, _, _, st = gpu_ctx.uncompress(order, c, c2, paf, pa, pb)
    assert st.tie_band == 0
    assert st.psd_decisions == sum(fx["leaves"])
    assert "".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob)) == fx["pairs_out"]
    assert len(oa) > 0
    sel = slice(0, 40)  # the brute force takes about 0.2 s per pair
    fa, fb, _ = gpu_ctx.filter(order, c2, oa[sel], ob[sel])
    got = [(tuple(int(v) for v in x), tuple(int(v) for v in y)) for x, y in zip(fa, fb)]
    assert got == brute_unique(oa[sel], ob[sel], True)


@pytest.mark.parametrize("seq,c,c2,paf,njobs", [
    # paf lifts the filter bound 2v - paf to 255 where the side's own bound passes nothing
    ((1, -1, 1, 5, 1, -1, -1), 5, 1, -2, 3),             # 2*10^5 leaves, split at depth 5 of 7
    ((3, -1, 1, -3, 1, 1, -1, 1, -1, 1), 3, 1, -195, 2),  # split at depth 8 of 10
    ((-4, 6, -4, -2, 2, 0), 10, 5, -135, 4),             # 3600 leaves < 4000: every position set
    ((1, 1, -3, 3, -1), 5, 1, 0, 7),                     # 5000 leaves < 7000: every position set
])
def test_uncompress_side_job_split_matches_oracle(gpu_ctx, seq, c, c2, paf, njobs):
    """The uncompress_gpu seam's PROC_ID / PROC_NUM (uncompress_gpu.h:8,
    uncompression.cpp:167-191): each job's candidate lines in DFS order are the
    oracle's, and the jobs together give the single-job lines."""
    seq = np.array(seq, np.int8)
    whole, _ = _oracle.uncompress_side(seq, c, c2, paf, 1)
    parts = []
    for j in range(njobs):
        for seqflag in (0, 1):
            ref_text, _ = _oracle.uncompress_side(seq, c, c2, paf, seqflag, j, njobs)
            cands, rot, st = gpu_ctx.uncompress_side(seq, c, c2, paf, seqflag, j, njobs)
            got = "".join(l + "\n" for l in cands.lines())
            assert got == ref_text, (j, seqflag)
            if seqflag:
                parts += got.splitlines()
    assert sorted(parts) == sorted(whole.splitlines())
    assert len(whole) > 0
    cands, _, _ = gpu_ctx.uncompress_side(seq, c, c2, paf, 1, njobs, njobs)   # no job of that id
    assert len(cands) == 0


@pytest.mark.parametrize("batch_leaves", [64, 1 << 14])
def test_uncompress_batch_graph_matches_eager(gpu_ctx, monkeypatch, batch_leaves):
    """Each batch's band recheck and re-match chain runs as a captured hipGraph that
    updates its slot's instantiated graph; GOLAY_UNC_GRAPH=0 launches the same
    kernels directly.  Small batches give many graph updates and launches, with the
    batch's eager work (uploads, filter launches, read-backs) and synchronizations
    between them: every run must give the oracle's pairs and temporary files."""
    order, compress, paf = 26, [2, 1], 0
    wd = _oracle_dir(order, compress, paf, ["gen", "sort", "match", "uncompress"])
    c, c2 = compress
    pa, pb = _read_pairs(os.path.join(wd, "%d-pairs-%d" % (order, c)), order // c)
    want = open(os.path.join(wd, "%d-pairs-%d" % (order, c2))).read()
    monkeypatch.setenv("GOLAY_UNC_BATCH_LEAVES", str(batch_leaves))
    for mode in ("1", "0", "1", "1"):
        monkeypatch.setenv("GOLAY_UNC_GRAPH", mode)
        oa, ob, ta, tb, st = gpu_ctx.uncompress(order, c, c2, paf, pa, pb, want_tmp=True)
        assert "".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob)) == want, mode
        assert "".join(l + "\n" for l in ta.lines()) == open(os.path.join(wd, "uncompress_tmp_a")).read()
        assert st.tie_band == 0
