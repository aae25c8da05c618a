"""Pins for the oracle (oracle/refcpu*.cpp), the CPU restatement of the reference.

Each check is independent of the restatement it tests:
  alphabet / decomps / coprimes   closed forms
  construct_generators            equals the affine group {i -> u i + t} (AGL or dihedral)
  is_canonical                    brute-force orbit minimum in numpy
  necklace enumeration            Burnside count; brute force over all words
  gen job split (-j/-n)           prefix rank mod jobs at the reference's prefix length,
                                  prenecklaces by a brute-force predicate
  PSD / filter / key text         numpy FFT; Parseval; Wiener-Khinchin with exact PAF
  check_if_pair                   numpy PAF
  gen                             brute force over all words (numpy FFT, orbit min)
  match                           nested-loop join on key text
  sort                            coreutils `sort | uniq` under LC_ALL=C on the gen files
  pair group (filter)             closure == parametric family of signed affine maps
  cache_filter                    brute-force orbits, lexicographically first input member
  uncompress                      every leaf in DFS order vs an itertools enumeration (best
                                  letter, first-position dihedral representatives with a
                                  Burnside count, split order, index layout, leaf count);
                                  the PROC_ID / PROC_NUM job split; compress(expanded) ==
                                  rotated original
  whole pipeline                  NOTES.md known answer v=34 -> 373 classes; NOTES.md v=26
                                  and v=72 fixtures; unique count independent of the
                                  compression route (v=16, 20, 26)
Parity unpinned (code-following only): the rotation amount of the in-place rotation loop
(uncompression.cpp:127-131), i.e. which occurrence of the best letter ends up first; FFTW's
last-ulp PSD values (a direct FP64 DFT stands in; the tie-band counter reports any value
within 1e-9 of a rounding or filter threshold).
"""
import itertools
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests import _oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def alphabet(c):
    return sorted({s * i for i in range(c % 2, c + 1, 2) for s in (1, -1)})


def units(n):
    return [u for u in range(1, n) if math.gcd(u, n) == 1] if n > 1 else [0]


def np_psd(seq):
    return np.abs(np.fft.fft(np.asarray(seq, dtype=np.float64))) ** 2


def orbit_min(seq, us):
    n = len(seq)
    best = None
    for u in us:
        for t in range(n):
            img = tuple(seq[(u * i + t) % n] for i in range(n))
            if best is None or img < best:
                best = img
    return best


def paf(seq, s):
    seq = np.asarray(seq)
    return int(np.dot(seq, np.roll(seq, -s)))


# ---------------------------------------------------------------- small math

def test_generator_group_is_affine(oracle):
    for n in range(3, 16):
        for flag in (0, 1):
            perms = _oracle.group_perms(flag, n)
            us = units(n) if flag == 0 else sorted({1, n - 1})
            expect = {tuple((u * i + t) % n + 1 for i in range(n)) for u in us for t in range(n)}
            assert {tuple(p) for p in perms} == expect, (n, flag)


def test_is_canonical_matches_orbit_min(oracle):
    rng = np.random.default_rng(7)
    for n in (5, 8, 9, 10, 12):
        for _ in range(200):
            seq = tuple(int(x) for x in rng.integers(-2, 3, size=n))
            for flag, us in ((0, units(n)), (1, sorted({1, n - 1}))):
                assert _oracle.is_canonical(seq, flag) == (orbit_min(seq, us) == seq)


def test_psd_against_numpy_and_identities(oracle):
    rng = np.random.default_rng(3)
    for n in (2, 3, 8, 9, 15, 17, 45, 50, 64):
        for _ in range(20):
            seq = rng.integers(-10, 11, size=n)
            p = _oracle.psd(seq)
            assert np.allclose(p, np_psd(seq), rtol=0, atol=1e-9 * max(1, np.abs(seq).sum() ** 2))
            assert p[0] == pytest.approx(seq.sum() ** 2)
            assert p.sum() == pytest.approx(n * (seq ** 2).sum())                         # Parseval
            pafs = np.array([paf(seq, s) for s in range(n)])
            wk = np.array([sum(pafs[s] * math.cos(2 * math.pi * k * s / n) for s in range(n)) for k in range(n)])
            assert np.allclose(p, wk, atol=1e-7)                                          # Wiener-Khinchin
            assert all(p[k] == p[n - k] for k in range(1, n))                              # bit-exact symmetry


def test_check_if_pair_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        n = int(rng.integers(2, 12))
        a = rng.choice([-1, 1], size=n)
        b = rng.choice([-1, 1], size=n)
        lam = int(rng.choice([0, -2]))
        c = int(rng.choice([1, 3]))
        expect = all(paf(a, s) + paf(b, s) == lam * c for s in range(1, n // 2 + 1))
        assert _oracle.check_if_pair(a, b, lam, c) == expect


# ---------------------------------------------------------------- gen

def brute_gen(order, c, lam):
    """All words of length order/c with row sum x (A) / y (B), numpy-FFT filter, orbit min."""
    n = order // c
    alpha = alphabet(c)
    S = 2 * order + (order - 1) * lam
    decomps = [(x, y) for x in range(S + 1) for y in range(x, S + 1) if x * x + y * y == S]
    bound = 2 * order - lam
    A, B = [], []
    for w in itertools.product(alpha, repeat=n):   # lexicographic = FKM order of necklaces
        rs = sum(w)
        for x, y in decomps:
            for target, us, out, inv in ((x, units(n), A, False), (y, sorted({1, n - 1}), B, True)):
                if rs != target:
                    continue
                p = np_psd(w)
                if any(p[k] > bound + 0.001 for k in range(1, n // 2 + 1)):
                    continue
                if orbit_min(w, us) != w:
                    continue
                key = "".join(str(bound - int(np.rint(p[k])) if inv else int(np.rint(p[k]))) for k in range(1, n))
                out.append("PSD" + key + " " + "".join("%d " % v for v in w))
    return "".join(l + "\n" for l in A), "".join(l + "\n" for l in B)


@pytest.mark.parametrize("order,c,lam", [(8, 1, 0), (10, 1, 0), (16, 2, 0), (18, 2, 0), (15, 3, -2), (21, 3, -2),
                                         (20, 2, 0)])
def test_gen_matches_brute_force(oracle, order, c, lam):
    rc, ta, tb, _ = _oracle.generate(order, c, lam)
    ba, bb = brute_gen(order, c, lam)
    assert rc == 0
    assert ta == ba and tb == bb


def test_necklace_count_and_job_split_partition(oracle, tmp_path):
    # jobs of a split cover the single-job output exactly once (lyndon.cpp:87-121)
    rc, ta, tb, orderly = _oracle.generate(26, 2, 0)
    n, k = 13, 3
    assert orderly == sum(k ** math.gcd(i, n) for i in range(1, n + 1)) // n            # Burnside
    parts_a, parts_b = [], []
    for j in range(3):
        _, a, b, _ = _oracle.generate(26, 2, 0, j, 3)
        parts_a += a.splitlines()
        parts_b += b.splitlines()
    assert sorted(parts_a) == sorted(ta.splitlines()) and sorted(parts_b) == sorted(tb.splitlines())


# ---------------------------------------------------------------- sort + match

def _cli(args, d):
    return subprocess.run([_oracle.CLI] + [str(a) for a in args] + ["-d", d], capture_output=True, text=True)


def test_match_is_nested_loop_join(oracle):
    with tempfile.TemporaryDirectory() as d:
        assert _cli([26, "-c", 2, "gen", "sort", "match"], d).returncode == 0
        wd = os.path.join(d, "order-26")
        la = open(os.path.join(wd, "26-filtered_sorted-a")).read().splitlines()
        lb = open(os.path.join(wd, "26-filtered_sorted-b")).read().splitlines()
        assert la == sorted(set(la)) and lb == sorted(set(lb))    # C-locale order + uniq
        out = []
        for x in la:                                              # A outer, B inner, file order
            for y in lb:
                if x.split()[0] != y.split()[0]:
                    continue
                a = [int(t) for t in x.split()[1:]]
                b = [int(t) for t in y.split()[1:]]
                if all(paf(a, s) + paf(b, s) == 0 for s in range(1, len(a) // 2 + 1)):
                    out.append("".join("%d " % v for v in a) + " " + "".join("%d " % v for v in b))
        assert open(os.path.join(wd, "26-pairs-2")).read() == "".join(l + "\n" for l in out)


def test_match_empty_file_fails(oracle):
    with tempfile.TemporaryDirectory() as d:
        wd = os.path.join(d, "order-16")
        os.makedirs(wd)
        open(os.path.join(wd, "16-filtered_sorted-a"), "w").write("")
        open(os.path.join(wd, "16-filtered_sorted-b"), "w").write("PSD1 0 0 0 0 0 0 0 0 \n")
        assert _cli([16, "-c", 2, "match"], d).returncode == 1            # match_pairs.cpp:32-38


# ---------------------------------------------------------------- filter group

def param_group(n, alt):
    """Signed affine maps (swap, u, e, kappa, sa, sb, ta, tb) as functions on pairs."""
    maps = []
    for sw in (0, 1):
        for u in units(n):
            for e in (1, -1):
                for kappa in ((0, 1) if alt else (0,)):
                    for sa in ((1, -1) if alt else (1,)):
                        for sb in ((1, -1) if alt else (1,)):
                            for ta in range(n):
                                for tb in range(n):
                                    maps.append((sw, u, e, kappa, sa, sb, ta, tb))
    return maps


def apply(m, a, b):
    sw, u, e, kappa, sa, sb, ta, tb = m
    n = len(a)
    X, Y = (b, a) if sw else (a, b)
    na = tuple(sa * (-1 if kappa and i % 2 else 1) * X[(u * i + ta) % n] for i in range(n))
    nb = tuple(sb * (-1 if kappa and i % 2 else 1) * Y[(e * u * i + tb) % n] for i in range(n))
    return na, nb


@pytest.mark.parametrize("n,alt", [(3, 0), (4, 0), (4, 1), (5, 0), (6, 1), (8, 1), (9, 0), (10, 0), (10, 1)])
def test_pair_group_is_signed_affine_family(oracle, n, alt):
    # generic pair: distinct nonzero magnitudes identify the transform from its image
    a = tuple(2 + i for i in range(n))
    b = tuple(2 + n + i for i in range(n))
    cls = {(tuple(r[:n]), tuple(r[n:])) for r in _oracle.class_pairs(n, alt, a, b)}
    fam = {apply(m, a, b) for m in param_group(n, alt)}
    assert cls == fam
    assert len(_oracle.pair_group(n, alt)) == len(fam)


def brute_filter(pairs, n, alt):
    maps = param_group(n, alt)
    seen, keep = set(), []
    for p in sorted(set(pairs)):
        if p in seen:
            continue
        keep.append(p)
        for m in maps:
            seen.add(apply(m, *p))
    return keep


@pytest.mark.parametrize("n,alt,compress", [(6, 1, 1), (7, 0, 1), (8, 1, 1), (5, 0, 3), (6, 0, 2)])
def test_cache_filter_keeps_first_of_each_orbit(oracle, n, alt, compress):
    rng = np.random.default_rng(n)
    base = [(tuple(rng.choice([-1, 1], n)), tuple(rng.choice([-1, 1], n))) for _ in range(6)]
    maps = param_group(n, alt)
    pairs = []
    for p in base:                                          # several members of a few orbits
        for _ in range(4):
            pairs.append(apply(maps[int(rng.integers(len(maps)))], *p))
    text = "".join("".join("%d " % v for v in a) + " " + "".join("%d " % v for v in b) + "\n" for a, b in pairs)
    with tempfile.TemporaryDirectory() as d:
        order = n * compress
        wd = os.path.join(d, "order-%d" % order)
        os.makedirs(wd)
        open(os.path.join(wd, "%d-pairs-%d" % (order, compress)), "w").write(text)
        assert _cli([order, "-c", compress, "filter"], d).returncode == 0
        got = open(os.path.join(wd, "%d-unique" % order)).read()
    want = "".join("".join("%d " % v for v in a) + " " + "".join("%d " % v for v in b) + "\n"
                   for a, b in brute_filter(pairs, n, alt))
    assert got == want


# ---------------------------------------------------------------- uncompress

@pytest.mark.parametrize("seq,c,c2", [((-2, 0, -2, 0, 2, 0, 0, 0, 0, 2, 2, 0, 2), 2, 1),
                                      ((3, -1, 1, -3, 1, 1, -1), 3, 1),
                                      ((5, -3, 1, 1, -1, 3), 5, 1),
                                      ((4, 0, -2, 2, 0, 6), 6, 2)])
def test_uncompress_expansions_compress_back(oracle, seq, c, c2):
    text, leaves = _oracle.uncompress_side(seq, c, c2, 0, 1)
    L, r = len(seq), c // c2
    rotations = {tuple(seq[i:] + seq[:i]) for i in range(L)}
    for line in text.splitlines():
        w = [int(t) for t in line.split()[1:]]
        comp = tuple(sum(w[j + L * i] for i in range(r)) for j in range(L))
        assert comp in rotations
        assert set(w) <= set(alphabet(c2))
    assert leaves > 0


def split_list(letter, r, fine):
    """Every r-tuple over the finer alphabet summing to `letter`: the nondecreasing
    tuples in lexicographic order, each followed by its distinct permutations in
    lexicographic order (the order in which the expansions are visited)."""
    out = []
    for comb in itertools.combinations_with_replacement(sorted(fine), r):
        if sum(comb) == letter:
            out += sorted(set(itertools.permutations(comb)))
    return out


def dihedral_orbits(tuples):
    """Orbits of the dihedral group D_r (rotations, reversal) on a set of r-tuples."""
    seen, orbits = set(), []
    for t in tuples:
        if t in seen:
            continue
        r = len(t)
        orb = {tt for s in range(r) for tt in (t[s:] + t[:s], tuple(reversed(t[s:] + t[:s])))}
        seen |= orb
        orbits.append(orb)
    return orbits


def burnside_dihedral(tuples):
    """Burnside count of the same orbits: the mean number of tuples fixed per group element."""
    r = len(tuples[0])
    fixed = 0
    for s in range(r):
        for rev in (False, True):
            for t in tuples:
                img = t[s:] + t[:s]
                fixed += (tuple(reversed(img)) if rev else img) == t
    assert fixed % (2 * r) == 0
    return fixed // (2 * r)


@pytest.mark.parametrize("seq,c,c2", [((3, -1, 1, -3, 1, 1, -1), 3, 1),
                                      ((1, -1, 3, 1, -1, 1, -3, 3), 3, 1),
                                      ((-1, 5, -3, 1, 1, -1), 5, 1),
                                      ((1, 1, -3, 3, -1), 5, 1),
                                      ((-2, 0, 2, 0, 0, 2, -2), 2, 1),
                                      ((4, 0, -2, 2, 0, 6), 6, 2),
                                      ((10, -2, 0, 4, -6, 2), 10, 5),
                                      ((-2, 6, 0, 0, 2, -8, 4), 10, 5)])
def test_uncompress_every_leaf_in_order(oracle, seq, c, c2):
    """With a filter bound no word reaches (paf = -10^6) every leaf is written, so the
    oracle's lines are the whole DFS (uncompression.cpp:81-210).  Checked against an
    independent enumeration:
      * the word is rotated so that it starts with the best letter, the smallest
        letter with the most splits (:113-131), and it stays a rotation of the input;
      * position 0 runs over one representative per dihedral orbit of its splits,
        the lexicographically smallest, in increasing order (:133-142); the orbit
        count equals Burnside's;
      * every other position j runs over all splits of its letter, nondecreasing
        tuples in lexicographic order, each with its permutations in lexicographic
        order (:91-111); position 0 outermost;
      * tuple element i lands at index j + L*i (:153,160);
      * the leaf count is (orbits at position 0) x prod(split counts).
    The rotation amount (which occurrence of the best letter ends up first) is the
    one rule here that only follows the code."""
    fine = alphabet(c2)
    r, L = c // c2, len(seq)
    text, leaves = _oracle.uncompress_side(seq, c, c2, -10 ** 6, 1)
    words = [tuple(int(t) for t in line.split()[1:]) for line in text.splitlines()]
    assert len(words) == leaves
    rot = tuple(sum(words[0][j + L * i] for i in range(r)) for j in range(L))
    assert rot in {tuple(seq[i:] + seq[:i]) for i in range(L)}
    counts = {x: len(split_list(x, r, fine)) for x in set(seq)}
    best = min(x for x in counts if counts[x] == max(counts.values()))
    assert rot[0] == best
    reps = sorted(min(o) for o in dihedral_orbits(split_list(best, r, fine)))
    assert len(reps) == burnside_dihedral(split_list(best, r, fine))
    opts = [reps] + [split_list(rot[j], r, fine) for j in range(1, L)]
    want = []
    for choice in itertools.product(*opts):
        w = [0] * (L * r)
        for j, t in enumerate(choice):
            for i in range(r):
                w[j + L * i] = t[i]
        want.append(tuple(w))
    assert words == want
    assert leaves == len(reps) * math.prod(len(o) for o in opts[1:])


@pytest.mark.parametrize("seq,c,c2,njobs", [((1, -1, 1, 5, 1, -1, -1), 5, 1, 3),
                                            ((3, -1, 1, -3, 1, 1, -1, 1, -1), 3, 1, 2),
                                            ((10, -2, 0, 4, -6, 2), 10, 5, 4),
                                            ((1, 1, -3, 3, -1), 5, 1, 7)])
def test_uncompress_job_split(oracle, seq, c, c2, njobs):
    """uncompression.cpp:167-191: with PROC_NUM > 1 the partial words are expanded
    position by position until there are >= 1000 * PROC_NUM of them (or all
    positions are set), and job j keeps the partials of index j mod PROC_NUM.  With
    every leaf written (paf = -10^6): the jobs partition the single-job lines, each
    job's lines keep the DFS order, and job j holds exactly the leaves whose prefix
    index at that depth is j mod PROC_NUM.  The depth is read off the radices."""
    fine = alphabet(c2)
    r, L = c // c2, len(seq)
    text, leaves = _oracle.uncompress_side(seq, c, c2, -10 ** 6, 1)
    whole = text.splitlines()
    assert len(whole) == leaves
    words = [tuple(int(t) for t in line.split()[1:]) for line in whole]
    rot = tuple(sum(words[0][j + L * i] for i in range(r)) for j in range(L))
    counts = {x: len(split_list(x, r, fine)) for x in set(seq)}
    radix = [len(dihedral_orbits(split_list(rot[0], r, fine)))] + [counts[rot[j]] for j in range(1, L)]
    depth, prod = 0, 1
    while depth < L and prod < 1000 * njobs:
        prod *= radix[depth]
        depth += 1
    suffix = math.prod(radix[depth:])
    seen = []
    for j in range(njobs):
        t, lv = _oracle.uncompress_side(seq, c, c2, -10 ** 6, 1, j, njobs)
        mine = t.splitlines()
        assert lv == len(mine)
        assert mine == [l for i, l in enumerate(whole) if (i // suffix) % njobs == j]
        seen += mine
    assert sorted(seen) == sorted(whole)
    # the pipeline's own calls are single-job (uncompression.cpp:65-66)
    assert _oracle.uncompress_side(seq, c, c2, 0, 0, 0, 1) == _oracle.uncompress_side(seq, c, c2, 0, 0)


def is_prenecklace(w):
    """Each suffix is >= the prefix of the same length (a prefix of some necklace)."""
    return all(w[j:] >= w[:len(w) - j] for j in range(1, len(w)))


def test_prenecklace_predicate_is_prefix_of_necklace():
    # the predicate against its definition on a small alphabet: the length-i prefixes of
    # the necklaces of lengths i..2i
    for k, i in ((2, 5), (3, 4)):
        neck = set()
        for m in range(i, 2 * i + 1):
            for w in itertools.product(range(k), repeat=m):
                if all(w[s:] + w[:s] >= w for s in range(m)):
                    neck.add(w[:i])
        assert neck == {w for w in itertools.product(range(k), repeat=i) if is_prenecklace(w)}


@pytest.mark.parametrize("order,c,lam,njobs", [(40, 5, 0, 2), (40, 5, 0, 3)])
def test_job_split_assigns_prefix_rank_mod_jobs(oracle, order, c, lam, njobs):
    """lyndon.cpp:87-121: the prefix length i is the least with necklace_count(i, k) >=
    10^4 * jobs; the prenecklaces of length i are counted in lexicographic order from
    1 and job j keeps those whose count % jobs == j.  Every line of job j must start
    with a prefix of that rank."""
    n, alpha = order // c, alphabet(c)
    k = len(alpha)
    i = next(m for m in range(1, n + 1)
             if sum(k ** math.gcd(t, m) for t in range(1, m + 1)) // m >= 10 ** 4 * njobs)
    assert i < n
    rank, cnt = {}, 0
    for w in itertools.product(alpha, repeat=i):     # lexicographic order
        if is_prenecklace(w):
            cnt += 1
            rank[w] = cnt
    total = 0
    for j in range(njobs):
        _, ta, tb, _ = _oracle.generate(order, c, lam, j, njobs)
        for line in (ta + tb).splitlines():
            w = tuple(int(t) for t in line.split()[1:])
            assert rank[w[:i]] % njobs == j
            total += 1
    _, ta, tb, _ = _oracle.generate(order, c, lam)
    assert total == len((ta + tb).splitlines()) > 0


def test_job_split_without_enough_prefixes_repeats_the_whole_set(oracle):
    """No prefix length reaches 10^4 * jobs necklaces below the word length, so the
    wrapper takes the single-job branch in every job (lyndon.cpp:95,107-109)."""
    whole = _oracle.generate(16, 2, 0)[1:3]
    for j in range(2):
        assert _oracle.generate(16, 2, 0, j, 2)[1:3] == whole


# ---------------------------------------------------------------- known answers

def test_notes_fixtures(oracle):
    # NOTES.md:455,459 -- the v=26 A-candidate and its 2-compression
    a26 = [-1, -1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, 1, -1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, 1, 1]
    comp = [a26[j] + a26[j + 13] for j in range(13)]
    assert comp == [-2, 0, -2, 0, 2, 0, 0, 0, 0, 2, 2, 0, 2]
    rc, ta, tb, _ = _oracle.generate(26, 2, 0)
    canon = orbit_min(tuple(comp), units(13))
    assert "".join("%d " % v for v in canon) in {l.split(" ", 1)[1] for l in ta.splitlines()}
    # NOTES.md:436 -- the v=72 6-compressed pair: PAF-complementary and key-matched
    a = [-6, 0, 0, 0, 0, 0, 6, 0, 0, 0, 0, 0]
    b = [0, 0, 0, 0, 0, 6, 0, 0, 0, 0, 0, 6]
    assert _oracle.check_if_pair(a, b, 0, 6)
    pa, pb = _oracle.psd(a), _oracle.psd(b)
    assert all(pa[k] <= 144.001 and pb[k] <= 144.001 for k in range(1, 7))
    assert [int(np.rint(pa[k])) for k in range(1, 12)] == [144 - int(np.rint(pb[k])) for k in range(1, 12)]


def test_c1_results_dir_matches_golden(oracle):
    with tempfile.TemporaryDirectory() as d:
        assert _cli([16, "--compress", 2, 1, "all"], d).returncode == 0
        wd = os.path.join(d, "order-16")
        gd = os.path.join(GOLDEN, "v16_c2_1")
        assert sorted(os.listdir(wd)) == sorted(os.listdir(gd))
        for f in os.listdir(gd):
            assert open(os.path.join(wd, f)).read() == open(os.path.join(gd, f)).read(), f


@pytest.mark.parametrize("order,routes", [(16, [[2, 1], [4, 2, 1]]), (20, [[2, 1], [5, 1], [10, 5, 1]]),
                                          ])
def test_unique_count_independent_of_compression_route(oracle, order, routes):
    outs = []
    for r in routes:
        with tempfile.TemporaryDirectory() as d:
            assert _cli([order, "-c"] + r + ["all"], d).returncode == 0
            outs.append(open(os.path.join(d, "order-%d" % order, "%d-unique" % order)).read().count("\n"))
    assert len(set(outs)) == 1, outs


@pytest.mark.skipif(not os.environ.get("GOLAY_SLOW"), reason="~8 min on one core; set GOLAY_SLOW=1")
def test_v34_filter_known_answer(oracle):
    """NOTES.md:574-575: 373 classes at v=34 (also the PGP 2022 count).  The
    golden 34-unique was written by this same oracle run (tests/golden/v34)."""
    with tempfile.TemporaryDirectory() as d:
        wd = os.path.join(d, "order-34")
        os.makedirs(wd)
        src = open(os.path.join(GOLDEN, "v34", "34-pairs-1")).read()
        open(os.path.join(wd, "34-pairs-1"), "w").write(src)
        r = _cli([34, "-c", 1, "filter"], d)
        assert r.returncode == 0
        got = open(os.path.join(wd, "34-unique")).read()
    assert got.count("\n") == 373
    assert got == open(os.path.join(GOLDEN, "v34", "34-unique")).read()


@pytest.mark.parametrize("args", [[26, "-c", 2], [45, "-c", 3, "--paf", -2]])
def test_sorted_files_equal_gnu_sort_uniq(oracle, args, tmp_path):
    # sort.cpp:17 runs `sort ... | uniq`; the oracle's sort must give GNU sort's
    # byte order under LC_ALL=C (coreutils sort and uniq are run here as the check)
    d = str(tmp_path)
    assert _cli(list(args) + ["gen", "sort"], d).returncode == 0
    wd = os.path.join(d, "order-%d" % args[0])
    env = dict(os.environ, LC_ALL="C")
    for ab in "ab":
        gen = os.path.join(wd, "%d-filtered-%s#0" % (args[0], ab))
        ref = subprocess.run("sort '%s' | uniq" % gen, shell=True, env=env, capture_output=True, check=True).stdout
        assert open(os.path.join(wd, "%d-filtered_sorted-%s" % (args[0], ab)), "rb").read() == ref
