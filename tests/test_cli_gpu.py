"""The drop-in CLI (golaypair_amd/bin/golaypair) vs the oracle CLI: every file of
the results directory byte-identical (src/main.cpp:14-251 file names and formats),
same exit codes, and stage-by-stage interop with files the oracle wrote."""
import filecmp
import os
import subprocess
import tempfile

import pytest

import golaypair_amd as g
from tests import _oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def run(exe, args, d):
    # -d before the subcommands: after `uncompress` it means --device (main.cpp:202,220)
    return subprocess.run([exe, "-d", d] + [str(a) for a in args], capture_output=True, text=True)


def same_dirs(d1, d2):
    f1, f2 = sorted(os.listdir(d1)), sorted(os.listdir(d2))
    assert f1 == f2
    for f in f1:
        assert open(os.path.join(d1, f), "rb").read() == open(os.path.join(d2, f), "rb").read(), f


@pytest.mark.parametrize("args", [
    [16, "--compress", 2, 1, "all"],
    [26, "--compress", 2, 1, "all"],
    [21, "--compress", 3, 1, "--paf", -2, "all"],
    [20, "--compress", 10, 5, 1, "all"],
    [16, "--compress", 4, 2, 1, "all"],
    [20, "-c", 5, 1, "all"],
    [27, "-c", 3, 1, "--paf", -2, "all"],
])
def test_cli_all_matches_oracle(args):
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        r1 = run(g.CLI_PATH, args, d1)
        r2 = run(_oracle.CLI, args, d2)
        assert r1.returncode == r2.returncode == 0, r1.stderr
        order = args[0]
        same_dirs(os.path.join(d1, "order-%d" % order), os.path.join(d2, "order-%d" % order))


@pytest.mark.parametrize("flags,select", [
    (["--range_begin", 1, "--range_step", 2], lambda i: i >= 1 and (i - 1) % 2 == 0),
    (["-b", 0, "-e", 3], lambda i: i < 3),
    (["--line", 2], lambda i: i == 1),
])
def test_cli_uncompress_slicing(flags, select):
    """--line (1-indexed) and --range_begin/--range_end/--range_step (0-indexed
    lines, main.cpp:110-113,192-196; script/uncompress.sh shards with them)
    uncompress only the selected lines of the input pairs file: the output is
    the oracle's for a file holding just those lines."""
    order, comp = 26, [2, 1]
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        assert run(_oracle.CLI, [order, "--compress"] + comp + ["gen", "sort", "match"], d2).returncode == 0
        wd1, wd2 = os.path.join(d1, "order-%d" % order), os.path.join(d2, "order-%d" % order)
        name = "%d-pairs-%d" % (order, comp[0])
        lines = open(os.path.join(wd2, name)).read().splitlines(True)
        assert len(lines) > 4
        os.makedirs(wd1)
        open(os.path.join(wd1, name), "w").write("".join(lines))
        r = run(g.CLI_PATH, [order, "--compress"] + comp + ["uncompress"] + flags, d1)
        assert r.returncode == 0, r.stderr
        open(os.path.join(wd2, name), "w").write("".join(l for i, l in enumerate(lines) if select(i)))
        assert run(_oracle.CLI, [order, "--compress"] + comp + ["uncompress"], d2).returncode == 0
        out = "%d-pairs-%d" % (order, comp[1])
        assert open(os.path.join(wd1, out)).read() == open(os.path.join(wd2, out)).read()


@pytest.mark.parametrize("args", [
    [16, "--compress", 2, 1, "all"],
    [21, "--compress", 3, 1, "--paf", -2, "all"],
    [20, "--compress", 10, 5, 1, "all"],
])
def test_cli_binary_candidate_files(args):
    """--binary writes the candidate files (gen shards, sorted files, the
    uncompression temporary files) in the compact binary form; read back they
    give the text files line for line, and the pair and unique files are the
    same bytes as in text mode."""
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        assert run(g.CLI_PATH, args + ["--binary"], d1).returncode == 0
        assert run(g.CLI_PATH, args, d2).returncode == 0
        order = args[0]
        w1, w2 = os.path.join(d1, "order-%d" % order), os.path.join(d2, "order-%d" % order)
        assert sorted(os.listdir(w1)) == sorted(os.listdir(w2))
        nbin = 0
        for f in os.listdir(w1):
            raw = open(os.path.join(w1, f), "rb").read()
            text = open(os.path.join(w2, f)).read()
            if raw[:4] == g.BINARY_MAGIC:
                nbin += 1
                got = "".join(l + "\n" for l in g.read_binary_cands(os.path.join(w1, f)).lines())
                assert got == text, f
                if text.count("\n") >= 16:  # the 16-byte header aside, records are smaller than lines
                    assert len(raw) < len(text.encode()), f
            else:
                assert raw == text.encode(), f
        assert nbin >= 4


def test_cli_c1_golden():
    with tempfile.TemporaryDirectory() as d:
        assert run(g.CLI_PATH, [16, "--compress", 2, 1, "all"], d).returncode == 0
        same_dirs(os.path.join(d, "order-16"), os.path.join(GOLDEN, "v16_c2_1"))


def test_cli_v34_known_answer():
    with tempfile.TemporaryDirectory() as d:
        r = run(g.CLI_PATH, [34, "--compress", 2, 1, "all"], d)
        assert r.returncode == 0, r.stderr
        got = open(os.path.join(d, "order-34", "34-unique")).read()
        assert got.count("\n") == 373
        assert got == open(os.path.join(GOLDEN, "v34", "34-unique")).read()


def test_cli_stage_interop_with_oracle_files():
    # oracle writes gen shards (-j/-n), our CLI sorts, matches, uncompresses, filters
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        for j in range(3):
            assert run(_oracle.CLI, [26, "-c", 2, 1, "-j", j, "-n", 3, "gen"], d1).returncode == 0
            assert run(g.CLI_PATH, [26, "-c", 2, 1, "-j", j, "-n", 3, "gen"], d2).returncode == 0
        wd1, wd2 = os.path.join(d1, "order-26"), os.path.join(d2, "order-26")
        same_dirs(wd1, wd2)
        for stage in ("sort", "match", "uncompress", "filter"):
            assert run(g.CLI_PATH, [26, "-c", 2, 1, "-n", 3, stage], d1).returncode == 0
            assert run(_oracle.CLI, [26, "-c", 2, 1, "-n", 3, stage], d2).returncode == 0
        same_dirs(wd1, wd2)


def test_cli_exit_codes():
    with tempfile.TemporaryDirectory() as d:
        assert run(g.CLI_PATH, [7, "gen"], d).returncode == 1          # no decomposition
        assert run(g.CLI_PATH, [16], d).returncode != 0                 # a subcommand is required
        assert run(g.CLI_PATH, [16, "-c", 2, "match"], d).returncode == 1   # no sorted files
        r = run(g.CLI_PATH, [16, "-c", 2, 1, "filter"], d)             # missing pairs file
        assert r.returncode == 0 and "Bad File." in r.stdout


def _sha(path):
    import hashlib
    data = open(path, "rb").read()
    return {"lines": data.count(b"\n"), "sha256": hashlib.sha256(data).hexdigest()}


@pytest.mark.parametrize("fixture,args", [
    ("level_v90_c10_5.json", [90, "--compress", 10, 5, "all"]),
    ("level_v45_c3_1_paf-2.json", [45, "--paf", -2, "--compress", 3, 1, "all"]),
])
def test_cli_whole_level_matches_oracle_digests(fixture, args):
    """Whole levels at full size: every matched pair of the coarse level is
    uncompressed (v=90: 156 pairs, 1.9e8 leaves at length 18; v=45: 7 577 pairs,
    2.9e9 leaves at length 45) and filtered.  The pairs files at both compressions
    and the unique file must be the oracle's bytes (sha256 and line counts written
    by tests/golden/make_golden.py)."""
    import json
    fx = json.load(open(os.path.join(GOLDEN, fixture)))
    with tempfile.TemporaryDirectory() as d:
        r = run(g.CLI_PATH, args, d)
        assert r.returncode == 0, r.stderr
        wd = os.path.join(d, "order-%d" % fx["order"])
        for name, want in fx["files"].items():
            assert _sha(os.path.join(wd, name)) == want, name


@pytest.mark.parametrize("order,route,classes", [
    (32, [2, 1], 838),          # NOTES.md:571-599 (838 after the altnegative fix, as in the 2022 paper)
    (40, [4, 2, 1], 9281),      # NOTES.md:605-616 (the reference's count, with and without compression)
])
def test_cli_known_class_counts(order, route, classes):
    """Equivalence-class counts the reference's own notes record for orders 32 and
    40, through the whole GPU pipeline (gen, sort, match, uncompress to +-1, filter
    with altnegative at length 32 / 40)."""
    with tempfile.TemporaryDirectory() as d:
        r = run(g.CLI_PATH, [order, "--compress"] + route + ["all"], d)
        assert r.returncode == 0, r.stderr
        got = open(os.path.join(d, "order-%d" % order, "%d-unique" % order)).read()
        assert got.count("\n") == classes
