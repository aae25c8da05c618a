"""Writes the committed golden fixtures from the oracle (oracle/build/refcpu).

    python tests/golden/make_golden.py [c1] [gen45] [gen90] [level90] [level45] [-jN]

c1       every file of `main 16 --compress 2 1 all` (config 1)   -> tests/golden/v16_c2_1/
gen45    sha256 + line counts of the gen files of v=45, c=3, paf -2 -> gen_v45_c3_paf-2.json
gen90    sha256 + line counts of the gen files of v=90, c=10       -> gen_v90_c10.json
level90  sha256 + line counts of every pairs file and the unique file of
         `main 90 --compress 10 5 all` (the whole 10 -> 5 level)   -> level_v90_c10_5.json
level45  the same for `main 45 --paf -2 --compress 3 1 all`        -> level_v45_c3_1_paf-2.json
slice50  a slice of the v=50 5 -> 1 level (config 4): the matched c=5 pairs
         with the fewest expansion leaves (-jN of them, one oracle process
         each): up to three that emit c=1 pairs and one that emits nothing, if
         any; input pair lines and the oracle's 50-pairs-1 lines
                                                                   -> slice_v50_c5_1.json
-jN      oracle processes for the uncompression stage (default 6)

The uncompression stage of a level is cut into contiguous line ranges of the
input pairs file, each range uncompressed by its own oracle process; the
reference appends each pair's matches in input order (uncompression.cpp:50-77),
so the concatenation of the ranges' outputs in range order is the file a single
run writes.
The fixtures are data (inputs and expected outputs); the oracle is the
restatement of the reference's algorithm that produced them.
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CLI = os.path.join(ROOT, "oracle", "build", "refcpu")


def run(args, d):
    # -d ahead of the subcommands: after `uncompress` it is --device (main.cpp:202,220)
    subprocess.run([CLI, "-d", d] + [str(a) for a in args], check=True, stdout=subprocess.DEVNULL)


def c1():
    out = os.path.join(HERE, "v16_c2_1")
    with tempfile.TemporaryDirectory() as d:
        run([16, "--compress", 2, 1, "all"], d)
        src = os.path.join(d, "order-16")
        if os.path.exists(out):
            shutil.rmtree(out)
        shutil.copytree(src, out)


def gen_digest(order, c, paf, name):
    with tempfile.TemporaryDirectory() as d:
        run([order, "--compress", c, "--paf", paf, "gen"], d)
        fx = {"order": order, "compress": c, "paf": paf}
        for side in "ab":
            p = os.path.join(d, "order-%d" % order, "%d-filtered-%s#0" % (order, side))
            data = open(p, "rb").read()
            fx[side] = {"lines": data.count(b"\n"), "sha256": hashlib.sha256(data).hexdigest(),
                        "head": data.decode().splitlines()[:20]}
        json.dump(fx, open(os.path.join(HERE, name), "w"), indent=1)


def digest(path):
    data = open(path, "rb").read()
    return {"lines": data.count(b"\n"), "sha256": hashlib.sha256(data).hexdigest()}


def level_digest(order, comp, paf, name, jobs, chunk_lines=8):
    """Digests of `main <order> --paf <paf> --compress <comp...> all`, with each
    uncompression level sharded over `jobs` oracle processes by input line range."""
    t0 = time.time()
    base = ["--paf", paf, "--compress"] + list(comp)
    fx = {"order": order, "compress": list(comp), "paf": paf, "files": {}}
    with tempfile.TemporaryDirectory() as d:
        run([order] + base + ["gen", "sort", "match"], d)
        wd = os.path.join(d, "order-%d" % order)
        for i in range(len(comp) - 1):
            src = os.path.join(wd, "%d-pairs-%d" % (order, comp[i]))
            lines = open(src).read().splitlines(True)
            chunks = [lines[k:k + chunk_lines] for k in range(0, len(lines), chunk_lines)]

            def one(k):
                cd = os.path.join(d, "chunk%d_%d" % (i, k))
                cw = os.path.join(cd, "order-%d" % order)
                os.makedirs(cw)
                open(os.path.join(cw, "%d-pairs-%d" % (order, comp[i])), "w").write("".join(chunks[k]))
                run([order, "--paf", paf, "--compress", comp[i], comp[i + 1], "uncompress"], cd)
                out = open(os.path.join(cw, "%d-pairs-%d" % (order, comp[i + 1]))).read()
                shutil.rmtree(cd)
                return out

            with ThreadPoolExecutor(max_workers=jobs) as ex:
                outs = list(ex.map(one, range(len(chunks))))
            open(os.path.join(wd, "%d-pairs-%d" % (order, comp[i + 1])), "w").write("".join(outs))
            print("level %d: %d -> %d done (%d input pairs, %.0f s)" % (order, comp[i], comp[i + 1], len(lines),
                                                                         time.time() - t0), flush=True)
        run([order] + base + ["filter"], d)
        for c in comp:
            fx["files"]["%d-pairs-%d" % (order, c)] = digest(os.path.join(wd, "%d-pairs-%d" % (order, c)))
        fx["files"]["%d-unique" % order] = digest(os.path.join(wd, "%d-unique" % order))
        u = open(os.path.join(wd, "%d-unique" % order)).read().splitlines()
        fx["unique_head"] = u[:8]
    fx["oracle_seconds"] = round(time.time() - t0)
    json.dump(fx, open(os.path.join(HERE, name), "w"), indent=1)


def leaf_estimate(word, r, splits):
    """Leaves of one side's expansion: the dihedral orbits of the first letter's
    splits times the split counts of the other letters; the word is rotated to a
    letter with the most splits first (uncompression.cpp:113-142)."""
    best = max(word, key=lambda x: (len(splits[x]), -x))
    orbits = {min(min(t[s:] + t[:s], tuple(reversed(t[s:] + t[:s]))) for s in range(r)) for t in splits[best]}
    n = len(orbits)
    rest = list(word)
    rest.remove(best)
    for x in rest:
        n *= len(splits[x])
    return n


def slice50(jobs, want_emitting=3):
    import itertools
    order, c, c2, paf = 50, 5, 1, 0
    r = c // c2
    splits = {}
    for x in range(-c, c + 1, 2):
        splits[x] = [t for t in itertools.product((-1, 1), repeat=r) if sum(t) == x]
    with tempfile.TemporaryDirectory() as d:
        run([order, "--compress", c, c2, "gen", "sort", "match"], d)
        lines = open(os.path.join(d, "order-%d" % order, "%d-pairs-%d" % (order, c))).read().splitlines(True)
    L = order // c
    est = []
    for i, l in enumerate(lines):
        v = [int(t) for t in l.split()]
        est.append((leaf_estimate(v[:L], r, splits) + leaf_estimate(v[L:], r, splits), i))
    est.sort()

    def one(item):
        n, i = item
        with tempfile.TemporaryDirectory() as cd:
            cw = os.path.join(cd, "order-%d" % order)
            os.makedirs(cw)
            open(os.path.join(cw, "%d-pairs-%d" % (order, c)), "w").write(lines[i])
            run([order, "--compress", c, c2, "uncompress"], cd)
            return open(os.path.join(cw, "%d-pairs-%d" % (order, c2))).read()

    # the cheapest `jobs` pairs at once (about 10^8 leaves each, tens of minutes per
    # oracle process); up to `want_emitting` that emit and one that emits nothing
    batch = est[:jobs]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(one, batch))
    picked = [(item, out) for item, out in zip(batch, outs) if out][:want_emitting]
    silent = [(item, out) for item, out in zip(batch, outs) if not out][:1]
    print("slice50: %d pairs tried, %d emitting" % (len(batch), len(picked) + 0), flush=True)
    chosen = sorted(picked + silent, key=lambda t: t[0][1])     # input file order
    fx = {"order": order, "compress": [c, c2], "paf": paf,
          "input_lines": [t[0][1] for t in chosen],
          "leaves": [t[0][0] for t in chosen],
          "pairs_in": "".join(lines[t[0][1]] for t in chosen),
          "pairs_out": "".join(t[1] for t in chosen)}
    json.dump(fx, open(os.path.join(HERE, "slice_v50_c5_1.json"), "w"), indent=1)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("-j")]
    jobs = int(([a[2:] for a in sys.argv[1:] if a.startswith("-j")] or ["6"])[0])
    what = args or ["c1", "gen45", "gen90"]
    if "c1" in what:
        c1()
    if "gen45" in what:
        gen_digest(45, 3, -2, "gen_v45_c3_paf-2.json")
    if "gen90" in what:
        gen_digest(90, 10, 0, "gen_v90_c10.json")
    if "level90" in what:
        level_digest(90, [10, 5], 0, "level_v90_c10_5.json", jobs)
    if "slice50" in what:
        slice50(jobs)
    if "level45" in what:
        level_digest(45, [3, 1], -2, "level_v45_c3_1_paf-2.json", jobs)
