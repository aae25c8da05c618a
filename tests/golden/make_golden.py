"""Writes the committed golden fixtures from the oracle (oracle/build/refcpu).

    python tests/golden/make_golden.py [c1] [gen45] [gen90]

c1     every file of `main 16 --compress 2 1 all` (config 1)   -> tests/golden/v16_c2_1/
gen45  sha256 + line counts of the gen files of v=45, c=3, paf -2 -> gen_v45_c3_paf-2.json
gen90  sha256 + line counts of the gen files of v=90, c=10       -> gen_v90_c10.json
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

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CLI = os.path.join(ROOT, "oracle", "build", "refcpu")


def run(args, d):
    subprocess.run([CLI] + [str(a) for a in args] + ["-d", d], check=True, stdout=subprocess.DEVNULL)


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


if __name__ == "__main__":
    what = sys.argv[1:] or ["c1", "gen45", "gen90"]
    if "c1" in what:
        c1()
    if "gen45" in what:
        gen_digest(45, 3, -2, "gen_v45_c3_paf-2.json")
    if "gen90" in what:
        gen_digest(90, 10, 0, "gen_v90_c10.json")
