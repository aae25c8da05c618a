#!/usr/bin/env python3
"""Benchmark: PSD-filtered candidates/s of the v=50 Golay-pair search on one MI355X.

Workload (one step): the exhaustive search of order v=50, PAF constant 0,
compression 5 -> 1, on the GPU through the C ABI:
  gen at c=5 (length-10 words over {-5,-3,-1,1,3,5}, row sums (0, 10) and (6, 8),
  canonical, PSD filter), sort + match at c=5 (4 691 pairs), uncompression 5 -> 1
  of every matched pair (9.2e11 length-50 words over {-1,1}: PSD filter + keys +
  re-match with exact PAF), equivalence filter at length 50 with altnegative.
value = PSD-filter decisions of the step / wall time of the step (split-BF16 MFMA
path).  Beside it: the plain FP32 path on a fixed slice of the v=50 pairs (same
output required), the mixed-length PSD microbenchmark, and the CPU oracle on P
processes.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ORDER, C0, C1, PAF = 50, 5, 1, 0
N1 = ORDER // C1
FLOPS_PER_LEAF = 4 * N1 * (N1 // 2) + 3 * (N1 // 2)   # algorithmic: real DFT k=1..25 of n=50 + |X|^2 = 5075
MFMA_PEAK_TFLOPS = 2500.0                             # MI355X dense BF16
HBM_PEAK = 8e12                                       # MI355X HBM3E, bytes/s
FP32_SLICE = (0, 64)                                  # matched c=5 pairs [0, 64) in file order
PROFILE = "k_unc_filter_v50_golay_search.md"


STAGE_MS = {}


def search_step(ctx, g):
    st = g.GpStats()
    t = [time.perf_counter()]
    a, b, s = ctx.generate(ORDER, C0, PAF)
    st.add(s)
    t.append(time.perf_counter())
    sa = ctx.sort_candidates(a)
    sb = ctx.sort_candidates(b)
    t.append(time.perf_counter())
    rc, pa, pb, s = ctx.match(ORDER, C0, PAF, sa, sb)
    st.add(s)
    t.append(time.perf_counter())
    ua, ub, _, _, su = ctx.uncompress(ORDER, C0, C1, PAF, pa, pb)
    st.add(su)
    t.append(time.perf_counter())
    fa, fb, s = ctx.filter(ORDER, C1, ua, ub)
    st.add(s)
    t.append(time.perf_counter())
    for i, name in enumerate(("gen", "sort", "match", "uncompress", "filter")):
        STAGE_MS[name] = (t[i + 1] - t[i]) * 1e3
    return st, su, len(pa), len(ua), len(fa)


def profiled_traffic(kernel="k_unc_filter"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (KiB per dispatch, scaled by what the same
    counters read for a plain copy of known size).  Not measured by this run."""
    path = os.path.join(ROOT, "profiles", PROFILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic %s fetch_kib=([\d.]+) fetch_scale=([\d.]+) write_kib=([\d.]+) write_scale=([\d.]+)"
                     % kernel, line)
        if m:
            f, fs, w, ws = map(float, m.groups())
            return (f * fs + w * ws) * 1024.0
    return None


def timed(ctx, g, steps, warmup):
    for _ in range(warmup):
        search_step(ctx, g)
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = search_step(ctx, g)
    t1 = time.perf_counter()
    return (t1 - t0) / steps, out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_worker(argv):
    """One single-threaded oracle process: job j of P of one side's uncompression
    (direct FP64 DFT per leaf, PSD filter, candidate line), timed."""
    from tests import _oracle
    seq = [int(x) for x in argv[0].split(",")]
    seqflag, job, njobs = int(argv[1]), int(argv[2]), int(argv[3])
    _oracle.load()
    t0 = time.perf_counter()
    text, leaves = _oracle.uncompress_side(seq, C0, C1, PAF, seqflag, job, njobs)
    print(json.dumps({"leaves": leaves, "seconds": time.perf_counter() - t0}))


def side_leaves(w):
    """Leaves of one c=5 side expanded to +-1: 10 splits of +-1, 5 of +-3, 1 of +-5,
    the first (best) letter's splits counted once per dihedral orbit (2, 1, 1)."""
    split = {1: 10, -1: 10, 3: 5, -3: 5, 5: 1, -5: 1}
    first = {1: 2, -1: 2, 3: 1, -3: 1, 5: 1, -5: 1}
    best = min(w, key=lambda x: (-split[int(x)], int(x)))
    n = first[int(best)]
    rest = [int(x) for x in w]
    rest.remove(int(best))
    for x in rest:
        n *= split[x]
    return n


def cpu_baseline(pa, pb):
    """The oracle on P single-threaded processes (P = the host cores this process
    may use, at most 16): each takes job 0 of 2 of one of the P cheapest sides of
    the v=50 matched pairs (uncompression.cpp:167-191 split), about 10-30 s each."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    P = max(1, min(16, cores))
    sides = sorted({(side_leaves(w), tuple(int(x) for x in w), f) for ws, f in ((pa, 0), (pb, 1)) for w in ws})[:P]
    t0 = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               ",".join(str(x) for x in w), str(f), "0", "2"], stdout=subprocess.PIPE, text=True)
             for _, w, f in sides]
    res = [json.loads(p.communicate()[0]) for p in procs]
    wall = time.perf_counter() - t0
    leaves = sum(r["leaves"] for r in res)
    return {"value": leaves / wall, "unit": "candidates/s", "cores": P, "kind": "port",
            "cpu_model": cpu_model(), "processes": P,
            "sample": "oracle uncompression 5->1 of the %d cheapest sides of the v=50 matched pairs, job 0 of 2 "
                      "each (uncompression.cpp:167-191), one single-threaded process per side: %d leaves in "
                      "%.1f s wall (slowest process %.1f s); direct FP64 DFT per leaf, FFTW absent"
                      % (P, leaves, wall, max(r["seconds"] for r in res))}


def microbench(ctx, g, mode):
    mix = np.array([[9, 10, 90, 0, 1 << 25], [10, 5, 50, 0, 1 << 24], [15, 3, 45, -2, 1 << 23],
                    [17, 3, 51, -2, 1 << 22], [18, 5, 90, 0, 1 << 21], [25, 2, 50, 0, 1 << 21]], np.int32)
    ctx.set_mode(mode)
    ctx.psd_microbench(0x601A7E5EED, mix)
    acc, digest, _, st = ctx.psd_microbench(0x601A7E5EED, mix)
    ctx.set_mode(g.GP_PSD_SPLIT_BF16)
    n = int(mix[:, 4].sum())
    # algorithmic bytes: n letters in and one decision out per candidate; the rows
    # are stored at 16 bytes per block of 16 letters
    alg = int(((mix[:, 0] + 1) * mix[:, 4]).sum())
    stored = int(((16 * ((mix[:, 0] + 15) // 16) + 1) * mix[:, 4]).sum())
    sec = st.psd_kernel_ms * 1e-3
    return {"candidates": n, "kernel_ms": st.psd_kernel_ms, "band_ms": st.band_kernel_ms,
            "candidates_per_s": n / sec, "band_fraction": st.band_rechecks / n,
            "accepted": int(acc.sum()), "digest": "%016x" % digest, "tie_band": st.tie_band,
            "algorithmic_bytes": alg, "stored_bytes": stored,
            "hbm_frac_algorithmic": alg / sec / HBM_PEAK, "hbm_frac_stored": stored / sec / HBM_PEAK}


MB_LENGTHS = [[9, 10, 90, 0], [10, 5, 50, 0], [15, 3, 45, -2], [16, 1, 16, 0], [17, 3, 51, -2], [18, 5, 90, 0],
              [25, 2, 50, 0], [9, 10, 90, 400]]


def microbench_by_length(count=1 << 26):
    """The microbenchmark's filter on `count` candidates of one word length at a
    time (n, c, v, paf as in the mix; the last entry lifts the bound so that every
    candidate lands in the FP64 band): filter and recheck kernel times, band
    fraction, rows per second."""
    import golaypair_amd as g
    ctx = g.Context(0)
    out = []
    for n, c, v, paf in MB_LENGTHS:
        mix = np.array([[n, c, v, paf, count]], np.int32)
        ctx.psd_microbench(0x601A7E5EED, mix)
        acc, _, _, st = ctx.psd_microbench(0x601A7E5EED, mix)
        out.append({"n": n, "c": c, "v": v, "paf": paf, "kernel_ms": st.psd_kernel_ms, "band_ms": st.band_kernel_ms,
                    "band_fraction": st.band_rechecks / count, "accepted_fraction": int(acc.sum()) / count,
                    "rows_per_s": count / (st.psd_kernel_ms * 1e-3),
                    "stored_bytes_per_s": count * (16 * ((n + 15) // 16) + 1) / (st.psd_kernel_ms * 1e-3)})
    ctx.close()
    return out


def sort_scale(sizes=(1 << 20, 1 << 22, 1 << 24), n=15, letters=(-3, -1, 1, 3)):
    """The device sort + uniq (gp_sort_candidates) on random candidate records of the
    v=45 shape (n = 15 letters of {-3,-1,1,3}, 7 key digits 0..92), each size sorted
    twice (the second timed): wall time of the call (host-to-device and back
    included) and records per second; rocprofv3 --kernel-trace gives the kernels."""
    import golaypair_amd as g
    ctx = g.Context(0)
    rng = np.random.default_rng(15)
    out = []
    for m in sizes:
        words = rng.choice(letters, size=(m, n)).astype(np.int8)
        keys = rng.integers(0, 93, size=(m, n // 2)).astype(np.uint8)
        c = g.Cands(words, keys)
        ctx.sort_candidates(c)
        t0 = time.perf_counter()
        s = ctx.sort_candidates(c)
        dt = time.perf_counter() - t0
        out.append({"records": m, "unique": len(s), "record_bytes": n + n // 2, "wall_ms": dt * 1e3,
                    "records_per_s": m / dt})
        print("sort %d records: %.2f ms" % (m, dt * 1e3), file=sys.stderr, flush=True)
    ctx.close()
    return out


def match_scale(m=1 << 21, n=15, letters=(-3, -1, 1, 3)):
    """The device merge-join (gp_match) on two random candidate sets of m records
    of the v=45 shape, half of the B keys copied from A rows so that key groups
    meet (their random words then fail the exact PAF check): joined twice, the
    second timed (wall time of the call, transfers included); rocprofv3
    --kernel-trace gives the kernels."""
    import golaypair_amd as g
    ctx = g.Context(0)
    rng = np.random.default_rng(45)
    wa = rng.choice(letters, size=(m, n)).astype(np.int8)
    ka = rng.integers(0, 93, size=(m, n // 2)).astype(np.uint8)
    wb = rng.choice(letters, size=(m, n)).astype(np.int8)
    kb = rng.integers(0, 93, size=(m, n // 2)).astype(np.uint8)
    kb[: m // 2] = ka[rng.integers(0, m, size=m // 2)]
    a = ctx.sort_candidates(g.Cands(wa, ka))
    b = ctx.sort_candidates(g.Cands(wb, kb))
    ctx.match(45, 3, -2, a, b)
    t0 = time.perf_counter()
    rc, pa, _, st = ctx.match(45, 3, -2, a, b)
    dt = time.perf_counter() - t0
    ctx.close()
    return {"records_a": len(a), "records_b": len(b), "record_bytes": n + n // 2, "pairs": len(pa),
            "pair_checks": st.pair_checks, "wall_ms": dt * 1e3}


def copy_probe(mib=1024, reps=4):
    """A plain device-to-device copy of known size (PMC calibration for FETCH_SIZE
    and WRITE_SIZE): `reps` copies of `mib` MiB."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": mib << 20, "reps": reps}))


E2E_CONFIGS = {  # name -> CLI arguments (the configurations of the reference's workloads)
    "v16": ["16", "--compress", "2", "1"],
    "v45": ["45", "--paf", "-2", "--compress", "3", "1"],
    "v90": ["90", "--compress", "10", "5"],
    "v50": ["50", "--compress", "5", "1"],
}


def e2e(names, workdir="/tmp/golay_e2e"):
    """End-to-end search time: wall clock of the drop-in CLI from start until
    {v}-unique is written (every results file written), per configuration."""
    import hashlib
    import shutil
    import subprocess
    cli = os.path.join(ROOT, "golaypair_amd", "bin", "golaypair")
    out = {}
    for name in names:
        args = E2E_CONFIGS[name]
        d = os.path.join(workdir, name)
        shutil.rmtree(d, ignore_errors=True)
        print("e2e %s: %s" % (name, " ".join(args)), file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        p = subprocess.run([cli] + args + ["-d", d, "all"], capture_output=True, text=True)
        dt = time.perf_counter() - t0
        uniq = os.path.join(d, "order-" + args[0], args[0] + "-unique")
        r = {"args": " ".join(args + ["-d", d, "all"]), "rc": p.returncode, "seconds": dt}
        if os.path.exists(uniq):
            data = open(uniq, "rb").read()
            r["unique_classes"] = data.count(b"\n")
            r["unique_sha256"] = hashlib.sha256(data).hexdigest()
        out[name] = r
    return out


def fp32_parity(order, c0, c1, paf, chunk=200):
    """The uncompression of every matched pair of a level on both PSD paths (in
    chunks of pairs, printing progress): outputs must be identical; reports the
    kernel-time ratio FP32 / split-BF16."""
    import hashlib
    import golaypair_amd as g
    ctx = g.Context(0)
    a, b, _ = ctx.generate(order, c0, paf)
    rc, pa, pb, _ = ctx.match(order, c0, paf, ctx.sort_candidates(a), ctx.sort_candidates(b))
    res = {"order": order, "compress": [c0, c1], "paf": paf, "pairs_in": len(pa)}
    texts = {}
    for mode, name in ((g.GP_PSD_SPLIT_BF16, "split_bf16"), (g.GP_PSD_FP32, "fp32")):
        ctx.set_mode(mode)
        parts, leaves, kern, t0 = [], 0, 0.0, time.perf_counter()
        for q in range(0, len(pa), chunk):
            oa, ob, _, _, st = ctx.uncompress(order, c0, c1, paf, pa[q:q + chunk], pb[q:q + chunk])
            parts.append("".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob)))
            leaves += st.psd_decisions
            kern += st.psd_kernel_ms
            print("%s: %d/%d pairs, %.3e leaves, %.1f s" % (name, min(q + chunk, len(pa)), len(pa), leaves,
                                                           time.perf_counter() - t0), file=sys.stderr, flush=True)
        texts[name] = "".join(parts)
        res[name] = {"leaves": leaves, "psd_kernel_ms": kern, "wall_s": time.perf_counter() - t0,
                     "pairs_out": texts[name].count("\n"),
                     "sha256": hashlib.sha256(texts[name].encode()).hexdigest(),
                     "candidates_per_s_kernel": leaves / (kern * 1e-3)}
    ctx.close()
    res["identical"] = texts["split_bf16"] == texts["fp32"]
    res["fp32_over_split_kernel_time"] = res["fp32"]["psd_kernel_ms"] / res["split_bf16"]["psd_kernel_ms"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-microbench", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="run only a plain 1 GiB copy (PMC calibration)")
    ap.add_argument("--microbench-only", action="store_true",
                    help="only: the mixed-length PSD microbenchmark (2^26 candidates) on both paths")
    ap.add_argument("--microbench-by-length", action="store_true",
                    help="only: the microbenchmark's filter on 2^26 candidates of each word length alone")
    ap.add_argument("--sort-scale", action="store_true",
                    help="only: the device candidate sort on 2^20 .. 2^24 random records")
    ap.add_argument("--match-scale", action="store_true",
                    help="only: the device merge-join on two random sets of 2^21 records")
    ap.add_argument("--e2e", nargs="*", metavar="CONFIG",
                    help="only: end-to-end CLI search time of %s (default all)" % ",".join(E2E_CONFIGS))
    ap.add_argument("--fp32-parity", type=int, metavar="ORDER", choices=(45, 50, 90),
                    help="only: uncompression of a whole level on both PSD paths, identical output and time ratio")
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        cpu_worker(sys.argv[2:])
        return
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.sort_scale:
        print(json.dumps(sort_scale()))
        return
    if args.match_scale:
        print(json.dumps(match_scale()))
        return
    if args.microbench_by_length:
        print(json.dumps(microbench_by_length()))
        return
    if args.microbench_only:
        import golaypair_amd as g
        ctx = g.Context(0)
        print(json.dumps({"split_bf16": microbench(ctx, g, g.GP_PSD_SPLIT_BF16),
                          "fp32": microbench(ctx, g, g.GP_PSD_FP32)}))
        ctx.close()
        return
    if args.e2e is not None:
        print(json.dumps(e2e(args.e2e or list(E2E_CONFIGS))))
        return
    if args.fp32_parity:
        cfg = {45: (45, 3, 1, -2), 50: (50, 5, 1, 0), 90: (90, 10, 5, 0)}[args.fp32_parity]
        print(json.dumps(fp32_parity(*cfg)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the search runs on a single MI355X (no multi-GPU path is built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import golaypair_amd as g
    ctx = g.Context(0)
    sec, (st, su, npairs0, npairs1, nunique) = timed(ctx, g, args.steps, args.warmup)
    value = st.psd_decisions / sec
    launches = max(1, su.psd_launches)
    kern_ms = su.psd_kernel_ms / launches
    achieved = su.psd_decisions * FLOPS_PER_LEAF / (su.psd_kernel_ms * 1e-3) / 1e12
    line = {
        "metric": "PSD-filtered candidates/s",
        "value": value,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "v=50 Golay-pair search (paf 0): gen at c=5, sort, match, uncompress 5->1 of all "
                               "matched pairs, equivalence filter at length 50 with altnegative; exhaustive, "
                               "no input data",
                   "order": ORDER, "compress": [C0, C1], "paf": PAF, "psd": "split-BF16 MFMA, FP32 accumulate, "
                   "FP64 band recheck and keys"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_PEAK_TFLOPS, "traffic": profiled_traffic(),
                     "traffic_source": "profiles/%s (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, "
                                       "not measured by this run)" % PROFILE,
                     "kernel": "k_unc_filter", "kernel_ms_per_launch": kern_ms, "launches_per_step": launches,
                     "flops_per_candidate": FLOPS_PER_LEAF, "candidates_per_launch": su.psd_decisions / launches},
        "end_to_end_s": sec,
        "stage_ms_last_step": dict(STAGE_MS),
        "search": {"pairs_c%d" % C0: npairs0, "pairs_c%d" % C1: npairs1, "unique": nunique,
                   "psd_decisions": st.psd_decisions, "leaves": su.psd_decisions, "band_rechecks": st.band_rechecks,
                   "tie_band": st.tie_band, "psd_pass": st.psd_pass, "pair_checks": st.pair_checks,
                   "kernel_ms": st.kernel_ms},
    }
    pa = pb = None
    if not (args.no_fp32 and args.no_cpu_baseline):
        a, b, _ = ctx.generate(ORDER, C0, PAF)
        rc, pa, pb, _ = ctx.match(ORDER, C0, PAF, ctx.sort_candidates(a), ctx.sort_candidates(b))
    if not args.no_fp32:
        q0, q1 = FP32_SLICE
        outs, res = {}, {}
        for mode, name in ((g.GP_PSD_SPLIT_BF16, "split_bf16"), (g.GP_PSD_FP32, "fp32")):
            ctx.set_mode(mode)
            t0 = time.perf_counter()
            oa, ob, _, _, s = ctx.uncompress(ORDER, C0, C1, PAF, pa[q0:q1], pb[q0:q1])
            dt = time.perf_counter() - t0
            outs[name] = "".join(g.pair_line(x, y) + "\n" for x, y in zip(oa, ob))
            res[name] = {"leaves": s.psd_decisions, "psd_kernel_ms": s.psd_kernel_ms, "wall_ms": dt * 1e3,
                         "candidates_per_s_kernel": s.psd_decisions / (s.psd_kernel_ms * 1e-3),
                         "pairs_out": len(oa)}
        ctx.set_mode(g.GP_PSD_SPLIT_BF16)
        line["fp32_path"] = {"slice": "uncompression 5->1 of the v=50 matched pairs [%d, %d) in file order" % (q0, q1),
                             "identical_output": outs["split_bf16"] == outs["fp32"],
                             "split_bf16": res["split_bf16"], "fp32": res["fp32"],
                             "split_bf16_speedup_kernel": res["fp32"]["psd_kernel_ms"] /
                             max(res["split_bf16"]["psd_kernel_ms"], 1e-9)}
    if not args.no_microbench:
        line["microbench"] = {"split_bf16": microbench(ctx, g, g.GP_PSD_SPLIT_BF16),
                              "fp32": microbench(ctx, g, g.GP_PSD_FP32)}
    ctx.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(pa, pb)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
