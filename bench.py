#!/usr/bin/env python3
"""Benchmark: PSD-filtered candidates/s of the v=45 Legendre-pair search on one MI355X.

Workload (one step): the whole search of order v=45, PAF constant -2,
compression 3 -> 1, on the GPU through the C ABI:
  gen at c=3 (length-15 words over {-3,-1,1,3}, row sum 1, canonical, PSD filter),
  sort + match at c=3, uncompression 3 -> 1 of every matched pair (length-45
  words over {-1,1}, PSD filter + keys + match with exact PAF), equivalence filter.
value = PSD-filter decisions of the step / wall time of the step (split-BF16 MFMA
path); the plain FP32 path and the CPU oracle are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ORDER, C0, C1, PAF = 45, 3, 1, -2
FLOPS_PER_LEAF = 4 * 45 * 22 + 3 * 22       # algorithmic: real DFT k=1..22 of n=45 + |X|^2
MFMA_PEAK_TFLOPS = 2500.0                    # MI355X dense BF16


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
    WRITE_SIZE summary (KiB per dispatch, scaled by a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", "k_unc_filter_v45_legendre_search.md")
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


def cpu_baseline(pa, pb, seconds_hint=20.0):
    """The oracle (direct FP64 DFT per leaf, one sequence per thread) on a bounded
    sample: the uncompression of the first matched pairs, both sides."""
    from tests import _oracle
    _oracle.load()
    threads = max(1, min(16, os.cpu_count() or 1))
    sides = []
    for q in range(len(pa)):
        sides.append((pa[q], 0))
        sides.append((pb[q], 1))
        if len(sides) >= threads:
            break
    leaves = [0] * len(sides)

    def work(i):
        _, leaves[i] = _oracle.uncompress_side(sides[i][0], C0, C1, PAF, sides[i][1])

    t0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(i,)) for i in range(len(sides))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    return {"value": sum(leaves) / dt, "unit": "candidates/s", "cores": len(sides), "kind": "port",
            "sample": "oracle uncompression 3->1 of the first %d sequences of the v=45 matched pairs "
                      "(%d leaves, %.1f s wall, one sequence per thread; direct FP64 DFT, FFTW absent)"
                      % (len(sides), sum(leaves), dt)}


def microbench(ctx, g, mode):
    mix = np.array([[9, 10, 90, 0, 1 << 25], [10, 5, 50, 0, 1 << 24], [15, 3, 45, -2, 1 << 23],
                    [17, 3, 51, -2, 1 << 22], [18, 5, 90, 0, 1 << 21], [25, 2, 50, 0, 1 << 21]], np.int32)
    ctx.set_mode(mode)
    ctx.psd_microbench(0x601A7E5EED, mix)
    acc, digest, _, st = ctx.psd_microbench(0x601A7E5EED, mix)
    ctx.set_mode(g.GP_PSD_SPLIT_BF16)
    n = int(mix[:, 4].sum())
    return {"candidates": n, "kernel_ms": st.psd_kernel_ms, "band_ms": st.band_kernel_ms,
            "candidates_per_s": n / (st.psd_kernel_ms * 1e-3), "band_fraction": st.band_rechecks / n,
            "accepted": int(acc.sum()), "digest": "%016x" % digest, "tie_band": st.tie_band}


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
    ap.add_argument("--e2e", nargs="*", metavar="CONFIG",
                    help="only: end-to-end CLI search time of %s (default all)" % ",".join(E2E_CONFIGS))
    ap.add_argument("--fp32-parity", type=int, metavar="ORDER", choices=(45, 50, 90),
                    help="only: uncompression of a whole level on both PSD paths, identical output and time ratio")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
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
    sec, (st, su, npairs3, npairs1, nunique) = timed(ctx, g, args.steps, args.warmup)
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
        "config": {"workload": "v=45 Legendre-pair search (paf -2): gen at c=3, match, uncompress 3->1 of all "
                               "matched pairs, equivalence filter; exhaustive, no input data",
                   "order": ORDER, "compress": [C0, C1], "paf": PAF, "psd": "split-BF16 MFMA, FP32 accumulate, "
                   "FP64 band recheck and keys"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_PEAK_TFLOPS, "traffic": profiled_traffic(),
                     "kernel": "k_unc_filter", "kernel_ms_per_launch": kern_ms,
                     "flops_per_candidate": FLOPS_PER_LEAF, "candidates_per_launch": su.psd_decisions / launches},
        "end_to_end_s": sec,
        "stage_ms_last_step": dict(STAGE_MS),
        "search": {"pairs_c3": npairs3, "pairs_c1": npairs1, "unique": nunique, "psd_decisions": st.psd_decisions,
                   "leaves": su.psd_decisions, "band_rechecks": st.band_rechecks, "tie_band": st.tie_band,
                   "psd_pass": st.psd_pass, "pair_checks": st.pair_checks, "kernel_ms": st.kernel_ms},
    }
    if not args.no_fp32:
        ctx.set_mode(g.GP_PSD_FP32)
        sec32, (st32, su32, *_rest) = timed(ctx, g, 1, 0)
        ctx.set_mode(g.GP_PSD_SPLIT_BF16)
        line["fp32_path"] = {"value": st32.psd_decisions / sec32, "ms_per_step": sec32 * 1e3,
                             "uncompress_kernel_ms": su32.psd_kernel_ms,
                             "split_bf16_speedup_kernel": su32.psd_kernel_ms / max(su.psd_kernel_ms, 1e-9),
                             "split_bf16_speedup_step": sec32 / sec}
    if not args.no_microbench:
        line["microbench"] = {"split_bf16": microbench(ctx, g, g.GP_PSD_SPLIT_BF16),
                              "fp32": microbench(ctx, g, g.GP_PSD_FP32)}
    if not args.no_cpu_baseline:
        a, b, _ = ctx.generate(ORDER, C0, PAF)
        rc, pa, pb, _ = ctx.match(ORDER, C0, PAF, ctx.sort_candidates(a), ctx.sort_candidates(b))
        line["cpu_baseline"] = cpu_baseline(pa, pb)
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
