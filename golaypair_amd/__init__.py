"""MI355X (gfx950) GolayPair search pipeline.

The compute lives in the C-ABI library ``lib/libgolaypair_gpu.so`` (declared in
``include/golay_gpu.h``); ``bin/golaypair`` is the drop-in for the reference's
``main`` command line.  This module is a thin ctypes binding for Python callers
(tests, bench.py).  There is no CPU fallback: if the library is missing or no
GPU is present, the calls raise.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgolaypair_gpu.so")
CLI_PATH = os.path.join(PKG_DIR, "bin", "golaypair")
HEADER_PATH = os.path.join(ROOT_DIR, "include", "golay_gpu.h")

GP_OK = 0
GP_NO_INPUT = 1
GP_ERR_ARG = -1
GP_ERR_HIP = -2
GP_ERR_CAPACITY = -3
GP_ERR_UNSUPPORTED = -4
GP_PSD_SPLIT_BF16 = 0
GP_PSD_FP32 = 1

_P = ctypes.POINTER


class GpCands(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nkey", ctypes.c_int32), ("count", ctypes.c_int64),
                ("capacity", ctypes.c_int64), ("words", _P(ctypes.c_int8)), ("keys", _P(ctypes.c_uint8))]


class GpPairs(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("count", ctypes.c_int64), ("capacity", ctypes.c_int64),
                ("a", _P(ctypes.c_int8)), ("b", _P(ctypes.c_int8))]


class GpStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in
                ("words", "psd_decisions", "psd_pass", "band_rechecks", "key_rechecks", "canonical",
                 "pair_checks", "pairs", "tie_band")] + [("kernel_ms", ctypes.c_double),
                                                         ("psd_kernel_ms", ctypes.c_double),
                                                         ("band_kernel_ms", ctypes.c_double),
                                                         ("psd_launches", ctypes.c_uint64)]

    def add(self, other):
        for k, _ in self._fields_:
            setattr(self, k, getattr(self, k) + getattr(other, k))
        return self

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GolayError(RuntimeError):
    pass


_lib = None

# name -> (restype, argtypes); the complete exported surface of golay_gpu.h
SIGNATURES = {
    "gp_create": (ctypes.c_int, [_P(ctypes.c_void_p), ctypes.c_int]),
    "gp_destroy": (None, [ctypes.c_void_p]),
    "gp_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "gp_set_psd_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gp_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "gp_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _P(ctypes.c_int64)]),
    "gp_generate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, _P(GpCands), _P(GpCands), _P(GpStats)]),
    "gp_sort_candidates": (ctypes.c_int, [ctypes.c_void_p, _P(GpCands)]),
    "gp_match": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(GpCands),
                                _P(GpCands), _P(GpPairs), _P(GpStats)]),
    "gp_uncompress_side": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int8), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _P(GpCands), _P(GpStats)]),
    "gp_uncompress": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     _P(GpPairs), _P(GpPairs), _P(GpCands), _P(GpCands), _P(GpStats)]),
    "gp_filter": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P(GpPairs), _P(GpPairs),
                                 _P(GpStats)]),
    "gp_psd_filter_batch": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int8), ctypes.c_int, _P(ctypes.c_int32),
                                           _P(ctypes.c_int32), _P(ctypes.c_int32), ctypes.c_int64,
                                           _P(ctypes.c_uint8), _P(ctypes.c_float), _P(GpStats)]),
    "gp_psd_filter_batch_static": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p,
                                                  _P(GpStats)]),
    "gp_psd_microbench": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, _P(ctypes.c_int32),
                                         ctypes.c_int, _P(ctypes.c_uint64), _P(ctypes.c_uint64),
                                         _P(ctypes.c_uint8), ctypes.c_int64, _P(GpStats)]),
    "gp_psd_keys": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int8), ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                   ctypes.c_int, ctypes.c_int, _P(ctypes.c_uint8), _P(GpStats)]),
    "gp_microbench_words": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, _P(ctypes.c_int32), ctypes.c_int,
                                           ctypes.c_int64, ctypes.c_int64, _P(ctypes.c_int32), _P(ctypes.c_int8),
                                           ctypes.c_int]),
}


def load_library(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GolayError("libgolaypair_gpu.so not built: run `python -m golaypair_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a, ctype):
    return a.ctypes.data_as(_P(ctype))


class Context:
    """One gp_ctx on a HIP device."""

    def __init__(self, device=0, mode=GP_PSD_SPLIT_BF16):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.gp_create(ctypes.byref(h), device)
        if rc != GP_OK:
            raise GolayError("gp_create(%d) failed with %d (no GPU?)" % (device, rc))
        self.h = h
        self.set_mode(mode)

    def close(self):
        if self.h:
            self.lib.gp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what):
        if rc < 0:
            raise GolayError("%s failed (%d): %s" % (what, rc, self.lib.gp_last_error(self.h).decode()))
        return rc

    def set_mode(self, mode):
        self.check(self.lib.gp_set_psd_mode(self.h, mode), "gp_set_psd_mode")

    # -- PSD filter -----------------------------------------------------------
    def psd_filter_batch(self, words, lens, orders, pafs, want_max=False):
        words = np.ascontiguousarray(words, dtype=np.int8)
        count, stride = words.shape
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        orders = np.ascontiguousarray(orders, dtype=np.int32)
        pafs = np.ascontiguousarray(pafs, dtype=np.int32)
        pas = np.zeros(count, dtype=np.uint8)
        mx = np.zeros(count, dtype=np.float32) if want_max else None
        st = GpStats()
        rc = self.lib.gp_psd_filter_batch(self.h, ptr(words, ctypes.c_int8), stride, ptr(lens, ctypes.c_int32),
                                          ptr(orders, ctypes.c_int32), ptr(pafs, ctypes.c_int32), count,
                                          ptr(pas, ctypes.c_uint8),
                                          ptr(mx, ctypes.c_float) if want_max else None, ctypes.byref(st))
        self.check(rc, "gp_psd_filter_batch")
        return pas, mx, st

    def psd_microbench(self, seed, mix, pass_cap=0):
        mix = np.ascontiguousarray(mix, dtype=np.int32).reshape(-1, 5)
        count = int(mix[:, 4].sum())
        acc = np.zeros(len(mix), dtype=np.uint64)
        dig = ctypes.c_uint64()
        pas = np.zeros(max(pass_cap, 1), dtype=np.uint8)
        st = GpStats()
        rc = self.lib.gp_psd_microbench(self.h, seed, count, ptr(mix, ctypes.c_int32), len(mix),
                                        ptr(acc, ctypes.c_uint64), ctypes.byref(dig),
                                        ptr(pas, ctypes.c_uint8) if pass_cap else None, pass_cap,
                                        ctypes.byref(st))
        self.check(rc, "gp_psd_microbench")
        return acc, dig.value, pas[:pass_cap], st


class Cands:
    """Host candidate set: words (m, n) int8 and key digits (m, n//2) uint8."""

    def __init__(self, words, keys):
        self.words = np.ascontiguousarray(words, dtype=np.int8)
        self.keys = np.ascontiguousarray(keys, dtype=np.uint8)

    @property
    def n(self):
        return self.words.shape[1]

    def __len__(self):
        return self.words.shape[0]

    def struct(self):
        m, n = self.words.shape
        c = GpCands(n, n // 2, m, m, ptr(self.words, ctypes.c_int8), ptr(self.keys, ctypes.c_uint8))
        return c

    def lines(self):
        return [cand_line(w, k) for w, k in zip(self.words, self.keys)]


def cand_line(word, key):
    """Candidate line text without newline (io.cpp:14-36)."""
    n = len(word)
    K = n // 2
    digits = "".join(str(int(key[(k if k <= K else n - k) - 1])) for k in range(1, n))
    return "PSD" + digits + " " + "".join("%d " % v for v in word)


BINARY_MAGIC = b"GPCB"


def read_binary_cands(path):
    """A candidate file written with the CLI's --binary option: "GPCB", version
    1, n, K = n // 2, a zero byte, the record count (u64 little endian), then
    per record n int8 letters and K uint8 key values (the numbers printed after
    "PSD" in the text form).  Returns Cands; Cands.lines() gives the text form."""
    data = open(path, "rb").read()
    if len(data) < 16 or data[:4] != BINARY_MAGIC or data[4] != 1:
        raise ValueError("%s: not a binary candidate file" % path)
    n, K = data[5], data[6]
    m = int.from_bytes(data[8:16], "little")
    if K != n // 2 or len(data) != 16 + m * (n + K):
        raise ValueError("%s: bad header or size" % path)
    rec = np.frombuffer(data, np.uint8, offset=16).reshape(m, n + K)
    return Cands(rec[:, :n].view(np.int8), rec[:, n:])


def pair_line(a, b):
    """Pair line text without newline (io.cpp:75-95)."""
    return "".join("%d " % v for v in a) + " " + "".join("%d " % v for v in b)


def _empty_cands(n, cap):
    return Cands(np.zeros((cap, n), np.int8), np.zeros((cap, n // 2), np.uint8))


def _ctx_methods():
    def generate(self, order, compress, paf=0, job=0, njobs=1):
        n = order // compress
        cap = 1 << 16
        while True:
            a, b = _empty_cands(n, cap), _empty_cands(n, cap)
            sa, sb = a.struct(), b.struct()
            st = GpStats()
            rc = self.lib.gp_generate(self.h, order, compress, paf, job, njobs, ctypes.byref(sa), ctypes.byref(sb),
                                      ctypes.byref(st))
            if rc == GP_ERR_CAPACITY:
                cap = max(sa.count, sb.count) + 1
                continue
            self.check(rc, "gp_generate")
            if rc == GP_NO_INPUT:
                return None, None, st
            return Cands(a.words[:sa.count], a.keys[:sa.count]), Cands(b.words[:sb.count], b.keys[:sb.count]), st

    def sort_candidates(self, c):
        c = Cands(c.words.copy(), c.keys.copy())
        s = c.struct()
        self.check(self.lib.gp_sort_candidates(self.h, ctypes.byref(s)), "gp_sort_candidates")
        return Cands(c.words[:s.count], c.keys[:s.count])

    def match(self, order, compress, paf, a, b):
        n = order // compress
        cap = max(1024, 4 * (len(a) + len(b)))
        while True:
            pa = np.zeros((cap, n), np.int8)
            pb = np.zeros((cap, n), np.int8)
            out = GpPairs(n, 0, cap, ptr(pa, ctypes.c_int8), ptr(pb, ctypes.c_int8))
            sa, sb = a.struct(), b.struct()
            st = GpStats()
            rc = self.lib.gp_match(self.h, order, compress, paf, ctypes.byref(sa), ctypes.byref(sb),
                                   ctypes.byref(out), ctypes.byref(st))
            if rc == GP_ERR_CAPACITY:
                cap = out.count + 1
                continue
            self.check(rc, "gp_match")
            return rc, pa[:out.count], pb[:out.count], st

    def psd_keys(self, words, bound, inverted):
        words = np.ascontiguousarray(words, dtype=np.int8)
        m, n = words.shape
        keys = np.zeros((m, n // 2), np.uint8)
        self.check(self.lib.gp_psd_keys(self.h, ptr(words, ctypes.c_int8), n, n, m, bound, int(inverted),
                                        ptr(keys, ctypes.c_uint8), None), "gp_psd_keys")
        return keys

    def uncompress_side(self, orig, compress, new_compress, paf, seqflag, proc_id=0, proc_num=1):
        """Candidates of one sequence (job proc_id of proc_num) in DFS order;
        returns (cands, rotated orig, stats)."""
        o = np.ascontiguousarray(orig, dtype=np.int8).copy()
        N = len(o) * (compress // new_compress)
        cap = 1 << 12
        while True:
            out = _empty_cands(N, cap)
            s = out.struct()
            st = GpStats()
            o2 = o.copy()
            rc = self.lib.gp_uncompress_side(self.h, ptr(o2, ctypes.c_int8), len(o2), compress, new_compress, paf,
                                             proc_id, proc_num, seqflag, ctypes.byref(s), ctypes.byref(st))
            if rc == GP_ERR_CAPACITY:
                cap = s.count + 1
                continue
            self.check(rc, "gp_uncompress_side")
            return Cands(out.words[:s.count], out.keys[:s.count]), o2, st

    def uncompress(self, order, compress, new_compress, paf, pa, pb, want_tmp=False):
        """Uncompress pair arrays (m, order/compress); returns (a, b, tmp_a, tmp_b, stats)."""
        pa = np.ascontiguousarray(pa, dtype=np.int8)
        pb = np.ascontiguousarray(pb, dtype=np.int8)
        L = order // compress
        N = order // new_compress
        inp = GpPairs(L, len(pa), len(pa), ptr(pa, ctypes.c_int8), ptr(pb, ctypes.c_int8))
        cap, tcap = max(1024, 4 * len(pa)), 1 << 16
        while True:
            oa = np.zeros((cap, N), np.int8)
            ob = np.zeros((cap, N), np.int8)
            out = GpPairs(N, 0, cap, ptr(oa, ctypes.c_int8), ptr(ob, ctypes.c_int8))
            ta, tb = _empty_cands(N, tcap), _empty_cands(N, tcap)
            sa, sb = ta.struct(), tb.struct()
            st = GpStats()
            rc = self.lib.gp_uncompress(self.h, order, compress, new_compress, paf, ctypes.byref(inp),
                                        ctypes.byref(out), ctypes.byref(sa) if want_tmp else None,
                                        ctypes.byref(sb) if want_tmp else None, ctypes.byref(st))
            if rc == GP_ERR_CAPACITY:
                cap = max(cap, out.count + 1)
                tcap = max(tcap, sa.count + 1, sb.count + 1)
                continue
            self.check(rc, "gp_uncompress")
            tmp = (Cands(ta.words[:sa.count], ta.keys[:sa.count]), Cands(tb.words[:sb.count], tb.keys[:sb.count])) \
                if want_tmp else (None, None)
            return oa[:out.count], ob[:out.count], tmp[0], tmp[1], st

    def filter(self, order, compress, pa, pb):
        """Equivalence filter of pair arrays; returns (a, b, stats) in (a, b) order."""
        pa = np.ascontiguousarray(pa, dtype=np.int8)
        pb = np.ascontiguousarray(pb, dtype=np.int8)
        n = order // compress
        m = len(pa)
        oa = np.zeros((max(m, 1), n), np.int8)
        ob = np.zeros((max(m, 1), n), np.int8)
        inp = GpPairs(n, m, m, ptr(pa, ctypes.c_int8), ptr(pb, ctypes.c_int8))
        out = GpPairs(n, 0, max(m, 1), ptr(oa, ctypes.c_int8), ptr(ob, ctypes.c_int8))
        st = GpStats()
        self.check(self.lib.gp_filter(self.h, order, compress, ctypes.byref(inp), ctypes.byref(out),
                                      ctypes.byref(st)), "gp_filter")
        return oa[:out.count], ob[:out.count], st

    for f in (generate, sort_candidates, match, psd_keys, uncompress_side, uncompress, filter):
        setattr(Context, f.__name__, f)


_ctx_methods()


def microbench_words(seed, mix, first, rows, stride=64):
    """Host copy of the microbenchmark generator: (kinds, words) of candidates
    first..first+rows-1 (words zero padded to `stride`)."""
    lib = load_library()
    mix = np.ascontiguousarray(mix, dtype=np.int32).reshape(-1, 5)
    count = int(mix[:, 4].sum())
    kinds = np.zeros(rows, dtype=np.int32)
    out = np.zeros((rows, stride), dtype=np.int8)
    rc = lib.gp_microbench_words(seed, count, ptr(mix, ctypes.c_int32), len(mix), first, rows,
                                 ptr(kinds, ctypes.c_int32), ptr(out, ctypes.c_int8), stride)
    if rc != GP_OK:
        raise GolayError("gp_microbench_words failed")
    return kinds, out
