"""ctypes binding of the oracle (oracle/refcpu*.cpp), for tests only.

The oracle is the CPU restatement of the reference's algorithm; it is the
checker, never the code under test.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "build", "librefcpu.so")
CLI = os.path.join(ORACLE_DIR, "build", "refcpu")
_P = ctypes.POINTER
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j4"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not (os.path.exists(LIB) and os.path.exists(CLI)):
        build()
    lib = ctypes.CDLL(LIB)
    lib.ref_psd.argtypes = [_P(ctypes.c_int32), ctypes.c_int, _P(ctypes.c_double)]
    lib.ref_psd_filter_batch.argtypes = [_P(ctypes.c_int8), ctypes.c_int, _P(ctypes.c_int32), _P(ctypes.c_int32),
                                         _P(ctypes.c_int32), ctypes.c_int64, _P(ctypes.c_uint8),
                                         _P(ctypes.c_double), _P(ctypes.c_double), ctypes.c_int]
    lib.ref_psd_filter_batch_timed.restype = ctypes.c_double
    lib.ref_psd_filter_batch_timed.argtypes = [_P(ctypes.c_int8), ctypes.c_int, _P(ctypes.c_int32),
                                               _P(ctypes.c_int32), _P(ctypes.c_int32), ctypes.c_int64,
                                               _P(ctypes.c_uint8), ctypes.c_int]
    lib.ref_is_canonical.argtypes = [_P(ctypes.c_int32), ctypes.c_int, ctypes.c_int]
    lib.ref_group_size.restype = ctypes.c_int64
    lib.ref_group_size.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.ref_group_perms.restype = ctypes.c_int64
    lib.ref_group_perms.argtypes = [ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32), ctypes.c_int64]
    lib.ref_pair_group.restype = ctypes.c_int64
    lib.ref_pair_group.argtypes = [ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32), ctypes.c_int64]
    lib.ref_class_pairs.restype = ctypes.c_int64
    lib.ref_class_pairs.argtypes = [ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32), _P(ctypes.c_int32),
                                    _P(ctypes.c_int32), ctypes.c_int64]
    lib.ref_check_if_pair.argtypes = [_P(ctypes.c_int32), _P(ctypes.c_int32), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
    lib.ref_generate.restype = ctypes.c_int64
    lib.ref_generate.argtypes = [ctypes.c_int] * 5 + [ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p,
                                                      ctypes.c_int64, _P(ctypes.c_int64)]
    lib.ref_uncompress_side.restype = ctypes.c_int64
    lib.ref_uncompress_side.argtypes = [_P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64,
                                        _P(ctypes.c_uint64)]
    lib.ref_tie_band.restype = ctypes.c_uint64
    _lib = lib
    return lib


def i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_P(ctypes.c_int32))


def psd(seq):
    lib = load()
    s, p = i32(seq)
    out = np.zeros(len(s), dtype=np.float64)
    lib.ref_psd(p, len(s), out.ctypes.data_as(_P(ctypes.c_double)))
    return out


def psd_filter_batch(words, lens, orders, pafs):
    lib = load()
    words = np.ascontiguousarray(words, dtype=np.int8)
    count, stride = words.shape
    lens, pl = i32(lens)
    orders, po = i32(orders)
    pafs, pp = i32(pafs)
    pas = np.zeros(count, dtype=np.uint8)
    mx = np.zeros(count, dtype=np.float64)
    lib.ref_psd_filter_batch(words.ctypes.data_as(_P(ctypes.c_int8)), stride, pl, po, pp, count,
                             pas.ctypes.data_as(_P(ctypes.c_uint8)), mx.ctypes.data_as(_P(ctypes.c_double)),
                             None, 0)
    return pas, mx


def is_canonical(seq, flag):
    s, p = i32(seq)
    return bool(load().ref_is_canonical(p, len(s), flag))


def group_perms(flag, n):
    lib = load()
    m = lib.ref_group_size(flag, n)
    out = np.zeros((m, n), dtype=np.int32)
    lib.ref_group_perms(flag, n, out.ctypes.data_as(_P(ctypes.c_int32)), m)
    return out


def pair_group(n, alt):
    lib = load()
    m = lib.ref_pair_group(n, int(alt), None, 0)
    out = np.zeros((m, 2 * n), dtype=np.int32)
    lib.ref_pair_group(n, int(alt), out.ctypes.data_as(_P(ctypes.c_int32)), m)
    return out


def class_pairs(n, alt, a, b):
    lib = load()
    a, pa = i32(a)
    b, pb = i32(b)
    m = lib.ref_class_pairs(n, int(alt), pa, pb, None, 0)
    out = np.zeros((m, 2 * n), dtype=np.int32)
    lib.ref_class_pairs(n, int(alt), pa, pb, out.ctypes.data_as(_P(ctypes.c_int32)), m)
    return out


def check_if_pair(a, b, paf, c):
    a, pa = i32(a)
    b, pb = i32(b)
    return bool(load().ref_check_if_pair(pa, pb, len(a), paf, c))


def generate(order, c, paf, job=0, njobs=1):
    lib = load()
    sizes = (ctypes.c_int64 * 3)()
    rc = lib.ref_generate(order, c, paf, job, njobs, None, 0, None, 0, sizes)
    ba = ctypes.create_string_buffer(sizes[0] + 1)
    bb = ctypes.create_string_buffer(sizes[1] + 1)
    rc = lib.ref_generate(order, c, paf, job, njobs, ba, sizes[0] + 1, bb, sizes[1] + 1, sizes)
    return rc, ba.value.decode(), bb.value.decode(), sizes[2]


def uncompress_side(orig, c, c2, paf, seqflag, proc_id=0, proc_num=1):
    lib = load()
    o, po = i32(orig)
    leaves = ctypes.c_uint64()
    cap = 1 << 24  # one pass when the text fits; the call reports the size it needs
    buf = ctypes.create_string_buffer(cap)
    n = lib.ref_uncompress_side(po, len(o), c, c2, paf, seqflag, proc_id, proc_num, buf, cap, ctypes.byref(leaves))
    if n + 1 > cap:
        o, po = i32(orig)
        buf = ctypes.create_string_buffer(n + 1)
        lib.ref_uncompress_side(po, len(o), c, c2, paf, seqflag, proc_id, proc_num, buf, n + 1,
                                ctypes.byref(leaves))
    return buf.value.decode(), leaves.value
