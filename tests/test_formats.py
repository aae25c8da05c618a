"""Host-side file formats (no GPU): the compact binary candidate form read by
golaypair_amd.read_binary_cands gives the reference's text lines (io.cpp:14-36)."""
import struct

import numpy as np
import pytest

import golaypair_amd as g


def _write(path, n, words, keys, version=1):
    m = len(words)
    hdr = g.BINARY_MAGIC + bytes([version, n, n // 2, 0]) + struct.pack("<Q", m)
    body = b"".join(np.asarray(w, np.int8).tobytes() + np.asarray(k, np.uint8).tobytes() for w, k in zip(words, keys))
    open(path, "wb").write(hdr + body)


def test_binary_cands_round_trip(tmp_path):
    # n = 5, K = 2: key digits d1 d2 d2 d1 (PSD symmetry), then the letters
    words = [[1, -1, 3, -3, 1], [-1, -1, -1, 1, 1]]
    keys = [[12, 7], [0, 255]]
    p = tmp_path / "c.bin"
    _write(p, 5, words, keys)
    c = g.read_binary_cands(p)
    assert c.lines() == ["PSD127712 1 -1 3 -3 1 ", "PSD02552550 -1 -1 -1 1 1 "]
    assert (c.words == np.array(words, np.int8)).all() and (c.keys == np.array(keys, np.uint8)).all()


def test_binary_cands_empty_and_bad(tmp_path):
    p = tmp_path / "e.bin"
    _write(p, 8, [], [])
    assert len(g.read_binary_cands(p)) == 0
    _write(p, 8, [[0] * 8], [[1, 2, 3, 4]], version=2)
    with pytest.raises(ValueError):
        g.read_binary_cands(p)
    open(p, "wb").write(b"PSD1 2 3 \n")
    with pytest.raises(ValueError):
        g.read_binary_cands(p)
