"""Committed golden vectors (tests/golden/ntt_vectors.json): the oracle still
reproduces them (CPU), and the GPU library matches them bit-exactly."""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ntt_vectors.json")))


def test_oracle_reproduces_golden(orc):
    for c in GOLD["transforms"]:
        x = np.array(c["input"], dtype=np.uint64)
        got = orc.transform(x, c["q"], bool(c["negacyclic"]), bool(c["inverse"]))
        assert list(map(int, got)) == c["output"], (c["q"], c["n"])
    for p in GOLD["negacyclic_products"]:
        got = orc.negacyclic_schoolbook(np.array(p["a"], dtype=np.uint64), np.array(p["b"], dtype=np.uint64), p["q"])
        assert list(map(int, got)) == p["product"]


def test_golden_ramp_kat():
    c = GOLD["transforms"][0]
    assert c["input"][:3] == [1, 2, 3] and c["output"][:5] == [2080, 945746745, 44582636, 962587333, 472683146]


@pytest.mark.gpu
def test_gpu_matches_golden(nt):
    ctx = nt.Context()
    for c in GOLD["transforms"]:
        n = c["n"]
        log_n = n.bit_length() - 1
        flags = (nt.FLAG_NEGACYCLIC if c["negacyclic"] else 0) | (nt.FLAG_INVERSE if c["inverse"] else 0)
        for u64 in (False, True):
            if not u64 and c["q"] >= (1 << 31):
                continue
            dt = np.uint64 if u64 else np.uint32
            buf = nt.DeviceBuffer.from_array(np.array(c["input"], dtype=dt))
            ctx.transform(buf.ptr, c["q"], log_n, flags | (nt.FLAG_U64 if u64 else 0))
            nt.device_sync()
            assert list(map(int, buf.download(dt, n))) == c["output"], (c["q"], n, flags, u64)
            buf.free()
    ctx.destroy()
    for p in GOLD["negacyclic_products"]:
        n = p["n"]
        cfg = nt.Config(n, p["q"])
        a = nt.DeviceBuffer.from_array(np.array(p["a"], dtype=np.uint64))
        b = nt.DeviceBuffer.from_array(np.array(p["b"], dtype=np.uint64))
        r = nt.DeviceBuffer(n * 8)
        cfg.negacyclic_multiply(r, a, b)
        nt.device_sync()
        assert list(map(int, r.download(np.uint64, n))) == p["product"]
        cfg.cleanup()
