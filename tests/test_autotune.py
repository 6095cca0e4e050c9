"""Autotuner: the saved table round-trips through load/save (CPU); on the
GPU it picks a split at least as fast as the heuristic and transforms made
with the tuned plans stay bit-exact."""
import os

import numpy as np
import pytest

BB = 2013265921


def test_tune_table_load_save_roundtrip(nt, tmp_path):
    src = tmp_path / "t.txt"
    src.write_text("0 24 24 2 12 12 0 0 14 3\n0 20 20 3 7 7 6 0 12 0\n1 18 18 2 9 9 0 0 13 0\n"
                   "0 20 20 2 9 9 0 0 12 0\n")  # last line does not cover n: ignored
    nt.autotune_clear()
    nt.autotune_load(str(src))
    dst = tmp_path / "u.txt"
    nt.autotune_save(str(dst))
    lines = sorted(dst.read_text().split("\n"))
    assert [l for l in lines if l] == sorted(["0 24 24 2 12 12 0 0 14 3", "0 20 20 3 7 7 6 0 12 0",
                                             "1 18 18 2 9 9 0 0 13 0"])
    nt.autotune_clear()
    with pytest.raises(nt.NttError):
        nt.autotune_load(str(tmp_path / "missing.txt"))


def test_split_lookup_falls_back_to_nearest_batch_size(nt, tmp_path):
    """An exact (arith, log n, elements in flight) entry wins; otherwise the
    entry for the same n at the nearest load within a factor of 4; otherwise
    the heuristic."""
    nt.autotune_clear()
    heur = {(ln, le): nt.split_for(BB, ln, 1 << le) for ln, le in ((17, 17), (18, 15), (16, 24))}
    src = tmp_path / "t.txt"
    src.write_text("0 18 24 2 10 8 0 0 13 3\n0 18 20 2 9 9 0 0 12 0\n0 16 16 2 8 8 0 0 12 3\n")
    nt.autotune_load(str(src))
    assert nt.split_for(BB, 18, 1 << 24) == {"split": [10, 8], "log_e": 13, "mfma_mask": 3}
    assert nt.split_for(BB, 18, 1 << 20) == {"split": [9, 9], "log_e": 12, "mfma_mask": 0}
    assert nt.split_for(BB, 18, 1 << 18) == {"split": [9, 9], "log_e": 12, "mfma_mask": 0}   # nearest: 2^20
    assert nt.split_for(BB, 18, 1 << 23) == {"split": [10, 8], "log_e": 13, "mfma_mask": 3}  # nearest: 2^24
    assert nt.split_for(BB, 18, 1 << 15) == heur[(18, 15)]                                   # 2^20 is too far
    assert nt.split_for(BB, 16, 1 << 24) == heur[(16, 24)]                                   # 2^16 is too far
    assert nt.split_for(BB, 17, 1 << 17) == heur[(17, 17)]                                   # no entry for n
    assert nt.split_for(BB, 18, 1 << 28) == {"split": [10, 8], "log_e": 13, "mfma_mask": 3}  # beyond 2^24
    nt.autotune_clear()


def test_split_lookup_beyond_the_largest_tuned_load(nt, tmp_path):
    """A load above every tuned one takes the largest tuned load's entry when
    that load fills the chip (2^22 elements or more), else the heuristic."""
    nt.autotune_clear()
    heur = {le: nt.split_for(BB, 16, 1 << le) for le in (19, 27)}
    src = tmp_path / "t.txt"
    src.write_text("1 22 24 2 12 10 0 0 13 0\n0 16 16 2 8 8 0 0 12 3\n")
    nt.autotune_load(str(src))
    Q62 = 4611686018326724609
    tuned = {"split": [12, 10], "log_e": 13, "mfma_mask": 0}
    assert nt.split_for(Q62, 22, 1 << 26) == tuned     # within a factor of 4
    assert nt.split_for(Q62, 22, 1 << 27) == tuned     # beyond: the 2^24 load's plan
    assert nt.split_for(Q62, 22, 1 << 30) == tuned
    assert nt.split_for(BB, 16, 1 << 19) == heur[19]   # 2^16 is tuned only at a load that leaves CUs idle
    assert nt.split_for(BB, 16, 1 << 27) == heur[27]
    nt.autotune_clear()


@pytest.mark.gpu
def test_autotune_then_transform_exact(nt, orc):
    ctx = nt.Context()
    res = nt.autotune(ctx, BB, [12, 16, 20], 1 << 20, reps=3)
    for r in res:
        assert r["candidates"] >= 3 and sum(r["split"]) == r["log_n"]
        assert r["us"] <= r["heuristic_us"] * 1.05 + 1.0
    for log_n in (16, 20):
        x = orc.random_poly(1 << log_n, BB, 5)
        buf = nt.DeviceBuffer.from_array(x.astype(np.uint32))
        b = nt.Batch(ctx, [(buf.ptr, BB, log_n, nt.FLAG_NEGACYCLIC)] * 1)
        b.launch()
        nt.device_sync()
        assert np.array_equal(buf.download(np.uint32, 1 << log_n).astype(np.uint64), orc.ntt_negacyclic(x, BB))
        b.destroy()
    nt.autotune_clear()
    ctx.destroy()
