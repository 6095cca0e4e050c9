"""The vectorised input generator and prime sets used by tests and bench.py
agree with the oracle's sequential generator and prime search."""
import numpy as np

import workloads as W


def test_generator_matches_oracle(orc):
    for q in (W.BABYBEAR, 1056440321, 4611686018326724609, 3):
        for seed in (0, 42, 0x5EED ^ 7):
            assert np.array_equal(W.random_poly(5000, q, seed), orc.random_poly(5000, q, seed)), (q, seed)


def test_prime_sets_match_oracle(orc):
    assert W.rns_primes() == orc.ntt_primes(1 << 31, 17, 24)
    assert W.p30_primes() == [p for p in orc.ntt_primes(1 << 30, 19, 200) if p > (1 << 29)]
    assert W.p62_primes() == orc.ntt_primes(1 << 62, 23, 6)


def test_uneven_batch_shape():
    jobs = W.uneven_batch()
    assert len(jobs) == 512
    assert {j["log_n"] for j in jobs} == set(range(10, 19))
    for j in jobs:
        assert (j["q"] - 1) % (1 << (j["log_n"] + 1)) == 0
    assert W.uneven_batch() == jobs  # deterministic


def test_serving_requests_shape():
    reqs = W.serving_requests(200, 1000.0)
    ts = [t for t, _ in reqs]
    assert ts == sorted(ts) and 0.1 < ts[-1] < 0.4
    for _, polys in reqs:
        assert 1 <= len(polys) <= 32
        for p in polys:
            assert 12 <= p["log_n"] <= 22 and (p["q"] - 1) % (1 << (p["log_n"] + 1)) == 0
