"""The sampled defining-sum check of tests/props.py against the oracle on
the CPU: it accepts the oracle's outputs in all four directions and rejects
an output with one coefficient changed."""
import numpy as np
import pytest

from props import sampled_sums_ok

BB = 2013265921


@pytest.mark.parametrize("nega", [False, True])
@pytest.mark.parametrize("inv", [False, True])
@pytest.mark.parametrize("log_n", [1, 4, 11])
def test_sampled_sums_accept_oracle_and_reject_changes(orc, nega, inv, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, BB, 31 + log_n)
    y = orc.transform(x, BB, nega, inv)
    g = orc.generator(BB)
    ks = sorted({0, 1 % n, n // 3, n - 1})
    assert sampled_sums_ok(y, x, BB, g, nega, inv, ks=ks)
    for k in ks:
        bad = y.copy()
        bad[k] = (bad[k] + 1) % BB
        assert not sampled_sums_ok(bad, x, BB, g, nega, inv, ks=ks), k
    # the neighbouring direction's output is not accepted either
    assert not sampled_sums_ok(orc.transform(x, BB, nega, not inv), x, BB, g, nega, inv, ks=ks) or n <= 2
