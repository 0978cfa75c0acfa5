"""Pins the oracle's restatement of thredholdHistMaterialAccCPU
(kernel.cu:8380-8418), the curl-histogram threshold (CPU only).

The reference holds no fixture for this function (its one call,
kernel.cu:8827, is commented out), so it is pinned by closed forms:

  * values on bin centres between a minimum of 0 and a maximum of 1: the
    threshold is k / 10000 for the bin k of the value of rank ceil(p V) from
    the top (a bottom-up walk, an upward walk from N-1 or an off-by-one bin
    fails it);
  * the 1e4 / -1e4 seeds of min / max: a constant field of 2e4 has min 1e4,
    so every value falls in the top bin;
  * a constant field (max == min) puts every voxel in bin 0 and returns min;
  * percentage above 1 is never reached and returns the maximum; percentage
    0 returns the top bin's lower edge;
  * the quantile property on random data: the count of values at or above
    the returned edge reaches p V, the count above the next edge does not.
"""
import numpy as np
import pytest

import oracle as O

N = 10000


def centred_field(bins, shape, rng):
    """0 and 1 plus values at the centres of the given bins, shuffled."""
    v = np.concatenate([[0.0, 1.0], (np.asarray(bins) + 0.5) / N]).astype(np.float32)
    assert v.size == np.prod(shape)
    rng.shuffle(v)
    return v.reshape(shape)


@pytest.mark.parametrize("p", [0.001, 0.1, 0.25, 0.5, 0.9, 0.999])
def test_bin_centres_give_the_rank_bin(p):
    rng = np.random.default_rng(3)
    shape = (10, 20, 25)
    bins = rng.integers(1, N - 2, size=int(np.prod(shape)) - 2)
    c = centred_field(bins, shape, rng)
    V = c.size
    need = np.float32(V) * np.float32(p)
    # bins of every value in descending order: 1 -> N-1, 0 -> 0
    allbins = np.sort(np.concatenate([[N - 1, 0], bins]))[::-1]
    rank = int(np.ceil(need))  # the running count first reaches need here
    k = allbins[rank - 1]
    assert O.threshold_hist(c, p) == np.float32(k) / np.float32(N)


def test_seeds_of_min_and_max():
    c = np.full((4, 5, 6), 2e4, np.float32)
    # min stays at its 1e4 seed, max = 2e4: every value in the top bin
    expect = np.float32(N - 1) / np.float32(N) * np.float32(1e4) + np.float32(1e4)
    assert O.threshold_hist(c, 0.5) == expect
    c = np.full((4, 5, 6), -3e4, np.float32)
    # max stays at -1e4, min = -3e4: every value in bin 0 -> the minimum
    assert O.threshold_hist(c, 0.5) == np.float32(-3e4)


def test_constant_field_returns_min():
    c = np.full((3, 4, 5), 0.75, np.float32)
    assert O.threshold_hist(c, 0.3) == np.float32(0.75)


def test_percentage_limits():
    rng = np.random.default_rng(5)
    c = rng.random((8, 9, 10)).astype(np.float32)
    lo, hi = np.float32(c.min()), np.float32(c.max())
    assert O.threshold_hist(c, 1.5) == hi
    top = np.float32(N - 1) / np.float32(N) * (hi - lo) + lo
    assert O.threshold_hist(c, 0.0) == top


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quantile_property(seed):
    rng = np.random.default_rng(seed)
    c = np.abs(rng.normal(size=(16, 17, 18))).astype(np.float32) * np.float32(3.0)
    V = c.size
    lo, hi = c.min(), c.max()
    for p in (0.01, 0.2, 0.7):
        t = O.threshold_hist(c, p)
        step = (float(hi) - float(lo)) / N
        # bin edges in float64: the count at or above the returned edge
        # reaches p V, above the next edge it does not (slack of 1% of a bin at
        # the edges, where the FP32 bin index may round across)
        assert np.count_nonzero(c >= t - 1e-2 * step) >= p * V
        assert np.count_nonzero(c >= t + step + 1e-2 * step) < p * V
