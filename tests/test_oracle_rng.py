"""Pins the oracle's Gaussian stream (Omega) against independent libraries.

Philox4x64-10 against numpy's Philox bit generator (numpy increments the
counter before generating, so numpy counter c yields block c+1); the inverse
normal CDF against scipy.special.ndtri; and the element -> (block, lane)
mapping and float rounding of the stream against a numpy re-derivation.
"""
import numpy as np
import pytest
from scipy.special import ndtri

from oracle import ref


def numpy_philox_block(key0: int, block: int) -> list[int]:
    ctr = np.array([(block - 1) % 2**64, 0, 0, 0], dtype=np.uint64)
    if block == 0:
        ctr = np.array([2**64 - 1, 2**64 - 1, 2**64 - 1, 2**64 - 1], dtype=np.uint64)
    bg = np.random.Philox(key=np.array([key0, 0], dtype=np.uint64), counter=ctr)
    return [int(x) for x in bg.random_raw(4)]


@pytest.mark.parametrize("key0", [0, 1, 777, 2**63 + 12345])
@pytest.mark.parametrize("block", [0, 1, 2, 1000, 2**40 + 7])
def test_philox_matches_numpy(key0, block):
    assert ref.philox4x64([block, 0, 0, 0], [key0, 0]) == numpy_philox_block(key0, block)


def test_philox_counter_words_matter():
    a = ref.philox4x64([5, 0, 0, 0], [9, 0])
    b = ref.philox4x64([5, 1, 0, 0], [9, 0])
    c = ref.philox4x64([5, 0, 0, 0], [9, 1])
    assert a != b and a != c and b != c


def test_ndtri_matches_scipy():
    rng = np.random.default_rng(0)
    us = np.concatenate([
        rng.random(20000),
        np.logspace(-300, -1, 400),
        1.0 - np.logspace(-16, -1, 200),
        np.array([0.5, 0.5 + 1e-17, 0.075, 0.925, 0.425, 1e-10]),
    ])
    mine = np.array([ref.ndtri(float(u)) for u in us])
    want = ndtri(us)
    err = np.abs(mine - want) / np.maximum(np.abs(want), 1e-300)
    assert err.max() < 5e-15, (err.max(), us[err.argmax()])


def test_ndtri_symmetry_and_known_values():
    assert ref.ndtri(0.5) == 0.0
    assert abs(ref.ndtri(0.975) - 1.959963984540054) < 1e-15
    assert abs(ref.ndtri(0.8413447460685429) - 1.0) < 1e-14
    # p and 1-p both exact in binary: the two branches must mirror
    for p in [2.0**-40, 2.0**-20, 0.0078125, 0.125, 0.3125]:
        assert abs(ref.ndtri(p) + ref.ndtri(1.0 - p)) <= 4e-16 * abs(ref.ndtri(p))


def stream_numpy(seed: int, offset: int, count: int) -> np.ndarray:
    out = np.empty(count)
    for i in range(count):
        e = offset + i
        w = numpy_philox_block(seed, e >> 2)[e & 3]
        u = ((w >> 11) + 0.5) * 2.0**-53
        out[i] = np.float64(np.float32(ndtri(u)))
    return out


@pytest.mark.parametrize("seed,offset,count", [(777, 0, 64), (1, 3, 37), (2**40, 10**9, 19)])
def test_stream_mapping(seed, offset, count):
    got = ref.gaussian(seed, offset, count)
    want = stream_numpy(seed, offset, count)
    # float rounding of an ndtri that differs by <= 1 ulp (double) can only
    # flip a float rounding at an exact tie; none occur in these draws
    assert np.array_equal(got, want)


def test_stream_is_float_valued_and_gaussian():
    x = ref.gaussian(12345, 0, 200000)
    assert np.array_equal(x, x.astype(np.float32).astype(np.float64))
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1.0) < 0.01
    # fourth moment of N(0,1) is 3
    assert abs(np.mean(x**4) - 3.0) < 0.1
