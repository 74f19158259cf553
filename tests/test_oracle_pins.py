"""Pins of the CPU oracle against independent implementations and known answers.

The oracle restates the reference; these tests check it against what is fixed
independently of it: numpy's MT19937 (seeding and its published test sets),
scipy's inverse normal CDF, numpy's symmetric eigensolver, closed forms.
"""
import math
import os

import numpy as np
import pytest
import scipy.special

import _oracle as O

NUMPY_MT_DATA = os.path.join(os.path.dirname(np.__file__), "random", "tests", "data")


def test_mt_seeding_matches_numpy_init_by_array():
    # commons-math3 setSeed(long 3141) -> init_by_array({0, 3141}); numpy seeds RandomState(array) the same way
    state = np.empty(624, dtype=np.uint32)
    O.lib().orc_mt_seeded_state(3141, state.ctypes.data_as(O.POINTER(O.c_uint32)))
    np_state = np.random.RandomState(np.array([0, 3141], dtype=np.uint32)).get_state()[1]
    assert np.array_equal(state, np_state)


@pytest.mark.parametrize("seed", [3141, 1352, 0, 7, -5, 2**31 - 1])
def test_mt_words_match_numpy(seed):
    words = O.mt_words(seed, 0, 5000)
    u = seed & 0xFFFFFFFFFFFFFFFF
    key = np.array([u >> 32, u & 0xFFFFFFFF], dtype=np.uint32)
    bg = np.random.MT19937()
    bg._legacy_seeding(key)
    assert np.array_equal(words, bg.random_raw(5000).astype(np.uint32))


@pytest.mark.parametrize("name", ["mt19937-testset-1.csv", "mt19937-testset-2.csv"])
def test_mt_twist_and_temper_replay_numpy_test_sets(name):
    # numpy's published MT19937 known answers: state = SeedSequence words with key[0] = 0x80000000, pos = 623
    path = os.path.join(NUMPY_MT_DATA, name)
    if not os.path.exists(path):
        pytest.skip("numpy test data not installed")
    rows = open(path).read().split("\n")
    seed = int(rows[0].split(",")[1], 0)
    ref = np.array([int(r.split(",")[1], 0) for r in rows[1:1001] if r], dtype=np.uint64).astype(np.uint32)
    st = np.random.SeedSequence(seed).generate_state(624, np.uint32)
    st[0] = 0x80000000
    out = np.empty(ref.size, dtype=np.uint32)
    O.lib().orc_mt_words_from_state(st.ctypes.data_as(O.POINTER(O.c_uint32)), 623, ref.size,
                                    out.ctypes.data_as(O.POINTER(O.c_uint32)))
    assert np.array_equal(out, ref)


def test_next_double_composition():
    # nextDouble = ((w0 >>> 6) << 26 | (w1 >>> 6)) * 2^-52 from consecutive words
    w = O.mt_words(3141, 0, 200).astype(np.uint64)
    expect = (((w[0::2] >> 6) << 26) | (w[1::2] >> 6)).astype(np.float64) * 2.0**-52
    assert np.array_equal(O.mt_doubles(3141, 0, 100), expect)
    assert np.array_equal(O.mt_doubles(3141, 37, 10), expect[37:47])


def test_icdf_against_scipy():
    ps = np.concatenate([np.geomspace(1e-300, 1e-3, 400), np.linspace(1e-3, 1 - 1e-3, 4000),
                         1 - np.geomspace(1e-3, 1e-16, 400)])
    v = O.icdf(ps)
    r = scipy.special.ndtri(ps)
    assert np.max(np.abs(v - r) / np.maximum(np.abs(r), 1e-300)) < 2e-15


def test_icdf_special_cases():
    assert O.lib().orc_icdf(0.5) == 0.0
    assert O.lib().orc_icdf(0.0) == 0.0        # reference quirk: p <= 0 returns 0, not -inf
    assert O.lib().orc_icdf(1.0) == 0.0
    ps = np.linspace(0.01, 0.99, 99)
    assert np.allclose(O.icdf(1 - ps), -O.icdf(ps), rtol=0, atol=1e-14)


def test_brownian_layout_and_scaling():
    times = np.arange(5) * 0.5
    dw = O.brownian(times, 3, 7, 3141)            # [nT][F][P]
    u = O.mt_doubles(3141, 0, 7 * 4 * 3).reshape(7, 4, 3)   # path, time, factor
    expect = np.transpose(O.icdf(u).reshape(7, 4, 3), (1, 2, 0)) * np.sqrt(0.5)
    assert np.array_equal(dw, expect)


def test_eigen_against_numpy():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((12, 12))
    a = a @ a.T
    ev, V = O.eigen_sym(a)
    order = np.argsort(ev)
    ref_ev, ref_V = np.linalg.eigh(a)
    assert np.allclose(ev[order], ref_ev, rtol=1e-12, atol=1e-12)
    for c in range(12):
        v, r = V[:, order[c]], ref_V[:, c]
        assert min(np.max(np.abs(v - r)), np.max(np.abs(v + r))) < 1e-10


def _factor_reduction_numpy(corr, m):
    """Textbook PCA with the reference's conventions (independent of the oracle's eigensolver)."""
    def fm(c):
        ev, V = np.linalg.eigh(c)
        idx = np.argsort(-ev, kind="stable")[:m]
        out = np.empty((c.shape[0], m))
        for f, i in enumerate(idx):
            sign = 1.0 if V[0, i] > 0 else -1.0
            out[:, f] = sign * np.sqrt(max(ev[i], 0.0) / np.sum(V[:, i] ** 2)) * V[:, i]
        return out
    f = fm(corr)
    f = f / np.sqrt(np.sum(f * f, axis=1))[:, None]
    return fm(f @ f.T)


@pytest.mark.parametrize("n,m,a", [(20, 3, 0.1), (40, 5, 0.05), (80, 10, 0.05)])
def test_factor_reduction_against_numpy(n, m, a):
    T = np.arange(n) * 0.5
    corr = np.exp(-a * np.abs(T[:, None] - T[None, :]))
    f = O.factor_reduction(corr, m)
    ref = _factor_reduction_numpy(corr, m)
    assert np.max(np.abs(f - ref)) < 1e-10
    # rows of the reduced factor matrix have (nearly) unit length; first entries positive
    assert np.allclose(np.sum(f * f, axis=1), 1.0, atol=1e-10)
    assert np.all(f[0, :] > 0)


def test_correlation_model_full_rank_reproduces_matrix():
    T = np.arange(6) * 0.5
    fac, corr = O.corr_expdecay(T, 0.1, 6)
    rho = np.exp(-0.1 * np.abs(T[:, None] - T[None, :]))
    assert np.allclose(fac @ fac.T, rho, atol=1e-12)
    assert np.allclose(np.diag(corr), 1.0)


def test_vol_fourparam_closed_form():
    times = np.arange(5) * 0.5
    T = np.arange(4) * 0.5
    v = O.vol_fourparam(times, T, 0.2, 0.05, 0.25, 0.1)
    for i, t in enumerate(times[:-1]):
        for j, Tj in enumerate(T):
            tau = Tj - t
            expect = 0.0 if tau <= 0 else (0.05 * tau + 0.2) * math.exp(-0.25 * tau) + 0.1
            assert abs(v[i, j] - expect) <= 2e-16 * abs(expect)


def test_discount_curve_closed_form():
    nt = np.arange(20) * 0.5
    nf = np.full(20, 0.03)
    for k in range(21):
        assert abs(O.df_from_forward_curve(nt, nf, 0.5, 0.5 * k) - (1 + 0.015) ** (-k)) < 1e-15


def test_round_time_survives_half_years():
    for k in range(200):
        assert O.lib().orc_round_time(0.5 * k) == 0.5 * k



def test_reference_quantile_test():
    # RandomVariableTest.testGetQuantile (src/test/java/net/finmath/montecarlo/RandomVariableTest.java:144-172):
    # commons-math3's MersenneTwister built from the int seed 3141 (init_genrand seeding, as
    # numpy's legacy RandomState(3141) seeds), 26+26-bit nextDouble, the inverse normal, and the
    # quantile rule of RandomVariableFromDoubleArray.getQuantile (:446-460)
    n = 10_000_000
    w = np.random.RandomState(3141).randint(0, 2**32, size=2 * n, dtype=np.uint64).astype(np.uint64)
    u = (((w[0::2] >> np.uint64(6)) << np.uint64(26)) | (w[1::2] >> np.uint64(6))).astype(np.float64) * 2.0**-52
    samples = np.sort(O.icdf(u))

    def quantile(q):
        return samples[min(max(int(math.floor((n + 1) * q - 1 + 0.5)), 0), n - 1)]

    assert quantile(0.0) == samples[0]
    assert abs(quantile(0.05) - (-1.645)) < 1e-3
    assert abs(quantile(0.5)) < 2e-4
    assert abs(quantile(0.95) - 1.645) < 1e-3
    assert abs(quantile(0.99) - 2.33) < 1e-2
