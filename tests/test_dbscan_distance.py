"""The peak-clustering kernel's neighbour test against the reference's.

The reference tests sqrt(dp * dp + di * di) < eps with di = 0
(dbscan, U-CUDA/cudaLibrary.cu:2088-2166; oracle/ucuda_oracle.c restates
it). The wave-per-cell kernel tests |dp| < eps when 2^-500 <= eps <= 2^500:
in binary floating point with rounding to nearest sqrt(fl(x * x)) = |x| as
long as x * x neither underflows nor overflows, and inside that eps range
the under- and overflowing squares give the same verdict as |x|. numpy's
sqrt and multiply are the correctly rounded IEEE operations the oracle's C
uses, so equal verdicts here are equal verdicts against the oracle.
"""
import numpy as np

LO, HI = 2.0 ** -500, 2.0 ** 500


def _distances(rng):
    """|dp| over the whole double range: random significands at every
    exponent (subnormals included), the edges around 2^-511 and 2^512 where
    the square leaves the normal range, powers of two and their neighbours,
    zeros, infinities and NaN, with both signs."""
    e = rng.integers(-1074, 1024, 4_000_000)
    x = np.ldexp(rng.uniform(1.0, 2.0, e.size), e)
    edges = np.ldexp(1.0, np.arange(-1074, 1024)).astype(np.float64)
    special = [0.0, np.inf, np.nan, np.finfo(np.float64).max, np.finfo(np.float64).tiny, 5e-324]
    near = [np.nextafter(edges, 0.0), edges, np.nextafter(edges, np.inf)]
    band = [np.ldexp(rng.uniform(1.0, 2.0, 200_000), rng.integers(lo, hi, 200_000)) for lo, hi in ((-515, -505),
                                                                                               (505, 515))]
    x = np.concatenate([x, *near, np.array(special), *band])
    return np.concatenate([x, -x])


def _eps_values(rng):
    inside = np.ldexp(rng.uniform(1.0, 2.0, 64), rng.integers(-500, 500, 64))
    return np.concatenate([[LO, HI, np.nextafter(LO, np.inf), np.nextafter(HI, 0.0), 0.1, 1e-3, 1.0, 1e150,
                            1e-150], inside])


def test_abs_equals_sqrt_of_square_for_eps_in_range():
    rng = np.random.default_rng(7)
    dp = _distances(rng)
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        ref_dist = np.sqrt(dp * dp + 0.0 * 0.0)
        for eps in _eps_values(rng):
            ref = ref_dist < eps
            got = np.abs(dp) < eps
            bad = np.flatnonzero(ref != got)
            assert bad.size == 0, (eps, dp[bad[:5]])


def test_the_range_is_needed():
    """Outside [2^-500, 2^500] the two tests part: just below eps = 1e160
    the squares overflow, so the kernel keeps the sqrt form there."""
    eps = 1e160
    dp = eps * (1.0 - np.linspace(1e-9, 1e-3, 1000))
    with np.errstate(over="ignore"):
        assert ((np.sqrt(dp * dp) < eps) != (np.abs(dp) < eps)).all()
    assert not (LO <= eps <= HI)
