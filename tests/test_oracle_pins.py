"""Pins of the CPU oracle against known answers and closed forms (no GPU).

PCG32: O'Neill's published pcg32-demo output (seed 42, stream 54).  Samplers and Green's
functions: distributions and integrals computed here independently (quadrature, KS tests),
not retyped formulas.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pcg32_known_answers(oracle_lib):
    # pcg32_srandom_r(&rng, 42u, 54u); first outputs of pcg32_random_r (pcg-c-basic pcg32-demo)
    with open(os.path.join(GOLDEN, "pcg32_demo_42_54.json")) as f:
        kat = json.load(f)
    want = [int(v, 16) for v in kat["u32"]]
    assert want[0] == 0xA15C02B7 and len(want) == 6
    out = oracle_lib.pcg32_u32(kat["seed"], kat["stream"], len(want))
    assert [int(v) for v in out] == want


def test_pcg32_float_is_mantissa_construction(oracle_lib):
    u = oracle_lib.pcg32_u32(42, 54, 1000)
    f = oracle_lib.pcg32_f32(42, 54, 1000)
    # nextFloat = float with mantissa (u >> 9) in [1, 2), minus 1: exactly (u >> 9) / 2^23
    np.testing.assert_array_equal(f, ((u >> 9).astype(np.float64) / 2 ** 23).astype(np.float32))
    assert f.min() >= 0.0 and f.max() < 1.0


def test_pcg32_streams_are_distinct_and_uniform(oracle_lib):
    a = oracle_lib.pcg32_u32(7, 1, 4096)
    b = oracle_lib.pcg32_u32(7, 2, 4096)
    assert (a != b).mean() > 0.99
    f = oracle_lib.pcg32_f32(7, 3, 200_000)
    assert stats.kstest(f, "uniform").pvalue > 1e-4


@pytest.mark.parametrize("dim", [2, 3])
def test_unit_sphere_sampler_moments(oracle_lib, dim):
    d = oracle_lib.sample_unit_sphere(dim, 5, 400_000).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    assert np.abs(d.mean(0)).max() < 5e-3
    np.testing.assert_allclose(d.T @ d / len(d), np.eye(dim) / dim, atol=5e-3)
    if dim == 3:  # Archimedes: z uniform on [-1, 1]
        assert stats.kstest(d[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-4


def test_harmonic_ball_radius_sampler_3d(oracle_lib):
    R = 0.7
    r, _ = oracle_lib.sample_ball_radius(3, R, 0.0, 9, 200_000)
    s = r / R
    # pdf 6 s (1 - s) on [0, 1] -> CDF 3 s^2 - 2 s^3
    assert stats.kstest(s, lambda x: 3 * x ** 2 - 2 * x ** 3).pvalue > 1e-4


def test_harmonic_ball_radius_sampler_2d(oracle_lib):
    R = 1.3
    r, _ = oracle_lib.sample_ball_radius(2, R, 0.0, 9, 200_000)
    s = np.clip(r / R, 1e-12, 1.0)
    # pdf 4 s ln(1/s) on [0, 1] -> CDF s^2 (1 - 2 ln s)
    assert stats.kstest(s, lambda x: np.where(x > 0, x ** 2 * (1 - 2 * np.log(np.maximum(x, 1e-300))), 0.0)).pvalue > 1e-4


@pytest.mark.parametrize("dim,R", [(2, 0.5), (2, 1.7), (3, 0.5), (3, 1.7)])
def test_harmonic_green_norm_is_integral_of_g(oracle_lib, dim, R):
    # norm() must equal the integral of G(r) over the ball (the source weight of one step)
    def g(r):
        return oracle_lib.greens_ball(dim, R, 0.0, r)["G"]
    surf = (lambda r: 2 * math.pi * r) if dim == 2 else (lambda r: 4 * math.pi * r * r)
    val, _ = integrate.quad(lambda r: g(r) * surf(r), 1e-9, R, limit=200)
    assert abs(oracle_lib.greens_ball(dim, R, 0.0, R / 2)["norm"] - val) < 1e-5 * max(1.0, val)


@pytest.mark.parametrize("lam,R", [(1.0, 0.5), (4.0, 1.0), (10.0, 0.3)])
def test_yukawa_3d_closed_forms(oracle_lib, lam, R):
    q = oracle_lib.greens_ball(3, R, lam, R / 3)
    mu = math.sqrt(lam) * R
    # probability of reaching the sphere: mu / sinh(mu); the dampening of a WoS step equals it
    assert abs(q["dampening"] - mu / math.sinh(mu)) < 2e-5
    assert abs(4 * math.pi * q["poisson_kernel"] - mu / math.sinh(mu)) < 2e-5
    # norm = integral of G over the ball = (1 - P) / lambda
    val, _ = integrate.quad(lambda r: oracle_lib.greens_ball(3, R, lam, r)["G"] * 4 * math.pi * r * r, 1e-9, R,
                            limit=200)
    assert abs(q["norm"] - val) < 2e-5 * max(1.0, val)
    assert abs(q["norm"] - (1 - mu / math.sinh(mu)) / lam) < 2e-5
    # Yukawa G is the harmonic one as lambda -> 0 and vanishes on the sphere
    assert abs(oracle_lib.greens_ball(3, R, lam, R * 0.999999)["G"]) < 1e-4


def test_yukawa_3d_radius_sampler_matches_density(oracle_lib):
    lam, R = 4.0, 0.8
    r, _ = oracle_lib.sample_ball_radius(3, R, lam, 21, 100_000)
    norm = oracle_lib.greens_ball(3, R, lam, R / 2)["norm"]

    def pdf(x):
        return oracle_lib.greens_ball(3, R, lam, max(x, 1e-7))["G"] * 4 * math.pi * x * x / norm
    grid = np.linspace(0, R, 401)
    dens = np.array([pdf(x) for x in grid])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    assert stats.kstest(r, lambda x: np.interp(x, grid, cdf)).pvalue > 1e-4


def test_offset_point_follows_rtg_ch6(oracle_lib):
    # far from the origin the offset moves each coordinate by integer ulps in the direction of n
    p = np.array([0.5, -0.75, 2.0], np.float32)
    n = np.array([1.0, 0.5, -0.25], np.float32)
    q = oracle_lib.offset_point(p, n)
    steps = q.view(np.int32).astype(np.int64) - p.view(np.int32).astype(np.int64)
    expect = (n * 256).astype(np.int64) * np.where(p < 0, -1, 1)
    np.testing.assert_array_equal(steps, expect)
    assert (np.sign(q - p) == np.sign(n)).all()
    # near the origin a fixed float offset of n / 65536 is used
    p0 = np.array([0.01, -0.02, 0.0], np.float32)
    q0 = oracle_lib.offset_point(p0, n)
    np.testing.assert_allclose(q0, p0 + n / 65536.0, rtol=0, atol=1e-9)


def test_yukawa_ball_terms_at_small_and_large_radii(oracle_lib):
    """The 3D Yukawa ball (distributions.h:835-1005) against float64 closed forms:
    G(r) = sinh(mu (R - r)) / (4 pi r sinh(mu R)), norm = (1 - mu R / sinh mu R) / lambda,
    dampening mu r (K(mu r) + I(mu r) e^{-mu R} / sinh mu R), gradient norm
    mu (K(mu r) - I(mu r) K(mu R) / I(mu R)) / (4 pi r^2) and Poisson-kernel gradient scale
    lambda / (4 pi I(mu R)), with K(x) = e^{-x} (1 + 1/x), I(x) = cosh x - sinh x / x -- down to
    mu R = 4e-4, where the exponential forms of these terms lose every digit in float."""
    import math
    O = oracle_lib

    def I(x):
        if x < 1e-2:
            x2 = x * x
            return x2 * (1 / 3 + x2 * (1 / 30 + x2 / 840))
        return math.cosh(x) - math.sinh(x) / x

    def K(x):
        return math.exp(-x) * (1 + 1 / x)

    for lam in (0.5, 4.0, 10.0):
        mu = math.sqrt(lam)
        for R in (1e-4, 2e-3, 0.05, 0.3, 1.5):
            x = mu * R
            for frac in (0.1, 0.5, 0.9):
                r = frac * R
                t = O.yukawa_ball(R, lam, r)
                want = dict(norm=(1 - x / math.sinh(x)) / lam if x > 1e-2 else R * R * (1 / 6 - 7 * x * x / 360),
                            dampening=mu * r * (K(mu * r) + I(mu * r) * math.exp(-x) / math.sinh(x)),
                            gradient_norm=mu * (K(mu * r) - I(mu * r) * K(x) / I(x)) / (4 * math.pi * r * r),
                            pk_gradient_scale=lam / I(x) / (4 * math.pi),
                            G=math.sinh(mu * (R - r)) / (4 * math.pi * r * math.sinh(x)),
                            P=x / (4 * math.pi * math.sinh(x)))
                for k, v in want.items():
                    assert math.isfinite(t[k]) and abs(t[k] - v) <= 2e-4 * abs(v) + 1e-30, (lam, R, r, k, t[k], v)
