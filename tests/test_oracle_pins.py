"""Pins three parts of the oracle that dort's renders cannot pin, because dort has no
counterpart for them: the conductor Fresnel term of the GGX conductor, the BRDF it
feeds, and the counter-based generator (include/dgpu_rng.h) that the GPU path and the
oracle's counter mode share.

* Conductor Fresnel: checked against the Fresnel equations evaluated in complex
  arithmetic (Born & Wolf, Principles of Optics, section 14.2: r_s and r_p of a wave
  from vacuum onto a medium of complex index n = eta + i k), which the oracle's
  real-valued form (a^2 + b^2 terms) must equal for every angle.  Also against its
  closed forms: ((eta-1)^2 + k^2) / ((eta+1)^2 + k^2) at normal incidence, 1 at grazing
  incidence, and dort's dielectric Fresnel (fresnel.cpp:5-47) when k = 0.
* GGX conductor BRDF: the Walter et al. 2007 microfacet form F D G1 G1 / (4 cos cos)
  with the GGX D and Smith G1 of microfacet.cpp:31-57 composed here in numpy.
* Philox4x32-10: the Random123 known-answer vectors (kat_vectors, 10 rounds), and an
  independent numpy implementation of the published round function on random counters.
"""
import math

import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import oracle


def fresnel_complex(cos_i, eta, k):
    """Unpolarised reflectance from vacuum onto index eta + i k (Born & Wolf)."""
    n = complex(eta, k)
    ci = np.asarray(cos_i, dtype=np.complex128)
    si2 = 1.0 - ci * ci
    ct = np.sqrt(1.0 - si2 / (n * n))
    rs = (ci - n * ct) / (ci + n * ct)
    rp = (n * ci - ct) / (n * ci + ct)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


METALS = sorted(scenes.CONDUCTORS)
COS = np.concatenate([np.linspace(0.0, 1.0, 401), [1e-6, 1e-4, 1e-3, 0.999999]])


@pytest.mark.parametrize("metal", METALS)
def test_conductor_fresnel_equals_complex_fresnel_equations(metal):
    for eta, k in zip(*scenes.CONDUCTORS[metal]):
        ref = fresnel_complex(COS, eta, k)
        np.testing.assert_allclose(oracle.fresnel_conductor(COS, eta, k, fp64=True), ref, rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(oracle.fresnel_conductor(COS, eta, k), ref, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("metal", METALS)
def test_conductor_fresnel_normal_and_grazing_incidence(metal):
    for eta, k in zip(*scenes.CONDUCTORS[metal]):
        normal = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
        for fp64, tol in ((True, 1e-14), (False, 1e-6)):
            f = oracle.fresnel_conductor([1.0, 0.0, 1e-4], eta, k, fp64=fp64)
            assert abs(f[0] - normal) <= tol * max(1.0, normal)
            assert abs(f[1] - 1.0) <= tol           # grazing: total reflection
            assert f[2] > 0.99                       # and continuously so
        # between the two the reflectance of a metal first dips, then rises to 1
        f = oracle.fresnel_conductor(COS[:401], eta, k, fp64=True)
        assert f.min() <= normal + 1e-12 and np.all(f <= 1.0 + 1e-12)


@pytest.mark.parametrize("eta", [1.1, 1.33, 1.5, 2.4])
def test_conductor_fresnel_with_k_zero_is_dorts_dielectric_fresnel(eta):
    """An absorption-free 'conductor' is a dielectric seen from outside: the complex form
    must reduce to fresnel_dielectric(cos_i, cos_t, 1, eta) of dort (fresnel.cpp:22-35)."""
    c = COS[COS > 0]
    diel = oracle.fresnel_dielectric(c, 1.0, eta, fp64=True)
    np.testing.assert_allclose(oracle.fresnel_conductor(c, eta, 0.0, fp64=True), diel, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(oracle.fresnel_conductor(c, eta, 0.0), oracle.fresnel_dielectric(c, 1.0, eta),
                               rtol=5e-5, atol=1e-6)
    # and dort's dielectric Fresnel itself is Born & Wolf's for real indices
    np.testing.assert_allclose(diel, fresnel_complex(c, eta, 0.0), rtol=1e-10, atol=1e-14)


def test_dielectric_fresnel_total_internal_reflection():
    """From inside (ior 1.5 into 1) beyond the critical angle the reflectance is 1."""
    crit = math.sqrt(1.0 - (1.0 / 1.5) ** 2)  # cos of the critical angle
    c = np.array([crit * 0.5, crit * 0.99, crit * 1.01, 0.9, 1.0])
    f = oracle.fresnel_dielectric(c, 1.5, 1.0, fp64=True)
    assert f[0] == 1.0 and f[1] == 1.0
    assert f[2] < 1.0 and abs(f[4] - 0.04) < 1e-12


def _ggx_d(alpha, cos_m):
    c2 = cos_m * cos_m
    t2 = (1 - c2) / c2
    return alpha ** 2 / (math.pi * c2 * c2 * (alpha ** 2 + t2) ** 2)


def _ggx_g1(alpha, cos_v):
    c2 = cos_v * cos_v
    t2 = (1 - c2) / c2
    return 2.0 / (1.0 + math.sqrt(1.0 + alpha ** 2 * t2))


@pytest.mark.parametrize("metal,alpha", [("gold", 0.1), ("copper", 0.3), ("aluminium", 0.5), ("silver", 0.8)])
def test_ggx_conductor_brdf_is_the_microfacet_form(metal, alpha):
    m = scenes.ggx_conductor(alpha, metal, tint=(0.9, 0.8, 0.7))
    rng = np.random.default_rng(3)
    n = 2000
    wi = rng.normal(size=(n, 3))
    wo = rng.normal(size=(n, 3))
    wi[:, 2] = np.abs(wi[:, 2]) + 0.05
    wo[:, 2] = np.abs(wo[:, 2]) + 0.05
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    f, _ = oracle.bsdf_eval(m, wi.astype(np.float32), wo.astype(np.float32))
    eta, k = scenes.CONDUCTORS[metal]
    tint = np.array([0.9, 0.8, 0.7])
    wi32, wo32 = wi.astype(np.float32).astype(np.float64), wo.astype(np.float32).astype(np.float64)
    for i in range(n):
        h = wi32[i] + wo32[i]
        h /= np.linalg.norm(h)
        ch = float(np.dot(wi32[i], h))
        dg = _ggx_d(alpha, h[2]) * _ggx_g1(alpha, wi32[i][2]) * _ggx_g1(alpha, wo32[i][2])
        ref = tint * np.array([fresnel_complex(ch, eta[c], k[c]) for c in range(3)]) * dg / (
            4 * wi32[i][2] * wo32[i][2])
        np.testing.assert_allclose(f[i], ref, rtol=2e-4, atol=1e-7)


# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key, expected output.
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert [int(x) for x in oracle.philox4x32_10(ctr, key)] == list(want)


def _philox_numpy(c, k):
    """Philox4x32-10 after Salmon et al. (SC'11), vectorised over rows of c (N x 4)."""
    c = c.astype(np.uint64)
    k0, k1 = np.uint64(k[0]), np.uint64(k[1])
    m0, m1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    w0, w1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    mask = np.uint64(0xFFFFFFFF)
    x0, x1, x2, x3 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    for _ in range(10):
        p0, p1 = m0 * x0, m1 * x2
        x0, x1, x2, x3 = ((p1 >> np.uint64(32)) ^ x1 ^ k0) & mask, p1 & mask, \
            ((p0 >> np.uint64(32)) ^ x3 ^ k1) & mask, p0 & mask
        k0, k1 = (k0 + w0) & mask, (k1 + w1) & mask
    return np.stack([x0, x1, x2, x3], axis=1).astype(np.uint32)


def test_philox_matches_independent_implementation():
    rng = np.random.default_rng(11)
    ctrs = rng.integers(0, 2 ** 32, size=(64, 4), dtype=np.uint64).astype(np.uint32)
    keys = rng.integers(0, 2 ** 32, size=(64, 2), dtype=np.uint64).astype(np.uint32)
    for c, k in zip(ctrs, keys):
        assert np.array_equal(oracle.philox4x32_10(c, k), _philox_numpy(c[None, :], k)[0])
