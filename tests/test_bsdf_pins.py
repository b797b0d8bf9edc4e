"""Pins of the oracle's DiffuseBsdf value (src/diffuse_bsdf.cc:57-89) for
roughness sigma > 0, against closed forms written from the reference's
formula: f = color (a + b t), a = 1 / (pi + (3 pi - 4) / 6 sigma) and
b = sigma a (Fujii's A and B; (3 pi - 4) / 6 = pi / 2 - 2 / 3),
s = V.L - max(N.L, 0) max(N.V, 0), t = s where s <= 0 and
t = max(N.L, N.V) + REAL_EPS (1e-5) where s > 0.

With N = +z and directions at polar angles (tv, tl) and azimuths (pv, pl)
above the horizon, s = sin tv sin tl cos(pv - pl): the sign of cos(pv - pl)
selects the branch, so both are checked against their closed forms in double
precision.  A slip in A or B, in the branch condition, in max vs min or a
dropped REAL_EPS moves the value by more than the tolerance.
"""
import numpy as np
import pytest

PI = np.pi
EPS = 1e-5
TOL = 1e-6  # relative to the scale color (a + b |t|) of the terms summed (fp32 evaluation)


def _close(got, want, scale):
    err = np.abs(got - want)
    assert (err <= TOL * scale).all(), (err / scale).max()


def _dirs(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def _ab(sigma):
    a = 1.0 / (PI + (PI / 2 - 2.0 / 3.0) * sigma)
    return a, sigma * a


def _cases(rng, k, same_side):
    tv, tl = rng.uniform(0.05, 1.5, k), rng.uniform(0.05, 1.5, k)
    pv = rng.uniform(0, 2 * PI, k)
    dphi = rng.uniform(-1.4, 1.4, k) if same_side else rng.uniform(1.65, 2 * PI - 1.65, k)
    pl = pv + dphi
    return tv, tl, pv, pl


@pytest.mark.parametrize("same_side", [False, True], ids=["s<=0", "s>0"])
def test_value_matches_the_closed_form(oracle, same_side):
    rng = np.random.default_rng(5 + same_side)
    k = 4000
    tv, tl, pv, pl = _cases(rng, k, same_side)
    V, L = _dirs(tv, pv), _dirs(tl, pl)
    N = np.tile([0.0, 0.0, 1.0], (k, 1))
    color = rng.uniform(0.1, 1.0, (k, 3))
    sigma = rng.uniform(0.05, 1.0, k)
    got = oracle.diffuse_bsdf(color, sigma, N, V, L).astype(np.float64)
    a, b = _ab(sigma)
    s = np.sin(tv) * np.sin(tl) * np.cos(pv - pl)
    assert (s > 1e-4).all() if same_side else (s < -1e-4).all()
    t = np.maximum(np.cos(tv), np.cos(tl)) + EPS if same_side else s
    want = color * (a + b * t)[:, None]
    _close(got, want, color * (a + b * np.abs(t))[:, None])
    if same_side:
        # where sigma > 0.8, dropping REAL_EPS from the override would move the
        # value by color * b * 1e-5, more than twice the tolerance
        rough = sigma > 0.8
        assert rough.sum() > 100
        without = color * (a + b * (t - EPS))[:, None]
        assert (np.abs(want - without)[rough] > 2 * TOL * np.abs(want)[rough]).all()


def test_zero_roughness_is_lambert(oracle):
    rng = np.random.default_rng(7)
    k = 1000
    tv, tl, pv, pl = rng.uniform(0.05, 1.5, k), rng.uniform(0.05, 1.5, k), rng.uniform(0, 6.28, k), \
        rng.uniform(0, 6.28, k)
    color = rng.uniform(0.1, 1.0, (k, 3))
    got = oracle.diffuse_bsdf(color, np.zeros(k), np.tile([0.0, 0.0, 1.0], (k, 1)), _dirs(tv, pv), _dirs(tl, pl))
    np.testing.assert_allclose(got, color / PI, rtol=2e-7)


def test_light_below_the_horizon_clamps_its_cosine(oracle):
    """N.L < 0: max(N.L, 0) = 0, so s = V.L; with V.L > 0 the override takes
    max(0, N.V) + REAL_EPS, with V.L <= 0 the value is a + b V.L."""
    rng = np.random.default_rng(9)
    k = 2000
    tv = rng.uniform(0.05, 1.5, k)
    tl = rng.uniform(1.65, 2.0, k)  # below the horizon
    pv = rng.uniform(0, 2 * PI, k)
    pl = pv + rng.uniform(-0.3, 0.3, k) + np.where(rng.uniform(size=k) < 0.5, 0.0, PI)
    V, L = _dirs(tv, pv), _dirs(tl, pl)
    color = rng.uniform(0.1, 1.0, (k, 3))
    sigma = rng.uniform(0.05, 1.0, k)
    got = oracle.diffuse_bsdf(color, sigma, np.tile([0.0, 0.0, 1.0], (k, 1)), V, L).astype(np.float64)
    a, b = _ab(sigma)
    vl = (V * L).sum(1)
    keep = np.abs(vl) > 1e-4
    t = np.where(vl > 0, np.cos(tv) + EPS, vl)
    want = color * (a + b * t)[:, None]
    _close(got[keep], want[keep], (color * (a + b * np.abs(t))[:, None])[keep])
    assert (vl[keep] > 0).any() and (vl[keep] < 0).any()


def test_reciprocity(oracle):
    """f(V, L) = f(L, V) bit for bit: every term of the formula is symmetric
    in V and L, including the override's max(N.L, N.V)."""
    rng = np.random.default_rng(11)
    k = 4000
    N = rng.normal(size=(k, 3))
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    V = rng.normal(size=(k, 3))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    L = rng.normal(size=(k, 3))
    L /= np.linalg.norm(L, axis=1, keepdims=True)
    color = rng.uniform(0.1, 1.0, (k, 3))
    sigma = rng.uniform(0.0, 1.0, k)
    np.testing.assert_array_equal(oracle.diffuse_bsdf(color, sigma, N, V, L), oracle.diffuse_bsdf(color, sigma, N, L, V))
