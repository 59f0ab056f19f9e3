"""Pins of the batched-con2prim oracle and the MLP reference (CPU only).

* mlp_ref.loss_grad against central finite differences in float64;
* bf16_round on hand-worked ties;
* orc4 Newton: converged W is a root of the 1DW residual (checked in
  extended precision with mpmath), round trip P -> U -> P, and independence
  of the converged P from the initial guess;
* features: sign-flip invariance in S, scale behaviour, and z against its
  closed form in the primitives (and that form's first-order term
  2 rho p Gamma^2 / ((gamma - 1) E^2) for p << rho).
"""
import numpy as np
import pytest

import mlp_ref
from verona_amd.c2p import c4_states, p2u

G = 4.0 / 3.0


@pytest.mark.parametrize("q_ref", [0.0, 0.5])
def test_mlp_reference_gradient_matches_finite_differences(q_ref):
    rng = np.random.default_rng(0)
    params = mlp_ref.he_init(3).astype(np.float64)
    params[mlp_ref.NPARAM - 2:] = [0.3, -0.2]
    X = rng.standard_normal((40, mlp_ref.NIN))
    T = rng.standard_normal((40, 2))
    loss, g = mlp_ref.loss_grad(params, X, T, w1=0.7, q_ref=q_ref)
    idx = rng.choice(mlp_ref.NPARAM, 300, replace=False)
    h = 1e-6
    lo = lambda p: mlp_ref.loss_only(p, X, T, 0.7, q_ref=q_ref)  # noqa: E731
    for k in idx:
        e = np.zeros_like(params)
        e[k] = h
        fd = (lo(params + e) - lo(params - e)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(fd)), (k, fd, g[k])
    assert abs(loss - lo(params)) < 1e-14 * max(1.0, loss)
    if q_ref > 0:
        # the weighted loss is the plain one with each sample's first error scaled
        s = 1 + (X[:, 2] ** 2 * np.exp(T[:, 0]) / q_ref) ** 2
        y = mlp_ref.forward(params, X)["y"]
        want = ((s * (y[:, 0] - T[:, 0]) ** 2 + 0.7 * (y[:, 1] - T[:, 1]) ** 2)).mean()
        assert abs(loss - want) <= 1e-12 * want


def test_bf16_round_ties_to_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1
    cases = {
        float(one): 1.0,
        float(one + ulp / 2): 1.0,               # tie -> even (1.0)
        float(one + 3 * ulp / 2): 1.0 + 2 * float(ulp),  # tie -> even (1 + 2 ulp)
        float(one + ulp / 2 + np.float32(2.0 ** -20)): 1.0 + float(ulp),
        -2.5: -2.5,
    }
    for x, want in cases.items():
        assert mlp_ref.bf16_round(np.array([x], np.float32))[0] == np.float32(want), x


def _moderate_states(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(-2, 1, n)
    p = 10 ** rng.uniform(-2, 1, n)
    lor = 10 ** rng.uniform(0, 1, n)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    P = np.empty((n, 5))
    P[:, 0], P[:, 1] = rho, p / (G - 1)
    P[:, 2:] = np.sqrt(lor ** 2 - 1)[:, None] * d
    return p2u(P, G), P


def test_orc4_newton_round_trip_and_guess_independence(oracle):
    U, P = _moderate_states(2000, 1)
    E = -U[:, 1]
    Pa, ita, ca = oracle.c2p4_solve(U, E, G, 100, 1e-12)
    assert ca.all()
    rel = np.abs(Pa - P) / np.maximum(np.abs(P), 1e-300)
    assert np.median(rel) < 1e-13 and rel[:, 0].max() < 1e-9
    W0g = oracle.c2p4_w0_from_prims(np.tile([1.0, 1.0, 0, 0, 0], (len(U), 1)), G)
    assert (W0g == 1.0 + G).all()
    Pg, itg, cg = oracle.c2p4_solve(U, W0g, G, 100, 1e-12)
    assert cg.all()
    d = np.abs(Pg - Pa).sum(0) / np.abs(Pa).sum(0)
    assert (d < 1e-12).all(), d
    assert itg.mean() > ita.mean()


def test_orc4_root_in_extended_precision(oracle):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    U, P = c4_states(24, seed=5, gamma=G)
    Pa, _, ca = oracle.c2p4_solve(U, -U[:, 1], G, 200, 1e-12)
    assert ca.all()
    c = mp.mpf(G - 1) / mp.mpf(G)
    for k in range(len(U)):
        D, Qt = (mp.mpf(float(x)) for x in U[k, :2])
        S2 = sum(mp.mpf(float(x)) ** 2 for x in U[k, 2:])

        def f(W):
            v2 = S2 / W ** 2
            return Qt + W - c * (W * (1 - v2) - D * mp.sqrt(1 - v2))

        lo, hi = mp.sqrt(S2) * (1 + mp.mpf("1e-30")), -Qt * 2 + 10
        W = mp.findroot(f, (lo, hi), solver="anderson")
        p = Pa[k]
        Wo = (p[0] + G * p[1]) * (1 + p[2] ** 2 + p[3] ** 2 + p[4] ** 2)
        # W from the recovered primitives carries the conditioning of the
        # recovery (Gamma^2); the root itself must match to 1e-12 relative
        lor2 = 1 + p[2] ** 2 + p[3] ** 2 + p[4] ** 2
        assert abs(mp.mpf(Wo) - W) / W < 1e-12 * lor2, (k, Wo, W)


def test_features_invariance_scale_and_closed_form(oracle):
    U, P = c4_states(5000, seed=9, gamma=G)
    X, E, z, ok = oracle.c2p4_features(U)
    assert ok.mean() > 0.99
    Uf = U.copy()
    Uf[:, 2:] *= -1.0
    Uf[:, 3] *= -1.0
    Xf, _, zf, okf = oracle.c2p4_features(Uf)
    assert (Xf == X).all() and (zf == z).all() and (okf == ok).all()
    # scaling (D, tau, S) by k leaves every (dimensionless) feature fixed
    k = 8.0
    Xs, _, zs, _ = oracle.c2p4_features(U * k)
    m = ok
    assert np.allclose(Xs[m], X[m], rtol=1e-6, atol=1e-7)
    # the features from the primitives: D/E = rho Gamma / E, |S|/E = w Gamma |u| / E,
    # sqrt(z) and ln(z)/10
    lorp = np.sqrt(1 + (P[:, 2:] ** 2).sum(1))
    wp = P[:, 0] + G * P[:, 1]
    assert np.allclose(X[m, 0], (P[:, 0] * lorp / E)[m], rtol=1e-6)
    assert np.allclose(X[m, 1], (wp * lorp * np.sqrt((P[:, 2:] ** 2).sum(1)) / E)[m], rtol=1e-6, atol=1e-7)
    assert np.allclose(X[m, 2] ** 2, z[m], rtol=1e-6)
    assert np.allclose(X[m, 3], (np.log(z) / 10)[m], rtol=1e-6, atol=1e-6)
    # closed form from the primitives: E^2 - S^2 - D^2 = Gamma^2 (w^2 - rho^2 - 2 w p) + p^2,
    # whose first-order term in p is 2 rho p Gamma^2 / (gamma - 1)
    rho, p = P[:, 0], (G - 1) * P[:, 1]
    w = rho + G * P[:, 1]
    lor2 = 1 + (P[:, 2:] ** 2).sum(1)
    zc = (lor2 * ((w - rho) * (w + rho) - 2 * w * p) + p * p) / E ** 2
    well = ok & (zc > 1e-6)
    assert well.sum() > 1000
    assert np.allclose(z[well], zc[well], rtol=1e-6)
    small = ok & (p / rho < 1e-5) & (zc > 1e-7)
    assert small.sum() > 20
    assert np.allclose(z[small], (2 * rho * p * lor2 / ((G - 1) * E ** 2))[small], rtol=1e-4)
