"""Pins the guiding oracle (oracle/guide_oracle.py) with checks that do not retype
its formulas: finite differences, normalisation on the sphere, closed-form vMF
marginals and moments, and Adam's closed-form first step."""
import numpy as np
import pytest

from oracle import guide_oracle as G


def _batch(rng, n):
    X = rng.normal(0, 1, (n, 32))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(0.1, 2.0, n)
    return X, d, w


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = G.init_like(3).astype(np.float64)
    X, d, w = _batch(rng, 24)
    loss, g = G.gradient(theta, X, d, w, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(o, o + 200), 6, replace=False)
                          for o in (0, 2048, 2112, 6208, 6272, 10368, 10432, 12850)] +
                         [np.arange(13056, 13097)])
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tp[i] += h
        tm = theta.copy()
        tm[i] -= h
        fd = (G.gradient(tp, X, d, w, emulate=False)[0] - G.gradient(tm, X, d, w, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_output_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    o = rng.normal(0, 1.5, (6, 41))
    o[:, 3:40:5] = rng.uniform(-3.0, 3.0, (6, 8))  # raw log-kappa parameter
    d = rng.normal(size=(6, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(0.5, 2, 6)
    cosn = np.array([0.3, 0.9, -0.2, 0.0, 0.6, 0.05])   # incl. directions below the surface
    _, g = G.loss_and_output_grad(o, d, w, cosn)
    h = 1e-6
    for j in range(41):
        op, om = o.copy(), o.copy()
        op[:, j] += h
        om[:, j] -= h
        fd = (G.loss_and_output_grad(op, d, w, cosn)[0] - G.loss_and_output_grad(om, d, w, cosn)[0]) / (2 * h)
        np.testing.assert_allclose(g[:, j], fd, rtol=1e-5, atol=1e-7)


def test_selection_output_moves_toward_the_better_technique():
    """Where the sampled direction has much more guide than cosine density, the
    loss gradient raises alpha; where it has none, it lowers it."""
    o = np.zeros((2, 41))
    o[:, 0:3] = [0.0, 0.0, 1.0]
    o[:, 3] = 2.0                       # a sharp lobe around +z
    d = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    _, g = G.loss_and_output_grad(o, d, np.ones(2), np.array([0.5, 0.5]))
    assert g[0, 40] < 0 < g[1, 40]    # gradient descent: alpha up for 0, down for 1


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mixture_pdf_integrates_to_one(seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(0, 1, (1, 41))
    o[0, 3:40:5] = rng.uniform(-3, -0.2, 8)   # kappa up to ~33: resolved by the grid below
    m = G.raw_to_mixture(o)
    w = _fib_sphere(400_000)
    p = G.mixture_pdf(np.repeat(m, len(w), 0), w)
    assert abs(p.mean() * 4 * np.pi - 1.0) < 2e-3


def test_single_lobe_sampler_matches_vmf_marginal_and_mean():
    rng = np.random.default_rng(4)
    n = 200_000
    for kap in (0.3, 4.0, 60.0):
        mu = rng.normal(size=3)
        mu /= np.linalg.norm(mu)
        m = np.zeros((n, 40))
        m[:, 0:3] = mu
        m[:, 3] = kap
        m[:, 4] = 1.0
        m[:, 8::5] = kap  # other lobes have zero weight
        m[:, 5:8] = mu
        u = rng.uniform(size=(n, 3))
        d = G.mixture_sample(m, u[:, 0], u[:, 1], u[:, 2])
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
        c = d @ mu
        # CDF of mu.w for a vMF on S^2: (e^{k(c-1)} - e^{-2k}) / (1 - e^{-2k})
        cs = np.sort(c)
        cdf = (np.exp(kap * (cs - 1)) - np.exp(-2 * kap)) / (1 - np.exp(-2 * kap))
        ks = np.abs(cdf - (np.arange(n) + 0.5) / n).max()
        assert ks < 5e-3, (kap, ks)
        mean = d.mean(0)
        A = 1 / np.tanh(kap) - 1 / kap
        np.testing.assert_allclose(mean, A * mu, atol=4e-3)


def test_cosine_branch_has_pdf_cos_over_pi_and_unit_weight():
    rng = np.random.default_rng(5)
    n = 5000
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    m = G.raw_to_mixture(rng.normal(size=(n, 41)), alpha_max=0.0)
    u = rng.uniform(size=(n, 4))
    d, q, wf = G.guided_sample(m, nrm, u)
    cosv = (d * nrm).sum(1)
    assert (cosv >= -1e-12).all()
    np.testing.assert_allclose(q, np.maximum(cosv, 0) / np.pi, rtol=1e-12)
    np.testing.assert_allclose(wf[cosv > 1e-9], 1.0, rtol=1e-9)
    # cosine-weighted: E[cos] = 2/3
    assert abs(cosv.mean() - 2 / 3) < 0.01


def test_guided_estimator_is_unbiased_for_a_known_integrand():
    """E[(cos/pi)/q * L(w)] over the MIS sample equals the cosine integral of L."""
    rng = np.random.default_rng(6)
    n = 400_000
    nrm = np.tile([0.0, 0.0, 1.0], (n, 1))
    o = np.zeros((1, 41))
    o[0, 0:3] = [0.3, 0.2, 1.0]
    o[0, 3] = 0.5
    o[0, 5:8] = [0.0, 1.0, 0.0]
    o[0, 8] = 1.0
    m = np.repeat(G.raw_to_mixture(o, alpha_max=1.2), n, 0)   # alpha = 1.2 sigmoid(0) = 0.6
    u = rng.uniform(size=(n, 4))
    d, q, wf = G.guided_sample(m, nrm, u)
    L = 1.0 + 3.0 * np.maximum(d[:, 0], 0) ** 2          # incident radiance
    est = (wf * L).mean()
    # int (cos/pi) (1 + 3 max(x,0)^2) dw over the upper hemisphere = 1 + 3 * (1/8)
    assert abs(est - (1 + 3 / 8)) < 0.01


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(7)
    th = rng.normal(size=100)
    g = rng.normal(size=100)
    t1, m, v = G.adam(th, np.zeros(100), np.zeros(100), g, 1, lr=1e-3)
    np.testing.assert_allclose(t1, th - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie -> even (mantissa ...0)
    assert r[2] == 1.0 + 2 ** -7         # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8


def test_encode_is_periodic_features_of_the_normalised_position():
    p = np.array([[1.0, -2.0, 0.5], [0.0, 0.0, 0.0]])
    f = G.encode(p, [[0, 0, 1]] * 2, [[0, 1, 0]] * 2, [0, 0, 0], 4.0)
    np.testing.assert_allclose(f[0, 0], np.sin(np.pi * 0.25))
    np.testing.assert_allclose(f[0, 8 + 3], np.cos(2 * np.pi * -0.5))   # y, octave 1
    np.testing.assert_allclose(f[0, 8 + 7], np.cos(8 * np.pi * -0.5))   # y, octave 3
    np.testing.assert_allclose(f[1, 0::2][:12], 0.0, atol=1e-15)
    assert f[0, 30] == 1.0 and f[0, 31] == 0.0
