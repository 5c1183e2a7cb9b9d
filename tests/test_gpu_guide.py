"""GPU parity of the learned guide: MFMA forward / gradient / Adam against the CPU
oracle, the encoding and the guided sample, online training, the unbiasedness of
the guided tracer against the plain one, and what the trained network gains."""
import numpy as np
import pytest

from mupsi_amd import scenes
from oracle import guide_oracle as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def guide(ctx):
    from mupsi_amd import Guide
    g = Guide(ctx, seed=5)
    yield g
    g.close()


def _batch(rng, n):
    X = rng.normal(0, 1, (n, 32)).astype(np.float32)
    X[:, 30] = 1.0
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.uniform(0.1, 2.0, n).astype(np.float32)
    return X, d, w


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_forward_matches_oracle(guide, dev):
    from mupsi_amd import batch
    theta = G.init_like(11)
    guide.set_params(theta)
    np.testing.assert_array_equal(guide.params(), theta)
    rng = np.random.default_rng(0)
    X, _, _ = _batch(rng, 1000)
    raw, mix = (t.cpu().numpy() for t in batch.guide_forward(guide, _t(X, dev)))
    ref, _ = G.forward(theta, X, emulate=True)
    err = np.abs(raw - ref) / (1.0 + np.abs(ref))
    print("forward: max", err.max(), "median", np.median(err))
    assert np.median(err) < 1e-5 and err.max() < 2e-2
    ref64, _ = G.forward(theta, X, emulate=False)       # BF16 vs FP64 network
    assert np.abs(raw - ref64).max() < 0.1 * (1 + np.abs(ref64).max())
    np.testing.assert_allclose(mix, G.raw_to_mixture(raw), rtol=2e-5, atol=2e-6)


@pytest.mark.parametrize("n", [64, 777])
def test_gradient_matches_oracle(guide, dev, n):
    from mupsi_amd import batch
    theta = G.init_like(12)
    guide.set_params(theta)
    rng = np.random.default_rng(n)
    X, d, w = _batch(rng, n)
    loss, g = batch.guide_gradient(guide, _t(X, dev), _t(d, dev), _t(w, dev))
    rl, rg = G.gradient(theta, X, d, w, emulate=True)
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print(f"n={n}: loss {loss} vs {rl}; gradient rel L2 error {rel:.2e}")
    assert abs(loss - rl) <= 1e-4 * (1 + abs(rl))
    assert rel < 1e-2
    for name, (a, b) in {"W1": (0, 2048), "b1": (2048, 2112), "W4": (10432, 13056),
                         "b4": (13056, 13097)}.items():
        assert np.linalg.norm(g[a:b] - rg[a:b]) <= 2e-2 * np.linalg.norm(rg[a:b]) + 1e-7, name
    _, g64 = G.gradient(theta, X, d, w, emulate=False)
    cos = g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64))
    assert cos > 0.995, cos


def test_train_step_is_adam_on_the_gradient(guide, dev):
    from mupsi_amd import batch
    theta = G.init_like(13)
    guide.set_params(theta)
    rng = np.random.default_rng(3)
    X, d, w = _batch(rng, 512)
    _, g = batch.guide_gradient(guide, _t(X, dev), _t(d, dev), _t(w, dev))
    batch.guide_train_step(guide, _t(X, dev), _t(d, dev), _t(w, dev))
    t1 = guide.params()
    exp, _, _ = G.adam(theta.astype(np.float64), 0.0, 0.0, g.astype(np.float64), 1, lr=1e-3)
    np.testing.assert_allclose(t1, exp, rtol=0, atol=2e-6)


def test_encode_matches_oracle(ctx, dev):
    from mupsi_amd import Guide, batch
    g = Guide(ctx, box_center=(1.0, -2.0, 0.5), box_radius=3.0)
    rng = np.random.default_rng(4)
    p = rng.normal(0, 3, (500, 3)).astype(np.float32)
    n = rng.normal(size=(500, 3)).astype(np.float32)
    wo = rng.normal(size=(500, 3)).astype(np.float32)
    f = batch.guide_encode(g, _t(p, dev), _t(n, dev), _t(wo, dev)).cpu().numpy()
    ref = G.encode(p, n, wo, [1.0, -2.0, 0.5], 3.0)
    np.testing.assert_allclose(f, ref, atol=2e-5, rtol=1e-5)
    g.close()


def test_guided_sample_matches_oracle(guide, dev):
    from mupsi_amd import batch
    rng = np.random.default_rng(5)
    n = 4000
    raw = rng.normal(0, 1, (n, 41))
    raw[:, 3:40:5] = rng.uniform(-3, 1.5, (n, 8))
    mix = G.raw_to_mixture(raw).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    u = rng.uniform(size=(n, 4)).astype(np.float32)
    d, q, w = (t.cpu().numpy() for t in batch.guide_sample(guide, _t(mix, dev), _t(nrm, dev), _t(u, dev)))
    rd, rq, rw = G.guided_sample(mix.astype(np.float64), nrm.astype(np.float64),
                                 u.astype(np.float64))
    np.testing.assert_allclose(d, rd, atol=2e-4)
    ok = rq > 1e-3
    np.testing.assert_allclose(q[ok], rq[ok], rtol=2e-3)
    np.testing.assert_allclose(w[ok], rw[ok], rtol=2e-3, atol=1e-6)


def test_training_learns_a_directional_distribution(ctx, dev):
    """Adam on samples of a position-dependent vMF: the network's log-likelihood of
    fresh samples rises from the uniform level to near the target's."""
    from mupsi_amd import Guide, batch
    g = Guide(ctx, seed=9)
    rng = np.random.default_rng(6)

    def draw(n):
        p = rng.uniform(-1, 1, (n, 3))
        mu = np.stack([np.sign(p[:, 0]) * 0.6, np.full(n, 0.8), np.zeros(n)], 1)
        m = np.zeros((n, 44))
        m[:, 0:3] = mu
        m[:, 3] = 40.0
        m[:, 4] = 1.0
        m[:, 8:40:5] = 1.0
        m[:, 5:8] = mu
        d = G.mixture_sample(m, rng.uniform(size=n), rng.uniform(size=n), rng.uniform(size=n))
        nrm = np.tile([0.0, 1.0, 0.0], (n, 1))
        X = G.encode(p, nrm, nrm, [0, 0, 0], 1.0).astype(np.float32)
        return X, d.astype(np.float32), np.ones(n, np.float32)

    Xt, dt, wt = draw(4096)
    l0, _ = batch.guide_gradient(g, _t(Xt, dev), _t(dt, dev), _t(wt, dev))
    for _ in range(300):
        X, d, w = draw(4096)
        batch.guide_train_step(g, _t(X, dev), _t(d, dev), _t(w, dev))
    l1, _ = batch.guide_gradient(g, _t(Xt, dev), _t(dt, dev), _t(wt, dev))
    # the loss is -log q of the combined density q = alpha p_guide + (1 - alpha) cos/pi:
    # from about -log(0.45/4pi + 0.55 cos/pi) > 1 untrained to below 0 once the guide
    # concentrates on the target (vMF(40) alone: -log(40/2pi) + 1 = -0.85)
    print("held-out loss", l0, "->", l1)
    assert l0 > 1.0 and l1 < 0.0
    g.close()


def _scene5(res):
    return scenes.preset_scene("guided_sphere_quad_1024", width=res, height=res)


def _chunk_means(s, mask, chunks, spp, k0=0):
    """GP-region mean of `chunks` independent renders of `spp` passes each."""
    out = []
    for c in range(chunks):
        s.reset()
        s.render_passes(k0 + c * spp, spp)
        out.append(s.image()[mask].mean())
    return np.array(out)


def test_guided_render_is_unbiased(ctx):
    """Same expectation as the plain tracer over the GP region, judged against the
    spread of independent chunks (no per-path replay is possible: the guided
    tracer draws different numbers)."""
    from mupsi_amd import Guide, Scene
    sc = _scene5(48)
    plain = Scene(ctx, sc)
    plain.render_passes(0, 64)
    sky = Scene(ctx, sc)
    sky.render_passes(64, 64)
    mask = np.abs(plain.image() - sky.image()).sum(2) > 0    # pixels whose paths are random
    assert mask.sum() > 50
    a = _chunk_means(plain, mask, 8, 512)
    guided = Scene(ctx, sc)
    g = Guide.for_scene(ctx, sc, train_passes=8, batch_size=1024)
    guided.set_guide(g)
    b = _chunk_means(guided, mask, 8, 512)
    st = g.stats()
    assert st["train_passes_done"] == 8 and st["train_steps"] > 0, st
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    z = abs(a.mean() - b.mean()) / se
    print(f"GP-region mean plain {a.mean():.5f} guided {b.mean():.5f} (se {se:.5f}, z {z:.2f}); {st}")
    assert z < 4.0 and abs(a.mean() - b.mean()) < 0.02 * a.mean()
    g.close()


def test_guided_tracer_with_alpha_zero_weights_like_the_plain_tracer(ctx):
    """alpha = 0: every guided sample is the cosine sample with weight albedo, so the
    image converges to the plain one; with the same passes the sky pixels agree."""
    from mupsi_amd import Guide, Scene
    sc = scenes.preset_scene("sphere_renewal_256", width=32, height=32)
    s = Scene(ctx, sc)
    g = Guide.for_scene(ctx, sc, alpha=0.0, train_passes=0)
    s.set_guide(g)
    s.render_passes(0, 512)
    b = s.image()
    p = Scene(ctx, sc)
    p.render_passes(0, 512)
    a = p.image()
    mask = np.abs(a - b).sum(2) > 0
    assert abs(a[mask].mean() - b[mask].mean()) <= 0.02 * a[mask].mean()
    g.close()


def test_trained_guide_reduces_error(ctx):
    """At equal sample count after training, the guided image is closer to a
    high-sample plain reference than the plain image is (relMSE with the
    eps = 0.01 denominator), both with the reference's Russian roulette and with
    the guided tracer's gentler one."""
    from mupsi_amd import Guide, Scene
    res = 256
    sc = _scene5(res)
    ref = Scene(ctx, sc)
    ref.render_passes(0, 4096)
    R = ref.image()

    def relmse(I):
        return float(np.mean((I - R) ** 2 / (R ** 2 + 0.01)))
    plain = Scene(ctx, sc)
    plain.render_passes(5000, 64)
    ep = relmse(plain.image())
    out = {}
    for boost in (1.0, 4.0):
        g = Guide.for_scene(ctx, sc, train_passes=32, rr_boost=boost)
        guided = Scene(ctx, sc)
        guided.set_guide(g)
        guided.render_passes(5000, 32)      # the 32 training passes
        guided.reset()
        guided.render_passes(5032, 64)      # 64 passes with the trained network
        out[boost] = relmse(guided.image())
        print(f"relMSE at 64 spp: plain {ep:.3e} guided(rr_boost {boost}) {out[boost]:.3e} "
              f"ratio {ep / out[boost]:.2f}; {g.stats()}")
        g.close()
    assert out[1.0] < 0.85 * ep      # the network alone
    assert out[4.0] < 0.5 * ep       # with the guided roulette
