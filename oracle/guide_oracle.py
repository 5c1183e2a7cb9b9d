"""CPU oracle of the learned path-guiding network (TEST INFRASTRUCTURE ONLY).

The reference has no guiding; this file restates, in numpy FP64, what the HIP
guide kernels compute, so the tests can check them:

- `encode`: the 32 input features of a guided vertex (p, n, wo);
- `forward` / `gradient`: the 32-64-64-64-41 MLP, its loss -w log q(dir) on
  the combined density q = alpha p_guide + (1 - alpha) cos/pi (alpha the
  network's selection output) and the parameter gradient, either in plain FP64
  (`emulate=False`) or with the kernels' BF16 rounding points reproduced
  (`emulate=True`: inputs, weights, hidden activations and output/hidden
  gradients rounded to BF16 round-to-nearest-even where the MFMA kernels round
  them, everything else in FP64);
- `adam`: one Adam step (beta1 0.9, beta2 0.999, eps 1e-8);
- the von Mises-Fisher mixture: parameters from raw outputs (44 floats: 8 x
  (mu, kappa, weight), alpha, 3 zeros), density, the three-uniform sampler and
  the guided one-sample-MIS Lambertian sample.

How this oracle is pinned (tests/test_oracle_guide.py): the FP64 gradient
against central finite differences of the FP64 loss; the mixture density
integrates to 1 on the sphere; the sampler's cosine distribution against the
closed-form vMF CDF and its mean against coth(k) - 1/k; the cosine branch
against pdf cos/pi and weight 1; Adam's first step against its closed form
lr * g / (|g| + eps).  Nothing in the product imports it.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT, N_LOBES, N_MIX = 32, 64, 41, 8, 44
OFF = {"W1": 0, "b1": 2048, "W2": 2112, "b2": 6208, "W3": 6272, "b3": 10368, "W4": 10432,
       "b4": 13056}
N_PARAMS = 13097
LOGK_MIN, LOGK_MAX = -3.0, 8.0
ALPHA_MAX = 0.9


def bf16(x) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(theta: np.ndarray) -> dict:
    t = np.asarray(theta, np.float64)
    return {"W1": t[0:2048].reshape(64, 32), "b1": t[2048:2112],
            "W2": t[2112:6208].reshape(64, 64), "b2": t[6208:6272],
            "W3": t[6272:10368].reshape(64, 64), "b3": t[10368:10432],
            "W4": t[10432:13056].reshape(41, 64), "b4": t[13056:13097]}


def pack(P: dict) -> np.ndarray:
    return np.concatenate([P[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")])


def encode(p, n, wo, center, radius) -> np.ndarray:
    """Features: sin/cos(2^i pi p_hat_a) at 8a + 2i (+1), then n, wo, 1, n.wo."""
    p = np.asarray(p, np.float64)
    ph = np.clip((p - np.asarray(center, np.float64)) / float(radius), -1.0, 1.0)
    n = np.asarray(n, np.float64)
    wo = np.asarray(wo, np.float64)
    f = np.zeros((p.shape[0], N_IN))
    for a in range(3):
        for i in range(4):
            ang = np.pi * ph[:, a] * (2.0 ** i)
            f[:, 8 * a + 2 * i] = np.sin(ang)
            f[:, 8 * a + 2 * i + 1] = np.cos(ang)
    f[:, 24:27] = n
    f[:, 27:30] = wo
    f[:, 30] = 1.0
    f[:, 31] = (n * wo).sum(1)
    return f


def forward(theta, X, emulate: bool = True):
    """Raw outputs (n, 40) and the hidden activations (for `gradient`)."""
    P = unpack(theta)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W = {k: r(P[k]) if k.startswith("W") else P[k] for k in P}
    x = r(X)
    h1 = r(np.maximum(x @ W["W1"].T + P["b1"], 0.0))
    h2 = r(np.maximum(h1 @ W["W2"].T + P["b2"], 0.0))
    h3 = r(np.maximum(h2 @ W["W3"].T + P["b3"], 0.0))
    o = h3 @ W["W4"].T + P["b4"]
    return o, (x, h1, h2, h3, W)


def _split(o):
    o = np.asarray(o, np.float64)
    lob = o[:, :40].reshape(-1, N_LOBES, 5)
    return lob[..., 0:3], lob[..., 3], lob[..., 4]


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-np.asarray(x, np.float64)))


def log_kappa(s):
    """Smooth bound of the concentration: log kappa = lo + (hi - lo) sigmoid(s)."""
    return LOGK_MIN + (LOGK_MAX - LOGK_MIN) / (1.0 + np.exp(-np.asarray(s, np.float64)))


def raw_to_mixture(o, alpha_max: float = ALPHA_MAX) -> np.ndarray:
    """(n, 41) raw -> (n, 44) = 8 x (mu.x, mu.y, mu.z, kappa, weight), alpha, 0, 0, 0."""
    o = np.asarray(o, np.float64)
    v, s, a = _split(o)
    ln = np.linalg.norm(v, axis=-1, keepdims=True)
    mu = np.where(ln > 1e-12, v / np.maximum(ln, 1e-300), np.array([0.0, 0.0, 1.0]))
    kap = np.exp(log_kappa(s))
    e = np.exp(a - a.max(-1, keepdims=True))
    pi = e / e.sum(-1, keepdims=True)
    out = np.zeros((o.shape[0], N_MIX))
    out[:, :40] = np.concatenate([mu, kap[..., None], pi[..., None]], -1).reshape(-1, 40)
    out[:, 40] = alpha_max * _sigmoid(o[:, 40])
    return out


def vmf_log_norm(k):
    return np.log(k) - np.log(2.0 * np.pi) - np.log1p(-np.exp(-2.0 * k))


def mixture_pdf(m, w) -> np.ndarray:
    m = np.asarray(m, np.float64)[:, :40].reshape(-1, N_LOBES, 5)
    w = np.asarray(w, np.float64)
    c = (m[..., 0:3] * w[:, None, :]).sum(-1)
    k = m[..., 3]
    return (m[..., 4] * np.exp(vmf_log_norm(k) + k * (c - 1.0))).sum(-1)


def _onb(mu):
    """Duff et al. 2017 basis (t, b) around unit vectors mu (n, 3)."""
    sg = np.where(mu[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sg + mu[:, 2])
    b = mu[:, 0] * mu[:, 1] * a
    t = np.stack([1.0 + sg * mu[:, 0] ** 2 * a, sg * b, -sg * mu[:, 0]], 1)
    bb = np.stack([b, sg + mu[:, 1] ** 2 * a, -mu[:, 1]], 1)
    return t, bb


def mixture_sample(m, u1, u2, u3) -> np.ndarray:
    m = np.asarray(m, np.float64)[:, :40].reshape(-1, N_LOBES, 5)
    cdf = np.cumsum(m[..., 4], -1)
    k = np.minimum((np.asarray(u1)[:, None] >= cdf[:, :N_LOBES - 1]).sum(-1), N_LOBES - 1)
    idx = np.arange(m.shape[0])
    mu = m[idx, k, 0:3]
    kap = m[idx, k, 3]
    u2 = np.asarray(u2, np.float64)
    w = 1.0 + np.log(u2 + (1.0 - u2) * np.exp(-2.0 * kap)) / kap
    w = np.clip(w, -1.0, 1.0)
    r = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    ph = 2.0 * np.pi * np.asarray(u3, np.float64)
    t, b = _onb(mu)
    d = mu * w[:, None] + t * (r * np.cos(ph))[:, None] + b * (r * np.sin(ph))[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def lambert_dir(n, r1, r2) -> np.ndarray:
    """The cosine sample of Lambertian::sample (bsdf.cpp:23-37) around n."""
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    phi = 2.0 * np.pi * np.asarray(r1, np.float64)
    r2 = np.asarray(r2, np.float64)
    ct, st = np.sqrt(1.0 - r2), np.sqrt(r2)
    t, b = _onb(n)
    return t * (st * np.cos(phi))[:, None] + b * (st * np.sin(phi))[:, None] + n * ct[:, None]


def guided_sample(m, n, u):
    """Direction, combined pdf q = alpha p_guide + (1-alpha) cos/pi (alpha = m[:, 40]),
    and the albedo-free weight (cos/pi)/q of the guided Lambertian sample."""
    u = np.asarray(u, np.float64)
    n = np.asarray(n, np.float64)
    alpha = np.asarray(m, np.float64)[:, 40]
    use_g = u[:, 0] < alpha
    d = np.where(use_g[:, None], mixture_sample(m, u[:, 1], u[:, 2], u[:, 3]),
                 lambert_dir(n, u[:, 1], u[:, 2]))
    cosv = (d * n).sum(1)
    q = alpha * mixture_pdf(m, d) + (1.0 - alpha) * np.maximum(cosv, 0.0) / np.pi
    wf = np.where((cosv > 0) & (q > 0), (cosv / np.pi) / np.where(q > 0, q, 1.0), 0.0)
    return d, q, wf


def loss_and_output_grad(o, dirs, w, cosn, alpha_max: float = ALPHA_MAX):
    """Per-sample loss -w log q(dir), q = alpha p_guide + (1 - alpha) max(cosn, 0)/pi,
    and d loss / d raw outputs (n, 41), FP64."""
    o = np.asarray(o, np.float64)
    v, s, a = _split(o)
    dirs = np.asarray(dirs, np.float64)
    w = np.asarray(w, np.float64)
    amax = a.max(-1, keepdims=True)
    lse_a = amax + np.log(np.exp(a - amax).sum(-1, keepdims=True))
    ln = np.linalg.norm(v, axis=-1)
    ok = ln > 1e-12
    mu = np.where(ok[..., None], v / np.maximum(ln, 1e-300)[..., None], np.array([0.0, 0.0, 1.0]))
    kap = np.exp(log_kappa(s))
    c = (mu * dirs[:, None, :]).sum(-1)
    ell = (a - lse_a) + vmf_log_norm(kap) + kap * (c - 1.0)
    lmax = ell.max(-1, keepdims=True)
    logpg = lmax[:, 0] + np.log(np.exp(ell - lmax).sum(-1))
    pg = np.exp(logpg)
    sa = _sigmoid(o[:, 40])
    alpha = alpha_max * sa
    pb = np.maximum(np.asarray(cosn, np.float64), 0.0) / np.pi
    q = alpha * pg + (1.0 - alpha) * pb
    wA = w * alpha * pg / q
    r = np.exp(ell - logpg[:, None])
    pi = np.exp(a - lse_a)
    gcoef = -wA[:, None] * r
    dlk = 1.0 / kap - 2.0 * np.exp(-2.0 * kap) / (-np.expm1(-2.0 * kap)) + (c - 1.0)
    sg = 1.0 / (1.0 + np.exp(-s))
    d = np.zeros(v.shape[:2] + (5,))
    d[..., 3] = gcoef * dlk * kap * (LOGK_MAX - LOGK_MIN) * sg * (1.0 - sg)
    d[..., 4] = -wA[:, None] * (r - pi)
    f = np.where(ok, gcoef * kap / np.maximum(ln, 1e-300), 0.0)
    d[..., 0:3] = f[..., None] * (dirs[:, None, :] - c[..., None] * mu)
    out = np.zeros((o.shape[0], N_OUT))
    out[:, :40] = d.reshape(-1, 40)
    out[:, 40] = -w * (pg - pb) / q * alpha_max * sa * (1.0 - sa)
    return -w * np.log(q), out


def gradient(theta, X, dirs, w, emulate: bool = True, alpha_max: float = ALPHA_MAX):
    """Mean loss and its gradient over the batch (the mupsi_guide_gradient contract);
    the vertex normal is read from the features (columns 24..26)."""
    o, (x, h1, h2, h3, W) = forward(theta, X, emulate)
    cosn = (np.asarray(X, np.float64)[:, 24:27] * np.asarray(dirs, np.float64)).sum(1)
    loss, dO = loss_and_output_grad(o, dirs, w, cosn, alpha_max)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    dO = r(dO)
    d3 = r((dO @ W["W4"]) * (h3 != 0))
    d2 = r((d3 @ W["W3"]) * (h2 != 0))
    d1 = r((d2 @ W["W2"]) * (h1 != 0))
    n = X.shape[0]
    G = {"W4": dO.T @ h3, "b4": dO.sum(0), "W3": d3.T @ h2, "b3": d3.sum(0),
         "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ x, "b1": d1.sum(0)}
    return loss.sum() / n, pack({k: G[k] / n for k in G})


def adam(theta, m, v, g, t, lr=1e-3):
    """One Adam step (t = 1 for the first) -> (theta, m, v)."""
    m = 0.9 * m + 0.1 * g
    v = 0.999 * v + 0.001 * g * g
    mh = m / (1.0 - 0.9 ** t)
    vh = v / (1.0 - 0.999 ** t)
    return theta - lr * mh / (np.sqrt(vh) + 1e-8), m, v


def init_like(seed: int = 0) -> np.ndarray:
    """A random parameter vector of the network's shape (for tests)."""
    rng = np.random.default_rng(seed)
    P = {"W1": rng.uniform(-1, 1, (64, 32)) * np.sqrt(6 / 32), "b1": rng.normal(0, 0.1, 64),
         "W2": rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 64), "b2": rng.normal(0, 0.1, 64),
         "W3": rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 64), "b3": rng.normal(0, 0.1, 64),
         "W4": rng.uniform(-1, 1, (41, 64)) * 0.3, "b4": rng.normal(0, 0.3, 41)}
    return pack(P).astype(np.float32)
