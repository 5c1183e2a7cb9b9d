"""Pinning the oracle's GP medium: the march against an independent t-grid scan
(gpmedium.cpp:46-129), Renewal+ against its closed-form invariants
(gpmedium.cpp:133-162), the mean functions against their definitions."""
import numpy as np
import pytest

from mupsi_amd import scenes
from oracle import oracle as O
from tests.test_oracle_psi import f32fma


def _scene(name):
    return scenes.preset_scene(name, width=16, height=16)


def scan(osc, ray, seed, cond=None, mode=0):
    """Brute-force restatement of sampleDistance/transmittance on the FP32 t grid."""
    o, d = ray[:3], ray[3:6]
    d = (d / np.sqrt(np.sum(d.astype(np.float64) ** 2))).astype(np.float32)  # Ray normalizes
    nt = np.float32(ray[6])
    ft = np.float32(min(ray[7], np.float32(2000.0)))
    dt = np.float32(min(np.float32(0.5), np.float32(ft - nt) / np.float32(100)))
    ts = []
    t = nt
    while t < ft and len(ts) < 5000:
        ts.append(t)
        t = np.float32(t + dt)
    ts = np.array(ts, np.float32)
    pts = np.stack([[f32fma(d[a], tt, o[a]) for a in range(3)] for tt in ts]) if len(ts) else \
        np.zeros((0, 3), np.float32)
    c = None if cond is None else np.repeat(cond[None], len(ts) + 1, 0)
    f = osc.field_eval(np.vstack([o[None], pts]), np.full(len(ts) + 1, seed, np.uint32), c,
                       fp64=False)[:, 0].astype(np.float32)
    fprev, fs = f[0], f[1:]
    for i, fc in enumerate(fs):
        if (fc > 0) != (fprev > 0) and abs(np.float32(fprev - fc)) > 1e-6:
            return True, ts[i], i + 1
    return False, None, len(ts)


def _rays(rng, n, far=False):
    c = np.array([0, 0, -500.0])
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = c + d0 * rng.uniform(75, 125, (n, 1))
    d = (c + rng.normal(size=(n, 3)) * 40) - o
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = rng.uniform(20, 200, n) if far else np.finfo(np.float32).max
    return rays


@pytest.mark.parametrize("name,k,mode,far", [("sphere_single_128", 1, 0, False),
                                             ("sphere_renewal_256", 2, 0, False),
                                             ("sphere_renewal_256", 2, 1, True)])
def test_march_equals_grid_scan(name, k, mode, far):
    osc = O.OracleScene(_scene(name))
    rng = np.random.default_rng(k * 10 + mode)
    rays = _rays(rng, 12, far)
    seeds = rng.integers(0, 2**32, 12, dtype=np.uint64).astype(np.uint32)
    cond = None
    if name == "sphere_renewal_256":
        cond = np.zeros((12, 8), np.float32)
        cond[:, 0] = 1
        cond[:, 1:4] = rays[:, :3] + rng.normal(0, 2, (12, 3))
        cond[:, 4] = rng.normal(0, 3, 12)
        cond[:, 5:8] = rng.normal(0, 2, (12, 3))
    m = osc.march(rays, seeds, cond, mode=mode, fp64=False, track=False)
    hits = 0
    for i in range(12):
        hit, t, steps = scan(osc, rays[i], int(seeds[i]), None if cond is None else cond[i], mode)
        assert bool(m["hit"][i]) == hit, i
        if hit:
            hits += 1
            assert m["t"][i] == t and m["steps"][i] == steps
    assert hits >= 3


def test_march_origin_value_is_fixed_and_zero_is_negative():
    """f_prev is f(origin) for the whole march; sign(0) is false (gpmedium.cpp:43,59)."""
    sc = _scene("sphere_single_128")
    osc = O.OracleScene(sc)
    # a ray that starts inside the mean sphere and leaves it: f_prev < 0, hit where f > 0
    rays = np.array([[0, 0, -500.0, 1, 0, 0, 0.0015, 3.4e38]], np.float32)
    m = osc.march(rays, np.array([42], np.uint32), None, fp64=False, track=False)
    assert m["hit"][0] == 1
    mu = osc.mean_eval(m["p"])[0, 0]
    assert 40 < mu + 70 < 100   # crosses near the sphere radius 70 +- psi


@pytest.mark.parametrize("fp64", [False, True])
def test_renewal_plus_invariants(fp64):
    """With the conditioning state of sampleCondition at C and the next seed:
    grad f_cond(C) = n exactly and f_cond(C) = mu(C)(1 - 1/sigma)."""
    sc = _scene("sphere_renewal_256")
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(4)
    n = 200
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    C3 = (np.array([0, 0, -500.0]) + d * rng.uniform(55, 85, (n, 1))).astype(np.float32)
    N3 = (d + rng.normal(0, 0.3, (n, 3))).astype(np.float32)
    N3 /= np.linalg.norm(N3, axis=1, keepdims=True)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    u, g = osc.condition(C3, N3, seeds, fp64=fp64)
    cond = np.concatenate([np.ones((n, 1)), C3, u[:, None], g], axis=1).astype(np.float32)
    f = osc.field_eval(C3, seeds, cond, fp64=True)
    mu = osc.mean_eval(C3)[:, 0]
    sigma = np.float32(sc["kernel"]["sigma"])
    _, l1 = O.psi_eval(sc["kernel"], C3, seeds)
    tol = 2e-5 * (1 + np.abs(mu) + l1)
    assert np.all(np.abs(f[:, 0] - mu * (1 - 1 / np.float64(sigma))) <= tol)
    assert np.all(np.abs(f[:, 1:] - N3).max(axis=1) <= 2e-5 * (1 + 18 * l1))
    # and the conditioning is actually doing work: the unconditioned field differs
    f0 = osc.field_eval(C3, seeds, None, fp64=True)
    assert np.mean(np.abs(f0[:, 1:] - N3).max(axis=1)) > 0.1


def test_single_realization_never_conditions():
    sc = _scene("sphere_single_128")
    osc = O.OracleScene(sc)
    C3 = np.array([[0, 0, -430.0]], np.float32)
    u, g = osc.condition(C3, np.array([[0, 0, 1.0]], np.float32), np.array([1], np.uint32), fp64=False)
    assert u[0] == 0 and np.all(g == 0)


def test_sphere_mean_definition():
    osc = O.OracleScene(_scene("sphere_renewal_256"))
    rng = np.random.default_rng(1)
    p = (np.array([0, 0, -500.0]) + rng.normal(0, 80, (100, 3))).astype(np.float32)
    m = osc.mean_eval(p)
    r = p.astype(np.float64) - np.array([0, 0, -500.0])
    nr = np.linalg.norm(r, axis=1)
    np.testing.assert_allclose(m[:, 0], nr - 70, atol=1e-4)
    np.testing.assert_allclose(m[:, 1:], r / nr[:, None], atol=1e-6)
