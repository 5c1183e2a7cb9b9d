"""GPU parity of the GP medium: field with conditioning, march hits, Renewal+ solve."""
import numpy as np
import pytest

from mupsi_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _scene(name):
    sc = scenes.preset_scene(name, width=32, height=32)
    return sc


def _cond_states(rng, n, center, radius, k):
    """Random active conditioning states near the mean surface."""
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    C = (np.asarray(center) + d * radius * rng.uniform(0.9, 1.1, (n, 1))).astype(np.float32)
    u = rng.normal(0, 3, n).astype(np.float32)
    g = rng.normal(0, 2, (n, 3)).astype(np.float32)
    act = (rng.uniform(size=n) < 0.7).astype(np.float32)
    cond = np.concatenate([act[:, None], C, u[:, None], g], axis=1).astype(np.float32)
    return cond


def _cond_dev(cond, dev):
    import torch
    c = torch.from_numpy(cond).to(dev)
    return {"active": c[:, 0].to(torch.uint8), "C": c[:, 1:4].contiguous(), "u": c[:, 4].contiguous(),
            "g": c[:, 5:8].contiguous()}


@pytest.mark.parametrize("name,k", [("sphere_single_128", 1), ("sphere_renewal_256", 2)])
def test_field_eval_conditioned(ctx, dev, name, k):
    import torch
    from mupsi_amd import Scene, batch
    sc = _scene(name)
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(k)
    n = 8000
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    p = (np.array([0, 0, -500.0]) + dirs * rng.uniform(40, 140, (n, 1))).astype(np.float32)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    cond = _cond_states(rng, n, [0, 0, -500.0], 70.0, sc["kernel"])
    cond[:, 1:4] = p + rng.normal(0, 1.5, (n, 3)).astype(np.float32)  # C near p: large correction
    ref = osc.field_eval(p, seeds, cond, fp64=True)
    got = batch.field_eval(s, torch.from_numpy(p).to(dev), torch.from_numpy(seeds.view(np.int32)).to(dev),
                           _cond_dev(cond, dev)).cpu().numpy().astype(np.float64)
    _, l1 = O.psi_eval(sc["kernel"], p, seeds)
    scale = 1.0 + np.abs(ref[:, 0] - 0) + l1 + np.abs(cond[:, 4]) + np.linalg.norm(cond[:, 5:8], axis=1)
    assert np.all(np.abs(got[:, 0] - ref[:, 0]) <= 1e-5 * scale)
    assert np.all(np.abs(got[:, 1:] - ref[:, 1:]).max(axis=1) <= 1e-5 * scale)


def _rays_into_medium(rng, n, center, r_enclose, r_mean):
    """Rays starting on/in the enclosing sphere aimed around the mean sphere."""
    c = np.asarray(center, np.float64)
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = c + d0 * r_enclose * rng.uniform(0.6, 1.0, (n, 1))
    tgt = c + rng.normal(size=(n, 3)) * r_mean * 0.8
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    flip = rng.uniform(size=n) < 0.25          # some rays aimed away (escape / exit)
    d[flip] *= -1
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = np.finfo(np.float32).max
    short = rng.uniform(size=n) < 0.1          # finite far_t (shadow-ray style dt)
    rays[short, 7] = rng.uniform(5, 300, short.sum())
    return rays


@pytest.mark.parametrize("name,k", [("sphere_single_128", 1), ("sphere_renewal_256", 2)])
def test_march_hits_bit_exact(ctx, dev, name, k):
    import torch
    from mupsi_amd import Scene, batch
    sc = _scene(name)
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(10 + k)
    n = 600
    rays = _rays_into_medium(rng, n, [0, 0, -500.0], 130.0, 70.0)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    cond = _cond_states(rng, n, [0, 0, -500.0], 70.0, sc["kernel"]) if name == "sphere_renewal_256" else None
    ref32 = osc.march(rays, seeds, cond, fp64=False, track=False)
    ref64 = osc.march(rays, seeds, cond, fp64=True, track=True)
    got = batch.march(s, torch.from_numpy(rays).to(dev), torch.from_numpy(seeds.view(np.int32)).to(dev),
                      _cond_dev(cond, dev) if cond is not None else None)
    hit = got["hit"].cpu().numpy()
    t = got["t"].cpu().numpy()
    nrm = got["n"].cpu().numpy()
    # a ray is ambiguous when some visited grid point has |f| within a few FP32 roundings of
    # zero (relative to |mu| + L1 mass).  Cell indices are FP32 decisions on both sides
    # (the FP64 oracle shares the FP32 grid points), so cell boundaries are not ambiguous.
    unamb = ref64["amb"] > 2e-6
    print(f"{name}: unambiguous {unamb.mean():.3f}, hits {int(hit.sum())}")
    # decisions: identical to the FP64 oracle on every unambiguous ray
    same_hit = hit == ref64["hit"]
    same_t = (~hit.astype(bool)) | (t.view(np.uint32) == ref64["t"].view(np.uint32))
    ok = same_hit & same_t
    assert ok[unamb].all(), f"{(~ok[unamb]).sum()} unambiguous rays differ"
    assert unamb.mean() > 0.95
    # and with the FP32-faithful oracle on every unambiguous ray (its own FP32
    # rounding can only move a decision whose FP64 margin is below the threshold)
    agree32 = (hit == ref32["hit"]) & ((~hit.astype(bool)) | (t == ref32["t"]))
    assert agree32[unamb].all(), f"{(~agree32[unamb]).sum()} unambiguous rays differ from FP32"
    # normals at common hits within 1e-5 of FP64 (normalized gradient)
    m = unamb & hit.astype(bool) & ref64["hit"].astype(bool)
    if m.any():
        err = np.abs(nrm[m] - ref64["n"][m]).max()
        assert err < 1e-5, err
    assert hit.sum() > 10


def test_renewal_solve_fp64(ctx, dev):
    import torch
    from mupsi_amd import Scene, batch
    sc = _scene("sphere_renewal_256")
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(5)
    n = 4000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    C3 = (np.array([0, 0, -500.0]) + d * rng.uniform(60, 80, (n, 1))).astype(np.float32)
    N3 = (d + rng.normal(0, 0.2, (n, 3))).astype(np.float32)
    N3 /= np.linalg.norm(N3, axis=1, keepdims=True)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    u_ref, g_ref = osc.condition(C3, N3, seeds, fp64=True)
    u, g = batch.condition(s, torch.from_numpy(C3).to(dev), torch.from_numpy(N3).to(dev),
                           torch.from_numpy(seeds.view(np.int32)).to(dev))
    u, g = u.cpu().numpy(), g.cpu().numpy()
    _, l1 = O.psi_eval(sc["kernel"], C3, seeds)
    assert np.all(np.abs(u - u_ref) <= 1e-5 * (1 + np.abs(u_ref) + l1))
    L2 = 18.0
    assert np.all(np.abs(g - g_ref).max(axis=1) <= 1e-5 * (1 + np.abs(g_ref).max(axis=1) + L2 * l1))


def test_renewal_invariants_on_gpu(ctx, dev):
    """grad f_cond(C) = n and f_cond(C) = mu(C)(1 - 1/sigma) for the GPU solve + field."""
    import torch
    from mupsi_amd import Scene, batch
    sc = _scene("sphere_renewal_256")
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(6)
    n = 2000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    C3 = (np.array([0, 0, -500.0]) + d * rng.uniform(60, 80, (n, 1))).astype(np.float32)
    N3 = d.astype(np.float32)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    Ct, Nt = torch.from_numpy(C3).to(dev), torch.from_numpy(N3).to(dev)
    st = torch.from_numpy(seeds.view(np.int32)).to(dev)
    u, g = batch.condition(s, Ct, Nt, st)
    cond = {"active": torch.ones(n, dtype=torch.uint8, device=dev), "C": Ct, "u": u.float(),
            "g": g.float()}
    f = batch.field_eval(s, Ct, st, cond).cpu().numpy().astype(np.float64)
    mu = osc.mean_eval(C3)[:, 0]
    sigma = sc["kernel"]["sigma"]
    _, l1 = O.psi_eval(sc["kernel"], C3, seeds)
    tol = 2e-5 * (1 + np.abs(mu) + l1 + np.abs(u.cpu().numpy()))
    assert np.all(np.abs(f[:, 0] - mu * (1 - 1 / sigma)) <= tol)
    assert np.all(np.abs(f[:, 1:] - N3).max(axis=1) <= 2e-5 * (1 + 18.0 * l1))
