"""mupsi_march_noalloc (the library's fixed device scratch, no allocation)
against mupsi_march on the same rays: hit flags, t, hit points, normals and
step counts identical, on two calls in a row (the second smaller, so the
scratch keeps the first call's contents beyond its rays), in both modes, and
a size above the scratch maximum refused before any launch."""
import numpy as np
import pytest

from mupsi_amd import scenes

pytestmark = pytest.mark.gpu


def _rays(rng, n):
    c = np.array([0, 0, -500.0])
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = c + d0 * 130.0 * rng.uniform(0.6, 1.0, (n, 1))
    d = (c + rng.normal(size=(n, 3)) * 56.0) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[rng.uniform(size=n) < 0.25] *= -1
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = np.finfo(np.float32).max
    short = rng.uniform(size=n) < 0.2
    rays[short, 7] = rng.uniform(5, 300, short.sum())
    return rays


def _cond(rng, n, dev):
    import torch
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    C3 = (np.array([0, 0, -500.0]) + d * 70.0 * rng.uniform(0.9, 1.1, (n, 1))).astype(np.float32)
    return {"active": torch.from_numpy((rng.uniform(size=n) < 0.7).astype(np.uint8)).to(dev),
            "C": torch.from_numpy(C3).to(dev),
            "u": torch.from_numpy(rng.normal(0, 3, n).astype(np.float32)).to(dev),
            "g": torch.from_numpy(rng.normal(0, 2, (n, 3)).astype(np.float32)).to(dev)}


def _same(a, b, mode):
    # transmittance (mode 1) computes no normals: nrm is left unwritten
    for key in ("hit", "t", "p", "n", "steps") if mode == 0 else ("hit", "t", "p", "steps"):
        x, y = a[key].cpu().numpy(), b[key].cpu().numpy()
        if key in ("t", "p", "n"):
            hit = a["hit"].cpu().numpy().astype(bool)
            x, y = x[hit], y[hit]
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert np.array_equal(x, y), key


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_march_equals_march(ctx, dev, mode):
    import torch
    from mupsi_amd import Scene, batch
    s = Scene(ctx, scenes.preset_scene("sphere_renewal_256", width=16, height=16))
    rng = np.random.default_rng(70 + mode)
    for n in (4000, 1500):   # second call smaller than the first
        rays = torch.from_numpy(_rays(rng, n)).to(dev)
        seeds = torch.from_numpy(rng.integers(0, 2**32, n, dtype=np.uint64)
                                 .astype(np.uint32).view(np.int32)).to(dev)
        cond = _cond(rng, n, dev)
        ref = batch.march(s, rays, seeds, cond, mode=mode)
        got = batch.march(s, rays, seeds, cond, mode=mode, noalloc=True)
        _same(ref, got, mode)
        assert int(ref["hit"].sum()) > 10
    s.close()


def test_noalloc_march_refuses_above_maximum(ctx, dev):
    import torch
    from mupsi_amd import MupsiError, Scene, batch, _abi
    s = Scene(ctx, scenes.preset_scene("sphere_renewal_256", width=16, height=16))
    n = _abi.MARCH_NOALLOC_MAX_RAYS + 1
    rays = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    seeds = torch.zeros(n, dtype=torch.int32, device=dev)
    with pytest.raises(MupsiError, match="scratch maximum"):
        batch.march(s, rays, seeds, noalloc=True)
    s.close()
