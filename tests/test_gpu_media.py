"""GPU parity of the SDF and external-GP (grid) media against the oracle: field
values and gradients, sphere-traced marches, whole paths, the scene.json route
and batched launches."""
import numpy as np
import pytest

from mupsi_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KINDS = [("sdf", "sphere"), ("sdf", "mesh"), ("grid", "sphere")]


def _scene(kind, mean, res=96):
    return scenes.medium_scene(kind, width=res, height=res, mean=mean)


def _rays(rng, n):
    """Rays from inside the enclosing sphere (r 130) toward the medium centre."""
    c = np.array([0.0, 0.0, -500.0])
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = c + d0 * rng.uniform(75.0, 125.0, (n, 1))
    d = (c + rng.normal(0, 50.0, (n, 3))) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = np.finfo(np.float32).max
    return rays


@pytest.mark.parametrize("kind,mean", KINDS)
def test_field_matches_oracle(ctx, dev, kind, mean):
    import torch
    from mupsi_amd import Scene, batch
    sc = _scene(kind, mean)
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(21)
    n = 4000
    p = (np.array([0.0, 0.0, -500.0]) + rng.normal(0, 60.0, (n, 3))).astype(np.float32)
    seeds = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    got = batch.field_eval(s, torch.from_numpy(p).to(dev),
                           torch.from_numpy(seeds.view(np.int32)).to(dev)).cpu().numpy()
    ref = osc.field_eval(p, seeds, fp64=False)
    same = got[:, 0].view(np.uint32) == ref[:, 0].astype(np.float32).view(np.uint32)
    assert same.mean() > 0.999, same.mean()
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(got[:, 1:], ref[:, 1:], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("kind,mean", KINDS)
def test_march_matches_oracle(ctx, dev, kind, mean):
    import torch
    from mupsi_amd import Scene, batch
    sc = _scene(kind, mean)
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(22)
    n = 600 if mean == "sphere" else 200
    rays = _rays(rng, n)
    seeds = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    ref = osc.march(rays, seeds, None, fp64=False, track=False)
    got = batch.march(s, torch.from_numpy(rays).to(dev), torch.from_numpy(seeds.view(np.int32)).to(dev))
    hit = got["hit"].cpu().numpy().astype(bool)
    t = got["t"].cpu().numpy()
    nrm = got["n"].cpu().numpy()
    rh = ref["hit"].astype(bool)
    agree = (hit == rh) & (~hit | (t.view(np.uint32) == ref["t"].view(np.uint32)))
    print(f"{kind}/{mean}: {agree.mean():.4f} of rays identical, hits {hit.sum()}")
    assert agree.mean() > 0.99 and hit.sum() > 20
    both = hit & rh & agree
    np.testing.assert_allclose(nrm[both], ref["n"][both], atol=1e-4)


@pytest.mark.parametrize("kind,mean", KINDS)
def test_path_replay_matches_oracle(ctx, kind, mean):
    from mupsi_amd import Scene
    sc = _scene(kind, mean)
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(23)
    n = 1500 if mean == "sphere" else 400
    W = sc["width"]
    x = rng.integers(int(0.3 * W), int(0.7 * W), n).astype(np.int32)
    y = rng.integers(int(0.3 * W), int(0.7 * W), n).astype(np.int32)
    k = rng.integers(0, 16, n).astype(np.int32)
    got = s.trace_paths(x, y, k)
    ref = osc.trace_paths(x, y, k)
    close = np.all(np.abs(got - ref) <= 1e-4 * (1 + np.abs(ref)), axis=1)
    print(f"{kind}/{mean}: {close.sum()}/{n} paths agree, nonzero {np.count_nonzero(ref.sum(1))}")
    assert close.mean() >= 0.97


@pytest.mark.parametrize("kind", ["sdf", "grid"])
def test_scene_json_route_matches_dict(ctx, tmp_path, kind):
    """scene.json with sdf_medium / external_gp_medium (grid files beside it)
    renders exactly what the scene dict renders."""
    from mupsi_amd import Scene
    from mupsi_amd.scene import write_scene_files
    sc = _scene(kind, "sphere", res=48)
    path = write_scene_files(sc, tmp_path)
    a = Scene(ctx, sc)
    b = Scene(ctx, str(path))
    a.render_passes(0, 2)
    b.render_passes(0, 2)
    np.testing.assert_array_equal(a.read_accum()[0], b.read_accum()[0])


def test_batch_with_sdf_and_grid_media_equals_single_renders(ctx):
    from mupsi_amd import Batch, Scene
    scs = [scenes.medium_scene("sdf", width=48, height=48),
           scenes.medium_scene("grid", width=40, height=40),
           scenes.preset_scene("sphere_renewal_256", width=32, height=32)]
    passes = [3, 2, 4]
    single = [Scene(ctx, sc) for sc in scs]
    for s, k in zip(single, passes):
        s.render_passes(5, k)
    batched = [Scene(ctx, sc) for sc in scs]
    b = Batch(ctx, batched, tuning=dict(paths_per_launch=1 << 13, march_blocks_per_cu=4,
                                        use_graph=True))
    b.render(5, passes)
    for s1, s2 in zip(single, batched):
        np.testing.assert_array_equal(s1.read_accum()[0], s2.read_accum()[0])
    b.close()
