"""GPU parity on the kernel parameters of the generated (batched) workload:
points_per_cell 1..8, cell sizes kr that are not powers of two (the cell
coordinate by division and the fmaf impulse positions), sphere and mesh
means, through the single-scene window marcher, whole-path replay of
generated scenes, and batched launches (each path's scene constants taken by
its scene id) against the oracle."""
import numpy as np
import pytest

from mupsi_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KRS = [0.05, 0.37, 1.3, 4.6]
AMB = 2e-6   # FP64 decision margin below which a ray or path is ambiguous


def _sigma(kr, ppc):
    """A sigma that gives psi an amplitude of about 2 (the mean's variation over
    a few grid steps), so that the march decisions depend on psi; at least the
    top of the generated range."""
    var = ppc * np.pi ** 1.5 / kr ** 3
    return float(max(2.7, 2.0 * np.sqrt(var / (27.0 * ppc))))


def _rays(rng, n, center, r_enclose, r_mean):
    c = np.asarray(center, np.float64)
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = c + d0 * r_enclose * rng.uniform(0.6, 1.0, (n, 1))
    tgt = c + rng.normal(size=(n, 3)) * r_mean * 0.8
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[rng.uniform(size=n) < 0.25] *= -1
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = np.finfo(np.float32).max
    short = rng.uniform(size=n) < 0.2
    rays[short, 7] = rng.uniform(0.5, 2.0, short.sum()) * r_enclose
    return rays


def _cond(rng, n, center, radius, dev):
    import torch
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    C3 = (np.asarray(center) + d * radius * rng.uniform(0.9, 1.1, (n, 1))).astype(np.float32)
    cond = np.zeros((n, 8), np.float32)
    cond[:, 0] = rng.uniform(size=n) < 0.7
    cond[:, 1:4] = C3
    cond[:, 4] = rng.normal(0, 3, n)
    cond[:, 5:8] = rng.normal(0, 2, (n, 3))
    c = torch.from_numpy(cond).to(dev)
    return cond, {"active": c[:, 0].to(torch.uint8), "C": c[:, 1:4].contiguous(),
                  "u": c[:, 4].contiguous(), "g": c[:, 5:8].contiguous()}


def _scene(mean, kr, ppc, sigma):
    if mean == "sphere":
        sc = scenes.preset_scene("sphere_renewal_256", width=16, height=16)
        geo = ([0, 0, -500.0], 130.0, 70.0)
    else:
        sc = scenes.preset_scene("icosphere_renewal_512", width=16, height=16, mesh_subdivisions=3)
        geo = ([0, 0, -5.0], 3.0, 2.0)
    sc["kernel"] = {"sigma": sigma, "length_scale": kr, "aniso": [1, 1, 1], "points_per_cell": ppc}
    return sc, geo


@pytest.mark.parametrize("ppc", [1, 2, 5, 6, 7, 8])
@pytest.mark.parametrize("mean", ["sphere", "mesh"])
def test_window_march_generated_kernels(ctx, dev, mean, ppc):
    """Hit flag and hit t bits equal to the FP64 oracle on every unambiguous ray,
    normals within 1e-5, for each cell size (conditioning on the sphere mean)."""
    import torch
    from mupsi_amd import Scene, batch
    for j, kr in enumerate(KRS):
        sig = _sigma(kr, ppc)
        sc, (c, r_enc, r_mean) = _scene(mean, kr, ppc, sig)
        s = Scene(ctx, sc)
        osc = O.OracleScene(sc)
        rng = np.random.default_rng(1000 * ppc + 10 * j + (mean == "mesh"))
        n = 240
        rays = _rays(rng, n, c, r_enc, r_mean)
        seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        cond, cdev = _cond(rng, n, c, r_mean, dev) if mean == "sphere" else (None, None)
        ref = osc.march(rays, seeds, cond, fp64=True, track=True)
        got = batch.march(s, torch.from_numpy(rays).to(dev),
                          torch.from_numpy(seeds.view(np.int32)).to(dev), cdev)
        hit = got["hit"].cpu().numpy()
        t = got["t"].cpu().numpy()
        nrm = got["n"].cpu().numpy()
        unamb = ref["amb"] > AMB
        ok = (hit == ref["hit"]) & ((~hit.astype(bool)) | (t.view(np.uint32) == ref["t"].view(np.uint32)))
        print(f"{mean} ppc {ppc} kr {kr} sigma {sig:.3g}: unambiguous {unamb.mean():.3f}, "
              f"hits {int(hit.sum())}, agree {ok[unamb].mean():.4f}")
        assert ok[unamb].all(), f"kr {kr}: {(~ok[unamb]).sum()} unambiguous rays differ"
        assert unamb.mean() > 0.9 and hit.sum() > 20
        m = unamb & hit.astype(bool)
        err = np.abs(nrm[m] - ref["n"][m]).max()
        assert err < 1e-5, (kr, err)
        s.close()


def _generated(names=None):
    """Three scenes of the generated workload at small size: a sphere mean and two
    mesh means, one of them with a cell size that is not a power of two."""
    scs = scenes.generated_batch(seed=5, n=12, max_res=40, max_subdivisions=3)
    mesh = [sc for sc in scs if sc["gp_medium"]["mean_type"] == "mesh"]
    sphere = [sc for sc in scs if sc["gp_medium"]["mean_type"] == "sphere"]
    picked = [sphere[0], mesh[0], mesh[1]]
    for sc in picked:
        sc["max_bounce"] = 6
    return picked


def test_generated_kernels_cover_non_power_of_two_cells():
    scs = _generated()
    krs = [sc["kernel"]["length_scale"] for sc in scs]
    assert any(sc["gp_medium"]["mean_type"] == "mesh" and np.frexp(k)[0] != 0.5
               for sc, k in zip(scs, krs))
    assert {sc["gp_medium"]["mean_type"] for sc in scs} == {"sphere", "mesh"}


def _compare_paths(got, ref, amb):
    unamb = amb > AMB
    close = np.all(np.abs(got - ref) <= 1e-4 * (1 + np.abs(ref)), axis=1)
    return unamb, close


@pytest.mark.parametrize("i", [0, 1, 2])
def test_generated_scene_replay(ctx, i):
    """Whole paths of generated scenes (single-scene launches) against the
    oracle: every path whose marches are unambiguous agrees."""
    from mupsi_amd import Scene
    sc = _generated()[i]
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(300 + i)
    n = 400
    W, H = sc["width"], sc["height"]
    x = rng.integers(0, W, n).astype(np.int32)
    y = rng.integers(0, H, n).astype(np.int32)
    k = rng.integers(0, 64, n).astype(np.int32)
    got = s.trace_paths(x, y, k)
    ref, amb = osc.trace_paths_amb(x, y, k)
    unamb, close = _compare_paths(got, ref, amb)
    print(f"{sc['name']} ({sc['gp_medium']['mean_type']}, kr {sc['kernel']['length_scale']:.3g}, "
          f"ppc {sc['kernel']['points_per_cell']}): unambiguous {unamb.sum()}/{n}, "
          f"agree {close[unamb].mean():.4f}, nonzero {np.count_nonzero(ref.sum(1))}")
    assert close[unamb].all(), np.flatnonzero(unamb & ~close)[:10]
    assert unamb.mean() > 0.9


def test_batched_render_matches_oracle(ctx):
    """The batched renderer (paths of three generated scenes in shared launches,
    scene constants by scene id) against the oracle's per-path radiance summed
    in pass order: every pixel whose paths are all unambiguous agrees."""
    from mupsi_amd import Batch, Scene
    scs = _generated()
    for sc in scs:
        sc["width"] = sc["height"] = 20
    passes = [3, 2, 3]
    bs = [Scene(ctx, sc) for sc in scs]
    b = Batch(ctx, bs, tuning=dict(paths_per_launch=1 << 11, march_blocks_per_cu=2, use_graph=True))
    b.render(4, passes)
    assert b.stats()["launches"] > 1
    for sc, s, kp in zip(scs, bs, passes):
        sums, cnt = s.read_accum()
        assert cnt == kp
        osc = O.OracleScene(sc)
        W, H = sc["width"], sc["height"]
        ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
        x = np.tile(xs.ravel(), kp).astype(np.int32)
        y = np.tile(ys.ravel(), kp).astype(np.int32)
        k = np.repeat(np.arange(4, 4 + kp), W * H).astype(np.int32)
        rgb, amb = osc.trace_paths_amb(x, y, k)
        ref = np.zeros((H * W, 3))
        for q in range(kp):   # FP64 sums in pass order, as the accumulator adds them
            ref += rgb[q * W * H:(q + 1) * W * H].astype(np.float64)
        pix_amb = (amb.reshape(kp, H * W) <= AMB).any(axis=0)
        got = sums.reshape(-1, 3)
        bad = np.any(np.abs(got - ref) > 1e-4 * (1 + np.abs(ref)), axis=1)
        print(f"{sc['name']}: {pix_amb.sum()} ambiguous pixels of {H * W}, "
              f"{(bad & ~pix_amb).sum()} unambiguous pixels differ")
        assert not (bad & ~pix_amb).any()
        assert pix_amb.mean() < 0.1
    b.close()


@pytest.mark.parametrize("prim", [0, -1])
def test_icosphere7_queries_bit_exact(ctx, dev, prim):
    """The 327,680-triangle mesh of the mesh-mean workload (its deep BVH and,
    for the mean mesh, its distance grid): triangle ids, t, u bits, occlusion,
    closest distance bits and triangle, inside bits against the oracle."""
    import torch
    from mupsi_amd import Scene, batch
    m = scenes.icosphere(7, displacement=0.05)
    xf = {"scale": 2.0, "position": [0.0, 0.0, -5.0]}
    sc = {"width": 8, "height": 8, "camera": dict(scenes.CAMERA), "max_bounce": 4,
          "gp_mode": "renewal_plus",
          "kernel": {"sigma": 0.3, "length_scale": 0.5, "aniso": [1, 1, 1], "points_per_cell": 4},
          "gp_medium": {"mean_type": "mesh", "mesh": m, "transform": xf},
          "primitives": [{"type": "mesh", "mesh": m, "bsdf": "lambertian", "transform": xf}]}
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(77)
    n = 20000
    c = np.array([0, 0, -5.0])
    o = (c + rng.normal(0, 3.0, (n, 3))).astype(np.float32)
    d = (c + rng.normal(0, 1.0, (n, 3)) - o).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 7] = np.finfo(np.float32).max
    rt = torch.from_numpy(rays).to(dev)
    tri, t, u, v = (x.cpu().numpy() for x in batch.mesh_intersect(s, prim, rt))
    otri, ot, ou, ov = osc.mesh_intersect(prim, rays)
    assert np.array_equal(tri, otri)
    h = tri >= 0
    assert h.sum() > 5000
    assert np.array_equal(t[h].view(np.uint32), ot[h].view(np.uint32))
    assert np.array_equal(u[h].view(np.uint32), ou[h].view(np.uint32))
    assert np.array_equal(v[h].view(np.uint32), ov[h].view(np.uint32))
    occ = batch.mesh_occluded(s, prim, rt).cpu().numpy()
    assert np.array_equal(occ, osc.mesh_occluded(prim, rays))
    # points around the surface (where the march queries them) and far away
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    p = (c + dirs * rng.uniform(1.5, 2.6, (n, 1))).astype(np.float32)
    p[: n // 10] = (c + rng.normal(0, 6.0, (n // 10, 3))).astype(np.float32)
    pt = torch.from_numpy(p).to(dev)
    dd, ctri, cl = (x.cpu().numpy() for x in batch.mesh_closest(s, prim, pt))
    od, otri2, ocl = osc.mesh_closest(prim, p)
    assert np.array_equal(dd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(ctri, otri2)
    assert np.array_equal(cl.view(np.uint32), ocl.view(np.uint32))
    ins = batch.mesh_inside(s, prim, pt).cpu().numpy()
    assert np.array_equal(ins, osc.mesh_inside(prim, p))
    assert 0.2 < ins.mean() < 0.8
