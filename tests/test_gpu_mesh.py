"""GPU parity of the BVH / mesh queries (bit-exact triangle ids) and of a mesh-mean march."""
import numpy as np
import pytest

from mupsi_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mesh_scene(sub=4):
    m = scenes.icosphere(sub, displacement=0.05)
    sc = {"width": 16, "height": 16, "camera": dict(scenes.CAMERA), "max_bounce": 4,
          "gp_mode": "renewal_plus",
          "kernel": {"sigma": 0.3, "length_scale": 0.5, "aniso": [1, 1, 1], "points_per_cell": 4},
          "gp_medium": {"mean_type": "mesh", "mesh": m,
                        "transform": {"scale": 2.0, "position": [0.0, 0.0, -5.0]}},
          "primitives": [{"type": "mesh", "mesh": m, "bsdf": "lambertian",
                          "transform": {"scale": 2.0, "position": [0.0, 0.0, -5.0]}}]}
    return sc


def _rays(rng, n, center, spread):
    o = (np.asarray(center) + rng.normal(0, spread, (n, 3))).astype(np.float32)
    d = (np.asarray(center) + rng.normal(0, spread * 0.3, (n, 3)) - o).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 0.0
    rays[:, 7] = np.finfo(np.float32).max
    return rays


@pytest.mark.parametrize("prim", [0, -1])
def test_mesh_queries_bit_exact(ctx, dev, prim):
    import torch
    from mupsi_amd import Scene, batch
    sc = _mesh_scene()
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(7)
    rays = _rays(rng, 20000, [0, 0, -5.0], 4.0)
    rt = torch.from_numpy(rays).to(dev)
    tri, t, u, v = (x.cpu().numpy() for x in batch.mesh_intersect(s, prim, rt))
    otri, ot, ou, ov = osc.mesh_intersect(prim, rays)
    assert np.array_equal(tri, otri)
    h = tri >= 0
    assert h.sum() > 5000
    assert np.array_equal(t[h].view(np.uint32), ot[h].view(np.uint32))
    assert np.array_equal(u[h].view(np.uint32), ou[h].view(np.uint32))
    occ = batch.mesh_occluded(s, prim, rt).cpu().numpy()
    assert np.array_equal(occ, osc.mesh_occluded(prim, rays))
    p = (np.array([0, 0, -5.0]) + rng.normal(0, 2.5, (20000, 3))).astype(np.float32)
    pt = torch.from_numpy(p).to(dev)
    d, ctri, cl = (x.cpu().numpy() for x in batch.mesh_closest(s, prim, pt))
    od, otri2, ocl = osc.mesh_closest(prim, p)
    assert np.array_equal(d.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(ctri, otri2)
    ins = batch.mesh_inside(s, prim, pt).cpu().numpy()
    assert np.array_equal(ins, osc.mesh_inside(prim, p))


def test_mesh_mean_march(ctx, dev):
    import torch
    from mupsi_amd import Scene, batch
    sc = _mesh_scene(sub=3)
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(8)
    n = 300
    c = np.array([0, 0, -5.0])
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = c + d0 * 2.9
    d = (c + rng.normal(0, 0.8, (n, 3))) - o
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = rng.uniform(3.0, 12.0, n)   # bounded march lengths keep the oracle fast
    # full-length rays (far_t = FLT_MAX, 4000 grid points) exercise the certain-sign skip,
    # including rays that run along -x through the parity ray's shadow slab
    full = np.arange(n) % 3 == 0
    rays[full, 7] = np.finfo(np.float32).max
    slab = np.arange(n) % 6 == 0
    rays[slab, 0:3] = c + np.array([1.5, 0.0, 0.0]) + rng.normal(0, 0.3, (slab.sum(), 3))
    rays[slab, 3:6] = np.array([-1.0, 0.0, 0.0]) + rng.normal(0, 0.02, (slab.sum(), 3))
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    ref64 = osc.march(rays, seeds, None, fp64=True, track=True)
    got = batch.march(s, torch.from_numpy(rays).to(dev), torch.from_numpy(seeds.view(np.int32)).to(dev))
    hit = got["hit"].cpu().numpy()
    t = got["t"].cpu().numpy()
    unamb = ref64["amb"] > 2e-6
    ok = (hit == ref64["hit"]) & ((~hit.astype(bool)) | (t.view(np.uint32) == ref64["t"].view(np.uint32)))
    assert ok[unamb].all()
    assert hit.sum() > 20


@pytest.mark.parametrize("prim", [0, -1])
def test_traversal_cost_equals_oracle(ctx, dev, prim):
    """The device traversal tests the same nodes and triangles as the oracle's
    restatement of bvh.cpp (same topology, same near/far order, same pruning)."""
    import torch
    from mupsi_amd import Scene, batch
    sc = _mesh_scene()
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(12)
    p = (np.array([0, 0, -5.0]) + rng.normal(0, 2.5, (20000, 3))).astype(np.float32)
    pt = torch.from_numpy(p).to(dev)
    for kind in (0, 1):
        nodes, tris = (x.cpu().numpy() for x in batch.mesh_query_cost(s, prim, kind, pt))
        onodes, otris = osc.mesh_query_cost(prim, kind, p)
        assert np.array_equal(nodes, onodes), kind
        assert np.array_equal(tris, otris), kind
        assert tris.sum() > 20000
    batch.mesh_closest(s, prim, pt)
    assert batch.last_kernel_ms(ctx) > 0.0
    # inside() as the library runs it: one traversal collecting the hits (kind 3)
    # tests fewer nodes in total than the reference's loop of casts (kind 1)
    n3, t3 = (x.cpu().numpy().astype(np.int64) for x in batch.mesh_query_cost(s, prim, 3, pt))
    n1, t1 = (x.cpu().numpy().astype(np.int64) for x in batch.mesh_query_cost(s, prim, 1, pt))
    assert 0 < n3.sum() < n1.sum() and t3.sum() > 0


def test_deep_tree_queries_bit_exact(ctx, dev):
    """A BVH deeper than 60 levels (a 70-triangle chain): inside() runs the
    reference's casts there (its 64-entry stacks may fill, and the walk without
    a stack is used only where they cannot); inside bits, hit triangles and t
    bits against the oracle, whose casts fill the same stack the same way."""
    import torch
    from mupsi_amd import Scene, batch
    m = scenes.chain_mesh()
    sc = {"width": 8, "height": 8, "camera": dict(scenes.CAMERA), "max_bounce": 2,
          "primitives": [{"type": "mesh", "mesh": m, "bsdf": "lambertian"}]}
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(21)
    n = 4000
    x = m["vertices"][::3, 0].astype(np.float64)
    px = x[rng.integers(0, len(x), n)] * rng.uniform(0.3, 1.2, n)
    p = np.stack([px, rng.uniform(-0.01, 0.05, n) * px, rng.uniform(-0.01, 0.05, n) * px], 1)
    p = p.astype(np.float32)
    pt = torch.from_numpy(p).to(dev)
    ins = batch.mesh_inside(s, 0, pt).cpu().numpy()
    assert np.array_equal(ins, osc.mesh_inside(0, p))
    assert 0.2 < ins.mean() < 0.8
    n1, t1 = (a.cpu().numpy() for a in batch.mesh_query_cost(s, 0, 1, pt))
    n3, t3 = (a.cpu().numpy() for a in batch.mesh_query_cost(s, 0, 3, pt))
    on1, ot1 = osc.mesh_query_cost(0, 1, p)
    assert np.array_equal(n1, on1) and np.array_equal(t1, ot1)
    assert np.array_equal(n3, n1) and np.array_equal(t3, t1)   # casts: no single walk here
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = p
    rays[:, 3] = 1.0
    rays[:, 4:6] = rng.normal(0, 0.01, (n, 2))
    rays[:, 7] = np.finfo(np.float32).max
    tri, t, u, v = (a.cpu().numpy() for a in batch.mesh_intersect(s, 0, torch.from_numpy(rays).to(dev)))
    otri, ot, ou, ov = osc.mesh_intersect(0, rays)
    assert np.array_equal(tri, otri)
    h = tri >= 0
    assert h.sum() > n // 4
    assert np.array_equal(t[h].view(np.uint32), ot[h].view(np.uint32))


def test_inside_leaf_bins_at_leaf_bounds(ctx, dev):
    """inside() over the leaf bins of a mean mesh at the points where a bin
    lookup could go wrong: y and z on a leaf's bounds and one ulp either side
    (identity transform, so the local point is the query point), x spread
    across the mesh.  Bits against the oracle's casts; the triangles tested equal
    the stackless walk's (kind 4) wherever neither falls back to the casts, and
    no more bin entries are read than the walk tests boxes."""
    import torch
    from mupsi_amd import Scene, batch
    m = scenes.icosphere(5, displacement=0.05)
    sc = {"width": 8, "height": 8, "camera": dict(scenes.CAMERA), "max_bounce": 2,
          "gp_mode": "renewal_plus",
          "kernel": {"sigma": 0.3, "length_scale": 0.5, "aniso": [1, 1, 1], "points_per_cell": 4},
          "gp_medium": {"mean_type": "mesh", "mesh": m},
          "primitives": [{"type": "mesh", "mesh": m, "bsdf": "lambertian"}]}
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    nodes, _, _ = O.bvh_build(m["vertices"], m["faces"])
    lb = nodes[(nodes.view(np.uint32)[:, 7] & 0x80000000) != 0, :6]
    rng = np.random.default_rng(31)
    i = rng.integers(0, len(lb), 6000)
    y = np.where(rng.random(6000) < 0.5, lb[i, 1], lb[i, 4]).astype(np.float32)
    z = np.where(rng.random(6000) < 0.5, lb[i, 2], lb[i, 5]).astype(np.float32)
    step = rng.integers(-1, 2, (2, 6000))
    y = np.where(step[0] < 0, np.nextafter(y, np.float32(-np.inf)),
                 np.where(step[0] > 0, np.nextafter(y, np.float32(np.inf)), y))
    z = np.where(step[1] < 0, np.nextafter(z, np.float32(-np.inf)),
                 np.where(step[1] > 0, np.nextafter(z, np.float32(np.inf)), z))
    x = rng.uniform(-1.2, 1.2, 6000).astype(np.float32)
    p = np.ascontiguousarray(np.stack([x, y, z], 1).astype(np.float32))
    pt = torch.from_numpy(p).to(dev)
    ins = batch.mesh_inside(s, -1, pt).cpu().numpy()
    assert np.array_equal(ins, osc.mesh_inside(-1, p))
    assert np.array_equal(batch.mesh_inside(s, 0, pt).cpu().numpy(), ins)   # no bins on a primitive
    n3, t3 = (a.cpu().numpy().astype(np.int64) for a in batch.mesh_query_cost(s, -1, 3, pt))
    n4, t4 = (a.cpu().numpy().astype(np.int64) for a in batch.mesh_query_cost(s, -1, 4, pt))
    n1, _ = (a.cpu().numpy().astype(np.int64) for a in batch.mesh_query_cost(s, -1, 1, pt))
    single = (n4 < n1) | (t4 == 0)
    print(f"bins: {n3.mean():.1f} entries + {t3.mean():.2f} tris per query; "
          f"walk {n4.mean():.1f} boxes + {t4.mean():.2f} tris; {single.mean():.4f} single pass")
    assert single.mean() > 0.9
    assert np.array_equal(t3[single], t4[single])
    assert n3.sum() < n4.sum()


@pytest.mark.parametrize("scale", [2.0, 0.35])
def test_closest_across_distance_grids(ctx, dev, scale):
    """Closest-point queries of a mean mesh at points in each of its distance
    grids (the fine grid around the box, the outer grid to 1.5 box lengths, the
    far grid to 8) and beyond all three: the search starts from the grid's upper
    bound there, so a bound below the true distance would change the result.
    Distance bits, triangle ids and closest points against the oracle's unbounded
    search."""
    import torch
    from mupsi_amd import Scene, batch
    m = scenes.icosphere(4, displacement=0.05)
    c = np.array([0.3, -0.2, -5.0])
    xf = {"scale": scale, "position": list(c)}
    sc = {"width": 8, "height": 8, "camera": dict(scenes.CAMERA), "max_bounce": 2,
          "gp_mode": "renewal_plus",
          "kernel": {"sigma": 0.3, "length_scale": 0.5, "aniso": [1, 1, 1], "points_per_cell": 4},
          "gp_medium": {"mean_type": "mesh", "mesh": m, "transform": xf}}
    s = Scene(ctx, sc)
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(41)
    ext = 2.1 * scale  # the box's length (world)
    pts = []
    for lo, hi in ((0.9, 1.3), (1.4, 2.2), (2.5, 4.5), (6.0, 16.0), (20.0, 40.0)):
        dirs = rng.normal(size=(3000, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        pts.append(c + dirs * (rng.uniform(lo, hi, (3000, 1)) * ext / 2))
    p = np.concatenate(pts).astype(np.float32)
    pt = torch.from_numpy(p).to(dev)
    d, tri, cl = (x.cpu().numpy() for x in batch.mesh_closest(s, -1, pt))
    od, otri, ocl = osc.mesh_closest(-1, p)
    assert np.array_equal(tri, otri)
    assert np.array_equal(d.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cl.view(np.uint32), ocl.view(np.uint32))
