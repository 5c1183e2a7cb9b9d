"""Empty and minimal inputs through the C ABI: every batch entry point accepts a
zero-length batch (and returns empty outputs without touching its arrays), a
render of zero passes leaves the accumulator as it was, and a 1x1 image renders."""
import numpy as np
import pytest

from mupsi_amd import scenes
from tests.test_gpu_mesh import _mesh_scene

pytestmark = pytest.mark.gpu


def test_empty_batches(ctx, dev):
    import torch
    from mupsi_amd import Scene, batch
    kernel = {"sigma": 1.5, "length_scale": 2.0, "aniso": [1, 1, 4], "points_per_cell": 3}
    p0 = torch.zeros((0, 3), dtype=torch.float32, device=dev)
    s0 = torch.zeros((0,), dtype=torch.int32, device=dev)
    out, l1 = batch.psi_eval(ctx, kernel, p0, s0)
    assert out.shape == (0, 4) and l1.shape == (0,)
    cs, pos, w = batch.cell_impulses(ctx, kernel, torch.zeros((0, 3), dtype=torch.int32, device=dev), s0)
    assert cs.shape == (0,) and pos.shape == (0, 3, 3) and w.shape == (0, 3)

    s = Scene(ctx, scenes.preset_scene("sphere_renewal_256", width=8, height=8))
    assert batch.field_eval(s, p0, s0).shape == (0, 4)
    r = batch.march(s, torch.zeros((0, 8), dtype=torch.float32, device=dev), s0)
    assert all(v.shape[0] == 0 for v in r.values())
    u, g = batch.condition(s, p0, p0, s0)
    assert u.shape == (0,) and g.shape == (0, 3)
    got = s.trace_paths(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert got.shape[0] == 0

    m = Scene(ctx, _mesh_scene())
    rays0 = torch.zeros((0, 8), dtype=torch.float32, device=dev)
    for prim in (0, -1):
        assert all(x.shape == (0,) for x in batch.mesh_intersect(m, prim, rays0))
        assert batch.mesh_occluded(m, prim, rays0).shape == (0,)
        d, tri, cl = batch.mesh_closest(m, prim, p0)
        assert d.shape == (0,) and tri.shape == (0,) and cl.shape == (0, 3)
        assert batch.mesh_inside(m, prim, p0).shape == (0,)
        nodes, tris = batch.mesh_query_cost(m, prim, 3, p0)
        assert nodes.shape == (0,) and tris.shape == (0,)
    m.close()
    s.close()


def test_zero_passes_and_one_pixel(ctx):
    from mupsi_amd import Scene
    s = Scene(ctx, scenes.preset_scene("sphere_renewal_256", width=8, height=8))
    s.render_passes(0, 2)
    before, n = s.read_accum()
    s.render_passes(2, 0)
    after, n2 = s.read_accum()
    assert n2 == n == 2
    np.testing.assert_array_equal(before, after)
    s.close()

    one = Scene(ctx, scenes.preset_scene("sphere_renewal_256", width=1, height=1))
    one.render_passes(0, 3)
    acc, k = one.read_accum()
    assert k == 3 and acc.size == 3 and np.isfinite(acc).all() and (acc >= 0).all()
    one.close()
