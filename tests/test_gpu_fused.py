"""Medium hits shaded inside the march launch (mupsi_scene_set_fused, the
default) give the accumulation, march statistics and replayed path radiance
that the shading kernel gives when every hit goes back to it (fused off): bit
for bit, for renewal_plus and single_realization, group sizes 8 and 16, a
non-power-of-two cell size, a low bounce cap, hits that certain-sign
classification decides at points that were not evaluated, with and without
the graph.  A mesh mean is not fused: the switch leaves it as it is."""
import numpy as np
import pytest

from mupsi_amd import scenes

pytestmark = pytest.mark.gpu


def _scene(name, **kw):
    sub = kw.pop("mesh_subdivisions", 7)
    sc = scenes.preset_scene(name, width=40, height=40, mesh_subdivisions=sub)
    sc["kernel"].update(kw.pop("kernel", {}))
    sc.update(kw)
    return sc


CASES = [
    ("sphere_renewal_256", {}),
    ("sphere_single_128", {}),
    ("sphere_renewal_256", {"kernel": {"points_per_cell": 1, "length_scale": 1.3}}),   # 8 lanes per ray
    ("sphere_renewal_256", {"kernel": {"points_per_cell": 5, "length_scale": 0.37}}),  # 16 lanes, kr not 2^k
    ("sphere_renewal_256", {"max_bounce": 3}),
    # |psi| <= 0.1 < dt / 2: certain-sign classification decides hits at points
    # that were not evaluated (the window moves to the hit point's cell first)
    ("sphere_renewal_256", {"kernel": {"sigma": 0.1, "points_per_cell": 1, "length_scale": 0.2}}),
    ("icosphere_renewal_512", {"mesh_subdivisions": 3}),                                # not fused
]


@pytest.mark.parametrize("name,kw", CASES)
def test_fused_equals_unfused(ctx, name, kw):
    from mupsi_amd import Scene
    sc = _scene(name, **kw)
    fused = Scene(ctx, sc)
    plain = Scene(ctx, sc)
    plain.set_fused(False)
    for k0, kc in ((0, 6), (17, 3)):
        for s in (fused, plain):
            s.reset()
            s.render_passes(k0, kc)
        a, na = fused.read_accum()
        b, nb = plain.read_accum()
        assert na == nb == kc
        np.testing.assert_array_equal(a, b)
        sa, sb = fused.stats(), plain.stats()
        for key in ("field_evals", "march_steps", "march_calls"):
            assert sa[key] == sb[key], (key, sa[key], sb[key])
        assert sa["field_evals"] > 0
    # the graph replays the fused launches
    fused.set_graph(True)
    for _ in range(2):
        fused.reset()
        fused.render_passes(0, 6)
        plain.reset()
        plain.render_passes(0, 6)
        np.testing.assert_array_equal(fused.read_accum()[0], plain.read_accum()[0])
    # replayed paths (explicit pixel and pass lists) through both
    rng = np.random.default_rng(5)
    x = rng.integers(0, 40, 500).astype(np.int32)
    y = rng.integers(0, 40, 500).astype(np.int32)
    k = rng.integers(0, 64, 500).astype(np.int32)
    np.testing.assert_array_equal(fused.trace_paths(x, y, k), plain.trace_paths(x, y, k))
    fused.close()
    plain.close()
