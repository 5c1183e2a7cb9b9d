"""Batched scenes: several scenes in shared wavefront launches give each scene
exactly the accumulation that rendering it alone gives (same passes), with and
without the captured graph, across several launches and uneven scenes."""
import numpy as np
import pytest

from mupsi_amd import scenes

pytestmark = pytest.mark.gpu


def _small_batch():
    scs = scenes.generated_batch(seed=11, n=6, max_res=48, max_subdivisions=3)
    for i, sc in enumerate(scs):       # cover ppc 1..8 on sphere and mesh means
        sc["kernel"]["points_per_cell"] = [1, 8, 3, 5, 2, 7][i]
    return scs


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_single_scene_renders(ctx, use_graph):
    from mupsi_amd import Batch, Scene
    scs = _small_batch()
    kinds = {sc["gp_medium"].get("mean_type", "sphere") for sc in scs}
    assert kinds == {"sphere", "mesh"}, kinds
    single = [Scene(ctx, sc) for sc in scs]
    passes = [3, 5, 2, 4, 6, 1]
    for s, k in zip(single, passes):
        s.render_passes(7, k)
    batched = [Scene(ctx, sc) for sc in scs]
    b = Batch(ctx, batched, tuning=dict(paths_per_launch=1 << 14, march_blocks_per_cu=2,
                                        use_graph=use_graph))
    b.render(7, passes)
    st = b.stats()
    assert st["launches"] > 1 and st["paths"] == sum(sc["width"] * sc["height"] * k
                                                     for sc, k in zip(scs, passes))
    if use_graph:
        assert st["graph_builds"] == 1
    for s1, s2, k in zip(single, batched, passes):
        a, na = s1.read_accum()
        c, nc = s2.read_accum()
        assert na == nc == k
        np.testing.assert_array_equal(a, c)
    b.close()


def test_batch_autotune_and_repeat(ctx):
    """Autotuning renders real passes; a second call continues the pass sequence."""
    from mupsi_amd import Batch, Scene
    scs = _small_batch()[:4]
    ref = [Scene(ctx, sc) for sc in scs]
    for s in ref:
        s.render_passes(0, 12)
    bs = [Scene(ctx, sc) for sc in scs]
    b = Batch(ctx, bs)
    b.render(0, 8)
    t = b.tuning()
    assert t["paths_per_launch"] > 0 and t["march_blocks_per_cu"] > 0
    b.render(8, 4)
    for s1, s2 in zip(ref, bs):
        np.testing.assert_array_equal(s1.read_accum()[0], s2.read_accum()[0])
    b.close()


def test_batch_probes_marcher_grid(ctx):
    """With a fixed launch size and the marcher grid left to the autotuner, the
    candidates are probed on real passes (several launches each) and the image
    stays that of the scenes rendered alone."""
    from mupsi_amd import Batch, Scene
    scs = _small_batch()[:3]
    ref = [Scene(ctx, sc) for sc in scs]
    for s in ref:
        s.render_passes(0, 16)
    bs = [Scene(ctx, sc) for sc in scs]
    cap = max(sc["width"] * sc["height"] for sc in scs)
    b = Batch(ctx, bs, tuning=dict(paths_per_launch=cap, march_blocks_per_cu=0, use_graph=True))
    b.render(0, 16)
    t = b.tuning()
    st = b.stats()
    assert t["paths_per_launch"] == cap and t["march_blocks_per_cu"] in (4, 8, 16)
    assert st["launches"] >= 6 and st["tune_ms"] > 0
    for s1, s2 in zip(ref, bs):
        np.testing.assert_array_equal(s1.read_accum()[0], s2.read_accum()[0])
    b.close()
