"""The plain renderer's bounce loop replayed from a captured hipGraph gives the
accumulation the launches issued one by one give, on every replay (with eager
work and synchronisations between replays), across launches of one size with
different first passes, after a settings change (a new capture), and with an
emitter (the deferred light-sample queue inside the graph)."""
import numpy as np
import pytest

from mupsi_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("light", [False, True])
def test_graph_replays_equal_eager(ctx, light):
    from mupsi_amd import Scene
    sc = scenes.sphere_light_scene(48, 48) if light else scenes.preset_scene(
        "sphere_renewal_256", width=48, height=48)
    eager = Scene(ctx, sc)
    graph = Scene(ctx, sc)
    graph.set_graph(True)
    for k0, kc in ((0, 8), (8, 4), (20, 12), (3, 4)):  # launches of 4 passes each
        eager.reset()
        eager.render_passes(k0, kc, passes_per_launch=4)
        a, na = eager.read_accum()
        graph.reset()
        graph.render_passes(k0, kc, passes_per_launch=4)
        b, nb = graph.read_accum()
        assert na == nb == kc
        np.testing.assert_array_equal(a, b)
        assert graph.stats()["field_evals"] == eager.stats()["field_evals"] > 0
        eager.render_passes(100, 1)            # eager work between replays
    assert graph.graph_builds() == 1
    # a settings change drops the capture; the next launch captures again
    for s in (eager, graph):
        s.set_settings(max_bounce=3, gp_mode="renewal_plus", skip_medium=False)
        s.reset()
        s.render_passes(0, 4, passes_per_launch=4)
    np.testing.assert_array_equal(eager.read_accum()[0], graph.read_accum()[0])
    assert graph.graph_builds() == 2
    # a cancel flag takes the launch-by-launch path (polled between bounces)
    import ctypes
    flag = ctypes.c_int32(0)
    graph.reset()
    graph.render_passes(0, 4, 4, flag)
    eager.reset()
    eager.render_passes(0, 4, 4)
    np.testing.assert_array_equal(eager.read_accum()[0], graph.read_accum()[0])
    assert graph.graph_builds() == 2
    eager.close()
    graph.close()
