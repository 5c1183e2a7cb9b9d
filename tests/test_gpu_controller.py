"""Controller semantics on the GPU renderer (src/controller/controller.cpp): the
progressive render's per-pass front/back buffer swap, cancel, the full render,
the GP setters and applyGpMedium (aniso reset to (1, 1, 1)), skip_medium, and
display / file output identical to the `mupsi` CLI.  Every image is compared
bit for bit with a freshly uploaded scene that has the same parameters."""
import copy
import subprocess
import time
from pathlib import Path

import numpy as np
import pytest

from mupsi_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def _fresh(ctx, sc, n):
    from mupsi_amd import Scene
    s = Scene(ctx, sc)
    s.render_passes(0, n)
    img = s.image().astype(np.float32)
    s.close()
    return img


def _edit(sc, kernel=None, **top):
    out = copy.deepcopy(sc)
    out["kernel"].update(kernel or {})
    out.update(top)
    return out


def _sphere_scene(res=48):
    return scenes.preset_scene("sphere_renewal_256", width=res, height=res)


def _controller(ctx, sc):
    from mupsi_amd import Controller
    c = Controller(ctx)
    assert c.load(sc)
    return c


def test_renderer_calls_back_after_every_pass(ctx):
    """Renderer::startRenderProgressive: on_spp(k) after pass k with the k-pass mean."""
    from mupsi_amd import Renderer, Scene
    sc = _sphere_scene(32)
    s = Scene(ctx, sc)
    r = Renderer(s)
    r.prepare_render()
    seen = []
    assert r.start_render_progressive(4, lambda k: seen.append((k, r.framebuffer().copy())))
    assert [k for k, _ in seen] == [1, 2, 3, 4] and r.progress() == 100
    for k, img in seen:
        np.testing.assert_array_equal(img.astype(np.float32), _fresh(ctx, sc, k))


def test_progressive_applies_gp_medium_with_isotropic_kernel(ctx):
    sc = _sphere_scene()
    assert sc["kernel"]["aniso"] != [1, 1, 1]
    c = _controller(ctx, sc)
    assert c.has_gp_medium() and c.gp_mode() == "renewal_plus" and c.gp_points_per_cell() == 3
    assert c.copy_display_to() is None          # nothing displayed yet
    c.start_progressive(5)
    observed = []
    t0 = time.time()
    while not c.join(0.01):
        observed.append(c.get_current_spp())
        assert time.time() - t0 < 90
    assert c.last_error is None
    assert observed == sorted(observed)
    assert c.get_current_spp() == 5 and c.is_spp_pass_done()
    c.ack_spp_pass()
    assert not c.is_spp_pass_done()
    raw = c.copy_display_raw_to()
    iso = _fresh(ctx, _edit(sc, {"aniso": [1, 1, 1]}), 5)
    np.testing.assert_array_equal(raw, iso)
    assert not np.array_equal(raw, _fresh(ctx, sc, 5))
    tone = c.copy_display_to()
    np.testing.assert_array_equal(tone, raw / (np.float32(1) + raw))
    flat = np.zeros(raw.size, np.float32)
    c.copy_display_to(flat)
    np.testing.assert_array_equal(flat.reshape(raw.shape), tone)
    c.close()


def test_full_render_keeps_the_loaded_medium_until_a_progressive_run(ctx):
    """start() renders spp passes without applyGpMedium; the medium a progressive
    run applied stays on the scene for the next full render."""
    sc = _sphere_scene(40)
    sc["spp"] = 3
    c = _controller(ctx, sc)
    assert c.get_spp() == 3
    c.start()
    assert c.join(90) and c.last_error is None and c.is_spp_pass_done()
    assert c.get_current_spp() == 0
    np.testing.assert_array_equal(c.copy_display_raw_to(), _fresh(ctx, sc, 3))
    c.start_progressive(1)
    assert c.join(90)
    c.start()
    assert c.join(90) and c.last_error is None
    np.testing.assert_array_equal(c.copy_display_raw_to(),
                                  _fresh(ctx, _edit(sc, {"aniso": [1, 1, 1]}), 3))
    c.close()


def test_cancel_keeps_the_last_completed_pass(ctx):
    sc = _sphere_scene(64)
    c = _controller(ctx, sc)
    c.start_progressive(1_000_000)
    t0 = time.time()
    while c.get_current_spp() < 3:
        assert time.time() - t0 < 60 and c.last_error is None
        time.sleep(0.002)
    c.stop()
    n = c.get_current_spp()
    assert 3 <= n < 1_000_000 and c.last_error is None
    np.testing.assert_array_equal(c.copy_display_raw_to(),
                                  _fresh(ctx, _edit(sc, {"aniso": [1, 1, 1]}), n))
    c.close()


def test_gp_setters_and_skip_medium_apply_on_the_next_run(ctx):
    sc = _sphere_scene()
    c = _controller(ctx, sc)
    c.set_gp_kernel_sigma(0.8)
    c.set_gp_kernel_length(3.0)
    c.set_gp_points_per_cell(5)
    c.set_gp_mode("single_realization")
    assert (c.gp_kernel_sigma(), c.gp_kernel_length(), c.gp_points_per_cell()) == (0.8, 3.0, 5)
    c.start_progressive(3)
    assert c.join(90) and c.last_error is None
    want = _edit(sc, {"sigma": 0.8, "length_scale": 3.0, "aniso": [1, 1, 1],
                      "points_per_cell": 5}, gp_mode="single_realization")
    np.testing.assert_array_equal(c.copy_display_raw_to(), _fresh(ctx, want, 3))
    c.start_progressive(2, skip_medium=True)
    assert c.join(90) and c.last_error is None
    np.testing.assert_array_equal(c.copy_display_raw_to(),
                                  _fresh(ctx, dict(want, skip_medium=True), 2))
    c.close()


def test_mesh_mean_from_scene_json_is_reloaded(ctx, tmp_path):
    """applyGpMedium re-reads the mean mesh file and rebuilds its device data."""
    from mupsi_amd.scene import write_scene_files
    sc = scenes.preset_scene("icosphere_renewal_512", width=40, height=40, mesh_subdivisions=4)
    path = write_scene_files(sc, tmp_path)
    c = _controller(ctx, str(path))
    assert c.config_path() == str(path)
    c.set_gp_points_per_cell(2)
    c.start_progressive(2)
    assert c.join(120) and c.last_error is None
    np.testing.assert_array_equal(c.copy_display_raw_to(),
                                  _fresh(ctx, _edit(sc, {"points_per_cell": 2}), 2))
    c.close()


def test_set_gp_medium_switches_mean_kind(ctx):
    """mupsi_scene_set_gp_medium from a sphere mean to a mesh mean and back."""
    from mupsi_amd import Scene
    sc = _sphere_scene(40)
    mesh = scenes.icosphere(3, displacement=0.05)
    tf = {"scale": 60.0, "position": [0.0, 0.0, -500.0]}
    s = Scene(ctx, sc)
    s.set_gp_medium("mesh", mesh=mesh, transform=tf, kernel=sc["kernel"])
    s.render_passes(0, 2)
    as_mesh = copy.deepcopy(sc)
    as_mesh["gp_medium"] = {"mean_type": "mesh", "mesh": mesh, "transform": tf}
    np.testing.assert_array_equal(s.image().astype(np.float32), _fresh(ctx, as_mesh, 2))
    s.reset()
    s.set_gp_medium("sphere", [0, 0, -500], 70.0, kernel=sc["kernel"])
    s.render_passes(0, 2)
    np.testing.assert_array_equal(s.image().astype(np.float32), _fresh(ctx, sc, 2))
    s.close()


def test_set_gp_medium_refusals_leave_the_scene_unchanged(ctx):
    from mupsi_amd import Batch, MupsiError, Scene
    sc = _sphere_scene(32)
    s = Scene(ctx, sc)
    with pytest.raises(MupsiError):
        s.set_gp_medium("sphere", [0, 0, -500], 70.0, kernel=dict(sc["kernel"], points_per_cell=0))
    with pytest.raises(MupsiError):
        s.set_gp_medium("sphere", [0, 0, -500], 70.0,
                        kernel=dict(sc["kernel"], aniso=[0, 0, 0]))
    b = Batch(ctx, [s], tuning=dict(paths_per_launch=1 << 12, march_blocks_per_cu=4,
                                    use_graph=False))
    with pytest.raises(MupsiError):
        s.set_gp_medium("sphere", [0, 0, -400], 50.0, kernel=sc["kernel"])
    with pytest.raises(MupsiError):
        s.set_settings(3, "single_realization", False)
    b.close()
    s.render_passes(0, 2)
    np.testing.assert_array_equal(s.image().astype(np.float32), _fresh(ctx, sc, 2))
    sdf = Scene(ctx, scenes.medium_scene("sdf", width=16, height=16))
    with pytest.raises(MupsiError):             # no gp_medium to replace
        sdf.set_gp_medium("sphere", [0, 0, -500], 70.0, kernel=sc["kernel"])
    sdf.close()
    s.close()


def test_display_files_equal_the_cli_output(ctx, tmp_path):
    from mupsi_amd.framebuffer import read_png
    from mupsi_amd.scene import write_scene_files
    exe = ROOT / "mupsi_amd" / "mupsi"
    assert exe.exists(), "the mupsi CLI is built with the library"
    sc = _sphere_scene(40)
    sc["spp"] = 3
    path = write_scene_files(sc, tmp_path)
    for ext in ("png", "pfm"):
        subprocess.run([str(exe), str(path), str(tmp_path / f"cli.{ext}")], check=True,
                       timeout=120, capture_output=True)
    c = _controller(ctx, str(path))
    c.start()
    assert c.join(90) and c.last_error is None
    c.save_display(tmp_path / "ctl.png")
    c.save_display(tmp_path / "ctl.pfm")
    np.testing.assert_array_equal(read_png(tmp_path / "ctl.png"), read_png(tmp_path / "cli.png"))
    assert (tmp_path / "ctl.pfm").read_bytes() == (tmp_path / "cli.pfm").read_bytes()
    c.close()
