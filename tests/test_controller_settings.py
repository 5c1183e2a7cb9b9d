"""Controller::load's kept settings (controller.cpp:40-137) on the host: the
reference's defaults and the keys applyGpMedium rebuilds from."""
import json

import pytest

from mupsi_amd import scenes
from mupsi_amd.controller import load_settings


def test_defaults_of_an_empty_scene_object():
    st = load_settings({}, spp=7)
    assert st == {"spp": 7, "max_bounce": 3, "output": "test.png",
                  "gp_mode": "single_realization", "has_gp_medium": False}


def test_sphere_gp_medium_and_kernel_keys():
    sc = scenes.preset_scene("sphere_renewal_256")
    st = load_settings(sc)
    assert st["spp"] == 64 and st["max_bounce"] == 10 and st["gp_mode"] == "renewal_plus"
    assert st["mean_type"] == "sphere" and st["mean_center"] == [0.0, 0.0, -500.0]
    assert st["mean_radius"] == 70.0
    assert (st["sigma"], st["length_scale"], st["points_per_cell"]) == (1.5, 2.0, 3)


def test_kernel_and_mean_defaults():
    st = load_settings({"gp_medium": {"mean_center": [1, 2, 3]}})
    assert st["mean_type"] == "sphere" and st["mean_radius"] == 70.0
    assert (st["sigma"], st["length_scale"], st["points_per_cell"]) == (1.0, 1.0, 3)


def test_mesh_mean_from_json(tmp_path):
    j = json.loads(json.dumps({"gp_medium": {"mean_type": "mesh", "file": "m.obj",
                                             "transform": {"scale": 2.0}},
                               "output": "x.png", "spp": 3}))
    st = load_settings(j)
    assert st["mean_mesh"] == "m.obj" and st["mean_mesh_scale"] == 2.0
    assert st["mean_mesh_pos"] == [0.0, 0.0, 0.0] and st["output"] == "x.png" and st["spp"] == 3


def test_unknown_mean_type_is_refused():
    with pytest.raises(ValueError):
        load_settings({"gp_medium": {"mean_type": "torus"}})
