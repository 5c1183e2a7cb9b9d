"""The named benchmark workloads: every preset builds, sizes and kernels are the
documented ones, unknown names are rejected, the generated batch is reproducible."""
import numpy as np
import pytest

from mupsi_amd import scenes


def test_presets_build_with_their_sizes():
    want = {"sphere_single_128": (128, 16, 1.0), "sphere_renewal_256": (256, 64, 2.0),
            "icosphere_renewal_512": (512, 64, 0.5), "guided_sphere_quad_1024": (1024, 256, 2.0)}
    assert set(scenes.PRESETS) == set(want)
    for name, (res, spp, kr) in want.items():
        sc = scenes.preset_scene(name, mesh_subdivisions=2)
        assert (sc["width"], sc["height"], sc["spp"]) == (res, res, spp), name
        assert sc["kernel"]["length_scale"] == kr, name
    small = scenes.preset_scene("sphere_renewal_256", width=8, height=6)
    assert (small["width"], small["height"]) == (8, 6)


def test_unknown_preset_is_rejected():
    with pytest.raises(ValueError, match="no preset scene"):
        scenes.preset_scene("configuration_2")


def test_generated_batch_is_reproducible():
    a = scenes.generated_batch(seed=3, n=5, max_res=64, max_subdivisions=2)
    b = scenes.generated_batch(seed=3, n=5, max_res=64, max_subdivisions=2)
    assert len(a) == 5
    for x, y in zip(a, b):
        assert x["kernel"] == y["kernel"] and x["width"] == y["width"]
        if x["gp_medium"]["mean_type"] == "mesh":
            np.testing.assert_array_equal(x["gp_medium"]["mesh"]["vertices"],
                                          y["gp_medium"]["mesh"]["vertices"])
