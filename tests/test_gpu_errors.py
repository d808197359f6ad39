"""Error behaviour of the C-ABI (status codes, no abort)."""
import numpy as np
import pytest

import petrichor_amd as P
from gpu_util import ctx
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu


def test_bad_material_id_is_rejected(gpu):
    sc = scenes.cornell_box()
    sc["tri_mat"] = sc["tri_mat"].copy()
    sc["tri_mat"][3] = 99
    with pytest.raises(P.PtError, match="material id out of range"):
        ctx().scene(sc)


def test_bad_render_settings_are_rejected(gpu):
    b = ctx().scene(scenes.cornell_box()).build_bvh()
    with pytest.raises(P.PtError):
        b.render(0, 16, 1, 8)
    with pytest.raises(P.PtError):
        b.render(16, 16, 1, 8, integrator=7)


def test_unsupported_width_is_rejected(gpu):
    s = ctx().scene(scenes.cornell_box())
    with pytest.raises(P.PtError, match="width"):
        s.build_bvh(width=3)


def test_empty_scene_renders_base_colour(gpu):
    sc = scenes.cornell_box()
    sc["tri_pos"] = np.zeros((0, 9), np.float32)
    sc["tri_mat"] = np.zeros(0, np.int32)
    sc["light_tris"] = np.zeros(0, np.int32)
    sc["env_param"] = np.array([0.5, 0.5, 0.5, 0, 1], np.float32)
    img, st = ctx().scene(sc).build_bvh().render(8, 8, 2, 4)
    assert st["queries"] == 8 * 8 * 2
    assert (img > 0).all()
