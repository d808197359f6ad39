"""The C++ host API renders through the C-ABI: a Cornell box built with Scene, Mesh,
Triangle, Lambert/Emission and Camera and rendered by Petrichor::Render equals the
Python binding's image of the same scene bit for bit (both integrators)."""
import numpy as np
import pytest

import cpp_api
from gpu_util import bits_equal, ctx, log
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("integrator", [0, 1])
def test_cpp_render_equals_python(gpu, integrator):
    log("C++ API render, integrator", integrator)
    sc = scenes.cornell_box()
    size = (64, 48, 4, 8, integrator, 5)
    cam, img = cpp_api.run(sc, size)
    ref, _ = ctx().scene(sc).build_bvh(16).render(64, 48, 4, 8, integrator, 5)
    assert bits_equal(cam, np.asarray(sc["camera"], np.float32))
    assert img.mean() > 0
    assert bits_equal(img, ref)
