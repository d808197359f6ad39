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


def test_cpp_texture_save_png_and_hdr(gpu):
    # Texture2D::Save: the PNG holds the oracle's tonemapped bytes of the rendered image;
    # the HDR file decodes to it within RGBE precision (1/128 of a pixel's largest channel)
    import oracle_lib as O
    log("C++ API Texture2D::Save")
    saved = {}
    _, img = cpp_api.run(scenes.cornell_box(), (40, 24, 4, 8, 0, 3), saved)
    assert np.array_equal(saved["png"], O.tonemap_rgb8(img))
    peak = img.max(axis=2, keepdims=True)
    assert np.all(np.abs(saved["hdr"] - img) <= peak / 128 + 1e-30)
