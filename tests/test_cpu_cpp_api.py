"""The C++ host API (include/petrichor/Petrichor.h) on the host: the Camera set-up
operations (Core/Camera.cpp:14-50, Camera.h:71-83) must leave the same state as the
Python restatement used by every scene, bit for bit."""
import os

import numpy as np
import pytest

import cpp_api
from petrichor_amd import scenes


@pytest.mark.skipif(not os.path.exists(cpp_api.DRIVER), reason="C++ driver not built")
def test_camera_state_matches_python_restatement():
    sc = scenes.cornell_box()
    cam, _ = cpp_api.run(sc)
    assert np.array_equal(cam.view(np.uint32), np.asarray(sc["camera"], np.float32).view(np.uint32))
