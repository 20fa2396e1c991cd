"""CPU-side checks of the product library: it loads, exports every symbol the
C ABI header declares, validates configuration like the reference, and its
synthetic workload generator is deterministic.  No GPU compute here."""
import os
import re

import numpy as np
import pytest

import curfil_amd as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "curfil_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cf_\w+)\(", text, re.M)))


def test_header_symbols_exported():
    names = _declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(cf.lib, n), n


def test_struct_layouts_match_header():
    assert cf.SAMPLE_DTYPE.itemsize == 16
    assert cf.NODE_DTYPE.itemsize == 56
    assert cf.NODE_DTYPE.fields["score"][1] == 32
    assert cf.NODE_DTYPE.fields["feature"][1] == 40
    assert cf.NODE_DTYPE.fields["is_leaf"][1] == 51


def test_configuration_validation_like_reference():
    # random_tree.h:263-272 and random_tree.cpp:114-124
    with pytest.raises(ValueError, match="maxImages"):
        cf.TrainingConfiguration(4711, 500, 100, 32, 8, 60, 10, 10, maxImages=2, imageCacheSize=5)
    with pytest.raises(ValueError, match="acceleration mode"):
        cf.TrainingConfiguration(4711, 500, 100, 32, 8, 60, 10, 10, accelerationMode="tpu")
    c = cf.TrainingConfiguration(4711, 500, 100, 32, 8, 60, 10, 10, ignored_labels=[0]).to_c()
    assert (c.random_seed, c.feature_count, c.box_radius, c.num_ignored, c.ignored_labels[0]) == (4711, 100, 60, 1, 0)


def test_synthetic_frames_deterministic_and_shaped():
    a = cf.synth_rgbd(4711, 160, 120, 12, 5, 0.05)
    b = cf.synth_rgbd(4711, 160, 120, 12, 5, 0.05)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    rgb, depth, labels = a
    assert rgb.shape == (120, 160, 3) and depth.shape == (120, 160) and labels.shape == (120, 160)
    assert labels.max() <= 5
    zero = np.mean(depth == 0)
    assert 0.03 < zero < 0.07
    valid = depth[depth > 0]
    assert valid.min() >= 500 and valid.max() <= 5000
    c = cf.synth_rgbd(4712, 160, 120, 12, 5, 0.05)
    assert not np.array_equal(c[0], rgb)


def test_context_without_device_fails_cleanly():
    import ctypes
    n = 0
    try:
        import torch
        n = torch.cuda.device_count()
    except Exception:
        pass
    if n > 0:
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    assert cf.lib.cf_ctx_create(0, ctypes.byref(h)) != cf.CF_OK
