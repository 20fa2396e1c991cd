"""The allocation-free prediction entry point (cf_predict_batch_device_noalloc) against
cf_predict_batch_device: two calls in a row, the second with fewer frames, give the same labels and
posteriors bit for bit; its refusals (forest not yet on the device, oversized batch) come back as
CF_ERR_ARG."""
import numpy as np
import pytest

import curfil_amd as cf
from helpers import config0_dataset

pytestmark = pytest.mark.gpu


def _bits(a):
    return a.view(np.uint32)


def test_noalloc_prediction_equals_device_prediction(ctx):
    rgb, depth, labels = config0_dataset()
    n, H, W, _ = rgb.shape
    im = cf.Images(ctx, n, W, H)
    for i in range(n):
        im.set_rgbd_u8(i, rgb[i], depth[i])
    forest = cf.synth_forest(ctx, im, depth.astype(np.int32), trees=3, depth=9, classes=5, seed=21)
    forest.normalizeHistograms(0.0, False)
    C = 5
    dp = cf.DeviceBuffer(ctx, 4 * C * H * W * 4)
    dl = cf.DeviceBuffer(ctx, 4 * H * W)
    # not yet on the device: refused, nothing written
    with pytest.raises(cf.CurfilError, match="not resident"):
        forest.predict_batch_device_noalloc(im, [0], dp.ptr, dl.ptr)
    for idx in ([3, 0, 2, 1], [1, 3]):
        k = len(idx)
        forest.predict_batch_device(im, idx, dp.ptr, dl.ptr)
        ref_p = dp.to_host(np.float32, (4, C, H, W))[:k].copy()
        ref_l = dl.to_host(np.uint8, (4, H, W))[:k].copy()
        forest.predict_batch_device_noalloc(im, idx, dp.ptr, dl.ptr)
        got_p = dp.to_host(np.float32, (4, C, H, W))[:k]
        got_l = dl.to_host(np.uint8, (4, H, W))[:k]
        assert np.array_equal(got_l, ref_l), idx
        assert np.array_equal(_bits(got_p), _bits(ref_p)), idx
        print(f"frames {idx}: labels {np.bincount(got_l.ravel(), minlength=C)}")
    with pytest.raises(cf.CurfilError, match="65536"):
        forest.predict_batch_device_noalloc(im, np.zeros(65537, np.int32), dp.ptr, dl.ptr)
    with pytest.raises(cf.CurfilError, match="out of range"):
        forest.predict_batch_device_noalloc(im, [4], dp.ptr, dl.ptr)
    dp.close()
    dl.close()
