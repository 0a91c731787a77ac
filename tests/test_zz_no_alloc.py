"""gbm3d_denoise_device_noalloc (workspace in the library's static scratch
array, no device allocation) equals gbm3d_denoise_device bit for bit, on two
calls in a row, the second with a smaller image (the array keeps the first
call's contents), and rejects an image whose workspace does not fit before
launching anything."""
import numpy as np
import pytest

import gbm3d_amd
from gbm3d_amd import _abi, synth
from gbm3d_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _run(ctx, fn, noisy, sigma, opts):
    m, n = noisy.shape
    dI = DeviceArray.from_host(noisy.astype(np.float32))
    dU, dU1 = DeviceArray((m, n)), DeviceArray((m, n))
    fn(dI.ptr, m, n, sigma, opts, dU.ptr, dU1.ptr)
    ctx.synchronize()
    return dU.to_host(), dU1.to_host()


@pytest.mark.parametrize("opts", [None, {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}],
                         ids=["default", "p8"])
def test_noalloc_equals_device_entry_point(ctx, opts):
    for seed, (m, n), sigma in ((710, (256, 200), 25.0), (711, (96, 80), 15.0)):   # second call smaller
        _, noisy = synth.make_pair(seed, m, n, sigma)
        U_ref, U1_ref = _run(ctx, ctx.denoise_device, noisy, sigma, opts)
        U, U1 = _run(ctx, ctx.denoise_device_noalloc, noisy, sigma, opts)
        assert np.isfinite(U).all()
        np.testing.assert_array_equal(U, U_ref)
        np.testing.assert_array_equal(U1, U1_ref)


def test_noalloc_rejects_images_beyond_the_scratch_array(ctx):
    need = gbm3d_amd.validate(None, 1024, 1024)
    assert need > (128 << 20)
    _, noisy = synth.make_pair(712, 1024, 1024, 25.0)
    with pytest.raises(gbm3d_amd.GBM3DError, match="scratch array") as e:
        _run(ctx, ctx.denoise_device_noalloc, noisy, 25.0, None)
    assert e.value.status == _abi.GBM3D_ERR_ARG
    synchronize()
    _, small = synth.make_pair(713, 64, 64, 25.0)
    U_ref, _ = _run(ctx, ctx.denoise_device, small, 25.0, None)
    U, _ = _run(ctx, ctx.denoise_device_noalloc, small, 25.0, None)
    np.testing.assert_array_equal(U, U_ref)
