"""resynth_image_synth_device_noalloc, the entry that allocates nothing, through the C
ABI: on consecutive calls (the second on a smaller image, over the scratch the first one
left) it returns what resynth_image_synth_device returns, byte for byte, in the exact
serial mode and in the batched mode with one visit per launch (both deterministic); the
largest image it accepts fits its scratch with room for the serial mode's PRNG stream; a
larger one is refused with RESYNTH_ERROR_TOO_LARGE."""
import ctypes

import numpy as np
import pytest

import resynthesizer_amd as R
from resynthesizer_amd import workloads as W

pytestmark = pytest.mark.gpu

BPP = {R.T_RGB: 3, R.T_RGBA: 4, R.T_Gray: 1, R.T_GrayA: 2}


def _call(entry, img, mask, fmt, params, opt):
    """Runs entry on device copies of img (H, W, bpp) and mask (H, W); returns
    (error, stats, synthesized image)."""
    import torch
    h, w = mask.shape
    d_img = torch.from_numpy(np.ascontiguousarray(img).reshape(h, -1)).cuda()
    d_mask = torch.from_numpy(np.ascontiguousarray(mask)).cuda()
    torch.cuda.synchronize()
    st = R.ResynthStats()
    cancel = ctypes.c_int(0)
    err = entry(d_img.data_ptr(), w, h, w * BPP[fmt], d_mask.data_ptr(), w, None, 0, fmt, ctypes.byref(params),
                R.PROGRESS(), None, ctypes.byref(cancel), ctypes.byref(opt), ctypes.byref(st))
    torch.cuda.synchronize()
    return err, st, d_img.cpu().numpy().reshape(img.shape)


def _case(h, w, fmt, seed, hole):
    img = W.grating_noise(h, w, BPP[fmt], seed)
    mask = np.zeros((h, w), np.uint8)
    y0, x0, y1, x1 = hole
    mask[y0:y1, x0:x1] = 255
    return img, mask


@pytest.mark.parametrize("mode,batch", [(R.MODE_SERIAL, 0), (R.MODE_BATCHED, 1)])
def test_noalloc_matches_device_entry_on_consecutive_calls(mode, batch):
    L = R.lib()
    cases = [
        # larger first, then smaller over the scratch the first call left
        (_case(120, 150, R.T_RGBA, 3, (30, 40, 80, 100)), R.T_RGBA, R.default_params(maxProbeCount=60)),
        (_case(70, 90, R.T_RGB, 4, (20, 25, 50, 60)), R.T_RGB,
         R.default_params(maxProbeCount=40, patchSize=20, matchContextType=2)),
    ]
    for (img, mask), fmt, params in cases:
        print("no-alloc case", img.shape, "mode", mode, "batch", batch, flush=True)
        opt = R.options(mode=mode, batch=batch)
        e0, s0, o0 = _call(L.resynth_image_synth_device, img, mask, fmt, params, opt)
        e1, s1, o1 = _call(L.resynth_image_synth_device_noalloc, img, mask, fmt, params, opt)
        assert e0 == e1 == 0
        assert np.array_equal(o0, o1)
        assert (s0.passes, s0.visits, s0.candidates, list(s0.betters)) == \
               (s1.passes, s1.visits, s1.candidates, list(s1.betters))
        assert s1.meanError == s0.meanError
        assert 0 < s1.scratchBytes <= R.NOALLOC_SCRATCH_BYTES
        sel = mask != 0
        assert np.array_equal(o1[~sel], img[~sel])


@pytest.mark.parametrize("order", [1, 2])
def test_noalloc_largest_image_fits_its_scratch(order):
    """1024 x 1024 RGBA (the maximum), nearly all of it a target (the largest target and
    ordering buffers), random order (sorted on the device) and inward order: it runs, and
    its scratch leaves room for the serial mode's 4 MiB PRNG stream."""
    h = w = 1024
    img, mask = _case(h, w, R.T_RGBA, 5, (16, 16, h - 16, w - 16))
    params = R.default_params(maxProbeCount=8, patchSize=10, matchContextType=order)
    opt = R.options(mode=R.MODE_BATCHED, max_passes=1)
    print("no-alloc largest image", flush=True)
    err, st, out = _call(R.lib().resynth_image_synth_device_noalloc, img, mask, R.T_RGBA, params, opt)
    assert err == 0 and st.nTargets == (h - 32) * (w - 32)
    sel = mask != 0
    assert np.array_equal(out[~sel], img[~sel]) and np.array_equal(out[..., 3], img[..., 3])
    print("scratch bytes used by the largest image:", st.scratchBytes, flush=True)
    assert st.scratchBytes + (4 << 20) + (1 << 16) <= R.NOALLOC_SCRATCH_BYTES, st.scratchBytes


def test_noalloc_refuses_images_beyond_its_maximum():
    opt = R.options(mode=R.MODE_BATCHED)
    st = R.ResynthStats()
    err = R.lib().resynth_image_synth_device_noalloc(None, 1025, 1024, 1025 * 3, None, 1025, None, 0, R.T_RGB,
                                                     ctypes.byref(R.default_params()), R.PROGRESS(), None, None,
                                                     ctypes.byref(opt), ctypes.byref(st))
    assert err == R.RESYNTH_ERROR_TOO_LARGE
