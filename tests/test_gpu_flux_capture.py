"""libprofit's flux-capturing suite (tests/test_model.h:445-551): a sky of
level 1, convolved or not with a flat PSF, evaluated without a mask and then
with a square mask at every position inside the model; the masked image must
equal the unmasked one on the mask (relative 1e-4, as the reference asserts)
and be exactly 0 outside it -- with no convolver, BRUTE and FFT, at
finesampling 1 and 2 (returned downsampled).  All mask positions of one
configuration run as one batch (pk_evaluate_batch); the first, middle and
last positions also through the single evaluation (pk_evaluate)."""
import numpy as np
import pytest

from libprofit_amd import abi
from libprofit_amd.model import evaluate_batch, evaluate_spec

pytestmark = pytest.mark.gpu

# (model size, mask size, psf size) of test_empty_mask, test_tiny,
# test_small_mask_{small,medium,full,bigger}_psf
CASES = [(20, 0, 5),
         (5, 2, 5), (5, 3, 5), (5, 4, 5), (5, 3, 4), (5, 4, 4), (5, 5, 4), (3, 1, 2),
         (20, 4, 5), (20, 5, 5), (20, 6, 5), (20, 5, 6), (20, 6, 6), (20, 7, 6),
         (20, 5, 9), (20, 5, 20), (20, 5, 30)]


def _spec(model_size, psf_size, conv, fs, mask=None):
    sky = abi.make_component("sky", bg=1.0, convolve=conv is not None)
    return abi.ModelSpec(model_size, model_size, [sky], psf=np.ones((psf_size, psf_size)), mask=mask,
                         finesampling=fs, return_finesampled=False,
                         convolver=conv if conv is not None else "brute")


def _check(masked, full, mask, what):
    assert masked.shape == full.shape, what
    expect = np.where(mask.astype(bool), full, 0.0)
    zero = expect == 0
    assert np.all(masked[zero] == 0), what  # EXPECT_0
    nz = ~zero
    assert np.all(np.abs(masked[nz] - expect[nz]) <= 1e-4 * np.abs(expect[nz])), what


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("model_size,mask_size,psf_size", CASES)
def test_flux_is_captured(ctx64, ctx32, prec, model_size, mask_size, psf_size):
    print("CASE flux capture", prec, model_size, mask_size, psf_size, flush=True)
    ctx = ctx64 if prec == "fp64" else ctx32
    for conv in (None, "brute", "fft"):
        for fs in (1, 2):
            full, _, _ = evaluate_spec(ctx, _spec(model_size, psf_size, conv, fs))
            masks = []
            for oy in range(model_size - mask_size + 1):
                for ox in range(model_size - mask_size + 1):
                    mk = np.zeros((model_size, model_size), np.uint8)
                    mk[oy:oy + mask_size, ox:ox + mask_size] = 1
                    masks.append(mk)
            specs = [_spec(model_size, psf_size, conv, fs, mk) for mk in masks]
            for k, ((img, _off), mk) in enumerate(zip(evaluate_batch(ctx, specs), masks)):
                _check(img, full, mk, (conv, fs, k, "batch"))
            for k in sorted({0, len(specs) // 2, len(specs) - 1}):
                img, _, _ = evaluate_spec(ctx, specs[k])
                _check(img, full, masks[k], (conv, fs, k, "single"))
