"""Device chi^2 and its adjoint gradient (pk_batch_evaluate_grad) against
central finite differences of the FP64 oracle with the refinement pattern
held fixed (relative step 1e-4), and against FD of the device forward path."""
import sys

import numpy as np
import pytest

from libprofit_amd import abi, workloads
from libprofit_amd.model import Batch
from grad_util import FitModel, bulge_disk256_model, chi2_of, fd_steps, oracle_fd_grad

pytestmark = pytest.mark.gpu

# Per parameter: |adjoint - FD| <= 2 err(FD) + RTOL |FD| + FLOOR max|FD|, FD the
# frozen-pattern central differences of the oracle converged over a ladder of
# steps (grad_util.oracle_fd_grad) and err(FD) their own error bound; RTOL is
# the precision of the adjoint: FP64 arithmetic, or FP32 profile derivatives
# whose sums can cancel (nser).  Differences at relative step 1e-4 are
# printed beside them.
RTOL = {"fp64": 1e-7, "fp32": 1e-4}
FLOOR = {"fp64": 0.0, "fp32": 0.0}


def announce(name):
    print("CASE", name, flush=True)
    sys.stdout.flush()


def small_fit(seed=0):
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)
    comps = [("sersic", dict(xcen=20.3, ycen=19.6, mag=15, re=3.0, nser=2.5, ang=30, axrat=0.7, box=0.0)),
             ("sersic", dict(xcen=20.1, ycen=19.9, mag=14.5, re=7.0, nser=0.8, ang=50, axrat=0.45, box=0.1)),
             ("sky", dict(bg=2e-9))]  # ~10 % of the galaxies' peak pixel (magzero 0)
    fm = FitModel(40, comps, psf)
    return fm


def make_data(oracle, fm_true, seed=0):
    img, _ = oracle.evaluate(fm_true.spec())
    rng = np.random.Generator(np.random.PCG64(seed))
    return workloads.noisy(img, rng)


# launch settings with each FP32 convolution form (VALU tiles, banded MFMA
# tiles, MFMA strips) for the forward convolution and the adjoint correlation
FORMS = [{"conv_tiles": t, "refine_occupancy": 100, "handover_div": 4, "claim_pixels": 64, "claim_cells": 16,
          "refine_passes": 5, "claim_spread": 0} for t in (0, 1, 2)]


def device_chi2_grad(ctx, fms, datas, sigmas, graph=False, tuning=None):
    b = Batch(ctx, [fm.spec() for fm in fms], graph=graph, grad=True)
    if tuning is not None:
        b.tuning = tuning
    b.set_data(datas, sigmas)
    b.evaluate_grad()
    chi2, grad = b.chi2_grad()
    out = []
    k = 0
    for fm in fms:
        nc = len(fm.kinds)
        out.append(fm.flat_grad(grad[k:k + nc]))
        k += nc
    b.close()
    return chi2, out


def check_close(adj, fd, prec, names, fd_err=None, fd_rel=None):
    floor = FLOOR[prec] * np.max(np.abs(fd))
    err = np.abs(adj - fd)
    lim = RTOL[prec] * np.abs(fd) + floor
    if fd_err is not None:
        lim = lim + 2 * fd_err
    if fd_rel is None:
        fd_rel = fd
    for n, a, f, f4, e, l in zip(names, adj, fd, fd_rel, err, lim):
        print("  %-10s adjoint % .10e  fd % .10e  rel %.2e  allowed %.2e  | fd(1e-4) rel %.2e"
              % (n, a, f, e / max(abs(f), 1e-300), l / max(abs(f), 1e-300), abs(f4 - a) / max(abs(a), 1e-300)),
              flush=True)
    bad = [n for n, e, l in zip(names, err, lim) if e > l]
    assert not bad, bad


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_small_fit_gradient_matches_frozen_fd(ctx64, ctx32, oracle, prec):
    announce("small fit grad %s" % prec)
    fm = small_fit()
    truth = FitModel(fm.size, [(k, dict(p)) for k, p in zip(fm.kinds, fm.base)], fm.psf)
    truth.base[0]["mag"] = 14.9
    truth.base[1]["re"] = 7.4
    data, sigma = make_data(oracle, truth)
    g_fd, fd_err, g_rel, c0 = oracle_fd_grad(oracle, fm, fm.theta(), data, sigma, with_error=True)
    ctx = ctx64 if prec == "fp64" else ctx32
    for tuning in ([None] if prec == "fp64" else FORMS):
        chi2, (g,) = device_chi2_grad(ctx, [fm], [data], [sigma], tuning=tuning)
        assert chi2[0] == pytest.approx(c0, rel=1e-12 if prec == "fp64" else 1e-5)
        check_close(g, g_fd, prec, fm.names(), fd_err, g_rel)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_bulge_disk256_gradient_matches_frozen_fd(ctx64, ctx32, oracle, prec):
    """256x256 bulge + disk with a 31x31 Moffat PSF, data = true model + 1 %
    noise, gradient over all 16 component parameters at a perturbed point."""
    announce("bulge_disk256 grad %s" % prec)
    fm, truth = bulge_disk256_model()
    data, sigma = make_data(oracle, truth)
    g_fd, fd_err, g_rel, c0 = oracle_fd_grad(oracle, fm, fm.theta(), data, sigma, with_error=True)
    ctx = ctx64 if prec == "fp64" else ctx32
    for tuning in ([None] if prec == "fp64" else FORMS):
        chi2, (g,) = device_chi2_grad(ctx, [fm], [data], [sigma], tuning=tuning)
        assert chi2[0] == pytest.approx(c0, rel=1e-12 if prec == "fp64" else 1e-5)
        check_close(g, g_fd, prec, fm.names(), fd_err, g_rel)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_masked_fit_gradient_matches_frozen_fd(ctx64, ctx32, oracle, prec):
    """A calculation mask (adjusted as Model::evaluate does: expanded by the
    PSF, applied again to the returned image): masked pixels carry no
    gradient, the chi^2 still counts them (model 0 there)."""
    announce("masked fit grad %s" % prec)
    fm = small_fit()
    mask = np.ones((fm.size, fm.size), np.uint8)
    mask[:, 24:] = 0          # a cut through the galaxy
    mask[5:12, 3:9] = 0       # and a hole in the outskirts
    fm.mask = mask
    truth = FitModel(fm.size, [(k, dict(p)) for k, p in zip(fm.kinds, fm.base)], fm.psf, mask=mask)
    truth.base[0]["mag"] = 14.9
    truth.base[1]["re"] = 7.4
    data, sigma = make_data(oracle, truth)
    g_fd, fd_err, g_rel, c0 = oracle_fd_grad(oracle, fm, fm.theta(), data, sigma, with_error=True)
    ctx = ctx64 if prec == "fp64" else ctx32
    for tuning in ([None] if prec == "fp64" else FORMS):
        chi2, (g,) = device_chi2_grad(ctx, [fm], [data], [sigma], tuning=tuning)
        assert chi2[0] == pytest.approx(c0, rel=1e-12 if prec == "fp64" else 1e-5)
        check_close(g, g_fd, prec, fm.names(), fd_err, g_rel)


@pytest.mark.parametrize("shape", [(8, 8), (10, 7)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_even_psf_gradient_matches_frozen_fd(ctx64, ctx32, oracle, prec, shape):
    """Even (and non-square) PSF sides: the adjoint correlation is shifted by
    one pixel on an even axis (the BRUTE convolution centres a kw-wide PSF at
    kw / 2)."""
    announce("even psf grad %s %s" % (prec, shape))
    fm = small_fit()
    w, h = shape
    yy, xx = np.mgrid[0:h, 0:w]
    psf = 1.0 / (1.0 + ((xx - (w - 1) / 2.0) ** 2 + (yy - (h - 1) / 2.0) ** 2) / 4.0) ** 2.5
    fm.psf = psf / psf.sum()
    truth = FitModel(fm.size, [(k, dict(p)) for k, p in zip(fm.kinds, fm.base)], fm.psf)
    truth.base[0]["xcen"] = 20.6
    truth.base[1]["re"] = 7.4
    data, sigma = make_data(oracle, truth)
    g_fd, fd_err, g_rel, c0 = oracle_fd_grad(oracle, fm, fm.theta(), data, sigma, with_error=True)
    ctx = ctx64 if prec == "fp64" else ctx32
    for tuning in ([None] if prec == "fp64" else FORMS):
        chi2, (g,) = device_chi2_grad(ctx, [fm], [data], [sigma], tuning=tuning)
        assert chi2[0] == pytest.approx(c0, rel=1e-12 if prec == "fp64" else 1e-5)
        check_close(g, g_fd, prec, fm.names(), fd_err, g_rel)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fft_convolved_fit_gradient(ctx64, ctx32, oracle, prec):
    """The FFT convolver's cropped result is the same-size linear convolution
    (up to the FFT's rounding, relative to the peak), so its adjoint is the
    direct correlation: against frozen-pattern differences of the oracle's
    FFT model."""
    announce("fft fit grad %s" % prec)
    fm = small_fit()
    fm.convolver = "fft"
    truth = FitModel(fm.size, [(k, dict(p)) for k, p in zip(fm.kinds, fm.base)], fm.psf, convolver="fft")
    truth.base[0]["mag"] = 14.9
    truth.base[1]["re"] = 7.4
    data, sigma = make_data(oracle, truth)
    g_fd, fd_err, g_rel, c0 = oracle_fd_grad(oracle, fm, fm.theta(), data, sigma, with_error=True)
    ctx = ctx64 if prec == "fp64" else ctx32
    for tuning in ([None] if prec == "fp64" else FORMS):
        chi2, (g,) = device_chi2_grad(ctx, [fm], [data], [sigma], tuning=tuning)
        assert chi2[0] == pytest.approx(c0, rel=1e-9 if prec == "fp64" else 1e-5)
        check_close(g, g_fd, prec, fm.names(), fd_err, g_rel)


def test_batch_of_fits(ctx32, oracle):
    """Several independent fits of different sizes and PSFs in one batch, in
    every convolution form (PSF heights 19 and 33: the strips' tail rows)."""
    announce("batch of fits")
    fms, datas, sigmas = [], [], []
    for i, (size, k) in enumerate([(48, 15), (40, 9), (56, 11), (52, 19), (66, 33)]):
        start, true = workloads.fit_batch(1, seed=100 + i, size=size, psf_size=k)
        img, _ = oracle.evaluate(true[0])
        d, s = workloads.noisy(img, np.random.Generator(np.random.PCG64(i)))
        comps = []
        for c in range(start[0].m.n_components):
            cp = start[0].m.components[c]
            if cp.kind == abi.PROFILE_KINDS["sky"]:
                comps.append(("sky", dict(bg=cp.p[0])))
            else:
                comps.append(("sersic", dict(xcen=cp.xcen, ycen=cp.ycen, mag=cp.mag, re=cp.p[0], nser=cp.p[1],
                                             ang=cp.ang, axrat=cp.axrat, box=cp.box)))
        fms.append(FitModel(size, comps, start[0]._psf, magzero=start[0].m.magzero))
        datas.append(d)
        sigmas.append(s)
    ref = [oracle_fd_grad(oracle, fm, fm.theta(), d, s, with_error=True) for fm, d, s in zip(fms, datas, sigmas)]
    for tuning in FORMS:
        chi2, gs = device_chi2_grad(ctx32, fms, datas, sigmas, tuning=tuning)
        for fm, c2, g, (g_fd, fd_err, g_rel, c0) in zip(fms, chi2, gs, ref):
            assert c2 == pytest.approx(c0, rel=1e-5), tuning["conv_tiles"]
            check_close(g, g_fd, "fp32", fm.names(), fd_err, g_rel)


def test_graph_replays_match_eager(ctx32, oracle):
    announce("grad graph replay")
    fm = small_fit()
    data, sigma = make_data(oracle, fm)
    specs = [fm.spec()]
    eager = Batch(ctx32, specs, graph=False, grad=True)
    eager.set_data([data], [sigma])
    eager.evaluate_grad()
    c_e, g_e = eager.chi2_grad()
    g = Batch(ctx32, specs, graph=True, grad=True)
    g.set_data([data], [sigma])
    for rep in range(3):
        g.evaluate_grad()
        c_g, g_g = g.chi2_grad()
        assert np.allclose(c_g, c_e, rtol=1e-13, atol=0), rep
        assert np.allclose(g_g, g_e, rtol=1e-11, atol=1e-11 * np.max(np.abs(g_e))), rep
        eager.evaluate_grad()
        eager.chi2_grad()
    g.close()
    eager.close()


def test_tuning_keeps_the_gradient(ctx32, oracle):
    """Launch settings of the adjoint: the refinement settings reorder only the
    FP64 atomic sums of the gradient; the matrix-core convolution forms change
    the FP32 rounding of the forward convolution and the PSF correlation."""
    announce("grad tuning")
    fm = small_fit()
    data, sigma = make_data(oracle, fm)
    b = Batch(ctx32, [fm.spec()], graph=True, grad=True)
    b.set_data([data], [sigma])
    b.evaluate_grad()
    c0, g0 = b.chi2_grad()
    scale = np.max(np.abs(g0), axis=1, keepdims=True)
    base = {"conv_tiles": 0, "refine_occupancy": 100, "handover_div": 4, "claim_pixels": 64, "claim_cells": 16,
            "refine_passes": 5, "claim_spread": 0}
    for setting, rtol in (({"handover_div": 2, "refine_occupancy": 50, "claim_pixels": 16, "refine_passes": 4,
                            "claim_spread": 16}, 1e-11),
                          ({"conv_tiles": 1}, 1e-5), ({"conv_tiles": 2}, 1e-5)):
        b.tuning = {**base, **setting}
        b.evaluate_grad()
        c1, g1 = b.chi2_grad()
        assert np.allclose(c1, c0, rtol=rtol, atol=0), setting
        assert np.all(np.abs(g1 - g0) <= rtol * scale), setting
    b.autotune()
    b.evaluate_grad()
    c2, g2 = b.chi2_grad()
    assert np.allclose(c2, c0, rtol=1e-5, atol=0)
    assert np.all(np.abs(g2 - g0) <= 1e-5 * scale)
    b.close()


def test_sky_gradient_closed_form(ctx32, oracle):
    """d chi^2 / d bg = sum_p -2 (data_p - model_p) / sigma^2 for an unconvolved sky."""
    announce("sky closed form")
    fm = small_fit()
    data, sigma = make_data(oracle, fm)
    data = data + 3 * sigma
    b = Batch(ctx32, [fm.spec()], graph=False, grad=True)
    b.set_data([data], [sigma])
    b.evaluate_grad()
    chi2, grad = b.chi2_grad()
    model = b.image(0)
    b.close()
    exp = float(np.sum(-2 * (data - model) / sigma ** 2))
    assert grad[2, 0] == pytest.approx(exp, rel=1e-5)
    assert chi2[0] == pytest.approx(chi2_of(model, data, sigma), rel=1e-6)


def test_fd_of_device_forward(ctx32, oracle):
    """Central differences of the device FP32 chi^2 itself.  Their precision
    is bounded by FP32 rounding of the images (eps ~ 6e-8) over the relative
    change a step makes in them: a relative step of 1e-4 in the sky level moves
    the pixels by ~1e-5 (FD noise ~0.6 %), so the step is 1e-3 and the
    tolerance 1 % (the FP64 checks above are the precise ones)."""
    announce("fd of device forward")
    fm = small_fit()
    truth = FitModel(fm.size, [(k, dict(p)) for k, p in zip(fm.kinds, fm.base)], fm.psf)
    truth.base[0]["xcen"] = 20.0
    data, sigma = make_data(oracle, truth)
    theta = fm.theta()
    chi2, (g,) = device_chi2_grad(ctx32, [fm], [data], [sigma])
    b = Batch(ctx32, [fm.spec()], graph=False, grad=True)
    b.set_data([data], [sigma])
    h = fd_steps(theta, fm.names(), rel=1e-3)
    names = fm.names()
    for k in [names.index(n) for n in ("0.xcen", "0.mag", "0.re", "1.nser", "1.axrat", "2.bg")]:
        vals = []
        for s in (+1, -1):
            t = theta.copy()
            t[k] += s * h[k]
            comps = [fm.spec(t).m.components[c] for c in range(len(fm.kinds))]
            b.set_components(comps)
            b.evaluate_grad()
            vals.append(b.chi2_grad()[0][0])
        fd = (vals[0] - vals[1]) / (2 * h[k])
        print("  %-8s adjoint % .8e fd % .8e" % (names[k], g[k], fd), flush=True)
        assert abs(fd - g[k]) <= 1e-2 * abs(g[k]), names[k]
    b.close()


def test_torch_autograd_function(ctx32, oracle):
    """The adjoint behind torch.autograd: theta.grad after chi2.backward()
    equals the frozen-pattern FD of the oracle."""
    announce("torch autograd")
    import torch
    from torch_chi2 import Chi2Function
    fm = small_fit()
    truth = FitModel(fm.size, [(k, dict(p)) for k, p in zip(fm.kinds, fm.base)], fm.psf)
    truth.base[1]["ang"] = 55.0
    data, sigma = make_data(oracle, truth)
    b = Batch(ctx32, [fm.spec()], graph=True, grad=True)
    b.set_data([data], [sigma])
    theta = torch.tensor(fm.theta(), dtype=torch.float64, requires_grad=True)
    chi2 = Chi2Function.apply(theta, b, fm)
    chi2.backward()
    g_fd, fd_err, g_rel, c0 = oracle_fd_grad(oracle, fm, fm.theta(), data, sigma, with_error=True)
    assert float(chi2) == pytest.approx(c0, rel=1e-5)
    check_close(theta.grad.numpy(), g_fd, "fp32", fm.names(), fd_err, g_rel)
    # a descent step -1e-4 |theta_k| sign(g_k) lowers chi^2 (the harness drives a fit)
    with torch.no_grad():
        scale = torch.where(theta == 0, torch.full_like(theta, 1e-3), theta.abs())
        step = -1e-4 * scale * torch.sign(theta.grad)
    chi2b = Chi2Function.apply((theta + step).detach(), b, fm)
    assert float(chi2b) < float(chi2), (float(chi2b), float(chi2))
    b.close()
