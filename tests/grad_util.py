"""Helpers of the chi^2-gradient tests: parameter vectors of bulge+disk(+sky)
models and central finite differences of chi^2 computed with the FP64
oracle, with the refinement pattern of the central point held fixed."""
import numpy as np

from libprofit_amd import abi

SLOTS = abi.GRAD_SLOTS


class FitModel:
    """A model whose components are given as (kind, params) with the
    parameters the gradient covers first (abi.GRAD_PARAMS order)."""

    def __init__(self, size, comps, psf, convolver="brute", magzero=0.0, mask=None):
        self.size = size
        self.mask = mask
        self.magzero = magzero
        self.kinds = [k for k, _ in comps]
        self.base = [dict(p) for _, p in comps]
        self.psf = psf
        self.convolver = convolver

    def theta(self):
        return np.array([self.base[i][n] for i, k in enumerate(self.kinds) for n in abi.GRAD_PARAMS[k]])

    def names(self):
        return ["%d.%s" % (i, n) for i, k in enumerate(self.kinds) for n in abi.GRAD_PARAMS[k]]

    def spec(self, theta=None):
        params = [dict(p) for p in self.base]
        if theta is not None:
            t = 0
            for i, k in enumerate(self.kinds):
                for n in abi.GRAD_PARAMS[k]:
                    params[i][n] = float(theta[t])
                    t += 1
        comps = []
        for k, p in zip(self.kinds, params):
            extra = {"convolve": True} if k == "sersic" and self.psf is not None else {}
            comps.append(abi.make_component(k, **extra, **p))
        return abi.ModelSpec(self.size, self.size, comps, psf=self.psf, convolver=self.convolver,
                             magzero=self.magzero, mask=self.mask)

    def flat_grad(self, grad):
        """Gradient slots (n_components x SLOTS) -> vector in theta order."""
        out = []
        for i, k in enumerate(self.kinds):
            out.extend(grad[i, : len(abi.GRAD_PARAMS[k])])
        return np.array(out)


def chi2_of(image, data, sigma):
    return float(np.sum(((data - image) / sigma) ** 2))


def fd_steps(theta, names, rel=1e-4, absolute=1e-4):
    """Central-difference steps: rel * |theta|, or `absolute` where theta is 0."""
    h = rel * np.abs(theta)
    h[h == 0] = absolute
    return h


def oracle_fd_grad(oracle, fm, theta, data, sigma, rel=1e-4, with_error=False):
    """d chi^2 / d theta by central differences of the FP64 oracle with the
    refinement pattern of `theta` held fixed, relative step `rel`.

    With with_error: returns (g, err, g_rel, chi2): g the differences at a
    relative step chosen per parameter (below), err = |g(h) - g(2h)| a bound
    of their truncation and rounding error (the larger difference to the
    steps 2h and h/2), g_rel those at `rel`.  The image
    of a Sersic profile with nser > 1 sampled on a fixed refinement pattern
    has a cusp that moves across sub-cells down to ~1e-9 px when its centre
    moves, so for those positions the differences converge to the
    derivative only at steps ~1e-9 relative (at 1e-4 they can be off by
    ~1e-3); every other parameter is converged at 1e-6."""
    oracle.pattern("record")
    img0, _ = oracle.evaluate(fm.spec(theta))
    c0 = chi2_of(img0, data, sigma)

    def central(k, step):
        vals = []
        for s in (+1, -1):
            t = theta.copy()
            t[k] += s * step
            oracle.pattern("replay")
            img, _ = oracle.evaluate(fm.spec(t))
            assert oracle.pattern_ok(), "replayed pattern did not match"
            vals.append(chi2_of(img, data, sigma))
        return (vals[0] - vals[1]) / (2 * step)

    names = fm.names()
    g_rel = np.zeros_like(theta)
    g = np.zeros_like(theta)
    err = np.zeros_like(theta)
    try:
        for k, name in enumerate(names):
            g_rel[k] = central(k, fd_steps(theta, None, rel)[k])
            if not with_error:
                continue
            ci, pname = name.split(".")
            cusp = fm.kinds[int(ci)] == "sersic" and pname in ("xcen", "ycen") and \
                fm.base[int(ci)]["nser"] > 1 if "nser" in fm.base[int(ci)] else False
            r = 1e-9 if cusp else 1e-6
            # relative step, but at least r in absolute terms: below that the
            # chi^2 differences of a small parameter (box ~ 0.05) are a few
            # ulps and can come out exactly proportional to the step
            h = r * max(abs(theta[k]), 1.0)
            g[k] = central(k, h)
            err[k] = max(abs(g[k] - central(k, 2 * h)), abs(g[k] - central(k, 0.5 * h)))
    finally:
        oracle.pattern("off")
    if with_error:
        return g, err, g_rel, c0
    return g_rel, c0


def bulge_disk256_model(size=256, sky=None):
    """The 256x256 differentiation workload at a perturbed starting point:
    (FitModel at the start, true parameter vector)."""
    from libprofit_amd import workloads
    comps, psf = workloads.bulge_disk_fit(size)
    true = []
    for c in comps:
        true.append({"xcen": c.xcen, "ycen": c.ycen, "mag": c.mag, "re": c.p[0], "nser": c.p[1], "ang": c.ang,
                     "axrat": c.axrat, "box": c.box})
    comps_t = [("sersic", p) for p in true]
    if sky is not None:
        comps_t.append(("sky", {"bg": sky}))
    fm_true = FitModel(size, comps_t, psf)
    start = [dict(p) for p in true]
    start[0].update(xcen=start[0]["xcen"] + 0.21, ycen=start[0]["ycen"] - 0.17, mag=17.1, re=8.5, nser=3.7,
                    ang=43.0, axrat=0.77, box=0.05)
    start[1].update(xcen=start[1]["xcen"] - 0.13, ycen=start[1]["ycen"] + 0.11, mag=15.95, re=28.0, nser=0.93,
                    ang=38.0, axrat=0.42, box=-0.04)
    comps_s = [("sersic", p) for p in start]
    if sky is not None:
        comps_s.append(("sky", {"bg": sky * 1.05}))
    return FitModel(size, comps_s, psf), fm_true
