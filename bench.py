"""Benchmark: batched FP32 galaxy-image generation on one MI355X.

Workload (one step): 4096 galaxies, square stamps with side ~U{32..512}, each
with 1-3 components drawn from {sersic, coresersic, moffat, ferrer, king,
brokenexp} (magnitudes U[14,20] against magzero 30) plus a sky level
U[0,10], convolved with a per-galaxy Moffat PSF of odd side ~U{15..41}; one
batched evaluation (FP64 per-component set-up on the device, FP32 profile
evaluation with FP64 refinement decisions near the threshold, FP32 direct
convolution), replayed as a hipGraph.  Inputs are resident in device memory
before the timed region.

Also reported on the same run:
  - the plain FP64 path (one Model::evaluate per galaxy) on a sample of the
    same galaxies, in images/s, and the speed-up of the batched FP32 path;
  - accuracy against the FP64 CPU oracle on a sample spanning every PSF side
    and the stamp sizes: per-pixel (above 1e-6 of the peak) and total-flux
    relative error of the batched FP32 and the plain FP64 images; the run
    fails (exit 3) when the FP32 images miss 1e-5 or a galaxy of the sample
    does not dominate some pixel of its stamp;
  - roofline of the dominant kernel (the convolution) and of the pixel pass,
    the refinement and the refinement adjoint, from HIP events on the launch
    stream and the kernels' own evaluation counters;
  - the CPU baseline: the FP64 oracle built -O3 -march=native on the host,
    OpenMP, on a bounded sample;
  - the 100x100 Sersic and the 400x400 Sersic + 25x25 PSF models (BRUTE and
    FFT): plain FP64 GPU, batched FP32 GPU and CPU images/s with their errors;
  - chi^2-gradient throughput of 1024 bulge+disk+sky fits (128x128, PSF 31)
    and of the 256x256 bulge+disk fit.

Prints ONE JSON line.  Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "images/s & chi2-grads/s: batched FP32 vs plain FP64 GPU at <=1e-5 rel flux err"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# FP32 VALU issue: 256 CUs x 64 lanes x 2.4 GHz lane-slots per second
VALU_PEAK_TSLOTS = 256 * 64 * 2.4e9 / 1e12
# algorithmic VALU slots: one FP32 evaluate_at (Sersic general form: 14 FP32
# operations + 4 transcendentals at quarter rate), the pixel pass's
# classification of one (pixel, radial component) pair (coordinate transform
# and radius test), and one evaluate_at of the adjoint (value and partials)
SLOTS_EVAL = 30
SLOTS_CLASSIFY = 8
SLOTS_EVAL_ADJ = 90
TOL_FP32 = 1e-5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_errors(got, ref, floor=1e-6):
    peak = np.max(np.abs(ref))
    sel = np.abs(ref) > floor * peak
    pix = float(np.max(np.abs(got[sel] - ref[sel]) / np.abs(ref[sel])))
    flux = float(abs(got.sum() - ref.sum()) / abs(ref.sum()))
    return pix, flux


def conv_flops(specs):
    """Algorithmic flops of the direct convolution of one batch evaluation:
    2 * kw * kh per returned pixel (the crop of the PSF padding keeps only
    those; the padding pixels are sources, not outputs)."""
    tot = 0
    for s in specs:
        m = s.m
        tot += 2 * m.psf_width * m.psf_height * m.width * m.height
    return tot


def conv_traffic(kernel):
    """HBM bytes per step (all launches of one evaluation) of the convolution
    kernel on this workload, as measured with rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE and recorded in profiles/<kernel>_galaxy4096_traffic.txt (null
    if absent)."""
    path = os.path.join(ROOT, "profiles", "%s_galaxy4096_traffic.txt" % kernel)
    try:
        for line in open(path):
            if line.startswith("hbm_bytes_per_step:"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def accuracy_sample(specs, n, seed=12345):
    """Indices of a sample of the batch for the plain FP64 timing and the
    accuracy checks: for every PSF side its smallest and its largest stamp,
    the rest drawn at random (seeded)."""
    by_k = {}
    for i, s in enumerate(specs):
        by_k.setdefault(int(s.m.psf_width), []).append(i)
    pick = set()
    for _k, idx in sorted(by_k.items()):
        idx = sorted(idx, key=lambda i: (specs[i].m.width, i))
        pick.update((idx[0], idx[-1]))
    rng = np.random.Generator(np.random.PCG64(seed))
    for i in rng.permutation(len(specs)).tolist():
        if len(pick) >= n:
            break
        pick.add(i)
    return sorted(pick)


def host_cpu():
    """What the CPU baseline ran on: logical CPUs, those this process may
    use, and the CPU model name."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def galaxy_share(O, spec, ref, L=None):
    """Largest share of a compared pixel that the non-sky profiles make up."""
    from libprofit_amd import abi
    comps = [abi.Component.from_buffer_copy(spec.m.components[k]) for k in range(spec.m.n_components)]
    comps = [c for c in comps if c.kind != abi.PROFILE_KINDS["sky"]]
    gal, _ = O.evaluate(abi.ModelSpec(spec.m.width, spec.m.height, comps, psf=spec._psf, magzero=spec.m.magzero,
                                      convolver=spec.m.convolver), L)
    sel = np.abs(ref) > 1e-6 * np.max(np.abs(ref))
    return float(np.max(np.abs(gal[sel]) / np.abs(ref[sel])))


def model_rows(args, Batch, evaluate_spec, ctx32, ctx64, O, threads):
    """The 100x100 Sersic (profit-cli defaults) and the 400x400 Sersic + 25x25
    Moffat PSF models (64 seeds, BRUTE and FFT): images/s of the plain FP64
    GPU path (one pk_evaluate per image, host buffers), of the batched FP32
    path (one batch of all seeds, graph) and of the CPU oracle, with the
    errors of both GPU paths against the oracle."""
    from libprofit_amd import abi, workloads
    rows = {}

    def timed(fn, reps):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps

    def errs(got, refs):
        e = [rel_errors(g, r) for g, r in zip(got, refs)]
        return {"max_pixel_rel_err": max(x[0] for x in e), "max_flux_rel_err": max(x[1] for x in e)}

    cases = [("sersic100", [abi.ModelSpec(100, 100, [abi.make_component(
        "sersic", xcen=50.3, ycen=49.7, mag=15, re=10, nser=4, ang=30, axrat=0.6, box=0)])], 200)]
    for conv in ("brute", "fft"):
        specs = []
        for seed in range(args.c2_seeds):
            sp = workloads.single_sersic_psf(seed)
            sp.m.convolver = abi.CONVOLVERS[conv]
            specs.append(sp)
        cases.append(("sersic400_psf25_" + conv, specs, 1))
    Lf = O.fast_lib() if O is not None else None
    for name, specs, reps in cases:
        row = {"images": len(specs)}
        # plain FP64 GPU: one pk_evaluate per image
        plain = []
        dt = timed(lambda: plain.__setitem__(slice(None), [evaluate_spec(ctx64, s)[0] for s in specs]), reps)
        row["plain_fp64_gpu_images_per_s"] = round(len(specs) / dt, 2)
        b = Batch(ctx32, specs, graph=True)
        b.evaluate()
        b.sync()
        dtb = timed(lambda: (b.evaluate(), b.sync()), max(reps, 20))
        row["batched_fp32_gpu_images_per_s"] = round(len(specs) / dtb, 2)
        got32 = [b.image(i) for i in range(len(specs))]
        b.close()
        if O is not None:
            refs = [O.evaluate(s)[0] for s in specs]
            row["plain_fp64_vs_oracle"] = errs(plain, refs)
            row["batched_fp32_vs_oracle"] = errs(got32, refs)
            ncpu = min(len(specs), 8)
            dtc = timed(lambda: [O.evaluate(s, Lf) for s in specs[:ncpu]], 1 if reps == 1 else 20)
            row["cpu_images_per_s"] = round(ncpu / dtc, 2)
            row["cpu_threads"] = threads
        log("%s: %s" % (name, row))
        rows[name] = row
    return rows


def fd_plain_gradient(ctx64, spec, data, sigma, evaluate_spec):
    """chi^2 gradient of one fit by central differences of the plain FP64 path
    (one pk_evaluate per point: 2 per parameter + 1), the way a caller of
    libprofit's Model::evaluate gets it."""
    from libprofit_amd import abi
    comps = [abi.Component.from_buffer_copy(spec.m.components[c]) for c in range(spec.m.n_components)]
    img, _, _ = evaluate_spec(ctx64, spec)
    chi0 = float(np.sum(((data - img) / sigma) ** 2))
    g = []
    for ci, c in enumerate(comps):
        names = abi.GRAD_PARAMS["sersic" if c.kind == abi.PROFILE_KINDS["sersic"] else "sky"]
        for n in names:
            vals = []
            for sgn in (1, -1):
                cc = [abi.Component.from_buffer_copy(x) for x in comps]
                if n in abi.PROFILE_PARAMS["sersic"] and c.kind == abi.PROFILE_KINDS["sersic"]:
                    idx = abi.PROFILE_PARAMS["sersic"].index(n)
                    h = 1e-4 * abs(cc[ci].p[idx]) or 1e-4
                    cc[ci].p[idx] += sgn * h
                elif n == "bg":
                    h = 1e-4 * abs(cc[ci].p[0]) or 1e-4
                    cc[ci].p[0] += sgn * h
                else:
                    v = getattr(cc[ci], n)
                    h = 1e-4 * abs(v) or 1e-4
                    setattr(cc[ci], n, v + sgn * h)
                sp = abi.ModelSpec(spec.m.width, spec.m.height, cc, psf=spec._psf, magzero=spec.m.magzero)
                im, _, _ = evaluate_spec(ctx64, sp)
                vals.append(float(np.sum(((data - im) / sigma) ** 2)))
            g.append((vals[0] - vals[1]) / (2 * h))
    return chi0, np.array(g)


def grad_bench(args, Batch, Context, evaluate_spec, ctx32, ctx64):
    """chi^2-gradient throughput: 1024 bulge+disk+sky fits of 128x128 with a
    31x31 PSF (one FP32 forward + adjoint of all fits per step, graph) and
    the 256x256 bulge+disk fit with a 31x31 PSF, beside the plain FP64 path
    (finite differences of pk_evaluate) on a sample of the fits."""
    from libprofit_amd import workloads
    out = {}
    start, true = workloads.fit_batch(args.fits, seed=args.seed)
    # data: the true models (FP64 batch) + 1 % noise
    tb = Batch(ctx64, true, graph=False)
    tb.evaluate()
    tb.sync()
    rng = np.random.Generator(np.random.PCG64(args.seed))
    datas, sigmas = [], []
    for i in range(len(true)):
        d, sg = workloads.noisy(tb.image(i), rng)
        datas.append(d)
        sigmas.append(sg)
    tb.close()
    fixed = json.loads(args.fits_tuning) if args.fits_tuning else None
    gb = Batch(ctx32, start, graph=True, grad=True, autotune=not (args.no_tune or fixed))
    if fixed:
        gb.tuning = fixed
    gb.set_data(datas, sigmas)
    for _ in range(args.warmup):
        gb.evaluate_grad()
    gb.chi2_grad()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gb.evaluate_grad()
    chi2_32, grad_32 = gb.chi2_grad()
    dt = (time.perf_counter() - t0) / args.steps
    out["fits_chi2_grads_per_s"] = round(len(start) / dt, 2)
    out["fits_ms_per_step"] = round(dt * 1e3, 4)
    log("fits: %d x (128^2 bulge+disk+sky, PSF 31) forward+gradient in %.3f ms -> %.1f chi2-grads/s"
        % (len(start), dt * 1e3, len(start) / dt))
    # stage split (eager launches, HIP events)
    eb = Batch(ctx32, start, graph=False, grad=True)
    eb.tuning = gb.tuning
    out["fits_tuning"] = gb.tuning.as_dict()
    eb.set_data(datas, sigmas)
    eb.evaluate_grad()
    eb.sync()
    eb.evaluate_grad()
    inf = eb.sync()
    out["fits_stage_ms"] = {"forward": round(inf.ms_total, 4), "adjoint": round(inf.ms_grad, 4),
                            "adjoint_refine": round(inf.ms_grad_refine, 4)}
    slots = SLOTS_EVAL_ADJ * inf.n_grad_subcells
    ach = slots / (inf.ms_grad_refine * 1e-3) / 1e12 if inf.ms_grad_refine > 0 else 0.0
    out["fits_adjoint_refine_roofline"] = {
        "kernel": "refine32_kernel<..., GRAD> (%d launches) + r32_resolve_kernel<GRAD> + grad_finalize_kernel, "
                  "per step" % out["fits_tuning"]["refine_passes"],
        "bound": "valu", "achieved": round(ach, 3), "peak": round(VALU_PEAK_TSLOTS, 2), "unit": "Tslot/s",
        "frac": round(ach / VALU_PEAK_TSLOTS, 4), "evaluations": int(inf.n_grad_subcells),
        "slots_per_evaluation": SLOTS_EVAL_ADJ, "ms": round(inf.ms_grad_refine, 4)}
    eb.close()
    # FP64 adjoint of the same fits: accuracy of the FP32 gradients
    b64 = Batch(ctx64, start, graph=False, grad=True)
    b64.set_data(datas, sigmas)
    b64.evaluate_grad()
    chi2_64, grad_64 = b64.chi2_grad()
    b64.close()
    # per fit: largest FP32 error over the largest FP64 gradient component,
    # and the error norm over the gradient norm (the full-size test bounds
    # both by 2e-4); per-element relative errors are large only for
    # components near zero (terms that cancel near their optimum)
    n = len(start)
    g32, g64 = grad_32.reshape(n, -1), grad_64.reshape(n, -1)
    top = np.abs(g64).max(axis=1)
    elem = np.abs(g32 - g64).max(axis=1) / top
    norm = np.linalg.norm(g32 - g64, axis=1) / np.linalg.norm(g64, axis=1)
    nz = np.abs(grad_64) > 0
    rel = np.abs(grad_32[nz] - grad_64[nz]) / np.abs(grad_64[nz])
    out["fits_fp32_vs_fp64_grad"] = {
        "max_error_over_largest_component": {"median": float(np.median(elem)), "max": float(elem.max())},
        "error_norm_over_gradient_norm": {"median": float(np.median(norm)), "max": float(norm.max())},
        "per_element_rel_err": {"median": float(np.median(rel)), "p99": float(np.percentile(rel, 99)),
                                "max": float(np.max(rel))},
        "bound": 2e-4}
    out["fits_fp32_vs_fp64_chi2_max_rel_err"] = float(np.max(np.abs(chi2_32 - chi2_64) / np.abs(chi2_64)))
    # plain FP64 path: central differences of pk_evaluate on a sample of fits
    ns = min(args.grad_sample, len(start))
    t0 = time.perf_counter()
    for i in range(ns):
        fd_plain_gradient(ctx64, start[i], datas[i], sigmas[i], evaluate_spec)
    dtp = (time.perf_counter() - t0) / ns
    out["fits_plain_fp64_fd_chi2_grads_per_s"] = round(1.0 / dtp, 3)
    out["fits_speedup_vs_plain_fp64"] = round(len(start) / dt * dtp, 2)
    log("fits: plain FP64 FD gradient %.2f ms/fit -> speed-up %.1fx" % (dtp * 1e3, len(start) / dt * dtp))
    gb.close()

    # the single 256x256 bulge + disk fit
    comps, psf = workloads.bulge_disk_fit(256)
    tspec = workloads.abi.ModelSpec(256, 256, comps, psf=psf)
    img, _, _ = evaluate_spec(ctx64, tspec)
    d, sg = workloads.noisy(img, np.random.Generator(np.random.PCG64(0)))
    scomps = []
    r2 = np.random.Generator(np.random.PCG64(1))
    for c in comps:
        p = workloads.perturbed(dict(xcen=c.xcen, ycen=c.ycen, mag=c.mag, re=c.p[0], nser=c.p[1], ang=c.ang,
                                     axrat=c.axrat, box=c.box), r2)
        scomps.append(workloads.abi.make_component("sersic", convolve=True, **p))
    sspec = workloads.abi.ModelSpec(256, 256, scomps, psf=psf)
    one = Batch(ctx32, [sspec], graph=True, grad=True, autotune=not args.no_tune)
    one.set_data([d], [sg])
    for _ in range(args.warmup):
        one.evaluate_grad()
    one.chi2_grad()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one.evaluate_grad()
    one.chi2_grad()
    dt1 = (time.perf_counter() - t0) / args.steps
    one.close()
    t0 = time.perf_counter()
    fd_plain_gradient(ctx64, sspec, d, sg, evaluate_spec)
    dtp1 = time.perf_counter() - t0
    out["c256_chi2_grads_per_s"] = round(1.0 / dt1, 2)
    out["c256_plain_fp64_fd_chi2_grads_per_s"] = round(1.0 / dtp1, 3)
    out["c256_speedup_vs_plain_fp64"] = round(dtp1 / dt1, 2)
    log("256^2 fit: %.3f ms per forward+gradient, plain FP64 FD %.1f ms -> %.1fx" % (dt1 * 1e3, dtp1 * 1e3,
                                                                                  dtp1 / dt1))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--galaxies", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sample", type=int, default=96,
                    help="galaxies timed on the plain FP64 path and checked against the oracle")
    ap.add_argument("--c2-seeds", type=int, default=64, help="seeds of the 400x400 Sersic + PSF models")
    ap.add_argument("--no-models", action="store_true", help="skip the 100x100 and 400x400 model rows")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fits", type=int, default=1024, help="fits of the chi^2-gradient workload")
    ap.add_argument("--grad-sample", type=int, default=8, help="fits timed on the plain FP64 FD gradient")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch settings instead of the autotuner's")
    ap.add_argument("--tuning", default=None, metavar="JSON",
                    help="fixed launch settings (pk_tuning fields) instead of the autotuner's")
    ap.add_argument("--fits-tuning", default=None, metavar="JSON",
                    help="fixed launch settings of the chi^2-gradient batch instead of the autotuner's")
    ap.add_argument("--fits-only", action="store_true",
                    help="only run the chi^2-gradient workload (prints its figures as one JSON line)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 3 times (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 3 copies of %d bytes" % (4 * n))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the batched path shards galaxies across GPUs with no exchange, "
              "and only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)

    from libprofit_amd import workloads
    from libprofit_amd.model import Batch, Context, evaluate_spec

    if args.fits_only:
        ctx32 = Context(0, "fp32")
        ctx64 = Context(0, "fp64")
        print(json.dumps(grad_bench(args, Batch, Context, evaluate_spec, ctx32, ctx64)), flush=True)
        return

    t0 = time.time()
    specs = workloads.galaxy_batch(args.galaxies, args.seed)
    npix = workloads.pixel_count(specs)
    log("workload: %d galaxies, %.3g output pixels (%.1fs to draw inputs)" % (len(specs), npix, time.time() - t0))

    ctx32 = Context(0, "fp32")
    ctx64 = Context(0, "fp64")

    # ---- batched FP32 path (graph) --------------------------------------------------
    t0 = time.time()
    fixed = json.loads(args.tuning) if args.tuning else None
    batch = Batch(ctx32, specs, graph=True, autotune=not (args.no_tune or fixed))
    if fixed:
        batch.tuning = fixed
    log("batch created in %.2fs" % (time.time() - t0))
    t0 = time.time()
    batch.evaluate()  # the autotuner runs here, before the graph is captured
    batch.sync()
    tuning = batch.tuning.as_dict()
    log("first evaluation (autotuning included) %.2fs; settings %s" % (time.time() - t0, tuning))
    for _ in range(args.warmup):
        batch.evaluate()
    info = batch.sync()
    import ctypes
    L = batch.L
    # timed region: device-resident inputs, K graph replays
    L.pk_batch_sync(batch.handle, None)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.evaluate()
    info = batch.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    images_per_s = len(specs) / ((t1 - t0) / args.steps)
    log("batched fp32: %.3f ms/step, %.1f images/s, %d refine entries, %d subcells" %
        (ms_per_step, images_per_s, info.n_refine_entries, info.n_subcells))

    # ---- per-kernel timing (same kernels, launched without the graph so that HIP
    # events sit between them on the launch stream) ---------------------------------
    ev_batch = Batch(ctx32, specs, graph=False)
    ev_batch.tuning = batch.tuning
    stages = {"setup": 0.0, "coarse": 0.0, "refine": 0.0, "conv": 0.0, "compose": 0.0, "total": 0.0}
    nrep = max(3, min(args.steps, 10))
    for _ in range(2):
        ev_batch.evaluate()
    ev_batch.sync()
    for _ in range(nrep):
        ev_batch.evaluate()
        inf = ev_batch.sync()
        for k in stages:
            stages[k] += getattr(inf, "ms_" + k) / nrep
    ev_batch.close()
    log("stages (ms): " + ", ".join("%s %.3f" % kv for kv in stages.items()))
    flops = conv_flops(specs)
    conv_ms = stages["conv"]
    achieved_tflops = flops / (conv_ms * 1e-3) / 1e12
    mfma = inf.conv_tiles in (1, 2)
    kernel, form = {1: ("conv_mfma_kernel", "FP32 MFMA tiles"),
                    2: ("conv_strip_kernel", "FP32 MFMA strips")}.get(inf.conv_tiles, ("conv_kernel", "FP32 FMA"))
    roofline = {
        "kernel": "%s (direct PSF convolution, %s, composition fused in; %d launches per step, one per "
                  "PSF-size class; figures per step)" % (kernel, form, inf.n_conv_launches),
        # 2*kw*kh flop per returned pixel over ~12 B per pixel: far right of the
        # HBM ridge, bound by the matrix cores (MFMA forms) or FP32 FMA issue
        "bound": "mfma" if mfma else "valu",
        "achieved": round(achieved_tflops, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4),
        "traffic": conv_traffic(kernel),
        "algorithmic_flops_per_launch": flops,
        "avg_launch_ms": round(conv_ms, 4),
        "launches_per_step": int(inf.n_conv_launches),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
    }
    # the profile-evaluation kernels, priced in FP32 VALU issue slots from
    # the evaluation counts of the same workload (pk_batch_work_counts)
    wc = batch.work_counts()
    pairs = int(wc.pixel_component_pairs)

    def valu_entry(kernel_name, slots, ms, **extra):
        ach = slots / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        e = {"kernel": kernel_name, "bound": "valu", "achieved": round(ach, 3), "peak": round(VALU_PEAK_TSLOTS, 2),
             "unit": "Tslot/s", "frac": round(ach / VALU_PEAK_TSLOTS, 4), "ms": round(ms, 4)}
        e.update(extra)
        return e

    kernels = {
        "conv": {k: roofline[k] for k in ("bound", "achieved", "peak", "unit", "frac", "avg_launch_ms")},
        "coarse": valu_entry("coarse32_kernel<float> (one launch per step)",
                             SLOTS_CLASSIFY * pairs + SLOTS_EVAL * wc.coarse_evaluations, stages["coarse"],
                             pixel_component_pairs=pairs, evaluations=int(wc.coarse_evaluations),
                             slots_per_pair=SLOTS_CLASSIFY, slots_per_evaluation=SLOTS_EVAL),
        "refine": valu_entry("refine32_kernel (%d launches) + r32_resolve_kernel + finalize32_kernel, per step"
                             % tuning["refine_passes"],
                             SLOTS_EVAL * inf.n_subcells, stages["refine"], evaluations=int(inf.n_subcells),
                             slots_per_evaluation=SLOTS_EVAL),
    }
    log("kernels: %s" % json.dumps(kernels))

    # ---- plain FP64 path on a sample of the same galaxies ---------------------------
    sample = accuracy_sample(specs, min(args.sample, len(specs)), seed=12345)
    evaluate_spec(ctx64, specs[sample[0]])  # warm
    t0 = time.perf_counter()
    ref64 = {}
    for i in sample:
        img, _, _ = evaluate_spec(ctx64, specs[i])
        ref64[i] = img
    t1 = time.perf_counter()
    sample_pix = sum(int(specs[i].m.width) * int(specs[i].m.height) for i in sample)
    plain_s_per_img = (t1 - t0) / len(sample)
    # normalise to the batch's mean image size (the plain path's cost scales with pixels)
    plain_images_per_s = 1.0 / plain_s_per_img * (sample_pix / len(sample)) / (npix / len(specs))
    log("plain fp64: %.3f ms/image on the sample -> %.1f images/s at the batch mean size" %
        (plain_s_per_img * 1e3, plain_images_per_s))

    # ---- CPU baseline and accuracy: the FP64 oracle ---------------------------------
    # The oracle (the checker) draws the sample's reference images; its
    # -O3 -march=native build on this host is timed on galaxies of the same batch
    # in a seeded order until --cpu-seconds of work.
    cpu, accuracy, model_row = None, None, {}
    accuracy_ok = True
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        host = host_cpu()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or host["usable"]
        O.set_threads(threads)
        Lf = O.fast_lib()
        Lf.or_set_threads(threads)
        worst = {"fp32": [0.0, 0.0], "fp64": [0.0, 0.0]}
        min_share, worst_at = 1.0, None
        for i in sample:
            ref, _ = O.evaluate(specs[i])
            p32, f32 = rel_errors(batch.image(i), ref)
            p64, f64 = rel_errors(ref64[i], ref)
            if p32 > worst["fp32"][0]:
                worst_at = i
            worst["fp32"] = [max(worst["fp32"][0], p32), max(worst["fp32"][1], f32)]
            worst["fp64"] = [max(worst["fp64"][0], p64), max(worst["fp64"][1], f64)]
            min_share = min(min_share, galaxy_share(O, specs[i], ref))
        accuracy = {
            "reference": "FP64 CPU oracle (QUADPACK normalisation as libprofit's)",
            "sample": "%d galaxies: every PSF side's smallest and largest stamp + seeded random" % len(sample),
            "batched_fp32_vs_oracle": {"max_pixel_rel_err": worst["fp32"][0], "max_flux_rel_err": worst["fp32"][1],
                                       "worst_galaxy": worst_at},
            "plain_fp64_vs_oracle": {"max_pixel_rel_err": worst["fp64"][0], "max_flux_rel_err": worst["fp64"][1]},
            "min_galaxy_share_of_a_pixel": min_share,
            "pixels_compared": "above 1e-6 of each image's peak",
        }
        accuracy_ok = worst["fp32"][0] <= TOL_FP32 and worst["fp32"][1] <= TOL_FP32 and min_share >= 1e3 * TOL_FP32
        log("accuracy: %s" % json.dumps(accuracy))
        order = np.random.Generator(np.random.PCG64(777)).permutation(len(specs)).tolist()
        done, tpix, t0 = 0, 0, time.perf_counter()
        for i in order:
            O.evaluate(specs[i], Lf)
            done += 1
            tpix += int(specs[i].m.width) * int(specs[i].m.height)
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu_ips = done / dt * (tpix / done) / (npix / len(specs))
        cpu = {"value": round(cpu_ips, 3), "unit": "images/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d galaxies in a seeded order, %.1f s: FP64 oracle built -O3 -march=native "
                         "on this host, OpenMP over pixels, BRUTE with the AVX dot-product order; images/s "
                         "normalised to the batch mean image size" % (done, len(specs), dt),
               "host": host}
        log("cpu baseline: %s" % cpu)
        if not args.no_models:
            model_row = model_rows(args, Batch, evaluate_spec, ctx32, ctx64, O, threads)
    elif not args.no_models:
        model_row = model_rows(args, Batch, evaluate_spec, ctx32, ctx64, None, 0)

    grad = {}
    if not args.no_grad:
        grad = grad_bench(args, Batch, Context, evaluate_spec, ctx32, ctx64)

    out = {
        "metric": METRIC,
        "value": round(images_per_s, 2),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "galaxy_batch: %d galaxies, stamps U{32..512}, 1-3 radial components (mag U[14,20], "
                        "magzero 30) + sky U[0,10], per-galaxy Moffat PSF U{15..41}, BRUTE, FP32 bulk / FP64 "
                        "set-up" % len(specs),
            "galaxies": len(specs),
            "output_pixels": npix,
            "seed": args.seed,
            "graph": True,
            "tuning": tuning,
        },
        "plain_fp64_images_per_s": round(plain_images_per_s, 2),
        "speedup_vs_plain_fp64": round(images_per_s / plain_images_per_s, 2),
        "accuracy": accuracy,
        "refine_entries": int(info.n_refine_entries),
        "refine_evaluations": int(info.n_subcells),
        "coarse_evaluations": int(wc.coarse_evaluations),
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "models": model_row,
    }
    out.update(grad)
    batch.close()
    print(json.dumps(out), flush=True)
    if not accuracy_ok:
        log("FAILED: batched FP32 images beyond %g of the oracle, or a galaxy not visible in its stamp" % TOL_FP32)
        sys.exit(3)


if __name__ == "__main__":
    main()
