"""Batched, device-resident evaluation (pk_batch_*): every image of a
heterogeneous batch against the FP64 oracle, graph replays against eager
launches, and parameter updates between evaluations."""
import numpy as np
import pytest

from libprofit_amd import abi, workloads
from libprofit_amd.model import Batch
from parity_util import TOL_FP32, TOL_FP64, rel_errors

pytestmark = pytest.mark.gpu


def small_batch(n=20, seed=3):
    return workloads.galaxy_batch(n, seed, smin=24, smax=96)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_matches_oracle(ctx64, ctx32, oracle, prec):
    print("CASE batch", prec, flush=True)
    specs = small_batch()
    ctx = ctx64 if prec == "fp64" else ctx32
    b = Batch(ctx, specs, graph=False)
    b.evaluate()
    b.sync()
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    for i, s in enumerate(specs):
        ref, _ = oracle.evaluate(s)
        got = b.image(i)
        pix, flux = rel_errors(got, ref)
        assert pix <= tol, (i, pix)
        assert flux <= tol, (i, flux)
    b.close()


def test_graph_replays_match_eager(ctx32):
    print("CASE graph replay", flush=True)
    specs = small_batch(12, 5)
    eager = Batch(ctx32, specs, graph=False)
    eager.evaluate()
    eager.sync()
    ref = [eager.image(i) for i in range(len(specs))]
    g = Batch(ctx32, specs, graph=True)
    for rep in range(3):
        g.evaluate()
        g.sync()
        for i in range(len(specs)):
            assert np.array_equal(g.image(i), ref[i]), (rep, i)
        # eager work of the program between replays
        eager.evaluate()
        eager.sync()
    g.close()
    eager.close()


def test_set_components(ctx64, oracle):
    print("CASE set_components", flush=True)
    specs = small_batch(6, 9)
    b = Batch(ctx64, specs, graph=True)
    b.evaluate()
    b.sync()
    # move every component and brighten it, as one step of a fit would
    new_specs, comps = [], []
    for s in specs:
        cs = []
        for k in range(s.m.n_components):
            c = abi.Component.from_buffer_copy(s.m.components[k])
            if c.kind != abi.PROFILE_KINDS["sky"]:
                c.xcen += 0.37
                c.ycen -= 0.21
                c.mag -= 0.5
            cs.append(c)
        comps += cs
        new_specs.append(abi.ModelSpec(s.m.width, s.m.height, cs, psf=s._psf))
    b.set_components(comps)
    b.evaluate()
    b.sync()
    for i, s in enumerate(new_specs):
        ref, _ = oracle.evaluate(s)
        pix, _ = rel_errors(b.image(i), ref)
        assert pix <= TOL_FP64, (i, pix)
    b.close()
