"""Shared helpers for GPU-vs-oracle parity tests (test infrastructure)."""
import numpy as np

from oracle.oracle import Oracle
from snac_amd import Context

FIELDS = ("coord", "velocity", "stress", "plastic_strain")


def run_both(problem, nsteps, threads=4, fields=FIELDS, variant=None):
    o = Oracle(problem, threads=threads)
    if problem.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    o.step(nsteps)
    ref = o.state()
    rsc = o.scalars()
    o.close()
    ctx = Context(problem).Construct().Build().Initialise()
    if variant is not None:
        ctx.set_elem_variant(variant)
    ctx.Execute(nsteps)
    got = ctx.state(fields=fields)
    gsc = ctx.scalars()
    ctx.Destroy()
    return ref, got, rsc, gsc


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(a)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def n_bit_equal(a, b):
    return int(np.sum(np.asarray(a) == np.asarray(b))), int(np.asarray(a).size)
