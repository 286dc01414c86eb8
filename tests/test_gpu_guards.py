"""Failure reporting: inputs the fixed-point force path cannot represent end
in PM_ERR_NUMERIC instead of silently wrong forces, and the context stays
usable.  The kernels only raise a flag (every access stays in bounds); the
host turns it into the error after the evaluation."""
import sys

import numpy as np
import pytest

from protomol_amd import gpu, water
from tests.cases import gpu_forces_c1

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ctx(make, s, precision):
    c = make(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.0)
    return c


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_nonfinite_position_is_an_error(gpu_ctx_factory, precision):
    say(f"case: NaN position water648 {precision}")
    s = water.named("water648", seed=13, jitter=0.3)
    c = _ctx(gpu_ctx_factory, s, precision)
    hs = gpu_forces_c1(c) + [c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                                     [24, 24, 24, 8.0, 4, 1e-6, -1.0])]
    f0, e0 = c.evaluate(hs, s.pos)
    bad = s.pos.copy()
    bad[17, 1] = np.nan
    with pytest.raises(gpu.PmError) as ei:
        c.evaluate(hs, bad)
    assert ei.value.code == gpu.PM_ERR_NUMERIC
    say("error:", ei.value.msg)
    # the flag is cleared: the next valid evaluation gives the same result
    f1, e1 = c.evaluate(hs, s.pos)
    np.testing.assert_array_equal(f0, f1)
    np.testing.assert_array_equal(e0, e1)


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_force_beyond_fixed_point_range_is_an_error(gpu_ctx_factory, precision):
    """Two oxygens 0.05 A apart: the LJ force (~1e23 kcal/mol/A) is far
    beyond the 64-bit fixed-point range of +-2^23 kcal/mol/A."""
    say(f"case: overlapping oxygens water648 {precision}")
    s = water.named("water648", seed=14, jitter=0.3)
    c = _ctx(gpu_ctx_factory, s, precision)
    hs = gpu_forces_c1(c)
    bad = s.pos.copy()
    bad[3] = bad[0] + np.array([0.05, 0.0, 0.0])
    with pytest.raises(gpu.PmError) as ei:
        c.evaluate(hs, bad)
    assert ei.value.code == gpu.PM_ERR_NUMERIC
    say("error:", ei.value.msg)
    c.evaluate(hs, s.pos)


def test_md_run_reports_nonfinite_state(gpu_ctx_factory):
    say("case: MD from a NaN velocity water648 mixed")
    s = water.named("water648", seed=15, jitter=0.3)
    c = _ctx(gpu_ctx_factory, s, "mixed")
    c.md_setup(gpu_forces_c1(c) + [c.force("Bond", []), c.force("Angle", [])], "leapfrog", dt=1.0)
    v = s.vel.copy()
    v[5, 2] = np.inf
    c.md_set_state(s.pos, v)
    c.md_run(3)    # queues the steps; the flag is read when the run is synchronized
    with pytest.raises(gpu.PmError) as ei:
        c.synchronize()
    assert ei.value.code == gpu.PM_ERR_NUMERIC
    say("error:", ei.value.msg)
    c.md_set_state(s.pos, s.vel)
    c.md_run(3)
    c.synchronize()


def test_molecular_virial_request_is_refused(gpu_ctx_factory):
    """A ScalarStructure with molecularVirial() on must not get an evaluation
    without it: the request is refused with PM_ERR_UNSUPPORTED."""
    say("case: molecular virial request water648")
    s = water.named("water648", seed=16, jitter=0.3)
    c = _ctx(gpu_ctx_factory, s, "mixed")
    with pytest.raises(gpu.PmError) as ei:
        c.evaluate(gpu_forces_c1(c), s.pos, virial=True, molecular_virial=True)
    assert ei.value.code == gpu.PM_ERR_UNSUPPORTED
