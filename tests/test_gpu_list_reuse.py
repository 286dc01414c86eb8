"""Pair-list reuse (PM_OPT_LIST_SKIN): lists kept between evaluations while
every atom stays within half the skin of its position at the build, rebuilt
on the device otherwise.  The evaluated pair set is the pairs within the
cutoff in both cases, so forces agree with a context that rebuilds at every
evaluation (FP64: to the order of accumulation), and an MD run follows the
same trajectory."""
import sys

import numpy as np
import pytest

from protomol_amd import water
from tests.cases import ALPHA_RC9, rms_rel

pytestmark = pytest.mark.gpu

FID = ("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
       [8.0, 9.0, ALPHA_RC9, 9.0, 9.0])
PME = ("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", [48, 48, 48, 9.0, 4, 1e-6, -1.0])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(factory, s, precision, skin):
    c = factory(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.5)
    c.set_option("list_skin", skin)
    return c, [c.force(*FID), c.force(*PME)]


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 2e-6)])
def test_reuse_and_rebuild_match_fresh_lists(gpu_ctx_factory, precision, tol):
    say(f"case: list reuse water8k {precision}")
    s = water.named("water8k", seed=3)
    c, hs = make(gpu_ctx_factory, s, precision, 600)
    ref, rhs = make(gpu_ctx_factory, s, precision, 0)
    assert c.get_option("list_skin") == 600 and ref.get_option("list_skin") == 0
    rng = np.random.default_rng(4)
    x = s.pos.copy()
    b0 = None
    for step, amp in enumerate([0.0, 0.05, 0.1, 0.1, 0.0]):
        x = x + amp * rng.uniform(-1, 1, x.shape) / np.sqrt(3)   # every move below 0.6 / 2 A in total
        f, e = c.evaluate(hs, x)
        fr, er = ref.evaluate(rhs, x)
        builds = c.stats()["list_builds"]
        if b0 is None:
            b0 = builds
        say(step, "rms", rms_rel(f, fr), "builds", builds)
        assert rms_rel(f, fr) <= tol
        assert abs(e[0] - er[0]) <= tol * abs(er[0]) and abs(e[1] - er[1]) <= tol * max(abs(er[1]), 1.0)
    # the moves above add up to less than half the skin: one build only
    assert c.stats()["list_builds"] == b0
    # one atom moved past half the skin: rebuilt, still the same forces
    x2 = x.copy()
    x2[7] += np.array([0.35, 0.0, 0.0])
    f, _ = c.evaluate(hs, x2)
    fr, _ = ref.evaluate(rhs, x2)
    say("after a 0.35 A move: rms", rms_rel(f, fr), "builds", c.stats()["list_builds"])
    assert c.stats()["list_builds"] == b0 + 1
    assert rms_rel(f, fr) <= tol


@pytest.mark.parametrize("integrator", ["leapfrog", "langevin"])
def test_md_with_reuse_follows_rebuild_every_step(gpu_ctx_factory, integrator):
    """FP64 leapfrog or Langevin, 50 steps of the 8,100-atom box with PME:
    the same trajectory (to accumulation order) with and without list reuse.
    With reuse the position update of each step also runs the displacement
    test that decides the next rebuild."""
    say(f"case: md 50 steps with/without list reuse, FP64, {integrator}")
    s = water.named("water8k", seed=3)
    out = {}
    for skin in (0, 600):
        c, hs = make(gpu_ctx_factory, s, "double", skin)
        hs = hs + [c.force("Bond", []), c.force("Angle", [])]
        c.md_setup(hs, integrator, dt=1.0, temperature=300.0, gamma=5.0, seed=11)
        c.md_set_state(s.pos, s.vel)
        b0 = c.stats()["list_builds"]
        c.md_run(50)
        x, v, e = c.md_get_state()
        out[skin] = (x, v, e, c.stats()["list_builds"] - b0)
    say("builds with skin 0.6 A:", out[600][3])
    assert 0 < out[600][3] < 48   # some steps reused their lists
    dx = np.abs(out[600][0] - out[0][0]).max()
    say("max |dx|", dx, "E", out[0][2][:2], out[600][2][:2])
    assert dx < 1e-8
