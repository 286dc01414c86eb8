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


def make_inner(factory, s, precision, skin, prune):
    c, hs = make(factory, s, precision, skin)
    c.set_option("prune_skin", prune)
    return c, hs


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 2e-6)])
def test_inner_lists_match_fresh_lists(gpu_ctx_factory, precision, tol):
    """Inner lists (PM_OPT_PRUNE_SKIN 0.2 A under a 1.0 A list skin): moves
    below half the inner skin reuse the pruned lists, a move past it prunes
    again without a build, a move past half the list skin builds; the forces
    equal a context that rebuilds at every evaluation throughout, and the
    pruned lists issue fewer lanes than the full ones."""
    say(f"case: inner lists water8k {precision}")
    s = water.named("water8k", seed=5)
    c, hs = make_inner(gpu_ctx_factory, s, precision, 1000, 200)
    full, fhs = make(gpu_ctx_factory, s, precision, 1000)
    ref, rhs = make(gpu_ctx_factory, s, precision, 0)
    assert c.get_option("prune_skin") == 200
    rng = np.random.default_rng(6)
    x = s.pos.copy()

    def check(xx, what):
        f, e = c.evaluate(hs, xx)
        fr, er = ref.evaluate(rhs, xx)
        ff, _ = full.evaluate(fhs, xx)
        st, sf = c.stats(), full.stats()
        say(what, "rms", rms_rel(f, fr), "builds", st["list_builds"], "prunes", st["list_prunes"],
            "lanes", st["lanes"], "full lists", sf["lanes"])
        assert rms_rel(f, fr) <= tol
        assert rms_rel(ff, fr) <= tol
        assert abs(e[0] - er[0]) <= tol * abs(er[0]) and abs(e[1] - er[1]) <= tol * max(abs(er[1]), 1.0)
        assert st["lanes"] < sf["lanes"]
        return st

    st0 = check(x, "start")
    for k in range(3):
        x = x + 0.03 * rng.uniform(-1, 1, x.shape) / np.sqrt(3)   # 0.09 A in total, below 0.1 A
        st = check(x, f"small move {k}")
        assert st["list_builds"] == st0["list_builds"] and st["list_prunes"] == st0["list_prunes"]
    x[11] += np.array([0.0, 0.15, 0.0])                           # past 0.1 A, below 0.5 A
    st = check(x, "0.15 A move")
    assert st["list_builds"] == st0["list_builds"] and st["list_prunes"] == st0["list_prunes"] + 1
    x[13] += np.array([0.0, 0.0, 0.6])                            # past 0.5 A
    st = check(x, "0.6 A move")
    assert st["list_builds"] == st0["list_builds"] + 1


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_md_with_inner_lists_follows_rebuild_every_step(gpu_ctx_factory, precision):
    """Leapfrog, 60 steps of the 8,100-atom box with PME, inner lists (0.3 A
    under a 1.0 A list skin) against lists rebuilt at every step: the same
    trajectory (FP64: to accumulation order; mixed: within the FP32 rounding
    of pair sums grouped differently), with prunes and builds both taking
    place."""
    say(f"case: md 60 steps inner lists vs rebuild every step, {precision}")
    s = water.named("water8k", seed=3)
    out = {}
    for skin, prune in ((0, 0), (1000, 300)):
        c, hs = make_inner(gpu_ctx_factory, s, precision, skin, prune)
        hs = hs + [c.force("Bond", []), c.force("Angle", [])]
        c.md_setup(hs, "leapfrog", dt=1.0)
        c.md_set_state(s.pos, s.vel)
        st0 = c.stats()
        c.md_run(60)
        x, v, e = c.md_get_state()
        st = c.stats()
        out[skin] = (x, v, e, st["list_builds"] - st0["list_builds"], st["list_prunes"] - st0["list_prunes"])
    say("builds", out[1000][3], "prunes", out[1000][4])
    assert 0 < out[1000][3] < 30 and out[1000][4] > 0
    dx = np.abs(out[1000][0] - out[0][0]).max()
    say("max |dx|", dx, "E", out[0][2][:2], out[1000][2][:2])
    assert dx < (1e-8 if precision == "double" else 1e-4)


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 2e-6)])
def test_exact_list_test_matches_bounding_box(gpu_ctx_factory, precision, tol):
    """PM_OPT_LIST_TEST 1 (the exact per-atom distance for the list, the
    default) against 0 (the bounding-box test, in a box where it applies):
    the same forces as each other and as lists rebuilt at every evaluation,
    with fewer entries, over moves that keep the lists and one that rebuilds
    them."""
    say(f"case: exact list test water8k {precision}")
    s = water.named("water8k", seed=8)
    c, hs = make(gpu_ctx_factory, s, precision, 800)
    c.set_option("list_test", 1)
    assert c.get_option("list_test") == 1
    bb, bhs = make(gpu_ctx_factory, s, precision, 800)
    bb.set_option("list_test", 0)
    assert bb.get_option("list_test") == 0
    ref, rhs = make(gpu_ctx_factory, s, precision, 0)
    rng = np.random.default_rng(9)
    x = s.pos.copy()
    for step, amp in enumerate([0.0, 0.1, 0.1, 0.5]):
        x = x + amp * rng.uniform(-1, 1, x.shape) / np.sqrt(3)
        f, e = c.evaluate(hs, x)
        fb, _ = bb.evaluate(bhs, x)
        fr, er = ref.evaluate(rhs, x)
        st, sb = c.stats(), bb.stats()
        say(step, "rms", rms_rel(f, fr), rms_rel(fb, fr), "entries", st["list_entries"], sb["list_entries"])
        assert rms_rel(f, fr) <= tol and rms_rel(fb, fr) <= tol
        # the group's Coulomb energy is a near cancellation of the real-space
        # and reciprocal parts: its tolerance follows the size of the parts
        assert abs(e[0] - er[0]) <= tol * max(abs(er[0]), 10 * abs(er[1]))
        assert st["list_entries"] < sb["list_entries"]
