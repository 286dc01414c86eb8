"""pm_md_autotune: times the MD step over cluster sizes and list skins,
restores the state it started from, and leaves the fastest options set; a
run after tuning is the run a fresh context with those options makes."""
import sys

import numpy as np
import pytest

from protomol_amd import water
from tests.cases import ALPHA_RC9

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(factory, s, integrator, m=None, skin=None):
    c = factory("mixed")
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.5)
    if m is not None:
        c.set_option("cluster_size", m)
        c.set_option("list_skin", skin)
    hs = [c.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
                  [8.0, 9.0, ALPHA_RC9, 9.0, 9.0]),
          c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", [48, 48, 48, 9.0, 4, 1e-6, -1.0]),
          c.force("Bond", []), c.force("Angle", [])]
    c.md_setup(hs, integrator, dt=1.0, temperature=300.0, gamma=5.0, seed=11)
    c.md_set_state(s.pos, s.vel)
    return c


@pytest.mark.parametrize("integrator", ["leapfrog", "langevin"])
def test_autotune_restores_state_and_keeps_the_choice(gpu_ctx_factory, integrator):
    say(f"case: autotune water8k {integrator}")
    s = water.named("water8k", seed=5)
    c = setup(gpu_ctx_factory, s, integrator)
    c.md_run(5)
    x0, v0, _ = c.md_get_state()
    m, skin, ms = c.md_autotune(trial_steps=12)
    say("chosen cluster", m, "skin", skin, "ms/step", ms)
    assert m in (8, 4) and skin in (0, 400, 800, 1200) and ms > 0
    assert c.get_option("cluster_size") == m and c.get_option("list_skin") == skin
    x1, v1, _ = c.md_get_state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    c.md_run(10)
    xa, va, ea = c.md_get_state()
    # a fresh context with the chosen options, from the same state
    f = setup(gpu_ctx_factory, s, integrator, m, skin)
    f.md_run(5)
    f.md_set_state(x0, v0)
    f.md_run(10)
    xb, vb, eb = f.md_get_state()
    say("max |dx|", float(np.abs(xa - xb).max()), "E", ea[:2], eb[:2])
    # the two runs may differ in their list layouts (FP32 rounding of the j
    # row sums, ~1e-8 A after 10 steps); a wrong Langevin noise counter would
    # move atoms by ~1e-3 A per step
    assert np.abs(xa - xb).max() <= 1e-6
