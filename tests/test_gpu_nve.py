"""NVE energy series of the 98,304-atom water box (LJ C2 9 -> 10 A + Ewald
real space, PME 96^3 order 4, bonds, angles; leapfrog, dt 1 fs) against the
series the compiled reference's forces produce under the same leapfrog
(tests/golden/nve_water98k_reference.json, written by
tests/golden/make_nve_reference.py).  The trajectories stay together until
round-off growth separates them (chaos), so the pointwise comparison stops at
step 201; the 1000-step drift slopes are compared as numbers."""
import json
import os
import sys

import numpy as np
import pytest

from protomol_amd import water
from tests.cases import ALPHA_RC10

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def series(ctx, s, chunks):
    hs = [ctx.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
                    [9.0, 10.0, ALPHA_RC10, 10.0, 10.0]),
          ctx.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                    [96, 96, 96, 10.0, 4, 1e-6, -1.0]),
          ctx.force("Bond", []), ctx.force("Angle", [])]
    ctx.md_setup(hs, "leapfrog", dt=1.0)
    ctx.md_set_state(s.pos, s.vel)
    out = []
    for k in range(chunks):
        ctx.md_run(1 if k == 0 else 100)
        _, v, en = ctx.md_get_state()
        out.append(0.5 * float((s.mass[:, None] * v * v).sum()) + float(en[:4].sum()))
    return np.array(out)


@pytest.mark.parametrize("precision,tol_e,tol_slope", [("double", 1e-6, 0.05), ("mixed", 0.5, 0.05)])
def test_nve_series_matches_reference(gpu_ctx_factory, precision, tol_e, tol_slope):
    with open(os.path.join(HERE, "golden", "nve_water98k_reference.json")) as fh:
        ref = json.load(fh)
    say(f"case: NVE water98k 1001 steps, {precision}")
    s = water.named("water98k", seed=1)
    c = gpu_ctx_factory(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 5.0)
    e = series(c, s, 11)
    r = np.array(ref["e_total"])
    say("E steps 1/101/201", e[:3].tolist(), "reference", r[:3].tolist())
    assert np.abs(e[:3] - r[:3]).max() <= tol_e
    t = np.array(ref["t_ns"])
    slope = float(np.polyfit(t, e, 1)[0]) / (3 * s.n)
    say("drift kcal/mol/ns/dof", slope, "reference", ref["kcal_mol_per_ns_per_dof"])
    assert abs(slope - ref["kcal_mol_per_ns_per_dof"]) <= tol_slope * abs(ref["kcal_mol_per_ns_per_dof"])
