"""The PME charge spread fused with the z transform (PM_OPT_PME_SPREAD 1, the
default) against the spread by bricks of cells into LDS tiles (0).  Both add
the same 64-bit fixed-point stencil values into each grid point, so the grid,
and with it every force and energy, is the same bit for bit: on single
evaluations (grids whose sides are not multiples of the fused kernel's 4 x 4
lines, orders 4, 6 and 8, cell grids of 2 to 7 cells per side, both
precisions; the fused spread also takes coarse cells with a fine grid, where
the bricks' tiles do not fit LDS: test_fused_spread_coarse_cells) and along MD runs with list reuse, where atoms have drifted from
the cells they were binned in at the last layout build.  The fused path
against the oracle itself is test_pme_small (it is the default there)."""
import sys

import numpy as np
import pytest

from protomol_amd import water
from tests.cases import ALPHA_RC9

pytestmark = pytest.mark.gpu

PME = "Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ctx(factory, s, precision, cell, spread):
    c = factory(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, cell)
    c.set_option("pme_spread", spread)
    assert c.get_option("pme_spread") == spread
    return c


# (cell 9.31, two cells per side: the brick path's tiles fit only the order-4 24^3 grid)
CASES = [(g, o, c) for g, o in [((24, 24, 24), 4), ((32, 30, 20), 6), ((25, 24, 18), 8), ((18, 20, 27), 4),
                                ((21, 28, 14), 4)]
         for c in (2.6, 4.0)] + [((24, 24, 24), 4, 9.31)]


@pytest.mark.parametrize("grid,order,cell", CASES)
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_fused_spread_equals_brick_spread(gpu_ctx_factory, grid, order, cell, precision):
    say(f"case: pme spread fused vs bricks water648 grid {grid} order {order} cell {cell} {precision}")
    s = water.named("water648", seed=5, jitter=0.4)
    out = []
    for spread in (1, 0):
        c = ctx(gpu_ctx_factory, s, precision, cell, spread)
        h = c.force(PME, [grid[0], grid[1], grid[2], 8.0, order, 1e-6, -1.0])
        f, e = c.evaluate([h], s.pos, virial=True)
        out.append((f.copy(), e.copy()))
    say("max |df|", np.abs(out[0][0] - out[1][0]).max(), "E", out[0][1][0], out[1][1][0])
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_fused_spread_md_with_list_reuse(gpu_ctx_factory, precision):
    """60 leapfrog steps of the 8,100-atom box (LJ-C2 + Ewald-real rc 9, PME
    48^3, bonds and angles) with the default list skin: the layout is kept
    over most steps while atoms drift up to half the skin from their cells,
    and the fused spread must still reach every stencil point; the two
    trajectories are the same bit for bit."""
    say(f"case: md 60 steps water8k fused vs brick spread {precision}")
    s = water.named("water8k", seed=3)
    out = {}
    for spread in (1, 0):
        c = ctx(gpu_ctx_factory, s, precision, 4.5, spread)
        hs = [c.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 "
                      "-switchingFunction Cutoff", [8.0, 9.0, ALPHA_RC9, 9.0, 9.0]),
              c.force(PME, [48, 48, 48, 9.0, 4, 1e-6, -1.0]), c.force("Bond", []), c.force("Angle", [])]
        c.md_setup(hs, "leapfrog", dt=1.0)
        c.md_set_state(s.pos, s.vel)
        b0 = c.stats()["list_builds"]
        c.md_run(60)
        x, v, e = c.md_get_state()
        out[spread] = (x, v, e, c.stats()["list_builds"] - b0)
    say("builds", out[1][3], out[0][3], "max |dx|", np.abs(out[1][0] - out[0][0]).max())
    assert 0 < out[1][3] < 50
    assert np.array_equal(out[1][0], out[0][0])
    assert np.array_equal(out[1][1], out[0][1])


def test_fused_spread_coarse_cells(gpu_ctx_factory, port):
    """Two cells per side with a 32 x 30 x 20 grid at order 6: the brick
    spread's LDS tile would exceed 64 KB (that path refuses it); the fused
    spread needs no tile and matches the oracle."""
    from tests.cases import rms_rel
    say("case: pme fused spread water648 cell 9.31 grid (32, 30, 20) order 6")
    s = water.named("water648", seed=5, jitter=0.4)
    for precision, tol in (("mixed", 5e-6), ("double", 1e-10)):
        c = ctx(gpu_ctx_factory, s, precision, 9.31, 1)
        f, e = c.evaluate([c.force(PME, [32, 30, 20, 8.0, 6, 1e-6, -1.0])], s.pos)
        o = port.pme(s, (32, 30, 20), 6, 8.0, 9.31, flags=6)
        say(precision, "rms rel force", rms_rel(f, o["forces"]))
        assert rms_rel(f, o["forces"]) < tol
    c = ctx(gpu_ctx_factory, s, "mixed", 9.31, 0)
    from protomol_amd.gpu import PmError
    with pytest.raises(PmError, match="spread tile"):
        c.evaluate([c.force(PME, [32, 30, 20, 8.0, 6, 1e-6, -1.0])], s.pos)


def test_fused_spread_long_z_lines(gpu_ctx_factory, port):
    """A 240-point z axis: 4 x 4 z lines no longer fit a block's LDS, so the
    fused spread owns 4 x 2 lines (mixed) or 4 x 1 (FP64); against the oracle."""
    from tests.cases import rms_rel
    say("case: pme fused spread water648 grid (20, 24, 240) order 4")
    s = water.named("water648", seed=5, jitter=0.4)
    for precision, tol in (("mixed", 5e-6), ("double", 1e-10)):
        c = ctx(gpu_ctx_factory, s, precision, 4.0, 1)
        f, e = c.evaluate([c.force(PME, [20, 24, 240, 8.0, 4, 1e-6, -1.0])], s.pos)
        o = port.pme(s, (20, 24, 240), 4, 8.0, 4.0, flags=6)
        say(precision, "rms rel force", rms_rel(f, o["forces"]))
        assert rms_rel(f, o["forces"]) < tol
