"""GPU parity on boxes narrower than twice the cutoff and on the 64-replica
workload, against the FP64 oracle.

Below L = 2 rc the reference enumerator maps its deltas to basis cells and
deduplicates them when the grid has fewer than three cells per axis
(topology/CellListEnumerator_periodicBoundaries.h:63-102,
ArrayCellListStructure.h:108-130), and every pair is counted once at its
minimal image (PeriodicBoundaryConditions.h:76-117).  The oracle follows
those rules bit for bit (tests/test_oracle.py::test_reference_small_box_semantics
pins it against the compiled reference); these tests hold the HIP path to it.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from protomol_amd import water
from tests.cases import ALPHA_RC8, c1_terms, gpu_forces_c1, replica_context, rms_rel

pytestmark = pytest.mark.gpu

# (GPU force id, values, oracle terms): the four force sets of the oracle's
# small-box test
CASES = {
    "lj_c2": ("LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0],
              [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)]),
    "coul_c2": ("Coulomb -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0],
                [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)]),
    "coul_c1": ("Coulomb -algorithm NonbondedCutoff -switchingFunction C1", [8.0, 8.0],
                [po.term(po.POT_COULOMB, po.SWF_C1, 8.0)]),
    "lj_ewr": ("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
               [7.0, 8.0, ALPHA_RC8, 8.0, 8.0],
               [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA_RC8)]),
}
TOL = {"mixed": 5e-6, "double": 1e-11}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup_ctx(make, s, cellsize, precision):
    c = make(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, cellsize)
    return c


@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("cellsize", [4.0, 3.0, 9.3])
def test_small_box_below_twice_cutoff(gpu_ctx_factory, port, cellsize, precision):
    """27 waters in a 9.3 A cube with cutoff 8: 2, 3 or 1 cells per axis.
    Per-cell candidate and in-cutoff pair counts are the reference
    enumerator's, bit for bit; forces, energies and the virial match the
    oracle for every force set."""
    s = water.water_box(27, 9.3, seed=4, jitter=0.2)
    tol = TOL[precision]
    c = setup_ctx(gpu_ctx_factory, s, cellsize, precision)
    for name, (fid, vals, terms) in CASES.items():
        say(f"case: small box 9.3 A cell {cellsize} {name} {precision}")
        o = port.cutoff(s, terms, 8.0, cellsize, virial=True, counts=True)
        d = c.cell_diagnostics(s.pos, 8.0, 8.0)
        np.testing.assert_array_equal(d["cand"], o["cand"])
        np.testing.assert_array_equal(d["inrc"], o["inrc"])
        f, e = c.evaluate([c.force(fid, vals)], s.pos, virial=True)
        err = rms_rel(f, o["forces"])
        say("rms rel force", err, "E", e[:2], o["energies"][:2])
        assert err < tol
        # the 27-water box's LJ energy is a near-total cancellation of
        # repulsion and attraction (about -0.06 kcal/mol): the energy check
        # has an absolute floor of tol x 10 x (|E_C| + |E_LJ| + 10 kcal/mol)
        np.testing.assert_allclose(e[:2], o["energies"][:2], rtol=tol * 10,
                                   atol=tol * 10 * (np.abs(o["energies"][:2]).sum() + 10.0))
        np.testing.assert_allclose(e[7:], o["energies"][7:], rtol=tol * 50,
                                   atol=tol * 50 * np.abs(o["energies"][7:]).max())


@pytest.mark.parametrize("cluster", [8, 4])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_smallest_replica_box(gpu_ctx_factory, port, precision, cluster):
    """The smallest replica of the 64-replica workload: 100 waters in a
    14.4 A cube (0.0334 molecules/A^3), LJ and Coulomb with C2 switching
    7 -> 8 A (L < 2 rc = 16 A), cell size 4."""
    n, edges = water.replica_sizes(1, 0, 100, 100)
    s = water.water_box(int(n[0]), float(edges[0]), seed=31, jitter=0.2)
    assert s.box[0] < 16.0
    say(f"case: replica box {s.box[0]:.2f} A cluster {cluster} {precision}")
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    c.set_option("cluster_size", cluster)
    f, e = c.evaluate(gpu_forces_c1(c), s.pos, virial=True)
    lj, co = c1_terms()
    a = port.cutoff(s, lj, 8.0, 4.0, virial=True, counts=True)
    b = port.cutoff(s, co, 8.0, 4.0, virial=True)
    fr = a["forces"] + b["forces"]
    er = a["energies"] + b["energies"]
    tol = TOL[precision]
    err = rms_rel(f, fr)
    say("rms rel force", err, "E", e[:2], er[:2])
    assert err < tol
    np.testing.assert_allclose(e[:2], er[:2], rtol=tol * 10, atol=1e-9)
    np.testing.assert_allclose(e[7:], er[7:], rtol=tol * 50, atol=tol * 50 * np.abs(er[7:]).max())
    d = c.cell_diagnostics(s.pos, 8.0, 8.0)
    np.testing.assert_array_equal(d["inrc"], a["inrc"])


def test_replica_workload_64(gpu_ctx_factory, port):
    """The 64-replica workload (uneven boxes of 100-1000 waters, 14.4-31 A,
    LJ + Coulomb C2 7 -> 8 A, bonds, angles, Langevin on the 300-450 K ladder)
    advanced as one batch of batched launches:
      * after the steps, every replica's LJ and Coulomb energies of the last
        force evaluation match the oracle at the replica's positions;
      * the forces of every replica at those positions (single-context
        evaluation) match the oracle;
      * the batched trajectories equal, bit for bit, each replica run alone."""
    from protomol_amd import gpu
    systems = water.replica_boxes(64, seed=0, lo=100, hi=1000)
    temps = water.replica_temperatures(64)
    small = sum(1 for s in systems if s.box[0] < 16.0)
    say(f"case: 64 replicas ({sum(s.n for s in systems)} atoms, {small} below 2 rc), batch 12 steps")
    ctxs = [replica_context(gpu_ctx_factory, s, t, seed=k + 1) for k, (s, t) in enumerate(zip(systems, temps))]
    b = gpu.Batch(ctxs)
    b.run(5)
    b.run(7)
    batched, why = b.info()
    say("batched launches:", batched, why)
    assert batched
    got = [c.md_get_state() for c in ctxs]
    b.close()
    lj, co = c1_terms()
    worst_e = 0.0
    worst_f = 0.0
    for k, (s, c) in enumerate(zip(systems, ctxs)):
        x, _, e = got[k]
        s2 = s.copy()
        s2.pos = x
        a = port.cutoff(s2, lj, 8.0, 4.0)
        o = port.cutoff(s2, co, 8.0, 4.0)
        er = a["energies"][:2] + o["energies"][:2]
        rel = np.abs(e[:2] - er) / np.abs(er)
        worst_e = max(worst_e, float(rel.max()))
        assert rel.max() < 5e-5, (k, s.n, s.box[0], e[:2], er)
        h = gpu_forces_c1(c)
        f, _ = c.evaluate(h, x)
        err = rms_rel(f, a["forces"] + o["forces"])
        worst_f = max(worst_f, err)
        assert err < 5e-6, (k, s.n, s.box[0], err)
    say("worst relative energy error", worst_e, "worst rms rel force", worst_f)
    for k, (s, t) in enumerate(zip(systems, temps)):
        c = replica_context(gpu_ctx_factory, s, t, seed=k + 1)
        c.set_option("graph", 0)
        c.md_run(5)
        c.md_run(7)
        ref = c.md_get_state()
        for a_, r_ in zip(got[k], ref):
            np.testing.assert_array_equal(a_, r_)
        c.close()


@pytest.mark.parametrize("precision,tol", [("mixed", 5e-6), ("double", 1e-10)])
def test_pme_charged_system(gpu_ctx_factory, port, precision, tol):
    """A net-charged box (216 waters and one Na+): PME reciprocal + correction
    with the charged-system energy and virial term
    (NonbondedPMEwaldSystemForce.h:300-305, 618-628) against the oracle."""
    say(f"case: pme charged water648 + ion {precision}")
    s = water.with_ion(water.named("water648", seed=26, jitter=0.3), 1.0, (0.3, -0.2, 0.1))
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    h = c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                [24, 24, 24, 8.0, 4, 1e-6, -1.0])
    f, e = c.evaluate([h], s.pos, virial=True)
    o = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=6, virial=True)
    err = rms_rel(f, o["forces"])
    scale = sum(abs(o["parts"][k]) for k in ("reciprocal", "intra", "self"))
    say("rms rel force", err, "E", e[0], o["energies"][0], "scale", scale)
    assert err < tol
    assert abs(e[0] - o["energies"][0]) < tol * 0.2 * scale
    np.testing.assert_allclose(e[7:], o["energies"][7:], rtol=tol * 100,
                               atol=tol * 100 * np.abs(o["energies"][7:]).max())


@pytest.mark.parametrize("prune", [0, 300])
def test_small_box_list_reuse_follows_rebuild_every_step(gpu_ctx_factory, prune):
    """List reuse in boxes below 2 (cutoff + skin + cell diagonal), whose lists
    take the exact per-atom test: FP64 Langevin, 60 steps of the smallest
    replica box (100 waters, 14.4 A cube, L < 2 rc) and of a 300-water box
    with list skin 0.6 A (and inner lists at 0.3 A) follow the same
    trajectory as lists rebuilt at every step, with fewer builds than
    steps."""
    for nmol in (100, 300):
        L = (nmol / 0.0334) ** (1.0 / 3.0)
        s = water.water_box(nmol, L, seed=40 + nmol, jitter=0.2)
        say(f"case: small box reuse {nmol} waters {L:.2f} A prune {prune}")
        out = {}
        for skin in (0, 600):
            c = setup_ctx(gpu_ctx_factory, s, 4.0, "double")
            c.set_option("list_skin", skin)
            c.set_option("prune_skin", prune if skin else 0)
            hs = gpu_forces_c1(c) + [c.force("Bond", []), c.force("Angle", [])]
            c.md_setup(hs, "langevin", dt=1.0, temperature=300.0, gamma=5.0, seed=3)
            c.md_set_state(s.pos, s.vel)
            st0 = c.stats()
            c.md_run(60)
            x, v, e = c.md_get_state()
            st = c.stats()
            out[skin] = (x, e, st["list_builds"] - st0["list_builds"], st["list_prunes"] - st0["list_prunes"])
        dx = np.abs(out[600][0] - out[0][0]).max()
        say("builds", out[600][2], "prunes", out[600][3], "max |dx|", dx, "E", out[0][1][:2], out[600][1][:2])
        assert 0 < out[600][2] < 60          # the lattice start heats the water: builds are frequent
        if prune:
            assert out[600][3] > 0
        assert dx < 1e-8


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-10), ("double", 1e-11)])
@pytest.mark.parametrize("nmol,edge", [(27, 9.3), (100, 14.4)])
def test_pme_real_term_lattice_images(gpu_ctx_factory, port, nmol, edge, precision, tol):
    """The PME real term with a box no wider than twice the cutoff (rc 8):
    pairs at every lattice image within the cutoff and each atom against its
    own images (NonbondedPMEwaldSystemForce.h:658-768), through k_ewald_lattice
    (FP64 in both precisions), every exclusion class; the real term alone and
    the full PME force against the oracle (test_oracle.py pins the oracle's
    lattice term against the compiled reference)."""
    from tests.cases import exclusion_mix
    say(f"case: pme real lattice images {nmol} waters {edge} A {precision}")
    s = exclusion_mix(water.water_box(nmol, edge, seed=6, jitter=0.3))
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    args = [16, 16, 16, 8.0, 4, 1e-6, -1.0]
    hr = c.force("Coulomb -algorithm PMEwald -real -interpolation BSpline", args)
    f, e = c.evaluate([hr], s.pos, virial=True)
    o = port.pme(s, (16, 16, 16), 4, 8.0, 4.0, flags=1, virial=True)
    say("real: rms rel force", rms_rel(f, o["forces"]), "E", e[0], o["energies"][0])
    assert rms_rel(f, o["forces"]) < tol
    np.testing.assert_allclose(e[0], o["energies"][0], rtol=1e-10)
    np.testing.assert_allclose(e[7:16], o["energies"][7:16], rtol=1e-9, atol=1e-9 * np.abs(o["energies"][7:16]).max())
    hf = c.force("CoulombPME", args)
    f, e = c.evaluate([hf], s.pos)
    o = port.pme(s, (16, 16, 16), 4, 8.0, 4.0, flags=7)
    full_tol = 5e-6 if precision == "mixed" else 1e-10
    say("full: rms rel force", rms_rel(f, o["forces"]))
    assert rms_rel(f, o["forces"]) < full_tol
    scale = sum(abs(o["parts"][k]) for k in ("real", "reciprocal", "intra", "self"))
    assert abs(e[0] - o["energies"][0]) < full_tol * 0.2 * scale
