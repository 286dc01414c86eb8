"""GPU parity: the HIP kernels through the C ABI against the FP64 oracle."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from protomol_amd import water
from tests.cases import ALPHA_RC8, ALPHA_RC10, c1_terms, exclusion_mix, gpu_forces_c1, rms_rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup_ctx(make, s, cellsize, precision="mixed"):
    c = make(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, cellsize)
    return c


def test_cell_binning_bitexact(gpu_ctx_factory, port):
    say("case: binning water648 cell 4 cutoff 8")
    s = water.named("water648", seed=1)
    c = setup_ctx(gpu_ctx_factory, s, 4.0)
    g = c.cell_diagnostics(s.pos, 8.0, 8.0)
    o = port.cells(s, 4.0)
    assert g["dims"] == o["dims"]
    np.testing.assert_array_equal(g["cell"], o["cell"])
    np.testing.assert_array_equal(g["head"], o["head"])
    np.testing.assert_array_equal(g["next"], o["next"])
    lj, _ = c1_terms()
    oc = port.cutoff(s, lj, 8.0, 4.0, counts=True)
    np.testing.assert_array_equal(g["cand"], oc["cand"])
    np.testing.assert_array_equal(g["inrc"], oc["inrc"])


def test_cell_binning_bitexact_large_grid(gpu_ctx_factory, port):
    """A grid of more than two 8192-cell chunks takes the multi-block cell
    scan: cells, head/next lists and per-cell counts as the reference's."""
    say("case: binning water8k cell 1.5 (27^3 cells) cutoff 4")
    s = water.named("water8k", seed=2)
    c = setup_ctx(gpu_ctx_factory, s, 1.5)
    g = c.cell_diagnostics(s.pos, 4.0, 4.0)
    o = port.cells(s, 1.5)
    assert g["dims"] == o["dims"]
    assert int(np.prod(g["dims"])) > 2 * 8192
    np.testing.assert_array_equal(g["cell"], o["cell"])
    np.testing.assert_array_equal(g["head"], o["head"])
    np.testing.assert_array_equal(g["next"], o["next"])
    lj, _ = c1_terms()
    oc = port.cutoff(s, [po_term_lj4()], 4.0, 1.5, counts=True)
    np.testing.assert_array_equal(g["cand"], oc["cand"])
    np.testing.assert_array_equal(g["inrc"], oc["inrc"])


def po_term_lj4():
    from oracle import pyoracle as po
    return po.term(po.POT_LJ, po.SWF_C2, 4.0, 3.0)


@pytest.mark.parametrize("precision,tol", [("mixed", 2e-6), ("double", 1e-11)])
def test_forces_large_grid(gpu_ctx_factory, port, precision, tol):
    """Forces with the multi-block cell scan: the 98,304-atom box with
    3.6 A cells (27^3 = 19,683 cells, three scan blocks)."""
    say(f"case: water98k LJ+Coulomb C2 cell 3.6 {precision}")
    s = water.named("water98k", seed=3)
    c = setup_ctx(gpu_ctx_factory, s, 3.6, precision)
    hs = gpu_forces_c1(c)
    f, e = c.evaluate(hs, s.pos)
    lj, co = c1_terms()
    o = port.cutoff(s, lj, 8.0, 3.6)
    o2 = port.cutoff(s, co, 8.0, 3.6)
    fr = o["forces"] + o2["forces"]
    err = rms_rel(f, fr)
    say("rms rel", err)
    assert err < tol


@pytest.mark.parametrize("cluster", [8, 4])
@pytest.mark.parametrize("precision,tol", [("mixed", 2e-6), ("double", 1e-11)])
def test_cutoff_c1(gpu_ctx_factory, port, precision, tol, cluster):
    say(f"case: water648 LJ+Coulomb C2 {precision} cluster {cluster}")
    s = water.named("water648", seed=1)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    c.set_option("cluster_size", cluster)
    assert c.get_option("cluster_size") == cluster
    f, e = c.evaluate(gpu_forces_c1(c), s.pos, virial=True)
    lj, co = c1_terms()
    a = port.cutoff(s, lj, 8.0, 4.0, virial=True)
    b = port.cutoff(s, co, 8.0, 4.0, virial=True)
    fr = a["forces"] + b["forces"]
    er = a["energies"] + b["energies"]
    say("rms rel force", rms_rel(f, fr), "E", e[:2], er[:2])
    assert rms_rel(f, fr) < tol
    np.testing.assert_allclose(e[:2], er[:2], rtol=tol * 5, atol=0)
    np.testing.assert_allclose(e[7:], er[7:], rtol=tol * 50, atol=tol * 50 * np.abs(er[7:]).max())


VARIANTS = [
    ("LennardJones -algorithm NonbondedCutoff -switchingFunction C1", [8.0, 8.0], [po.term(po.POT_LJ, po.SWF_C1, 8.0)]),
    ("LennardJones -algorithm NonbondedCutoff -switchingFunction Cutoff", [8.0, 8.0], [po.term(po.POT_LJ, po.SWF_CUTOFF, 8.0)]),
    ("Coulomb -algorithm NonbondedCutoff -switchingFunction C1", [8.0, 8.0], [po.term(po.POT_COULOMB, po.SWF_C1, 8.0)]),
    ("Coulomb -algorithm NonbondedCutoff -switchingFunction Cutoff", [8.0, 8.0], [po.term(po.POT_COULOMB, po.SWF_CUTOFF, 8.0)]),
    ("CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction Cutoff", [ALPHA_RC8, 8.0, 8.0],
     [po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA_RC8)]),
    ("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
     [7.0, 8.0, ALPHA_RC8, 8.0, 8.0],
     [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA_RC8)]),
    ("LennardJones Coulomb -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction C1",
     [7.0, 8.0, 8.0, 8.0],
     [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0), po.term(po.POT_COULOMB, po.SWF_C1, 8.0)]),
    # the remaining registered OneAtomPairTwo combinations
    # (factories/registerForceExemplarsCutoff.cpp:76, 91, 131)
    ("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C1 -switchingFunction Cutoff",
     [8.0, ALPHA_RC8, 8.0, 8.0],
     [po.term(po.POT_LJ, po.SWF_C1, 8.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA_RC8)]),
    ("LennardJones Coulomb -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
     [7.0, 8.0, 8.0, 8.0],
     [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0), po.term(po.POT_COULOMB, po.SWF_CUTOFF, 8.0)]),
    ("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction Cutoff -switchingFunction Cutoff",
     [8.0, ALPHA_RC8, 8.0, 8.0],
     [po.term(po.POT_LJ, po.SWF_CUTOFF, 8.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA_RC8)]),
]


def _variant_id(fid):
    w = fid.split()
    sw = [w[i + 1] for i in range(len(w) - 1) if w[i] == "-switchingFunction"]
    return fid.split(" -")[0].replace(" ", "+") + "/" + "/".join(sw)


@pytest.mark.parametrize("fid,vals,terms", VARIANTS, ids=[_variant_id(v[0]) for v in VARIANTS])
@pytest.mark.parametrize("precision,tol", [("mixed", 5e-6), ("double", 1e-11)])
def test_cutoff_variants(gpu_ctx_factory, port, fid, vals, terms, precision, tol):
    say(f"case: variant {fid} {precision}")
    s = water.named("water648", seed=2, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    f, e = c.evaluate([c.force(fid, vals)], s.pos)
    o = port.cutoff(s, terms, vals[-1], 4.0)
    say("rms rel force", rms_rel(f, o["forces"]))
    assert rms_rel(f, o["forces"]) < tol
    np.testing.assert_allclose(e[:2], o["energies"][:2], rtol=tol * 10, atol=1e-9)


@pytest.mark.parametrize("cluster", [8, 4])
@pytest.mark.parametrize("precision,tol", [("mixed", 5e-6), ("double", 1e-11)])
def test_exclusion_rules(gpu_ctx_factory, port, precision, tol, cluster):
    """MODIFIED (A14/B14, csf) and NONE overrides, duplicate and cross-molecule
    exclusion pairs, through the two-term cutoff kernels and the PME correction."""
    say(f"case: exclusion rules water648 {precision} cluster {cluster}")
    s = exclusion_mix(water.named("water648", seed=8, jitter=0.4))
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    c.set_option("cluster_size", cluster)
    for fid, vals, terms in VARIANTS[5:7]:
        f, e = c.evaluate([c.force(fid, vals)], s.pos, virial=True)
        o = port.cutoff(s, terms, vals[-1], 4.0, virial=True)
        say(fid, "rms rel force", rms_rel(f, o["forces"]))
        assert rms_rel(f, o["forces"]) < tol
        np.testing.assert_allclose(e[:2], o["energies"][:2], rtol=tol * 10, atol=1e-9)
    h = c.force("Coulomb -algorithm PMEwald -correction -interpolation BSpline", [24, 24, 24, 8.0, 4, 1e-6, -1.0])
    f, e = c.evaluate([h], s.pos, virial=True)
    o = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=4, virial=True)
    say("correction rms rel force", rms_rel(f, o["forces"]))
    assert rms_rel(f, o["forces"]) < 1e-11
    np.testing.assert_allclose(e[0], o["energies"][0], rtol=1e-11)


@pytest.mark.parametrize("grid,order", [((24, 24, 24), 4), ((32, 30, 20), 6), ((25, 24, 18), 8), ((28, 21, 14), 4),
                                        ((24, 24, 24), 2), ((24, 24, 24), 10), ((32, 30, 20), 12)])
@pytest.mark.parametrize("precision,tol", [("mixed", 5e-6), ("double", 1e-10)])
@pytest.mark.parametrize("fft", [1, 2])
def test_pme_small(gpu_ctx_factory, port, grid, order, precision, tol, fft):
    """PME reciprocal + correction against the oracle, with the y and x
    transforms as separate passes (fft 1) and as one kernel per kz plane
    (fft 2); one grid has a factor 7 on every side (the radix-7 pass,
    lengths that ZFFT's fftfact_ splits by 7, base/fft.c:846-906), and the
    orders run over the even orders 2 to 12 (order 2: BSpline.cpp:74-79)."""
    say(f"case: pme water648 grid {grid} order {order} {precision} fft option {fft}")
    s = water.named("water648", seed=3, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    c.set_option("pme_fft", fft)
    h = c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                [grid[0], grid[1], grid[2], 8.0, order, 1e-6, -1.0])
    f, e = c.evaluate([h], s.pos, virial=True)
    assert c.get_option("pme_fft") == fft
    o = port.pme(s, grid, order, 8.0, 4.0, flags=6, virial=True)
    say("rms rel force", rms_rel(f, o["forces"]), "E", e[0], o["energies"][0])
    assert rms_rel(f, o["forces"]) < tol
    # the total is a small difference of large terms: scale the tolerance by them
    scale = sum(abs(o["parts"][k]) for k in ("reciprocal", "intra", "self"))
    assert abs(e[0] - o["energies"][0]) < tol * 0.2 * scale
    np.testing.assert_allclose(e[7:], o["energies"][7:], rtol=tol * 100, atol=tol * 100 * np.abs(o["energies"][7:]).max())


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_pme_gather_tail_matches_separate_terms(gpu_ctx_factory, precision):
    """With one PME reciprocal term the gather kernel also adds each atom's
    correction and bond/angle forces.  The same terms evaluated one handle
    at a time (separate kernels) and with the correction as its own handle
    give the same forces to the fixed-point granularity, and the same
    energies."""
    say(f"case: pme gather tail water648 {precision}")
    s = water.named("water648", seed=3, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    args = [24, 24, 24, 8.0, 4, 1e-6, -1.0]
    hp = c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", args)
    hr = c.force("Coulomb -algorithm PMEwald -reciprocal -interpolation BSpline", args)
    hc = c.force("Coulomb -algorithm PMEwald -correction -interpolation BSpline", args)
    hb = c.force("Bond", [])
    ha = c.force("Angle", [])
    f_all, e_all = c.evaluate([hp, hb, ha], s.pos)          # correction and bonded in the gather
    parts = [c.evaluate([h], s.pos) for h in (hr, hc, hb, ha)]   # each term on its own
    f_sep = sum(p[0] for p in parts)
    e_sep = sum(p[1] for p in parts)
    f_split, e_split = c.evaluate([hr, hc, hb, ha], s.pos)  # bonded in the gather, correction apart
    fmax = np.abs(f_all).max()
    say("max |df| separate", np.abs(f_all - f_sep).max(), "split", np.abs(f_all - f_split).max(), "of", fmax)
    assert np.abs(f_all - f_sep).max() <= 1e-11 * fmax
    assert np.abs(f_all - f_split).max() <= 1e-11 * fmax
    np.testing.assert_allclose(e_all, e_sep, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(e_all, e_split, rtol=1e-12, atol=1e-9)


def test_pme_full_alias(gpu_ctx_factory, port):
    say("case: CoulombPME alias water648 grid 24")
    s = water.named("water648", seed=4, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, "double")
    h = c.force("CoulombPME", [24, 24, 24, 8.0, 4, 1e-6, -1.0])
    f, e = c.evaluate([h], s.pos)
    o = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=7)
    assert rms_rel(f, o["forces"]) < 1e-10
    np.testing.assert_allclose(e[0], o["energies"][0], rtol=1e-10)


@pytest.mark.parametrize("fft", [1, 2])
def test_water98k_pme_mixed_accuracy(gpu_ctx_factory, port, fft):
    """98,304-atom water, LJ-C2 + Ewald-real rc 10 plus PME 96^3: RMS relative
    force error of the mixed-precision path against the FP64 oracle, with
    either arrangement of the y and x transforms."""
    say(f"case: water98k full force field mixed, fft option {fft}")
    s = water.named("water98k", seed=1, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 5.0, "mixed")
    c.set_option("pme_fft", fft)
    h1 = c.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
                 [9.0, 10.0, ALPHA_RC10, 10.0, 10.0])
    h2 = c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                 [96, 96, 96, 10.0, 4, 1e-6, -1.0])
    f, e = c.evaluate([h1, h2], s.pos)
    t = [po.term(po.POT_LJ, po.SWF_C2, 10.0, 9.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 10.0, alpha=ALPHA_RC10)]
    a = port.cutoff(s, t, 10.0, 5.0)
    b = port.pme(s, (96, 96, 96), 4, 10.0, 5.0, flags=6)
    fr = a["forces"] + b["forces"]
    er = a["energies"] + b["energies"]
    err = rms_rel(f, fr)
    # E_coulomb = real + reciprocal + intra + self: relative to the sum of magnitudes
    scale = abs(a["energies"][0]) + sum(abs(b["parts"][k]) for k in ("reciprocal", "intra", "self"))
    say("rms rel force", err, "E", e[:2], er[:2], "dE_coul/scale", (e[0] - er[0]) / scale,
        "rel E_lj", (e[1] - er[1]) / er[1])
    assert err < 5e-6
    assert abs(e[0] - er[0]) < 1e-7 * scale
    assert abs(e[1] - er[1]) < 5e-6 * abs(er[1])


def test_determinism(gpu_ctx_factory):
    say("case: determinism water648 mixed")
    s = water.named("water648", seed=5)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, "mixed")
    hs = gpu_forces_c1(c) + [c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                                     [24, 24, 24, 8.0, 4, 1e-6, -1.0])]
    f1, e1 = c.evaluate(hs, s.pos)
    f2, e2 = c.evaluate(hs, s.pos)
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(e1, e2)


def test_leapfrog_double(gpu_ctx_factory, port):
    say("case: leapfrog water648 double 5 steps")
    s = water.named("water648", seed=6, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, "double")
    hs = gpu_forces_c1(c)
    c.md_setup(hs, "leapfrog", dt=0.5)
    c.md_set_state(s.pos, s.vel)
    c.md_run(5)
    xg, vg, eg = c.md_get_state()
    # oracle: LeapfrogIntegrator::run(5)
    x = s.pos.copy()
    v = s.vel.copy()
    lj, co = c1_terms()

    def forces(xx):
        s2 = s.copy()
        s2.pos = xx
        return port.cutoff(s2, lj, 8.0, 4.0)["forces"] + port.cutoff(s2, co, 8.0, 4.0)["forces"]

    import ctypes as C
    lib = port.lib
    D = C.POINTER(C.c_double)

    def kd(fac, f):
        lib.orc_leapfrog_kickdrift(s.n, C.c_double(0.5), C.c_double(fac), s.mass.ctypes.data_as(D),
                                   f.ctypes.data_as(D), v.ctypes.data_as(D), x.ctypes.data_as(D))

    f = forces(x)
    kd(0.5, f)
    f = forces(x)
    for _ in range(4):
        kd(1.0, f)
        f = forces(x)
    lib.orc_half_kick(s.n, C.c_double(0.5), s.mass.ctypes.data_as(D), f.ctypes.data_as(D), v.ctypes.data_as(D))
    say("max |dx|", np.abs(xg - x).max(), "max |dv|", np.abs(vg - v).max())
    np.testing.assert_allclose(xg, x, atol=1e-9, rtol=0)
    np.testing.assert_allclose(vg, v, atol=1e-9 * np.abs(v).max(), rtol=0)


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_bonded(gpu_ctx_factory, port, precision):
    say(f"case: bond+angle water648 {precision}")
    s = water.named("water648", seed=7, jitter=0.4)
    rng = np.random.default_rng(7)
    s.pos = s.pos + rng.uniform(-0.1, 0.1, size=s.pos.shape)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, precision)
    f, e = c.evaluate([c.force("Bond", []), c.force("Angle", [])], s.pos, virial=True)
    o = port.bonded(s, 3, virial=True)
    say("rms rel force", rms_rel(f, o["forces"]), "E", e[2:4], o["energies"][2:4])
    assert rms_rel(f, o["forces"]) < 1e-11
    np.testing.assert_allclose(e[2:4], o["energies"][2:4], rtol=1e-11)
    np.testing.assert_allclose(e[7:], o["energies"][7:], rtol=1e-9, atol=1e-9 * np.abs(o["energies"][7:]).max())


def test_langevin_double(gpu_ctx_factory, port):
    """Three Langevin leapfrog steps (StandardIntegrator::run with the BBK half
    kick) against the oracle fed with the same counter-based deviates."""
    say("case: langevin water648 double 3 steps")
    from tests.cases import langevin_noise
    import ctypes as C
    s = water.named("water648", seed=9, jitter=0.4)
    c = setup_ctx(gpu_ctx_factory, s, 4.0, "double")
    hs = gpu_forces_c1(c)
    dt, temp, gamma, seed = 0.5, 300.0, 5.0, 77
    c.md_setup(hs, "langevin", dt=dt, temperature=temp, gamma=gamma, seed=seed)
    c.md_set_state(s.pos, s.vel)
    c.md_run(3)
    xg, vg, _ = c.md_get_state()
    lj, co = c1_terms()

    def forces(xx):
        s2 = s.copy()
        s2.pos = xx
        return port.cutoff(s2, lj, 8.0, 4.0)["forces"] + port.cutoff(s2, co, 8.0, 4.0)["forces"]

    D = C.POINTER(C.c_double)
    x = s.pos.copy()
    v = s.vel.copy()
    h = dt * water.INV_TIMEFACTOR

    def kick(f, counter):
        xi = np.ascontiguousarray(langevin_noise(seed, counter, 3 * s.n))
        port.lib.orc_langevin_half_kick(s.n, C.c_double(dt), C.c_double(temp), C.c_double(gamma),
                                        s.mass.ctypes.data_as(D), f.ctypes.data_as(D), xi.ctypes.data_as(D),
                                        v.ctypes.data_as(D))

    f = forces(x)
    counter = 0
    for _ in range(3):
        kick(f, counter); counter += 1
        x += h * v
        f = forces(x)
        kick(f, counter); counter += 1
    say("max |dx|", np.abs(xg - x).max(), "max |dv|", np.abs(vg - v).max())
    np.testing.assert_allclose(xg, x, atol=1e-10, rtol=0)
    np.testing.assert_allclose(vg, v, atol=1e-9 * np.abs(v).max(), rtol=0)


def test_langevin_temperature(gpu_ctx_factory):
    """Langevin dynamics holds the kinetic temperature of 648 atoms at the bath
    temperature (mean over 1500 steps after 1500 of equilibration, 2 %)."""
    say("case: langevin temperature water648 mixed")
    s = water.named("water648", seed=10, jitter=0.3)
    s.vel = s.vel * 0.5          # start cold: the thermostat must heat the system
    c = setup_ctx(gpu_ctx_factory, s, 4.0, "mixed")
    hs = gpu_forces_c1(c) + [c.force("Bond", []), c.force("Angle", [])]
    c.md_setup(hs, "langevin", dt=1.0, temperature=300.0, gamma=20.0, seed=5)
    c.md_set_state(s.pos, s.vel)
    c.md_run(1500)
    temps = []
    for _ in range(30):
        c.md_run(50)
        _, v, _ = c.md_get_state()
        ke = 0.5 * (s.mass[:, None] * v * v).sum()
        temps.append(2.0 * ke / (3 * s.n * water.BOLTZMANN))
    t = float(np.mean(temps))
    say("mean kinetic temperature", t)
    assert abs(t - 300.0) < 6.0


def test_water1m_pme_mixed_accuracy(gpu_ctx_factory, port):
    """999,999-atom water (215.4 A cube), LJ-C2 + Ewald-real rc 10 plus PME
    240^3: RMS relative force error of the mixed path against the FP64 oracle,
    and the in-cutoff pair count against the reference-order enumeration."""
    say("case: water1m full force field mixed")
    s = water.named("water1m", seed=1, jitter=0.3)
    c = setup_ctx(gpu_ctx_factory, s, 5.0, "mixed")
    h1 = c.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
                 [9.0, 10.0, ALPHA_RC10, 10.0, 10.0])
    h2 = c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                 [240, 240, 240, 10.0, 4, 1e-6, -1.0])
    f, e = c.evaluate([h1, h2], s.pos)
    t = [po.term(po.POT_LJ, po.SWF_C2, 10.0, 9.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 10.0, alpha=ALPHA_RC10)]
    a = port.cutoff_omp(s, t, 10.0, 5.0, threads=16)
    b = port.pme(s, (240, 240, 240), 4, 10.0, 5.0, flags=6)
    fr = a["forces"] + b["forces"]
    err = rms_rel(f, fr)
    say("rms rel force", err, "E", e[:2], (a["energies"] + b["energies"])[:2])
    assert err < 5e-6
    d = c.cell_diagnostics(s.pos, 10.0, 10.0)
    say("pairs within the cutoff", int(d["inrc"].sum()))
    assert 2.0e8 < int(d["inrc"].sum()) < 2.2e8


@pytest.mark.parametrize("integrator", ["leapfrog", "langevin"])
def test_graph_replay_matches_direct(gpu_ctx_factory, integrator):
    """MD steps replayed from captured HIP graphs give the same trajectory,
    bit for bit, as the same steps launched one kernel at a time."""
    say(f"case: graph replay vs direct {integrator} water648 mixed 29 steps")
    s = water.named("water648", seed=12, jitter=0.4)
    out = []
    for graph in (1, 0):
        c = setup_ctx(gpu_ctx_factory, s, 4.0, "mixed")
        c.set_option("graph", graph)
        hs = gpu_forces_c1(c) + [c.force("Bond", []), c.force("Angle", []),
                                 c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                                         [24, 24, 24, 8.0, 4, 1e-6, -1.0])]
        c.md_setup(hs, integrator, dt=1.0, temperature=300.0, gamma=5.0, seed=3)
        c.md_set_state(s.pos, s.vel)
        c.md_run(20)
        c.md_run(9)
        out.append(c.md_get_state())
        c.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mode", ["auto", "chains"])
def test_batch_matches_single_replicas(gpu_ctx_factory, mode):
    """Replicas of uneven size advanced as one batch (auto: one launch per
    kernel over all replicas; chains: one captured chain per replica) follow
    the same trajectories, bit for bit, as when each is run alone."""
    say(f"case: batch of 5 uneven replicas vs single runs, 23 steps, mode {mode}")
    from tests.cases import replica_context
    from protomol_amd import gpu
    systems = water.replica_boxes(5, seed=3, lo=100, hi=300)
    temps = water.replica_temperatures(5)
    ctxs = [replica_context(gpu_ctx_factory, s, t, seed=k + 1) for k, (s, t) in enumerate(zip(systems, temps))]
    b = gpu.Batch(ctxs)
    b.set_mode(mode)
    b.run(11)
    batched, why = b.info()
    say("batched launches:", batched, why)
    assert batched == (mode == "auto")
    b.run(12)
    got = [c.md_get_state() for c in ctxs]
    b.close()
    for k, (s, t) in enumerate(zip(systems, temps)):
        c = replica_context(gpu_ctx_factory, s, t, seed=k + 1)
        c.set_option("graph", 0)
        c.md_run(11)
        c.md_run(12)
        ref = c.md_get_state()
        for a, r in zip(got[k], ref):
            np.testing.assert_array_equal(a, r)


def test_batch_exchange_rule(gpu_ctx_factory, port):
    """Replica-exchange attempts: deltas from calcDelta (oracle), Metropolis
    acceptance on the seeded uniform, swapped temperatures and scaled velocities."""
    say("case: batch exchange rule, 6 replicas")
    from tests.cases import replica_context, rem_uniform
    from protomol_amd import gpu
    systems = water.replica_boxes(6, seed=4, lo=100, hi=160)
    temps = water.replica_temperatures(6, 300.0, 330.0)
    ctxs = [replica_context(gpu_ctx_factory, s, t, seed=k + 7) for k, (s, t) in enumerate(zip(systems, temps))]
    b = gpu.Batch(ctxs)
    seed = 99
    accepted_any = False
    for attempt in range(6):
        b.run(20)
        lad0, tl = b.ladder()
        states = [c.md_get_state() for c in ctxs]
        epot = [float(st[2][:7].sum()) for st in states]
        acc, dl = b.exchange(attempt, seed)
        lad1, _ = b.ladder()
        exp_lad = list(lad0)
        for k in range(attempt % 2, len(ctxs) - 1, 2):
            ra, rb = lad0[k], lad0[k + 1]
            d = port.rem_delta(tl[k], tl[k + 1], epot[ra], epot[rb])
            assert dl[k] == pytest.approx(d, rel=1e-12, abs=1e-12)
            w = 1.0 if d <= 0 else np.exp(-d)
            want = rem_uniform(seed, attempt, k) <= w
            assert bool(acc[k]) == want
            if want:
                accepted_any = True
                exp_lad[k], exp_lad[k + 1] = exp_lad[k + 1], exp_lad[k]
                va = ctxs[ra].md_get_state()[1]
                np.testing.assert_allclose(va, states[ra][1] * np.sqrt(tl[k + 1] / tl[k]), rtol=1e-14)
        assert list(lad1) == exp_lad
    b.close()
    say("accepted at least one swap:", accepted_any)
    assert accepted_any
