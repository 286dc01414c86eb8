"""CPU tests that pin the oracle (oracle/pmoracle.c).

Three independent anchors:
  * the golden fixtures written by tests/golden/make_golden.py from the
    compiled reference (always run);
  * the compiled reference itself, call for call (run where oracle/_ref is
    built, i.e. where the reference tree exists);
  * closed forms and brute force that do not depend on the reference.
"""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from protomol_amd import water
from tests.cases import exclusion_mix

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "water648_reference.npz")
HGOLD = os.path.join(HERE, "golden", "water81_hessian_reference.npz")
ALPHA8 = 0.394670228244391

CASES = {
    "lj_c2": [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)],
    "coul_c2": [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)],
    "coul_c1": [po.term(po.POT_COULOMB, po.SWF_C1, 8.0)],
    "lj_ewr": [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA8)],
}


def system_from(g):
    keys = ["pos", "q", "type", "molecule", "ex_a", "ex_b", "ex_cls", "ljA", "ljB", "ljA14", "ljB14", "box",
            "origin", "mass", "bond_a", "bond_b", "bond_k", "bond_r0", "angle_a", "angle_b", "angle_c",
            "angle_k", "angle_t0", "angle_kub", "angle_rub"]
    d = {k: np.array(g[k]) for k in keys}
    return water.WaterSystem(vel=np.zeros_like(d["pos"]), **d)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


# ------------------------------------------------------------- golden vectors
def test_golden_cells(port, gold):
    s = system_from(gold)
    c = port.cells(s, 4.0)
    assert tuple(c["dims"]) == tuple(gold["dims"])
    np.testing.assert_array_equal(c["cell"], gold["cell"])
    np.testing.assert_array_equal(c["head"], gold["head"])
    np.testing.assert_array_equal(c["next"], gold["next"])


@pytest.mark.parametrize("name", list(CASES))
def test_golden_cutoff(port, gold, name):
    s = system_from(gold)
    r = port.cutoff(s, CASES[name], 8.0, 4.0, virial=True, counts=True)
    np.testing.assert_array_equal(r["cand"], gold[f"{name}_cand"])
    np.testing.assert_array_equal(r["inrc"], gold[f"{name}_inrc"])
    # the restatement follows the reference's loop order: agreement to rounding
    np.testing.assert_allclose(r["forces"], gold[f"{name}_forces"], rtol=0, atol=1e-12 * np.abs(gold[f"{name}_forces"]).max())
    np.testing.assert_allclose(r["energies"], gold[f"{name}_energies"], rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("flags", [2, 4, 7])
def test_golden_pme(port, gold, flags):
    s = system_from(gold)
    r = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=flags, virial=True)
    f = gold[f"pme{flags}_forces"]
    np.testing.assert_allclose(r["forces"], f, rtol=0, atol=1e-11 * np.abs(f).max())
    np.testing.assert_allclose(r["energies"], gold[f"pme{flags}_energies"], rtol=1e-12, atol=1e-9)


def test_golden_bonded(port, gold):
    s = system_from(gold)
    r = port.bonded(s, 3, virial=True)
    np.testing.assert_allclose(r["forces"], gold["bonded_forces"], rtol=0, atol=1e-12 * np.abs(gold["bonded_forces"]).max())
    np.testing.assert_allclose(r["energies"], gold["bonded_energies"], rtol=1e-13, atol=1e-12)


def test_golden_hessian(port):
    g = np.load(HGOLD)
    s = system_from(g)
    terms = {"lj_c2": [po.term(po.POT_LJ, po.SWF_C2, 4.5, 3.5)],
             "coul_c1": [po.term(po.POT_COULOMB, po.SWF_C1, 4.5)],
             "coul_c2": [po.term(po.POT_COULOMB, po.SWF_C2, 4.5, 3.5)]}
    for name, t in terms.items():
        hv = port.hessian_vec(s, t, g["v"])
        ref = g[f"{name}_hv"]
        np.testing.assert_allclose(hv, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


# ----------------------------------------------------- the compiled reference
@pytest.mark.parametrize("seed", [21, 22])
def test_reference_cutoff_bitexact(port, ref, seed):
    s = water.named("water648", seed=seed, jitter=0.3)
    for terms in CASES.values():
        a = port.cutoff(s, terms, 8.0, 4.0, virial=True, counts=True)
        b = ref.cutoff(s, terms, 8.0, 4.0, virial=True, counts=True)
        np.testing.assert_array_equal(a["forces"], b["forces"])
        np.testing.assert_array_equal(a["energies"], b["energies"])
        np.testing.assert_array_equal(a["cand"], b["cand"])
        np.testing.assert_array_equal(a["inrc"], b["inrc"])


def test_reference_small_box_semantics(port, ref):
    """Boxes below twice the cutoff: cell pairs deduplicated through basis
    cells (n < 3) and minimal-image pairs, as in the reference enumerator."""
    s = water.water_box(27, 9.3, seed=4, jitter=0.2)
    for cellsize in (4.0, 3.0, 9.3):
        for terms in CASES.values():
            a = port.cutoff(s, terms, 8.0, cellsize, counts=True)
            b = ref.cutoff(s, terms, 8.0, cellsize, counts=True)
            np.testing.assert_array_equal(a["forces"], b["forces"])
            np.testing.assert_array_equal(a["cand"], b["cand"])


def test_reference_exclusion_rules(port, ref):
    """MODIFIED / NONE overrides, duplicates and cross-molecule pairs."""
    s = exclusion_mix(water.named("water648", seed=24, jitter=0.3))
    for name in ("lj_c2", "coul_c1", "lj_ewr"):
        a = port.cutoff(s, CASES[name], 8.0, 4.0, virial=True, counts=True)
        b = ref.cutoff(s, CASES[name], 8.0, 4.0, virial=True, counts=True)
        np.testing.assert_array_equal(a["forces"], b["forces"])
        np.testing.assert_array_equal(a["energies"], b["energies"])
        np.testing.assert_array_equal(a["inrc"], b["inrc"])
    a = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=4, virial=True)
    b = ref.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=4, virial=True)
    np.testing.assert_allclose(a["forces"], b["forces"], rtol=0, atol=1e-12 * np.abs(b["forces"]).max())
    np.testing.assert_allclose(a["energies"], b["energies"], rtol=1e-12, atol=1e-9)


def test_reference_pme(port, ref):
    s = water.named("water648", seed=23, jitter=0.3)
    for grid, order in (((20, 24, 30), 4), ((32, 32, 32), 6), ((25, 25, 25), 8), ((24, 24, 24), 2),
                        ((24, 24, 24), 10), ((24, 24, 24), 12)):
        for flags in (1, 2, 4, 7):
            a = port.pme(s, grid, order, 8.0, 4.0, flags=flags, virial=True)
            b = ref.pme(s, grid, order, 8.0, 4.0, flags=flags, virial=True)
            np.testing.assert_allclose(a["forces"], b["forces"], rtol=0, atol=1e-11 * max(1.0, np.abs(b["forces"]).max()))
            np.testing.assert_allclose(a["energies"], b["energies"], rtol=1e-12, atol=1e-9)


def test_reference_pme_real_lattice_images(port, ref):
    """The PME real term in boxes no wider than twice the cutoff, where pairs
    meet lattice images beyond the minimal one (and atoms their own images):
    the 27-water 9.3 A box and a 100-water 14.4 A box at rc 8, every exclusion
    class, energy and virial, against the compiled reference."""
    boxes = [exclusion_mix(water.water_box(27, 9.3, seed=4, jitter=0.2)),
             exclusion_mix(water.water_box(100, 14.4, seed=6, jitter=0.3))]
    for s in boxes:
        for cellsize in (4.0, 9.3):
            a = port.pme(s, (16, 16, 16), 4, 8.0, cellsize, flags=1, virial=True)
            b = ref.pme(s, (16, 16, 16), 4, 8.0, cellsize, flags=1, virial=True)
            np.testing.assert_allclose(a["forces"], b["forces"], rtol=0, atol=1e-11 * np.abs(b["forces"]).max())
            np.testing.assert_allclose(a["energies"], b["energies"], rtol=1e-11, atol=1e-9)


def test_reference_pme_charged_system(port, ref):
    """Net-charged box (one Na+ added to 216 waters): the charged-system energy
    -pi/(2 V alpha^2) (sum q)^2 and its virial diagonal
    (NonbondedPMEwaldSystemForce.h:300-305, 618-628) against the compiled
    reference, and the term itself in closed form: the ion has no exclusion
    pairs, so the correction-only energy of the charged box exceeds the
    neutral box's by the ion's self energy -alpha/sqrt(pi) q^2 plus the term,
    and its virial diagonal by the term alone."""
    base = water.named("water648", seed=25, jitter=0.3)
    s = water.with_ion(base, 1.0, (0.3, -0.2, 0.1))
    for flags in (4, 6, 7):
        a = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=flags, virial=True)
        b = ref.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=flags, virial=True)
        np.testing.assert_allclose(a["forces"], b["forces"], rtol=0, atol=1e-11 * np.abs(b["forces"]).max())
        np.testing.assert_allclose(a["energies"], b["energies"], rtol=1e-12, atol=1e-9)
    n = port.pme(base, (24, 24, 24), 4, 8.0, 4.0, flags=4, virial=True)
    c = port.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=4, virial=True)
    alpha = port.pme_alpha(8.0, 1e-6)
    vol = float(np.prod(s.box))
    qsum = float(s.q.sum())
    qion = float(s.q[-1])
    term = -math.pi / (2.0 * vol * alpha * alpha) * qsum * qsum
    assert abs(float(base.q.sum()) * 0.00268283) < 1e-5 < abs(qsum * 0.00268283)
    np.testing.assert_allclose(c["energies"][0] - n["energies"][0], -alpha / math.sqrt(math.pi) * qion * qion + term,
                               rtol=1e-10)
    for k in (7, 11, 15):
        np.testing.assert_allclose(c["energies"][k] - n["energies"][k], term, rtol=1e-9)
    for k in (8, 9, 10, 12, 13, 14):
        assert c["energies"][k] == n["energies"][k]


def test_reference_bspline(port, ref):
    for order in (2, 4, 6, 8, 10):
        for w in (0.0, 0.13, 0.5, 0.999):
            a = port.bspline(order, w)
            b = ref.bspline(order, w)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])


# ------------------------------------------------------------- closed forms
def two_atoms(r, qa=1.0, qb=-1.0, ta=0, tb=0, L=40.0):
    s = water.water_box(1, L, seed=0)
    pos = np.array([[0.1, 0.2, 0.3], [0.1 + r, 0.2, 0.3]])
    A, B = s.ljA, s.ljB
    return water.WaterSystem(pos=pos, vel=np.zeros((2, 3)), q=np.array([qa, qb]), type=np.array([ta, tb], np.int32),
                             mass=np.ones(2), molecule=np.array([0, 1], np.int32), ex_a=np.zeros(0, np.int32),
                             ex_b=np.zeros(0, np.int32), ex_cls=np.zeros(0, np.int32), ljA=A, ljB=B, ljA14=A,
                             ljB14=B, box=np.array([L, L, L]), origin=np.zeros(3))


def c2(r2, s, c):
    if r2 > c * c:
        return 0.0, 0.0
    if r2 < s * s:
        return 1.0, 0.0
    d = (c * c - s * s) ** 3
    v = (c * c - r2) ** 2 * (c * c + 2 * r2 - 3 * s * s) / d
    dv = 12 * (c * c - r2) * (s * s - r2) / d   # 2 dS/d(r^2)
    return v, dv


@pytest.mark.parametrize("r", [3.0, 6.5, 7.3, 7.9, 8.1])
def test_closed_form_lj_c2(port, r):
    s = two_atoms(r)
    A, B = s.ljA[0, 0], s.ljB[0, 0]
    out = port.cutoff(s, [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0)
    e = A / r ** 12 - B / r ** 6
    dedr = -12 * A / r ** 13 + 6 * B / r ** 7
    sv, dv = c2(r * r, 7.0, 8.0)
    E = e * sv
    # F along +x on atom 1 = -dE/dr, dS/dr = 2 r dS/d(r^2) = r * dv
    dSdr = r * dv
    F1 = -(dedr * sv + e * dSdr)
    assert out["energies"][po.E_LJ] == pytest.approx(E, rel=1e-12, abs=1e-15)
    assert out["forces"][1, 0] == pytest.approx(F1, rel=1e-10, abs=1e-14)
    assert out["forces"][0, 0] == pytest.approx(-F1, rel=1e-10, abs=1e-14)


@pytest.mark.parametrize("r", [2.0, 5.0, 7.9])
def test_closed_form_coulomb_c1(port, r):
    s = two_atoms(r, qa=10.0, qb=-7.0)
    out = port.cutoff(s, [po.term(po.POT_COULOMB, po.SWF_C1, 8.0)], 8.0, 4.0)
    c = 8.0
    S = 1 - 1.5 * r / c + 0.5 * r ** 3 / c ** 3
    e = -70.0 / r
    assert out["energies"][po.E_COULOMB] == pytest.approx(e * S, rel=1e-12)
    dSdr = -1.5 / c + 1.5 * r * r / c ** 3
    F1 = -((70.0 / r ** 2) * S + e * dSdr)
    assert out["forces"][1, 0] == pytest.approx(F1, rel=1e-10)


@pytest.mark.parametrize("r", [1.0, 4.0, 7.5])
def test_closed_form_ewald_real(port, r):
    """CoulombEwaldReal uses the A&S 7.1.26 erfc: within 1.5e-7 of qq erfc(ar)/r."""
    s = two_atoms(r, qa=10.0, qb=-7.0)
    out = port.cutoff(s, [po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA8)], 8.0, 4.0)
    exact = -70.0 * math.erfc(ALPHA8 * r) / r
    assert abs(out["energies"][po.E_COULOMB] - exact) < 70.0 * 1.5e-7 / r * 1.01


def brute_pairs_energy(s, cut, fn):
    n = s.n
    L = s.box
    e = 0.0
    for i in range(n):
        d = s.pos[i + 1:] - s.pos[i]
        d -= L * np.rint(d / L)
        r2 = (d * d).sum(1)
        js = np.nonzero(r2 <= cut * cut)[0] + i + 1
        for j, rr in zip(js, r2[js - i - 1]):
            if s.molecule[i] == s.molecule[j]:
                continue   # water: every intramolecular pair is excluded (FULL)
            e += fn(i, j, rr)
    return e


def test_brute_force_pair_set(port):
    """Cell list + enumerator == O(N^2) minimal-image enumeration (L >= 2 rc)."""
    s = water.named("water648", seed=31, jitter=0.3)
    A = s.ljA

    def lj(i, j, r2):
        a, b = A[s.type[i], s.type[j]], s.ljB[s.type[i], s.type[j]]
        v, _ = c2(r2, 7.0, 8.0)
        return (a / r2 ** 6 - b / r2 ** 3) * v

    def coul(i, j, r2):
        v, _ = c2(r2, 7.0, 8.0)
        return s.q[i] * s.q[j] / math.sqrt(r2) * v

    out_lj = port.cutoff(s, [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0, counts=True)
    out_c = port.cutoff(s, [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0)
    assert out_lj["energies"][po.E_LJ] == pytest.approx(brute_pairs_energy(s, 8.0, lj), rel=1e-11)
    assert out_c["energies"][po.E_COULOMB] == pytest.approx(brute_pairs_energy(s, 8.0, coul), rel=1e-11)
    # number of in-cutoff, non-excluded pairs
    n = s.n
    cnt = 0
    for i in range(n):
        d = s.pos[i + 1:] - s.pos[i]
        d -= s.box * np.rint(d / s.box)
        ok = (d * d).sum(1) <= 64.0
        ok &= s.molecule[i + 1:] != s.molecule[i]
        cnt += int(ok.sum())
    assert int(out_lj["inrc"].sum()) == cnt


def test_bspline_identities(port):
    for order in (4, 6, 8):
        for w in np.linspace(0.0, 0.999, 7):
            th, dth = port.bspline(order, w)
            assert th.sum() == pytest.approx(1.0, abs=1e-14)      # partition of unity
            assert dth.sum() == pytest.approx(0.0, abs=1e-14)
    # cubic (order 4) closed form: weights of grid points i0-1 .. i0+2
    w = 0.3
    th, dth = port.bspline(4, w)
    np.testing.assert_allclose(th, [(1 - w) ** 3 / 6, (3 * w ** 3 - 6 * w ** 2 + 4) / 6,
                                    (-3 * w ** 3 + 3 * w ** 2 + 3 * w + 1) / 6, w ** 3 / 6], rtol=1e-14)
    np.testing.assert_allclose(dth, [-(1 - w) ** 2 / 2, (3 * w ** 2 - 4 * w) / 2,
                                     (-3 * w ** 2 + 2 * w + 1) / 2, w ** 2 / 2], rtol=1e-13)


@pytest.mark.parametrize("shape", [(12, 12, 12), (24, 20, 15), (7, 9, 10), (16, 30, 25)])
def test_fft_against_numpy(port, shape):
    rng = np.random.default_rng(0)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = port.fft3d(a, -1)
    np.testing.assert_allclose(f, np.fft.fftn(a), rtol=0, atol=1e-10 * np.abs(a).sum())
    b = port.fft3d(a, +1)
    np.testing.assert_allclose(b, np.fft.ifftn(a) * a.size, rtol=0, atol=1e-10 * np.abs(a).sum())


def test_alpha_bisection(port):
    for rc, acc, expect in ((10.0, 1e-6, 0.3123413274), (9.0, 1e-6, 0.3488316179), (8.0, 1e-6, 0.3946702282)):
        a = port.pme_alpha(rc, acc)
        assert a == pytest.approx(expect, abs=1e-9)
        assert math.erfc(a * rc) / rc == pytest.approx(acc, rel=1e-9)


def test_pme_against_direct_ewald(port):
    """Reciprocal PME energy and forces converge to the direct Ewald sum."""
    s = water.water_box(27, 10.0, seed=5, jitter=0.3)
    alpha = 0.35
    L = s.box[0]
    V = L ** 3
    kmax = 12
    ks = np.arange(-kmax, kmax + 1)
    kx, ky, kz = np.meshgrid(ks, ks, ks, indexing="ij")
    m = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], 1) / L
    m = m[(m != 0).any(1)]
    m2 = (m * m).sum(1)
    phase = 2 * np.pi * s.pos @ m.T
    S = (s.q[:, None] * np.exp(1j * phase)).sum(0)
    w = np.exp(-np.pi ** 2 * m2 / alpha ** 2) / (np.pi * V * m2)
    e_dir = 0.5 * (w * np.abs(S) ** 2).sum()
    # F_i = -dE/dr_i
    f_dir = np.zeros((s.n, 3))
    for c in range(3):
        f_dir[:, c] = (s.q[:, None] * (w * 2 * np.pi * m[:, c])[None, :] *
                       np.imag(np.exp(1j * phase) * np.conj(S)[None, :])).sum(1)
    r = port.pme(s, (48, 48, 48), 8, 5.0, 3.0, alpha=alpha, flags=2)
    assert r["parts"]["reciprocal"] == pytest.approx(e_dir, rel=1e-7)
    err = np.sqrt(((r["forces"] - f_dir) ** 2).sum() / (f_dir ** 2).sum())
    assert err < 1e-5


def test_hessian_against_finite_differences(port):
    s = water.water_box(27, 9.4, seed=8, jitter=0.3)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((s.n, 3))
    v /= np.linalg.norm(v)
    for terms in ([po.term(po.POT_LJ, po.SWF_C2, 4.5, 3.5)], [po.term(po.POT_COULOMB, po.SWF_C1, 4.5)]):
        hv = port.hessian_vec(s, terms, v)
        h = 1e-4
        sp, sm = s.copy(), s.copy()
        sp.pos = s.pos + h * v
        sm.pos = s.pos - h * v
        fp = port.cutoff(sp, terms, 4.5, 2.0)["forces"]
        fm = port.cutoff(sm, terms, 4.5, 2.0)["forces"]
        fd = -(fp - fm) / (2 * h)
        # pairs crossing the cutoff during the displacement spoil the FD at a
        # handful of atoms; the RMS is what pins the analytic Hessian
        assert np.sqrt(((hv - fd) ** 2).sum() / (fd ** 2).sum()) < 1e-5


def test_leapfrog_constant_force(port):
    """One particle under a constant force: leapfrog reproduces x(t) = x0 + v0 t + F t^2 / 2m."""
    import ctypes as C
    D = C.POINTER(C.c_double)
    x = np.array([[0.0, 0.0, 0.0]])
    v = np.array([[0.01, 0.0, 0.0]])
    f = np.array([[2.0, -1.0, 0.5]])
    m = np.array([3.0])
    dt = 2.0
    h = dt * water.INV_TIMEFACTOR
    port.lib.orc_leapfrog_kickdrift(1, C.c_double(dt), C.c_double(0.5), m.ctypes.data_as(D), f.ctypes.data_as(D),
                                    v.ctypes.data_as(D), x.ctypes.data_as(D))
    for _ in range(9):
        port.lib.orc_leapfrog_kickdrift(1, C.c_double(dt), C.c_double(1.0), m.ctypes.data_as(D), f.ctypes.data_as(D),
                                        v.ctypes.data_as(D), x.ctypes.data_as(D))
    port.lib.orc_half_kick(1, C.c_double(dt), m.ctypes.data_as(D), f.ctypes.data_as(D), v.ctypes.data_as(D))
    t = 10 * h
    np.testing.assert_allclose(x[0], 0.01 * np.array([1, 0, 0]) * t + f[0] / m[0] * t * t / 2, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(v[0], np.array([0.01, 0, 0]) + f[0] / m[0] * t, rtol=1e-12)


def test_rem_delta_boltzmann_ratio(port):
    """exp(-delta) of a swap is the ratio of the two replicas' joint Boltzmann
    weights after and before the configurations trade temperatures,
    exp(-E_j/kT_i - E_i/kT_j) / exp(-E_i/kT_i - E_j/kT_j), with ProtoMol's
    Boltzmann constant (base/pmconstants.cpp:74); rem.cpp:941-946 accepts
    when delta <= 0 or the uniform is <= exp(-delta).  A sign slip (E_i - E_j)
    or a swapped temperature pair inverts the ratio and fails here."""
    kb = 0.001987191
    cases = ((300.0, 310.0, -100.0, -95.0), (450.0, 300.0, 5.0, -7.5), (300.0, 301.5, -2.5e4, -2.49e4),
             (350.0, 300.0, -12.0, -12.0))
    for ti, tj, ei, ej in cases:
        before = math.exp(-ei / (kb * ti)) * math.exp(-ej / (kb * tj)) if abs(ei) < 1e3 else None
        d = port.rem_delta(ti, tj, ei, ej)
        if before is not None:
            after = math.exp(-ej / (kb * ti)) * math.exp(-ei / (kb * tj))
            assert math.exp(-d) == pytest.approx(after / before, rel=1e-12)
        else:
            # large energies: the same ratio taken in logarithms
            log_ratio = (-ej / (kb * ti) - ei / (kb * tj)) - (-ei / (kb * ti) - ej / (kb * tj))
            assert -d == pytest.approx(log_ratio, rel=1e-9)
    # a colder replica holding the higher energy always swaps (delta < 0)
    assert port.rem_delta(300.0, 330.0, -90.0, -100.0) < 0.0
    assert port.rem_delta(300.0, 330.0, -100.0, -90.0) > 0.0
    assert port.rem_delta(300.0, 330.0, -50.0, -50.0) == 0.0


def test_rem_driver_rounds_and_credits(port):
    """The driver restatement (rem.cpp:646-720) on hand-made attempts: rounds
    before EqSwitches attempt nothing, round i tests the pairs (k, k+1) with k
    of i's parity only (getNextAddress :949-957), a pair whose colder slot
    holds the higher energy always swaps (delta < 0), one with a huge positive
    delta never does, and each switch is credited to the hotter slot k + 1
    (:814-815), the rates dividing by NumSwitches - EqSwitches."""
    from oracle import pyoracle
    temps = [300.0, 310.0, 320.0, 330.0]
    swap, keep = (-100.0, -200.0), (-200.0, 1e6)
    attempts = {2: [*swap, *swap], 3: [-1.0, *keep, -1.0], 4: [*keep, *swap]}
    flags, rates = pyoracle.rem_driver(port, temps, attempts, 11, 5, 2)
    assert sorted(flags) == [2, 3, 4]
    assert flags[2] == {0: 1, 2: 1}
    assert flags[3] == {1: 0}
    assert flags[4] == {0: 0, 2: 1}
    assert rates == [0.0, 1 / 3, 0.0, 2 / 3]
    # the uniform is in [0, 1) and differs between attempts and pairs
    us = {pyoracle.rem_uniform(11, i, k) for i in range(20) for k in range(8)}
    assert len(us) == 160 and min(us) >= 0.0 and max(us) < 1.0


def _langevin_kick(port, n, dt, temp, gamma, mass, f, xi, v):
    import ctypes as C
    D = C.POINTER(C.c_double)
    port.lib.orc_langevin_half_kick(n, C.c_double(dt), C.c_double(temp), C.c_double(gamma), mass.ctypes.data_as(D),
                                    f.ctypes.data_as(D), xi.ctypes.data_as(D), v.ctypes.data_as(D))


def test_langevin_half_kick_deterministic_part(port):
    """Without noise (T = 0) a half kick integrates dv/dt = F/m - gamma v over
    half a step exactly: v = v0 e^(-gamma t) + F/(m gamma) (1 - e^(-gamma t)),
    t = dt/2 in AKMA time (dt_fs * INV_TIMEFACTOR) and gamma converted from
    1/ps to 1/AKMA-time (LangevinLeapfrogIntegrator.cpp:36, :58-79).  The
    solution is checked against a fine RK4 integration of the ODE, and the
    gamma -> 0 limit against the plain half kick."""
    n = 4
    rng = np.random.default_rng(3)
    mass = np.array([1.008, 15.9994, 12.011, 35.45])
    f = rng.normal(0.0, 20.0, (n, 3))
    v0 = rng.normal(0.0, 0.02, (n, 3))
    xi = np.zeros((n, 3))
    for dt, gamma in ((1.0, 5.0), (2.0, 91.0), (0.5, 1000.0)):
        v = v0.copy()
        _langevin_kick(port, n, dt, 0.0, gamma, mass, f, xi, v)
        t = 0.5 * dt * water.INV_TIMEFACTOR
        g = gamma / 1000.0 / water.INV_TIMEFACTOR      # 1/ps -> 1/fs -> 1/AKMA time
        a = f / mass[:, None]
        # RK4 of dv/dt = a - g v with 2000 steps
        w = v0.copy()
        hstep = t / 2000
        for _ in range(2000):
            k1 = a - g * w
            k2 = a - g * (w + 0.5 * hstep * k1)
            k3 = a - g * (w + 0.5 * hstep * k2)
            k4 = a - g * (w + hstep * k3)
            w = w + hstep / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        np.testing.assert_allclose(v, w, rtol=1e-9, atol=1e-12)
    # gamma -> 0: the Langevin half kick becomes the leapfrog half kick
    import ctypes as C
    D = C.POINTER(C.c_double)
    v = v0.copy()
    # (gamma t = 5e-7: the friction's relative effect is gamma t / 2 and the
    # rounding of 1 - e^(-gamma t) about 1e-16 / (gamma t))
    _langevin_kick(port, n, 1.0, 0.0, 1e-3, mass, f, xi, v)
    w = v0.copy()
    port.lib.orc_half_kick(n, C.c_double(1.0), mass.ctypes.data_as(D), f.ctypes.data_as(D), w.ctypes.data_as(D))
    np.testing.assert_allclose(v, w, rtol=1e-6)


def test_langevin_half_kick_noise_is_ornstein_uhlenbeck(port):
    """With F = 0 the velocity is an Ornstein-Uhlenbeck process, whose exact
    transition over t = dt/2 is v1 = v0 e^(-gamma t) + N(0, kT/m (1 - e^(-2 gamma t))):
    from v0 ~ N(0, kT/m) the regression slope of v1 on v0 is e^(-gamma t), the
    residual variance kT/m (1 - e^(-2 gamma t)) and the variance of v1 stays
    kT/m (the Maxwell distribution is stationary).  A wrong friction or noise
    amplitude (fdt, vdt, ndt, the 2 k T gamma constant) moves one of the three."""
    n = 200_000
    rng = np.random.default_rng(11)
    temp, gamma, dt = 310.0, 50.0, 2.0
    mass = np.full(n, 15.9994)
    kt_m = water.BOLTZMANN * temp / mass[0]
    v0 = rng.standard_normal((n, 3)) * math.sqrt(kt_m)
    xi = rng.standard_normal((n, 3))
    v = v0.copy()
    _langevin_kick(port, n, dt, temp, gamma, mass, np.zeros((n, 3)), xi, v)
    t = 0.5 * dt * water.INV_TIMEFACTOR
    g = gamma / 1000.0 / water.INV_TIMEFACTOR
    x, y = v0.reshape(-1), v.reshape(-1)
    slope = (x * y).sum() / (x * x).sum()
    resid = y - slope * x
    m = x.size
    assert slope == pytest.approx(math.exp(-g * t), abs=4.0 / math.sqrt(m))
    assert resid.var() == pytest.approx(kt_m * (1 - math.exp(-2 * g * t)), rel=6.0 * math.sqrt(2.0 / m))
    assert y.var() == pytest.approx(kt_m, rel=6.0 * math.sqrt(2.0 / m))
