"""GPU adjoint: Hessian-vector products and parameter gradients through the C
ABI, against the oracle's reference-pinned dense pair Hessian, central finite
differences of the GPU's own forces and energies, and closed properties."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from protomol_amd import water
from tests.cases import ALPHA_RC8, c1_terms, gpu_forces_c1, rms_rel

pytestmark = pytest.mark.gpu

ALPHA_RC9 = 0.3488316179034193


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def unit(rng, shape):
    v = rng.standard_normal(shape)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("precision,tol", [("double", 1e-9), ("mixed", 2e-5)])
def test_hvp_against_reference_hessian(gpu_ctx_factory, port, precision, tol):
    """LJ + Coulomb with C2 switching (the reference's Hessian covers these):
    H v of the GPU against the oracle's dense pair Hessian, which reproduces
    integrators/Hessian.cpp on the golden fixture."""
    say(f"case: hvp water648 LJ+Coulomb C2 {precision}")
    s = water.named("water648", seed=21, jitter=0.4)
    c = gpu_ctx_factory(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.0)
    hs = gpu_forces_c1(c)
    v = unit(np.random.default_rng(1), (s.n, 3))
    hv = c.hessian_vec(hs, s.pos, v)
    lj, co = c1_terms()
    ref = port.hessian_vec(s, lj, v) + port.hessian_vec(s, co, v)
    say("rms rel", rms_rel(hv, ref))
    assert rms_rel(hv, ref) < tol


def away_from_cutoff(s, v, rc, margin=2e-4):
    """v with the rows of atoms in a pair whose minimum-image distance lies
    within `margin` of rc set to zero.  Switching functions with a kink or a
    jump at rc (C1: second derivative; Cutoff: the force itself) make a central
    difference across rc measure the jump instead of the derivative; with these
    rows fixed, no pair crosses rc over a step of h |v| << margin."""
    from scipy.spatial import cKDTree
    L = np.asarray(s.box, float).reshape(-1)
    L = np.array([L[0], L[4], L[8]]) if L.size == 9 else L[:3]
    x = np.mod(np.asarray(s.pos) - np.asarray(s.origin).reshape(-1)[:3], L)
    tree = cKDTree(x, boxsize=L)
    outer = tree.query_pairs(rc + margin, output_type="ndarray")
    d = x[outer[:, 0]] - x[outer[:, 1]]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(1))
    near = outer[np.abs(r - rc) < margin]
    v = v.copy()
    v[np.unique(near.reshape(-1))] = 0.0
    return v / np.linalg.norm(v), len(near)


@pytest.mark.parametrize("fid,vals", [
    ("LennardJones Coulomb -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction C1", [7.0, 8.0, 8.0, 8.0]),
    ("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
     [7.0, 8.0, ALPHA_RC8, 8.0, 8.0]),
])
def test_hvp_cutoff_finite_differences(gpu_ctx_factory, fid, vals):
    """H v against central differences of the GPU's FP64 forces, h = 1e-4 A,
    on the 8,100-atom box (its 4 A cells give no wrapped duplicate stencil
    entries) with v zero on the atoms of pairs within 2e-4 A of rc."""
    say(f"case: hvp vs FD {fid.split(' -')[0]}")
    s = water.named("water8k", seed=22, jitter=0.4)
    c = gpu_ctx_factory("double")
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.0)
    hs = [c.force(fid, vals)]
    v, nnear = away_from_cutoff(s, unit(np.random.default_rng(2), (s.n, 3)), 8.0)
    say("pairs held off rc", nnear)
    hv = c.hessian_vec(hs, s.pos, v)
    h = 1e-4
    fp, _ = c.evaluate(hs, s.pos + h * v)
    fm, _ = c.evaluate(hs, s.pos - h * v)
    fd = -(fp - fm) / (2 * h)
    say("rms rel vs FD", rms_rel(hv, fd))
    assert rms_rel(hv, fd) < 2e-6


@pytest.mark.parametrize("precision,tol", [("double", 2e-6), ("mixed", 5e-5)])
def test_hvp_pme_finite_differences(gpu_ctx_factory, precision, tol):
    """Nonbonded + PME (reciprocal and correction): H v against central
    differences of the FP64 forces, on the 8,100-atom perturbed box (rc 9,
    PME 48^3)."""
    say(f"case: hvp water8k perturbed, PME 48^3, {precision}")
    s, sigma, eps = water.perturbed_box("water8k", seed=1)
    cd = gpu_ctx_factory("double")
    c = cd if precision == "double" else gpu_ctx_factory(precision)
    for ctx in {id(cd): cd, id(c): c}.values():
        ctx.set_topology(s)
        ctx.set_boundary(s.box, s.origin, 4.5)
        ctx.set_lj_per_atom(sigma, eps)

    def forces(ctx):
        return [ctx.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
                          [8.0, 9.0, ALPHA_RC9, 9.0, 9.0]),
                ctx.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                          [48, 48, 48, 9.0, 4, 1e-6, -1.0])]
    hd = forces(cd)
    hs = hd if c is cd else forces(c)
    v = unit(np.random.default_rng(3), (s.n, 3))
    hv = c.hessian_vec(hs, s.pos, v)
    h = 1e-4
    fp, _ = cd.evaluate(hd, s.pos + h * v)
    fm, _ = cd.evaluate(hd, s.pos - h * v)
    fd = -(fp - fm) / (2 * h)
    say("rms rel vs FD", rms_rel(hv, fd))
    assert rms_rel(hv, fd) < tol
    # symmetry of H: u . (H v) = v . (H u)
    u = unit(np.random.default_rng(4), (s.n, 3))
    hu = c.hessian_vec(hs, s.pos, u)
    a, b = float((u * hv).sum()), float((v * hu).sum())
    say("u.Hv", a, "v.Hu", b)
    assert abs(a - b) <= 1e-6 * max(abs(a), 1.0) if precision == "double" else abs(a - b) <= 1e-4 * max(abs(a), 1.0)


@pytest.mark.parametrize("precision,tol,tolq", [("double", 1e-6, 1e-9), ("mixed", 2e-4, 2e-5)])
def test_param_gradients_finite_differences(gpu_ctx_factory, precision, tol, tolq):
    """dE/dq, dE/dsigma, dE/deps of nonbonded + PME on the perturbed 8,100-atom
    box: directional derivatives against central differences of the FP64
    energy (neutral charge direction, h = 0.05 e: the energy is quadratic in
    the charges, so the central difference carries roundoff only; relative
    steps 1e-5 and 1e-6 for LJ)."""
    say(f"case: parameter gradients water8k {precision}")
    s, sigma, eps = water.perturbed_box("water8k", seed=2)
    ids = [("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
            [8.0, 9.0, ALPHA_RC9, 9.0, 9.0]),
           ("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", [48, 48, 48, 9.0, 4, 1e-6, -1.0])]

    def ctx_for(prec, sys_, sg, ep):
        c = gpu_ctx_factory(prec)
        c.set_topology(sys_)
        c.set_boundary(sys_.box, sys_.origin, 4.5)
        c.set_lj_per_atom(sg, ep)
        return c, [c.force(i, v) for i, v in ids]

    c, hs = ctx_for(precision, s, sigma, eps)
    dq, ds, de = c.param_gradients(hs, s.pos)

    def energy(sys_, sg, ep):
        cc, hh = ctx_for("double", sys_, sg, ep)
        _, e = cc.evaluate(hh, sys_.pos)
        cc.close()
        return float(e[0] + e[1])

    rng = np.random.default_rng(5)
    w = rng.standard_normal(s.n)
    w -= w.mean()
    w /= np.linalg.norm(w)
    hq = 0.05 * water.SQRT_COULOMB
    sp, sm = s.copy(), s.copy()
    sp.q = s.q + hq * w
    sm.q = s.q - hq * w
    fdq = (energy(sp, sigma, eps) - energy(sm, sigma, eps)) / (2 * hq)
    say("dE/dq . w", float(dq @ w), "FD", fdq)
    assert abs(float(dq @ w) - fdq) <= tolq * abs(fdq)
    o = s.type == 0
    ws = np.where(o, rng.standard_normal(s.n), 0.0)
    hs_ = 1e-5
    fds = (energy(s, sigma + hs_ * ws, eps) - energy(s, sigma - hs_ * ws, eps)) / (2 * hs_)
    say("dE/dsigma . w", float(ds @ ws), "FD", fds)
    assert abs(float(ds @ ws) - fds) <= tol * abs(fds) + 1e-6
    we = np.where(o, rng.standard_normal(s.n), 0.0)
    he = 1e-6
    fde = (energy(s, sigma, eps + he * we) - energy(s, sigma, eps - he * we)) / (2 * he)
    say("dE/deps . w", float(de @ we), "FD", fde)
    assert abs(float(de @ we) - fde) <= tol * abs(fde) + 1e-5


def test_per_atom_lj_matches_type_table(gpu_ctx_factory, port):
    """Per-atom sigma/epsilon (CHARMM combination in the kernel) against the
    oracle with one atom type per atom, on a 150-atom box."""
    say("case: per-atom LJ vs per-atom type table, 150 atoms")
    s = water.water_box(50, 11.5, seed=6, jitter=0.3)
    rng = np.random.default_rng(6)
    sigma = np.where(s.type == 0, water.SIGMA_O, 0.4) * (1 + 0.02 * rng.standard_normal(s.n))
    eps = np.where(s.type == 0, water.EPS_O, 0.046) * (1 + 0.02 * rng.standard_normal(s.n))
    A, B = water.per_atom_lj_tables(sigma, eps)
    st = s.copy()
    st.type = np.arange(s.n, dtype=np.int32)
    st.ljA, st.ljB, st.ljA14, st.ljB14 = A, B, A, B
    terms = [po.term(po.POT_LJ, po.SWF_C2, 5.0, 4.0)]
    o = port.cutoff(st, terms, 5.0, 2.5)
    for precision, tol in (("double", 1e-11), ("mixed", 5e-6)):
        c = gpu_ctx_factory(precision)
        c.set_topology(s)
        c.set_boundary(s.box, s.origin, 2.5)
        c.set_lj_per_atom(sigma, eps)
        f, e = c.evaluate([c.force("LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [4.0, 5.0, 5.0])], s.pos)
        say(precision, "rms rel", rms_rel(f, o["forces"]), "E", e[1], o["energies"][1])
        assert rms_rel(f, o["forces"]) < tol
        assert e[1] == pytest.approx(o["energies"][1], rel=tol * 10)
