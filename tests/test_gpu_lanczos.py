"""Device-resident Lanczos (pm_lanczos) against the numpy Lanczos checker
driven by the oracle's reference-pinned Hessian-vector products, against the
spectrum of the dense oracle Hessian (Cauchy interlacing), and on the
8,100-atom PME workload (30 steps, 20 lowest mass-weighted modes, FP64 and
mixed): orthonormal Ritz vectors and residual bounds that agree with
explicit residuals."""
import sys
import time

import numpy as np
import pytest

from oracle import pyoracle as po
from protomol_amd import water
from tests.cases import ALPHA_RC9, c1_terms, gpu_forces_c1, lanczos_reference, tridiagonal_eigen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mass_weighted", [False, True])
def test_lanczos_matches_checker(gpu_ctx_factory, port, mass_weighted):
    say(f"case: lanczos water648 LJ+Coulomb C2, 30 steps, mass_weighted={mass_weighted}")
    s = water.named("water648", seed=21, jitter=0.4)
    c = gpu_ctx_factory("double")
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.0)
    hs = gpu_forces_c1(c)
    v0 = np.random.default_rng(8).standard_normal((s.n, 3))
    vals, vecs, res, al, be = c.lanczos(hs, s.pos, v0, steps=30, nvec=5, mass_weighted=mass_weighted)
    lj, co = c1_terms()
    w = 1.0 / np.sqrt(np.repeat(s.mass, 3)) if mass_weighted else np.ones(3 * s.n)

    def mv(x):
        y = (w * x).reshape(-1, 3)
        return w * (port.hessian_vec(s, lj, y) + port.hessian_vec(s, co, y)).reshape(-1)

    ral, rbe, V = lanczos_reference(mv, v0.reshape(-1), 30)
    scale = np.abs(ral).max()
    say("max |alpha diff| / scale", float(np.abs(al - ral).max() / scale),
        "max |beta diff| / scale", float(np.abs(be - rbe).max() / scale))
    assert np.abs(al - ral).max() <= 1e-8 * scale
    assert np.abs(be - rbe).max() <= 1e-8 * scale
    th, S = tridiagonal_eigen(ral, rbe)
    assert np.abs(vals - th).max() <= 1e-8 * np.abs(th).max()
    Y = S[:, :5].T @ V[:30]
    for i in range(5):   # eigenvectors carry a sign
        sgn = np.sign(Y[i] @ vecs[i].reshape(-1))
        assert np.abs(sgn * vecs[i].reshape(-1) - Y[i]).max() <= 1e-6
    assert np.abs(res - np.abs(rbe[-1] * S[-1])).max() <= 1e-8 * scale


def test_lanczos_interlaces_the_dense_spectrum(gpu_ctx_factory, port):
    """81 atoms (the dense-Hessian fixture system), LJ + Coulomb C2 3.5 -> 4.5:
    every Ritz value of a 60-step run lies inside Cauchy's interlacing bounds
    of the dense oracle Hessian's eigenvalues, and the extreme ones converge."""
    say("case: lanczos interlacing, 81 atoms, 60 steps")
    s = water.water_box(27, 9.4, seed=3, jitter=0.3)
    terms = [po.term(po.POT_LJ, po.SWF_C2, 4.5, 3.5), po.term(po.POT_COULOMB, po.SWF_C2, 4.5, 3.5)]
    n3 = 3 * s.n
    H = np.zeros((n3, n3))
    for k in range(n3):
        e = np.zeros(n3)
        e[k] = 1.0
        H[:, k] = sum(port.hessian_vec(s, [t], e.reshape(-1, 3), cellsize=2.0) for t in terms).reshape(-1)
    lam = np.linalg.eigvalsh(0.5 * (H + H.T))
    c = gpu_ctx_factory("double")
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 2.0)
    hs = [c.force("LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [3.5, 4.5, 4.5]),
          c.force("Coulomb -algorithm NonbondedCutoff -switchingFunction C2", [3.5, 4.5, 4.5])]
    m = 60
    vals, _, _, _, _ = c.lanczos(hs, s.pos, np.random.default_rng(9).standard_normal((s.n, 3)), steps=m)
    eps = 1e-9 * np.abs(lam).max()
    say("extremes", vals[0], lam[0], vals[-1], lam[-1])
    for k in range(m):
        assert vals[k] >= lam[k] - eps
        assert vals[m - 1 - k] <= lam[n3 - 1 - k] + eps
    assert abs(vals[-1] - lam[-1]) <= 1e-6 * abs(lam[-1])
    assert abs(vals[0] - lam[0]) <= 1e-6 * abs(lam).max()


def test_lanczos_c5_pme_modes(gpu_ctx_factory):
    """The 8,100-atom perturbed box, LJ (per-atom) + Ewald real space rc 9 and
    PME 48^3: 30 mass-weighted steps, the 20 lowest Ritz pairs."""
    s, sigma, eps = water.perturbed_box("water8k", seed=1)
    out = {}
    for precision in ("double", "mixed"):
        say(f"case: lanczos water8k PME 48^3 {precision}")
        c = gpu_ctx_factory(precision)
        c.set_topology(s)
        c.set_boundary(s.box, s.origin, 4.5)
        c.set_lj_per_atom(sigma, eps)
        hs = [c.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
                      [8.0, 9.0, ALPHA_RC9, 9.0, 9.0]),
              c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                      [48, 48, 48, 9.0, 4, 1e-6, -1.0])]
        v0 = np.random.default_rng(10).standard_normal((s.n, 3))
        c.lanczos(hs, s.pos, v0, steps=30, nvec=20, mass_weighted=True)   # warm (allocations, FFT plans)
        c.synchronize()
        t0 = time.perf_counter()
        vals, vecs, res, al, be = c.lanczos(hs, s.pos, v0, steps=30, nvec=20, mass_weighted=True)
        ms = (time.perf_counter() - t0) * 1e3
        say(precision, "30-step Lanczos + 20 Ritz vectors: %.2f ms" % ms, "lowest", vals[:3], "highest", vals[-1])
        Y = vecs.reshape(20, -1)
        assert np.abs(Y @ Y.T - np.eye(20)).max() <= (1e-10 if precision == "double" else 1e-5)
        w = 1.0 / np.sqrt(np.repeat(s.mass, 3))
        for i in (0, 1, 19):
            ay = w * c.hessian_vec(hs, s.pos, (w * Y[i]).reshape(-1, 3)).reshape(-1)
            r = np.linalg.norm(ay - vals[i] * Y[i])
            say(" ritz", i, vals[i], "residual", r, "bound", res[i])
            assert abs(r - res[i]) <= (1e-6 if precision == "double" else 1e-3) * np.abs(vals).max()
        out[precision] = vals
    top = np.abs(out["double"]).max()
    say("max |mixed - double| / max|theta|", float(np.abs(out["mixed"] - out["double"]).max() / top))
    assert np.abs(out["mixed"] - out["double"]).max() <= 1e-3 * top
