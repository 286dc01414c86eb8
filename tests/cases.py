"""Force groups and systems shared by the tests (the named workloads)."""
import numpy as np

from protomol_amd import water
from oracle import pyoracle as po

ALPHA_RC10 = 0.3123413274340875   # erfc bisection at rc = 10, accuracy 1e-6
ALPHA_RC9 = 0.3488316179           # rc = 9 (NonbondedPMEwaldSystemForce.h:564-586)
ALPHA_RC8 = 0.394670228244391


def c1_terms():
    """648-atom water: LJ and Coulomb, C2 switching 7 -> 8, algorithm cutoff 8."""
    return ([po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)], [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)])


def gpu_forces_c1(ctx):
    h1 = ctx.force("LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0])
    h2 = ctx.force("Coulomb -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0])
    return [h1, h2]


def rms_rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def exclusion_mix(s):
    """A copy of a water system whose exclusion table exercises every rule of
    ExclusionTable / OneAtomPair: H atoms carry LJ, 1-4 tables differ from the
    plain ones, csf = 0.5; some H-H pairs are re-added as MODIFIED and some O-H
    pairs as NONE (the last add wins), some pairs are added twice, and some
    pairs join atoms of different molecules (never applied by OneAtomPair,
    kept in the PME correction table)."""
    s = s.copy()
    A, B = water._lj_tables([(water.SIGMA_O, water.EPS_O), (0.4, 0.046)])
    s.ljA, s.ljB = A, B
    s.ljA14, s.ljB14 = 0.7 * A, 0.6 * B
    s.csf = 0.5
    nm = s.n // 3
    ea, eb, ec = list(s.ex_a), list(s.ex_b), list(s.ex_cls)
    for m in range(nm):
        o, h1, h2 = 3 * m, 3 * m + 1, 3 * m + 2
        if m % 3 == 0:
            ea.append(h2); eb.append(h1); ec.append(water.EXCL_MODIFIED)
        if m % 5 == 1:
            ea.append(o); eb.append(h1); ec.append(water.EXCL_NONE)
        if m % 4 == 2:
            ea.append(o); eb.append(h2); ec.append(water.EXCL_FULL)
        if m % 7 == 3 and m + 1 < nm:
            ea.append(o); eb.append(o + 3); ec.append(water.EXCL_FULL)
    s.ex_a = np.array(ea, np.int32)
    s.ex_b = np.array(eb, np.int32)
    s.ex_cls = np.array(ec, np.int32)
    return s


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def langevin_noise(seed, counter, n3):
    """The Gaussian deviates of the device Langevin half kick number `counter`
    (SplitMix64 of (seed, counter, index), Box-Muller), restated in numpy."""
    with np.errstate(over="ignore"):
        idx = np.arange(n3, dtype=np.uint64)
        key = np.uint64(counter) * np.uint64(0x100000001B3) + idx * np.uint64(2)
        h1 = _mix64(np.uint64(seed) ^ _mix64(key))
        h2 = _mix64(h1 ^ np.uint64(0x5851F42D4C957F2D))
    u1 = ((h1 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740993.0)
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def rem_uniform(seed, attempt, k):
    """The exchange test's uniform deviate for (seed, attempt, ladder pair k)."""
    with np.errstate(over="ignore"):
        h = _mix64(np.uint64(seed) ^ _mix64(np.uint64(attempt) * np.uint64(0x100000001B3) + np.uint64(2 * k + 1)))
    return float(h >> np.uint64(11)) * (1.0 / 9007199254740992.0)


def replica_context(make, s, temperature, seed, precision="mixed"):
    """A context for one replica of the batched workload: LJ and Coulomb with C2
    switching 7 -> 8 A (one fused pass), bonds, angles, Langevin at `temperature`."""
    c = make(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.0)
    hs = gpu_forces_c1(c) + [c.force("Bond", []), c.force("Angle", [])]
    c.md_setup(hs, "langevin", dt=1.0, temperature=float(temperature), gamma=5.0, seed=seed)
    c.md_set_state(s.pos, s.vel)
    return c


def lanczos_reference(matvec, v0, steps):
    """Lanczos with full reorthogonalisation (two classical Gram-Schmidt
    passes, alpha_j = the v_j coefficients of both), restated in numpy as the
    checker of the device iteration.  Returns (alpha, beta, V[steps + 1])."""
    n = v0.size
    V = np.zeros((steps + 1, n))
    V[0] = v0.reshape(-1) / np.linalg.norm(v0)
    alpha = np.zeros(steps)
    beta = np.zeros(steps)
    for j in range(steps):
        w = np.asarray(matvec(V[j]), float).reshape(-1)
        for _ in range(2):
            c = V[:j + 1] @ w
            w = w - c @ V[:j + 1]
            alpha[j] += c[j]
        beta[j] = np.linalg.norm(w)
        V[j + 1] = w / beta[j] if beta[j] > 1e-300 else 0.0
    return alpha, beta, V


def tridiagonal_eigen(alpha, beta):
    m = len(alpha)
    T = np.diag(alpha) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
    return np.linalg.eigh(T)
