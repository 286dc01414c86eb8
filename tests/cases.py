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
