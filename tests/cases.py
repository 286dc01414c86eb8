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
