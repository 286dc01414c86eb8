"""The reference's own periodic PME example on the GPU:
examples/bpti_water_14281/bpti.pme.opt.conf (BPTI in water, 14,281 atoms,
63 x 50 x 50 A periodic box, cell size 6; Bond, Angle, Dihedral, Improper;
LennardJones C2 (switch-on 0.1 A) + CoulombEwaldReal in one NonbondedCutoff
force at 10 A; PME reciprocal + correction on a 60 x 48 x 48 grid, order 4),
the only in-scope example with CHARMM scaled 1-4 (MODIFIED) pairs.

pmrun runs the file unchanged (its step count and energy output are set from
the command line, as the reference's parseCommandLine allows); its step-0
allEnergies line is compared with the compiled reference build's energies at
the PDB positions (oracle/_ref: the reference's own force code over the
topology its own readers and buildTopology make, which the front end
reproduces bitwise).  The forces of the same force set through the C ABI are
compared with the reference build's forces: mixed precision to the 5e-6 RMS
relative bar, FP64 to 1e-10."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from protomol_amd import front
from protomol_amd.build import PMRUN
from tests.cases import rms_rel

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples", "bpti")
FILES = ["bpti.pme.opt.conf", "bpti.psf", "bpti.par", "bpti.pdb"]
BOX, ORIGIN, CELL = (63.0, 50.0, 50.0), (0.0, 0.0, 0.0), 6.0
ALPHA = 0.312341327434087
LJ_EWALD = ("LennardJones CoulombEwaldReal -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff "
            "-switchon 0.1 -cutoff 10.0 -alpha 0.312341327434087 -cutoff 10 -cutoff 10")
PME = ("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline -gridsize 60 48 48 -cutoff 10 "
       "-order 4 -accuracy 1e-6")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def system():
    t = front.load_system(os.path.join(G, "bpti.psf"), os.path.join(G, "bpti.par"))
    x = front.read_coords(os.path.join(G, "bpti.pdb"))
    return t, t.system(x, box=BOX, origin=ORIGIN)


def reference(ref, t, s):
    """Forces and energies of the example's force set from the reference build."""
    from oracle import pyoracle as po
    terms = [po.term(po.POT_LJ, po.SWF_C2, 10.0, 0.1), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 10.0, alpha=ALPHA)]
    parts = [ref.cutoff(s, terms, 10.0, CELL), ref.pme(s, (60, 48, 48), 4, 10.0, CELL, flags=6), ref.bonded(s, 3)]
    for imp, tors in ((0, t.dihedrals), (1, t.impropers)):
        parts.append(ref.torsions(s, tors, imp))
    f = sum(p["forces"] for p in parts)
    e = sum(p["energies"] for p in parts)
    return f, e


@pytest.fixture(scope="module")
def bpti(ref):
    t, s = system()
    f, e = reference(ref, t, s)
    return t, s, f, e


@pytest.mark.parametrize("precision,rtol", [("double", 1e-9), ("mixed", 2e-6)])
def test_pmrun_bpti_pme_first_energy_line(tmp_path, bpti, precision, rtol):
    t, s, fr, er = bpti
    tmp = str(tmp_path)
    for f in FILES:
        shutil.copy(os.path.join(G, f), tmp)
    say(f"case: pmrun bpti.pme.opt.conf {precision}")
    r = subprocess.run([PMRUN, "bpti.pme.opt.conf", "--precision", precision, "--numsteps", "10",
                        "--allEnergiesFile", "bpti.energy"], cwd=tmp, capture_output=True, text=True, timeout=110)
    say(r.stdout[-2000:], r.stderr[-1500:])
    assert r.returncode == 0, r.stderr
    lines = open(os.path.join(tmp, "bpti.energy")).read().split("\n")
    got = [float(w) for w in lines[0].split()]
    # allEnergies columns: time, E_potential, E_kinetic, E_total, T, E_bond,
    # E_angle, E_dihedral, E_improper, E_VdW, E_coulomb, E_other, volume
    want = {5: er[2], 6: er[3], 7: er[4], 8: er[5], 9: er[1], 10: er[0], 1: er[:7].sum()}
    say("pmrun", got, "reference", want)
    assert got[0] == 0.0 and got[2] == 0.0            # step 0, temperature 0: no kinetic energy
    for col, v in want.items():
        # printed to eight significant digits; bonded terms are FP64 in both modes
        tol = max(rtol, 6e-8) if col in (1, 9, 10) else 6e-8
        assert abs(got[col] - v) <= tol * abs(v) + 1e-6, (col, got[col], v)
    assert len([ln for ln in lines if ln.strip()]) == 2      # steps 0 and 10


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("mixed", 5e-6)])
def test_bpti_forces_match_reference_build(gpu_ctx_factory, bpti, precision, tol):
    t, s, fr, er = bpti
    say(f"case: bpti forces through the C ABI {precision}")
    c = gpu_ctx_factory(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, CELL)
    hs = [c.force(*front.force_values(d)) for d in (LJ_EWALD, PME)]
    hs += [c.force(k, []) for k in ("Bond", "Angle", "Dihedral", "Improper")]
    f, e = c.evaluate(hs, s.pos)
    err = rms_rel(f, fr)
    say("rms rel force", err, "energies", e[:6], er[:6])
    assert err < tol
    np.testing.assert_allclose(e[:6], er[:6], rtol=max(tol, 1e-11), atol=1e-6)
