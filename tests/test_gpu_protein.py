"""Bonded forces of a protein on the GPU: the WW domain (551 atoms,
par_all27) read by the host front end, against the energies the reference
wrote for it and against the oracle.

examples/REGRESSION/Hessian/wwd.energy (and NormalMode/wwd.energy), line 1,
holds E_bond 19.256258, E_angle 80.007526, E_dihedral 165.01607 and
E_improper 2.9485567 for wwd.min.pdb: the PSF/PAR readers, the topology
builder and the four bonded forces together must reproduce them to the
eight digits printed.  Forces and virials are checked against the oracle's
restatement (itself bitwise with the reference build on these inputs, see
tests/test_oracle.py).  The dihedrals include planar peptide bonds, so both
branches of calcTorsion (|sin phi| > 0.1 and the small-sine one) run.
"""
import os
import sys

import numpy as np
import pytest

from protomol_amd import front

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples", "wwd")
GOLDEN = {"bond": 19.256258, "angle": 80.007526, "dihedral": 165.01607, "improper": 2.9485567}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def protein():
    t = front.load_system(os.path.join(G, "wwd_mineq.psf"), os.path.join(G, "par_all27_prot_lipid.inp"))
    x = front.read_coords(os.path.join(G, "wwd.min.pdb"))
    ext = x.max(0) - x.min(0)
    box = np.full(3, 2 * ext.max() + 20.0)          # a periodic box no pair of the molecule wraps in
    return t, t.system(x, box=box, origin=x.mean(0))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_protein_bonded_energies_match_reference_output(gpu_ctx_factory, port, precision):
    say(f"case: wwd bonded {precision}")
    t, s = protein()
    c = gpu_ctx_factory(precision)
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 8.0)
    hs = [c.force(k, []) for k in ("Bond", "Angle", "Dihedral", "Improper")]
    f, e = c.evaluate(hs, s.pos, virial=True)
    got = dict(bond=e[2], angle=e[3], dihedral=e[4], improper=e[5])
    say("energies", got)
    for k, v in GOLDEN.items():
        assert float(f"{got[k]:.8g}") == v, (k, got[k], v)
    # forces and virial against the oracle (FP64 terms in both precisions)
    fo = np.zeros_like(f)
    eo = np.zeros(16)
    b = port.bonded(s, 3, virial=True)
    fo += b["forces"]
    eo += b["energies"]
    for imp, tors in ((0, t.dihedrals), (1, t.impropers)):
        r = port.torsions(s, tors, imp, virial=True)
        fo += r["forces"]
        eo += r["energies"]
    err = np.abs(f - fo).max() / np.abs(fo).max()
    say("max |df| / max |f|", err)
    assert err < 1e-12
    np.testing.assert_allclose(e[2:6], eo[2:6], rtol=1e-12)
    np.testing.assert_allclose(e[7:16], eo[7:16], rtol=1e-9, atol=1e-9 * np.abs(eo[7:16]).max())


def test_protein_torsions_in_pme_gather_tail(gpu_ctx_factory):
    """With one PME term the gather thread of an atom adds its bonded forces,
    torsions included: the same forces and energies as evaluating the terms
    one handle at a time."""
    say("case: wwd pme gather tail with torsions")
    t, s = protein()
    c = gpu_ctx_factory("mixed")
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 8.0)
    n = 80                                        # about 1 A per grid point, radix 2 and 5
    args = [n, n, n, 9.0, 4, 1e-6, -1.0]
    hp = c.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", args)
    hr = c.force("Coulomb -algorithm PMEwald -reciprocal -interpolation BSpline", args)
    hc = c.force("Coulomb -algorithm PMEwald -correction -interpolation BSpline", args)
    hb = [c.force(k, []) for k in ("Bond", "Angle", "Dihedral", "Improper")]
    f_all, e_all = c.evaluate([hp] + hb, s.pos)
    parts = [c.evaluate([h], s.pos) for h in [hr, hc] + hb]
    f_sep = sum(p[0] for p in parts)
    e_sep = sum(p[1] for p in parts)
    fmax = np.abs(f_all).max()
    say("max |df|", np.abs(f_all - f_sep).max(), "of", fmax)
    assert np.abs(f_all - f_sep).max() <= 1e-11 * fmax
    np.testing.assert_allclose(e_all, e_sep, rtol=1e-12, atol=1e-9)


def test_torsion_input_errors(gpu_ctx_factory):
    t, s = protein()
    c = gpu_ctx_factory("double")
    c.set_topology(s)
    bad = front.Torsions(np.array([[0, 1, 1, 2]], np.int32), np.array([0, 1], np.int32), np.ones(1),
                         np.ones(1, np.int32), np.zeros(1))
    from protomol_amd.gpu import PmError
    with pytest.raises(PmError, match="repeats an atom"):
        c.set_torsions(0, bad)
    bad.atoms[0] = [0, 1, 2, s.n]
    with pytest.raises(PmError, match="names atom"):
        c.set_torsions(1, bad)
