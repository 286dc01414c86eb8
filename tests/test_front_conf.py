"""The .conf layer of the host front end (CPU).

  * force definitions -> device force ids and positional values, following
    ForceFactoryDetails::doMake's keyword matching (each parameter looked up
    from where the previous one ended, wrapping once; a keyword-less value
    follows the previous one; defaults filled in).  The expectations are the
    values the reference prints for such definitions (the "-cutoff" given
    once serves the switching function and the algorithm, as in
    examples/REGRESSION/Hessian/output.correct) and the bpti example's
    three "-cutoff" words taken in order;
  * start-up momentum removal and random velocities against the reference
    build (removeLinearMomentum, removeAngularMomentum, randomVelocity);
  * the closed forms: no linear momentum left in the solute, no angular
    momentum left about its centre of mass.
"""
import os

import numpy as np
import pytest

from protomol_amd import front

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples")
W216 = (os.path.join(G, "water_216/water.216.psf"), os.path.join(G, "water_216/water.216.par"))
WWD = (os.path.join(G, "wwd/wwd_mineq.psf"), os.path.join(G, "wwd/par_all27_prot_lipid.inp"))


@pytest.mark.parametrize("definition,fid,values", [
    ("Angle", "Angle", []),
    ("LennardJones -algorithm NonbondedCutoff -switchingFunction C2 -switchon 11 -cutoff 13",
     "LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [11, 13, 13]),
    ("Coulomb -algorithm NonbondedCutoff -switchingFunction C1 -cutoff 13",
     "Coulomb -algorithm NonbondedCutoff -switchingFunction C1", [13, 13]),
    # examples/bpti_water_14281/bpti.pme.opt.conf
    ("time LennardJones CoulombEwaldReal -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff "
     "-switchon 0.1 -cutoff 10.0 -alpha 0.312341327434087 -cutoff 10 -cutoff 10",
     "LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction Cutoff",
     [0.1, 10.0, 0.312341327434087, 10, 10]),
    ("time Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline -gridsize 60 48 48 -cutoff 10 "
     "-order 4 -accuracy 1e-6",
     "Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline", [60, 48, 48, 10, 4, 1e-6, -1.0]),
    ("Coulomb -algorithm PMEwald -interpolation BSpline -cutoff 9 -gridsize 24 24 32",
     "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline",
     [24, 24, 32, 9, 4, 1e-6, -1.0]),
])
def test_force_values(definition, fid, values):
    got_id, got = front.force_values(definition)
    assert got_id == fid
    np.testing.assert_array_equal(got, values)


def test_force_values_errors():
    with pytest.raises(front.FrontError, match="-switchon is not defined"):
        front.force_values("LennardJones -algorithm NonbondedCutoff -switchingFunction C2 -cutoff 8")
    with pytest.raises(front.FrontError, match="not implemented on the GPU"):
        front.force_values("Coulomb -algorithm NonbondedFull")


def test_momentum_removal_matches_reference(ref):
    from oracle import pyoracle as po
    x = front.read_coords(os.path.join(G, "wwd/wwd.min.pdb"))
    rng = np.random.default_rng(4)
    v = rng.normal(0, 0.01, x.shape)
    r = po.ref_topology(*WWD)
    try:
        for lin, ang in ((1, 0), (0, 1), (1, 1)):
            mine = front.remove_momentum(*WWD, x, v, lin, ang)
            theirs = po.ref_remove_momentum(r, x, v, lin, ang)
            np.testing.assert_allclose(mine, theirs, rtol=0, atol=1e-15 * np.abs(v).max())
    finally:
        po.ref_topology_free(r)


def test_momentum_removal_closed_form():
    t = front.load_system(*WWD)
    x = front.read_coords(os.path.join(G, "wwd/wwd.min.pdb"))
    v = np.random.default_rng(5).normal(0, 0.01, x.shape)
    m = t.mass[:, None]
    out = front.remove_momentum(*WWD, x, v, True, True)
    assert np.abs((m * out).sum(0)).max() < 1e-12 * np.abs(m * v).sum()
    com = (m * x).sum(0) / m.sum()
    L = (m * np.cross(x - com, out)).sum(0)
    assert np.abs(L).max() < 1e-10 * np.abs(m * np.cross(x - com, v)).sum()
    # SPC water (types SPCO/SPCH) is not recognised as water: every atom is solute
    xw = front.read_coords(os.path.join(G, "water_216/water.216.equil.pdb"))
    vw = np.random.default_rng(6).normal(0, 0.01, xw.shape)
    tw = front.load_system(*W216)
    ow = front.remove_momentum(*W216, xw, vw, True, False)
    assert np.abs((tw.mass[:, None] * ow).sum(0)).max() < 1e-12


def test_random_velocity_matches_reference(ref):
    """randomVelocity seeds the process-wide drand48 generator on its first
    call only; this is the one call of the reference's in this process."""
    from oracle import pyoracle as po
    r = po.ref_topology(*WWD)
    try:
        theirs = po.ref_random_velocity(r, 310.0, 11)
    finally:
        po.ref_topology_free(r)
    mine = front.random_velocity(*WWD, 310.0, 11)
    np.testing.assert_array_equal(mine, theirs)
    # Maxwell-Boltzmann: the kinetic temperature of 1653 components is
    # 300 K within a few percent
    t = front.load_system(*WWD)
    v300 = front.random_velocity(*WWD, 300.0, 11)
    T = (t.mass[:, None] * v300 ** 2).sum() / (0.001987191 * 3 * t.n)
    assert abs(T - 300.0) < 0.1 * 300.0
