"""The host front end (libpmfront.so) against the reference's own readers and
topology builder, and against what the reference's regression files fix.

The inputs are the reference's example and regression files, kept as data
under tests/golden/examples/:
  water_216/  SPC water box in CHARMM22 PAR format (examples/water_216)
  water_648/  TIP3 (X-PLOR / CHARMM19 PAR) and SPC/E (CHARMM22) waters
              (examples/water_648)
  wwd/        the WW-domain protein with par_all27_prot_lipid.inp and the
              allEnergies files the reference wrote for it
              (examples/REGRESSION/Hessian, NormalMode, Minimization)

Pins, strongest first:
  * bitwise equality with the reference's PSFReader + PARReader +
    buildTopology compiled here (oracle/_ref), for all four systems;
  * the counts the reference printed for the protein in
    examples/REGRESSION/Hessian/output.correct (551 atoms, 32 types, 561
    bonds, 1003 angles, 1485 dihedrals with 1558 terms, 89 impropers, one
    molecule, 1650 degrees of freedom, 3004 exclusion pairs);
  * the CHARMM combination rule on the PAR values, by hand;
  * the allEnergies line format, by rewriting every line of the reference's
    own energy files byte for byte.
"""
import os

import numpy as np
import pytest

from protomol_amd import front

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples")
SYSTEMS = {
    "water216": ("water_216/water.216.psf", "water_216/water.216.par"),
    "water648_tip3": ("water_648/water648.tip3.psf", "water_648/water648.tip3.par"),
    "water648_spce": ("water_648/water648.spce.psf", "water_648/water648.spce.par"),
    "wwd": ("wwd/wwd_mineq.psf", "wwd/par_all27_prot_lipid.inp"),
}


def paths(name):
    return tuple(os.path.join(G, p) for p in SYSTEMS[name])


def last_class(t):
    last = {}
    for a, b, c in zip(t.ex_a, t.ex_b, t.ex_cls):
        last[(int(a), int(b))] = int(c)
    return last


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("exclude", ["scaled1-4", "1-3"])
def test_topology_matches_reference_build(ref, name, exclude):
    from oracle import pyoracle as po
    psf, par = paths(name)
    t = front.load_system(psf, par, exclude)
    r = po.ref_topology(psf, par, exclude)
    try:
        for k, v in r["info"].items():
            assert t.info[k] == v, k
        assert t.type_names == r["type_names"]
        for k in ["q", "mass", "type", "molecule", "ljA", "ljB", "ljA14", "ljB14", "bond_a", "bond_b", "bond_k",
                  "bond_r0", "angle_a", "angle_b", "angle_c", "angle_k", "angle_t0", "angle_kub", "angle_rub"]:
            np.testing.assert_array_equal(getattr(t, k), r[k], err_msg=k)
        for key in ["dihedrals", "impropers"]:
            for f in ["atoms", "term_row", "k", "period", "phase"]:
                np.testing.assert_array_equal(getattr(getattr(t, key), f), r[key][f], err_msg=f"{key}.{f}")
        last = last_class(t)
        pairs = np.array(sorted(last)) if last else np.zeros((0, 2), np.int32)
        if len(pairs):
            np.testing.assert_array_equal(po.ref_exclusion_class(r, pairs[:, 0], pairs[:, 1]),
                                          [last[tuple(p)] for p in pairs])
    finally:
        po.ref_topology_free(r)


def test_protein_counts_match_reference_printout():
    """examples/REGRESSION/Hessian/output.correct, the topology block."""
    t = front.load_system(*paths("wwd"))
    i = t.info
    assert (i["atoms"], i["types"], i["bonds"], i["angles"]) == (551, 32, 561, 1003)
    assert (i["dihedrals"], i["dihedral_terms"], i["impropers"]) == (1485, 1558, 89)
    assert (i["molecules"], i["dof"], i["par_dialect"], i["dihedral_mult_psf"]) == (1, 1650, 28, 0)
    assert len(last_class(t)) == 3004


def test_water_dialects_and_lj_by_hand():
    """TIP3 parameters in the X-PLOR (CHARMM19) file: sigma is converted
    to Rmin/2 by 2^(-5/6); SPC/E in the CHARMM22 file is Rmin/2 already."""
    t = front.load_system(*paths("water648_tip3"))
    assert t.info["par_dialect"] == 19 and t.info["dihedral_mult_psf"] == 1
    o = t.type_names.index("OT")
    s = front.load_system(*paths("water648_spce"))
    assert s.info["par_dialect"] == 28
    so = s.type_names.index("SPCO")
    r, e = 2 * 1.774, 0.15525                      # SPCO: eps -0.15525, Rmin/2 1.774 (water648.spce.par)
    np.testing.assert_allclose(s.ljA[so, so], r ** 12 * e, rtol=1e-15)
    np.testing.assert_allclose(s.ljB[so, so], 2 * r ** 6 * e, rtol=1e-15)
    assert t.ljA[o, o] > 0 and np.allclose(t.ljA, t.ljA.T) and np.allclose(s.ljB, s.ljB.T)
    # charges scaled by sqrt(332.0636); masses as in the PSF
    np.testing.assert_allclose(s.q[:3] / 18.2226123264, [-0.8476, 0.4238, 0.4238], rtol=1e-15)
    np.testing.assert_allclose(s.mass[:3], [15.9994, 1.008, 1.008])


def test_exclusion_classes_scaled14():
    """scaled1-4: bonds and 1-3 pairs full, dihedral ends modified unless a
    shorter path made them full first... the last add wins, as in
    ExclusionTable::add."""
    t = front.load_system(*paths("wwd"))
    last = last_class(t)
    bonds = {(min(a, b), max(a, b)) for a, b in zip(t.bond_a, t.bond_b)}
    assert all(last[p] == 2 for p in bonds if p not in
               {(min(d[0], d[3]), max(d[0], d[3])) for d in t.dihedrals.atoms})
    ends = [(min(d[0], d[3]), max(d[0], d[3])) for d in t.dihedrals.atoms]
    assert all(last[p] == 1 for p in ends)
    t13 = front.load_system(*paths("wwd"), exclude="1-3")
    assert set(last_class(t13).values()) == {2}


def test_read_coords_pdb_and_xyz():
    x = front.read_coords(os.path.join(G, "water_216/water.216.equil.pdb"))
    assert x.shape == (648, 3)
    np.testing.assert_array_equal(x[0], [0.40208, -7.0032, 0.77401])
    np.testing.assert_array_equal(x[-1], [3.86304, 11.8601, 3.15403])
    v = front.read_coords(os.path.join(G, "water_216/water.216.equil.vel.xyz"))
    assert v.shape == (648, 3)
    np.testing.assert_array_equal(v[0], [0.0527719709945324, 0.0510948551807236, 0.207482722122328])
    p = front.read_coords(os.path.join(G, "water_648/water648.fluid.pos.xyz"))
    assert p.shape == (648, 3)
    np.testing.assert_array_equal(p[0], [-3.98414434604143, 2.26027991298031, -7.89295144965026])


@pytest.mark.parametrize("fname", ["wwd.energy", "wwd.anm.energy", "wwd.min.energy"])
def test_energies_line_rewrites_reference_files(fname):
    with open(os.path.join(G, "wwd", fname)) as f:
        lines = f.readlines()
    assert lines
    for line in lines:
        vals = [float(w) for w in line.split()]
        assert len(vals) == 13
        assert front.energies_line(vals) == line


def test_pdb_writer_round_trip(tmp_path):
    src = os.path.join(G, "water_216/water.216.equil.pdb")
    x = front.read_coords(src)
    out = str(tmp_path / "out.pdb")
    front.write_pdb(out, src, x, "round trip")
    with open(src) as a, open(out) as b:
        ra = [l.rstrip("\n") for l in a if l.startswith("ATOM")]
        rb = [l.rstrip("\n") for l in b if l.startswith("ATOM")]
    # the source lines were written by the reference's PDBWriter: the same
    # columns outside the coordinates (81-column layout included), and the
    # coordinates read back the same ("-13.180" there is "-13.18" after a
    # round trip through toString)
    assert len(ra) == len(rb) == 648
    for a, b in zip(ra, rb):
        assert a[:31] == b[:31] and a[55:].rstrip() == b[55:].rstrip()
    np.testing.assert_array_equal(front.read_coords(out), x)


def test_errors_are_reported(tmp_path):
    bad = tmp_path / "bad.psf"
    bad.write_text("PSF\n\n       2 !NATOM\n 1 W 1 TIP3 OH2 OT -0.8 15.9 0\n")
    with pytest.raises(front.FrontError, match="ends inside"):
        front.load_system(str(bad), paths("water216")[1])
    with pytest.raises(front.FrontError, match="exclude"):
        front.load_system(*paths("water216"), exclude="2-5")
    psf, _ = paths("wwd")
    with pytest.raises(front.FrontError, match="could not find bond"):
        front.load_system(psf, paths("water216")[1])


def test_dcd_and_xyz_writers_match_reference_writers(ref, tmp_path):
    """Byte for byte against DCDTrajectoryWriter and XYZWriter compiled from
    the reference, except the free-text remark lines (build date, user)."""
    import ctypes as C
    psf, par = paths("water216")
    x = front.read_coords(os.path.join(G, "water_216/water.216.equil.pdb"))
    rng = np.random.default_rng(2)
    frames = np.stack([x + 0.01 * k * rng.standard_normal(x.shape) for k in range(3)])
    mine, theirs = str(tmp_path / "a.dcd"), str(tmp_path / "b.dcd")
    front.write_dcd(mine, frames)
    f = np.ascontiguousarray(frames)
    assert ref.lib.ref_write_dcd(theirs.encode(), x.shape[0], 3, f.ctypes.data_as(C.POINTER(C.c_double))) == 0
    a, b = open(mine, "rb").read(), open(theirs, "rb").read()
    assert len(a) == len(b)
    title = slice(100, 260)                  # the two 80-character remarks
    assert a[:title.start] == b[:title.start] and a[title.stop:] == b[title.stop:]
    t = front.load_system(psf, par)
    xm, xt = str(tmp_path / "a.xyz"), str(tmp_path / "b.xyz")
    front.write_xyz(xm, psf, par, frames[2], "c")
    names = "\n".join(t.type_names[k] for k in t.type).encode()
    f2 = np.ascontiguousarray(frames[2])
    assert ref.lib.ref_write_xyz(xt.encode(), x.shape[0], f2.ctypes.data_as(C.POINTER(C.c_double)), names) == 0
    la, lb = open(xm).read().splitlines(), open(xt).read().splitlines()
    assert la[0] == lb[0] and la[2:] == lb[2:]
