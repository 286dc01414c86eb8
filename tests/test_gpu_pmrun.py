"""pmrun runs a reference .conf file unchanged on the GPU
(examples/water_216/water.216.lf.conf: 648 SPC atoms in a 30.9 x 27.8 x
30.7 A box with a shifted origin, Bond + Angle, LennardJones C2 and
Coulomb C1 cutoffs at 13 A, Leapfrog 0.1 fs, PDB positions and XYZ
velocities, linear and angular momentum removed at start-up).  The step
count is shortened from the command line as the reference's
parseCommandLine allows ("--numsteps 40 --outputfreq 10").

Checked against the oracle: the same start-up (front end's momentum
removal, pinned against the reference build in test_front_conf.py), then
the oracle's FP64 forces and leapfrog, compared at every allEnergies line
(eight significant digits as printed) and on the final velocities of
finXYZVelFile; the DCD trajectory's frames are the oracle's positions in
float."""
import os
import shutil
import subprocess
import sys
import ctypes as C

import numpy as np
import pytest

from protomol_amd import front
from protomol_amd.build import PMRUN

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples", "water_216")
FILES = ["water.216.lf.conf", "water.216.psf", "water.216.par", "water.216.equil.pdb", "water.216.equil.vel.xyz"]
TF = 48.88821290839616


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_pmrun(tmp, *args):
    for f in FILES:
        shutil.copy(os.path.join(G, f), tmp)
    cmd = [PMRUN, "water.216.lf.conf", *args]
    say("case: pmrun", " ".join(cmd))
    r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=120)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    return r


def oracle_run(port, nsteps, every):
    """Start-up as protomol.cpp does it, then LeapfrogIntegrator::run."""
    from oracle import pyoracle as po
    psf, par = (os.path.join(G, f) for f in FILES[1:3])
    t = front.load_system(psf, par)
    x = front.read_coords(os.path.join(G, FILES[3]))
    v = front.read_coords(os.path.join(G, FILES[4]))
    v = front.remove_momentum(psf, par, x, v, True, True)
    box = np.array([30.8927, 27.8485, 30.7217])
    s = t.system(x, v, box=box, origin=(0.98035, -0.08475, 0.19285))
    lj = [po.term(po.POT_LJ, po.SWF_C2, 13.0, 11.0)]
    co = [po.term(po.POT_COULOMB, po.SWF_C1, 13.0)]
    D = C.POINTER(C.c_double)
    lib = port.lib
    dt = 0.1

    def forces():
        a = port.cutoff(s, lj, 13.0, 7.0)
        b = port.cutoff(s, co, 13.0, 7.0)
        c = port.bonded(s, 3)
        return a["forces"] + b["forces"] + c["forces"], a["energies"] + b["energies"] + c["energies"]

    def row(step, e):
        ke = 0.5 * (t.mass[:, None] * s.vel ** 2).sum()
        pe = e[:7].sum()
        return [step * dt, pe, ke, pe + ke, 2 * ke / (0.001987191 * t.info["dof"]), e[2], e[3], e[4], e[5], e[1],
                e[0], e[6], box.prod()]

    f, e = forces()
    rows, frames = [row(0, e)], [s.pos.copy()]
    step = 0
    while step < nsteps:
        for k in range(every):
            fac = 0.5 if k == 0 else 1.0
            lib.orc_leapfrog_kickdrift(s.n, C.c_double(dt), C.c_double(fac), s.mass.ctypes.data_as(D),
                                       f.ctypes.data_as(D), s.vel.ctypes.data_as(D), s.pos.ctypes.data_as(D))
            f, e = forces()
        lib.orc_half_kick(s.n, C.c_double(dt), s.mass.ctypes.data_as(D), f.ctypes.data_as(D), s.vel.ctypes.data_as(D))
        step += every
        rows.append(row(step, e))
        frames.append(s.pos.copy())
    return np.array(rows), s.vel.copy(), frames


def read_dcd(path, n):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:92], np.int32)
    assert hdr[0] == 84 and raw[4:8] == b"CORD"
    nsets = hdr[2]
    off = 92 + 4 + 164 + 4 + 12
    frames = []
    for _ in range(nsets):
        xyz = []
        for _ in range(3):
            assert np.frombuffer(raw[off:off + 4], np.int32)[0] == 4 * n
            xyz.append(np.frombuffer(raw[off + 4:off + 4 + 4 * n], np.float32))
            off += 8 + 4 * n
        frames.append(np.stack(xyz, 1))
    assert off == len(raw)
    return frames


def test_pmrun_water216_leapfrog_matches_oracle(tmp_path, port):
    tmp = str(tmp_path)
    run_pmrun(tmp, "--numsteps", "40", "--outputfreq", "10", "--dcdFile", "traj.dcd", "--precision", "double")
    rows, v_end, frames = oracle_run(port, 40, 10)
    with open(os.path.join(tmp, "water.216.energies.lf")) as f:
        lines = f.readlines()
    assert len(lines) == 5                     # steps 0, 10, 20, 30 and the final output at 40
    got = np.array([[float(w) for w in l.split()] for l in lines])
    say("pmrun first line:", lines[0].rstrip())
    say("oracle first line:", front.energies_line(rows[0]).rstrip())
    np.testing.assert_allclose(got[:, 0], rows[:, 0], atol=1e-9)
    # printed to 8 significant digits: within one unit of the 8th digit
    np.testing.assert_allclose(got[:, 1:], rows[:, 1:], rtol=1.2e-7, atol=1e-7)
    same = sum(l == front.energies_line(r) for l, r in zip(lines, rows))
    say("lines identical to the oracle's as printed:", same, "of", len(lines))
    assert lines[0] == front.energies_line(rows[0])
    with open(os.path.join(tmp, "water.216.energies.lf.header")) as f:
        assert f.read().split() == ["Time(fs)", "E_potential", "E_kinetic", "E_total", "Temperature", "E_bond",
                                    "E_angle", "E_dihedral", "E_improper", "E_VdW", "E_coulomb", "E_other",
                                    "Volume(A^3)"]
    vel = front.read_coords(os.path.join(tmp, "water.216.fin.vel.lf.xyz"))
    say("max |dv|", np.abs(vel - v_end).max(), "of", np.abs(v_end).max())
    np.testing.assert_allclose(vel, v_end, atol=1e-9 * np.abs(v_end).max(), rtol=0)
    pdb = front.read_coords(os.path.join(tmp, "water.216.fin.lf.pdb"))
    assert np.abs(pdb - frames[-1]).max() < 1e-3       # 7 printed characters per coordinate
    dcd = read_dcd(os.path.join(tmp, "traj.dcd"), 648)
    assert len(dcd) == 5
    for a, b in zip(dcd, frames):
        np.testing.assert_allclose(a, b.astype(np.float32), atol=1e-4)


def test_pmrun_reports_unsupported(tmp_path):
    tmp = str(tmp_path)
    for f in FILES:
        shutil.copy(os.path.join(G, f), tmp)
    r = subprocess.run([PMRUN, "water.216.lf.conf", "--boundaryConditions", "Vacuum"], cwd=tmp,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "periodic" in r.stderr


REM_CONF = """
numsteps 20
NumSwitches 7
EqSwitches 1
REMTemperatureFile temps.txt
REMExchangeRatesFile rates.out
psfFile water.216.psf
parFile water.216.par
posFile water.216.equil.pdb
temperature 300
seed 7
boundaryConditions Periodic
cellBasisVector1 30.8927 0 0
cellBasisVector2 0 27.8485 0
cellBasisVector3 0 0 30.7217
cellorigin 0.98035 -0.08475 0.19285
cellManager Cubic
cellsize 7
exclude scaled1-4
Integrator {
  level 0 LangevinLeapfrog {
    timestep 1
    temperature 300
    gamma 5
    seed 7
    force Angle, Bond
    force LennardJones -algorithm NonbondedCutoff -switchingFunction C2 -switchon 11 -cutoff 13
    force Coulomb -algorithm NonbondedCutoff -switchingFunction C1 -cutoff 13
  }
}
"""
TEMPS = [300.0, 303.0, 306.0, 309.0, 312.0]


def test_pmrun_replica_exchange_rates(tmp_path, port):
    """REM mode against the oracle's restatement of rem.cpp's driver
    (applications/rem-app/rem.cpp:646-720): pmrun prints, for every attempt,
    the potential energies by ladder slot and the accepted flags; the oracle
    recomputes each flag from those energies (calcDelta via orc_rem_delta,
    the Metropolis test with the documented uniform), the parity of the
    attempted pairs, the first attempted round (EqSwitches) and the rates
    file's credits to the higher-temperature partner.  Accepted switches
    redraw velocities (doSwitch with redrawVelocities = 1)."""
    from oracle import pyoracle
    tmp = str(tmp_path)
    for f in FILES[1:4]:
        shutil.copy(os.path.join(G, f), tmp)
    open(os.path.join(tmp, "rem.conf"), "w").write(REM_CONF)
    open(os.path.join(tmp, "temps.txt"), "w").write("\n".join(str(t) for t in TEMPS) + "\n")
    say("case: pmrun REM water216 5 replicas 7 switches")
    r = subprocess.run([PMRUN, "rem.conf", "--precision", "double"], cwd=tmp, capture_output=True, text=True,
                       timeout=120)
    say(r.stdout[-2500:], r.stderr[-1500:])
    assert r.returncode == 0, r.stderr
    rates = np.loadtxt(os.path.join(tmp, "rates.out"))
    R, nsw, eq = len(TEMPS), 7, 1
    attempts, flags = {}, {}
    for line in r.stdout.splitlines():
        if not line.startswith("REM attempt"):
            continue
        w = line.split()
        i = int(w[2])
        attempts[i] = [float(x) for x in w[3:3 + R]]
        flags[i] = {k: int(f) for k, f in zip(range(i & 1, R - 1, 2), w[3 + R:])}
        assert len(w) == 3 + R + len(flags[i])
    assert sorted(attempts) == list(range(eq, nsw))      # one attempt per round from EqSwitches on
    want_flags, want_rates = pyoracle.rem_driver(port, TEMPS, attempts, 7, nsw, eq)
    say("flags", flags, "oracle", want_flags, "rates", rates, "oracle", want_rates)
    assert flags == want_flags
    assert rates[0] == 0
    assert sum(sum(f.values()) for f in flags.values()) > 0     # neighbours 3 K apart do exchange
    np.testing.assert_allclose(rates, want_rates, rtol=1e-5)      # printed with 6 significant digits
    # every replica's energy stays finite and in the liquid range after redraws
    assert all(np.isfinite(e).all() and max(e) < 0 for e in attempts.values())


PME_CONF = """
numsteps 10
outputfreq 10
allEnergiesFile pme.energies
psfFile water.216.psf
parFile water.216.par
posFile water.216.equil.pdb
velFile water.216.equil.vel.xyz
boundaryConditions Periodic
cellBasisVector1 30.8927 0 0
cellBasisVector2 0 27.8485 0
cellBasisVector3 0 0 30.7217
cellorigin 0.98035 -0.08475 0.19285
cellManager Cubic
cellsize 5
exclude scaled1-4
Integrator {
  level 0 Leapfrog {
    timestep 0.5
    force Bond, Angle
    force LennardJones CoulombEwaldReal -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff
      -switchon 9 -cutoff 10 -alpha ALPHA -cutoff 10 -cutoff 10
    force Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline
      -gridsize 32 30 32 -cutoff 10 -order 4 -accuracy 1e-6
  }
}
"""


def test_pmrun_pme_conf_first_line_matches_oracle(tmp_path, port):
    """The main workload's force set (LJ C2 + Ewald real in one cutoff force,
    PME reciprocal + correction, bonds, angles) written as a .conf for the
    216-water box: pmrun's energies at step 0 against the oracle."""
    from oracle import pyoracle as po
    tmp = str(tmp_path)
    for f in FILES[1:]:
        shutil.copy(os.path.join(G, f), tmp)
    alpha = port.pme_alpha(10.0, 1e-6)
    open(os.path.join(tmp, "pme.conf"), "w").write(PME_CONF.replace("ALPHA", repr(alpha)))
    say("case: pmrun PME conf water216 grid 32x30x32")
    r = subprocess.run([PMRUN, "pme.conf", "--precision", "double"], cwd=tmp, capture_output=True, text=True,
                       timeout=120)
    say(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0, r.stderr
    got = [float(w) for w in open(os.path.join(tmp, "pme.energies")).readline().split()]
    psf, par = (os.path.join(G, f) for f in FILES[1:3])
    t = front.load_system(psf, par)
    x = front.read_coords(os.path.join(G, FILES[3]))
    s = t.system(x, box=(30.8927, 27.8485, 30.7217), origin=(0.98035, -0.08475, 0.19285))
    terms = [po.term(po.POT_LJ, po.SWF_C2, 10.0, 9.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 10.0, alpha=alpha)]
    a = port.cutoff(s, terms, 10.0, 5.0)["energies"]
    b = port.pme(s, (32, 30, 32), 4, 10.0, 5.0, flags=6)["energies"]
    c = port.bonded(s, 3)["energies"]
    e = a + b + c
    say("pmrun", got[1], got[9], got[10], "oracle", e[:7].sum(), e[1], e[0])
    np.testing.assert_allclose(got[9], e[1], rtol=1.2e-7)          # E_VdW
    np.testing.assert_allclose(got[10], e[0], rtol=1.2e-7)         # E_coulomb: real + reciprocal + correction
    np.testing.assert_allclose(got[1], e[:7].sum(), rtol=1.2e-7)   # E_potential
