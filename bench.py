"""Benchmark: ns/day of the 98,304-atom PME water box on one MI355X.

Workload (one "step" = one leapfrog MD step, forces re-evaluated every step):
  32,768 three-site waters (98,304 atoms) on a 32^3 lattice in a 99.3 A cube,
  seeded orientations and Maxwell-Boltzmann velocities at 300 K;
  LennardJones (C2 switching 9 -> 10 A) + CoulombEwaldReal (cutoff 10 A) in one
  NonbondedCutoff force, cell size 5 A;
  PME reciprocal + correction, grid 96^3, order 4, alpha from accuracy 1e-6;
  harmonic Bond and Angle terms holding the waters together (TIP3 flexible);
  leapfrog, dt = 1 fs.
The timed steps run in mixed precision (FP32 pair/PME arithmetic, 64-bit
fixed-point force accumulation); the same workload in FP64 is reported beside
it.  Positions and velocities are resident in HBM during the timed region.

Prints one JSON line (see the metric definition in the output).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DT_FS = 1.0
RC = 10.0
SWITCHON = 9.0
GRID = (96, 96, 96)
ORDER = 4
# cell size of the workload's NonbondedCutoff force (the reference path's
# cellsize parameter; pairs and forces do not depend on it): 4.0 A, 24 cells
# of 4.14 A per side of the 99.3 A box (cell 3.0 / 3.6 / 4.0 / 4.5 / 5.0 /
# 6.5 A: 382 / 390 / 400-401 / 395 / 387 / 373 ns/day)
CELL = 4.0
PEAK_FP32 = 157.3e12      # MI355X vector FP32, FLOP/s
PEAK_HBM = 8.0e12         # MI355X HBM3E, B/s
FLOPS_PER_PAIR = 69       # LJ-C2 + Ewald-real per in-cutoff pair (algorithmic count)
# HBM bytes per launch of the pair kernel at the default options: FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0 (KiB per dispatch from rocprofv3 --pmc, scaled by what the
# counters read for a 512 MB device copy), profiles/k_pair_jl_water98k_hbm_traffic.md
PAIR_TRAFFIC_BYTES = (28791.9 * 2.0 + 137466.0 * 1.0) * 1024


def build_system():
    from protomol_amd import water
    return water.named("water98k", seed=1)


def make_forces(ctx, alpha):
    h_nb = ctx.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff "
                     "-switchingFunction C2 -switchingFunction Cutoff",
                     [SWITCHON, RC, alpha, RC, RC])
    h_pme = ctx.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                      [GRID[0], GRID[1], GRID[2], RC, ORDER, 1e-6, -1.0])
    h_b = ctx.force("Bond", [])
    h_a = ctx.force("Angle", [])
    return [h_nb, h_pme, h_b, h_a]


def pme_alpha(rc, accuracy):
    # erfc bisection of the PME set-up (the same arithmetic as the library)
    from math import erfc
    alpha = 1.0
    while erfc(alpha * rc) / rc >= accuracy:
        alpha *= 2.0
    lo, hi = 0.0, alpha
    for _ in range(100):
        alpha = 0.5 * (lo + hi)
        if erfc(alpha * rc) / rc >= accuracy:
            lo = alpha
        else:
            hi = alpha
    return alpha


# library options given on the command line (--opt name=value), applied to
# every context of the main workload
EXTRA_OPTIONS = {}


def run_md(s, precision, steps, warmup, timing=(), graph=True, skin=-1, cluster=0, autotune=False):
    from protomol_amd import gpu
    ctx = gpu.Context(0, precision)
    ctx.set_topology(s)
    ctx.set_boundary(s.box, s.origin, CELL)
    ctx.set_option("graph", 1 if graph else 0)
    ctx.set_option("list_skin", skin)
    ctx.set_option("cluster_size", cluster)
    for k, v in EXTRA_OPTIONS.items():
        ctx.set_option(k, v)
    hs = make_forces(ctx, pme_alpha(RC, 1e-6))
    ctx.md_setup(hs, "leapfrog", dt=DT_FS)
    ctx.md_set_state(s.pos, s.vel)
    tuned = None
    if autotune:
        # layout options chosen by timing short runs (the state is restored)
        tuned = ctx.md_autotune(20)
    ctx.md_run(max(warmup, 1))
    ctx.synchronize()
    ctx.set_timing(timing)
    builds0 = ctx.stats()["list_builds"]
    t0 = time.perf_counter()
    ctx.md_run(steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    builds = ctx.stats()["list_builds"] - builds0
    skin_used = ctx.get_option("list_skin")
    tim = ctx.timing()
    x, v, e = ctx.md_get_state()
    # in-cutoff, non-excluded pairs at the final positions (reference enumerator count)
    diag = ctx.cell_diagnostics(x, RC, RC)
    pairs = int(diag["inrc"].sum())
    stats = ctx.stats()
    cluster_used = ctx.get_option("cluster_size")
    ctx.close()
    return dict(seconds=t1 - t0, timing=tim, pairs=pairs, energies=e, x=x, v=v, stats=stats, builds=builds,
                skin=skin_used * 1e-3, skin_milli=skin_used, cluster=cluster_used, tuned=tuned)


def c3_bench(steps=200, warmup=20, accuracy=True):
    """The million-atom workload: 333,333 waters (999,999 atoms) in a 215.4 A
    cube, the force set of the main workload with PME 240^3 (order 4, rc 10),
    leapfrog 1 fs, mixed precision, library defaults; ns/day over `steps`
    steps after `warmup`, then a pass with HIP events around the pair and PME
    kernels (pair interactions/s against the FP32 VALU peak, PME compulsory
    bytes against HBM) and the force error against the FP64 oracle."""
    from protomol_amd import water
    s = water.named("water1m", seed=1)
    from protomol_amd import gpu
    grid = (240, 240, 240)

    def setup(timing=()):
        ctx = gpu.Context(0, "mixed")
        ctx.set_topology(s)
        ctx.set_boundary(s.box, s.origin, CELL)
        alpha = pme_alpha(RC, 1e-6)
        hs = [ctx.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff "
                        "-switchingFunction C2 -switchingFunction Cutoff", [SWITCHON, RC, alpha, RC, RC]),
              ctx.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                        [grid[0], grid[1], grid[2], RC, ORDER, 1e-6, -1.0]),
              ctx.force("Bond", []), ctx.force("Angle", [])]
        ctx.md_setup(hs, "leapfrog", dt=DT_FS)
        ctx.md_set_state(s.pos, s.vel)
        ctx.md_run(warmup)
        ctx.synchronize()
        ctx.set_timing(timing)
        return ctx
    ctx = setup()
    t0 = time.perf_counter()
    ctx.md_run(steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    builds = ctx.stats()["list_builds"]
    ctx.close()
    out = {"workload": "water1m: 999,999-atom TIP3 water, 215.4 A cube, LJ-C2(9-10 A) + Ewald-real rc 10 A, "
                       "PME 240^3 order 4, bonds+angles, leapfrog dt 1 fs", "atoms": int(s.n), "steps": steps,
           "ms_per_step": dt * 1e3 / steps, "ns_per_day": steps / dt * DT_FS * 86400 * 1e-6,
           "atom_steps_per_s": s.n * steps / dt, "list_builds": builds}
    # kernel-event pass (events launch kernels directly)
    ev_steps = max(20, steps // 4)
    ctx = setup(("pair", "pme"))
    ctx.md_run(ev_steps)
    ctx.synchronize()
    tim = ctx.timing()
    x, _, _ = ctx.md_get_state()
    pairs = int(ctx.cell_diagnostics(x, RC, RC)["inrc"].sum())
    ctx.close()
    pair_ms, pair_n = tim["pair"]
    pme_ms, pme_n = tim["pme"]
    pair_s = pair_ms * 1e-3 / max(pair_n, 1)
    pme_s = pme_ms * 1e-3 / max(pme_n, 1)
    ng = grid[0] * grid[1] * grid[2]
    ngc = grid[0] * grid[1] * (grid[2] // 2 + 1)
    n = s.n
    pme_bytes = (16 * n + 4 * ng) + (4 * ng + 8 * ngc) + (8 * ngc + 4 * ng) + (16 * n + 4 * ng + 24 * n)
    out.update({"pairs_per_step": pairs, "pair_us_per_step": pair_s * 1e6,
                "pair_interactions_per_s": pairs / pair_s,
                "pair_tflops": pairs * FLOPS_PER_PAIR / pair_s / 1e12,
                "valu_frac": pairs * FLOPS_PER_PAIR / pair_s / PEAK_FP32,
                "pme_us_per_step": pme_s * 1e6, "pme_grid_points_per_s": ng / pme_s,
                "pme_compulsory_bytes": pme_bytes, "pme_hbm_frac": (pme_bytes / pme_s) / PEAK_HBM})
    if accuracy:
        out["accuracy"] = accuracy_vs_oracle(s, x, pme_alpha(RC, 1e-6), grid, CELL)
    return out


def rigid_bench(steps=300, warmup=30, dt_fs=2.0, eps=1e-5, max_iter=30):
    """The main workload with rigid water: the same force set, SHAKE after
    every drift (protomol's default shakeEpsilon 1e-5, shakeMaxIter 30; every
    O-H bond plus the H-H distance of each water), leapfrog at 2 fs, mixed
    precision, library defaults."""
    from protomol_amd import gpu
    s = build_system()
    ctx = gpu.Context(0, "mixed")
    ctx.set_topology(s)
    ctx.set_boundary(s.box, s.origin, CELL)
    hs = make_forces(ctx, pme_alpha(RC, 1e-6))
    ctx.md_setup(hs, "leapfrog", dt=dt_fs)
    ctx.md_set_state(s.pos, s.vel)
    ctx.md_set_constraints(eps, max_iter, 0.0, 0)
    ctx.md_run(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.md_run(steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    x, _, _ = ctx.md_get_state()
    nc = ctx.constraint_count()
    ctx.close()
    # SHAKE's own error measure (ModifierMetaShake::calcError) at the end
    b = np.asarray(s.bond_r0)
    ba, bb = np.asarray(s.bond_a), np.asarray(s.bond_b)
    d = x[bb] - x[ba]
    err_b = np.abs((d * d).sum(1) - b * b) / (2 * b * b)
    return {"workload": "water98k rigid: the main workload with SHAKE (eps %g, maxIter %d; %d constraints), "
                        "leapfrog dt %g fs" % (eps, max_iter, nc, dt_fs),
            "steps": steps, "dt_fs": dt_fs, "ms_per_step": dt * 1e3 / steps,
            "ns_per_day": steps / dt * dt_fs * 86400 * 1e-6, "constraints": nc,
            "bond_error_mean_end": float(err_b.mean())}


EXAMPLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "examples")


def tile_system(s, t, reps, spacing):
    """reps^3 copies of a system (front.Topology t, system s) on a cubic
    lattice of the given spacing, in one periodic box."""
    from protomol_amd import front
    n = s.n
    k = reps ** 3
    g = np.stack(np.meshgrid(*[np.arange(reps)] * 3, indexing="ij"), -1).reshape(-1, 3) * spacing
    pos = (s.pos[None, :, :] - s.pos.mean(0) + g[:, None, :]).reshape(-1, 3)
    off = (np.arange(k) * n)[:, None]

    def idx(a):
        return (np.asarray(a)[None, :] + off).reshape(-1).astype(np.int32)

    def tors(T):
        m = len(T.atoms)
        u = len(T.k)
        row = (T.term_row[None, :-1] + (np.arange(k) * u)[:, None]).reshape(-1)
        return front.Torsions((T.atoms[None] + off[:, :, None]).reshape(-1, 4).astype(np.int32),
                              np.append(row, k * u).astype(np.int32), np.tile(T.k, k),
                              np.tile(T.period, k).astype(np.int32), np.tile(T.phase, k))
    big = t.system(pos, box=np.full(3, reps * spacing), origin=np.zeros(3))
    big.q, big.mass, big.type = np.tile(s.q, k), np.tile(s.mass, k), np.tile(s.type, k).astype(np.int32)
    big.molecule = (np.asarray(s.molecule)[None, :] + (np.arange(k) * (int(s.molecule.max()) + 1))[:, None]
                    ).reshape(-1).astype(np.int32)
    big.ex_a, big.ex_b, big.ex_cls = idx(s.ex_a), idx(s.ex_b), np.tile(s.ex_cls, k).astype(np.int32)
    big.bond_a, big.bond_b = idx(s.bond_a), idx(s.bond_b)
    big.bond_k, big.bond_r0 = np.tile(s.bond_k, k), np.tile(s.bond_r0, k)
    big.angle_a, big.angle_b, big.angle_c = idx(s.angle_a), idx(s.angle_b), idx(s.angle_c)
    for f in ("angle_k", "angle_t0", "angle_kub", "angle_rub"):
        setattr(big, f, np.tile(getattr(s, f), k))
    big.dihedrals, big.impropers = tors(t.dihedrals), tors(t.impropers)
    return big


def bonded_probe(evaluations=10):
    """Only the bonded evaluations of the 8x8x8 tiled WW domain (for
    rocprofv3 runs of the bonded kernels)."""
    from protomol_amd import front, gpu
    t = front.load_system(os.path.join(EXAMPLES, "wwd", "wwd_mineq.psf"),
                          os.path.join(EXAMPLES, "wwd", "par_all27_prot_lipid.inp"))
    x = front.read_coords(os.path.join(EXAMPLES, "wwd", "wwd.min.pdb"))
    big = tile_system(t.system(x, box=np.full(3, 80.0), origin=x.mean(0)), t, 8, 60.0)
    ctx = gpu.Context(0, "double")
    ctx.set_topology(big)
    ctx.set_boundary(big.box, big.origin, 12.0)
    hs = [ctx.force(k, []) for k in ("Bond", "Angle", "Dihedral", "Improper")]
    for _ in range(evaluations):
        _, e = ctx.evaluate(hs, big.pos)
    ctx.close()
    return {"workload": "bonded probe: 512 copies of the WW domain (%d atoms), %d evaluations" % (big.n, evaluations),
            "energies": [float(v) for v in e[2:6]]}


def conf_bench():
    """The reference's own example files through the front end: the whole
    examples/water_216/water.216.lf.conf run by pmrun (FP64, 5000 steps of
    0.1 fs), and one evaluation of the WW-domain protein's bonded terms
    (Bond, Angle, Dihedral, Improper; FP64) -- each beside the reference
    build's CPU time for the same force evaluation (one core)."""
    import re
    import shutil
    import subprocess
    import tempfile
    from protomol_amd import front, gpu
    from protomol_amd.build import PMRUN
    out = {}
    w = os.path.join(EXAMPLES, "water_216")
    with tempfile.TemporaryDirectory() as tmp:
        for f in os.listdir(w):
            shutil.copy(os.path.join(w, f), tmp)
        r = subprocess.run([PMRUN, "water.216.lf.conf"], cwd=tmp, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError("pmrun failed: " + r.stderr[-500:])
        m = re.search(r"Run: (\d+) steps of ([\d.]+) fs on (\d+) atoms in ([\d.]+) s \(([\d.]+) ms/step, ([\d.]+) ns/day",
                      r.stdout)
        lines = open(os.path.join(tmp, "water.216.energies.lf")).read().splitlines()
        etot = [float(l.split()[3]) for l in lines]
    out["water216_lf_conf"] = {
        "workload": "examples/water_216/water.216.lf.conf unchanged: 648 atoms, LJ C2 11-13 A + Coulomb C1 13 A "
                    "cutoffs, bonds + angles, leapfrog 0.1 fs, FP64",
        "steps": int(m.group(1)), "ms_per_step": float(m.group(5)), "ns_per_day": float(m.group(6)),
        "seconds": float(m.group(4)), "energy_lines": len(lines),
        "total_energy_range_kcal_mol": max(etot) - min(etot)}
    t = front.load_system(os.path.join(EXAMPLES, "wwd", "wwd_mineq.psf"),
                          os.path.join(EXAMPLES, "wwd", "par_all27_prot_lipid.inp"))
    x = front.read_coords(os.path.join(EXAMPLES, "wwd", "wwd.min.pdb"))
    s = t.system(x, box=np.full(3, 80.0), origin=x.mean(0))
    ctx = gpu.Context(0, "double")
    ctx.set_topology(s)
    ctx.set_boundary(s.box, s.origin, 8.0)
    hs = [ctx.force(k, []) for k in ("Bond", "Angle", "Dihedral", "Improper")]
    f = np.zeros((s.n, 3))
    e = np.zeros(16)
    for _ in range(5):
        ctx.evaluate(hs, s.pos, forces=f, energies=e)
    reps = 200
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.evaluate(hs, s.pos, forces=f, energies=e)
    gpu_ms = (time.perf_counter() - t0) * 1e3 / reps
    _, e1 = ctx.evaluate(hs, s.pos)
    ctx.close()
    out["wwd_bonded"] = {"workload": "WW domain (551 atoms, par_all27): Bond + Angle + Dihedral + Improper, one "
                                     "evaluation through pm_evaluate (host round trip included), FP64",
                         "ms_per_evaluation": gpu_ms,
                         "energies": {"bond": float(e1[2]), "angle": float(e1[3]), "dihedral": float(e1[4]),
                                      "improper": float(e1[5])}}
    # the bonded kernels at a size that fills the chip: 8x8x8 copies of the
    # protein 60 A apart (282,112 atoms, 760,320 dihedrals), kernel time from
    # HIP events around the bonded launches
    reps = 8
    big = tile_system(s, t, reps, 60.0)
    ctx = gpu.Context(0, "double")
    ctx.set_topology(big)
    ctx.set_boundary(big.box, big.origin, 12.0)
    hs = [ctx.force(k, []) for k in ("Bond", "Angle", "Dihedral", "Improper")]
    ctx.set_timing(("bonded",))
    for _ in range(3):
        ctx.evaluate(hs, big.pos)
    ctx.set_timing(("bonded",))
    nev = 20
    for _ in range(nev):
        ctx.evaluate(hs, big.pos)
    ms, cnt = ctx.timing()["bonded"]
    ctx.close()
    nt = {k: len(getattr(big, k)) for k in ("bond_a", "angle_a")}
    nd, ni = len(big.dihedrals.atoms), len(big.impropers.atoms)
    # algorithmic bytes of one force + energy evaluation: positions read and
    # forces written once per atom (FP64), every term's atoms and parameters
    # read once (bond 2 int + 2 double, angle 3 int + 4 double, torsion 4 int
    # + term row + per term double k, int n, double phase)
    nterm_d, nterm_i = len(big.dihedrals.k), len(big.impropers.k)
    algo = (big.n * 48 + nt["bond_a"] * 24 + nt["angle_a"] * 44 + (nd + ni) * 20 + (nterm_d + nterm_i) * 20)
    us = ms * 1e3 / nev
    out["protein_bonded_tiled"] = {
        "workload": "%d copies of the WW domain (%d atoms; %d bonds, %d angles, %d dihedrals, %d impropers), "
                    "Bond + Angle + Dihedral + Improper force and energy kernels, FP64" %
                    (reps ** 3, big.n, nt["bond_a"], nt["angle_a"], nd, ni),
        "us_per_evaluation": us, "atoms_per_s": big.n / (us * 1e-6),
        "algorithmic_bytes": algo, "achieved_GBps": algo / (us * 1e-6) / 1e9, "hbm_peak_GBps": 8000.0}
    try:
        from oracle import pyoracle as po
        kind = "reference" if po.available("reference") else "port"
        o = po.Oracle(kind)
        t0 = time.perf_counter()
        for _ in range(20):
            o.bonded(s, 3)
            o.torsions(s, t.dihedrals, 0)
            o.torsions(s, t.impropers, 1)
        out["wwd_bonded"]["cpu_ms_per_evaluation"] = (time.perf_counter() - t0) * 1e3 / 20
        ws = front.load_system(os.path.join(w, "water.216.psf"), os.path.join(w, "water.216.par"))
        xw = front.read_coords(os.path.join(w, "water.216.equil.pdb"))
        sw = ws.system(xw, box=(30.8927, 27.8485, 30.7217), origin=(0.98035, -0.08475, 0.19285))
        t0 = time.perf_counter()
        for _ in range(10):
            o.cutoff(sw, [po.term(po.POT_LJ, po.SWF_C2, 13.0, 11.0)], 13.0, 7.0)
            o.cutoff(sw, [po.term(po.POT_COULOMB, po.SWF_C1, 13.0)], 13.0, 7.0)
            o.bonded(sw, 3)
        cpu_ms = (time.perf_counter() - t0) * 1e3 / 10
        out["water216_lf_conf"]["cpu_ms_per_step"] = cpu_ms
        out["water216_lf_conf"]["cpu_ns_per_day"] = 0.1 * 1e-6 / (cpu_ms * 1e-3 / 86400)
        out["cpu_kind"] = kind
    except Exception as ex:  # informational only
        out["cpu_error"] = str(ex)
    return out


def force_errors(f, fr, mass, e, er):
    """Accuracy of forces f against the FP64 oracle's fr: the RMS relative
    force error sqrt(sum |F - F_ref|^2 / sum |F_ref|^2), and the statistics
    of the reference's CompareForce (forces/CompareForce.cpp:95-119): mass-
    weighted average and maximum relative errors and the absolute maximum;
    with the relative error of the potential energy (sum of the terms)."""
    f = np.asarray(f).reshape(-1, 3)
    fr = np.asarray(fr).reshape(-1, 3)
    d2 = ((f - fr) ** 2).sum(1)
    rm = 1.0 / np.asarray(mass)
    errf = np.sqrt(d2 * rm)
    fmag = np.sqrt((f ** 2).sum(1) * rm)
    tot = fmag.sum()
    return {"rms_rel_force": float(np.sqrt(d2.sum() / (fr ** 2).sum())),
            "rel_energy": float(abs(e - er) / abs(er)),
            "compareforce_rel_avg": float(errf.sum() / tot),
            "compareforce_rel_max": float(len(f) * errf.max() / tot),
            "compareforce_abs_max": float(np.sqrt(d2.max()))}


def oracle_threads():
    """Host threads of this run: the box's CPU share (OMP_NUM_THREADS where it
    is set), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def accuracy_vs_oracle(s, x, alpha, grid, cell, rc=RC, switchon=SWITCHON):
    """Forces and potential energy of the main force set (LJ-C2 + Ewald-real,
    PME reciprocal + correction, bonds + angles) at positions x through
    pm_evaluate in mixed and FP64 precision, against the FP64 oracle (the C
    restatement pinned to the compiled reference), outside any timed region."""
    from oracle import pyoracle as po
    from protomol_amd import gpu
    s2 = s.copy()
    s2.pos = np.ascontiguousarray(x)
    o = po.Oracle("port")
    terms = [po.term(po.POT_LJ, po.SWF_C2, rc, switchon), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, rc, alpha=alpha)]
    a = o.cutoff_omp(s2, terms, rc, cell, threads=oracle_threads())
    b = o.pme(s2, grid, ORDER, rc, cell, flags=6)
    c = o.bonded(s2, 3)
    fr = a["forces"] + b["forces"] + c["forces"]
    er = float((a["energies"] + b["energies"] + c["energies"])[:4].sum())
    out = {}
    for precision in ("mixed", "double"):
        ctx = gpu.Context(0, precision)
        ctx.set_topology(s2)
        ctx.set_boundary(s2.box, s2.origin, cell)
        hs = [ctx.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff "
                        "-switchingFunction C2 -switchingFunction Cutoff", [switchon, rc, alpha, rc, rc]),
              ctx.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                        [grid[0], grid[1], grid[2], rc, ORDER, 1e-6, -1.0]),
              ctx.force("Bond", []), ctx.force("Angle", [])]
        f, e = ctx.evaluate(hs, s2.pos)
        ctx.close()
        out["fp64" if precision == "double" else "mixed"] = force_errors(f, fr, s2.mass, float(e[:4].sum()), er)
    out["oracle"] = "FP64 C restatement (bit-exact with the compiled reference on the cutoff path, 1e-11 on PME)"
    return out


def dropin_bench(s, x, alpha, n=20, warmup=3):
    """The drop-in route: one pm_evaluate of the main force set per MD step,
    as ProtoMol's ForceGroup::evaluateSystemForces would call it through the
    GpuSystemForce adapter, with host position, force and energy buffers (the
    PCIe copies in and out included) and energies every call; at the end
    positions of the timed run, so the kept pair lists are reused between
    calls as on most MD steps."""
    from protomol_amd import gpu
    out = {"workload": "water98k: pm_evaluate of LJ-C2 + Ewald-real, PME 96^3 reciprocal + correction, bonds + "
                       "angles from host buffers, energies every call", "evaluations": n}
    xs = np.ascontiguousarray(x)
    for precision in ("mixed", "double"):
        ctx = gpu.Context(0, precision)
        ctx.set_topology(s)
        ctx.set_boundary(s.box, s.origin, CELL)
        hs = make_forces(ctx, alpha)
        f = np.zeros((s.n, 3))
        e = np.zeros(gpu.ENERGY_COUNT)
        for _ in range(warmup):
            ctx.evaluate(hs, xs, forces=f, energies=e)
        t0 = time.perf_counter()
        for _ in range(n):
            ctx.evaluate(hs, xs, forces=f, energies=e)
        dt = (time.perf_counter() - t0) / n
        ctx.close()
        out["fp64" if precision == "double" else "mixed"] = {
            "ms_per_evaluation": dt * 1e3, "ns_per_day_at_one_evaluation_per_step": DT_FS * 86400 * 1e-6 / dt}
    return out


def c1_bench(steps=100, warmup=10):
    """The plain-method workload: 216 waters (648 atoms) in an 18.62 A cube,
    LennardJones and Coulomb with C2 switching 7 -> 8 A (two NonbondedCutoff
    forces sharing one cell build), cell size 4, bonds + angles, leapfrog
    1 fs, FP64, 100 steps; the reference build's time for the same force
    evaluation on one host core beside it, and the force error of both
    precisions against the FP64 oracle."""
    from protomol_amd import gpu, water
    from oracle import pyoracle as po
    s = water.named("water648", seed=1, jitter=0.3)

    def forces(ctx):
        return [ctx.force("LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0]),
                ctx.force("Coulomb -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0]),
                ctx.force("Bond", []), ctx.force("Angle", [])]
    out = {"workload": "water648: 648-atom TIP3 water, 18.62 A cube, LJ + Coulomb C2 7-8 A (cell 4), bonds + "
                       "angles, leapfrog dt 1 fs", "atoms": int(s.n), "steps": steps}
    for precision in ("double", "mixed"):
        ctx = gpu.Context(0, precision)
        ctx.set_topology(s)
        ctx.set_boundary(s.box, s.origin, 4.0)
        hs = forces(ctx)
        ctx.md_setup(hs, "leapfrog", dt=DT_FS)
        ctx.md_set_state(s.pos, s.vel)
        ctx.md_run(warmup)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.md_run(steps)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        f, e = ctx.evaluate(hs[:2], s.pos)
        pairs = int(ctx.cell_diagnostics(s.pos, 8.0, 8.0)["inrc"].sum())
        ctx.close()
        key = "fp64" if precision == "double" else "mixed"
        out[key] = {"ms_per_step": dt * 1e3 / steps, "ns_per_day": steps / dt * DT_FS * 86400 * 1e-6,
                    "steps_per_s": steps / dt, "pairs_per_step": pairs}
        o = po.Oracle("port")
        a = o.cutoff(s, [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0)
        b = o.cutoff(s, [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0)
        out[key]["accuracy"] = force_errors(f, a["forces"] + b["forces"], s.mass, float(e[:2].sum()),
                                            float((a["energies"] + b["energies"])[:2].sum()))
    try:
        kind = "reference" if po.available("reference") else "port"
        o = po.Oracle(kind)
        t0 = time.perf_counter()
        reps = 20
        for _ in range(reps):
            o.cutoff(s, [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0)
            o.cutoff(s, [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)], 8.0, 4.0)
            o.bonded(s, 3)
        cpu = (time.perf_counter() - t0) / reps
        out["cpu_one_core"] = {"kind": kind, "ms_per_step": cpu * 1e3,
                               "ns_per_day": DT_FS * 86400 * 1e-6 / cpu}
    except Exception as ex:  # informational only
        out["cpu_error"] = str(ex)
    return out


def c5_bench(reps=20):
    """The differentiation workload: 2,700 waters (8,100 atoms) in a 43.2 A
    cube with seeded per-atom charge and LJ noise, LJ-C2 (8 -> 9 A, per-atom
    sigma/epsilon) + Ewald-real rc 9 and PME 48^3 order 4: Hessian-vector
    products per second (adjoint kernels, every term), the parameter
    gradients dE/dq, dE/dsigma, dE/deps, and 30 mass-weighted Lanczos steps
    with the 20 lowest Ritz vectors, in FP64 and mixed precision."""
    from protomol_amd import gpu, water
    s, sigma, eps = water.perturbed_box("water8k", seed=1)
    alpha = pme_alpha(9.0, 1e-6)
    out = {"workload": "water8k perturbed: 8,100 atoms, 43.2 A cube, LJ-C2 8-9 A (per-atom sigma/eps) + "
                       "Ewald-real rc 9 + PME 48^3 order 4", "atoms": int(s.n)}
    v = np.random.default_rng(5).standard_normal((s.n, 3))
    v /= np.linalg.norm(v)
    for precision in ("double", "mixed"):
        ctx = gpu.Context(0, precision)
        ctx.set_topology(s)
        ctx.set_boundary(s.box, s.origin, 4.5)
        ctx.set_lj_per_atom(sigma, eps)
        hs = [ctx.force("LennardJones CoulombEwaldReal -algorithm NonbondedCutoff -switchingFunction C2 "
                        "-switchingFunction Cutoff", [8.0, 9.0, alpha, 9.0, 9.0]),
              ctx.force("Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
                        [48, 48, 48, 9.0, 4, 1e-6, -1.0])]
        hv = ctx.hessian_vec(hs, s.pos, v)
        t0 = time.perf_counter()
        for _ in range(reps):
            hv = ctx.hessian_vec(hs, s.pos, v)
        t_hvp = (time.perf_counter() - t0) / reps
        ctx.param_gradients(hs, s.pos)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.param_gradients(hs, s.pos)
        t_grad = (time.perf_counter() - t0) / reps
        v0 = np.random.default_rng(10).standard_normal((s.n, 3))
        ctx.lanczos(hs, s.pos, v0, steps=30, nvec=20, mass_weighted=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        vals, _, res, _, _ = ctx.lanczos(hs, s.pos, v0, steps=30, nvec=20, mass_weighted=True)
        t_lz = time.perf_counter() - t0
        # central difference of the forces along v (h = 1e-4 A) as the check of H v
        h = 1e-4
        fp, _ = ctx.evaluate(hs, s.pos + h * v)
        fm, _ = ctx.evaluate(hs, s.pos - h * v)
        fd = -(fp - fm) / (2 * h)
        ctx.close()
        key = "fp64" if precision == "double" else "mixed"
        out[key] = {"hvp_per_s": 1.0 / t_hvp, "hvp_ms": t_hvp * 1e3, "param_gradients_ms": t_grad * 1e3,
                    "lanczos_30_steps_20_vectors_ms": t_lz * 1e3,
                    "lowest_ritz_values": [float(x) for x in vals[:3]],
                    "hvp_vs_central_difference_rms_rel": float(np.sqrt(((hv - fd) ** 2).sum() / (fd ** 2).sum()))}
    return out


def copy_probe(mb):
    """Ten device-to-device copies of `mb` megabytes: the kernel whose
    FETCH_SIZE / WRITE_SIZE counters calibrate the HBM-traffic runs."""
    import torch
    n = mb * 1000 * 1000 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "copies": 10}))


def nve_drift(s, precision, steps=1000, chunk=100, equilibrate=2000):
    """Total-energy drift of a leapfrog NVE run after `equilibrate` Langevin
    steps at 300 K (the lattice start relaxes violently over the first
    picosecond): E = kinetic (synchronous velocities after the closing half
    kick) + potential, sampled every `chunk` steps; the least-squares slope in
    kcal/mol per ns and per degree of freedom, and the RMS fluctuation."""
    from protomol_amd import gpu
    ctx = gpu.Context(0, precision)
    ctx.set_topology(s)
    ctx.set_boundary(s.box, s.origin, CELL)
    hs = make_forces(ctx, pme_alpha(RC, 1e-6))
    ctx.md_setup(hs, "langevin", dt=DT_FS, temperature=300.0, gamma=5.0, seed=2024)
    ctx.md_set_state(s.pos, s.vel)
    ctx.md_run(equilibrate)
    x, v, _ = ctx.md_get_state()
    ctx.md_setup(hs, "leapfrog", dt=DT_FS)
    ctx.md_set_state(x, v)
    t, e = [], []
    for k in range(steps // chunk + 1):
        ctx.md_run(chunk if k else 1)
        x, v, en = ctx.md_get_state()
        ke = 0.5 * float((s.mass[:, None] * v * v).sum())
        t.append((k * chunk + (0 if k else 1)) * DT_FS * 1e-6)   # ns
        e.append(ke + float(en[0] + en[1] + en[2] + en[3]))
    ctx.close()
    coef = np.polyfit(np.array(t), np.array(e), 1)
    resid = np.array(e) - np.polyval(coef, np.array(t))
    return {"steps": steps, "equilibration_steps": equilibrate,
            "kcal_mol_per_ns_per_dof": float(coef[0]) / (3 * s.n),
            "rms_fluctuation_kcal_mol": float(np.sqrt((resid ** 2).mean())),
            "e_start": e[0], "e_end": e[-1], "samples": len(e)}


def replica_bench(count=64, steps=100, segments=2, seed=0):
    """Batched replicas: `count` water boxes of 100-1000 molecules at 0.0334
    molecules/A^3, LJ + Coulomb C2 7 -> 8 A (no PME), bonds and angles, Langevin
    at a geometric ladder 300 -> 450 K, an exchange attempt every `steps` steps.
    Replica-steps/s of the batch (one launch per kernel over all replicas,
    replayed from a captured graph) against the batch as one captured chain
    per replica and against the same steps launched replica after replica."""
    from protomol_amd import gpu, water
    systems = water.replica_boxes(count, seed=seed)
    temps = water.replica_temperatures(count)
    ctxs = []
    for k, (s, t) in enumerate(zip(systems, temps)):
        c = gpu.Context(0, "mixed")
        c.set_topology(s)
        c.set_boundary(s.box, s.origin, 4.0)
        hs = [c.force("LennardJones -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0]),
              c.force("Coulomb -algorithm NonbondedCutoff -switchingFunction C2", [7.0, 8.0, 8.0]),
              c.force("Bond", []), c.force("Angle", [])]
        c.md_setup(hs, "langevin", dt=DT_FS, temperature=float(t), gamma=5.0, seed=k + 1)
        c.md_set_state(s.pos, s.vel)
        ctxs.append(c)
    def timed_batch(mode, first_attempt):
        b = gpu.Batch(ctxs)
        b.set_mode(mode)
        b.run(10)
        for c in ctxs:
            c.synchronize()
        acc_n = 0
        t0 = time.perf_counter()
        for seg in range(segments):
            b.run(steps)
            acc, _ = b.exchange(first_attempt + seg, 12345)
            acc_n += int((acc == 1).sum())
        for c in ctxs:
            c.synchronize()
        dt = time.perf_counter() - t0
        batched, why = b.info()
        b.close()
        return dt, acc_n, batched, why

    t_batch, accepted, batched, why = timed_batch("auto", 0)
    t_chain, _, _, _ = timed_batch("chains", segments)
    for c in ctxs:
        c.set_option("graph", 0)
    t0 = time.perf_counter()
    for seg in range(segments):
        for c in ctxs:
            c.md_run(steps)
    for c in ctxs:
        c.synchronize()
    t_seq = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    rs = count * steps * segments
    return {"replicas": count, "atoms": int(sum(s.n for s in systems)),
            "atoms_min": int(min(s.n for s in systems)), "atoms_max": int(max(s.n for s in systems)),
            "steps": steps * segments, "exchange_attempts": segments, "accepted_swaps": accepted,
            "replica_steps_per_s": rs / t_batch, "sequential_replica_steps_per_s": rs / t_seq,
            "speedup_vs_sequential": t_seq / t_batch,
            "batched_launches": batched, "not_batched_because": why or None,
            "chains_replica_steps_per_s": rs / t_chain}


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(s, alpha):
    """The CPU path beside the GPU (a reported baseline only), per the plan of
    timing the reference and its restatement on the same host:
      * t_ref: the reference itself (oracle/_ref, compiled from the ProtoMol
        sources; single-threaded like ProtoMol) for one force evaluation of
        the main force set (LJ-C2 + Ewald-real cutoff, PME 96^3 reciprocal +
        correction, bonds + angles) of the 98,304-atom box;
      * t_1, t_N: the C restatement of the same evaluation on one thread and
        with its pair loop on all N host threads (per-thread force buffers;
        its PME stays on one thread, as ZFFT is);
    value = the reference's all-core equivalent t_ref * t_N / t_1 in ns/day."""
    from oracle import pyoracle as po
    nthr = oracle_threads()
    terms = [po.term(po.POT_LJ, po.SWF_C2, RC, SWITCHON), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, RC, alpha=alpha)]

    def ns_day(sec):
        return (1.0 / sec) * DT_FS * 86400 * 1e-6
    p = po.Oracle("port")
    t0 = time.perf_counter()
    p.cutoff(s, terms, RC, CELL)
    p.pme(s, GRID, ORDER, RC, CELL, flags=6)
    p.bonded(s, 3)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = p.cutoff_omp(s, terms, RC, CELL, threads=nthr)
    p.pme(s, GRID, ORDER, RC, CELL, flags=6)
    p.bonded(s, 3)
    tn = time.perf_counter() - t0
    used = int(r["threads"])
    if po.available("reference"):
        o = po.Oracle("reference")
        t0 = time.perf_counter()
        o.cutoff(s, terms, RC, CELL)
        o.pme(s, GRID, ORDER, RC, CELL, flags=6)
        o.bonded(s, 3)
        tref = time.perf_counter() - t0
        kind = "reference"
        sample = ("one force evaluation of the 98,304-atom box (LJ-C2 + Ewald-real cutoff, PME 96^3 reciprocal + "
                  "correction, bonds + angles): the reference build on 1 core %.2f s, scaled by the C "
                  "restatement's speedup on %d threads (%.2f s -> %.2f s)" % (tref, used, t1, tn))
        value = ns_day(tref * tn / t1)
    else:
        tref = None
        kind = "port"
        sample = ("one force evaluation of the 98,304-atom box by the C restatement on %d threads, %.2f s "
                  "(%.2f s on one thread)" % (used, tn, t1))
        value = ns_day(tn)
    out = {"value": value, "unit": "ns/day", "cores": used, "kind": kind, "sample": sample,
           "reference_1core_ns_per_day": ns_day(tref) if tref else None,
           "port_1thread_ns_per_day": ns_day(t1), "port_all_threads_ns_per_day": ns_day(tn),
           "R_port_over_reference": (t1 / tref) if tref else None}
    out.update(cpu_info())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-drift", action="store_true", help="skip the 1000-step NVE drift runs")
    ap.add_argument("--no-replicas", action="store_true", help="skip the batched replica-exchange measurement")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: MD steps replay captured HIP graphs (default); 0: direct launches (for counter runs)")
    ap.add_argument("--list-skin", type=int, default=-1, metavar="MILLI_A",
                    help="pair-list skin in 1e-3 A (-1 automatic, 0 rebuild every step)")
    ap.add_argument("--no-c3", action="store_true", help="skip the million-atom measurement")
    ap.add_argument("--no-c1", action="store_true", help="skip the 648-atom plain-method measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the differentiation (HVP, Lanczos) measurement")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the force-error comparisons against the FP64 oracle")
    ap.add_argument("--no-rigid", action="store_true", help="skip the rigid-water (SHAKE, 2 fs) measurement")
    ap.add_argument("--no-conf", action="store_true", help="skip the example .conf / protein measurements")
    ap.add_argument("--bonded-probe", action="store_true",
                    help="only the tiled-protein bonded evaluations (for profiling the bonded kernels)")
    ap.add_argument("--only-replicas", action="store_true",
                    help="run only the batched replica-exchange measurement and print its record")
    ap.add_argument("--autotune", action="store_true",
                    help="let pm_md_autotune choose the layout options (default: the library's defaults)")
    ap.add_argument("--opt", action="append", default=[], metavar="NAME=VALUE",
                    help="a library option (pm_set_option) for the main workload's contexts, e.g. pme_priority=0")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run 10 device copies of MB megabytes (the calibration kernel of HBM counters) and exit")
    args = ap.parse_args()
    for o in args.opt:
        k, v = o.split("=", 1)
        EXTRA_OPTIONS[k] = int(v)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.only_replicas:
        print(json.dumps({"replica_exchange": replica_bench()}))
        return
    if args.bonded_probe:
        print(json.dumps(bonded_probe()))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X; --gpus {args.gpus} is not run "
              f"(this build measures the one-GPU share of a sharded run).", file=sys.stderr)
        sys.exit(2)
    steps, warmup = args.steps, args.warmup
    s = build_system()
    alpha = pme_alpha(RC, 1e-6)

    # the throughput pass: no instrumentation (captured graphs unless --graph 0)
    run = run_md(s, "mixed", steps, warmup, graph=bool(args.graph), skin=args.list_skin,
                 autotune=args.autotune)
    ms_step = run["seconds"] * 1e3 / steps
    ns_day = (steps / run["seconds"]) * DT_FS * 86400 * 1e-6
    # the roofline pass: the same steps with HIP events around the pair and PME
    # kernels on the streams they run on (event timing launches directly)
    mixed = run_md(s, "mixed", steps, warmup, timing=("pair", "pme"), skin=run["skin_milli"], cluster=run["cluster"])
    ms_step_events = mixed["seconds"] * 1e3 / steps

    fp64_steps = args.fp64_steps if args.fp64_steps is not None else max(10, steps // 10)
    fp64 = run_md(s, "double", fp64_steps, max(2, warmup // 10), skin=args.list_skin)
    ns_day_fp64 = (fp64_steps / fp64["seconds"]) * DT_FS * 86400 * 1e-6

    # roofline of the dominant kernel (the pair kernel), from live HIP events
    pair_ms, pair_n = mixed["timing"]["pair"]
    pair_avg_s = pair_ms * 1e-3 / max(pair_n, 1)
    flops = mixed["pairs"] * FLOPS_PER_PAIR
    achieved = flops / pair_avg_s
    pme_ms, pme_n = mixed["timing"]["pme"]
    pme_avg_s = pme_ms * 1e-3 / max(pme_n, 1)
    ng = GRID[0] * GRID[1] * GRID[2]
    ngc = GRID[0] * GRID[1] * (GRID[2] // 2 + 1)
    n = s.n
    # compulsory PME bytes per step (FP32 data): spread, R2C, convolution (fused: 0), C2R, gather
    pme_bytes = (16 * n + 4 * ng) + (4 * ng + 8 * ngc) + (8 * ngc + 4 * ng) + (16 * n + 4 * ng + 24 * n)
    line = {
        "metric": "ns/day @98k-atom PME water, mixed vs FP64 on GPU; replica-steps/s",
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_step,
        "ms_per_step_with_kernel_events": ms_step_events,
        "graph": bool(args.graph),
        "list_skin_A": run["skin"],
        "cluster_size": run["cluster"],
        "options": dict(EXTRA_OPTIONS),
        "autotuned": run["tuned"] is not None,
        "list_builds_per_step": run["builds"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "water98k: 98,304-atom TIP3 water, 99.3 A cube, LJ-C2(9-10 A) + Ewald-real rc 10 A, "
                               "PME 96^3 order 4, bonds+angles, leapfrog dt 1 fs",
                   "atoms": n, "pme_grid": list(GRID), "cutoff": RC, "cellsize": CELL, "dt_fs": DT_FS},
        "fp64_ns_per_day": ns_day_fp64,
        "fp64_steps": fp64_steps,
        "pair_interactions_per_s": mixed["pairs"] / pair_avg_s,
        "pairs_per_step": mixed["pairs"],
        "pme_grid_points_per_s": ng / pme_avg_s,
        "pme_us_per_step": pme_avg_s * 1e6,
        "pair_us_per_step": pair_avg_s * 1e6,
        "pme_hbm_frac": (pme_bytes / pme_avg_s) / PEAK_HBM,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": PAIR_TRAFFIC_BYTES if run["skin_milli"] == 2200 and run["cluster"] == 8 and not args.opt
                                else None,
                     "traffic_note": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated) "
                                     "at the default options (cell 4.0 A, list skin 2.2 A, exact list test, cluster 8); "
                                     "null for other options",
                     "kernel": "k_pair_jl (LJ-C2 + Ewald-real, %d flops/pair)" % FLOPS_PER_PAIR,
                     "note": "FP32 vector arithmetic (the kernel issues no MFMA); its time is set by the "
                             "64-bit fixed-point j-force atomics more than by VALU issue; "
                             "peak = MI355X vector FP32 157.3 TFLOP/s (counts packed FP32 dual issue, "
                             "which this build leaves off: the non-packed rate is 78.6)",
                     "frac_of_nonpacked_fp32": achieved / (PEAK_FP32 / 2)},
        "energies_end": {"coulomb": float(mixed["energies"][0]), "lj": float(mixed["energies"][1]),
                         "bond": float(mixed["energies"][2]), "angle": float(mixed["energies"][3])},
        "list_lanes_per_pair": mixed["stats"]["lanes"] / max(mixed["pairs"], 1),
    }
    if not args.no_accuracy:
        # the main workload's force error at the end of the timed run, outside the timed region
        line["accuracy"] = accuracy_vs_oracle(s, run["x"], alpha, GRID, CELL)
    line["dropin_evaluate"] = dropin_bench(s, run["x"], alpha)
    if not args.no_c1:
        line["c1_water648"] = c1_bench()
    if not args.no_replicas:
        line["replica_exchange"] = replica_bench()
    if not args.no_c3:
        line["c3_million_atoms"] = c3_bench(accuracy=not args.no_accuracy)
    if not args.no_c5:
        line["c5_differentiation"] = c5_bench()
    if not args.no_rigid:
        line["rigid_water_shake"] = rigid_bench()
    if not args.no_conf:
        line["reference_examples"] = conf_bench()
    if not args.no_drift:
        line["nve_drift"] = {"mixed": nve_drift(s, "mixed"), "fp64": nve_drift(s, "double")}
    if not args.no_cpu_baseline:
        cb = cpu_baseline(s, alpha)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu_baseline"] = ns_day / cb["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
