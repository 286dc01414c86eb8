"""Synthetic rigid-geometry 3-site water boxes (the benchmark and test inputs).

The molecule, units and topology conventions follow ProtoMol:
  * charges are scaled by sqrt(332.0636) = 18.2226123264
    (frontend/buildTopology.cpp:148, base/pmconstants.cpp:72);
  * Lennard-Jones A/B tables use the CHARMM mixing rule
    A = (Ri+Rj)^12 sqrt(ei ej), B = 2 (Ri+Rj)^6 sqrt(ei ej) with R = Rmin/2
    (frontend/buildTopology.cpp:643-664);
  * the three intramolecular pairs of a TIP3 water are fully excluded
    (bond + angle under exclude scaled1-4, frontend/buildTopology.cpp:963-1011);
  * velocities are in A per AKMA time unit so that x += v * dt * INV_TIMEFACTOR
    (integrators/LeapfrogIntegrator.cpp:96-100).
Positions are generated from a seed: molecule oxygens on a cubic lattice in a
box centred on the origin, uniformly random orientations, Maxwell-Boltzmann
velocities.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SQRT_COULOMB = 18.2226123264
BOLTZMANN = 0.001987191
INV_TIMEFACTOR = 0.02045482828087295

EXCL_NONE, EXCL_MODIFIED, EXCL_FULL = 0, 1, 2

Q_O, Q_H = -0.834, 0.417
M_O, M_H = 15.9994, 1.008
R_OH = 0.9572
THETA_HOH = 104.52
SIGMA_O, EPS_O = 3.1507, 0.1521
# CHARMM TIP3 flexible terms (OT-HT bond, HT-OT-HT angle, no Urey-Bradley)
K_BOND, K_ANGLE = 450.0, 55.0


@dataclasses.dataclass
class WaterSystem:
    """Plain arrays describing one periodic system (AoS positions, Angstrom)."""

    pos: np.ndarray          # (n, 3) float64
    vel: np.ndarray          # (n, 3) float64
    q: np.ndarray            # (n,) scaled charges
    type: np.ndarray         # (n,) int32 atom type
    mass: np.ndarray         # (n,) amu
    molecule: np.ndarray     # (n,) int32
    ex_a: np.ndarray         # exclusion pairs (int32)
    ex_b: np.ndarray
    ex_cls: np.ndarray
    ljA: np.ndarray          # (T, T) float64
    ljB: np.ndarray
    ljA14: np.ndarray
    ljB14: np.ndarray
    box: np.ndarray          # (3,) box edges (orthogonal)
    origin: np.ndarray       # (3,) box centre
    csf: float = 1.0         # coulombScalingFactor
    # bonded terms (BondSystemForce / AngleSystemForce inputs)
    bond_a: np.ndarray = None
    bond_b: np.ndarray = None
    bond_k: np.ndarray = None
    bond_r0: np.ndarray = None
    angle_a: np.ndarray = None
    angle_b: np.ndarray = None
    angle_c: np.ndarray = None
    angle_k: np.ndarray = None
    angle_t0: np.ndarray = None
    angle_kub: np.ndarray = None
    angle_rub: np.ndarray = None

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def ntypes(self) -> int:
        return int(self.ljA.shape[0])

    def copy(self) -> "WaterSystem":
        return dataclasses.replace(self, **{f.name: (np.array(getattr(self, f.name), copy=True)
                                                     if isinstance(getattr(self, f.name), np.ndarray)
                                                     else getattr(self, f.name))
                                            for f in dataclasses.fields(self)})


def _lj_tables(sigma_eps):
    """CHARMM mixing: (Ri+Rj)^12 * sqrt(ei*ej) with R = sigma*2^(-5/6)."""
    t = len(sigma_eps)
    A = np.zeros((t, t))
    B = np.zeros((t, t))
    for i, (si, ei) in enumerate(sigma_eps):
        for j, (sj, ej) in enumerate(sigma_eps):
            r = si * 2.0 ** (-5.0 / 6.0) + sj * 2.0 ** (-5.0 / 6.0)
            e = math.sqrt(ei * ej)
            A[i, j] = r ** 12 * e
            B[i, j] = 2.0 * r ** 6 * e
    return A, B


def _random_rotations(rng, m):
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(m), rng.random(m), rng.random(m)
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    qw = a * np.sin(2 * np.pi * u2)
    qx = a * np.cos(2 * np.pi * u2)
    qy = b * np.sin(2 * np.pi * u3)
    qz = b * np.cos(2 * np.pi * u3)
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    R[:, 0, 1] = 2 * (qx * qy - qz * qw)
    R[:, 0, 2] = 2 * (qx * qz + qy * qw)
    R[:, 1, 0] = 2 * (qx * qy + qz * qw)
    R[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    R[:, 1, 2] = 2 * (qy * qz - qx * qw)
    R[:, 2, 0] = 2 * (qx * qz - qy * qw)
    R[:, 2, 1] = 2 * (qy * qz + qx * qw)
    R[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    return R


def water_box(n_molecules: int, box: float, seed: int = 0, temperature: float = 300.0,
              lattice: int | None = None, jitter: float = 0.0) -> WaterSystem:
    """Build a periodic cube of `n_molecules` waters with edge `box` (Angstrom).

    Oxygens sit on an `lattice`^3 cubic lattice (default: the smallest cube
    holding n_molecules sites; a seeded subset is used when it holds more),
    displaced by a seeded uniform offset in [-jitter, jitter] per axis.
    """
    rng = np.random.default_rng(seed)
    side = lattice or int(math.ceil(n_molecules ** (1.0 / 3.0) - 1e-9))
    while side ** 3 < n_molecules:
        side += 1
    nsites = side ** 3
    sites = np.arange(nsites)
    if nsites > n_molecules:
        sites = np.sort(rng.choice(nsites, size=n_molecules, replace=False))
    a = box / side
    ix, iy, iz = sites // (side * side), (sites // side) % side, sites % side
    centres = np.stack([ix, iy, iz], axis=1) * a + 0.5 * a - 0.5 * box
    if jitter > 0.0:
        centres = centres + rng.uniform(-jitter, jitter, size=centres.shape)
    half = math.radians(THETA_HOH) / 2.0
    local = np.array([[0.0, 0.0, 0.0],
                      [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                      [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]])
    rot = _random_rotations(rng, n_molecules)
    pos = centres[:, None, :] + np.einsum("mij,aj->mai", rot, local)
    pos = pos.reshape(-1, 3)
    n = 3 * n_molecules
    q = np.tile([Q_O, Q_H, Q_H], n_molecules) * SQRT_COULOMB
    typ = np.tile([0, 1, 1], n_molecules).astype(np.int32)
    mass = np.tile([M_O, M_H, M_H], n_molecules)
    mol = np.repeat(np.arange(n_molecules, dtype=np.int32), 3)
    base = 3 * np.arange(n_molecules, dtype=np.int32)
    ex_a = np.stack([base, base, base + 1], axis=1).reshape(-1).astype(np.int32)
    ex_b = np.stack([base + 1, base + 2, base + 2], axis=1).reshape(-1).astype(np.int32)
    ex_c = np.full(ex_a.shape, EXCL_FULL, dtype=np.int32)
    A, B = _lj_tables([(SIGMA_O, EPS_O), (0.0, 0.0)])
    vel = rng.standard_normal((n, 3)) * np.sqrt(BOLTZMANN * temperature / mass)[:, None]
    nb = 2 * n_molecules
    bond_a = np.stack([base, base], axis=1).reshape(-1).astype(np.int32)
    bond_b = np.stack([base + 1, base + 2], axis=1).reshape(-1).astype(np.int32)
    return WaterSystem(pos=np.ascontiguousarray(pos), vel=vel, q=q, type=typ, mass=mass,
                       molecule=mol, ex_a=ex_a, ex_b=ex_b, ex_cls=ex_c, ljA=A, ljB=B,
                       ljA14=A.copy(), ljB14=B.copy(), box=np.array([box, box, box], dtype=float),
                       origin=np.zeros(3),
                       bond_a=bond_a, bond_b=bond_b, bond_k=np.full(nb, K_BOND), bond_r0=np.full(nb, R_OH),
                       angle_a=(base + 1).astype(np.int32), angle_b=base.astype(np.int32),
                       angle_c=(base + 2).astype(np.int32), angle_k=np.full(n_molecules, K_ANGLE),
                       angle_t0=np.full(n_molecules, math.radians(THETA_HOH)),
                       angle_kub=np.zeros(n_molecules), angle_rub=np.zeros(n_molecules))


# Named workloads (molecule count, box edge in Angstrom).
CONFIGS = {
    "water648": (216, 18.62),            # 648 atoms, cutoff 8 with C2 from 7, cell 4
    "water98k": (32768, 99.3),           # 98,304 atoms, PME 96^3, rc 10
    "water1m": (333333, 215.4),          # 999,999 atoms, PME 240^3, rc 10
    "water8k": (2700, 43.2),             # 8,100 atoms, PME 48^3, rc 9
}


def named(name: str, seed: int = 0, jitter: float = 0.0) -> WaterSystem:
    nmol, box = CONFIGS[name]
    lattice = {"water1m": 70, "water8k": 14}.get(name)
    return water_box(nmol, box, seed=seed, lattice=lattice, jitter=jitter)


def replica_sizes(count: int = 64, seed: int = 0, lo: int = 100, hi: int = 1000):
    """Uneven replica sizes (molecules) and cube edges at 0.0334 molecules/A^3."""
    rng = np.random.default_rng(seed)
    n = rng.integers(lo, hi + 1, size=count)
    edges = (n / 0.0334) ** (1.0 / 3.0)
    return n, edges


def replica_temperatures(count: int = 64, lo: float = 300.0, hi: float = 450.0):
    """Geometric temperature ladder lo * (hi/lo)^(k/(count-1))."""
    if count == 1:
        return np.array([lo])
    return lo * (hi / lo) ** (np.arange(count) / (count - 1))


def replica_boxes(count: int = 64, seed: int = 0, lo: int = 100, hi: int = 1000, jitter: float = 0.2):
    """The uneven replica workload: `count` water boxes of seeded sizes (lo..hi
    molecules) at 0.0334 molecules/A^3, each generated from its own seed."""
    n, edges = replica_sizes(count, seed, lo, hi)
    return [water_box(int(m), float(L), seed=1000 * seed + k, jitter=jitter) for k, (m, L) in enumerate(zip(n, edges))]


def perturbed_box(name: str = "water8k", seed: int = 0, dq: float = 0.01, dlj: float = 0.02, jitter: float = 0.3):
    """A water box with seeded per-atom parameter noise: charges + N(0, dq) e,
    shifted back to a neutral total; LJ sigma and epsilon x (1 + dlj N(0, 1))
    per atom (zero parameters stay zero).  Returns the system and the per-atom
    sigma and epsilon arrays."""
    s = named(name, seed=seed, jitter=jitter)
    rng = np.random.default_rng(10_000 + seed)
    q = s.q / SQRT_COULOMB + rng.normal(0.0, dq, s.n)
    q -= q.mean()
    s.q = q * SQRT_COULOMB
    base_sigma = np.where(s.type == 0, SIGMA_O, 0.0)
    base_eps = np.where(s.type == 0, EPS_O, 0.0)
    sigma = base_sigma * (1.0 + dlj * rng.standard_normal(s.n))
    eps = base_eps * (1.0 + dlj * rng.standard_normal(s.n))
    return s, sigma, eps


def with_ion(s: WaterSystem, charge_e: float = 1.0, position=(0.0, 0.0, 0.0), mass: float = 22.99) -> WaterSystem:
    """A copy of `s` with one more atom: an ion of `charge_e` elementary
    charges (scaled as every charge) at `position`, in a molecule of its own,
    with no exclusion pairs, no bonded terms and the type of the water H (no
    Lennard-Jones).  The box then carries a net charge."""
    t = s.copy()
    t.pos = np.vstack([t.pos, np.asarray(position, float)[None, :]])
    t.vel = np.vstack([t.vel, np.zeros((1, 3))])
    t.q = np.append(t.q, charge_e * SQRT_COULOMB)
    t.type = np.append(t.type, np.int32(1)).astype(np.int32)
    t.mass = np.append(t.mass, mass)
    t.molecule = np.append(t.molecule, np.int32(int(t.molecule.max()) + 1)).astype(np.int32)
    return t


def per_atom_lj_tables(sigma, eps):
    """Type tables with one type per atom that reproduce per-atom CHARMM
    combination (for checkers on small systems)."""
    R = np.asarray(sigma) * 2.0 ** (-5.0 / 6.0)
    e = np.sqrt(np.asarray(eps))
    rs = R[:, None] + R[None, :]
    ee = e[:, None] * e[None, :]
    return rs ** 12 * ee, 2.0 * rs ** 6 * ee
