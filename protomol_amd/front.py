"""ctypes bindings of the host front end (libpmfront.so, include/pmfront.h).

load_system() reads a PSF and a PAR file into a Topology (the arrays the
device library takes); Topology.system() pairs it with positions,
velocities and a box as the object gpu.Context.set_topology() accepts.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmfront.so")

INFO_NAMES = ["atoms", "types", "bonds", "angles", "dihedrals", "dihedral_terms", "impropers",
              "improper_terms", "exclusions", "molecules", "dof", "par_dialect", "dihedral_mult_psf"]

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_lib = None


class FrontError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FrontError(f"{path} not built (python -m protomol_amd.build)")
    lib = C.CDLL(path)
    lib.pmf_last_error.restype = C.c_char_p
    lib.pmf_load_system.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
    lib.pmf_free_system.argtypes = [C.c_void_p]
    lib.pmf_free_system.restype = None
    lib.pmf_system_info.argtypes = [C.c_void_p, _I]
    lib.pmf_system_atoms.argtypes = [C.c_void_p, _D, _D, _I, _I]
    lib.pmf_system_lj.argtypes = [C.c_void_p, _D, _D, _D, _D]
    lib.pmf_system_bonds.argtypes = [C.c_void_p, _I, _I, _D, _D]
    lib.pmf_system_angles.argtypes = [C.c_void_p, _I, _I, _I, _D, _D, _D, _D]
    lib.pmf_system_torsions.argtypes = [C.c_void_p, C.c_int, _I, _I, _D, _I, _D]
    lib.pmf_system_exclusions.argtypes = [C.c_void_p, _I, _I, _I]
    lib.pmf_system_type_name.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
    lib.pmf_read_coords.argtypes = [C.c_char_p, _D, C.c_int, _I]
    lib.pmf_energies_line.argtypes = [_D, C.c_char_p, C.c_int]
    lib.pmf_write_xyz.argtypes = [C.c_char_p, C.c_void_p, _D, C.c_int, C.c_char_p]
    lib.pmf_write_pdb.argtypes = [C.c_char_p, C.c_char_p, _D, C.c_int, C.c_char_p]
    lib.pmf_write_dcd.argtypes = [C.c_char_p, C.c_int, C.c_int, _D, C.c_int, C.c_double, C.c_char_p]
    lib.pmf_force_values.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _D, C.c_int, _I]
    lib.pmf_remove_momentum.argtypes = [C.c_void_p, _D, _D, C.c_int, C.c_int]
    lib.pmf_random_velocity.argtypes = [C.c_void_p, C.c_double, C.c_uint, _D]
    _lib = lib
    return lib


def _check(lib, rc):
    if rc < 0:
        raise FrontError(lib.pmf_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


@dataclasses.dataclass
class Torsions:
    atoms: np.ndarray       # (m, 4) int32
    term_row: np.ndarray    # (m + 1,) int32
    k: np.ndarray
    period: np.ndarray      # int32
    phase: np.ndarray       # radians


@dataclasses.dataclass
class Topology:
    info: dict
    type_names: list
    q: np.ndarray
    mass: np.ndarray
    type: np.ndarray
    molecule: np.ndarray
    ljA: np.ndarray
    ljB: np.ndarray
    ljA14: np.ndarray
    ljB14: np.ndarray
    bond_a: np.ndarray
    bond_b: np.ndarray
    bond_k: np.ndarray
    bond_r0: np.ndarray
    angle_a: np.ndarray
    angle_b: np.ndarray
    angle_c: np.ndarray
    angle_k: np.ndarray
    angle_t0: np.ndarray
    angle_kub: np.ndarray
    angle_rub: np.ndarray
    dihedrals: Torsions
    impropers: Torsions
    ex_a: np.ndarray
    ex_b: np.ndarray
    ex_cls: np.ndarray

    @property
    def n(self) -> int:
        return int(self.q.shape[0])

    def system(self, pos, vel=None, box=(1000.0, 1000.0, 1000.0), origin=(0.0, 0.0, 0.0)):
        """The object gpu.Context.set_topology / set_bonded / set_torsions take."""
        from protomol_amd.water import WaterSystem
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        s = WaterSystem(pos=pos, vel=vel, q=self.q.copy(), type=self.type.copy(), mass=self.mass.copy(),
                        molecule=self.molecule.copy(), ex_a=self.ex_a.copy(), ex_b=self.ex_b.copy(),
                        ex_cls=self.ex_cls.copy(), ljA=self.ljA, ljB=self.ljB, ljA14=self.ljA14, ljB14=self.ljB14,
                        box=np.asarray(box, np.float64), origin=np.asarray(origin, np.float64), csf=1.0,
                        bond_a=self.bond_a, bond_b=self.bond_b, bond_k=self.bond_k, bond_r0=self.bond_r0,
                        angle_a=self.angle_a, angle_b=self.angle_b, angle_c=self.angle_c, angle_k=self.angle_k,
                        angle_t0=self.angle_t0, angle_kub=self.angle_kub, angle_rub=self.angle_rub)
        s.dihedrals = self.dihedrals
        s.impropers = self.impropers
        return s


def load_system(psf: str, par: str, exclude: str = "scaled1-4", dihedral_mult_psf: int = -1) -> Topology:
    lib = load_library()
    h = C.c_void_p()
    _check(lib, lib.pmf_load_system(psf.encode(), par.encode(), exclude.encode(), dihedral_mult_psf, C.byref(h)))
    try:
        info_a = np.zeros(len(INFO_NAMES), np.int32)
        _check(lib, lib.pmf_system_info(h, _ip(info_a)))
        info = dict(zip(INFO_NAMES, (int(x) for x in info_a)))
        n, nt = info["atoms"], info["types"]
        q, mass = np.zeros(n), np.zeros(n)
        typ, mol = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib, lib.pmf_system_atoms(h, _dp(q), _dp(mass), _ip(typ), _ip(mol)))
        lj = [np.zeros((nt, nt)) for _ in range(4)]
        _check(lib, lib.pmf_system_lj(h, *[_dp(x) for x in lj]))
        nb, na = info["bonds"], info["angles"]
        ba, bb = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        bk, br = np.zeros(nb), np.zeros(nb)
        _check(lib, lib.pmf_system_bonds(h, _ip(ba), _ip(bb), _dp(bk), _dp(br)))
        aa, ab, ac = (np.zeros(na, np.int32) for _ in range(3))
        ak, at, au, ar = (np.zeros(na) for _ in range(4))
        _check(lib, lib.pmf_system_angles(h, _ip(aa), _ip(ab), _ip(ac), _dp(ak), _dp(at), _dp(au), _dp(ar)))
        tors = []
        for imp, (m, u) in enumerate([(info["dihedrals"], info["dihedral_terms"]),
                                      (info["impropers"], info["improper_terms"])]):
            t = Torsions(np.zeros((m, 4), np.int32), np.zeros(m + 1, np.int32), np.zeros(u),
                         np.zeros(u, np.int32), np.zeros(u))
            _check(lib, lib.pmf_system_torsions(h, imp, _ip(t.atoms), _ip(t.term_row), _dp(t.k), _ip(t.period),
                                                _dp(t.phase)))
            tors.append(t)
        ne = info["exclusions"]
        ea, eb, ec = (np.zeros(ne, np.int32) for _ in range(3))
        _check(lib, lib.pmf_system_exclusions(h, _ip(ea), _ip(eb), _ip(ec)))
        names = []
        buf = C.create_string_buffer(256)
        for k in range(nt):
            _check(lib, lib.pmf_system_type_name(h, k, buf, 256))
            names.append(buf.value.decode())
    finally:
        lib.pmf_free_system(h)
    return Topology(info, names, q, mass, typ, mol, *lj, ba, bb, bk, br, aa, ab, ac, ak, at, au, ar,
                    tors[0], tors[1], ea, eb, ec)


def force_values(definition: str):
    """(force id, positional values) of a .conf force definition."""
    lib = load_library()
    buf = C.create_string_buffer(512)
    vals = np.zeros(32)
    n = C.c_int()
    _check(lib, lib.pmf_force_values(definition.encode(), buf, 512, _dp(vals), 32, C.byref(n)))
    return buf.value.decode(), vals[:n.value].copy()


def _with_system(psf, par, exclude, fn):
    lib = load_library()
    h = C.c_void_p()
    _check(lib, lib.pmf_load_system(psf.encode(), par.encode(), exclude.encode(), -1, C.byref(h)))
    try:
        return fn(lib, h)
    finally:
        lib.pmf_free_system(h)


def remove_momentum(psf: str, par: str, x, v, linear=True, angular=True, exclude="scaled1-4") -> np.ndarray:
    xx = np.ascontiguousarray(x, np.float64)
    vv = np.array(v, np.float64, copy=True, order="C")
    _with_system(psf, par, exclude,
                 lambda lib, h: _check(lib, lib.pmf_remove_momentum(h, _dp(xx), _dp(vv), int(linear), int(angular))))
    return vv


def random_velocity(psf: str, par: str, temperature: float, seed: int, exclude="scaled1-4") -> np.ndarray:
    def go(lib, h):
        n = C.c_int()
        info = np.zeros(len(INFO_NAMES), np.int32)
        _check(lib, lib.pmf_system_info(h, _ip(info)))
        v = np.zeros((int(info[0]), 3))
        _check(lib, lib.pmf_random_velocity(h, float(temperature), int(seed), _dp(v)))
        return v
    return _with_system(psf, par, exclude, go)


def read_coords(path: str) -> np.ndarray:
    lib = load_library()
    n = C.c_int()
    _check(lib, lib.pmf_read_coords(path.encode(), None, 0, C.byref(n)))
    x = np.zeros((n.value, 3))
    _check(lib, lib.pmf_read_coords(path.encode(), _dp(x), n.value, C.byref(n)))
    return x


def energies_line(values) -> str:
    lib = load_library()
    v = np.ascontiguousarray(values, np.float64)
    if v.shape != (13,):
        raise ValueError("energies_line takes 13 values")
    buf = C.create_string_buffer(512)
    _check(lib, lib.pmf_energies_line(_dp(v), buf, 512))
    return buf.value.decode()


def write_dcd(path: str, frames, first_step: int = 1, timestep_fs: float = 1.0, comment: str = "") -> None:
    lib = load_library()
    x = np.ascontiguousarray(frames, np.float64)
    nf, n = x.shape[0], x.shape[1]
    _check(lib, lib.pmf_write_dcd(path.encode(), n, nf, _dp(x), first_step, timestep_fs, comment.encode()))


def write_xyz(path: str, psf: str, par: str, xyz, comment: str = "") -> None:
    xx = np.ascontiguousarray(xyz, np.float64)
    _with_system(psf, par, "scaled1-4", lambda lib, h: _check(
        lib, lib.pmf_write_xyz(path.encode(), h, _dp(xx), xx.shape[0], comment.encode())))


def write_pdb(path: str, template_pdb: str, xyz, comment: str = "") -> None:
    lib = load_library()
    x = np.ascontiguousarray(xyz, np.float64)
    _check(lib, lib.pmf_write_pdb(path.encode(), template_pdb.encode(), _dp(x), x.shape[0], comment.encode()))
