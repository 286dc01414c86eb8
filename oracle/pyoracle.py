"""ctypes bindings for the parity checkers (TEST INFRASTRUCTURE ONLY).

  Oracle("port")      -> oracle/libpmoracle.so, the C restatement
  Oracle("reference") -> oracle/_ref/libpmref.so, the ProtoMol sources compiled
                         here (absent where the reference tree is absent)

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "libpmoracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libpmref.so")

POT_LJ, POT_COULOMB, POT_EWALDREAL = 0, 1, 2
SWF_CUTOFF, SWF_C1, SWF_C2 = 0, 1, 2
E_COULOMB, E_LJ, E_BOND, E_ANGLE, E_DIHEDRAL, E_IMPROPER = 0, 1, 2, 3, 4, 5
E_VIRIAL = 7
E_COUNT = 16


class Term(C.Structure):
    _fields_ = [("pot", C.c_int), ("swf", C.c_int), ("alpha", C.c_double),
                ("switchon", C.c_double), ("cutoff", C.c_double)]


def term(pot, swf, cutoff, switchon=0.0, alpha=0.0):
    return Term(pot, swf, alpha, switchon, cutoff)


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_longlong)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


def available(kind: str) -> bool:
    return os.path.exists(PORT_LIB if kind == "port" else REF_LIB)


class Oracle:
    def __init__(self, kind: str = "port"):
        self.kind = kind
        path = PORT_LIB if kind == "port" else REF_LIB
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle{' ref' if kind != 'port' else ''})")
        self.lib = C.CDLL(path)
        p = "orc_" if kind == "port" else "ref_"
        self._cells = getattr(self.lib, p + "cells")
        self._cutoff = getattr(self.lib, p + "cutoff")
        self._pme = getattr(self.lib, p + "pme")
        self._bspline = getattr(self.lib, p + "bspline")
        self._hess = getattr(self.lib, p + "hessian_vec")
        self._bonded = getattr(self.lib, p + "bonded")
        self._torsions = getattr(self.lib, p + "torsions")
        self._cutoff_omp = getattr(self.lib, "orc_cutoff_omp", None) if kind == "port" else None
        self._delta = getattr(self.lib, "orc_delta_list", None) if kind == "port" else None
        self._fft = getattr(self.lib, "orc_fft3d", None) if kind == "port" else None
        self._alpha = getattr(self.lib, "orc_pme_alpha", None) if kind == "port" else None
        if self._alpha is not None:
            self._alpha.restype = C.c_double
        rem = getattr(self.lib, "orc_rem_delta", None) if kind == "port" else None
        if rem is not None:
            rem.restype = C.c_double
        self._rem = rem

    # ---------------------------------------------------------------- cells
    def cells(self, sys, cellsize):
        pos, pp = _d(sys.pos)
        box, bp = _d(sys.box)
        org, op = _d(sys.origin)
        dims = (C.c_int * 3)()
        real = (C.c_double * 3)()
        nc = self._cells(sys.n, pp, bp, op, C.c_double(cellsize), dims, real, None, None, None, 0)
        nc = dims[0] * dims[1] * dims[2]
        cell = np.zeros(sys.n, np.int32)
        head = np.zeros(nc, np.int32)
        nxt = np.zeros(sys.n, np.int32)
        r = self._cells(sys.n, pp, bp, op, C.c_double(cellsize), dims, real,
                        cell.ctypes.data_as(_I), head.ctypes.data_as(_I), nxt.ctypes.data_as(_I), nc)
        assert r == nc, r
        return dict(dims=tuple(dims), realcell=tuple(real), cell=cell, head=head, next=nxt)

    def delta_list(self, sys, cellsize, cutoff):
        box, bp = _d(sys.box)
        org, op = _d(sys.origin)
        out = np.zeros(3 * 4096, np.int32)
        n = self._delta(bp, op, C.c_double(cellsize), C.c_double(cutoff), out.ctypes.data_as(_I), 4096)
        assert n >= 0
        return out[: 3 * n].reshape(-1, 3)

    # --------------------------------------------------------------- cutoff
    def _sysargs(self, sys):
        keep = []

        def d(a):
            a2, p = _d(a)
            keep.append(a2)
            return p

        def i(a):
            a2, p = _i(a)
            keep.append(a2)
            return p

        args = [sys.n, d(sys.pos), d(sys.q), i(sys.type), sys.ntypes, d(sys.ljA), d(sys.ljB),
                d(sys.ljA14), d(sys.ljB14), i(sys.molecule), len(sys.ex_a), i(sys.ex_a),
                i(sys.ex_b), i(sys.ex_cls), C.c_double(sys.csf), d(sys.box), d(sys.origin)]
        return args, keep

    def cutoff(self, sys, terms, algo_cutoff, cellsize, virial=False, counts=False):
        args, keep = self._sysargs(sys)
        tarr = (Term * len(terms))(*terms)
        f = np.zeros((sys.n, 3))
        e = np.zeros(E_COUNT)
        nc = None
        cand = inrc = None
        if counts:
            info = self.cells(sys, cellsize)
            nc = int(np.prod(info["dims"]))
            cand = np.zeros(nc, np.int64)
            inrc = np.zeros(nc, np.int64)
        r = self._cutoff(*args, C.c_double(cellsize), len(terms), tarr, C.c_double(algo_cutoff),
                         int(virial), f.ctypes.data_as(_D), e.ctypes.data_as(_D),
                         cand.ctypes.data_as(_L) if counts else None,
                         inrc.ctypes.data_as(_L) if counts else None)
        if r != 0:
            raise ValueError(f"{self.kind} cutoff failed ({r})")
        out = dict(forces=f, energies=e)
        if counts:
            out.update(cand=cand, inrc=inrc)
        return out

    def cutoff_omp(self, sys, terms, algo_cutoff, cellsize, threads=0):
        args, keep = self._sysargs(sys)
        tarr = (Term * len(terms))(*terms)
        f = np.zeros((sys.n, 3))
        e = np.zeros(E_COUNT)
        used = self._cutoff_omp(threads, *args, C.c_double(cellsize), len(terms), tarr,
                                C.c_double(algo_cutoff), f.ctypes.data_as(_D), e.ctypes.data_as(_D))
        return dict(forces=f, energies=e, threads=used)

    # ------------------------------------------------------------------ PME
    def pme(self, sys, grid, order, rc, cellsize, accuracy=1e-6, alpha=-1.0, flags=6,
            virial=False, want_grid=False):
        keep = []

        def d(a):
            a2, p = _d(a)
            keep.append(a2)
            return p

        def i(a):
            a2, p = _i(a)
            keep.append(a2)
            return p

        f = np.zeros((sys.n, 3))
        e = np.zeros(E_COUNT)
        common = [sys.n, d(sys.pos), d(sys.q), i(sys.molecule), len(sys.ex_a), i(sys.ex_a),
                  i(sys.ex_b), i(sys.ex_cls), C.c_double(sys.csf), d(sys.box), d(sys.origin),
                  C.c_double(cellsize), grid[0], grid[1], grid[2], order, C.c_double(rc),
                  C.c_double(accuracy), C.c_double(alpha), flags, int(virial),
                  f.ctypes.data_as(_D), e.ctypes.data_as(_D)]
        out = dict(forces=f, energies=e)
        if self.kind == "port":
            parts = np.zeros(6)
            qg = np.zeros(int(np.prod(grid))) if want_grid else None
            r = self._pme(*common, parts.ctypes.data_as(_D),
                          qg.ctypes.data_as(_D) if want_grid else None)
            out["parts"] = dict(zip(["real", "reciprocal", "intra", "self", "charged", "alpha"], parts))
            if want_grid:
                out["qgrid"] = qg.reshape(grid)
        else:
            r = self._pme(*common)
        if r != 0:
            raise ValueError(f"{self.kind} pme failed ({r})")
        return out

    def bspline(self, order, w):
        th = np.zeros(order)
        dth = np.zeros(order)
        self._bspline(order, C.c_double(w), th.ctypes.data_as(_D), dth.ctypes.data_as(_D))
        return th, dth

    def pme_alpha(self, rc, accuracy):
        return self._alpha(C.c_double(rc), C.c_double(accuracy))

    def fft3d(self, data, sign):
        a = np.ascontiguousarray(data, dtype=np.complex128).copy()
        nx, ny, nz = a.shape
        self._fft(a.ctypes.data_as(_D), nx, ny, nz, sign)
        return a

    def rem_delta(self, ti, tj, ei, ej):
        return self._rem(C.c_double(ti), C.c_double(tj), C.c_double(ei), C.c_double(ej))

    # --------------------------------------------------------------- bonded
    def bonded(self, sys, which=3, virial=False):
        keep = []

        def d(a):
            a2, p = _d(a)
            keep.append(a2)
            return p

        def i(a):
            a2, p = _i(a)
            keep.append(a2)
            return p

        f = np.zeros((sys.n, 3))
        e = np.zeros(E_COUNT)
        r = self._bonded(sys.n, d(sys.pos), d(sys.box), d(sys.origin), len(sys.bond_a), i(sys.bond_a),
                         i(sys.bond_b), d(sys.bond_k), d(sys.bond_r0), len(sys.angle_a), i(sys.angle_a),
                         i(sys.angle_b), i(sys.angle_c), d(sys.angle_k), d(sys.angle_t0), d(sys.angle_kub),
                         d(sys.angle_rub), which, int(virial), f.ctypes.data_as(_D), e.ctypes.data_as(_D))
        return dict(forces=f, energies=e)

    def torsions(self, sys, tors, improper, virial=False):
        """Dihedral (improper=False) or improper system force over tors
        (atoms (m, 4), term_row, k, period, phase in radians)."""
        g = (lambda name: getattr(tors, name)) if not isinstance(tors, dict) else tors.__getitem__
        pos, box, org = _d(sys.pos), _d(sys.box), _d(sys.origin)
        at, row, per = _i(g("atoms")), _i(g("term_row")), _i(g("period"))
        k, ph = _d(g("k")), _d(g("phase"))
        f = np.zeros((sys.n, 3))
        e = np.zeros(E_COUNT)
        self._torsions(sys.n, pos[1], box[1], org[1], len(at[0]), at[1], row[1], k[1], per[1], ph[1],
                       int(improper), int(virial), f.ctypes.data_as(_D), e.ctypes.data_as(_D))
        return dict(forces=f, energies=e)

    # -------------------------------------------------------------- Hessian
    def hessian_vec(self, sys, terms, v, cellsize=4.0):
        args, keep = self._sysargs(sys)
        tarr = (Term * len(terms))(*terms)
        v2, vp = _d(np.asarray(v).reshape(-1))
        hv = np.zeros(3 * sys.n)
        if self.kind == "port":
            r = self._hess(*args, len(terms), tarr, vp, hv.ctypes.data_as(_D))
        else:
            r = self._hess(*args, C.c_double(cellsize), len(terms), tarr, vp, hv.ctypes.data_as(_D))
        if r != 0:
            raise ValueError(f"{self.kind} hessian failed ({r})")
        return hv.reshape(-1, 3)

    # ------------------------------------------------------ SHAKE / RATTLE
    def constraint_list(self, sys):
        """buildRattleShakeBondConstraintList: (atom1, atom2, rest length)."""
        keep = []

        def d(a):
            a2, p = _d(a)
            keep.append(a2)
            return p

        def i(a):
            a2, p = _i(a)
            keep.append(a2)
            return p

        nb, na = len(sys.bond_a), len(sys.angle_a)
        c1 = np.zeros(nb + na, np.int32)
        c2 = np.zeros(nb + na, np.int32)
        r0 = np.zeros(nb + na)
        m = self.lib.orc_constraint_list(sys.n, d(sys.mass), nb, i(sys.bond_a), i(sys.bond_b), d(sys.bond_r0), na,
                                         i(sys.angle_a), i(sys.angle_b), i(sys.angle_c), d(sys.angle_t0),
                                         c1.ctypes.data_as(_I), c2.ctypes.data_as(_I), r0.ctypes.data_as(_D))
        return c1[:m], c2[:m], r0[:m]

    def shake(self, cons, mass, x, v, x_last, dt_fs, eps, max_iter):
        """ModifierShake::doExecute on copies: (x, v, x_last, sweeps)."""
        c1, c2, r0 = (np.ascontiguousarray(cons[0], np.int32), np.ascontiguousarray(cons[1], np.int32),
                      np.ascontiguousarray(cons[2], np.float64))
        m = np.ascontiguousarray(mass, np.float64)
        x = np.array(x, np.float64, order="C")
        v = np.array(v, np.float64, order="C")
        xl = np.array(x_last, np.float64, order="C")
        self.lib.orc_shake.argtypes = [C.c_int, _I, _I, _D, _D, C.c_double, C.c_double, C.c_int, _D, _D, _D]
        k = self.lib.orc_shake(len(c1), c1.ctypes.data_as(_I), c2.ctypes.data_as(_I), r0.ctypes.data_as(_D),
                               m.ctypes.data_as(_D), dt_fs, eps, max_iter, x.ctypes.data_as(_D),
                               v.ctypes.data_as(_D), xl.ctypes.data_as(_D))
        return x, v, xl, k

    def rattle(self, cons, mass, x, v, dt_fs, eps, max_iter):
        """ModifierRattle::doExecute on a copy of v: (v, sweeps)."""
        c1, c2 = np.ascontiguousarray(cons[0], np.int32), np.ascontiguousarray(cons[1], np.int32)
        m = np.ascontiguousarray(mass, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        v = np.array(v, np.float64, order="C")
        self.lib.orc_rattle.argtypes = [C.c_int, _I, _I, _D, C.c_double, C.c_double, C.c_int, _D, _D]
        k = self.lib.orc_rattle(len(c1), c1.ctypes.data_as(_I), c2.ctypes.data_as(_I), m.ctypes.data_as(_D),
                                dt_fs, eps, max_iter, x.ctypes.data_as(_D), v.ctypes.data_as(_D))
        return v, k


# ---------------------------------------------------------------- topology
_TOPO_INFO = ["atoms", "types", "bonds", "angles", "dihedrals", "dihedral_terms", "impropers",
              "improper_terms", "molecules", "dof", "par_dialect", "dihedral_mult_psf"]


def ref_topology(psf: str, par: str, exclude: str = "scaled1-4", dihedral_mult_psf: int = -1) -> dict:
    """The reference's PSFReader + PARReader + buildTopology (compiled here)
    on the two files: the arrays of the GenericTopology it builds."""
    lib = Oracle("reference").lib
    lib.ref_topo_load.restype = C.c_void_p
    lib.ref_topo_load.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, _I]
    for f in ("ref_topo_free", "ref_topo_atoms", "ref_topo_lj", "ref_topo_bonds", "ref_topo_angles",
              "ref_topo_torsions", "ref_topo_exclusion_class", "ref_topo_type_name", "ref_topo_remove_momentum",
              "ref_topo_random_velocity"):
        getattr(lib, f).argtypes = None
    info_a = np.zeros(12, np.int32)
    h = lib.ref_topo_load(psf.encode(), par.encode(), exclude.encode(), dihedral_mult_psf,
                          info_a.ctypes.data_as(_I))
    if not h:
        raise RuntimeError("reference readers failed on %s / %s" % (psf, par))
    h = C.c_void_p(h)
    info = dict(zip(_TOPO_INFO, (int(x) for x in info_a)))
    n, nt = info["atoms"], info["types"]
    out = {"info": info, "handle": h, "lib": lib}
    q, m = np.zeros(n), np.zeros(n)
    t, mol, wat = (np.zeros(n, np.int32) for _ in range(3))
    lib.ref_topo_atoms(h, _d(q)[1], _d(m)[1], _i(t)[1], _i(mol)[1], _i(wat)[1])
    out.update(q=q, mass=m, type=t, molecule=mol, water=wat)
    lj = [np.zeros((nt, nt)) for _ in range(4)]
    lib.ref_topo_lj(h, *[x.ctypes.data_as(_D) for x in lj])
    out.update(ljA=lj[0], ljB=lj[1], ljA14=lj[2], ljB14=lj[3])
    nb, na = info["bonds"], info["angles"]
    ba, bb = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    bk, br = np.zeros(nb), np.zeros(nb)
    lib.ref_topo_bonds(h, ba.ctypes.data_as(_I), bb.ctypes.data_as(_I), bk.ctypes.data_as(_D), br.ctypes.data_as(_D))
    out.update(bond_a=ba, bond_b=bb, bond_k=bk, bond_r0=br)
    aa, ab, ac = (np.zeros(na, np.int32) for _ in range(3))
    ak, at, au, ar = (np.zeros(na) for _ in range(4))
    lib.ref_topo_angles(h, *[x.ctypes.data_as(_I) for x in (aa, ab, ac)], *[x.ctypes.data_as(_D) for x in (ak, at, au, ar)])
    out.update(angle_a=aa, angle_b=ab, angle_c=ac, angle_k=ak, angle_t0=at, angle_kub=au, angle_rub=ar)
    for imp, (key, mk, uk) in enumerate([("dihedrals", "dihedrals", "dihedral_terms"),
                                          ("impropers", "impropers", "improper_terms")]):
        mm, uu = info[mk], info[uk]
        atoms, row = np.zeros((mm, 4), np.int32), np.zeros(mm + 1, np.int32)
        k, per, ph = np.zeros(uu), np.zeros(uu, np.int32), np.zeros(uu)
        lib.ref_topo_torsions(h, imp, atoms.ctypes.data_as(_I), row.ctypes.data_as(_I), k.ctypes.data_as(_D),
                              per.ctypes.data_as(_I), ph.ctypes.data_as(_D))
        out[key] = dict(atoms=atoms, term_row=row, k=k, period=per, phase=ph)
    names = []
    buf = C.create_string_buffer(256)
    for k in range(nt):
        lib.ref_topo_type_name(h, k, buf, 256)
        names.append(buf.value.decode())
    out["type_names"] = names
    return out


def ref_exclusion_class(topo: dict, a, b) -> np.ndarray:
    a, b = np.ascontiguousarray(a, np.int32), np.ascontiguousarray(b, np.int32)
    cls = np.zeros(len(a), np.int32)
    topo["lib"].ref_topo_exclusion_class(topo["handle"], len(a), a.ctypes.data_as(_I), b.ctypes.data_as(_I),
                                         cls.ctypes.data_as(_I))
    return cls


def ref_remove_momentum(topo: dict, pos, vel, linear=True, angular=True) -> np.ndarray:
    p = np.ascontiguousarray(pos, np.float64)
    v = np.array(vel, np.float64, copy=True, order="C")
    topo["lib"].ref_topo_remove_momentum(topo["handle"], p.ctypes.data_as(_D), v.ctypes.data_as(_D),
                                         int(linear), int(angular))
    return v


def ref_random_velocity(topo: dict, temperature: float, seed: int) -> np.ndarray:
    """randomVelocity of the reference build; its generator is seeded on the
    first call in a process only."""
    v = np.zeros((topo["info"]["atoms"], 3))
    topo["lib"].ref_topo_random_velocity(topo["handle"], C.c_double(temperature), C.c_uint(seed),
                                         v.ctypes.data_as(_D))
    return v


def ref_topology_free(topo: dict) -> None:
    if topo.get("handle"):
        topo["lib"].ref_topo_free(topo["handle"])
        topo["handle"] = None


# ------------------------------------------------------ replica-exchange driver
_M64 = (1 << 64) - 1


def _splitmix(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def rem_uniform(seed, attempt, k):
    """The exchange test's uniform deviate of attempt `attempt`, ladder pair
    (k, k+1), as pmgpu.h documents it (the reference draws drand48 here,
    rem.cpp:865)."""
    h = _splitmix((seed ^ _splitmix((attempt * 0x100000001B3 + 2 * k + 1) & _M64)) & _M64)
    return (h >> 11) * (1.0 / 9007199254740992.0)


def rem_driver(oracle, temps, attempts, seed, nswitches, eqswitches):
    """applications/rem-app/rem.cpp:646-720 over recorded attempts.

    temps: the ladder temperatures; attempts: {round: energies by ladder slot
    before that round's attempt}.  Every round i >= EqSwitches attempts the
    pairs (k, k+1) with k of i's parity (getNextAddress :949-957); delta is
    calcDelta (:941-946), accepted when delta <= 0 or u <= exp(-delta)
    (canDoEvenSwitch :841-880); an accepted switch is counted by the replica
    of the higher temperature (doEvenSwitch/doOddSwitch :783-838: the node
    whose partner address is lower records it) and the rates are the counts
    over NumSwitches - EqSwitches (:708-712).  Returns ({round: accepted flags
    by k}, rates by ladder slot)."""
    import math
    R = len(temps)
    credit = [0.0] * R
    flags = {}
    for i in range(eqswitches, nswitches):
        e = attempts[i]
        f = {}
        for k in range(i & 1, R - 1, 2):
            d = oracle.rem_delta(temps[k], temps[k + 1], e[k], e[k + 1])
            w = 1.0 if d <= 0 else math.exp(-d)
            f[k] = 1 if rem_uniform(seed, i, k) <= w else 0
            if f[k]:
                credit[k + 1] += 1.0
        flags[i] = f
    return flags, [c / (nswitches - eqswitches) for c in credit]
