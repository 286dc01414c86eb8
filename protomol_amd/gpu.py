"""ctypes binding of the C ABI in include/pmgpu.h (libpmgpu.so).

The library is required: importing this module on a machine without the
built library, or creating a Context without a HIP device, raises.  There is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmgpu.so")

PM_OK, PM_ERR_ARGUMENT, PM_ERR_STATE, PM_ERR_DEVICE, PM_ERR_CAPACITY, PM_ERR_UNSUPPORTED, PM_ERR_NUMERIC = range(7)
PRECISION = {"mixed": 0, "double": 1}
PM_EVAL_VIRIAL = 1
PM_EVAL_MOLVIRIAL = 2
INTEGRATOR = {"leapfrog": 0, "langevin": 1}
ENERGY_COUNT = 16
COULOMB, LENNARDJONES, BOND, ANGLE, DIHEDRAL, IMPROPER = 0, 1, 2, 3, 4, 5
VIRIAL = 7

# exported symbols, kept in sync with include/pmgpu.h (tested)
EXPORTS = ["pm_abi_version", "pm_last_error", "pm_create", "pm_destroy", "pm_set_precision",
           "pm_set_topology", "pm_set_bonded", "pm_set_torsions", "pm_set_boundary", "pm_force_create", "pm_force_uncache",
           "pm_evaluate", "pm_cell_diagnostics", "pm_md_setup", "pm_md_set_state", "pm_md_run",
           "pm_md_get_state", "pm_last_stats", "pm_set_timing", "pm_get_timing", "pm_set_option", "pm_get_option",
           "pm_synchronize", "pm_batch_create", "pm_batch_destroy", "pm_batch_last_error", "pm_batch_run",
           "pm_batch_exchange", "pm_batch_ladder", "pm_set_lj_per_atom", "pm_hessian_vec", "pm_param_gradients",
           "pm_set_charges", "pm_lanczos", "pm_md_autotune", "pm_batch_set_mode", "pm_batch_info",
           "pm_md_set_constraints", "pm_md_constraint_count", "pm_shake", "pm_rattle",
           "pm_batch_set_exchange_velocities", "pm_shake_noalloc"]
OPTIONS = {"cluster_size": 0, "graph": 1, "list_skin": 2, "pme_fft": 3, "pme_priority": 4, "prune_skin": 5, "list_test": 6,
           "pme_spread": 7, "pair_lanes": 8}
TIMING = ["bin", "list", "pair", "pme", "correction", "bonded", "reduce", "integrate"]


class PmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[pm error {code}] {msg}")
        self.code = code
        self.msg = msg


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run `python -m protomol_amd.build`")
    lib = C.CDLL(path)
    D = C.POINTER(C.c_double)
    I = C.POINTER(C.c_int)
    L = C.POINTER(C.c_longlong)
    P = C.c_void_p
    lib.pm_abi_version.restype = C.c_int
    lib.pm_last_error.restype = C.c_char_p
    lib.pm_last_error.argtypes = [P]
    lib.pm_create.argtypes = [C.c_int, C.POINTER(P)]
    lib.pm_destroy.argtypes = [P]
    lib.pm_destroy.restype = None
    lib.pm_set_precision.argtypes = [P, C.c_int]
    lib.pm_set_topology.argtypes = [P, C.c_int, D, I, C.c_int, D, D, D, D, I, D, C.c_int, I, I, I, C.c_double]
    lib.pm_set_bonded.argtypes = [P, C.c_int, I, I, D, D, C.c_int, I, I, I, D, D, D, D]
    lib.pm_set_torsions.argtypes = [P, C.c_int, C.c_int, I, I, D, I, D]
    lib.pm_set_boundary.argtypes = [P, D, D, D, D, C.c_double]
    lib.pm_force_create.argtypes = [P, C.c_char_p, D, C.c_int, I]
    lib.pm_force_uncache.argtypes = [P, C.c_int]
    lib.pm_evaluate.argtypes = [P, I, C.c_int, D, D, D, C.c_int]
    lib.pm_cell_diagnostics.argtypes = [P, D, C.c_double, C.c_double, I, I, I, I, L, L]
    lib.pm_md_setup.argtypes = [P, I, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_ulonglong]
    lib.pm_md_set_state.argtypes = [P, D, D]
    lib.pm_md_run.argtypes = [P, C.c_int]
    lib.pm_md_get_state.argtypes = [P, D, D, D]
    lib.pm_last_stats.argtypes = [P, L, C.c_int]
    lib.pm_set_timing.argtypes = [P, C.c_uint]
    lib.pm_get_timing.argtypes = [P, D, L, C.c_int]
    lib.pm_set_option.argtypes = [P, C.c_int, C.c_int]
    lib.pm_get_option.argtypes = [P, C.c_int, I]
    lib.pm_synchronize.argtypes = [P]
    lib.pm_batch_create.argtypes = [C.POINTER(P), C.c_int, C.POINTER(P)]
    lib.pm_batch_destroy.argtypes = [P]
    lib.pm_batch_destroy.restype = None
    lib.pm_batch_last_error.argtypes = [P]
    lib.pm_batch_last_error.restype = C.c_char_p
    lib.pm_batch_run.argtypes = [P, C.c_int]
    lib.pm_batch_exchange.argtypes = [P, C.c_int, C.c_ulonglong, I, D]
    lib.pm_batch_ladder.argtypes = [P, I, D]
    lib.pm_set_lj_per_atom.argtypes = [P, D, D, D, D]
    lib.pm_hessian_vec.argtypes = [P, I, C.c_int, D, D, D]
    lib.pm_param_gradients.argtypes = [P, I, C.c_int, D, D, D, D]
    lib.pm_set_charges.argtypes = [P, D]
    lib.pm_md_autotune.argtypes = [P, C.c_int, I, I, D]
    lib.pm_batch_set_mode.argtypes = [P, C.c_int]
    lib.pm_batch_set_exchange_velocities.argtypes = [P, C.c_int]
    lib.pm_batch_info.argtypes = [P, I, C.c_char_p, C.c_int]
    lib.pm_lanczos.argtypes = [P, I, C.c_int, D, D, C.c_int, C.c_int, C.c_int, D, D, D, D]
    lib.pm_md_set_constraints.argtypes = [P, C.c_double, C.c_int, C.c_double, C.c_int]
    lib.pm_md_constraint_count.argtypes = [P, I]
    lib.pm_shake.argtypes = [P, D, D, D, C.c_double, C.c_double, C.c_int, I]
    lib.pm_rattle.argtypes = [P, D, D, C.c_double, C.c_double, C.c_int, I]
    lib.pm_shake_noalloc.argtypes = [P, D, D, D, C.c_double, C.c_double, C.c_int, I]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _lp(a):
    return a.ctypes.data_as(C.POINTER(C.c_longlong))


class Context:
    """One HIP device context: topology, boundary, forces, MD state."""

    def __init__(self, device: int = 0, precision: str = "mixed"):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.pm_create(device, C.byref(h))
        if rc != PM_OK:
            raise PmError(rc, f"pm_create(device={device}) failed (no HIP device?)")
        self.h = h
        self.n = 0
        self._keep = []
        self.set_precision(precision)

    def close(self):
        if self.h:
            self.lib.pm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != PM_OK:
            raise PmError(rc, self.lib.pm_last_error(self.h).decode())

    def set_precision(self, precision: str):
        self._check(self.lib.pm_set_precision(self.h, PRECISION[precision]))
        self.precision = precision

    def set_topology(self, s):
        q = np.ascontiguousarray(s.q, np.float64)
        t = np.ascontiguousarray(s.type, np.int32)
        A, B = np.ascontiguousarray(s.ljA, np.float64), np.ascontiguousarray(s.ljB, np.float64)
        A14, B14 = np.ascontiguousarray(s.ljA14, np.float64), np.ascontiguousarray(s.ljB14, np.float64)
        mol = np.ascontiguousarray(s.molecule, np.int32)
        mass = np.ascontiguousarray(s.mass, np.float64)
        ea, eb, ec = (np.ascontiguousarray(x, np.int32) for x in (s.ex_a, s.ex_b, s.ex_cls))
        self._check(self.lib.pm_set_topology(self.h, len(q), _dp(q), _ip(t), A.shape[0], _dp(A), _dp(B),
                                             _dp(A14), _dp(B14), _ip(mol), _dp(mass), len(ea), _ip(ea),
                                             _ip(eb), _ip(ec), float(s.csf)))
        self.n = len(q)
        if getattr(s, "bond_a", None) is not None and (len(s.bond_a) or len(s.angle_a)):
            self.set_bonded(s)
        for kind, name in enumerate(("dihedrals", "impropers")):
            t = getattr(s, name, None)
            if t is not None and len(t.atoms):
                self.set_torsions(kind, t)

    def set_bonded(self, s):
        ba, bb = (np.ascontiguousarray(x, np.int32) for x in (s.bond_a, s.bond_b))
        bk, br = (np.ascontiguousarray(x, np.float64) for x in (s.bond_k, s.bond_r0))
        aa, ab, ac = (np.ascontiguousarray(x, np.int32) for x in (s.angle_a, s.angle_b, s.angle_c))
        ak, at, au, ar = (np.ascontiguousarray(x, np.float64) for x in (s.angle_k, s.angle_t0, s.angle_kub, s.angle_rub))
        self._keep = [ba, bb, bk, br, aa, ab, ac, ak, at, au, ar]
        self._check(self.lib.pm_set_bonded(self.h, len(ba), _ip(ba), _ip(bb), _dp(bk), _dp(br), len(aa),
                                           _ip(aa), _ip(ab), _ip(ac), _dp(ak), _dp(at), _dp(au), _dp(ar)))

    def set_torsions(self, kind: int, t):
        """kind 0 dihedrals, 1 impropers; t has atoms (m, 4), term_row, k, period, phase."""
        at = np.ascontiguousarray(t.atoms, np.int32).reshape(-1)
        row, per = (np.ascontiguousarray(x, np.int32) for x in (t.term_row, t.period))
        k, ph = (np.ascontiguousarray(x, np.float64) for x in (t.k, t.phase))
        self._check(self.lib.pm_set_torsions(self.h, kind, len(at) // 4, _ip(at), _ip(row), _dp(k), _ip(per), _dp(ph)))

    def set_boundary(self, box, origin, cellsize):
        e1 = np.array([box[0], 0, 0], np.float64)
        e2 = np.array([0, box[1], 0], np.float64)
        e3 = np.array([0, 0, box[2]], np.float64)
        o = np.ascontiguousarray(origin, np.float64)
        self._check(self.lib.pm_set_boundary(self.h, _dp(e1), _dp(e2), _dp(e3), _dp(o), float(cellsize)))

    def force(self, force_id: str, values) -> int:
        v = np.ascontiguousarray(values, np.float64)
        h = C.c_int()
        self._check(self.lib.pm_force_create(self.h, force_id.encode(), _dp(v), len(v), C.byref(h)))
        return h.value

    def uncache(self, handle: int):
        self._check(self.lib.pm_force_uncache(self.h, handle))

    def evaluate(self, handles, positions, virial=False, forces=None, energies=None, molecular_virial=False):
        hs = np.ascontiguousarray(handles, np.int32)
        x = np.ascontiguousarray(positions, np.float64).reshape(-1)
        f = np.zeros(3 * self.n) if forces is None else forces.reshape(-1)
        e = np.zeros(ENERGY_COUNT) if energies is None else energies
        flags = (PM_EVAL_VIRIAL if virial else 0) | (PM_EVAL_MOLVIRIAL if molecular_virial else 0)
        self._check(self.lib.pm_evaluate(self.h, _ip(hs), len(hs), _dp(x), _dp(f), _dp(e), flags))
        return f.reshape(-1, 3), e

    def cell_diagnostics(self, positions, algorithm_cutoff, pair_cutoff):
        dims = np.zeros(3, np.int32)
        self._check(self.lib.pm_cell_diagnostics(self.h, None, 0.0, 0.0, _ip(dims), None, None, None, None, None))
        nc = int(np.prod(dims))
        x = np.ascontiguousarray(positions, np.float64).reshape(-1)
        cell = np.zeros(self.n, np.int32)
        head = np.zeros(nc, np.int32)
        nxt = np.zeros(self.n, np.int32)
        cand = np.zeros(nc, np.int64)
        inrc = np.zeros(nc, np.int64)
        self._check(self.lib.pm_cell_diagnostics(self.h, _dp(x), float(algorithm_cutoff), float(pair_cutoff),
                                                 _ip(dims), _ip(cell), _ip(head), _ip(nxt), _lp(cand), _lp(inrc)))
        return dict(dims=tuple(int(d) for d in dims), cell=cell, head=head, next=nxt, cand=cand, inrc=inrc)

    def md_setup(self, handles, integrator="leapfrog", dt=1.0, temperature=300.0, gamma=0.0, seed=0):
        hs = np.ascontiguousarray(handles, np.int32)
        self._check(self.lib.pm_md_setup(self.h, _ip(hs), len(hs), INTEGRATOR[integrator], float(dt),
                                         float(temperature), float(gamma), int(seed)))

    def md_set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        self._check(self.lib.pm_md_set_state(self.h, _dp(x), _dp(v)))

    def md_run(self, nsteps: int):
        self._check(self.lib.pm_md_run(self.h, int(nsteps)))

    def md_autotune(self, trial_steps=20):
        """Time the MD step over cluster sizes and list skins, keep the fastest
        options (the state is restored).  Returns (cluster_size, skin_milli_A,
        ms_per_step)."""
        m, sk, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.pm_md_autotune(self.h, int(trial_steps), C.byref(m), C.byref(sk), C.byref(ms)))
        return m.value, sk.value, ms.value

    def md_get_state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        e = np.zeros(ENERGY_COUNT)
        self._check(self.lib.pm_md_get_state(self.h, _dp(x), _dp(v), _dp(e)))
        return x, v, e

    def md_set_constraints(self, shake_eps=1e-5, shake_max_iter=30, rattle_eps=0.0, rattle_max_iter=0):
        """SHAKE after every drift and RATTLE after every step of md_run (an
        epsilon or max_iter of 0 turns a modifier off)."""
        self._check(self.lib.pm_md_set_constraints(self.h, float(shake_eps), int(shake_max_iter),
                                                   float(rattle_eps), int(rattle_max_iter)))

    def constraint_count(self):
        k = C.c_int()
        self._check(self.lib.pm_md_constraint_count(self.h, C.byref(k)))
        return k.value

    def shake(self, x, v, x_last, dt_fs, eps=1e-5, max_iter=30):
        """One SHAKE application: returns (x, v, sweeps) corrected copies."""
        x = np.array(x, np.float64, order="C").reshape(-1, 3)
        v = np.array(v, np.float64, order="C").reshape(-1, 3)
        xl = np.ascontiguousarray(x_last, np.float64).reshape(-1, 3)
        k = C.c_int()
        self._check(self.lib.pm_shake(self.h, _dp(x), _dp(v), _dp(xl), float(dt_fs), float(eps), int(max_iter),
                                      C.byref(k)))
        return x, v, k.value

    def rattle(self, x, v, dt_fs, eps=1e-5, max_iter=30):
        """One RATTLE application: returns (v, sweeps), a corrected copy."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        v = np.array(v, np.float64, order="C").reshape(-1, 3)
        k = C.c_int()
        self._check(self.lib.pm_rattle(self.h, _dp(x), _dp(v), float(dt_fs), float(eps), int(max_iter), C.byref(k)))
        return v, k.value

    def stats(self):
        s = np.zeros(6, np.int64)
        self._check(self.lib.pm_last_stats(self.h, _lp(s), 6))
        return dict(list_entries=int(s[0]), lanes=int(s[1]), clusters=int(s[2]), slots=int(s[3]),
                    list_builds=int(s[4]), list_prunes=int(s[5]))

    def set_timing(self, categories=()):
        mask = 0
        for c in categories:
            mask |= 1 << TIMING.index(c)
        self._check(self.lib.pm_set_timing(self.h, mask))

    def timing(self):
        ms = np.zeros(len(TIMING))
        cnt = np.zeros(len(TIMING), np.int64)
        self._check(self.lib.pm_get_timing(self.h, _dp(ms), _lp(cnt), len(TIMING)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(TIMING)}

    def set_lj_per_atom(self, sigma, epsilon, sigma14=None, epsilon14=None):
        """Per-atom LJ sigma/epsilon (CHARMM combination per pair); None clears."""
        if sigma is None:
            self._check(self.lib.pm_set_lj_per_atom(self.h, None, None, None, None))
            return
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (sigma, epsilon, sigma14, epsilon14)]
        self._keep_lj = arrs
        self._check(self.lib.pm_set_lj_per_atom(self.h, *[None if a is None else _dp(a) for a in arrs]))

    def hessian_vec(self, handles, positions, v):
        hs = np.ascontiguousarray(handles, np.int32)
        x = np.ascontiguousarray(positions, np.float64).reshape(-1)
        vv = np.ascontiguousarray(v, np.float64).reshape(-1)
        hv = np.zeros(3 * self.n)
        self._check(self.lib.pm_hessian_vec(self.h, _ip(hs), len(hs), _dp(x), _dp(vv), _dp(hv)))
        return hv.reshape(-1, 3)

    def param_gradients(self, handles, positions, lj=True):
        hs = np.ascontiguousarray(handles, np.int32)
        x = np.ascontiguousarray(positions, np.float64).reshape(-1)
        dq = np.zeros(self.n)
        ds = np.zeros(self.n) if lj else None
        de = np.zeros(self.n) if lj else None
        self._check(self.lib.pm_param_gradients(self.h, _ip(hs), len(hs), _dp(x), _dp(dq),
                                                _dp(ds) if lj else None, _dp(de) if lj else None))
        return dq, ds, de

    def set_charges(self, q):
        """Replace the scaled charges (double[n])."""
        qq = np.ascontiguousarray(q, np.float64).reshape(-1)
        if qq.size != self.n:
            raise ValueError(f"{qq.size} charges for {self.n} atoms")
        self._check(self.lib.pm_set_charges(self.h, _dp(qq)))

    def lanczos(self, handles, positions, v0, steps=30, nvec=0, mass_weighted=False):
        """`steps` Lanczos steps (full reorthogonalisation, on the device) on H or
        M^-1/2 H M^-1/2.  Returns (ritz_values[steps] ascending,
        ritz_vectors[nvec, n, 3] of the lowest, residual bounds[steps],
        alpha[steps], beta[steps])."""
        hs = np.ascontiguousarray(handles, np.int32)
        x = np.ascontiguousarray(positions, np.float64).reshape(-1)
        v = np.ascontiguousarray(v0, np.float64).reshape(-1)
        vals = np.zeros(steps)
        vecs = np.zeros((max(nvec, 1), 3 * self.n))
        res = np.zeros(steps)
        ab = np.zeros(2 * steps)
        self._check(self.lib.pm_lanczos(self.h, _ip(hs), len(hs), _dp(x), _dp(v), int(steps), int(bool(mass_weighted)),
                                        int(nvec), _dp(vals), _dp(vecs) if nvec else None, _dp(res), _dp(ab)))
        return vals, vecs[:nvec].reshape(nvec, self.n, 3), res, ab[:steps], ab[steps:]

    def set_option(self, name: str, value: int):
        self._check(self.lib.pm_set_option(self.h, OPTIONS[name], int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int()
        self._check(self.lib.pm_get_option(self.h, OPTIONS[name], C.byref(v)))
        return v.value

    def synchronize(self):
        self._check(self.lib.pm_synchronize(self.h))


class Batch:
    """Replicas advanced together (pm_batch_*): contexts set up for Langevin MD."""

    def __init__(self, contexts):
        self.lib = load_library()
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        h = C.c_void_p()
        rc = self.lib.pm_batch_create(arr, len(self.contexts), C.byref(h))
        if rc != PM_OK:
            raise PmError(rc, "pm_batch_create failed (every context needs pm_md_setup with Langevin)")
        self.h = h

    def _check(self, rc):
        if rc != PM_OK:
            raise PmError(rc, self.lib.pm_batch_last_error(self.h).decode())

    def run(self, nsteps: int):
        self._check(self.lib.pm_batch_run(self.h, int(nsteps)))

    def exchange(self, attempt: int, seed: int):
        n = len(self.contexts)
        acc = np.full(n, -1, np.int32)
        dl = np.full(n, np.nan)
        self._check(self.lib.pm_batch_exchange(self.h, int(attempt), int(seed), _ip(acc), _dp(dl)))
        return acc, dl

    def set_exchange_velocities(self, mode: str):
        """'rescale' (default): sqrt(T_new / T_old); 'keep': left for the caller
        to redraw (rem.cpp's doSwitch redraws them)."""
        self._check(self.lib.pm_batch_set_exchange_velocities(self.h, {"rescale": 0, "keep": 1}[mode]))

    def set_mode(self, mode: str):
        """"auto" (batched launches when the replicas allow them) or "chains"."""
        self._check(self.lib.pm_batch_set_mode(self.h, {"auto": 0, "chains": 1}[mode]))

    def info(self):
        """(batched, why): whether the last run used batched launches, and why not."""
        b = C.c_int()
        why = C.create_string_buffer(256)
        self._check(self.lib.pm_batch_info(self.h, C.byref(b), why, 256))
        return bool(b.value), why.value.decode()

    def ladder(self):
        n = len(self.contexts)
        lad = np.zeros(n, np.int32)
        t = np.zeros(n)
        self._check(self.lib.pm_batch_ladder(self.h, _ip(lad), _dp(t)))
        return lad, t

    def close(self):
        if self.h:
            self.lib.pm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
