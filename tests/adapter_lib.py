"""ctypes access to oracle/_ref/libpmadapter.so: the GpuSystemForce adapter
(protomol_amd/adapter/) driven through the reference's own force factory and
ForceGroup (test infrastructure, built where the reference tree exists)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libpmadapter.so")

# the periodic ids the adapter serves (registerForceExemplarsCutoff.cpp and
# registerForceExemplarsFastElectrostatic.cpp register each of them)
GPU_IDS = [
    "Coulomb -switchingFunction C1 -algorithm NonbondedCutoff",
    "Coulomb -switchingFunction C2 -algorithm NonbondedCutoff",
    "Coulomb -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "CoulombEwaldReal -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "LennardJones -switchingFunction C1 -algorithm NonbondedCutoff",
    "LennardJones -switchingFunction C2 -algorithm NonbondedCutoff",
    "LennardJones -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "LennardJones CoulombEwaldReal -switchingFunction C1 -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "LennardJones Coulomb -switchingFunction C2 -switchingFunction C1 -algorithm NonbondedCutoff",
    "LennardJones CoulombEwaldReal -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "LennardJones Coulomb -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "LennardJones CoulombEwaldReal -switchingFunction Cutoff -switchingFunction Cutoff -algorithm NonbondedCutoff",
    "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline",
    "Coulomb -algorithm PMEwald -real -interpolation BSpline",
    "Coulomb -algorithm PMEwald -real -correction -interpolation BSpline",
    "Coulomb -algorithm PMEwald -reciprocal -interpolation BSpline",
    "Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
    "Coulomb -algorithm PMEwald -correction -interpolation BSpline",
]


def available():
    return os.path.exists(LIB)


class Adapter:
    def __init__(self):
        lib = C.CDLL(LIB)
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)
        lib.adp_register.argtypes = [C.c_int]
        lib.adp_describe.argtypes = [C.c_char_p, I, I, C.c_char_p, C.c_int, C.c_char_p, C.c_int, I, C.c_char_p, C.c_int]
        lib.adp_make.argtypes = [C.c_char_p, D, C.c_int, I, C.c_char_p, C.c_int]
        lib.adp_evaluate.argtypes = [C.c_int, D, D, I, C.c_int, D, D, D, D, I, D, C.c_int, I, I, I, C.c_double, D, D,
                                     C.c_double, C.c_int, C.POINTER(C.c_char_p), I, D, C.c_int, D, D, I, C.c_char_p,
                                     C.c_int]
        lib.adp_reset.argtypes = []
        lib.adp_set_precision.argtypes = [C.c_int]
        self.lib = lib

    def register(self, gpu):
        return self.lib.adp_register(1 if gpu else 0)

    def describe(self, fid):
        g, npar, cnp = C.c_int(), C.c_int(), C.c_int()
        idb, cid, kw = C.create_string_buffer(512), C.create_string_buffer(512), C.create_string_buffer(128)
        ok = self.lib.adp_describe(fid.encode(), C.byref(g), C.byref(npar), idb, 512, cid, 512, C.byref(cnp), kw, 128)
        if not ok:
            return None
        return dict(gpu=bool(g.value), nparams=npar.value, id=idb.value.decode(), cpu_id=cid.value.decode(),
                    cpu_nparams=cnp.value, keyword=kw.value.decode())

    def make(self, fid, values):
        v = np.ascontiguousarray(values, np.float64)
        g = C.c_int(-1)
        err = C.create_string_buffer(4096)
        ok = self.lib.adp_make(fid.encode(), v.ctypes.data_as(C.POINTER(C.c_double)), len(v), C.byref(g), err, 4096)
        return bool(ok), g.value == 1, err.value.decode()

    def evaluate(self, s, cellsize, forces, virial=False):
        """forces: [(id, values)]; one ForceGroup::evaluateSystemForces."""
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)

        def d(a):
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            return a.ctypes.data_as(D)

        def i(a):
            a = np.ascontiguousarray(a, np.int32)
            keep.append(a)
            return a.ctypes.data_as(I)
        keep = []
        ids = (C.c_char_p * len(forces))(*[f[0].encode() for f in forces])
        off = np.cumsum([0] + [len(f[1]) for f in forces]).astype(np.int32)
        vals = np.concatenate([np.asarray(f[1], float) for f in forces]) if forces else np.zeros(1)
        out_f = np.zeros(3 * s.n)
        out_e = np.zeros(16)
        ng = C.c_int()
        err = C.create_string_buffer(4096)
        r = self.lib.adp_evaluate(s.n, d(s.pos), d(s.q), i(s.type), s.ntypes, d(s.ljA), d(s.ljB), d(s.ljA14),
                                  d(s.ljB14), i(s.molecule), d(s.mass), len(s.ex_a), i(s.ex_a), i(s.ex_b), i(s.ex_cls),
                                  float(s.csf), d(s.box), d(s.origin), float(cellsize), len(forces), ids, i(off),
                                  d(vals), 1 if virial else 0, out_f.ctypes.data_as(D), out_e.ctypes.data_as(D),
                                  C.byref(ng), err, 4096)
        if r != 0:
            raise RuntimeError(err.value.decode())
        return out_f.reshape(-1, 3), out_e, ng.value

    def reset(self):
        self.lib.adp_reset()

    def set_precision(self, precision):
        self.lib.adp_set_precision({"mixed": 0, "double": 1}[precision])
