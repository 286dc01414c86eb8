"""Write the golden fixtures of tests/golden/ from the compiled reference.

Run in a tree where oracle/_ref/libpmref.so is built (make -C oracle ref,
which compiles the ProtoMol sources of the reference tree).  Each fixture
holds the inputs (positions and parameters) and the reference's outputs for
one evaluation; the tests replay the inputs through the C restatement and
through the GPU path.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from protomol_amd import water  # noqa: E402

ALPHA8 = 0.394670228244391


def sys_arrays(s):
    return dict(pos=s.pos, q=s.q, type=s.type, molecule=s.molecule, ex_a=s.ex_a, ex_b=s.ex_b,
                ex_cls=s.ex_cls, ljA=s.ljA, ljB=s.ljB, ljA14=s.ljA14, ljB14=s.ljB14, box=s.box,
                origin=s.origin, mass=s.mass, bond_a=s.bond_a, bond_b=s.bond_b, bond_k=s.bond_k,
                bond_r0=s.bond_r0, angle_a=s.angle_a, angle_b=s.angle_b, angle_c=s.angle_c,
                angle_k=s.angle_k, angle_t0=s.angle_t0, angle_kub=s.angle_kub, angle_rub=s.angle_rub)


def main():
    ref = po.Oracle("reference")
    s = water.named("water648", seed=11, jitter=0.4)
    rng = np.random.default_rng(11)
    s.pos = s.pos + rng.uniform(-0.05, 0.05, size=s.pos.shape)
    out = sys_arrays(s)
    c = ref.cells(s, 4.0)
    out.update(cell=c["cell"], head=c["head"], next=c["next"], dims=np.array(c["dims"]))
    cases = {
        "lj_c2": [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0)],
        "coul_c2": [po.term(po.POT_COULOMB, po.SWF_C2, 8.0, 7.0)],
        "coul_c1": [po.term(po.POT_COULOMB, po.SWF_C1, 8.0)],
        "lj_ewr": [po.term(po.POT_LJ, po.SWF_C2, 8.0, 7.0), po.term(po.POT_EWALDREAL, po.SWF_CUTOFF, 8.0, alpha=ALPHA8)],
    }
    for name, terms in cases.items():
        r = ref.cutoff(s, terms, 8.0, 4.0, virial=True, counts=True)
        out[f"{name}_forces"] = r["forces"]
        out[f"{name}_energies"] = r["energies"]
        out[f"{name}_cand"] = r["cand"]
        out[f"{name}_inrc"] = r["inrc"]
    for flags in (2, 4, 7):
        r = ref.pme(s, (24, 24, 24), 4, 8.0, 4.0, flags=flags, virial=True)
        out[f"pme{flags}_forces"] = r["forces"]
        out[f"pme{flags}_energies"] = r["energies"]
    r = ref.bonded(s, 3, virial=True)
    out["bonded_forces"] = r["forces"]
    out["bonded_energies"] = r["energies"]
    np.savez_compressed(os.path.join(HERE, "water648_reference.npz"), **out)

    # small system for the dense reference Hessian (integrators/Hessian.cpp)
    h = water.water_box(27, 9.4, seed=3, jitter=0.3)
    v = np.random.default_rng(5).standard_normal((h.n, 3))
    hout = sys_arrays(h)
    hout["v"] = v
    for name, terms in {"lj_c2": [po.term(po.POT_LJ, po.SWF_C2, 4.5, 3.5)],
                        "coul_c1": [po.term(po.POT_COULOMB, po.SWF_C1, 4.5)],
                        "coul_c2": [po.term(po.POT_COULOMB, po.SWF_C2, 4.5, 3.5)]}.items():
        hout[f"{name}_hv"] = ref.hessian_vec(h, terms, v, cellsize=2.0)
    np.savez_compressed(os.path.join(HERE, "water81_hessian_reference.npz"), **hout)
    print("wrote fixtures")


if __name__ == "__main__":
    main()
