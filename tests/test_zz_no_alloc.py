"""pm_shake_noalloc, the allocation-free variant of pm_shake, called through
the C ABI: on two calls in a row -- an 8,100-atom water box, then a
648-atom one on the same GPU (so the second call finds the first call's
contents in the shared work space beyond its own range) -- its positions,
velocities and sweep count equal pm_shake's bit for bit (the same FP64
kernels on the same inputs), and a work space request beyond the fixed size,
or a call before the constraint list exists, is refused with an error."""
import ctypes as C
import sys

import numpy as np
import pytest

from protomol_amd import gpu, water

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(ctx, x, v, xl, dt, eps, max_iter):
    lib = gpu.load_library()
    lib.pm_shake_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int)]
    x = np.array(x, np.float64, order="C").reshape(-1, 3)
    v = np.array(v, np.float64, order="C").reshape(-1, 3)
    xl = np.ascontiguousarray(xl, np.float64).reshape(-1, 3)
    k = C.c_int(-1)
    rc = lib.pm_shake_noalloc(ctx.h, gpu._dp(x), gpu._dp(v), gpu._dp(xl), float(dt), float(eps), int(max_iter),
                              C.byref(k))
    return rc, x, v, k.value


def _case(make, name, seed, amp=0.03):
    s = water.named(name, seed=seed)
    rng = np.random.default_rng(seed)
    xl = s.pos.copy()
    x = s.pos + amp * rng.standard_normal(s.pos.shape)
    c = make("double")
    c.set_topology(s)
    c.set_boundary(s.box, s.origin, 4.0)
    return s, c, x, s.vel.copy(), xl


def test_shake_noalloc_matches_shake(gpu_ctx_factory):
    made = []
    for name, seed, eps, mi in (("water8k", 5, 1e-10, 60), ("water648", 6, 1e-5, 30)):
        say(f"case: pm_shake_noalloc {name} eps {eps} max_iter {mi}")
        s, c, x, v, xl = _case(gpu_ctx_factory, name, seed)
        # no constraint list yet: refused (building it would allocate)
        rc, _, _, _ = _noalloc(c, x, v, xl, 1.0, eps, mi)
        assert rc == gpu.PM_ERR_STATE, rc
        assert c.constraint_count() == s.n      # builds the list (set-up, allocating)
        xr, vr, kr = c.shake(x, v, xl, 1.0, eps, mi)
        rc, xn, vn, kn = _noalloc(c, x, v, xl, 1.0, eps, mi)
        say("rc", rc, "sweeps", kn, kr, "max |dx|", np.abs(xn - xr).max(), "max |dv|", np.abs(vn - vr).max())
        assert rc == 0
        assert kn == kr and kr > 1
        np.testing.assert_array_equal(xn, xr)
        np.testing.assert_array_equal(vn, vr)
        made.append((c, x, v, xl))
    # beyond the fixed work space (8,100 atoms: 44 waves x 100,001 sweep
    # counts of error sums = 35 MB): refused, nothing written
    c, x, v, xl = made[0]
    rc, xn, _, kn = _noalloc(c, x, v, xl, 1.0, 1e-5, 100000)
    assert rc == gpu.PM_ERR_ARGUMENT and kn == -1
    np.testing.assert_array_equal(xn, x.reshape(-1, 3))
