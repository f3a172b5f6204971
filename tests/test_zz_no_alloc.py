"""The allocation-free entry point pmg_mresid_mixed_noalloc: on device
arrays the caller owns, two calls in a row (the second on a smaller grid,
so the fixed scratch holds the first call's partials beyond the second's)
give exactly the residual and L1 norm of pmg_mresid_mixed, which stages
the same arrays itself; a grid beyond the fixed scratch is refused before
any launch.  Device memory comes from the HIP runtime the library itself
uses (hipMalloc through ctypes): a second HIP runtime in the process, such
as the one a framework bundles, must not be loaded beside it."""
import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
H2D, D2H = 1, 2


class Hip:
    def __init__(self, mg):
        mg.lib()  # the library's runtime is the one loaded under this soname
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.live = []

    def put(self, a):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    ac, cc = H.random_op_vec(n, 7, rng)
    m = np.zeros((n, n, n))
    m[1:-1, 1:-1, 1:-1] = 1
    f = rng.standard_normal(n ** 3) * m.ravel()
    x = rng.standard_normal(n ** 3)
    return ac, cc, f, x


def _staged(mg, n, ac, cc, f, x):
    r = np.full(n ** 3, np.nan, np.float32)  # boundary entries are left as given
    l1 = ctypes.c_double()
    rc = mg.lib().pmg_mresid_mixed(n, n, n, ac.ctypes.data_as(_D), cc.ctypes.data_as(_D),
                                   f.ctypes.data_as(_D), x.ctypes.data_as(_D),
                                   r.ctypes.data_as(_F), ctypes.byref(l1))
    assert rc == 0
    return r, l1.value


def test_noalloc_matches_staged_entry_point(mg):
    hip = Hip(mg)
    try:
        for n, seed in ((129, 1), (41, 2)):
            print(f"no-alloc residual n={n}", flush=True)
            ac, cc, f, x = _problem(n, seed)
            d = [hip.put(a) for a in (ac, cc, f, x)]
            r0 = np.full(n ** 3, np.nan, np.float32)
            l10 = np.full(1, np.nan)
            dr, dl1 = hip.put(r0), hip.put(l10)
            rc = mg.lib().pmg_mresid_mixed_noalloc(n, n, n, *d, dr, dl1, None)
            assert rc == 0
            r = hip.get(dr, r0)
            l1 = float(hip.get(dl1, l10)[0])
            r_ref, l1_ref = _staged(mg, n, ac, cc, f, x)
            assert np.array_equal(r, r_ref, equal_nan=True), n
            assert np.isfinite(r.reshape(n, n, n)[1:-1, 1:-1, 1:-1]).all()
            assert l1 == l1_ref, (n, l1, l1_ref)
    finally:
        hip.free()


def test_noalloc_refuses_beyond_scratch(mg):
    hip = Hip(mg)
    try:
        p = hip.put(np.zeros(16))
        # 1025^3 needs more block partials than the fixed scratch holds:
        # refused on the host, before any kernel could touch the arrays
        rc = mg.lib().pmg_mresid_mixed_noalloc(1025, 1025, 1025, p, p, p, p, p, p, None)
        assert rc == mg.PMG_ESCRATCH
        rc = mg.lib().pmg_mresid_mixed_noalloc(9, 9, 9, p, p, p, p, p, None, None)
        assert rc == mg.PMG_EBADPARM
    finally:
        hip.free()
