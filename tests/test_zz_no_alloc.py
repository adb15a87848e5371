"""The allocation-free damping-coefficient entry point against the allocating one, through the C ABI:
the same result bit for bit on two calls in a row, the second with a smaller input (the shared
device array keeps the first call's contents), and sizes beyond its fixed array rejected."""
import ctypes as C

import numpy as np
import pytest

from bb_pd_amd import _lib

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T.reshape(-1)
                                                 if np.ndim(a) == 2 else a)).cuda()


def _arrays(n, seed):
    rng = np.random.default_rng(seed)
    F = np.asfortranarray(rng.standard_normal((n, 3)))
    vp = np.asfortranarray(rng.standard_normal((n, 3)) * (rng.uniform(size=(n, 3)) < 0.8))
    # F - Fp opposite to vp: every term of cn1 = -sum u^2 (F - Fp) / (m DT vp) is positive, so the
    # damping coefficient 2 sqrt(cn1 / cn2) is neither 0 nor the clamp
    Fp = np.asfortranarray(F + 0.1 * vp * rng.uniform(0.5, 1.5, (n, 3)))
    u = np.asfortranarray(rng.standard_normal((n, 3)) * 1e-3)
    m = rng.uniform(1, 2, n)
    return [_dev(a) for a in (F, m, Fp, vp, u)]


def _call(fn, arrs, n):
    import torch
    F, m, Fp, vp, u = arrs
    cn = C.c_double(-1.0)
    s = torch.cuda.current_stream().cuda_stream
    rc = fn(C.c_void_p(F.data_ptr()), C.c_void_p(m.data_ptr()), C.c_void_p(Fp.data_ptr()), 1.0,
            C.c_void_p(vp.data_ptr()), C.c_void_p(u.data_ptr()), n, C.byref(cn), C.c_void_p(s))
    return rc, cn.value


def test_noalloc_damping_equals_allocating(gpu):
    L = _lib.lib()
    for n, seed in ((300_000, 1), (70_001, 2)):
        print(f"case: damping coefficient, {n} nodes", flush=True)
        arrs = _arrays(n, seed)
        rc_a, cn_a = _call(L.bbpd_calculatedampingcoefficient, arrs, n)
        rc_b, cn_b = _call(L.bbpd_calculatedampingcoefficient_noalloc, arrs, n)
        assert rc_a == 0 and rc_b == 0, L.bbpd_last_error()
        assert 0 < cn_a < 1.9
        assert np.float64(cn_b).view(np.int64) == np.float64(cn_a).view(np.int64), (n, cn_a, cn_b)


def test_noalloc_damping_rejects_oversize(gpu):
    L = _lib.lib()
    cn = C.c_double(0.0)
    rc = L.bbpd_calculatedampingcoefficient_noalloc(None, None, None, 1.0, None, None,
                                                    _lib.NOALLOC_MAX_NODES + 1, C.byref(cn), None)
    assert rc == -1
    assert b"BBPD_NOALLOC_MAX_NODES" in L.bbpd_last_error()
