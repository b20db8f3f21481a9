"""chaos_dbscan_noalloc (the allocation-free clustering entry point) against
chaos_dbscan on the same device inputs: labels and cluster counts equal,
on two calls in a row, the second with a smaller input (the shared scratch
keeps the first call's contents), for the z-scored 3-D DBSCAN and for the
2-D single linkage with the fixed-point sign rule."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


@pytest.fixture(scope="module")
def lib(chaos):
    for name in ("chaos_dbscan", "chaos_dbscan_noalloc"):
        fn = getattr(chaos.lib, name)
        fn.restype = C.c_int
        fn.argtypes = [C.POINTER(chaos.Options), C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int,
                       C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p, _I]
    return chaos.lib


class Dev:
    """Device copies of host arrays (chaos_device_malloc / memcpy)."""

    def __init__(self, chaos):
        self.chaos, self.ptrs = chaos, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.chaos._check(self.chaos.lib.chaos_device_malloc(C.byref(p), max(a.nbytes, 8), 0))
        self.chaos._check(self.chaos.lib.chaos_memcpy_h2d(p, a.ctypes.data, a.nbytes))
        self.ptrs.append(p)
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        self.chaos._check(self.chaos.lib.chaos_memcpy_d2h(out.ctypes.data, p, out.nbytes))
        return out

    def free(self):
        for p in self.ptrs:
            self.chaos.lib.chaos_device_free(p)


def blobs(n, dim, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-3, 3, (6, dim))
    k = rng.integers(0, 7, n)
    pts = np.where((k < 6)[:, None], centres[np.minimum(k, 5)] + rng.normal(0, 0.25, (n, dim)),
                   rng.uniform(-4, 4, (n, dim)))
    return np.ascontiguousarray(pts)


def run_both(chaos, lib, pts, eps, min_pts, sign_rule, include):
    n, dim = pts.shape
    dev = Dev(chaos)
    cols = [dev.put(pts[:, d]) for d in range(dim)]
    z = cols[2] if dim == 3 else None
    inc = dev.put(include.astype(np.int32)) if include is not None else None
    lab_a, lab_b = dev.put(np.zeros(n, np.int32)), dev.put(np.zeros(n, np.int32))
    opt = chaos.make_options(device_buffers=True)
    ka, kb = C.c_int(-1), C.c_int(-1)
    chaos._check(lib.chaos_dbscan(C.byref(opt), cols[0], cols[1], z, n, dim, eps, min_pts, sign_rule, inc, lab_a,
                                  C.byref(ka)))
    chaos._check(lib.chaos_dbscan_noalloc(C.byref(opt), cols[0], cols[1], z, n, dim, eps, min_pts, sign_rule, inc,
                                          lab_b, C.byref(kb)))
    a, b = dev.get(lab_a, n, np.int32), dev.get(lab_b, n, np.int32)
    dev.free()
    return a, ka.value, b, kb.value


@pytest.mark.parametrize("n1,n2", [(300000, 40000)])
def test_noalloc_equals_dbscan_3d_twice(chaos, lib, n1, n2):
    for call, n in enumerate((n1, n2)):
        pts = blobs(n, 3, 5 + call)
        include = (np.random.default_rng(call).random(n) > 0.05).astype(np.int32)
        print(f"no-alloc 3-D call {call}: {n} points", flush=True)
        a, ka, b, kb = run_both(chaos, lib, pts, 0.2, 12, 0, include)
        assert ka == kb and ka >= 3
        assert np.array_equal(a, b)
        assert (a[include == 0] == 0).all()


def test_noalloc_equals_dbscan_sign_rule_twice(chaos, lib):
    for call, n in enumerate((120000, 9000)):
        pts = blobs(n, 2, 11 + call)
        rng = np.random.default_rng(21 + call)
        include = rng.choice([1, 1, 1, 1, 0, -1], n).astype(np.int32)
        print(f"no-alloc sign-rule call {call}: {n} points", flush=True)
        a, ka, b, kb = run_both(chaos, lib, pts, 0.05, 1, 1, include)
        assert ka == kb
        assert np.array_equal(a, b)
        assert (a < 0).any() and (a > 0).any()


def test_noalloc_rejects_more_points_than_its_scratch(chaos, lib):
    opt = chaos.make_options(device_buffers=True)
    k = C.c_int()
    # the size check comes before any device access, so the pointers are never read
    rc = lib.chaos_dbscan_noalloc(C.byref(opt), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), (1 << 20) + 1, 3,
                                  0.5, 20, 0, None, C.c_void_p(16), C.byref(k))
    assert rc == -5  # CHAOS_E_CAPACITY
