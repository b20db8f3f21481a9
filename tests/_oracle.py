"""ctypes bindings to oracle/liboracle.so (the CPU oracle; test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product (chaoslibrary_amd) never does.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

SYS = {"chameleon": 0, "rossler_cd": 1, "basins": 2, "basins_2": 3, "damir": 4,
       "rossler_rk4": 5, "lorenz_rk4": 6}
SEM_REFERENCE, SEM_CONFIG = 0, 1
CHK_BIF_REF, CHK_BASINS_REF, CHK_CONFIG = 0, 1, 2

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_dims": (C.c_int, [C.c_int, _I, _I]),
        "orc_sin": (C.c_double, [C.c_double]),
        "orc_cos": (C.c_double, [C.c_double]),
        "orc_exp": (C.c_double, [C.c_double]),
        "orc_atan": (C.c_double, [C.c_double]),
        "orc_step": (None, [C.c_int, _D, _D, C.c_double]),
        "orc_steps": (None, [C.c_int, _D, _D, C.c_double, C.c_longlong]),
        "orc_integrate": (None, [C.c_int, C.c_int, _D, _D, C.c_longlong, C.c_double]),
        "orc_grid_value": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double]),
        "orc_grid_value_pow": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]),
        "orc_loop": (C.c_int, [C.c_int, _D, _D, C.c_double, C.c_int, C.c_int, C.c_int,
                               C.c_double, C.c_int, _D]),
        "orc_diverged": (C.c_int, [C.c_double, C.c_double, C.c_int]),
        "orc_peak_finder": (C.c_int, [_D, C.c_int, _D, _D, C.c_double]),
        "orc_bif1d": (C.c_longlong, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D, C.c_int,
                                     _D, _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int,
                                     _I, _D, _D, _D, C.c_longlong]),
        "orc_bif2d": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D, C.c_int, _D,
                                _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int,
                                C.c_double, _I, C.c_int, C.c_int]),
        "orc_dbscan_ref": (C.c_int, [_D, _D, _D, C.c_int, C.c_int, C.c_double]),
        "orc_dbscan1d_count": (C.c_int, [_D, C.c_int, C.c_double]),
        "orc_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
        "orc_unit_vector": (None, [C.c_uint64, C.c_longlong, C.c_int, _D]),
        "orc_lle_point": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, _D, C.c_int, _D, C.c_double, C.c_uint64,
                                       C.c_longlong]),
        "orc_lle2d": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                C.c_double, _D, C.c_int, _D, _I, C.c_double, _D, C.c_int,
                                C.c_uint64, _D, C.c_int, C.c_int]),
        "orc_lle2d_omp": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                    C.c_double, _D, C.c_int, _D, _I, C.c_double, _D, C.c_int,
                                    C.c_uint64, _D, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_basins_features": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D,
                                          C.c_int, _D, _I, C.c_int, C.c_double, C.c_double, _D,
                                          C.c_int, C.c_int, _D, _D, _D, _I, C.c_int, C.c_int]),
        "orc_basins_cluster_ref": (C.c_int, [_D, _D, _I, C.c_int, C.c_double, _I]),
        "orc_det_sum": (C.c_double, [_D, _I, C.c_longlong]),
        "orc_zscore": (C.c_int, [_D, _D, _D, _I, C.c_longlong, _D]),
        "orc_dbscan_bruteforce": (C.c_int, [_D, C.c_int, C.c_int, C.c_double, C.c_int, _I, _I]),
        "orc_dbscan_grid": (C.c_int, [_D, C.c_int, C.c_int, C.c_double, C.c_int, _I, _I]),
        "orc_legacy_lle_point": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.c_double, C.c_double, _D, C.c_int, _D, C.c_double,
                                              C.c_double, C.c_uint64, C.c_longlong]),
        "orc_legacy_lle": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.c_double, C.c_double, _D, C.c_int, _D, _I, C.c_double, C.c_double,
                                     _D, C.c_int, C.c_uint64, _D, C.c_int, C.c_int]),
        "orc_gram_schmidt": (None, [_D, _D, C.c_int, _D]),
        "orc_legacy_ls_point": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                          _D, C.c_int, _D, C.c_double, C.c_double, C.c_uint64,
                                          C.c_longlong, _D]),
        "orc_legacy_ls": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                    C.c_double, _D, C.c_int, _D, _I, C.c_double, C.c_double, _D, C.c_int,
                                    C.c_uint64, _D, C.c_int, C.c_int]),
        "orc_time_domain": (C.c_longlong, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D, C.c_int,
                                           _D, _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int,
                                           _D, _I]),
        "orc_legacy_bif1d_ic": (C.c_longlong, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D, C.c_int,
                                               _D, _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int,
                                               _I, _D, C.c_longlong]),
        "orc_legacy_bif2d_ic": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _D, C.c_int, _D,
                                          _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, C.c_double,
                                          _I]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def dptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_I) if a is not None else None


def dims(sysid):
    nx, na = C.c_int(), C.c_int()
    assert lib.orc_dims(sysid, C.byref(nx), C.byref(na)) == 0
    return nx.value, na.value


def steps(sysid, x, a, h, n):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    a = np.ascontiguousarray(a, dtype=np.float64)
    lib.orc_steps(sysid, dptr(x), dptr(a), h, n)
    return x


def integrate(sysid, params, states, nsteps, h):
    params = np.ascontiguousarray(params, dtype=np.float64)
    states = np.ascontiguousarray(states, dtype=np.float64).copy()
    lib.orc_integrate(sysid, states.shape[0], dptr(params), dptr(states), nsteps, h)
    return states


def loop(sysid, x, a, h, amount, pre, wv, maxv, check, record=True):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    a = np.ascontiguousarray(a, dtype=np.float64)
    data = np.zeros(max(amount, 1)) if record else None
    flag = lib.orc_loop(sysid, dptr(x), dptr(a), h, amount, pre, wv, maxv, check, dptr(data))
    return flag, x, data


def peak_finder(data, h):
    data = np.ascontiguousarray(data, dtype=np.float64)
    n = data.size
    pk = np.zeros(max(n, 1))
    tp = np.zeros(max(n, 1))
    P = lib.orc_peak_finder(dptr(data), n, dptr(pk), dptr(tp), h)
    return pk[:P].copy(), tp[:P].copy()


def dbscan_ref(intervals, eps, peaks=None):
    iv = np.ascontiguousarray(intervals, dtype=np.float64)
    P = iv.size
    data = np.zeros(max(P, 1)) if peaks is None else np.ascontiguousarray(peaks, dtype=np.float64).copy()
    helpful = np.zeros(max(2 * P, 1))
    return lib.orc_dbscan_ref(dptr(data), dptr(iv), dptr(helpful), P, helpful.size, eps)


def dbscan1d_count(values, eps):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return lib.orc_dbscan1d_count(dptr(v), v.size, eps)


def unit_vector(seed, point, nx=3):
    v = np.zeros(nx)
    lib.orc_unit_vector(seed, point, nx, dptr(v))
    return v


def lle_point(sysid, sem, tMax, NT, h, eps, ic, a, tT, seed, point):
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib.orc_lle_point(sysid, sem, tMax, NT, h, eps, dptr(ic), ic.size, dptr(a), tT, seed, point)


def lle2d(sysid, sem, tMax, NT, nPts, h, eps, ic, ranges, idx, tT, values, seed,
          row_begin=0, row_end=None):
    row_end = nPts if row_end is None else row_end
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    ranges = np.ascontiguousarray(ranges, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros((row_end - row_begin) * nPts)
    rc = lib.orc_lle2d(sysid, sem, tMax, NT, nPts, h, eps, dptr(ic), ic.size, dptr(ranges), iptr(idx),
                       tT, dptr(values), values.size, seed, dptr(out), row_begin, row_end)
    assert rc == 0, rc
    return out.reshape(row_end - row_begin, nPts)


def bif1d(sysid, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre, capacity=None):
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    ranges = np.ascontiguousarray(ranges, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    counts = np.zeros(nPts, dtype=np.int32)
    first = np.zeros(nPts)
    cap = capacity or max(1, int(nPts * (tMax / h / pre) // 2 + 1))
    pk = np.zeros(cap)
    iv = np.zeros(cap)
    tot = lib.orc_bif1d(sysid, sem, tMax, nPts, h, dptr(ic), ic.size, dptr(ranges), iptr(idx), wv, maxv,
                        tT, dptr(values), values.size, pre, iptr(counts), dptr(first), dptr(pk),
                        dptr(iv), cap)
    assert tot >= 0, tot
    return counts, first, pk[:tot].copy(), iv[:tot].copy()


def bif2d(sysid, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre, eps,
          row_begin=0, row_end=None):
    row_end = nPts if row_end is None else row_end
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    ranges = np.ascontiguousarray(ranges, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros((row_end - row_begin) * nPts, dtype=np.int32)
    rc = lib.orc_bif2d(sysid, sem, tMax, nPts, h, dptr(ic), ic.size, dptr(ranges), iptr(idx), wv, maxv,
                       tT, dptr(values), values.size, pre, eps, iptr(out), row_begin, row_end)
    assert rc == 0, rc
    return out.reshape(row_end - row_begin, nPts)


def basins_features(sysid, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre,
                    point_begin=0, point_end=None):
    point_end = nPts * nPts if point_end is None else point_end
    n = point_end - point_begin
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    ranges = np.ascontiguousarray(ranges, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    f0, f1, f2 = np.zeros(n), np.zeros(n), np.zeros(n)
    flags = np.zeros(n, dtype=np.int32)
    rc = lib.orc_basins_features(sysid, sem, tMax, nPts, h, dptr(ic), ic.size, dptr(ranges), iptr(idx),
                                 wv, maxv, tT, dptr(values), values.size, pre, dptr(f0), dptr(f1),
                                 dptr(f2), iptr(flags), point_begin, point_end)
    assert rc == 0, rc
    return f0, f1, f2, flags


def basins_cluster_ref(f0, f1, flags, eps):
    f0 = np.ascontiguousarray(f0, dtype=np.float64)
    f1 = np.ascontiguousarray(f1, dtype=np.float64)
    flags = np.ascontiguousarray(flags, dtype=np.int32)
    labels = np.zeros(f0.size, dtype=np.int32)
    assert lib.orc_basins_cluster_ref(dptr(f0), dptr(f1), iptr(flags), f0.size, eps, iptr(labels)) == 0
    return labels


def det_sum(x, mask=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    return lib.orc_det_sum(dptr(x), iptr(m), x.size)


def zscore(f0, f1, f2, flags):
    f0, f1, f2 = (np.ascontiguousarray(f, dtype=np.float64).copy() for f in (f0, f1, f2))
    flags = np.ascontiguousarray(flags, dtype=np.int32)
    mom = np.zeros(6)
    assert lib.orc_zscore(dptr(f0), dptr(f1), dptr(f2), iptr(flags), f0.size, dptr(mom)) == 0
    return f0, f1, f2, mom


def dbscan_bruteforce(pts, eps, min_pts, include=None):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n, dim = pts.shape
    inc = None if include is None else np.ascontiguousarray(include, dtype=np.int32)
    labels = np.zeros(n, dtype=np.int32)
    k = lib.orc_dbscan_bruteforce(dptr(pts), n, dim, eps, min_pts, iptr(inc), iptr(labels))
    assert k >= 0
    return labels


def dbscan_grid(pts, eps, min_pts, include=None):
    """orc_dbscan_grid: the brute force's clustering in O(n) neighbour work."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n, dim = pts.shape
    inc = None if include is None else np.ascontiguousarray(include, dtype=np.int32)
    labels = np.zeros(n, dtype=np.int32)
    k = lib.orc_dbscan_grid(dptr(pts), n, dim, eps, min_pts, iptr(inc), iptr(labels))
    assert k >= 0, k
    return labels


def _arr(x, dt=np.float64):
    return np.ascontiguousarray(x, dtype=dt)


def legacy_lle(sysid, sem, form, dims, tMax, NT, nPts, h, eps, ic, ranges, idx, maxv, tT, values, seed,
               point_begin=0, point_end=None):
    n = nPts ** dims
    point_end = n if point_end is None else point_end
    ic, ranges, values = _arr(ic), _arr(ranges), _arr(values)
    idx = _arr(idx, np.int32)
    out = np.zeros(point_end - point_begin)
    rc = lib.orc_legacy_lle(sysid, sem, form, dims, tMax, NT, nPts, h, eps, dptr(ic), ic.size, dptr(ranges),
                            iptr(idx), maxv, tT, dptr(values), values.size, seed, dptr(out), point_begin,
                            point_end)
    assert rc == 0, rc
    return out


def gram_schmidt(a):
    a = _arr(a)
    n = a.shape[0]
    b = np.zeros_like(a)
    den = np.zeros(n)
    lib.orc_gram_schmidt(dptr(a), dptr(b), n, dptr(den))
    return b, den


def legacy_ls(sysid, sem, dims, tMax, NT, nPts, h, eps, ic, ranges, idx, maxv, tT, values, seed,
              point_begin=0, point_end=None):
    n = nPts ** dims
    point_end = n if point_end is None else point_end
    ic, ranges, values = _arr(ic), _arr(ranges), _arr(values)
    idx = _arr(idx, np.int32)
    out = np.zeros((point_end - point_begin) * ic.size)
    rc = lib.orc_legacy_ls(sysid, sem, dims, tMax, NT, nPts, h, eps, dptr(ic), ic.size, dptr(ranges),
                           iptr(idx), maxv, tT, dptr(values), values.size, seed, dptr(out), point_begin,
                           point_end)
    assert rc == 0, rc
    return out.reshape(point_end - point_begin, ic.size)


def time_domain(sysid, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre):
    ic, ranges, values = _arr(ic), _arr(ranges), _arr(values)
    idx = _arr(idx, np.int32)
    q = tMax / h / pre
    B = int(q) if sem == SEM_REFERENCE else int(np.floor(q + 0.5))
    out = np.zeros(nPts * max(B, 0))
    flags = np.zeros(nPts, dtype=np.int32)
    got = lib.orc_time_domain(sysid, sem, tMax, nPts, h, dptr(ic), ic.size, dptr(ranges), iptr(idx), wv, maxv,
                              tT, dptr(values), values.size, pre, dptr(out), iptr(flags))
    assert got == B, (got, B)
    return out.reshape(nPts, B), flags


def legacy_bif1d_ic(sysid, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre, capacity=1 << 22):
    ic, ranges, values = _arr(ic), _arr(ranges), _arr(values)
    idx = _arr(idx, np.int32)
    counts = np.zeros(nPts, dtype=np.int32)
    peaks = np.zeros(capacity)
    total = lib.orc_legacy_bif1d_ic(sysid, sem, tMax, nPts, h, dptr(ic), ic.size, dptr(ranges), iptr(idx), wv,
                                    maxv, tT, dptr(values), values.size, pre, iptr(counts), dptr(peaks), capacity)
    assert total >= 0, total
    return counts, peaks[:total]


def legacy_bif2d_ic(sysid, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre, eps):
    ic, ranges, values = _arr(ic), _arr(ranges), _arr(values)
    idx = _arr(idx, np.int32)
    out = np.zeros(nPts * nPts, dtype=np.int32)
    rc = lib.orc_legacy_bif2d_ic(sysid, sem, tMax, nPts, h, dptr(ic), ic.size, dptr(ranges), iptr(idx), wv, maxv,
                                 tT, dptr(values), values.size, pre, eps, iptr(out))
    assert rc == 0, rc
    return out
