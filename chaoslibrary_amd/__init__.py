"""chaoslibrary_amd -- MI355X (gfx950) ChaosLibrary hot path over a C-ABI.

This module binds ``libchaos.so`` (hand-written HIP kernels + C-ABI, built by
``__graft_entry__.build()`` / ``make -C chaoslibrary_amd/csrc``) with ctypes
and mirrors the reference's C++ host API names:

    bifurcation1D, bifurcation2D   <- Bifurcation/include/bifurcationHOST.h:16-50
    LLE2D                          <- LLE/include/LLEHost.h:25-41
    basinsOfAttraction_2           <- Basins/include/basinsHOST.h:15-30

plus value-returning variants (``bif1d``, ``bif2d``, ``lle2d``, ``basins2d``,
``integrate``).  There is no CPU fallback: every call runs the HIP kernels,
and importing fails loudly when the library is missing.  Errors raise
``ChaosError`` (the reference prints and exits, debri/cudaMacros.cu:3-10).
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "ChaosError", "Options", "Stats", "SYSTEMS", "SEM_REFERENCE", "SEM_CONFIG", "lib", "LIB_PATH",
    "system_dims", "integrate", "lle2d", "LLE2D", "bif1d", "bifurcation1D", "bif2d",
    "bifurcation2D", "basins2d", "basinsOfAttraction_2", "device_count",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchaos.so")

SYSTEMS = {"chameleon": 0, "rossler_cd": 1, "basins": 2, "basins_2": 3, "damir": 4,
           "rossler_rk4": 5, "lorenz_rk4": 6}
SEM_REFERENCE, SEM_CONFIG = 0, 1

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class ChaosError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("analysis_ms", C.c_double), ("total_ms", C.c_double),
                ("rk4_steps", C.c_double), ("peaks", C.c_longlong), ("launches", C.c_int),
                ("reserved", C.c_int)]


class Options(C.Structure):
    _fields_ = [("system", C.c_int), ("semantics", C.c_int), ("seed", C.c_uint64),
                ("device", C.c_int), ("device_buffers", C.c_int), ("stream", C.c_void_p),
                ("min_pts", C.c_int), ("block", C.c_int), ("stats", C.POINTER(Stats)),
                ("output", C.c_int), ("reserved", C.c_int)]

OUT_CSV, OUT_BINARY = 1, 2


# HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
# default).  The job server integrates on up to four streams and finishes jobs
# on two more; with four queues a finishing stream shares a queue with an
# integration stream, and its kernels wait behind whole persistent rounds
# (on a saturated stream the finishing falls behind and the median latency
# doubles).  HIP reads the variable when it starts, so it is raised here,
# before the library loads, unless the process asked for at least as many.
HW_QUEUES = 16


def _raise_hw_queues():
    v = os.environ.get("GPU_MAX_HW_QUEUES", "")
    if not v.isdigit() or int(v) < HW_QUEUES:
        os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)


_raise_hw_queues()


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"chaoslibrary_amd: {LIB_PATH} is missing; build it with "
                          "`python -c 'import __graft_entry__ as g; g.build()'`")
    return C.CDLL(LIB_PATH)


lib = _load()

_SIGS = {
    "chaos_options_init": (None, [C.POINTER(Options)]),
    "chaos_api_version": (C.c_int, []),
    "chaos_last_error": (C.c_char_p, []),
    "chaos_system_dims": (C.c_int, [C.c_int, _I, _I]),
    "chaos_set_default_options": (C.c_int, [C.POINTER(Options)]),
    "chaos_device_count": (C.c_int, [_I]),
    "chaos_device_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.c_int]),
    "chaos_device_free": (C.c_int, [C.c_void_p]),
    "chaos_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "chaos_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "chaos_device_synchronize": (C.c_int, [C.c_int]),
    "chaos_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "chaos_integrate": (C.c_int, [C.POINTER(Options), C.c_int, _D, _D, C.c_longlong, C.c_double]),
    "chaos_lle2d": (C.c_int, [C.POINTER(Options), C.c_double, C.c_double, C.c_int, C.c_double,
                              C.c_double, _D, C.c_int, _D, _I, C.c_int, C.c_double, C.c_double, _D,
                              C.c_int, C.c_void_p]),
    "chaos_LLE2D": (C.c_int, [C.POINTER(Options), C.c_double, C.c_double, C.c_int, C.c_double,
                              C.c_double, _D, C.c_int, _D, _I, C.c_int, C.c_double, C.c_double, _D,
                              C.c_int, C.c_char_p]),
    "chaos_dbscan": (C.c_int, [C.POINTER(Options), _D, _D, _D, C.c_longlong, C.c_int, C.c_double, C.c_int,
                               C.c_int, _I, _I, _I]),
}


def _bind():
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_bind()


def _check(rc):
    if rc != 0:
        raise ChaosError(rc, lib.chaos_last_error().decode(errors="replace"))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def make_options(system="chameleon", semantics=SEM_REFERENCE, seed=12345, device=0, block=0,
                 min_pts=20, stats=None, device_buffers=False, output=OUT_CSV):
    o = Options()
    lib.chaos_options_init(C.byref(o))
    o.system = SYSTEMS[system] if isinstance(system, str) else int(system)
    o.semantics = int(semantics)
    o.seed = int(seed)
    o.device = int(device)
    o.block = int(block)
    o.min_pts = int(min_pts)
    o.device_buffers = 1 if device_buffers else 0
    o.output = int(output)
    if stats is not None:
        o.stats = C.pointer(stats)
    return o


def system_dims(system):
    nx, na = C.c_int(), C.c_int()
    sid = SYSTEMS[system] if isinstance(system, str) else int(system)
    _check(lib.chaos_system_dims(sid, C.byref(nx), C.byref(na)))
    return nx.value, na.value


def device_count():
    c = C.c_int()
    _check(lib.chaos_device_count(C.byref(c)))
    return c.value


def integrate(system, params, states, nsteps, h, **kw):
    """Advance states[n][NX] by nsteps of the system map; returns the new states."""
    o = make_options(system, **kw)
    params = _f64(params)
    states = _f64(states).copy()
    n = states.shape[0]
    _check(lib.chaos_integrate(C.byref(o), n, _dp(params), _dp(states), int(nsteps), float(h)))
    return states


def lle2d(tMax, NT, nPts, h, eps, initialConditions, ranges, indicesOfMutVars, transientTime,
          values, writableVar=0, maxValue=0.0, system="chameleon", semantics=SEM_REFERENCE,
          out_device_ptr=None, **kw):
    """LLE map; returns out[j, i] = LLE(A_i, B_j) (row = second swept parameter)."""
    o = make_options(system, semantics, device_buffers=out_device_ptr is not None, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    if out_device_ptr is None:
        out = np.empty((nPts, nPts), dtype=np.float64)
        ptr = out.ctypes.data
    else:
        out, ptr = None, out_device_ptr
    _check(lib.chaos_lle2d(C.byref(o), tMax, NT, nPts, h, eps, _dp(ic), ic.size, _dp(rg), _ip(ix),
                           writableVar, maxValue, transientTime, _dp(vals), vals.size, C.c_void_p(ptr)))
    return out


def LLE2D(tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
          indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
          OUT_FILE_PATH, options=None):
    """LLE::LLE2D (LLE/include/LLEHost.h:25-41): writes the reference CSV."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    o = options if options is not None else None
    _check(lib.chaos_LLE2D(C.byref(o) if o is not None else None, tMax, NT, nPts, h, eps, _dp(ic),
                           amountOfInitialConditions, _dp(rg), _ip(ix), writableVar, maxValue,
                           transientTime, _dp(vals), amountOfValues, str(OUT_FILE_PATH).encode()))


_SIGS2 = {
    "chaos_bif1d": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D, _I,
                              C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, _I, _D, _D, _D,
                              C.c_longlong, C.POINTER(C.c_longlong)]),
    "chaos_bifurcation1D": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D,
                                      _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, C.c_char_p]),
    "chaos_bif2d": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D, _I,
                              C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, C.c_double, C.c_void_p]),
    "chaos_bifurcation2D": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D,
                                      _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, C.c_double,
                                      C.c_char_p]),
    "chaos_basins2d": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D, _I,
                                 C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, C.c_double,
                                 C.c_void_p, C.c_void_p, C.c_void_p]),
    "chaos_peak_finder": (C.c_int, [C.POINTER(Options), _D, C.c_int, C.c_int, C.c_double, _I, _D, _D]),
    "chaos_basinsOfAttraction_2": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int,
                                             _D, _D, _I, C.c_int, C.c_double, C.c_double, _D, C.c_int,
                                             C.c_int, C.c_double, C.c_char_p]),
}
for _n, (_r, _a) in _SIGS2.items():
    _fn = getattr(lib, _n)
    _fn.restype = _r
    _fn.argtypes = _a


def _opt_ref(options):
    return C.byref(options) if options is not None else None


def peak_finder(data, h, **kw):
    """Reference peakFinder on each row of data (GPU, one lane per row):
    returns a list of (peaks, intervals) per row."""
    o = make_options(**kw)
    d = _f64(np.atleast_2d(data))
    rows, n = d.shape
    counts = np.zeros(rows, dtype=np.int32)
    pk = np.zeros((rows, max(n, 1)))
    iv = np.zeros((rows, max(n, 1)))
    _check(lib.chaos_peak_finder(C.byref(o), _dp(d), rows, n, h, _ip(counts), _dp(pk), _dp(iv)))
    return [(pk[r, :counts[r]].copy(), iv[r, :counts[r]].copy()) for r in range(rows)]


def bif1d(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue,
          transientTime, values, preScaller=1, system="chameleon", semantics=SEM_REFERENCE, **kw):
    """Values of Bifurcation::bifurcation1D: (counts, first, peaks, intervals).

    counts[p] = kept peaks of point p, or -1 for a fixed point (first[p] is
    then its first recorded sample); peaks/intervals are concatenated in
    point order."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    counts = np.zeros(nPts, dtype=np.int32)
    first = np.zeros(nPts)
    total = C.c_longlong(0)
    args = (C.byref(o), tMax, nPts, h, ic.size, _dp(ic), _dp(rg), _ip(ix), writableVar, maxValue,
            transientTime, _dp(vals), vals.size, preScaller, _ip(counts), _dp(first))
    # one sweep when the peak buffers are large enough (one peak per 32
    # samples); a second, exactly sized one only when they are not
    # (an invalid h reaches the library, which reports it)
    q = tMax / h / max(1, preScaller) if h > 0 else 0.0
    B = max(1, int(round(q))) if np.isfinite(q) else 1
    cap = max(1, nPts * min(B // 2 + 1, max(256, B // 32)))
    pk, iv = np.zeros(cap), np.zeros(cap)
    rc = lib.chaos_bif1d(*args, _dp(pk), _dp(iv), cap, C.byref(total))
    if rc == -5:
        cap = max(1, total.value)
        pk, iv = np.zeros(cap), np.zeros(cap)
        rc = lib.chaos_bif1d(*args, _dp(pk), _dp(iv), cap, C.byref(total))
    _check(rc)
    return counts, first, pk[:total.value], iv[:total.value]


def bifurcation1D(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                  writableVar, maxValue, transientTime, values, amountOfValues, preScaller, OUT_FILE_PATH,
                  options=None):
    """Bifurcation::bifurcation1D (Bifurcation/include/bifurcationHOST.h:16-30)."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(lib.chaos_bifurcation1D(_opt_ref(options), tMax, nPts, h, amountOfInitialConditions, _dp(ic), _dp(rg),
                                   _ip(ix), writableVar, maxValue, transientTime, _dp(vals), amountOfValues,
                                   preScaller, str(OUT_FILE_PATH).encode()))


def bif2d(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue, transientTime,
          values, preScaller=1, eps=1e-2, system="chameleon", semantics=SEM_REFERENCE, out_device_ptr=None, **kw):
    """Values of Bifurcation::bifurcation2D: out[j, i] (row = second swept parameter)."""
    o = make_options(system, semantics, device_buffers=out_device_ptr is not None, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    if out_device_ptr is None:
        out = np.zeros((nPts, nPts), dtype=np.int32)
        ptr = out.ctypes.data
    else:
        out, ptr = None, out_device_ptr
    _check(lib.chaos_bif2d(C.byref(o), tMax, nPts, h, ic.size, _dp(ic), _dp(rg), _ip(ix), writableVar, maxValue,
                           transientTime, _dp(vals), vals.size, preScaller, eps, C.c_void_p(ptr)))
    return out


def bifurcation2D(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                  writableVar, maxValue, transientTime, values, amountOfValues, preScaller, eps, OUT_FILE_PATH,
                  options=None):
    """Bifurcation::bifurcation2D (Bifurcation/include/bifurcationHOST.h:35-50)."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(lib.chaos_bifurcation2D(_opt_ref(options), tMax, nPts, h, amountOfInitialConditions, _dp(ic), _dp(rg),
                                   _ip(ix), writableVar, maxValue, transientTime, _dp(vals), amountOfValues,
                                   preScaller, eps, str(OUT_FILE_PATH).encode()))


def basins2d(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue, transientTime,
             values, preScaller=1, eps=0.5, system="chameleon", semantics=SEM_REFERENCE, **kw):
    """Values of Basins::basinsOfAttraction_2: (labels[j,i], features[k,j,i], flags[j,i])."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    labels = np.zeros((nPts, nPts), dtype=np.int32)
    flags = np.zeros((nPts, nPts), dtype=np.int32)
    nf = 2 if int(semantics) == SEM_REFERENCE else 3
    feats = np.zeros((nf, nPts, nPts))
    _check(lib.chaos_basins2d(C.byref(o), tMax, nPts, h, ic.size, _dp(ic), _dp(rg), _ip(ix), writableVar,
                              maxValue, transientTime, _dp(vals), vals.size, preScaller, eps,
                              C.c_void_p(labels.ctypes.data), C.c_void_p(feats.ctypes.data),
                              C.c_void_p(flags.ctypes.data)))
    return labels, feats, flags


def dbscan(points, eps, min_pts=20, sign_rule=False, include=None, **kw):
    """chaos_dbscan over points[n, dim] (dim 2 or 3): (labels[n], K)."""
    pts = _f64(points)
    n, dim = pts.shape
    cols = [np.ascontiguousarray(pts[:, d]) for d in range(dim)]
    inc = None if include is None else _i32(include)
    labels = np.zeros(n, dtype=np.int32)
    k = C.c_int(0)
    o = make_options(**kw)
    _check(lib.chaos_dbscan(C.byref(o), _dp(cols[0]), _dp(cols[1]), _dp(cols[2]) if dim == 3 else None, n, dim,
                            float(eps), int(min_pts), 1 if sign_rule else 0, None if inc is None else _ip(inc),
                            _ip(labels), C.byref(k)))
    return labels, k.value


def basinsOfAttraction_2(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                         writableVar, maxValue, transientTime, values, amountOfValues, preScaller, eps,
                         OUT_FILE_PATH, options=None):
    """Basins::basinsOfAttraction_2 (Basins/include/basinsHOST.h:15-30)."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(lib.chaos_basinsOfAttraction_2(_opt_ref(options), tMax, nPts, h, amountOfInitialConditions, _dp(ic),
                                          _dp(rg), _ip(ix), writableVar, maxValue, transientTime, _dp(vals),
                                          amountOfValues, preScaller, eps, str(OUT_FILE_PATH).encode()))


# ---------------------------------------------------------------------------
# legacy entry points (namespace old_library)
# ---------------------------------------------------------------------------

lib.chaos_map_bin_to_csv.restype = C.c_int
lib.chaos_map_bin_to_csv.argtypes = [C.c_char_p, C.c_char_p]


def map_bin_to_csv(bin_path, csv_path):
    """chaos_map_bin_to_csv: the CSV text of a map sidecar (<path>.bin)."""
    _check(lib.chaos_map_bin_to_csv(str(bin_path).encode(), str(csv_path).encode()))


_LLE_STYLE = [C.POINTER(Options), C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, _D, C.c_int, _D, _I,
              C.c_int, C.c_double, C.c_double, _D, C.c_int]
_SIGS_LEGACY = {
    "chaos_legacy_lle": (C.c_int, _LLE_STYLE + [C.c_int, C.c_int, C.c_void_p]),
    "chaos_legacy_ls": (C.c_int, _LLE_STYLE + [C.c_int, C.c_void_p]),
    "chaos_time_domain": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D, _I,
                                    C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_longlong, C.POINTER(C.c_longlong)]),
    "chaos_TimeDomainCalculation": (C.c_int, [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D,
                                              _D, _I, C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_int,
                                              C.c_char_p]),
}
_BIF_STYLE = [C.POINTER(Options), C.c_double, C.c_int, C.c_double, C.c_int, _D, _D, _I, C.c_int, C.c_double,
              C.c_double, _D, C.c_int, C.c_int]
_SIGS_LEGACY["chaos_legacy_bif1d_ic"] = (C.c_int, _BIF_STYLE + [C.c_void_p, C.c_void_p, C.c_longlong,
                                                               C.POINTER(C.c_longlong)])
_SIGS_LEGACY["chaos_legacy_bif2d_ic"] = (C.c_int, _BIF_STYLE + [C.c_double, C.c_void_p])
_SIGS_LEGACY["chaos_bifurcation1DIC"] = (C.c_int, _BIF_STYLE + [C.c_char_p])
_SIGS_LEGACY["chaos_bifurcation2DIC"] = (C.c_int, _BIF_STYLE + [C.c_double, C.c_char_p])
for _n in ("chaos_LLE1D", "chaos_LLE1DIC", "chaos_LLE2DIC", "chaos_LS1D", "chaos_LS2D"):
    _SIGS_LEGACY[_n] = (C.c_int, _LLE_STYLE + [C.c_char_p])
for _name, (_res, _args) in _SIGS_LEGACY.items():
    getattr(lib, _name).restype = _res
    getattr(lib, _name).argtypes = _args


def _lle_style_args(o, tMax, NT, nPts, h, eps, ic, ranges, idx, writableVar, maxValue, transientTime, vals):
    return (C.byref(o), tMax, NT, nPts, h, eps, _dp(ic), ic.size, _dp(ranges), _ip(idx), writableVar, maxValue,
            transientTime, _dp(vals), vals.size)


def legacy_lle(tMax, NT, nPts, h, eps, initialConditions, ranges, indicesOfMutVars, transientTime, values,
               dims=1, sweep_ic=False, maxValue=0.0, writableVar=0, system="chameleon", semantics=SEM_REFERENCE,
               **kw):
    """chaos_legacy_lle: LLE over nPts**dims points (p = k0 + nPts*k1)."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    out = np.zeros(nPts ** dims)
    _check(lib.chaos_legacy_lle(*_lle_style_args(o, tMax, NT, nPts, h, eps, ic, rg, ix, writableVar, maxValue,
                                                 transientTime, vals), dims, 1 if sweep_ic else 0,
                                C.c_void_p(out.ctypes.data)))
    return out


def legacy_ls(tMax, NT, nPts, h, eps, initialConditions, ranges, indicesOfMutVars, transientTime, values,
              dims=1, maxValue=0.0, writableVar=0, system="chameleon", semantics=SEM_REFERENCE, **kw):
    """chaos_legacy_ls: Lyapunov spectrum, shape (nPts**dims, NX)."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    out = np.zeros((nPts ** dims, ic.size))
    _check(lib.chaos_legacy_ls(*_lle_style_args(o, tMax, NT, nPts, h, eps, ic, rg, ix, writableVar, maxValue,
                                                transientTime, vals), dims, C.c_void_p(out.ctypes.data)))
    return out


def time_domain(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue, transientTime,
                values, preScaller=1, system="chameleon", semantics=SEM_REFERENCE, **kw):
    """chaos_time_domain: (samples[nPts, B], flags[nPts])."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    B = C.c_longlong(0)
    args = (C.byref(o), tMax, nPts, h, ic.size, _dp(ic), _dp(rg), _ip(ix), writableVar, maxValue, transientTime,
            _dp(vals), vals.size, preScaller)
    _check(lib.chaos_time_domain(*args, None, None, 0, C.byref(B)))
    out = np.zeros((nPts, B.value))
    flags = np.zeros(nPts, dtype=np.int32)
    _check(lib.chaos_time_domain(*args, C.c_void_p(out.ctypes.data), C.c_void_p(flags.ctypes.data), out.size,
                                 C.byref(B)))
    return out, flags


def legacy_bif1d_ic(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue,
                    transientTime, values, preScaller=1, system="chameleon", semantics=SEM_REFERENCE, **kw):
    """chaos_legacy_bif1d_ic: (counts[nPts], kept peak values concatenated)."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    counts = np.zeros(nPts, dtype=np.int32)
    total = C.c_longlong(0)
    args = (C.byref(o), tMax, nPts, h, ic.size, _dp(ic), _dp(rg), _ip(ix), writableVar, maxValue, transientTime,
            _dp(vals), vals.size, preScaller, C.c_void_p(counts.ctypes.data))
    # (an invalid h reaches the library, which reports it)
    q = tMax / h / max(1, preScaller) if h > 0 else 0.0
    B = max(1, int(round(q))) if np.isfinite(q) else 1
    cap = max(1, nPts * min(B // 2 + 1, max(256, B // 32)))
    pk = np.zeros(cap)
    rc = lib.chaos_legacy_bif1d_ic(*args, C.c_void_p(pk.ctypes.data), cap, C.byref(total))
    if rc == -5:
        cap = max(1, total.value)
        pk = np.zeros(cap)
        rc = lib.chaos_legacy_bif1d_ic(*args, C.c_void_p(pk.ctypes.data), cap, C.byref(total))
    _check(rc)
    return counts, pk[:total.value]


def legacy_bif2d_ic(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue,
                    transientTime, values, preScaller=1, eps=0.1, system="chameleon", semantics=SEM_REFERENCE, **kw):
    """chaos_legacy_bif2d_ic: int map over nPts^2 initial conditions (p = k0 + nPts*k1)."""
    o = make_options(system, semantics, **kw)
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    out = np.zeros(nPts * nPts, dtype=np.int32)
    _check(lib.chaos_legacy_bif2d_ic(C.byref(o), tMax, nPts, h, ic.size, _dp(ic), _dp(rg), _ip(ix), writableVar,
                                     maxValue, transientTime, _dp(vals), vals.size, preScaller, eps,
                                     C.c_void_p(out.ctypes.data)))
    return out


def bifurcation1DIC(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                    writableVar, maxValue, transientTime, values, amountOfValues, preScaller, OUT_FILE_PATH,
                    options=None):
    """old_library::bifurcation1DIC (debri/hostLibrary.cuh:109-123)."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(lib.chaos_bifurcation1DIC(_opt_ref(options), tMax, nPts, h, amountOfInitialConditions, _dp(ic), _dp(rg),
                                     _ip(ix), writableVar, maxValue, transientTime, _dp(vals), amountOfValues,
                                     preScaller, str(OUT_FILE_PATH).encode()))


def bifurcation2DIC(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                    writableVar, maxValue, transientTime, values, amountOfValues, preScaller, eps, OUT_FILE_PATH,
                    options=None):
    """old_library::bifurcation2DIC (debri/hostLibrary.cuh:219-234)."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(lib.chaos_bifurcation2DIC(_opt_ref(options), tMax, nPts, h, amountOfInitialConditions, _dp(ic), _dp(rg),
                                     _ip(ix), writableVar, maxValue, transientTime, _dp(vals), amountOfValues,
                                     preScaller, eps, str(OUT_FILE_PATH).encode()))


def _legacy_file(fn, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
                 indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues, OUT_FILE_PATH,
                 options):
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(getattr(lib, fn)(_opt_ref(options), tMax, NT, nPts, h, eps, _dp(ic), amountOfInitialConditions, _dp(rg),
                            _ip(ix), writableVar, maxValue, transientTime, _dp(vals), amountOfValues,
                            str(OUT_FILE_PATH).encode()))


def LLE1D(*args, options=None):
    """old_library::LLE1D (debri/hostLibrary.cuh:241-256)."""
    _legacy_file("chaos_LLE1D", *args, options=options)


def LLE1DIC(*args, options=None):
    """old_library::LLE1DIC (debri/hostLibrary.cuh:263-278)."""
    _legacy_file("chaos_LLE1DIC", *args, options=options)


def LLE2DIC(*args, options=None):
    """old_library::LLE2DIC (debri/hostLibrary.cuh:339-354)."""
    _legacy_file("chaos_LLE2DIC", *args, options=options)


def LS1D(*args, options=None):
    """old_library::LS1D (debri/hostLibrary.cuh:377-392)."""
    _legacy_file("chaos_LS1D", *args, options=options)


def LS2D(*args, options=None):
    """old_library::LS2D (debri/hostLibrary.cuh:416-431)."""
    _legacy_file("chaos_LS2D", *args, options=options)


def TimeDomainCalculation(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                          writableVar, maxValue, transientTime, values, amountOfValues, preScaller, OUT_FILE_PATH,
                          options=None):
    """old_library::TimeDomainCalculation (debri/hostLibrary.cuh:69-83)."""
    ic, rg, ix, vals = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(lib.chaos_TimeDomainCalculation(_opt_ref(options), tMax, nPts, h, amountOfInitialConditions, _dp(ic),
                                           _dp(rg), _ip(ix), writableVar, maxValue, transientTime, _dp(vals),
                                           amountOfValues, preScaller, str(OUT_FILE_PATH).encode()))


# ---------------------------------------------------------------------------
# batched engine / job server / mixed-job stream
# ---------------------------------------------------------------------------

JOB_LLE2D, JOB_BIF1D, JOB_BIF2D, JOB_BASINS2D = 0, 1, 2, 3


class Job(C.Structure):
    _fields_ = [("kind", C.c_int), ("system", C.c_int), ("semantics", C.c_int), ("nPts", C.c_int),
                ("tMax", C.c_double), ("NT", C.c_double), ("h", C.c_double), ("eps", C.c_double),
                ("maxValue", C.c_double), ("transientTime", C.c_double), ("preScaller", C.c_int),
                ("writableVar", C.c_int), ("min_pts", C.c_int), ("amountOfInitialConditions", C.c_int),
                ("seed", C.c_uint64), ("initialConditions", C.c_double * 4), ("values", C.c_double * 8),
                ("amountOfValues", C.c_int), ("indicesOfMutVars", C.c_int * 2), ("reserved", C.c_int),
                ("ranges", C.c_double * 4)]


class EngineParams(C.Structure):
    _fields_ = [("block", C.c_int), ("waves_per_simd", C.c_int), ("steps_per_launch", C.c_int),
                ("max_jobs", C.c_int), ("max_tiles", C.c_int), ("use_graph", C.c_int),
                ("hbm_fraction", C.c_double), ("points_per_tile", C.c_int), ("reserved", C.c_int)]


class JobInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("kind", C.c_int), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double), ("rk4_steps", C.c_double),
                ("footprint_bytes", C.c_longlong), ("n_points", C.c_longlong), ("result_len", C.c_longlong),
                ("reran", C.c_int), ("reserved", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("launches", C.c_longlong), ("graph_replays", C.c_longlong), ("tiles", C.c_longlong),
                ("busy_ms", C.c_double), ("finalize_ms", C.c_double), ("rk4_steps", C.c_double),
                ("active_jobs", C.c_int), ("pending_jobs", C.c_int), ("hbm_in_use", C.c_longlong),
                ("hbm_budget", C.c_longlong), ("finishing_jobs", C.c_int), ("reserved", C.c_int)]


class StreamConfig(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("duration_s", C.c_double), ("horizon_scale", C.c_double),
                ("seed", C.c_uint64), ("bif_min", C.c_int), ("bif_max", C.c_int), ("lle_min", C.c_int),
                ("lle_max", C.c_int), ("basins_min", C.c_int), ("basins_max", C.c_int), ("mode", C.c_int),
                ("max_jobs", C.c_int), ("progress_s", C.c_double)]


class StreamReport(C.Structure):
    _fields_ = [("jobs", C.c_int), ("jobs_kind", C.c_int * 3), ("makespan_s", C.c_double),
                ("rk4_steps", C.c_double), ("steps_per_s", C.c_double), ("jobs_per_s", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("p99_kind_ms", C.c_double * 3), ("busy_ms", C.c_double), ("launches", C.c_longlong),
                ("finalize_ms", C.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["jobs_kind"] = list(self.jobs_kind)
        d["p99_kind_ms"] = list(self.p99_kind_ms)
        return d


_SIGS3 = {
    "chaos_engine_params_default": (C.c_int, [C.POINTER(EngineParams)]),
    "chaos_server_create": (C.c_int, [C.c_int, C.POINTER(EngineParams), C.POINTER(C.c_void_p)]),
    "chaos_server_destroy": (C.c_int, [C.c_void_p]),
    "chaos_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_longlong)]),
    "chaos_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "chaos_server_drain": (C.c_int, [C.c_void_p]),
    "chaos_server_job_info": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(JobInfo)]),
    "chaos_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "chaos_server_result_lle": (C.c_int, [C.c_void_p, C.c_longlong, _D]),
    "chaos_server_result_bif2d": (C.c_int, [C.c_void_p, C.c_longlong, _I]),
    "chaos_server_result_bif1d": (C.c_int, [C.c_void_p, C.c_longlong, _I, _D, _D, _D, C.c_longlong,
                                            C.POINTER(C.c_longlong)]),
    "chaos_server_result_basins": (C.c_int, [C.c_void_p, C.c_longlong, _I, _D, _I]),
    "chaos_server_write_result": (C.c_int, [C.c_void_p, C.c_longlong, C.c_char_p, C.c_int]),
    "chaos_server_release": (C.c_int, [C.c_void_p, C.c_longlong]),
    "chaos_stream_config_default": (C.c_int, [C.POINTER(StreamConfig)]),
    "chaos_stream_jobs": (C.c_int, [C.POINTER(StreamConfig), C.POINTER(Job), _D, C.c_int, _I]),
    "chaos_stream_run": (C.c_int, [C.c_int, C.POINTER(StreamConfig), C.POINTER(EngineParams),
                                   C.POINTER(StreamReport), _D, C.c_int]),
    "chaos_autotune": (C.c_int, [C.c_int, C.c_char_p, C.POINTER(EngineParams), C.POINTER(C.c_double)]),
    "chaos_autotune_load": (C.c_int, [C.c_char_p, C.POINTER(EngineParams)]),
}
for _n, (_r, _a) in _SIGS3.items():
    _fn = getattr(lib, _n)
    _fn.restype = _r
    _fn.argtypes = _a


def engine_params(**kw):
    p = EngineParams()
    _check(lib.chaos_engine_params_default(C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def make_job(kind, system, semantics, nPts, initialConditions, values, ranges, indicesOfMutVars, h=0.01,
             tMax=0.0, NT=0.5, eps=0.0, maxValue=0.0, transientTime=0.0, preScaller=1, writableVar=0,
             min_pts=20, seed=12345):
    j = Job()
    j.kind = kind
    j.system = SYSTEMS[system] if isinstance(system, str) else int(system)
    j.semantics = semantics
    j.nPts = nPts
    j.tMax, j.NT, j.h, j.eps, j.maxValue, j.transientTime = tMax, NT, h, eps, maxValue, transientTime
    j.preScaller, j.writableVar, j.min_pts, j.seed = preScaller, writableVar, min_pts, seed
    j.amountOfInitialConditions = len(initialConditions)
    for i, v in enumerate(initialConditions):
        j.initialConditions[i] = v
    j.amountOfValues = len(values)
    for i, v in enumerate(values):
        j.values[i] = v
    for i, v in enumerate(ranges):
        j.ranges[i] = v
    for i, v in enumerate(indicesOfMutVars):
        j.indicesOfMutVars[i] = v
    return j


class Server:
    """Job server over the batched engine (chaos_server_*)."""

    def __init__(self, device=0, params=None):
        self.h = C.c_void_p()
        _check(lib.chaos_server_create(device, C.byref(params) if params is not None else None, C.byref(self.h)))

    def close(self):
        if self.h:
            _check(lib.chaos_server_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, job):
        jid = C.c_longlong()
        _check(lib.chaos_server_submit(self.h, C.byref(job), C.byref(jid)))
        return jid.value

    def step(self):
        c = C.c_int()
        _check(lib.chaos_server_step(self.h, C.byref(c)))
        return c.value

    def drain(self):
        _check(lib.chaos_server_drain(self.h))

    def info(self, jid):
        i = JobInfo()
        _check(lib.chaos_server_job_info(self.h, jid, C.byref(i)))
        return i

    def stats(self):
        s = ServerStats()
        _check(lib.chaos_server_stats_get(self.h, C.byref(s)))
        return s

    def result(self, jid, job):
        """Results in the layout of the one-shot wrappers for the job's kind."""
        n = job.nPts
        if job.kind == JOB_LLE2D:
            out = np.zeros((n, n))
            _check(lib.chaos_server_result_lle(self.h, jid, _dp(out)))
            return out
        if job.kind == JOB_BIF2D:
            out = np.zeros((n, n), dtype=np.int32)
            _check(lib.chaos_server_result_bif2d(self.h, jid, _ip(out)))
            return out
        if job.kind == JOB_BIF1D:
            info = self.info(jid)
            cap = max(1, info.result_len)
            counts, first = np.zeros(n, dtype=np.int32), np.zeros(n)
            pk, iv = np.zeros(cap), np.zeros(cap)
            tot = C.c_longlong()
            _check(lib.chaos_server_result_bif1d(self.h, jid, _ip(counts), _dp(first), _dp(pk), _dp(iv), cap,
                                                 C.byref(tot)))
            return counts, first, pk[:tot.value], iv[:tot.value]
        labels, flags = np.zeros((n, n), dtype=np.int32), np.zeros((n, n), dtype=np.int32)
        nf = 2 if job.semantics == SEM_REFERENCE else 3
        feats = np.zeros((nf, n, n))
        _check(lib.chaos_server_result_basins(self.h, jid, _ip(labels), _dp(feats), _ip(flags)))
        return labels, feats, flags

    def write_result(self, jid, path, output=0):
        """Writes the finished job's files as the one-shot file-writing entry
        point would (output: OUT_CSV and/or OUT_BINARY, 0 = CSV)."""
        _check(lib.chaos_server_write_result(self.h, jid, os.fsencode(path), output))

    def release(self, jid):
        _check(lib.chaos_server_release(self.h, jid))


def stream_config(**kw):
    c = StreamConfig()
    _check(lib.chaos_stream_config_default(C.byref(c)))
    for k, v in kw.items():
        setattr(c, "lambda_" if k == "lam" else k, v)
    return c


def stream_jobs(cfg):
    """The stream's jobs in arrival order: (list of Job, arrival_ms array)."""
    cnt = C.c_int(0)
    _check(lib.chaos_stream_jobs(C.byref(cfg), None, None, 0, C.byref(cnt)))
    n = cnt.value
    jobs = (Job * max(1, n))()
    arr = np.zeros(max(1, n))
    _check(lib.chaos_stream_jobs(C.byref(cfg), jobs, _dp(arr), n, C.byref(cnt)))
    return [jobs[k] for k in range(n)], arr[:n].copy()


def stream_run(cfg, params=None, device=0):
    """Run the seeded mixed-job stream; returns (report, latencies_ms)."""
    rep = StreamReport()
    cap = 1 << 20
    lat = np.zeros(cap)
    _check(lib.chaos_stream_run(device, C.byref(cfg), C.byref(params) if params is not None else None,
                                C.byref(rep), _dp(lat), cap))
    return rep, lat[:rep.jobs].copy()


def autotune(cache_path=None, device=0):
    p = EngineParams()
    rate = C.c_double()
    _check(lib.chaos_autotune(device, cache_path.encode() if cache_path else None, C.byref(p), C.byref(rate)))
    return p, rate.value


def autotune_load(cache_path):
    p = EngineParams()
    _check(lib.chaos_autotune_load(cache_path.encode(), C.byref(p)))
    return p
