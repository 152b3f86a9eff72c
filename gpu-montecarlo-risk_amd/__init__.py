"""MI355X (gfx950) Monte Carlo crash-risk model: Python view of the C ABI.

The product is ``_lib/libmcr_hip.so`` (built from ``csrc/`` by ``build()``), whose
C ABI is declared in ``include/mcr_hip.h``.  This module only binds it with
ctypes; there is no Python or CPU fallback, and every entry point raises when
the library is missing or a call fails.

Reference API mirrored here (same names, argument meaning and return values):
  monte_carlo_crash_simulation(x, N, k, mu, Sigma, M)  -> (P_hat, count, S_values)
      src/01-python-serial/python_serial.py:11-75
  compute_accuracy(P_hat, M) -> (std_error, ci_lower, ci_upper)
      src/01-python-serial/python_serial.py:77-105
  get_hip_model() -> ModelFunctions {name, run_model}
      src/model_interface.h:33-36, src/05-GPU/monte_carlos_cuda.cu:165-171
"""
from __future__ import annotations

import ctypes
import math
import os
import re
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MCR_LIB_PATH selects another build of the same library (the host-sanitizer
# variant of tests/sanitize.sh); by default the in-tree build
LIB_PATH = os.environ.get("MCR_LIB_PATH") or os.path.join(_HERE, "_lib", "libmcr_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mcr_hip.h")
HEADER_PATHS = [HEADER_PATH, os.path.join(os.path.dirname(_HERE), "include", "mcr_io.h")]

# constants from include/mcr_hip.h
MCR_OK = 0
MCR_E_ARG, MCR_E_UNSUPPORTED, MCR_E_DEVICE, MCR_E_NOMEM, MCR_E_FACTOR = -1, -2, -3, -4, -5
MCR_MAX_ALPHA = 4
MCR_MAX_N, MCR_MAX_N_BATCH, MCR_MAX_N_NOALLOC = 1024, 767, 608
PREC_FP32, PREC_BF16X3, PREC_AUTO = 0, 1, 2
KERNEL_AUTO, KERNEL_VALU_REG, KERNEL_VALU_TILE, KERNEL_MFMA, KERNEL_MFMA_LDS = 0, 1, 2, 3, 4
KERNEL_VALU_CHUNK = 5
F_VAR, F_NEAR, F_SHIST, F_TEST_NARROW, F_TEST_WIDE = 0x1, 0x2, 0x4, 0x100, 0x200
KERNEL_NAMES = {KERNEL_VALU_REG: "valu_reg", KERNEL_VALU_TILE: "valu_tile", KERNEL_MFMA: "mfma",
                KERNEL_MFMA_LDS: "mfma_lds", KERNEL_VALU_CHUNK: "valu_chunk"}


class McrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mcr error {code}: {msg}")
        self.code = code


def build(jobs: int = 8) -> str:
    """Compile csrc/ for gfx950 into _lib/libmcr_hip.so (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", os.path.join(_HERE, "csrc")], check=True)
    return LIB_PATH


_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class Query(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("mu", _f64p), ("sigma", _f64p), ("ld_sigma", ctypes.c_int),
                ("weights", _f64p), ("x", ctypes.c_double), ("k", ctypes.c_int),
                ("paths", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("path_offset", ctypes.c_uint64), ("precision", ctypes.c_int),
                ("kernel", ctypes.c_int), ("flags", ctypes.c_int), ("n_alpha", ctypes.c_int),
                ("alpha", ctypes.c_double * MCR_MAX_ALPHA), ("S_out", _i32p),
                ("loss_out", _f32p), ("s_hist", _i64p)]


class Result(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64), ("near_threshold", ctypes.c_int64),
                ("p_hat", ctypes.c_double), ("std_error", ctypes.c_double),
                ("ci_lower", ctypes.c_double), ("ci_upper", ctypes.c_double),
                ("var", ctypes.c_double * MCR_MAX_ALPHA), ("es", ctypes.c_double * MCR_MAX_ALPHA),
                ("tail_sum_fixed", ctypes.c_int64 * MCR_MAX_ALPHA), ("var_passes", ctypes.c_int),
                ("loss_mean", ctypes.c_double), ("loss_sd", ctypes.c_double),
                ("jitter", ctypes.c_double), ("factor_attempts", ctypes.c_int),
                ("factor_clamped", ctypes.c_int), ("kernel", ctypes.c_int),
                ("ms_kernel", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("ms_device", ctypes.c_double)]


class FactorInfo(ctypes.Structure):
    _fields_ = [("attempts", ctypes.c_int), ("jitter", ctypes.c_double), ("clamped", ctypes.c_int)]


class ModelExport(ctypes.Structure):
    _fields_ = [("L64", _f64p), ("L32", _f32p), ("mu32", _f32p), ("v32", _f32p),
                ("mp32", ctypes.c_float), ("thr32", ctypes.c_float), ("loss_sd", ctypes.c_double)]


# GSL-layout structs of include/model_interface.h (for the plug-in boundary)
class GslVector(ctypes.Structure):
    _fields_ = [("size", ctypes.c_size_t), ("stride", ctypes.c_size_t), ("data", _f64p),
                ("block", ctypes.c_void_p), ("owner", ctypes.c_int)]


class GslMatrix(ctypes.Structure):
    _fields_ = [("size1", ctypes.c_size_t), ("size2", ctypes.c_size_t), ("tda", ctypes.c_size_t),
                ("data", _f64p), ("block", ctypes.c_void_p), ("owner", ctypes.c_int)]


class MonteCarloResult(ctypes.Structure):
    _fields_ = [("P_hat", ctypes.c_double), ("count", ctypes.c_int), ("S_values", _i32p),
                ("std_error", ctypes.c_double), ("ci_lower", ctypes.c_double),
                ("ci_upper", ctypes.c_double)]


class MonteCarloParams(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("N", ctypes.c_int), ("k", ctypes.c_int),
                ("mu", ctypes.POINTER(GslVector)), ("Sigma", ctypes.POINTER(GslMatrix)),
                ("M", ctypes.c_int), ("random_seed", ctypes.c_ulong)]


ModelRunFunc = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(MonteCarloParams),
                                ctypes.POINTER(MonteCarloResult))


class ModelFunctions(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("run_model", ModelRunFunc)]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded product library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run build())")
        L = ctypes.CDLL(LIB_PATH)
        L.mcr_init.argtypes = [ctypes.c_int]
        L.mcr_last_error.restype = ctypes.c_char_p
        L.mcr_run.argtypes = [ctypes.POINTER(Query), ctypes.POINTER(Result)]
        L.mcr_run_noalloc.argtypes = [ctypes.POINTER(Query), ctypes.POINTER(Result)]
        L.mcr_factor.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _f64p,
                                 ctypes.POINTER(FactorInfo)]
        L.mcr_prepare.argtypes = [ctypes.POINTER(Query), ctypes.POINTER(ModelExport)]
        L.mcr_normals.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int,
                                  _f32p]
        L.mcr_synth_factor_model.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_uint64, _f64p, _f64p, _f64p]
        L.hip_simulate.argtypes = [ctypes.POINTER(MonteCarloParams),
                                   ctypes.POINTER(MonteCarloResult)]
        L.get_hip_model.restype = ModelFunctions
        L.mcr_run_batch.argtypes = [ctypes.POINTER(Query), ctypes.c_int, ctypes.POINTER(Result),
                                    ctypes.POINTER(BatchOpts)]
        L.mcr_autotune_batch.argtypes = [ctypes.POINTER(Query), ctypes.c_int,
                                         ctypes.POINTER(BatchOpts), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def declared_functions() -> list[str]:
    """Function names declared in include/mcr_hip.h and include/mcr_io.h (the
    exported C ABI)."""
    text = "\n".join(open(p).read() for p in HEADER_PATHS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*([a-z_][a-z0-9_]*)\s*\(",
                      text, flags=re.M)


def _check(code: int) -> None:
    if code != MCR_OK:
        raise McrError(code, lib().mcr_last_error().decode())


def _square(sigma) -> np.ndarray:
    """sigma as a contiguous n x n float64 array: a square 2-D array, or n*n
    values flat (numpy's covariance of one variable is 0-d, i.e. 1 x 1)."""
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    if s.ndim == 1:
        n = math.isqrt(s.size)
        if n * n != s.size:
            raise ValueError(f"sigma with {s.size} values is not n x n")
        return s.reshape(n, n)
    if s.ndim != 2 or s.shape[0] != s.shape[1]:
        raise ValueError(f"sigma must be square, not {s.shape}")
    return s


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


@dataclass
class RiskResult:
    count: int
    near_threshold: int
    p_hat: float
    std_error: float
    ci_lower: float
    ci_upper: float
    var: list = field(default_factory=list)
    es: list = field(default_factory=list)
    tail_sum_fixed: list = field(default_factory=list)
    var_passes: int = 0
    loss_mean: float = 0.0
    loss_sd: float = 0.0
    jitter: float = 0.0
    factor_attempts: int = 0
    factor_clamped: int = 0
    kernel: str = ""
    ms_kernel: float = 0.0
    ms_total: float = 0.0
    ms_device: float = 0.0
    S: np.ndarray | None = None
    loss: np.ndarray | None = None
    s_hist: np.ndarray | None = None


def _make_query(mu, sigma, x, k, paths, seed, weights=None, precision=PREC_FP32,
                kernel=KERNEL_AUTO, alphas=(), near=False, want_S=False, want_loss=False,
                want_hist=False, path_offset=0, extra_flags=0):
    """A filled Query plus the arrays it points into (keep them alive)."""
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    if mu.ndim != 1:
        raise ValueError(f"mu must be one-dimensional (shape {mu.shape})")
    n = int(mu.shape[0])
    # the library copies n rows of ld_sigma doubles and n weights: the arrays
    # must hold them
    if sigma.ndim == 2:
        if sigma.shape[0] < n or sigma.shape[1] < n:
            raise ValueError(f"sigma of shape {sigma.shape} is smaller than {n} x {n}")
    elif sigma.ndim != 1 or sigma.size < n * n:
        raise ValueError(f"sigma must be {n} x {n} (or hold {n * n} values flat)")
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    if w is not None and w.shape != (n,):
        raise ValueError(f"weights must have shape ({n},), not {w.shape}")
    S = np.zeros(paths, dtype=np.int32) if want_S else None
    loss = np.zeros(paths, dtype=np.float32) if want_loss else None
    hist = np.zeros(n + 1, dtype=np.int64) if want_hist else None
    flags = extra_flags
    if alphas:
        flags |= F_VAR
    if near:
        flags |= F_NEAR
    if want_hist:
        flags |= F_SHIST
    q = Query()
    q.n = n
    q.mu = _ptr(mu, _f64p)
    q.sigma = _ptr(sigma, _f64p)
    q.ld_sigma = int(sigma.shape[1]) if sigma.ndim == 2 else n
    q.weights = _ptr(w, _f64p)
    q.x = float(x)
    q.k = int(k)
    q.paths = int(paths)
    q.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    q.path_offset = int(path_offset)
    q.precision = int(precision)
    q.kernel = int(kernel)
    q.flags = flags
    q.n_alpha = len(alphas)
    for i, a in enumerate(alphas):
        q.alpha[i] = float(a)
    q.S_out = _ptr(S, _i32p)
    q.loss_out = _ptr(loss, _f32p)
    q.s_hist = _ptr(hist, _i64p)
    return q, (mu, sigma, w, S, loss, hist)


def _result(r: Result, na: int, keep=None) -> RiskResult:
    S, loss, hist = (keep[3], keep[4], keep[5]) if keep else (None, None, None)
    return RiskResult(count=r.count, near_threshold=r.near_threshold, p_hat=r.p_hat,
                      std_error=r.std_error, ci_lower=r.ci_lower, ci_upper=r.ci_upper,
                      var=list(r.var)[:na], es=list(r.es)[:na],
                      tail_sum_fixed=list(r.tail_sum_fixed)[:na], var_passes=r.var_passes,
                      loss_mean=r.loss_mean, loss_sd=r.loss_sd, jitter=r.jitter,
                      factor_attempts=r.factor_attempts, factor_clamped=r.factor_clamped,
                      kernel=KERNEL_NAMES.get(r.kernel, str(r.kernel)), ms_kernel=r.ms_kernel,
                      ms_total=r.ms_total, ms_device=r.ms_device, S=S, loss=loss,
                      s_hist=hist)


def run(mu, sigma, x: float, k: int, paths: int, seed: int = 0, *, weights=None,
        precision: int = PREC_FP32, kernel: int = KERNEL_AUTO, alphas=(), near: bool = False,
        want_S: bool = False, want_loss: bool = False, want_hist: bool = False,
        path_offset: int = 0, extra_flags: int = 0) -> RiskResult:
    """One risk query on the GPU (mcr_run)."""
    q, keep = _make_query(mu, sigma, x, k, paths, seed, weights, precision, kernel, alphas, near,
                          want_S, want_loss, want_hist, path_offset, extra_flags)
    r = Result()
    _check(lib().mcr_run(ctypes.byref(q), ctypes.byref(r)))
    return _result(r, len(alphas), keep)


def init(device: int = 0) -> None:
    """Initialises the library on a device (mcr_init); idempotent."""
    _check(lib().mcr_init(int(device)))


def run_noalloc(mu, sigma, x: float, k: int, paths: int, seed: int = 0, *, weights=None,
                precision: int = PREC_FP32, kernel: int = KERNEL_AUTO, alphas=(),
                near: bool = False, want_S: bool = False, want_loss: bool = False,
                want_hist: bool = False, path_offset: int = 0,
                extra_flags: int = 0) -> RiskResult:
    """One risk query through the allocation-free entry point (mcr_run_noalloc):
    the library's static device scratch, no allocation; needs init() first."""
    q, keep = _make_query(mu, sigma, x, k, paths, seed, weights, precision, kernel, alphas, near,
                          want_S, want_loss, want_hist, path_offset, extra_flags)
    r = Result()
    _check(lib().mcr_run_noalloc(ctypes.byref(q), ctypes.byref(r)))
    return _result(r, len(alphas), keep)


BATCH_PERSISTENT, BATCH_PER_QUERY = 0, 1
ORDER_LPT, ORDER_NATURAL = 0, 1


class BatchOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("chunk_paths", ctypes.c_int), ("order", ctypes.c_int), ("grid", ctypes.c_int),
                ("ms_device", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("chunk_used", ctypes.c_int), ("order_used", ctypes.c_int),
                ("graph_replayed", ctypes.c_int), ("work_items", ctypes.c_int64),
                ("streams", ctypes.c_int), ("shapes", ctypes.c_int * 4),
                ("streams_used", ctypes.c_int), ("big_shapes", ctypes.c_int * 4),
                ("host_plan_ms", ctypes.c_double), ("host_stage_ms", ctypes.c_double),
                ("host_wait_ms", ctypes.c_double), ("host_finalize_ms", ctypes.c_double)]


def _queries(portfolios):
    """portfolios: dicts with mu, sigma, x, k, paths, seed and optional weights,
    alphas, precision, kernel, path_offset, extra_flags."""
    qs = (Query * len(portfolios))()
    keep = []
    for i, p in enumerate(portfolios):
        q, kp = _make_query(p["mu"], p["sigma"], p["x"], p["k"], p["paths"], p.get("seed", 0),
                            p.get("weights"), p.get("precision", PREC_BF16X3),
                            p.get("kernel", KERNEL_AUTO), tuple(p.get("alphas", ())),
                            path_offset=p.get("path_offset", 0),
                            extra_flags=p.get("extra_flags", 0))
        qs[i] = q
        keep.append(kp)
    return qs, keep


def run_batch(portfolios, *, mode: int = BATCH_PERSISTENT, use_graph: bool = True,
              chunk_paths: int = 0, order: int = ORDER_LPT, grid: int = 0, streams: int = 0,
              shapes=(0, 0, 0, 0)):
    """Independent queries in one batch (mcr_run_batch): (results, BatchOpts).
    shapes: per row-tile range, 0 = the tuned workgroup shape, s > 0 = shape s - 1."""
    qs, keep = _queries(portfolios)
    rs = (Result * len(portfolios))()
    o = BatchOpts(mode, int(use_graph), chunk_paths, order, grid)
    o.streams = int(streams)
    for r, sh in enumerate(shapes):
        o.shapes[r] = int(sh)
    _check(lib().mcr_run_batch(qs, len(portfolios), rs, ctypes.byref(o)))
    out = [_result(rs[i], len(portfolios[i].get("alphas", ()))) for i in range(len(portfolios))]
    del keep
    return out, o


def autotune_batch(portfolios):
    """Times the persistent batch's candidate settings; the fastest becomes the
    library default.  Returns (BatchOpts of the winner, candidates timed)."""
    qs, keep = _queries(portfolios)
    o = BatchOpts()
    nc = ctypes.c_int(0)
    _check(lib().mcr_autotune_batch(qs, len(portfolios), ctypes.byref(o), ctypes.byref(nc)))
    del keep
    return o, nc.value


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("cache_entries", ctypes.c_int),
                ("max_active", ctypes.c_int), ("round_ms", ctypes.c_double),
                ("grid", ctypes.c_int), ("streams", ctypes.c_int),
                ("age_limit_ms", ctypes.c_double)]


class TicketInfo(ctypes.Structure):
    _fields_ = [("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
                ("rounds", ctypes.c_int), ("cache_hit", ctypes.c_int), ("status", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("failed", ctypes.c_int64), ("rounds", ctypes.c_int64),
                ("cache_hits", ctypes.c_int64), ("cache_misses", ctypes.c_int64),
                ("paths_done", ctypes.c_int64), ("device_ms", ctypes.c_double),
                ("peak_bytes", ctypes.c_int64), ("budget_bytes", ctypes.c_int64),
                ("active", ctypes.c_int), ("queued", ctypes.c_int),
                ("host_admit_ms", ctypes.c_double), ("host_plan_ms", ctypes.c_double),
                ("host_enqueue_ms", ctypes.c_double), ("host_wait_ms", ctypes.c_double),
                ("host_finalize_ms", ctypes.c_double)]


def _server_lib():
    L = lib()
    if not getattr(L, "_server_sigs", False):
        vp = ctypes.c_void_p
        L.mcr_server_footprint.argtypes = [ctypes.POINTER(Query), ctypes.POINTER(ctypes.c_int64)]
        L.mcr_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerOpts),
                                        ctypes.POINTER(vp)]
        L.mcr_server_submit.argtypes = [vp, ctypes.POINTER(Query), ctypes.POINTER(ctypes.c_uint64)]
        L.mcr_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(Result),
                                      ctypes.POINTER(TicketInfo)]
        L.mcr_server_poll.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(Result),
                                      ctypes.POINTER(TicketInfo)]
        L.mcr_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.mcr_server_destroy.argtypes = [vp]
        L._server_sigs = True
    return L


def server_footprint(portfolio) -> int:
    """Device bytes a query holds while admitted (mcr_server_footprint; host only)."""
    qs, keep = _queries([portfolio])
    b = ctypes.c_int64(0)
    _check(_server_lib().mcr_server_footprint(ctypes.byref(qs[0]), ctypes.byref(b)))
    del keep
    return b.value


class Server:
    """Serving loop (mcr_server_*): submit queries as they arrive, wait for
    their results.  Portfolios are dicts as for run_batch."""

    def __init__(self, device: int = 0, *, hbm_budget: int = 0, cache_entries: int = 0,
                 max_active: int = 0, round_ms: float = 0.0, grid: int = 0, streams: int = 0,
                 age_limit_ms: float = 0.0):
        self._L = _server_lib()
        self._h = ctypes.c_void_p()
        o = ServerOpts(hbm_budget, cache_entries, max_active, round_ms, grid, streams, age_limit_ms)
        _check(self._L.mcr_server_create(device, ctypes.byref(o), ctypes.byref(self._h)))
        self._na = {}

    def submit(self, portfolio) -> int:
        qs, keep = _queries([portfolio])
        t = ctypes.c_uint64(0)
        _check(self._L.mcr_server_submit(self._h, ctypes.byref(qs[0]), ctypes.byref(t)))
        del keep  # the server copied the inputs
        self._na[t.value] = len(portfolio.get("alphas", ()))
        return t.value

    def wait(self, ticket: int):
        """(RiskResult, TicketInfo) of a ticket; raises McrError on failure."""
        r, info = Result(), TicketInfo()
        code = self._L.mcr_server_wait(self._h, ticket, ctypes.byref(r), ctypes.byref(info))
        na = self._na.pop(ticket, 0)
        _check(code)
        return _result(r, na), info

    def poll(self, ticket: int):
        """None while pending, else as wait()."""
        r, info = Result(), TicketInfo()
        code = self._L.mcr_server_poll(self._h, ticket, ctypes.byref(r), ctypes.byref(info))
        if code == 0:
            return None
        na = self._na.pop(ticket, 0)
        _check(0 if code == 1 else code)
        return _result(r, na), info

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(self._L.mcr_server_get_stats(self._h, ctypes.byref(s)))
        return s

    def close(self) -> None:
        if self._h:
            _check(self._L.mcr_server_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def factor(sigma):
    """GPU FP64 Cholesky with the jitter policy: (L, attempts, jitter, clamped)."""
    sigma = _square(sigma)
    n = sigma.shape[0]
    L = np.zeros((n, n), dtype=np.float64)
    info = FactorInfo()
    _check(lib().mcr_factor(n, _ptr(sigma, _f64p), n, _ptr(L, _f64p), ctypes.byref(info)))
    return L, info.attempts, info.jitter, info.clamped


def prepare(mu, sigma, x: float, weights=None) -> dict:
    """The fp32 model the path kernels read (factor computed on the GPU)."""
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    sigma = _square(sigma)
    n = mu.shape[0]
    if mu.ndim != 1 or sigma.shape != (n, n):
        raise ValueError(f"mu must be ({n},) and sigma {n} x {n} (got {mu.shape}, {sigma.shape})")
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    if w is not None and w.shape != (n,):
        raise ValueError(f"weights must have shape ({n},), not {w.shape}")
    q = Query()
    q.n, q.mu, q.sigma, q.ld_sigma = n, _ptr(mu, _f64p), _ptr(sigma, _f64p), n
    q.weights, q.x, q.k, q.paths = _ptr(w, _f64p), float(x), 1, 1
    q.precision, q.kernel = PREC_FP32, KERNEL_AUTO
    L64 = np.zeros((n, n), dtype=np.float64)
    L32 = np.zeros((n, n), dtype=np.float32)
    mu32 = np.zeros(n, dtype=np.float32)
    v32 = np.zeros(n, dtype=np.float32)
    ex = ModelExport()
    ex.L64, ex.L32 = _ptr(L64, _f64p), _ptr(L32, _f32p)
    ex.mu32, ex.v32 = _ptr(mu32, _f32p), _ptr(v32, _f32p)
    _check(lib().mcr_prepare(ctypes.byref(q), ctypes.byref(ex)))
    return {"L64": L64, "L32": L32, "mu32": mu32, "v32": v32, "mp32": np.float32(ex.mp32),
            "thr32": np.float32(ex.thr32), "loss_sd": ex.loss_sd}


def normals(seed: int, path0: int, npaths: int, n: int) -> np.ndarray:
    """The device normal stream, z[p, j] for paths path0 .. path0+npaths-1."""
    out = np.zeros((npaths, n), dtype=np.float32)
    _check(lib().mcr_normals(int(seed) & 0xFFFFFFFFFFFFFFFF, int(path0), int(npaths), int(n),
                             _ptr(out, _f32p)))
    return out


def synth_factor_model(n: int, K: int, T: int, seed: int, drift: float = 2e-4,
                       fvol: float = 0.01, ivol: float = 0.01, want_returns: bool = False):
    """Seeded factor-model returns -> (mu, Sigma[, returns]) (host computation)."""
    mu = np.zeros(n, dtype=np.float64)
    sigma = np.zeros((n, n), dtype=np.float64)
    R = np.zeros((T, n), dtype=np.float64) if want_returns else None
    _check(lib().mcr_synth_factor_model(n, K, T, drift, fvol, ivol, int(seed), _ptr(mu, _f64p),
                                        _ptr(sigma, _f64p), _ptr(R, _f64p)))
    return (mu, sigma, R) if want_returns else (mu, sigma)


# ---------------------------------------------------------------------------
# Mirror of the reference's Python prototype (python_serial.py) on the GPU.
# ---------------------------------------------------------------------------
def monte_carlo_crash_simulation(x, N, k, mu, Sigma, M, seed: int = 0,
                                 precision: int = PREC_FP32):
    """P(at least k of N assets fall below -x), as python_serial.py:11-75.

    Returns (P_hat, count, S_values) with S_values the per-path crash counts.
    """
    mu = np.asarray(mu, dtype=np.float64)
    if mu.shape[0] != N:
        raise ValueError(f"N = {N} does not match mu ({mu.shape[0]})")
    r = run(mu, Sigma, x, k, M, seed, precision=precision, want_S=True)
    return r.p_hat, r.count, r.S


def compute_accuracy(P_hat, M):
    """Standard error and clamped 95% CI, as python_serial.py:77-105."""
    std_error = math.sqrt(P_hat * (1 - P_hat) / M)
    margin = 1.96 * std_error
    return std_error, max(0, P_hat - margin), min(1, P_hat + margin)


def get_hip_model() -> ModelFunctions:
    """The plug-in model record {name: "hip", run_model: hip_simulate}."""
    return lib().get_hip_model()


def make_params(mu, Sigma, x: float, k: int, M: int, seed: int, *, mu_stride: int = 1,
                sigma_tda: int | None = None):
    """MonteCarloParams with GSL-layout views (keep the returned tuple alive)."""
    mu = np.asarray(mu, dtype=np.float64)
    Sigma = np.asarray(Sigma, dtype=np.float64)
    n = mu.shape[0]
    mbuf = np.zeros(n * mu_stride, dtype=np.float64)
    mbuf[::mu_stride] = mu
    tda = sigma_tda or n
    sbuf = np.zeros(n * tda, dtype=np.float64)
    for i in range(n):
        sbuf[i * tda:i * tda + n] = Sigma[i]
    gv = GslVector(n, mu_stride, mbuf.ctypes.data_as(_f64p), None, 0)
    gm = GslMatrix(n, n, tda, sbuf.ctypes.data_as(_f64p), None, 0)
    p = MonteCarloParams(float(x), n, int(k), ctypes.pointer(gv), ctypes.pointer(gm), int(M),
                         int(seed))
    return p, (mbuf, sbuf, gv, gm)


# ---------------------------------------------------------------- data I/O ---
class _Index(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("file", ctypes.c_char_p)]


class _Config(ctypes.Structure):
    _fields_ = [("start", ctypes.c_char_p), ("end", ctypes.c_char_p), ("x", ctypes.c_double),
                ("N", ctypes.c_int), ("M", ctypes.c_int), ("k", ctypes.c_int),
                ("random_seed", ctypes.c_ulong), ("train_ratio", ctypes.c_double),
                ("indices", ctypes.POINTER(_Index)), ("num_indices", ctypes.c_int),
                ("models", ctypes.POINTER(ctypes.c_char_p)), ("num_models", ctypes.c_int),
                ("comment", ctypes.c_char_p)]


class CsvRow(ctypes.Structure):
    _fields_ = [("iteration_id", ctypes.c_int), ("timestamp", ctypes.c_long),
                ("execution_time_ms", ctypes.c_long), ("comment", ctypes.c_char_p),
                ("start_date", ctypes.c_char_p), ("end_date", ctypes.c_char_p),
                ("train_ratio", ctypes.c_double), ("M", ctypes.c_int), ("k", ctypes.c_int),
                ("x", ctypes.c_double), ("model_name", ctypes.c_char_p),
                ("seed", ctypes.c_ulong), ("nodes", ctypes.c_int), ("threads", ctypes.c_int),
                ("processes", ctypes.c_int), ("indices", ctypes.POINTER(_Index)),
                ("num_indices", ctypes.c_int), ("actual_freq", ctypes.c_double),
                ("P_hat", ctypes.c_double), ("count", ctypes.c_int),
                ("std_error", ctypes.c_double), ("ci_lower", ctypes.c_double),
                ("ci_upper", ctypes.c_double)]


RUNNER_PATH = os.path.join(_HERE, "_lib", "mcr_runner")


def _io_lib():
    L = lib()
    if not getattr(L, "_io_sigs", False):
        dpp = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
        ip = ctypes.POINTER(ctypes.c_int)
        L.mcr_read_mu_bin.argtypes = [ctypes.c_char_p, dpp, ip]
        L.mcr_read_sigma_bin.argtypes = [ctypes.c_char_p, dpp, ip, ip]
        L.mcr_read_params_bin.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        L.mcr_write_bins.argtypes = [ctypes.c_char_p, _f64p, ctypes.c_int, _f64p,
                                     ctypes.c_double]
        L.mcr_free.argtypes = [ctypes.c_void_p]
        L.mcr_load_config.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Config)]
        L.mcr_free_config.argtypes = [ctypes.POINTER(_Config)]
        L.mcr_csv_header.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.mcr_csv_format_row.argtypes = [ctypes.POINTER(CsvRow), ctypes.c_char_p, ctypes.c_int]
        L.mcr_append_csv.argtypes = [ctypes.c_char_p, ctypes.POINTER(CsvRow)]
        L._io_sigs = True
    return L


def read_mu_bin(path) -> np.ndarray:
    """mu.bin (int32 N + N float64), as load_mu_binary reads it."""
    L = _io_lib()
    p, n = ctypes.POINTER(ctypes.c_double)(), ctypes.c_int(0)
    _check(L.mcr_read_mu_bin(os.fsencode(path), ctypes.byref(p), ctypes.byref(n)))
    out = np.ctypeslib.as_array(p, shape=(n.value,)).copy()
    L.mcr_free(p)
    return out


def read_sigma_bin(path) -> np.ndarray:
    """sigma.bin (int32 rows, int32 cols + row-major float64)."""
    L = _io_lib()
    p, r, c = ctypes.POINTER(ctypes.c_double)(), ctypes.c_int(0), ctypes.c_int(0)
    _check(L.mcr_read_sigma_bin(os.fsencode(path), ctypes.byref(p), ctypes.byref(r),
                                ctypes.byref(c)))
    out = np.ctypeslib.as_array(p, shape=(r.value, c.value)).copy()
    L.mcr_free(p)
    return out


def read_params_bin(path) -> float:
    """params.bin (one float64 in [0, 1])."""
    f = ctypes.c_double(0.0)
    _check(_io_lib().mcr_read_params_bin(os.fsencode(path), ctypes.byref(f)))
    return f.value


def write_bins(out_dir, mu, sigma, actual_freq: float) -> None:
    """out_dir/mu.bin, sigma.bin, params.bin in the reference's format."""
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    n = int(mu.shape[0])
    if sigma.shape != (n, n):
        raise ValueError(f"sigma must be {n} x {n}")
    _check(_io_lib().mcr_write_bins(os.fsencode(out_dir), _ptr(mu, _f64p), n, _ptr(sigma, _f64p),
                                    float(actual_freq)))


def load_config(path) -> dict:
    """config.yaml as the C loader reads it (mcr_load_config)."""
    L = _io_lib()
    c = _Config()
    _check(L.mcr_load_config(os.fsencode(path), ctypes.byref(c)))
    dec = lambda b: None if b is None else b.decode()  # noqa: E731
    out = {"start": dec(c.start), "end": dec(c.end), "x": c.x, "N": c.N, "M": c.M, "k": c.k,
           "random_seed": c.random_seed, "train_ratio": c.train_ratio,
           "indices": {c.indices[i].name.decode(): c.indices[i].file.decode()
                       for i in range(c.num_indices)},
           "models": [c.models[i].decode() for i in range(c.num_models)],
           "comment": dec(c.comment)}
    L.mcr_free_config(ctypes.byref(c))
    return out


def _csv_row(fields: dict):
    enc = lambda s: None if s is None else str(s).encode()  # noqa: E731
    idx = list((fields.get("indices") or {}).items())
    arr = (_Index * max(1, len(idx)))()
    for i, (nm, fl) in enumerate(idx):
        arr[i] = _Index(enc(nm), enc(fl))
    row = CsvRow()
    for name, _t in CsvRow._fields_:
        if name in ("indices", "num_indices") or name not in fields:
            continue
        v = fields[name]
        setattr(row, name, enc(v) if name in ("comment", "start_date", "end_date",
                                               "model_name") else v)
    row.indices = arr if idx else None
    row.num_indices = len(idx)
    return row, arr


def csv_header() -> str:
    buf = ctypes.create_string_buffer(1024)
    n = _io_lib().mcr_csv_header(buf, len(buf))
    if n < 0:
        _check(n)
    return buf.value.decode()


def csv_format_row(**fields) -> str:
    """The results-CSV line for the given fields (names of mcr_csv_row)."""
    row, keep = _csv_row(fields)
    buf = ctypes.create_string_buffer(1 << 16)
    n = _io_lib().mcr_csv_format_row(ctypes.byref(row), buf, len(buf))
    if n < 0:
        _check(n)
    del keep
    return buf.value.decode()


def append_csv(path, **fields) -> None:
    row, keep = _csv_row(fields)
    _check(_io_lib().mcr_append_csv(os.fsencode(path), ctypes.byref(row)))
    del keep


def preprocessing():
    """The real-data preprocessing module (preprocess.py)."""
    import importlib

    return importlib.import_module(__name__ + ".preprocess")


def workloads():
    """The seeded benchmark workloads (workloads.py)."""
    import importlib

    return importlib.import_module(__name__ + ".workloads")

