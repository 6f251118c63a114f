"""ctypes binding of libpgpu.so (the C ABI in include/pgpu.h).

The product path runs only through the HIP library: if libpgpu.so is missing
or no gfx950 device can be opened, every call raises PgError.  There is no
CPU fallback here.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgpu.so")

PG_OK, PG_EINVAL, PG_ENOMEM, PG_EDEVICE, PG_ERANGE, PG_ENOINIT, PG_EINTERNAL = 0, -1, -2, -3, -4, -5, -6
PATH_AUTO, PATH_NTT, PATH_FFT, PATH_NTT_REF = 0, 1, 2, 3
P64 = ctypes.POINTER(ctypes.c_uint64)
SZ = ctypes.c_size_t

# Every symbol include/pgpu.h declares.
EXPORTS = ["pg_init", "pg_shutdown", "pg_last_error", "pg_mul", "pg_buf_alloc", "pg_buf_free",
           "pg_buf_capacity", "pg_buf_upload", "pg_buf_download", "pg_buf_ptr", "pg_mul_dev",
           "pg_mul_batch", "pg_batch_plan", "pg_batch_run", "pg_batch_errors", "pg_batch_free",
           "pg_sync", "pg_timing_enable", "pg_timing_reset", "pg_timing_read",
           "pg_div", "pg_div_dev", "pg_recip", "pg_recip_dev", "pg_sqrt", "pg_to_decimal",
           "pg_chud_split", "pg_pi", "pg_server_start", "pg_server_submit", "pg_server_wait",
           "pg_server_drain", "pg_server_read_stats", "pg_server_stop", "pg_server_clock",
           "pg_ntt_autotune", "pg_ntt_plan_get", "pg_ntt_plan_set", "pg_ntt_batch_plan_get",
           "pg_ntt_batch_plan_set", "pg_host_staging",
           "pg_mul_dev_static", "pg_from_decimal", "pg_debug_perturb"]
NTT_VARIANTS = 16
OP_MUL, OP_DIV = 0, 1


class Job(ctypes.Structure):
    """pg_job (include/pgpu.h)."""
    _fields_ = [("a", ctypes.c_void_p), ("an", SZ), ("b", ctypes.c_void_p), ("bn", SZ),
                ("r", ctypes.c_void_p), ("r_cap", SZ), ("r_len", SZ), ("path", ctypes.c_int),
                ("status", ctypes.c_int), ("fft_max_round", ctypes.c_double)]


class Request(ctypes.Structure):
    """pg_request (include/pgpu.h)."""
    _fields_ = [("op", ctypes.c_int), ("a", ctypes.c_void_p), ("an", SZ), ("b", ctypes.c_void_p), ("bn", SZ),
                ("r", ctypes.c_void_p), ("r_cap", SZ), ("r2", ctypes.c_void_p), ("r2_cap", SZ),
                ("r_len", SZ), ("r2_len", SZ), ("status", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rounds", ctypes.c_uint64),
                ("products", ctypes.c_uint64), ("max_active", ctypes.c_uint64)]


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pgpu error %d: %s" % (code, msg))
        self.code = code


class StageTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("launches", ctypes.c_uint64),
                ("total_ms", ctypes.c_double), ("bytes", ctypes.c_double)]


_lib = None


def load(path=None):
    """Loads libpgpu.so (no device call).  PGPU_LIB overrides the path."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("PGPU_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise PgError(PG_ENOINIT, "libpgpu.so not built (python -m preclib_amd.build)")
    L = ctypes.CDLL(path)
    L.pg_init.argtypes = [ctypes.c_int, SZ]
    L.pg_last_error.restype = ctypes.c_char_p
    L.pg_mul.argtypes = [P64, SZ, ctypes.POINTER(SZ), P64, SZ, P64, SZ, ctypes.c_int,
                         ctypes.POINTER(ctypes.c_double)]
    L.pg_buf_alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), SZ]
    L.pg_buf_free.argtypes = [ctypes.c_void_p]
    L.pg_buf_capacity.argtypes = [ctypes.c_void_p]
    L.pg_buf_capacity.restype = SZ
    L.pg_buf_ptr.argtypes = [ctypes.c_void_p]
    L.pg_buf_ptr.restype = ctypes.c_void_p
    L.pg_buf_upload.argtypes = [ctypes.c_void_p, P64, SZ]
    L.pg_buf_download.argtypes = [ctypes.c_void_p, P64, SZ]
    L.pg_mul_dev.argtypes = [ctypes.c_void_p, ctypes.POINTER(SZ), ctypes.c_void_p, SZ, ctypes.c_void_p,
                             SZ, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.pg_mul_batch.argtypes = [ctypes.POINTER(Job), SZ]
    L.pg_batch_plan.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Job), SZ]
    L.pg_batch_run.argtypes = [ctypes.c_void_p]
    L.pg_batch_errors.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), SZ]
    L.pg_batch_free.argtypes = [ctypes.c_void_p]
    L.pg_div.argtypes = [P64, SZ, ctypes.POINTER(SZ), P64, SZ, ctypes.POINTER(SZ), P64, SZ, P64, SZ]
    L.pg_div_dev.argtypes = [ctypes.c_void_p, SZ, ctypes.POINTER(SZ), ctypes.c_void_p, SZ, ctypes.POINTER(SZ),
                             ctypes.c_void_p, SZ, ctypes.c_void_p, SZ]
    L.pg_recip.argtypes = [P64, SZ, ctypes.POINTER(SZ), P64, SZ, SZ]
    L.pg_recip_dev.argtypes = [ctypes.c_void_p, SZ, ctypes.POINTER(SZ), ctypes.c_void_p, SZ, SZ]
    L.pg_sqrt.argtypes = [P64, SZ, ctypes.POINTER(SZ), P64, SZ]
    L.pg_to_decimal.argtypes = [ctypes.c_char_p, SZ, ctypes.POINTER(SZ), P64, SZ]
    L.pg_chud_split.argtypes = [P64, SZ, ctypes.POINTER(SZ), P64, SZ, ctypes.POINTER(SZ), P64, SZ,
                                ctypes.POINTER(SZ), SZ]
    L.pg_pi.argtypes = [ctypes.c_char_p, SZ, ctypes.POINTER(SZ), SZ, ctypes.POINTER(ctypes.c_double)]
    L.pg_server_start.argtypes = [ctypes.POINTER(ctypes.c_void_p), SZ, SZ]
    L.pg_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
    L.pg_server_wait.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
    L.pg_server_drain.argtypes = [ctypes.c_void_p]
    L.pg_server_read_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
    L.pg_server_stop.argtypes = [ctypes.c_void_p]
    L.pg_server_clock.restype = ctypes.c_double
    L.pg_ntt_autotune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_int)]
    L.pg_ntt_plan_get.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.pg_ntt_plan_set.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.pg_ntt_batch_plan_get.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.pg_ntt_batch_plan_set.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.pg_timing_enable.argtypes = [ctypes.c_int]
    L.pg_timing_read.argtypes = [ctypes.POINTER(StageTime), ctypes.c_int]
    _lib = L
    return L


def _check(code):
    if code != PG_OK:
        msg = _lib.pg_last_error()
        raise PgError(code, msg.decode() if msg else "")
    return code


_inited = False


def init(device=0, budget=0):
    global _inited
    L = load()
    _check(L.pg_init(device, budget))
    _inited = True


def _ptr(a):
    return a.ctypes.data_as(P64)


def _trimmed(x):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    n = len(x)
    while n and x[n - 1] == 0:
        n -= 1
    return x[:n]


def mul(a, b, path=PATH_AUTO, want_round=False):
    """Product of two limb arrays on the GPU (host buffers in and out)."""
    if not _inited:
        init()
    a, b = _trimmed(a), _trimmed(b)
    r = np.zeros(len(a) + len(b), np.uint64)
    rl = SZ(0)
    err = ctypes.c_double(0.0)
    _check(_lib.pg_mul(_ptr(r), len(r), ctypes.byref(rl), _ptr(a), len(a), _ptr(b), len(b), path,
                       ctypes.byref(err)))
    r = r[: rl.value]
    return (r, err.value) if want_round else r


class DevBuf:
    """A limb array resident in HBM (pg_buf handle)."""

    def __init__(self, limbs):
        if not _inited:
            init()
        h = ctypes.c_void_p()
        _check(_lib.pg_buf_alloc(ctypes.byref(h), limbs))
        self.h = h
        self.cap = limbs

    @classmethod
    def of(cls, x):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        d = cls(max(len(x), 1))
        d.upload(x)
        return d

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        _check(_lib.pg_buf_upload(self.h, _ptr(x), len(x)))

    def download(self, n):
        out = np.zeros(n, np.uint64)
        _check(_lib.pg_buf_download(self.h, _ptr(out), n))
        return out

    def free(self):
        if self.h:
            _lib.pg_buf_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mul_dev(r, a, an, b, bn, path=PATH_AUTO, want_len=True):
    rl = SZ(0)
    err = ctypes.c_double(0.0)
    _check(_lib.pg_mul_dev(r.h, ctypes.byref(rl) if want_len else None, a.h, an, b.h, bn, path,
                           ctypes.byref(err)))
    return rl.value, err.value


def sync():
    _check(_lib.pg_sync())


def mul_dev_static(r, a, an, b, bn):
    """pg_mul_dev_static on DevBuf handles' device pointers (asynchronous)."""
    L = load()
    L.pg_mul_dev_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, SZ, ctypes.c_void_p, SZ]
    _check(L.pg_mul_dev_static(L.pg_buf_ptr(r.h), L.pg_buf_ptr(a.h), an, L.pg_buf_ptr(b.h), bn))


def from_decimal(s):
    """The integer of a decimal string as limbs (pg_from_decimal)."""
    L = load()
    if not _inited:
        init()
    b = s.encode() if isinstance(s, str) else bytes(s)
    cap = len(b) * 3322 // 64000 + 2
    out = np.zeros(max(cap, 1), np.uint64)
    n = SZ(0)
    L.pg_from_decimal.argtypes = [P64, SZ, ctypes.POINTER(SZ), ctypes.c_char_p, SZ]
    _check(L.pg_from_decimal(_ptr(out), cap, ctypes.byref(n), b, len(b)))
    return out[: n.value].copy()


def debug_perturb(what, amount):
    """pg_debug_perturb: one-shot perturbation of the next division (tests)."""
    L = load()
    L.pg_debug_perturb.argtypes = [ctypes.c_int, ctypes.c_int64]
    _check(L.pg_debug_perturb(what, amount))


def host_staging():
    """(held, spare) bytes of page-locked descriptor images (pg_host_staging)."""
    L = load()
    held, spare = SZ(0), SZ(0)
    _check(L.pg_host_staging(ctypes.byref(held), ctypes.byref(spare)))
    return held.value, spare.value


def timing(on=True):
    load().pg_timing_enable(1 if on else 0)


def timing_reset():
    load().pg_timing_reset()


def timing_read():
    L = load()
    n = L.pg_timing_read(None, 0)
    arr = (StageTime * max(n, 1))()
    n = L.pg_timing_read(arr, n)
    return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms, arr[i].bytes) for i in range(n)}


def divmod_(a, b):
    """(floor(a / b), a mod b) of limb arrays on the GPU (div_mulinv semantics:
    b == 0 gives two empty arrays)."""
    if not _inited:
        init()
    a, b = _trimmed(a), _trimmed(b)
    q = np.zeros(max(len(a) - len(b) + 1, 1), np.uint64)
    r = np.zeros(max(len(b), 1), np.uint64)
    ql, rl = SZ(0), SZ(0)
    _check(_lib.pg_div(_ptr(q), len(q), ctypes.byref(ql), _ptr(r), len(r), ctypes.byref(rl),
                       _ptr(a), len(a), _ptr(b), len(b)))
    return q[:ql.value].copy(), r[:rl.value].copy()


def recip(b, bits):
    """floor(2^bits / b) on the GPU (precn_reciprocal_newton semantics)."""
    if not _inited:
        init()
    b = _trimmed(b)
    x = np.zeros(max(bits // 64 + 2 - max(len(b) - 1, 0), 1), np.uint64)
    xl = SZ(0)
    _check(_lib.pg_recip(_ptr(x), len(x), ctypes.byref(xl), _ptr(b), len(b), bits))
    return x[:xl.value].copy()


def ntt_autotune(lo=12, hi=25, reps=5):
    """Times every NTT plan variant of 2^lo .. 2^hi; returns {logn: (chosen, [ms per variant])}."""
    if not _inited:
        init()
    k = hi - lo + 1
    ms = (ctypes.c_double * (k * NTT_VARIANTS))()
    ch = (ctypes.c_int * k)()
    _check(_lib.pg_ntt_autotune(lo, hi, reps, ms, ch))
    return {lo + i: (ch[i], [ms[i * NTT_VARIANTS + v] for v in range(NTT_VARIANTS)]) for i in range(k)}


def ntt_plans():
    if not _inited:
        init()
    t = (ctypes.c_int * 27)()
    _check(_lib.pg_ntt_plan_get(t, 27))
    return list(t)


def set_ntt_plans(table):
    if not _inited:
        init()
    t = (ctypes.c_int * len(table))(*table)
    _check(_lib.pg_ntt_plan_set(t, len(table)))


def ntt_batch_plans():
    """Plan variants of large batched launches per logn (-1: the single-product one)."""
    if not _inited:
        init()
    t = (ctypes.c_int * 27)()
    _check(_lib.pg_ntt_batch_plan_get(t, 27))
    return list(t)


def set_ntt_batch_plans(table):
    if not _inited:
        init()
    t = (ctypes.c_int * len(table))(*table)
    _check(_lib.pg_ntt_batch_plan_set(t, len(table)))


def isqrt(a):
    """floor(sqrt(a)) on the GPU (precn_sqrt semantics)."""
    if not _inited:
        init()
    a = _trimmed(a)
    x = np.zeros(len(a) // 2 + 2, np.uint64)
    xl = SZ(0)
    _check(_lib.pg_sqrt(_ptr(x), len(x), ctypes.byref(xl), _ptr(a), len(a)))
    return x[:xl.value].copy()


def to_decimal(a):
    """Decimal digits of a limb array (most significant first; '' for zero)."""
    if not _inited:
        init()
    a = _trimmed(a)
    cap = int(len(a) * 19.3) + 2
    buf = np.empty(cap, np.uint8)   # (not zero-filled; the library writes the n bytes)
    n = SZ(0)
    _check(_lib.pg_to_decimal(buf.ctypes.data_as(ctypes.c_char_p), cap, ctypes.byref(n), _ptr(a), len(a)))
    return str(memoryview(buf)[: n.value], "ascii")   # decoded in place, one copy


def chud_split(nterms, need_p=True):
    """(P, Q, T) of the Chudnovsky series over [0, nterms) (T as magnitude)."""
    if not _inited:
        init()
    cap = max(4 * nterms + 8, 16)
    p, q, t = (np.zeros(cap, np.uint64) for _ in range(3))
    pl, ql, tl = SZ(0), SZ(0), SZ(0)
    _check(_lib.pg_chud_split(_ptr(p) if need_p else None, cap, ctypes.byref(pl), _ptr(q), cap, ctypes.byref(ql),
                              _ptr(t), cap, ctypes.byref(tl), nterms))
    return (p[:pl.value].copy() if need_p else None), q[:ql.value].copy(), t[:tl.value].copy()


def pi(digits, want_phases=False):
    """The pi program's digit string for `digits` digits (no trailing newline)."""
    if not _inited:
        init()
    cap = digits + 64
    buf = np.empty(cap, np.uint8)   # (not zero-filled; the library writes the n bytes)
    n = SZ(0)
    ph = (ctypes.c_double * 4)()
    _check(_lib.pg_pi(buf.ctypes.data_as(ctypes.c_char_p), cap, ctypes.byref(n), digits, ph))
    s = str(memoryview(buf)[: n.value], "ascii")   # decoded in place, one copy
    return (s, list(ph)) if want_phases else s


def div_dev(q, r, a, an, b, bn):
    """Device-resident division: q, r (DevBuf or None), a, b DevBufs.
    Returns (q_len, r_len)."""
    ql, rl = SZ(0), SZ(0)
    _check(_lib.pg_div_dev(q.h if q else None, q.cap if q else 0, ctypes.byref(ql),
                           r.h if r else None, r.cap if r else 0, ctypes.byref(rl), a.h, an, b.h, bn))
    return ql.value, rl.value


def mul_batch(pairs, path=PATH_AUTO, want_round=False):
    """Products of many (a, b) pairs in one submission (host buffers)."""
    if not _inited:
        init()
    keep, jobs = [], (Job * max(len(pairs), 1))()
    outs = []
    for i, (a, b) in enumerate(pairs):
        a, b = _trimmed(a), _trimmed(b)
        r = np.zeros(len(a) + len(b), np.uint64)
        keep += [a, b]
        outs.append(r)
        jobs[i] = Job(a.ctypes.data, len(a), b.ctypes.data, len(b), r.ctypes.data, len(r), 0, path, 0, 0.0)
    _check(_lib.pg_mul_batch(jobs, len(pairs)))
    res = [outs[i][: jobs[i].r_len] for i in range(len(pairs))]
    if want_round:
        return res, [jobs[i].fft_max_round for i in range(len(pairs))]
    return res


class DevBatch:
    """A planned device-resident batch (pg_batch): products of DevBuf pairs."""

    def __init__(self, triples, path=PATH_AUTO):
        """triples: list of (r: DevBuf, a: DevBuf, an, b: DevBuf, bn)."""
        if not _inited:
            init()
        self.n = len(triples)
        jobs = (Job * max(self.n, 1))()
        for i, (r, a, an, b, bn) in enumerate(triples):
            jobs[i] = Job(_lib.pg_buf_ptr(a.h), an, _lib.pg_buf_ptr(b.h), bn, _lib.pg_buf_ptr(r.h), r.cap, 0,
                          path, 0, 0.0)
        h = ctypes.c_void_p()
        _check(_lib.pg_batch_plan(ctypes.byref(h), jobs, self.n))
        self.h = h

    def run(self):
        _check(_lib.pg_batch_run(self.h))

    def errors(self):
        e = (ctypes.c_double * max(self.n, 1))()
        _check(_lib.pg_batch_errors(self.h, e, self.n))
        return [e[i] for i in range(self.n)]

    def free(self):
        if self.h:
            _lib.pg_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ServerRequest:
    """One submitted request; its arrays stay referenced until it is read."""

    def __init__(self, op, a, b):
        self.a, self.b = _trimmed(a), _trimmed(b)
        an, bn = len(self.a), len(self.b)
        if op == OP_MUL:
            self.r, self.r2 = np.zeros(max(an + bn, 1), np.uint64), None
        else:
            self.r, self.r2 = np.zeros(max(an - bn + 1, 1), np.uint64), np.zeros(max(bn, an, 1), np.uint64)
        self.req = Request(op, self.a.ctypes.data if an else None, an, self.b.ctypes.data if bn else None, bn,
                           self.r.ctypes.data, len(self.r),
                           self.r2.ctypes.data if self.r2 is not None else None,
                           len(self.r2) if self.r2 is not None else 0, 0, 0, 0, 0.0, 0.0, 0.0)

    def result(self):
        _check(self.req.status)
        q = self.r[: self.req.r_len].copy()
        if self.r2 is None:
            return q
        return q, self.r2[: self.req.r2_len].copy()

    @property
    def latency(self):
        return self.req.t_done - self.req.t_submit


class Server:
    """The continuous-batching request loop (pg_server_*)."""

    def __init__(self, budget=0, max_active=0):
        if not _inited:
            init()
        h = ctypes.c_void_p()
        _check(_lib.pg_server_start(ctypes.byref(h), budget, max_active))
        self.h = h
        # the worker writes into a request's struct and arrays until it
        # finishes: every submitted request stays referenced until then
        self._live = []

    def submit(self, op, a, b):
        r = ServerRequest(op, a, b)
        _check(_lib.pg_server_submit(self.h, ctypes.byref(r.req)))
        self._live.append(r)
        return r

    def wait(self, r):
        _lib.pg_server_wait(self.h, ctypes.byref(r.req))
        return r.result()

    def drain(self):
        _check(_lib.pg_server_drain(self.h))
        self._live = []

    def stats(self):
        st = ServerStats()
        _check(_lib.pg_server_read_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def stop(self):
        if self.h:
            _lib.pg_server_stop(self.h)   # finishes queued work first
            self.h = None
            self._live = []

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass

    @staticmethod
    def clock():
        return load().pg_server_clock()
