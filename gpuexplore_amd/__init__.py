"""GPUexplore for MI355X: Python view of the C ABI in include/gpuexplore.h.

The library (gpuexplore_amd/lib/libgpuexplore.so, built by
``gpuexplore_amd.build``) does all the work; this module only marshals
arguments with ctypes.  Nothing here falls back to a CPU implementation: if
the library is missing, :func:`lib` raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgpuexplore.so")
CLI_PATH = os.path.join(HERE, "bin", "GPUexplore")

GX_OK = 0
ERRORS = {
    -1: "NOFILE", -2: "FORMAT", -3: "RANGE", -4: "NOMEM", -5: "HIP", -6: "NODEVICE",
    -7: "ARG", -8: "LIMIT", -9: "STATE", -10: "CAPACITY", -11: "SCRATCH",
}
NOALLOC_MAX_KEYS = 262144  # GX_NOALLOC_MAX_KEYS
PROP_NONE, PROP_DEADLOCK, PROP_SAFETY, PROP_LIVENESS = 0, 1, 2, 3
POR_CAMPLE, POR_AMPLE, POR_STUBBORN = 0, 1, 2  # gx_options.por_method


class GxError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {ERRORS.get(status, status)} ({status})")


class Options(ctypes.Structure):
    _fields_ = [
        ("blocks", ctypes.c_uint32),
        ("threads", ctypes.c_uint32),
        ("kernel_iters", ctypes.c_uint32),
        ("bucket_bytes", ctypes.c_uint32),
        ("q_ints", ctypes.c_uint64),
        ("property", ctypes.c_int32),
        ("por", ctypes.c_int32),
        ("cycle_proviso", ctypes.c_int32),
        ("verbosity", ctypes.c_int32),
        ("hash_seed", ctypes.c_uint64),
        ("cand_capacity", ctypes.c_uint64),
        ("frontier_capacity", ctypes.c_uint64),
        ("max_states", ctypes.c_uint64),
        ("expand_kernel", ctypes.c_uint32),
        ("insert_blocks", ctypes.c_uint32),
        ("por_method", ctypes.c_uint32),
        ("dedup", ctypes.c_uint32),
        ("reserved2", ctypes.c_uint32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_uint64),
        ("transitions", ctypes.c_uint64),
        ("levels", ctypes.c_uint64),
        ("lookups", ctypes.c_uint64),
        ("inserts", ctypes.c_uint64),
        ("hash_launches", ctypes.c_uint64),
        ("table_bytes", ctypes.c_uint64),
        ("violation", ctypes.c_int32),
        ("table_full", ctypes.c_int32),
        ("truncated", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("t_explore_s", ctypes.c_double),
        ("t_hash_s", ctypes.c_double),
        ("t_expand_s", ctypes.c_double),
        ("t_init_s", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("threads", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32),
        ("levels_per_sync", ctypes.c_uint32),
        ("bucket_bytes", ctypes.c_uint32),
        ("q_ints", ctypes.c_uint64),
        ("hash_seed", ctypes.c_uint64),
        ("cand_capacity", ctypes.c_uint64),
        ("use_graph", ctypes.c_uint32),
        ("insert_blocks", ctypes.c_uint32),
        ("time_kernels", ctypes.c_uint32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("t_explore_s", ctypes.c_double),
        ("t_init_s", ctypes.c_double),
        ("rounds", ctypes.c_uint64),
        ("graph_launches", ctypes.c_uint64),
        ("lookups", ctypes.c_uint64),
        ("states", ctypes.c_uint64),
        ("transitions", ctypes.c_uint64),
        ("cand_entries", ctypes.c_uint64),
        ("t_expand_s", ctypes.c_double),
        ("t_insert_s", ctypes.c_double),
        ("insert_launches", ctypes.c_uint64),
    ]


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("key_words", ctypes.c_uint32),
        ("max_procs", ctypes.c_uint32),
        ("threads", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32),
        ("rounds_per_step", ctypes.c_uint32),
        ("table_bytes", ctypes.c_uint64),
        ("page_bytes", ctypes.c_uint64),
        ("cand_capacity", ctypes.c_uint64),
        ("hash_seed", ctypes.c_uint64),
        ("model_words", ctypes.c_uint32),
        ("por_jobs", ctypes.c_uint32),
        ("insert_blocks", ctypes.c_uint32),
        ("large_slots", ctypes.c_uint32),
        ("large_states", ctypes.c_uint64),
        ("large_cus", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class JobOptions(ctypes.Structure):
    """gx_job_options: property, POR, cycle proviso and POR method of one job."""
    _fields_ = [
        ("property", ctypes.c_int32),
        ("por", ctypes.c_int32),
        ("cycle_proviso", ctypes.c_int32),
        ("por_method", ctypes.c_uint32),
    ]


def job_options(prop: int = 0, por: bool = False, cycle_proviso: bool = False,
                por_method: int = 0) -> JobOptions:
    return JobOptions(prop, int(bool(por)), int(bool(cycle_proviso)), por_method)


class ServerStatus(ctypes.Structure):
    _fields_ = [
        ("running", ctypes.c_uint32),
        ("waiting", ctypes.c_uint32),
        ("finished_pending", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("steps", ctypes.c_uint64),
        ("rounds", ctypes.c_uint64),
        ("pages_used", ctypes.c_uint64),
        ("pages_total", ctypes.c_uint64),
    ]


class JobResult(ctypes.Structure):
    _fields_ = [
        ("job_id", ctypes.c_uint64),
        ("result", Result),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("status", ctypes.c_int32),
        ("restarts", ctypes.c_int32),
    ]


class ModelInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "nr_procs", "bits_act", "proc_nrstates", "bits_statevector", "sv_nints", "liveness",
        "max_buf_ints", "nr_trans", "state_bits", "key_words")]


# exported symbols and their (restype, argtypes)
_P = ctypes.c_void_p
SIGNATURES = {
    "gx_model_load_gpf": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_P)]),
    "gx_model_load_gpf_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_P)]),
    "gx_model_free": (None, [_P]),
    "gx_model_info_get": (ctypes.c_int, [_P, ctypes.POINTER(ModelInfo)]),
    "gx_options_default": (None, [ctypes.POINTER(Options)]),
    "gx_explore": (ctypes.c_int, [_P, ctypes.POINTER(Options), ctypes.POINTER(Result)]),
    "gx_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "gx_explorer_create": (ctypes.c_int, [_P, ctypes.POINTER(Options), ctypes.POINTER(_P)]),
    "gx_explorer_run": (ctypes.c_int, [_P, ctypes.POINTER(Result)]),
    "gx_explorer_dump": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gx_explorer_free": (None, [_P]),
    "gx_batch_options_default": (None, [ctypes.POINTER(BatchOptions)]),
    "gx_server_options_default": (None, [ctypes.POINTER(ServerOptions)]),
    "gx_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOptions), ctypes.POINTER(_P)]),
    "gx_server_submit": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "gx_server_submit_job": (ctypes.c_int, [_P, _P, ctypes.POINTER(JobOptions), ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "gx_server_step": (ctypes.c_int, [_P, ctypes.POINTER(ServerStatus)]),
    "gx_server_poll": (ctypes.c_int, [_P, ctypes.POINTER(JobResult), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gx_server_free": (None, [_P]),
    "gx_explore_batch": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_uint64), ctypes.c_int,
                                        ctypes.POINTER(BatchOptions), ctypes.POINTER(Result),
                                        ctypes.POINTER(BatchStats)]),
    "gx_explore_batch_jobs": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(JobOptions),
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.c_int,
                                             ctypes.POINTER(BatchOptions), ctypes.POINTER(Result),
                                             ctypes.POINTER(BatchStats)]),
    "gx_table_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                       ctypes.POINTER(_P)]),
    "gx_table_free": (None, [_P]),
    "gx_table_insert_host": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint64)]),
    "gx_table_insert_device": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_float)]),
    "gx_table_insert_host_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                                    ctypes.POINTER(ctypes.c_uint8),
                                                    ctypes.POINTER(ctypes.c_uint64)]),
    "gx_table_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "gx_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "gx_device_alloc": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    "gx_device_free": (ctypes.c_int, [_P]),
    "gx_copy_to_device": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "gx_copy_to_host": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "gx_bench_copy": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "gx_bench_read": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "gx_bench_random": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]),
}

_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Loads the in-tree libgpuexplore.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run gpuexplore_amd.build.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != GX_OK:
        raise GxError(rc, what)


def device_count() -> int:
    n = ctypes.c_int(0)
    lib().gx_device_count(ctypes.byref(n))
    return n.value


class Model:
    """A loaded gpf model (gx_model handle)."""

    def __init__(self, handle: int):
        self.h = ctypes.c_void_p(handle)

    @classmethod
    def load(cls, path_noext: str, por: bool = False) -> "Model":
        h = ctypes.c_void_p()
        _check(lib().gx_model_load_gpf(path_noext.encode(), int(por), ctypes.byref(h)), "load " + path_noext)
        return cls(h.value)

    @classmethod
    def from_text(cls, text: str, por: bool = False) -> "Model":
        h = ctypes.c_void_p()
        b = text.encode()
        _check(lib().gx_model_load_gpf_text(b, len(b), int(por), ctypes.byref(h)), "load text")
        return cls(h.value)

    def info(self) -> dict:
        i = ModelInfo()
        _check(lib().gx_model_info_get(self.h, ctypes.byref(i)), "info")
        return {k: getattr(i, k) for k, _ in i._fields_}

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.gx_model_free(self.h)
            self.h = None


def _set_fields(o: ctypes.Structure, kw: dict) -> None:
    """Sets option fields by name; an unknown name is an error, not a new
    Python attribute that the library would never see."""
    names = {f for f, _ in type(o)._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"{type(o).__name__} has no field {k!r}")
        setattr(o, k, v)


def options(**kw) -> Options:
    o = Options()
    lib().gx_options_default(ctypes.byref(o))
    _set_fields(o, kw)
    return o


def explore(model: Model, **kw) -> dict:
    """gx_explore with keyword options (field names of gx_options)."""
    o = options(**kw)
    r = Result()
    rc = lib().gx_explore(model.h, ctypes.byref(o), ctypes.byref(r))
    d = r.as_dict()
    d["status"] = rc
    if rc not in (GX_OK, -10):
        raise GxError(rc, "explore")
    return d


class Explorer:
    """gx_explorer: device buffers allocated once, reused by every run."""

    def __init__(self, model: Model, **kw):
        self.model = model  # keeps the model handle alive
        self.opts = options(**kw)
        h = ctypes.c_void_p()
        _check(lib().gx_explorer_create(model.h, ctypes.byref(self.opts), ctypes.byref(h)), "explorer")
        self.h = h

    def run(self) -> dict:
        r = Result()
        rc = lib().gx_explorer_run(self.h, ctypes.byref(r))
        d = r.as_dict()
        d["status"] = rc
        if rc not in (GX_OK, -10):
            raise GxError(rc, "explorer run")
        return d

    def dump(self, path: str) -> None:
        """gx_explorer_dump: the last run's visited set in the --dump format."""
        _check(lib().gx_explorer_dump(self.h, path.encode()), "dump")

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.gx_explorer_free(self.h)
            self.h = None


def explore_batch(models, properties=None, weights=None, jobs=None, **kw):
    """gx_explore_batch (or gx_explore_batch_jobs when ``jobs`` -- a list of
    JobOptions, see job_options() -- is given).  Returns (list of
    per-instance result dicts, stats dict, status)."""
    n = len(models)
    o = BatchOptions()
    lib().gx_batch_options_default(ctypes.byref(o))
    _set_fields(o, kw)
    hs = (ctypes.c_void_p * n)(*[m.h.value for m in models])
    w = (ctypes.c_uint64 * n)(*weights) if weights is not None else None
    wp = ctypes.cast(w, ctypes.POINTER(ctypes.c_uint64)) if w is not None else None
    res = (Result * n)()
    st = BatchStats()
    mp = ctypes.cast(hs, ctypes.POINTER(ctypes.c_void_p))
    if jobs is not None:
        js = (JobOptions * n)(*jobs)
        rc = lib().gx_explore_batch_jobs(mp, js, wp, n, ctypes.byref(o), res, ctypes.byref(st))
    else:
        props = (ctypes.c_int32 * n)(*(properties or [0] * n))
        rc = lib().gx_explore_batch(mp, props, wp, n, ctypes.byref(o), res, ctypes.byref(st))
    if rc not in (GX_OK, -10):
        raise GxError(rc, "explore_batch")
    return [r.as_dict() for r in res], {k: getattr(st, k) for k, _ in st._fields_}, rc


class Server:
    """gx_server: continuous batching of model-checking jobs."""

    def __init__(self, **kw):
        o = ServerOptions()
        lib().gx_server_options_default(ctypes.byref(o))
        _set_fields(o, kw)
        h = ctypes.c_void_p()
        _check(lib().gx_server_create(ctypes.byref(o), ctypes.byref(h)), "server")
        self.h = h
        self._models = {}  # job id -> Model (kept alive until the job is polled)

    def submit(self, model: Model, prop: int = 0, est_states: int = 1, por: bool = False,
               cycle_proviso: bool = False, por_method: int = 0) -> int:
        """gx_server_submit_job with the job's property and POR options."""
        jid = ctypes.c_uint64(0)
        j = job_options(prop, por, cycle_proviso, por_method)
        _check(lib().gx_server_submit_job(self.h, model.h, ctypes.byref(j), est_states, ctypes.byref(jid)),
               "submit")
        self._models[jid.value] = model
        return jid.value

    def step(self) -> dict:
        st = ServerStatus()
        rc = lib().gx_server_step(self.h, ctypes.byref(st))
        _check(rc, "step")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def poll(self, max_results: int = 4096) -> list:
        buf = (JobResult * max_results)()
        n = ctypes.c_int(0)
        _check(lib().gx_server_poll(self.h, buf, max_results, ctypes.byref(n)), "poll")
        out = []
        for i in range(n.value):
            j = buf[i]
            d = j.result.as_dict()
            d.update(job_id=j.job_id, t_submit=j.t_submit, t_admit=j.t_admit, t_done=j.t_done,
                     status=j.status, restarts=j.restarts)
            self._models.pop(j.job_id, None)
            out.append(d)
        return out

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.gx_server_free(self.h)
            self.h = None


class Table:
    """Visited set used on its own (FIND-OR-PUT of 63-bit-chunk keys)."""

    def __init__(self, q_ints: int, key_words: int, bucket_bytes: int = 128, seed: int = 1):
        self.key_words = key_words
        h = ctypes.c_void_p()
        _check(lib().gx_table_create(q_ints, key_words, bucket_bytes, seed, ctypes.byref(h)), "table")
        self.h = h

    def insert(self, keys: Sequence[int]):
        """keys: flat sequence of n*key_words ints (< 2^63).  Returns
        (is_new bytes, inserted count, status)."""
        import numpy as np
        a = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
        n = a.size // self.key_words
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        ins = ctypes.c_uint64(0)
        rc = lib().gx_table_insert_host(
            self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
            flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(ins))
        return flags[:n], ins.value, rc

    def insert_noalloc(self, keys: Sequence[int]):
        """gx_table_insert_host_noalloc (at most NOALLOC_MAX_KEYS keys, staged
        in the library's fixed scratch).  Returns (is_new, inserted, status)."""
        import numpy as np
        a = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
        n = a.size // self.key_words
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        ins = ctypes.c_uint64(0)
        rc = lib().gx_table_insert_host_noalloc(
            self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
            flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(ins))
        return flags[:n], ins.value, rc

    def insert_device(self, keys: Sequence[int]):
        """gx_table_insert_device on a device copy of the keys (the copy is
        made and freed here).  Returns (inserted, status, kernel ms)."""
        import numpy as np
        a = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
        n = a.size // self.key_words
        d = ctypes.c_void_p()
        _check(lib().gx_device_alloc(max(8, a.nbytes), ctypes.byref(d)), "device alloc")
        try:
            _check(lib().gx_copy_to_device(d, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "copy")
            ins = ctypes.c_uint64(0)
            ms = ctypes.c_float(0)
            rc = lib().gx_table_insert_device(self.h, d, n, ctypes.byref(ins), ctypes.byref(ms))
        finally:
            lib().gx_device_free(d)
        return ins.value, rc, ms.value

    def count(self) -> int:
        c = ctypes.c_uint64(0)
        _check(lib().gx_table_count(self.h, ctypes.byref(c)), "count")
        return c.value

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.gx_table_free(self.h)
            self.h = None
