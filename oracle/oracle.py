"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
imports this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Optional, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

PROP_NONE, PROP_DEADLOCK, PROP_SAFETY, PROP_LIVENESS = 0, 1, 2, 3
POR_CAMPLE, POR_AMPLE, POR_STUBBORN = 0, 1, 2


class _Opts(ctypes.Structure):
    _fields_ = [
        ("property", ctypes.c_int),
        ("por", ctypes.c_int),
        ("cycle_proviso", ctypes.c_int),
        ("max_states", ctypes.c_uint64),
        ("hash_seed", ctypes.c_uint64),
        ("por_method", ctypes.c_int),
        ("expected_states", ctypes.c_uint64),
        ("progress", ctypes.c_int),
    ]


class _Res(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_uint64),
        ("transitions", ctypes.c_uint64),
        ("levels", ctypes.c_uint64),
        ("full_successors", ctypes.c_uint64),
        ("violation", ctypes.c_int),
        ("truncated", ctypes.c_int),
        ("seconds", ctypes.c_double),
    ]


class _Model(ctypes.Structure):
    _fields_ = [
        ("liveness", ctypes.c_uint32),
        ("nr_procs", ctypes.c_uint32),
        ("bits_act", ctypes.c_uint32),
        ("proc_nrstates", ctypes.c_uint32),
        ("bits_statevector", ctypes.c_uint32),
        ("sv_nints", ctypes.c_uint32),
        ("firstbit", ctypes.POINTER(ctypes.c_uint32)),
        ("bits_state", ctypes.POINTER(ctypes.c_uint32)),
        ("nbits_offset", ctypes.c_uint32),
        ("max_buf_ints", ctypes.c_uint32),
        ("proc_offsets_start", ctypes.POINTER(ctypes.c_uint32)),
        ("proc_offsets", ctypes.POINTER(ctypes.c_uint32)),
        ("nr_trans", ctypes.c_uint32),
        ("proc_trans", ctypes.POINTER(ctypes.c_uint32)),
        ("nbits_syncbits_offset", ctypes.c_uint32),
        ("nr_syncbits_offsets", ctypes.c_uint32),
        ("syncbits_offsets", ctypes.POINTER(ctypes.c_uint32)),
        ("nr_syncbits", ctypes.c_uint32),
        ("syncbits", ctypes.POINTER(ctypes.c_uint32)),
        ("words", ctypes.c_uint32),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_load_gpf.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.POINTER(_Model))]
        L.orc_load_gpf.restype = ctypes.c_int
        L.orc_free.argtypes = [ctypes.POINTER(_Model)]
        L.orc_explore.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(_Opts), ctypes.POINTER(_Res)]
        L.orc_explore.restype = ctypes.c_int
        L.orc_explore_mt.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(_Opts), ctypes.c_int,
                                     ctypes.POINTER(_Res)]
        L.orc_explore_mt.restype = ctypes.c_int
        L.orc_successors.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_int)]
        L.orc_successors.restype = ctypes.c_long
        L.orc_get.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]
        L.orc_get.restype = ctypes.c_uint32
        L.orc_set.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                              ctypes.c_uint32]
        L.orc_por_cluster.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(ctypes.c_uint32)]
        L.orc_por_cluster.restype = ctypes.c_uint64
        L.orc_por_select.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
        L.orc_por_select.restype = ctypes.c_uint64
        _lib = L
    return _lib


class OracleModel:
    def __init__(self, path_noext: str, por: bool = False):
        self.ptr = ctypes.POINTER(_Model)()
        rc = lib().orc_load_gpf(path_noext.encode(), int(por), ctypes.byref(self.ptr))
        if rc != 0:
            raise OSError(f"oracle failed to load {path_noext}.gpf: {rc}")
        self.m = self.ptr.contents

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_free(self.ptr)
            self.ptr = None

    @property
    def words(self) -> int:
        return self.m.words

    @property
    def nprocs(self) -> int:
        return self.m.nr_procs

    def arrays(self) -> dict:
        m = self.m
        n = m.nr_procs

        def arr(p, k):
            return [p[i] for i in range(k)]

        return dict(
            liveness=m.liveness, nr_procs=n, bits_act=m.bits_act, proc_nrstates=m.proc_nrstates,
            bits_statevector=m.bits_statevector, sv_nints=m.sv_nints,
            firstbit=arr(m.firstbit, n + 1), bits_state=arr(m.bits_state, n),
            nbits_offset=m.nbits_offset, max_buf_ints=m.max_buf_ints,
            proc_offsets_start=arr(m.proc_offsets_start, n + 1),
            proc_offsets=arr(m.proc_offsets, m.proc_offsets_start[n]),
            nr_trans=m.nr_trans, proc_trans=arr(m.proc_trans, m.nr_trans),
            nbits_syncbits_offset=m.nbits_syncbits_offset,
            syncbits_offsets=arr(m.syncbits_offsets, m.nr_syncbits_offsets),
            syncbits=arr(m.syncbits, m.nr_syncbits))

    def _vec(self, words: List[int]):
        a = (ctypes.c_uint32 * 9)()
        for i, w in enumerate(words):
            a[i] = w
        return a

    def get(self, words: List[int], p: int) -> int:
        return lib().orc_get(self.ptr, self._vec(words), p)

    def set(self, words: List[int], p: int, x: int) -> List[int]:
        a = self._vec(words)
        lib().orc_set(self.ptr, a, p, x)
        return [a[i] for i in range(len(words))]

    def successors(self, words: List[int]) -> Tuple[List[Tuple[int, ...]], bool]:
        W = self.words
        cap = 1 << 16
        out = (ctypes.c_uint32 * (cap * W))()
        dl = ctypes.c_int(0)
        n = lib().orc_successors(self.ptr, self._vec(words), out, cap, ctypes.byref(dl))
        if n < 0:
            raise OSError(f"orc_successors failed: {n}")
        if n > cap:
            raise OSError("too many successors")
        return [tuple(out[i * W:(i + 1) * W]) for i in range(n)], bool(dl.value)

    def por_cluster(self, words: List[int]) -> int:
        return lib().orc_por_cluster(self.ptr, self._vec(words))

    def por_select(self, words: List[int], method: int) -> int:
        return lib().orc_por_select(self.ptr, self._vec(words), method)

    def explore(self, prop: int = PROP_NONE, por: bool = False, proviso: bool = False,
                max_states: int = 0, seed: int = 0, method: int = POR_CAMPLE, expected_states: int = 0) -> dict:
        o = _Opts(prop, int(por), int(proviso), max_states, seed, method, expected_states)
        r = _Res()
        rc = lib().orc_explore(self.ptr, ctypes.byref(o), ctypes.byref(r))
        if rc != 0:
            raise OSError(f"orc_explore failed: {rc}")
        return dict(states=r.states, transitions=r.transitions, levels=r.levels,
                    full=r.full_successors, violation=bool(r.violation),
                    truncated=bool(r.truncated), seconds=r.seconds)

    def explore_mt(self, threads: int, prop: int = PROP_NONE, max_states: int = 0, seed: int = 0,
                   expected_states: int = 0, progress: bool = False) -> dict:
        """orc_explore_mt: the plain method on `threads` threads (the visited
        set presized for `expected_states` when given; with `progress`, a
        line on stderr after each level)."""
        o = _Opts(prop, 0, 0, max_states, seed, POR_CAMPLE, expected_states, int(progress))
        r = _Res()
        rc = lib().orc_explore_mt(self.ptr, ctypes.byref(o), threads, ctypes.byref(r))
        if rc != 0:
            raise OSError(f"orc_explore_mt failed: {rc}")
        return dict(states=r.states, transitions=r.transitions, levels=r.levels,
                    full=r.full_successors, violation=bool(r.violation),
                    truncated=bool(r.truncated), seconds=r.seconds)


def explore(path_noext: str, prop: int = PROP_NONE, por: bool = False, proviso: bool = False,
            max_states: int = 0, seed: int = 0, method: int = 0, expected_states: int = 0) -> dict:
    return OracleModel(path_noext, por).explore(prop, por, proviso, max_states, seed, method, expected_states)


def explore_mt(path_noext: str, threads: int, prop: int = PROP_NONE, max_states: int = 0, seed: int = 0,
               expected_states: int = 0) -> dict:
    return OracleModel(path_noext, False).explore_mt(threads, prop, max_states, seed, expected_states)
