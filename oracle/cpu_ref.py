"""Build recipe and ctypes binding of the C++ / OpenMP CPU restatement
(oracle/cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/ (checked against the numpy oracle and
the reference's golden files) and by bench.py's cpu_baseline leg (the timed
CPU baseline).  The product path never loads it.
"""
from __future__ import annotations

import ctypes as _c
import os
import subprocess

import numpy as np

from . import slaph_oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_ref.cpp")
LIB = os.path.join(HERE, "_build", "libslaph_cpuref.so")
# x86-64-v3 (AVX2 + FMA): portable to the hosts that run the library, not
# only to the one that built it.  -fcx-limited-range: complex products by the
# plain formula, as Eigen forms them (no inf / NaN recovery calls)
CMD = ["g++", "-O3", "-march=x86-64-v3", "-fcx-limited-range", "-fopenmp", "-std=c++17", "-shared", "-fPIC", SRC,
       "-o", LIB]


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    r = subprocess.run(CMD, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cpu_ref build failed: %s\n%s" % (" ".join(CMD), r.stderr))
    return LIB


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = _c.CDLL(LIB)
        P, D, I = _c.c_void_p, _c.POINTER(_c.c_double), _c.POINTER(_c.c_int)
        lib.cref_new.restype = P
        lib.cref_new.argtypes = [_c.c_int] * 5
        lib.cref_free.argtypes = [P]
        lib.cref_set_gamma.argtypes = [P, _c.c_int, I, D]
        lib.cref_set_peram.argtypes = [P, _c.c_int, D, _c.c_longlong, _c.c_longlong]
        lib.cref_set_rnd.argtypes = [P, _c.c_int, D]
        lib.cref_set_nvdv.argtypes = [P, _c.c_int]
        lib.cref_set_vdv.argtypes = [P, _c.c_int, _c.c_int, D]
        lib.cref_add_ql.argtypes = [P, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, I]
        lib.cref_add_trace_id.restype = _c.c_int
        lib.cref_add_trace_id.argtypes = [P, _c.c_char_p, _c.c_int, I]
        lib.cref_add_diagram.restype = _c.c_int
        lib.cref_add_diagram.argtypes = [P, _c.c_char_p]
        lib.cref_add_request.restype = _c.c_int
        lib.cref_add_request.argtypes = [P, _c.c_int, _c.c_int, _c.POINTER(_c.c_char_p), I, I, I]
        lib.cref_contract.restype = _c.c_double
        lib.cref_contract.argtypes = [P, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, D, _c.c_char_p, _c.c_int]
        lib.cref_last_stats.argtypes = [P, D]
        lib.cref_vdaggerv.restype = _c.c_double
        lib.cref_vdaggerv.argtypes = [D, D, _c.c_longlong, _c.c_int, D]
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_c.POINTER(_c.c_double))


def _ip(a):
    return a.ctypes.data_as(_c.POINTER(_c.c_int))


class CpuRef:
    """One configuration's inputs in the C++ restatement.  vdv: {(id, t): N x N}
    (a callable (id, t) -> matrix also works); perams / rnds as the oracle's."""

    def __init__(self, cfg, lut, vdv, perams, rnds):
        lib = _load()
        q0 = cfg["quarks"][0]
        self.Lt, self.N = cfg["Lt"], cfg["number_of_eigen_vec"]
        self.lut = lut
        self.h = lib.cref_new(self.Lt, self.N, q0["dilT"], q0["dilE"], len(perams))
        for g, (rows, vals) in enumerate(o.GAMMA):
            r = np.array(rows, dtype=np.int32)
            v = np.array([x for z in vals for x in (z.real, z.imag)], dtype=np.float64)
            lib.cref_set_gamma(self.h, g, _ip(r), _dp(v))
        for r, P in enumerate(perams):
            P = np.ascontiguousarray(P, dtype=np.complex128)
            lib.cref_set_peram(self.h, r, _dp(P), P.shape[0], P.shape[1])
        for r, rho in enumerate(rnds):
            rho = np.ascontiguousarray(rho, dtype=np.complex128)
            lib.cref_set_rnd(self.h, r, _dp(rho))
        nvdv = len(lut["vdv_lookup"])
        lib.cref_set_nvdv(self.h, nvdv)
        get = vdv if callable(vdv) else (lambda i, t: vdv[(i, t)])
        for vid in range(nvdv):
            for t in range(self.Lt):
                M = np.ascontiguousarray(get(vid, t), dtype=np.complex128)
                lib.cref_set_vdv(self.h, vid, t, _dp(M))
        for q, name in enumerate(("Q0", "Q1", "Q2")):
            for vid, dag, gam, pairs in lut["ql_lookup"][name]:
                pr = np.array([x for p in pairs for x in p], dtype=np.int32)
                lib.cref_add_ql(self.h, q, vid, int(dag), gam[0], len(pairs), _ip(pr))
        for name, lst in lut["trace_indices"].items():
            for ids in lst:
                a = np.array(ids, dtype=np.int32)
                if lib.cref_add_trace_id(self.h, name.encode(), len(ids), _ip(a)) < 0:
                    raise RuntimeError("unknown trace " + name)
        self.diagrams = []
        for dname, reqs in lut["requests"].items():
            if not reqs:
                continue
            d = lib.cref_add_diagram(self.h, dname.encode())
            self.diagrams.append((dname, [ds for _, ds in reqs]))
            for trs, _ in reqs:
                names = (_c.c_char_p * len(trs))(*[n.encode() for n, _, _ in trs])
                ids = np.array([i for _, i, _ in trs], dtype=np.int32)
                nloc = np.array([len(loc) for _, _, loc in trs], dtype=np.int32)
                sinks = np.array([1 if x == "sink" else 0 for _, _, loc in trs for x in loc] or [0], dtype=np.int32)
                if lib.cref_add_request(self.h, d, len(trs), names, _ip(ids), _ip(nloc), _ip(sinks)) != 0:
                    raise RuntimeError("bad request in " + dname)

    def __del__(self):
        if getattr(self, "h", None):
            _load().cref_free(self.h)
            self.h = None

    def contract(self, threads=1, kahan=False, first_block_pair=0, block_pairs=-1, slice_pairs=-1):
        """Correlators {diagram: [(dataset, complex[Lt])]} and wall seconds."""
        total = sum(len(ds) for _, ds in self.diagrams) * self.Lt
        out = np.zeros(2 * max(1, total), dtype=np.float64)
        err = _c.create_string_buffer(512)
        secs = _load().cref_contract(self.h, int(threads), int(kahan), int(first_block_pair), int(block_pairs),
                                     int(slice_pairs), _dp(out), err, 512)
        if secs < 0:
            raise RuntimeError(err.value.decode())
        vals = out[:2 * total].view(np.complex128)
        res, at = {}, 0
        for dname, dsets in self.diagrams:
            lst = []
            for ds in dsets:
                lst.append((ds, vals[at:at + self.Lt].copy()))
                at += self.Lt
            res[dname] = lst
        return res, secs

    def last_stats(self):
        """Of the last contract: {"Q0": (seconds, keys), "Q1": ..., "Q2": ...,
        "wall": seconds, "busy": thread-seconds in the block-pair loop};
        factor seconds are summed over threads."""
        out = np.zeros(8)
        _load().cref_last_stats(self.h, _dp(out))
        return {"Q0": (out[0], int(out[1])), "Q1": (out[2], int(out[3])), "Q2": (out[4], int(out[5])),
                "wall": out[6], "busy": out[7]}


def factor_keys(lut, dilT, slice_pairs):
    """The distinct factor keys ("Q0", (t,)), ("Q1", (t, b)), ("Q2", (b1, t, b2))
    that the trace recipes (DilutedTraceFactory.cpp:13-238, as in
    slaph_oracle.Traces) read for the requests of `lut` over `slice_pairs`."""
    B = lambda t: t // dilT  # noqa: E731
    keys = set()
    traces = set()
    for reqs in lut["requests"].values():
        for trs, _ in reqs:
            for name, _, locs in trs:
                traces.add((name, locs))
    for t_so, t_si in slice_pairs:
        for name, locs in traces:
            k = [t_so if l == "source" else t_si for l in locs]
            if name == "trQ1":
                keys.add(("Q1", (k[0], B(k[0]))))
            elif name == "trQ1Q1":
                keys |= {("Q1", (k[0], B(k[1]))), ("Q1", (k[1], B(k[0])))}
            elif name == "trQ0Q2":
                keys |= {("Q0", (k[1],)), ("Q2", (B(k[1]), k[0], B(k[1])))}
            elif name == "trQ1Q1Q1":
                keys |= {("Q1", (k[0], B(k[1]))), ("Q1", (k[1], B(k[2]))), ("Q1", (k[2], B(k[0])))}
            elif name == "trQ1Q0Q2":
                keys |= {("Q0", (k[0],)), ("Q2", (B(k[0]), k[0], B(k[1]))), ("Q1", (k[1], B(k[2])))}
            elif name == "trQ1Q1Q1Q1":
                keys |= {("Q1", (k[i], B(k[(i + 1) % 4]))) for i in range(4)}
            elif name == "trQ2Q0Q2Q0":
                keys |= {("Q0", (k[0],)), ("Q2", (B(k[0]), k[1], B(k[2]))), ("Q0", (k[2],)),
                         ("Q2", (B(k[2]), k[3], B(k[0])))}
            elif name == "trQ2Q0Q2Q0Q2Q0":
                b1, b3, b5 = B(k[1]), B(k[3]), B(k[5])
                keys |= {("Q2", (b5, k[0], b1)), ("Q0", (k[1],)), ("Q2", (b1, k[2], b3)), ("Q0", (k[3],)),
                         ("Q2", (b3, k[4], b5)), ("Q0", (k[5],))}
            else:
                raise KeyError(name)
    return keys


def vdaggerv(V, phase):
    """V^+ diag(phase[k / 3]) V by the C++ loop nest (one thread): (matrix, seconds)."""
    V = np.asfortranarray(V, dtype=np.complex128)
    ph = np.ascontiguousarray(phase, dtype=np.complex128)
    K, N = V.shape
    out = np.zeros((N, N), dtype=np.complex128)
    secs = _load().cref_vdaggerv(_dp(V.reshape(-1, order="F")), _dp(ph), K, N, _dp(out))
    return out, secs
