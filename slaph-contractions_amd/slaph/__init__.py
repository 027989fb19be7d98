"""Python binding of the MI355X contraction path of sLapH-contractions.

Thin ctypes layer over the C ABI in include/slaph_amd.h, shaped like the
reference's host flow (main/contract.cpp:28-107):

    ctx = Context(device=0)
    ctx.configure(open("test_all.ini").read())   # read_parameters + init_lookup_tables
    ctx.read_inputs(base_dir, config=1000)       # Perambulator / RandomVector / EigenVector
    ops = OperatorFactory(ctx); ops.create_operators()
    contract(ctx)                                # Correlators.cpp:40-127
    ctx.correlators("C2c")                       # {dataset name: complex[Lt]}

All arithmetic runs in the HIP library `libslaph_amd.so` built next to this
package; importing this module fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libslaph_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "slaph_amd.h")

if not os.path.exists(LIB_PATH):
    raise ImportError("libslaph_amd.so is not built: run `python slaph-contractions_amd/build.py` "
                      "(or __graft_entry__.build()) first")

_lib = ctypes.CDLL(LIB_PATH)

_c = ctypes
_P = _c.c_void_p
_D = _c.POINTER(_c.c_double)
_I = _c.POINTER(_c.c_int)

_SIGS = {
    "slc_ctx_create": (_c.c_int, [_c.c_int, _c.POINTER(_P)]),
    "slc_ctx_destroy": (None, [_P]),
    "slc_last_error": (_c.c_char_p, [_P]),
    "slc_configure": (_c.c_int, [_P, _c.c_char_p]),
    "slc_query": (_c.c_int, [_P, _c.c_char_p, _c.POINTER(_c.c_longlong)]),
    "slc_set_eigenvectors": (_c.c_int, [_P, _c.c_int, _D]),
    "slc_set_perambulator": (_c.c_int, [_P, _c.c_int, _D]),
    "slc_set_random_vector": (_c.c_int, [_P, _c.c_int, _D]),
    "slc_read_inputs": (_c.c_int, [_P, _c.c_char_p, _c.c_int]),
    "slc_synthetic_inputs": (_c.c_int, [_P, _c.c_ulonglong, _c.c_int]),
    "slc_get_input": (_c.c_int, [_P, _c.c_int, _c.c_int, _D]),
    "slc_create_operators": (_c.c_int, [_P]),
    "slc_create_operators_config": (_c.c_int, [_P, _c.c_int]),
    "slc_set_timeslice_streaming": (_c.c_int, [_P, _c.c_int]),
    "slc_set_graphs": (_c.c_int, [_P, _c.c_int]),
    "slc_autotune": (_c.c_int, [_P, _c.c_int]),
    "slc_autotune_report": (_c.c_int, [_P, _c.c_char_p, _c.c_int]),
    "slc_get_tuning": (_c.c_int, [_P, _c.c_char_p, _c.c_int]),
    "slc_set_tuning": (_c.c_int, [_P, _c.c_char_p]),
    "slc_set_gauge_links": (_c.c_int, [_P, _c.c_int, _D]),
    "slc_get_gauge_links": (_c.c_int, [_P, _c.c_int, _D]),
    "slc_synthetic_gauge": (_c.c_int, [_P, _c.c_ulonglong, _c.c_double]),
    "slc_smear_stout": (_c.c_int, [_P, _c.c_double, _c.c_int]),
    "slc_smear_hyp": (_c.c_int, [_P, _c.c_double, _c.c_double, _c.c_int]),
    "slc_read_gauge_field": (_c.c_int, [_P, _c.c_char_p, _c.c_char_p, _c.c_int, _c.c_int, _c.c_int]),
    "slc_plaquette": (_c.c_int, [_P, _c.c_int, _D]),
    "slc_read_lime_gauge": (_c.c_int, [_c.c_char_p] + [_c.c_int] * 6 + [_D]),
    "slc_vdaggerv_filename": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_int, _I, _c.c_char_p, _c.c_int, _c.c_char_p,
                                         _c.c_int]),
    "slc_query_string": (_c.c_int, [_P, _c.c_char_p, _c.c_char_p, _c.c_int]),
    "slc_return_vdaggerv": (_c.c_int, [_P, _c.c_int, _c.c_int, _D]),
    "slc_diluted_factor": (_c.c_int, [_P, _c.c_int, _I, _c.c_int, _D, _I, _c.c_int, _I]),
    "slc_contract": (_c.c_int, [_P, _c.c_int, _c.c_int]),
    "slc_inner_product": (_c.c_int, [_P, _c.c_int, _c.POINTER(_D), _c.POINTER(_c.POINTER(_c.c_uint16)), _I,
                                     _c.c_int, _D]),
    "slc_diluted_trace": (_c.c_int, [_P, _c.c_char_p, _c.c_int, _c.c_int, _I, _c.c_int, _I, _c.c_int, _D,
                                     _c.POINTER(_c.c_uint16), _c.c_int, _I, _I]),
    "slc_diagram_count": (_c.c_int, [_P, _I]),
    "slc_diagram_name": (_c.c_int, [_P, _c.c_int, _c.c_char_p, _c.c_int]),
    "slc_correlator_count": (_c.c_int, [_P, _c.c_char_p, _I]),
    "slc_correlator_name": (_c.c_int, [_P, _c.c_char_p, _c.c_int, _c.c_char_p, _c.c_int]),
    "slc_fetch_correlators": (_c.c_int, [_P, _c.c_char_p, _D]),
    "slc_stage_time": (_c.c_int, [_P, _c.c_char_p, _D]),
    "slc_kernel_stat": (_c.c_int, [_P, _c.c_char_p, _c.c_char_p, _D]),
    "slc_set_chunk_budget": (_c.c_int, [_P, _c.c_longlong]),
    "slc_set_precision": (_c.c_int, [_P, _c.c_int, _c.c_int]),
    "slc_ahb_device": (_c.c_int, [_P, _P, _P, _P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_longlong,
                                  _c.c_longlong, _c.c_longlong, _c.c_int, _c.c_longlong, _c.c_longlong,
                                  _c.c_longlong]),
    "slc_ahb_device_noalloc": (_c.c_int, [_P, _P, _P, _P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_longlong,
                                          _c.c_longlong, _c.c_longlong, _c.c_int, _c.c_longlong, _c.c_longlong,
                                          _c.c_longlong]),
    "slc_ahb_host": (_c.c_int, [_P, _D, _D, _D, _D, _c.c_int, _c.c_int, _c.c_int]),
    "slc_lookup_dump": (_c.c_int, [_c.c_char_p, _c.c_char_p, _c.c_int]),
    "slc_dilution_text": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_char_p, _c.c_int]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class SlaphError(RuntimeError):
    """Error reported through the C ABI (message of the reference's runtime_error)."""


def exported_symbols():
    return sorted(_SIGS)


def _dptr(a):
    return a.ctypes.data_as(_D)


def lookup_dump(infile_text: str) -> str:
    """init_lookup_tables as text; host only (no device needed)."""
    need = _lib.slc_lookup_dump(infile_text.encode(), None, 0)
    buf = _c.create_string_buffer(abs(need) + 1 if need > 0 else 1 << 20)
    rc = _lib.slc_lookup_dump(infile_text.encode(), buf, len(buf))
    if rc < 0:
        raise SlaphError(buf.value.decode())
    return buf.value.decode()


def vdaggerv_filename(kind: str, path: str, config: int, mom, displacement: str = "", t: int = 0) -> str:
    """The reference's VdaggerV file names; kind "write", "read" or "liuming"."""
    k = {"write": 0, "read": 1, "liuming": 2}[kind]
    m = (_c.c_int * 3)(*[int(x) for x in mom])
    args = (k, path.encode(), int(config), m, displacement.encode(), int(t))
    need = _lib.slc_vdaggerv_filename(*args, None, 0)
    if need < 0:
        raise SlaphError("invalid VdaggerV file name arguments")
    buf = _c.create_string_buffer(need)
    _lib.slc_vdaggerv_filename(*args, buf, need)
    return buf.value.decode()


def read_lime_gauge(path: str, L, Lt: int, slice_i: int = 0, slice_f: int = None) -> np.ndarray:
    """Links of timeslices slice_i..slice_f of a lime / ILDG file on the host,
    [t - slice_i, site, mu, row, col] (site = (x Ly + y) Lz + z, mu = x, y, z),
    by the library's reader (no context or device needed)."""
    Lx, Ly, Lz = (int(v) for v in L)
    if slice_f is None:
        slice_f = Lt - 1
    nt = max(0, int(slice_f) - int(slice_i) + 1)
    out = np.zeros((max(1, nt), Lx * Ly * Lz, 3, 3, 3), dtype=np.complex128)
    rc = _lib.slc_read_lime_gauge(path.encode(), Lx, Ly, Lz, int(Lt), int(slice_i), int(slice_f), _dptr(out.reshape(-1)))
    if rc != 0:
        raise SlaphError(_lib.slc_last_error(None).decode())
    return out.transpose(0, 1, 2, 4, 3).copy()


def dilution_text(num_slice: int, num_block: int, interlace: bool) -> str:
    need = _lib.slc_dilution_text(num_slice, num_block, int(interlace), None, 0)
    if need < 0:
        raise SlaphError("invalid dilution scheme")
    buf = _c.create_string_buffer(need)
    _lib.slc_dilution_text(num_slice, num_block, int(interlace), buf, need)
    return buf.value.decode()


class Context:
    """One device context (C ABI slc_ctx)."""

    def __init__(self, device: int = 0):
        h = _P()
        rc = _lib.slc_ctx_create(device, _c.byref(h))
        if rc != 0:
            raise SlaphError("cannot create a context on HIP device %d (code %d): %s"
                             % (device, rc, _lib.slc_last_error(None).decode()))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        self._h = None
        if h and _lib is not None:
            try:
                _lib.slc_ctx_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise SlaphError(_lib.slc_last_error(self._h).decode())
        return rc

    # GlobalData ---------------------------------------------------------
    def configure(self, infile_text: str):
        self._check(_lib.slc_configure(self._h, infile_text.encode()))

    def query(self, what: str) -> int:
        v = _c.c_longlong()
        self._check(_lib.slc_query(self._h, what.encode(), _c.byref(v)))
        return v.value

    def query_string(self, what: str) -> str:
        need = self._check(_lib.slc_query_string(self._h, what.encode(), None, 0))
        buf = _c.create_string_buffer(need)
        self._check(_lib.slc_query_string(self._h, what.encode(), buf, need))
        return buf.value.decode()

    # inputs ---------------------------------------------------------------
    def _expect_shape(self, what, a, shape):
        # the C side copies a fixed element count; a wrong shape would be
        # misread or read past its end
        if tuple(a.shape) != tuple(shape):
            raise SlaphError("%s must have shape %s, got %s" % (what, tuple(shape), tuple(a.shape)))

    def set_eigenvectors(self, t: int, V: np.ndarray):
        V = np.asfortranarray(V, dtype=np.complex128)
        K = 3 * self.query("Lx") * self.query("Ly") * self.query("Lz")
        self._expect_shape("eigenvectors (3 Lx Ly Lz, nev)", V, (K, self.query("nev")))
        self._check(_lib.slc_set_eigenvectors(self._h, t, _dptr(V.reshape(-1, order="F"))))

    def set_perambulator(self, r: int, P: np.ndarray):
        P = np.asfortranarray(P, dtype=np.complex128)
        self._expect_shape("perambulator (rows, cols)", P, (self.query("peram_rows"), self.query("peram_cols")))
        self._check(_lib.slc_set_perambulator(self._h, r, _dptr(P.reshape(-1, order="F"))))

    def set_random_vector(self, r: int, rho: np.ndarray):
        rho = np.ascontiguousarray(rho, dtype=np.complex128)
        self._expect_shape("random vector (Lt nev 4,)", rho, (self.query("Lt") * self.query("nev") * 4,))
        self._check(_lib.slc_set_random_vector(self._h, r, _dptr(rho)))

    def read_inputs(self, base_dir: str, config: int):
        self._check(_lib.slc_read_inputs(self._h, base_dir.encode(), config))

    def synthetic_inputs(self, seed: int, orthonormalise: bool = True):
        self._check(_lib.slc_synthetic_inputs(self._h, seed, int(orthonormalise)))

    def set_timeslice_streaming(self, on: bool = True):
        """One timeslice of eigenvectors resident; V(t) produced per timeslice."""
        self._check(_lib.slc_set_timeslice_streaming(self._h, int(on)))

    def _gauge_shape(self):
        return (self.query("Lx") * self.query("Ly") * self.query("Lz"), 3, 3, 3)

    def set_gauge_links(self, t: int, U: np.ndarray):
        """U[site, mu] = 3 x 3 link matrix (numpy, row/column as usual)."""
        U = np.asarray(U, dtype=np.complex128)
        self._expect_shape("gauge links (sites, 3, 3, 3)", U, self._gauge_shape())
        flat = np.ascontiguousarray(U.transpose(0, 1, 3, 2))  # column-major 3 x 3 per link
        self._check(_lib.slc_set_gauge_links(self._h, int(t), _dptr(flat.reshape(-1))))

    def get_gauge_links(self, t: int) -> np.ndarray:
        out = np.zeros(self._gauge_shape(), dtype=np.complex128)
        self._check(_lib.slc_get_gauge_links(self._h, int(t), _dptr(out.reshape(-1))))
        return out.transpose(0, 1, 3, 2).copy()  # back to [site, mu, row, col]

    def synthetic_gauge(self, seed: int, eps: float):
        self._check(_lib.slc_synthetic_gauge(self._h, int(seed), float(eps)))

    def read_gauge_field(self, config: int, slice_i: int = 0, slice_f: int = None, config_path: str = None,
                         base_dir: str = "."):
        """GaugeField::read_gauge_field: timeslices slice_i..slice_f (default
        all) of <config_path>/conf.%04d (default: path_config of the infile)."""
        if slice_f is None:
            slice_f = self.query("Lt") - 1
        self._check(_lib.slc_read_gauge_field(self._h, base_dir.encode(),
                                              None if config_path is None else config_path.encode(),
                                              int(config), int(slice_i), int(slice_f)))

    def plaquette(self, t: int) -> float:
        """GaugeField::plaque_ts of the stored links of timeslice t."""
        out = _c.c_double()
        self._check(_lib.slc_plaquette(self._h, int(t), _c.byref(out)))
        return out.value

    def smear_stout(self, rho: float, iterations: int):
        self._check(_lib.slc_smear_stout(self._h, float(rho), int(iterations)))

    def smear_hyp(self, alpha1: float, alpha2: float, iterations: int):
        self._check(_lib.slc_smear_hyp(self._h, float(alpha1), float(alpha2), int(iterations)))

    def get_input(self, kind: str, index: int) -> np.ndarray:
        """Resident input back on the host: 'eigenvectors' (K x N), 'perambulator'
        (rows x cols), 'random_vector' (Lt*N*4)."""
        k = {"eigenvectors": 0, "perambulator": 1, "random_vector": 2}[kind]
        Lt, N = self.query("Lt"), self.query("nev")
        if k == 0:
            shape = (3 * self.query("Lx") * self.query("Ly") * self.query("Lz"), N)
        elif k == 1:
            shape = (self.query("peram_rows"), self.query("peram_cols"))
        else:
            shape = (Lt * N * 4,)
        out = np.zeros(int(np.prod(shape)), dtype=np.complex128)
        self._check(_lib.slc_get_input(self._h, k, index, _dptr(out)))
        return out.reshape(shape, order="F") if len(shape) == 2 else out

    # results ----------------------------------------------------------------
    def diagrams(self):
        n = _c.c_int()
        self._check(_lib.slc_diagram_count(self._h, _c.byref(n)))
        out = []
        for i in range(n.value):
            buf = _c.create_string_buffer(256)
            self._check(_lib.slc_diagram_name(self._h, i, buf, 256))
            out.append(buf.value.decode())
        return out

    def correlators(self, diagram: str):
        n = _c.c_int()
        d = diagram.encode()
        self._check(_lib.slc_correlator_count(self._h, d, _c.byref(n)))
        Lt = self.query("Lt")
        vals = np.zeros(n.value * Lt, dtype=np.complex128)
        self._check(_lib.slc_fetch_correlators(self._h, d, _dptr(vals)))
        vals = vals.reshape(n.value, Lt)
        out = {}
        for i in range(n.value):
            need = self._check(_lib.slc_correlator_name(self._h, d, i, None, 0))
            buf = _c.create_string_buffer(need)
            self._check(_lib.slc_correlator_name(self._h, d, i, buf, need))
            out[buf.value.decode()] = vals[i]
        return out

    def stage_time(self, what: str) -> float:
        v = _c.c_double()
        self._check(_lib.slc_stage_time(self._h, what.encode(), _c.byref(v)))
        return v.value

    def kernel_stat(self, kernel: str, what: str) -> float:
        """Launch statistics of one kernel in the last contract or
        create_operators: what = "ms" (summed device time), "launches",
        "bytes" (summed algorithmic bytes) or "flops" (summed operations of
        the kernel's own algorithm)."""
        v = _c.c_double()
        self._check(_lib.slc_kernel_stat(self._h, kernel.encode(), what.encode(), _c.byref(v)))
        return v.value

    def _string(self, fn):
        need = self._check(fn(self._h, None, 0))
        buf = _c.create_string_buffer(max(1, need))
        self._check(fn(self._h, buf, need))
        return buf.value.decode()

    def autotune(self, repeats: int = 2) -> dict:
        """Time the schedule / launch-shape candidates on the current inputs
        and keep the fastest; returns the report (every candidate's ms)."""
        import json
        self._check(_lib.slc_autotune(self._h, int(repeats)))
        return json.loads(self._string(_lib.slc_autotune_report))

    def tuning(self) -> str:
        return self._string(_lib.slc_get_tuning)

    def set_tuning(self, text: str):
        self._check(_lib.slc_set_tuning(self._h, text.encode()))

    def set_graphs(self, on: bool = True):
        """HIP graph capture and replay of contract's launch sequence (default on)."""
        self._check(_lib.slc_set_graphs(self._h, int(on)))

    def set_chunk_budget(self, nbytes: int):
        self._check(_lib.slc_set_chunk_budget(self._h, int(nbytes)))

    def set_precision(self, vdaggerv: str = "fp64", factors: str = "fp64"):
        """Precision forms: vdaggerv "fp64" | "fp32" (single-pass FP32 MFMA) |
        "bf16x3" (three-term split BF16, FP32 accumulation) | "ozaki" (int8
        slices, exact slice products, FP64 combination); factors "fp64" |
        "fp32" | "fp48" (factor slabs and product tables stored in FP32 or in
        48-bit FP64 -- 36 mantissa bits --, arithmetic and accumulation FP64)."""
        vmodes = {"fp64": 0, "fp32": 1, "bf16x3": 2, "ozaki": 3}
        fmodes = {"fp64": 0, "fp32": 1, "fp48": 2}
        if vdaggerv not in vmodes or factors not in fmodes:
            raise ValueError("unknown precision form %r / %r" % (vdaggerv, factors))
        self._check(_lib.slc_set_precision(self._h, vmodes[vdaggerv], fmodes[factors]))

    def ahb_host(self, A: np.ndarray, B: np.ndarray, phase=None) -> np.ndarray:
        """A^H diag(phase[k/3]) B on the device for host matrices A (K x M), B (K x N)."""
        K, M = A.shape
        N = B.shape[1]
        a = np.asfortranarray(A, dtype=np.complex128).reshape(-1, order="F")
        b = np.asfortranarray(B, dtype=np.complex128).reshape(-1, order="F")
        c = np.zeros(M * N, dtype=np.complex128)
        p = None if phase is None else np.ascontiguousarray(phase, dtype=np.complex128)
        self._check(_lib.slc_ahb_host(self._h, _dptr(a), _dptr(b), _dptr(c), None if p is None else _dptr(p), M, N, K))
        return c.reshape(M, N, order="F")

    def ahb_device(self, A_ptr, B_ptr, C_ptr, phase_ptr, M, N, K, lda, ldb, ldc, batch=1, sA=0, sB=0, sC=0):
        """C = A^H diag(phase[k/3]) B on device pointers (ints)."""
        self._check(_lib.slc_ahb_device(self._h, A_ptr, B_ptr, C_ptr, phase_ptr, M, N, K, lda, ldb, ldc,
                                        batch, sA, sB, sC))

    def ahb_device_noalloc(self, A_ptr, B_ptr, C_ptr, phase_ptr, M, N, K, lda, ldb, ldc, batch=1, sA=0, sB=0,
                           sC=0):
        """ahb_device without any device allocation (job table in a fixed
        per-GPU array; batch <= 1024; overlapping calls on one GPU race)."""
        self._check(_lib.slc_ahb_device_noalloc(self._h, A_ptr, B_ptr, C_ptr, phase_ptr, M, N, K, lda, ldb, ldc,
                                                batch, sA, sB, sC))


class OperatorFactory:
    """OperatorFactory (include/OperatorsForMesons.hpp:39-69), handling_vdaggerv=build."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def create_operators(self, config=None):
        """create_operators for the configuration of the last read_inputs, or
        `config` (it names the VdaggerV files of handling_vdaggerv
        write / read / liuming)."""
        if config is None:
            self.ctx._check(_lib.slc_create_operators(self.ctx._h))
        else:
            self.ctx._check(_lib.slc_create_operators_config(self.ctx._h, int(config)))

    def return_vdaggerv(self, index: int, t: int) -> np.ndarray:
        N = self.ctx.query("nev")
        out = np.zeros(N * N, dtype=np.complex128)
        self.ctx._check(_lib.slc_return_vdaggerv(self.ctx._h, index, t, _dptr(out)))
        return out.reshape(N, N, order="F")


class DilutedFactorFactory:
    """DilutedFactorFactory<Q0|Q1|Q2> (include/DilutedFactorFactory.hpp:28-65):
    factory[key][op] -> list of (matrix D x D, (r1, r2))."""

    _QT = {"Q0": 0, "Q1": 1, "Q2": 2}

    def __init__(self, ctx: Context, qtype: str):
        self.ctx = ctx
        self.q = self._QT[qtype]

    def get(self, key, op: int):
        key = (_c.c_int * len(key))(*key)
        D = 4 * self.ctx.query("dilE")
        cap = 256
        out = np.zeros(cap * D * D, dtype=np.complex128)
        ric = np.zeros(2 * cap, dtype=np.int32)
        n = _c.c_int()
        self.ctx._check(_lib.slc_diluted_factor(self.ctx._h, self.q, key, op, _dptr(out),
                                                ric.ctypes.data_as(_I), cap, _c.byref(n)))
        mats = out[: n.value * D * D].reshape(n.value, D, D).transpose(0, 2, 1)  # column-major
        return [(mats[i], (int(ric[2 * i]), int(ric[2 * i + 1]))) for i in range(n.value)]


class DilutedTraceFactory:
    """AbstractDilutedTraceFactory (include/DilutedTraceFactory.hpp:21-29) of one
    trace type: get((t_source, t_sink), locations) -> {trace id: [(value, used mask)]},
    the reference's DilutedTracesMap, computed on the device."""

    def __init__(self, ctx: Context, trace: str):
        self.ctx = ctx
        self.trace = trace

    def get(self, slice_pair, locations):
        locs = [0 if l in (0, "source") else 1 for l in locations]
        lp = (_c.c_int * max(1, len(locs)))(*locs)
        n_ids, total = _c.c_int(), _c.c_int()
        cap_ids, cap = 64, 4096
        for _ in range(2):
            counts = np.zeros(cap_ids, dtype=np.int32)
            vals = np.zeros(2 * cap)
            used = np.zeros(cap, dtype=np.uint16)
            rc = _lib.slc_diluted_trace(self.ctx._h, self.trace.encode(), int(slice_pair[0]), int(slice_pair[1]), lp,
                                        len(locs), counts.ctypes.data_as(_I), cap_ids, _dptr(vals),
                                        used.ctypes.data_as(_c.POINTER(_c.c_uint16)), cap, _c.byref(n_ids),
                                        _c.byref(total))
            if rc == 0:
                break
            if n_ids.value <= cap_ids and total.value <= cap:
                self.ctx._check(rc)
            cap_ids, cap = max(cap_ids, n_ids.value), max(cap, total.value)
        self.ctx._check(rc)
        out, at = {}, 0
        for i in range(n_ids.value):
            c = int(counts[i])
            out[i] = [(complex(vals[2 * k], vals[2 * k + 1]), int(used[k])) for k in range(at, at + c)]
            at += c
        return out


def inner_product(ctx: Context, *traces, kahan: bool = False) -> complex:
    """inner_product(DilutedTraces, ...) (DilutedTrace.cpp:5-82) of 2 or 3
    trace vectors, or the mean of one (resolve_request1, Diagram.cpp:9-27),
    on the device.  Each trace vector is a sequence of (value, used_mask)."""
    k = len(traces)
    if not 1 <= k <= 3:
        raise ValueError("1 to 3 trace vectors")
    vals = [np.ascontiguousarray([complex(v) for v, _ in tr], dtype=np.complex128) for tr in traces]
    used = [np.ascontiguousarray([int(m) for _, m in tr], dtype=np.uint16) for tr in traces]
    vp = (_D * k)(*[_dptr(v) for v in vals])
    up = (_c.POINTER(_c.c_uint16) * k)(*[u.ctypes.data_as(_c.POINTER(_c.c_uint16)) for u in used])
    n = (_c.c_int * k)(*[len(tr) for tr in traces])
    out = np.zeros(2)
    ctx._check(_lib.slc_inner_product(ctx._h, k, vp, up, n, int(kahan), _dptr(out)))
    return complex(out[0], out[1])


def contract(ctx: Context, single_time_slice_combination: int = -1, kahan: bool = False):
    """contract() of Correlators.cpp:40-127 on the device."""
    ctx._check(_lib.slc_contract(ctx._h, single_time_slice_combination, int(kahan)))
