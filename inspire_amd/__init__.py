"""inspire_amd: MI355X (gfx950) ReinsPIRe online PIR answer path.

Python view of ``libpirgpu.so`` (the C ABI declared in include/pirgpu.h).
``PirServer`` mirrors the reference's ``PirServer<Db, uint64_t, Mat>``
(crypto/pir/server/pir_server.h:41-111): ``create`` takes the preprocessed
material, ``process_response`` answers a ``PirRequest`` with a ``PirResponse``,
and the stage-time getters report the DB / packing / poly-eval split. Errors
raise ``PirError`` carrying the reference's status code and message.

The library is built in-tree by ``inspire_amd.build``; importing this package
never falls back to a CPU implementation: without the library or a GPU, calls
fail loudly.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpirgpu.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None

OK = 0
INVALID_ARGUMENT = 3
DEADLINE_EXCEEDED = 4
RESOURCE_EXHAUSTED = 8
FAILED_PRECONDITION = 9
INTERNAL = 13
UNAVAILABLE = 14


class PirError(RuntimeError):
    """A non-OK status from libpirgpu (code = absl::StatusCode number)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def lib() -> ctypes.CDLL:
    """Loads libpirgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m inspire_amd.build`")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.pirg_version.restype = ctypes.c_char_p
        _lib.pirg_server_fft_margin.restype = ctypes.c_double
        _lib.pirg_server_destroy.restype = None
        _lib.pirg_packer_destroy.restype = None
        _lib.pirg_vpir_server_destroy.restype = None
        _lib.pirg_serve_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                           ctypes.c_char_p, ctypes.c_size_t]
        _lib.pirg_serve_poll.argtypes = [ctypes.c_void_p, ctypes.c_int, _u64p, _u64p,
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_char_p,
                                         ctypes.c_size_t]
        _lib.pirg_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_char_p, ctypes.c_size_t]
        _lib.pirg_host_free.argtypes = [ctypes.c_void_p]
        _lib.pirg_host_free.restype = None
        _lib.pirg_request_record_words.argtypes = [ctypes.c_void_p]
        _lib.pirg_request_record_words.restype = ctypes.c_size_t
        _lib.pirg_serve_set_memory_limit.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        _lib.pirg_serve_workspace_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.pirg_serve_workspace_bytes.restype = ctypes.c_size_t
    return _lib


class HostRecords:
    """n request records of a server's layout in pinned host memory
    (pirg_host_alloc), as the (n, words) uint64 array `array`. Records written
    with PirServer.write_record are submitted without a host copy by
    PirServer.submit_records; keep them (and this object) until their
    responses are polled."""

    def __init__(self, server: "PirServer", n: int):
        self.words = int(lib().pirg_request_record_words(server._h))
        p = ctypes.c_void_p()
        err = _err()
        _check(lib().pirg_host_alloc(n * self.words * 8, ctypes.byref(p), err, len(err)), err)
        self._p = p.value
        buf = (ctypes.c_uint64 * (n * self.words)).from_address(self._p)
        self.array = np.ctypeslib.as_array(buf).reshape(n, self.words)

    def close(self) -> None:
        if self._p is not None:
            self.array = None
            lib().pirg_host_free(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _CServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("answered", ctypes.c_longlong),
                ("batches", ctypes.c_longlong), ("mean_batch", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("slots", ctypes.c_int), ("max_batch", ctypes.c_int)]


class _CParams(ctypes.Structure):
    _fields_ = [
        ("log_degree", ctypes.c_int),
        ("log_q", ctypes.c_int),
        ("log_q1", ctypes.c_int),
        ("log_q2", ctypes.c_int),
        ("variance", ctypes.c_int),
        ("plaintext_modulus", ctypes.c_uint64),
        ("pack_log_digit", ctypes.c_int),
        ("pack_num_digits", ctypes.c_int),
        ("eval_log_digit", ctypes.c_int),
        ("eval_num_digits", ctypes.c_int),
        ("num_entries", ctypes.c_int),
        ("interp_degree", ctypes.c_int),
        ("entry_size_multiple", ctypes.c_int),
        ("db_bytes", ctypes.c_int),
        ("mat_bytes", ctypes.c_int),
        ("coeff_bits", ctypes.c_int),
    ]


class _CView(ctypes.Structure):
    _fields_ = [
        ("num_db_matrices", ctypes.c_int),
        ("db_cdata", ctypes.POINTER(_u64p)),
        ("db_rows", ctypes.POINTER(ctypes.c_int)),
        ("db_cols", ctypes.POINTER(ctypes.c_int)),
        ("num_outputs", ctypes.c_int),
        ("outputs_len", ctypes.POINTER(ctypes.c_int)),
        ("pack", ctypes.POINTER(ctypes.c_void_p)),
        ("a_tilde", ctypes.POINTER(_u64p)),
        ("t_vec_h", ctypes.POINTER(_u64p)),
        ("second_dim_a_len", ctypes.c_int),
        ("second_dim_a", _u64p),
    ]


class _CRequest(ctypes.Structure):
    _fields_ = [
        ("first_dim", _u64p),
        ("first_dim_len", ctypes.c_size_t),
        ("second_dim", _u64p),
        ("second_dim_polys", ctypes.c_size_t),
        ("packing_key", _u64p),
        ("packing_key_polys", ctypes.c_size_t),
    ]


@dataclasses.dataclass(frozen=True)
class PirParams:
    """PirParams<uint64_t> (crypto/pir/pir_params.h:36-105) with the
    reference's template types as fields (db_bytes 2 = uint16_t DB, 1 = uint8_t;
    mat_bytes 4 = int32_t pack matrix, 2 = int16_t; coeff_bits 64 = uint64_t
    coefficients, 32 = uint32_t)."""

    log_degree: int = 11
    log_q: int = 52
    log_q1: int = 29
    log_q2: int = 18
    variance: int = 40
    plaintext_modulus: int = 65535
    pack_log_digit: int = 19
    pack_num_digits: int = 2
    eval_log_digit: int = 19
    eval_num_digits: int = 3
    num_entries: int = 4096
    interp_degree: int = 2
    entry_size_multiple: int = 1
    db_bytes: int = 2
    mat_bytes: int = 4
    coeff_bits: int = 64

    @property
    def degree(self) -> int:
        return 1 << self.log_degree

    @property
    def interpolated_rows(self) -> int:
        """PirParams::InterpolatedRows (pir_params.h:71)."""
        return self.num_entries // self.interp_degree

    def c(self) -> _CParams:
        return _CParams(*[getattr(self, f.name) for f in dataclasses.fields(self)])


def metrics_tool_params(n: int, log_d: int, t: int, ell: int = 1) -> PirParams:
    """pir_metrics_tool defaults (crypto/pir/benchmarking/pir_metrics_tool.cc:79-87,111-114)."""
    return PirParams(log_degree=log_d, log_q=52, log_q1=16 + log_d + 2, log_q2=18,
                     variance=40, plaintext_modulus=65535, pack_log_digit=19,
                     pack_num_digits=2, eval_log_digit=19, eval_num_digits=3,
                     num_entries=n, interp_degree=t, entry_size_multiple=ell)


@dataclasses.dataclass
class PirRequest:
    """PirRequest<uint64_t> (crypto/pir/pir_params.h:196-208)."""

    first_dimension_query: np.ndarray          # r .. K values
    second_dimension_query: np.ndarray         # 2e x d (empty when t == 1)
    packing_key: np.ndarray                    # 2k x d


@dataclasses.dataclass
class PirResponse:
    """PirResponse<uint64_t> (crypto/pir/pir_params.h:211-216): ell RLWE
    ciphertexts, as an (ell, 2, d) array of (a, b)."""

    ciphertexts: np.ndarray


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _u64_2d(a) -> np.ndarray:
    # Convert before reshaping: numpy infers float64 for Python ints >= 2^63.
    a = _u64(a)
    return a.reshape(1, -1) if a.ndim == 1 else a


def _check(rc: int, err) -> None:
    if rc != 0:
        raise PirError(rc, err.value.decode())


def _err():
    return ctypes.create_string_buffer(1024)


class _RequestArrays:
    """Keeps the numpy buffers of a request alive while C reads them."""

    def __init__(self, req: PirRequest, d: int):
        self.fd = _u64(req.first_dimension_query).reshape(-1)
        sd = _u64(req.second_dimension_query)
        self.sd = sd.reshape(-1) if sd.size else np.zeros(1, dtype=np.uint64)
        self.sd_polys = sd.shape[0] if sd.ndim == 2 else (sd.size // d)
        pk = _u64(req.packing_key)
        self.pk = pk.reshape(-1)
        self.pk_polys = pk.shape[0] if pk.ndim == 2 else (pk.size // d)

    def view(self) -> _CRequest:
        return _CRequest(self.fd.ctypes.data_as(_u64p), self.fd.size,
                         self.sd.ctypes.data_as(_u64p), self.sd_polys,
                         self.pk.ctypes.data_as(_u64p), self.pk_polys)


class PirServer:
    """PirServer<uint16_t|uint8_t, uint64_t, int32_t|int16_t> on the GPU."""

    def __init__(self, handle, params: PirParams):
        self._h = handle
        self.params = params

    @classmethod
    def create(cls, params: PirParams, db_matrices: Sequence[np.ndarray],
               db_cols: int, pack: Sequence[Sequence[np.ndarray]],
               a_tilde: Sequence[Sequence[np.ndarray]],
               t_vec_h: Sequence[Sequence[np.ndarray]],
               second_dim_a: Optional[np.ndarray],
               db_rows: Optional[Sequence[int]] = None,
               db_cols_each: Optional[Sequence[int]] = None) -> "PirServer":
        """PirServer::Create (crypto/pir/server/pir_server.cc:52-124).

        db_matrices[k]: condensed words of shard k; pack[k][i], a_tilde[k][i],
        t_vec_h[k][i]: the PreprocessMatrixPackOutput fields of (shard k, i)."""
        keep = []
        dbs = [_u64(m) for m in db_matrices]
        keep += dbs
        n_db = len(dbs)
        db_arr = (_u64p * max(1, n_db))(*[m.ctypes.data_as(_u64p) for m in dbs])
        rows = list(db_rows) if db_rows is not None else [m.shape[0] for m in dbs]
        cols = list(db_cols_each) if db_cols_each is not None else [db_cols] * n_db
        rows_arr = (ctypes.c_int * max(1, n_db))(*rows)
        cols_arr = (ctypes.c_int * max(1, n_db))(*cols)
        n_out = len(pack)
        olen = (ctypes.c_int * max(1, n_out))(*[len(x) for x in pack])
        flat_pk, flat_at, flat_th = [], [], []
        for k in range(n_out):
            for i in range(len(pack[k])):
                pk = np.ascontiguousarray(pack[k][i])
                at = _u64(a_tilde[k][i])
                th = _u64(t_vec_h[k][i])
                keep += [pk, at, th]
                flat_pk.append(pk.ctypes.data_as(ctypes.c_void_p))
                flat_at.append(at.ctypes.data_as(_u64p))
                flat_th.append(th.ctypes.data_as(_u64p))
        n_flat = max(1, len(flat_pk))
        pk_arr = (ctypes.c_void_p * n_flat)(*flat_pk)
        at_arr = (_u64p * n_flat)(*flat_at)
        th_arr = (_u64p * n_flat)(*flat_th)
        if second_dim_a is not None and np.asarray(second_dim_a).size:
            sda = _u64(second_dim_a)
            keep.append(sda)
            sda_len = sda.shape[0]
            sda_ptr = sda.ctypes.data_as(_u64p)
        else:
            sda_len, sda_ptr = 0, None
        view = _CView(n_db, db_arr, rows_arr, cols_arr, n_out, olen, pk_arr, at_arr,
                      th_arr, sda_len, sda_ptr)
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_create(ctypes.byref(cp), ctypes.byref(view),
                                      ctypes.byref(handle), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    @classmethod
    def create_flat(cls, params: PirParams, db: np.ndarray, db_cols: int, pack: np.ndarray,
                    a_tilde: np.ndarray, t_vec_h: np.ndarray,
                    second_dim_a: Optional[np.ndarray]) -> "PirServer":
        """Create from the concatenated material (pirg_server_create_flat)."""
        db = _u64(db)
        pack = np.ascontiguousarray(pack)
        a_tilde = _u64(a_tilde)
        t_vec_h = _u64(t_vec_h)
        sda = _u64(second_dim_a) if second_dim_a is not None else np.zeros(1, np.uint64)
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_create_flat(
            ctypes.byref(cp), db.ctypes.data_as(_u64p), db_cols,
            pack.ctypes.data_as(ctypes.c_void_p), a_tilde.ctypes.data_as(_u64p),
            t_vec_h.ctypes.data_as(_u64p), sda.ctypes.data_as(_u64p), ctypes.byref(handle),
            err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    @classmethod
    def create_from_proto(cls, params: PirParams, data: bytes) -> "PirServer":
        """PirServer::Create on a serialized proto::PirPreprocessedData
        (pirg_server_create_from_proto)."""
        buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_create_from_proto(
            ctypes.byref(cp), buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(len(data)),
            ctypes.byref(handle), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    @classmethod
    def create_from_directory(cls, params: PirParams, path: str) -> "PirServer":
        """PirServer::Create on SavePirPreprocessedDataToDirectory's files
        (pirg_server_create_from_directory)."""
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_create_from_directory(
            ctypes.byref(cp), path.encode(), ctypes.byref(handle), err,
            ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    @classmethod
    def load_from_buffer(cls, params: PirParams, buffer: bytes) -> "PirServer":
        """PirServer::LoadFromBuffer (pir_server.cc:594-793) from the byte image
        SaveToFile writes (pirg_server_load_from_buffer)."""
        buf = np.frombuffer(buffer, dtype=np.uint8) if len(buffer) else np.zeros(1, np.uint8)
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_load_from_buffer(
            ctypes.byref(cp), buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(len(buffer)),
            ctypes.byref(handle), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    @classmethod
    def load_from_file(cls, params: PirParams, path: str) -> "PirServer":
        """PirServer::LoadFromFile (pir_server.cc:794-804)."""
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_load_from_file(ctypes.byref(cp), os.fsencode(path),
                                              ctypes.byref(handle), err,
                                              ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    def save_to_file(self, path: str) -> None:
        """PirServer::SaveToFile (pir_server.cc:433-592), written from the
        resident material."""
        err = _err()
        _check(lib().pirg_server_save_to_file(self._h, os.fsencode(path), err,
                                              ctypes.c_size_t(len(err))), err)

    def image_size(self) -> int:
        """Bytes of the SaveToFile image (no device work)."""
        n = ctypes.c_size_t()
        err = _err()
        _check(lib().pirg_server_save_to_buffer(self._h, None, ctypes.c_size_t(0), ctypes.byref(n),
                                                err, ctypes.c_size_t(len(err))), err)
        return n.value

    def save_to_buffer(self) -> np.ndarray:
        """The SaveToFile image as a uint8 array (pirg_server_save_to_buffer)."""
        size = self.image_size()
        out = np.empty(max(size, 1), dtype=np.uint8)
        n = ctypes.c_size_t()
        err = _err()
        _check(lib().pirg_server_save_to_buffer(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                                ctypes.c_size_t(size), ctypes.byref(n), err,
                                                ctypes.c_size_t(len(err))), err)
        return out[:size]

    @classmethod
    def create_preprocessed(cls, params: PirParams, db_raw: np.ndarray,
                            first_dim_a: np.ndarray, second_dim_a: Optional[np.ndarray],
                            pack_a: np.ndarray) -> "PirServer":
        """PreprocessingServer::Preprocess then PirServer::Create, both on the
        GPU (pirg_server_create_preprocessed); the material stays in HBM.
        The stage times of the preprocessing are in `preprocess_ms`."""
        raw = np.ascontiguousarray(db_raw, dtype=np.uint16 if params.db_bytes == 2 else np.uint8)
        fa, pa = _u64(first_dim_a), _u64(pack_a)
        sda = _u64(second_dim_a) if second_dim_a is not None else None
        ms = (ctypes.c_double * 4)()
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_server_create_preprocessed(
            ctypes.byref(cp), raw.ctypes.data_as(ctypes.c_void_p), fa.ctypes.data_as(_u64p),
            sda.ctypes.data_as(_u64p) if sda is not None else None, pa.ctypes.data_as(_u64p),
            ms, ctypes.byref(handle), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        srv = cls(handle, params)
        srv.preprocess_ms = dict(zip(("interpolate", "hints", "preprocess_pack", "total"),
                                     list(ms)))
        return srv

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().pirg_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def _out_shape(self):
        p = self.params
        return (p.entry_size_multiple, 2, p.degree)

    def process_response(self, request: PirRequest) -> PirResponse:
        """PirServer::ProcessResponse (crypto/pir/server/pir_server.cc:152-306)."""
        ra = _RequestArrays(request, self.params.degree)
        out = np.zeros(self._out_shape, dtype=np.uint64)
        err = _err()
        rc = lib().pirg_process(self._h, ra.fd.ctypes.data_as(_u64p), ctypes.c_size_t(ra.fd.size),
                                ra.sd.ctypes.data_as(_u64p), ctypes.c_size_t(ra.sd_polys),
                                ra.pk.ctypes.data_as(_u64p), ctypes.c_size_t(ra.pk_polys),
                                out.ctypes.data_as(_u64p), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return PirResponse(out)

    def process_batch(self, requests: Sequence[PirRequest]):
        """Answers many requests in one batched pass; returns (responses,
        statuses, first error message); responses[i] is None when statuses[i] != 0."""
        n = len(requests)
        ras = [_RequestArrays(r, self.params.degree) for r in requests]
        views = (_CRequest * max(1, n))(*[ra.view() for ra in ras])
        outs = [np.zeros(self._out_shape, dtype=np.uint64) for _ in range(n)]
        out_ptrs = (_u64p * max(1, n))(*[o.ctypes.data_as(_u64p) for o in outs])
        statuses = (ctypes.c_int * max(1, n))()
        err = _err()
        rc = lib().pirg_process_batch(self._h, n, views, out_ptrs, statuses, err,
                                      ctypes.c_size_t(len(err)))
        _check(rc, err)
        st = list(statuses)[:n]
        resp: List[Optional[PirResponse]] = [PirResponse(o) if s == 0 else None
                                             for o, s in zip(outs, st)]
        return resp, st, err.value.decode()

    # Device-resident batches (inputs staged in HBM before the timed region).
    def batch_stage(self, requests: Sequence[PirRequest]) -> None:
        n = len(requests)
        self._staged = [_RequestArrays(r, self.params.degree) for r in requests]
        views = (_CRequest * max(1, n))(*[ra.view() for ra in self._staged])
        err = _err()
        rc = lib().pirg_batch_stage(self._h, n, views, err, ctypes.c_size_t(len(err)))
        _check(rc, err)

    def batch_run(self, n: int) -> None:
        err = _err()
        rc = lib().pirg_batch_run(self._h, n, err, ctypes.c_size_t(len(err)))
        _check(rc, err)

    def batch_fetch(self, n: int) -> List[PirResponse]:
        outs = [np.zeros(self._out_shape, dtype=np.uint64) for _ in range(n)]
        out_ptrs = (_u64p * max(1, n))(*[o.ctypes.data_as(_u64p) for o in outs])
        err = _err()
        rc = lib().pirg_batch_fetch(self._h, n, out_ptrs, err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return [PirResponse(o) for o in outs]

    def stage_times_ms(self):
        """(GetDbMultiplyTime, GetPackingTime, GetPolyEvalTime) in ms
        (crypto/pir/server/pir_server.h:70-72)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().pirg_server_stage_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def scan_times_ms(self, reset: bool = False):
        """(DB-scan ms, pack-scan ms, batches) summed since the last reset."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        lib().pirg_server_scan_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n),
                                     int(reset))
        return a.value, b.value, n.value

    def scans_fused(self, batch: int) -> bool:
        """Whether batches of this size run the DB and pack scans as one
        launch (scan_times_ms then reports it as the DB scan's time)."""
        f = ctypes.c_int()
        _check(lib().pirg_server_scans_fused(self._h, int(batch), ctypes.byref(f)), _err())
        return bool(f.value)

    def fft_margin(self, reset: bool = False) -> float:
        """Largest |x - round(x)| the FFT products have rounded so far."""
        return lib().pirg_server_fft_margin(self._h, int(reset))

    def autotune(self, widths: Sequence[int]) -> str:
        """Times the scan configurations for each batch width and keeps the
        fastest; returns the per-candidate report."""
        arr = (ctypes.c_int * max(1, len(widths)))(*widths)
        rep = ctypes.create_string_buffer(1 << 16)
        rc = lib().pirg_server_autotune(self._h, arr, len(widths), rep, ctypes.c_size_t(len(rep)))
        _check(rc, rep)
        return rep.value.decode()

    def use_graphs(self, enable: bool = True) -> None:
        """process_batch / batch_run replay one captured hipGraph per batch size."""
        _check(lib().pirg_server_use_graphs(self._h, int(enable)), _err())

    def time_graph_scans(self, enable: bool = True) -> None:
        """Graphs captured from now on carry event nodes around the two scans,
        so scan_times_ms also counts graph replays."""
        _check(lib().pirg_server_time_graph_scans(self._h, int(enable)), _err())

    # Serving loop (continuous batching; thread-safe submit / poll).
    def serve_start(self, max_batch: int = 64, slots: int = 2,
                    max_outstanding: int = 4096) -> None:
        err = _err()
        rc = lib().pirg_serve_start(self._h, max_batch, slots, max_outstanding, err,
                                    ctypes.c_size_t(len(err)))
        _check(rc, err)

    def serve_set_memory_limit(self, nbytes: int) -> None:
        """Device memory serve_start may give its slots (0: 80% of the free
        memory at start)."""
        _check(lib().pirg_serve_set_memory_limit(self._h, int(nbytes)), _err())

    def serve_workspace_bytes(self, batch: int) -> int:
        """Device bytes one serving slot takes for batches of `batch`."""
        return int(lib().pirg_serve_workspace_bytes(self._h, int(batch)))

    def submit(self, request: PirRequest, tag: int) -> None:
        """Queues a request (validated like ProcessResponse); raises PirError
        (RESOURCE_EXHAUSTED when max_outstanding requests await polling)."""
        ra = _RequestArrays(request, self.params.degree)
        view = ra.view()
        err = _err()
        rc = lib().pirg_serve_submit(self._h, ctypes.byref(view), ctypes.c_uint64(tag), err,
                                     ctypes.c_size_t(len(err)))
        _check(rc, err)

    def submit_batch(self, requests: Sequence[PirRequest], tags: Sequence[int]) -> List[int]:
        """Queues many requests in one call (pirg_serve_submit_batch); returns
        each one's validation status (0 = queued)."""
        n = len(requests)
        ras = [_RequestArrays(r, self.params.degree) for r in requests]
        views = (_CRequest * max(1, n))(*[ra.view() for ra in ras])
        tg = (ctypes.c_uint64 * max(1, n))(*tags)
        st = (ctypes.c_int * max(1, n))()
        err = _err()
        rc = lib().pirg_serve_submit_batch(self._h, n, views, tg, st, err,
                                           ctypes.c_size_t(len(err)))
        _check(rc, err)
        return list(st)[:n]

    def record_words(self) -> int:
        """Words of one request record (pirg_request_record_words)."""
        return int(lib().pirg_request_record_words(self._h))

    def write_record(self, request: PirRequest, record: np.ndarray) -> None:
        """Validates a request as ProcessResponse does and writes its record
        into `record` (a row of HostRecords.array)."""
        if record.dtype != np.uint64 or not record.flags.c_contiguous or \
                record.size < self.record_words():
            raise ValueError("record must be a contiguous uint64 row of record_words() words")
        ra = _RequestArrays(request, self.params.degree)
        view = ra.view()
        err = _err()
        _check(lib().pirg_request_record(self._h, ctypes.byref(view),
                                         record.ctypes.data_as(_u64p), err,
                                         ctypes.c_size_t(len(err))), err)

    def submit_records(self, records: Sequence[np.ndarray], tags: Sequence[int]) -> None:
        """Queues records written by write_record into HostRecords memory,
        with no host copy (pirg_serve_submit_records): all or none."""
        n = len(records)
        ptrs = (_u64p * max(1, n))(*[r.ctypes.data_as(_u64p) for r in records])
        tg = (ctypes.c_uint64 * max(1, n))(*tags)
        err = _err()
        _check(lib().pirg_serve_submit_records(self._h, n, ptrs, tg, err,
                                               ctypes.c_size_t(len(err))), err)

    def poll_batch(self, max_results: int, timeout_ms: int = 1000):
        """Up to max_results finished requests as (tags, responses array
        (count, ell, 2, d), statuses, latencies); None when none finished in
        time."""
        outs = np.zeros((max_results,) + self._out_shape, dtype=np.uint64)
        ptrs = (_u64p * max_results)(*[outs[i].ctypes.data_as(_u64p)
                                        for i in range(max_results)])
        tags = (ctypes.c_uint64 * max_results)()
        st = (ctypes.c_int * max_results)()
        lat = (ctypes.c_double * max_results)()
        count = ctypes.c_int()
        err = _err()
        rc = lib().pirg_serve_poll_batch(self._h, max_results, int(timeout_ms), tags, ptrs, st, lat,
                                         ctypes.byref(count), err, ctypes.c_size_t(len(err)))
        if rc == DEADLINE_EXCEEDED:
            return None
        _check(rc, err)
        n = count.value
        return list(tags)[:n], outs[:n], list(st)[:n], list(lat)[:n]

    def poll(self, timeout_ms: int = 1000):
        """Next finished request as (tag, response or None, status, latency_ms,
        message), or None when nothing finished within timeout_ms."""
        out = np.zeros(self._out_shape, dtype=np.uint64)
        tag = ctypes.c_uint64()
        status = ctypes.c_int()
        lat = ctypes.c_double()
        err = _err()
        rc = lib().pirg_serve_poll(self._h, int(timeout_ms), ctypes.byref(tag),
                                   out.ctypes.data_as(_u64p), ctypes.byref(status),
                                   ctypes.byref(lat), err, ctypes.c_size_t(len(err)))
        if rc == DEADLINE_EXCEEDED:
            return None
        _check(rc, err)
        resp = PirResponse(out) if status.value == 0 else None
        return tag.value, resp, status.value, lat.value, err.value.decode()

    def serve_stats(self, reset: bool = False) -> dict:
        st = _CServeStats()
        _check(lib().pirg_serve_stats(self._h, ctypes.byref(st), int(reset)), _err())
        return {f: getattr(st, f) for f, _ in _CServeStats._fields_}

    def serve_stop(self) -> None:
        lib().pirg_serve_stop(self._h)


class PirPacker:
    """Batched MatrixPack (crypto/lwes_to_rlwe.cc:481-508) over independent
    jobs with resident PreprocessMatrixPackOutput material."""

    def __init__(self, handle, params: PirParams, njobs: int):
        self._h = handle
        self.params = params
        self.njobs = njobs

    @classmethod
    def create(cls, params: PirParams, pack: Sequence[np.ndarray],
               a_tilde: Sequence[np.ndarray], t_vec_h: Sequence[np.ndarray]) -> "PirPacker":
        J = len(pack)
        if not (len(a_tilde) == len(t_vec_h) == J):
            raise ValueError("pack, a_tilde and t_vec_h need one entry per job")
        packs = [np.ascontiguousarray(p) for p in pack]
        ats = [_u64(a) for a in a_tilde]
        ths = [_u64_2d(t) for t in t_vec_h]
        pk_arr = (ctypes.c_void_p * J)(*[p.ctypes.data for p in packs])
        at_arr = (_u64p * J)(*[a.ctypes.data_as(_u64p) for a in ats])
        th_arr = (_u64p * J)(*[t.ctypes.data_as(_u64p) for t in ths])
        cp = params.c()
        handle = ctypes.c_void_p()
        err = _err()
        rc = lib().pirg_packer_create(ctypes.byref(cp), J, pk_arr, at_arr, th_arr,
                                      ctypes.byref(handle), err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params, J)

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().pirg_packer_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def use_graph(self, enable: bool = True) -> None:
        lib().pirg_packer_use_graph(self._h, int(enable))

    def run(self, b: Sequence[np.ndarray], keys: Sequence[np.ndarray]) -> np.ndarray:
        """MatrixPack for every job: b[j] (m_j <= d LWE b-values), keys[j] the
        packing key y_vec_g || y_vec_h (2k x d). Returns (njobs, 2, d)."""
        J, d = self.njobs, self.params.degree
        bs = [_u64(x) for x in b]
        ks = [_u64_2d(x) for x in keys]
        out = np.zeros((J, 2, d), dtype=np.uint64)
        b_arr = (_u64p * J)(*[x.ctypes.data_as(_u64p) for x in bs])
        m_arr = (ctypes.c_int * J)(*[x.size for x in bs])
        k_arr = (_u64p * J)(*[x.ctypes.data_as(_u64p) for x in ks])
        o_arr = (_u64p * J)(*[out[j].ctypes.data_as(_u64p) for j in range(J)])
        err = _err()
        rc = lib().pirg_packer_run(self._h, b_arr, m_arr, k_arr, o_arr, err,
                                   ctypes.c_size_t(len(err)))
        _check(rc, err)
        return out

    def times_ms(self, reset: bool = False):
        """(total batch ms, pack-scan ms of eager runs, runs)."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        lib().pirg_packer_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n),
                                int(reset))
        return a.value, b.value, n.value


class _CVpirRequest(ctypes.Structure):
    _fields_ = [("first_dim", ctypes.c_void_p), ("first_dim_polys", ctypes.c_size_t),
                ("packing_key", ctypes.c_void_p), ("packing_key_polys", ctypes.c_size_t)]


class VpirServer:
    """VpirServer's online path (crypto/vpir/server/vpir_server.h:47-121) over
    libpirgpu (pirg_vpir_*)."""

    def __init__(self, handle, params: PirParams):
        self._h = handle
        self.params = params

    @classmethod
    def create(cls, params: PirParams, local_finalize: bool, db: np.ndarray, db_rows: int,
               db_cols: int, pack: np.ndarray, a_tilde: np.ndarray,
               t_vec_h: np.ndarray) -> "VpirServer":
        db = _u64(db)
        pack = np.ascontiguousarray(pack)
        a_tilde = _u64(a_tilde)
        t_vec_h = _u64(t_vec_h)
        handle = ctypes.c_void_p()
        err = _err()
        cp = params.c()
        rc = lib().pirg_vpir_server_create(
            ctypes.byref(cp), int(local_finalize), db.ctypes.data_as(_u64p), db_rows, db_cols,
            pack.ctypes.data_as(ctypes.c_void_p), a_tilde.shape[0] if a_tilde.ndim > 1 else 1,
            a_tilde.ctypes.data_as(_u64p), t_vec_h.ctypes.data_as(_u64p), ctypes.byref(handle),
            err, ctypes.c_size_t(len(err)))
        _check(rc, err)
        return cls(handle, params)

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().pirg_vpir_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _views(self, requests):
        d = self.params.degree
        keep, views = [], (_CVpirRequest * max(len(requests), 1))()
        for i, (fd, pk) in enumerate(requests):
            fd = _u64(fd).reshape(-1)
            pk = _u64(pk).reshape(-1)
            keep += [fd, pk]
            views[i] = _CVpirRequest(fd.ctypes.data, fd.size // d, pk.ctypes.data, pk.size // d)
        return keep, views

    def process_batch(self, requests):
        """[(first_dim, packing_key)] -> ([ell x d b arrays], statuses, message)."""
        p = self.params
        n = len(requests)
        keep, views = self._views(requests)
        outs = [np.zeros((p.entry_size_multiple, p.degree), dtype=np.uint64) for _ in range(n)]
        ptrs = (_u64p * max(n, 1))(*[o.ctypes.data_as(_u64p) for o in outs])
        st = (ctypes.c_int * max(n, 1))()
        err = _err()
        rc = lib().pirg_vpir_process_batch(self._h, n, views, ptrs, st, err,
                                           ctypes.c_size_t(len(err)))
        _check(rc, err)
        return outs, list(st)[:n], err.value.decode()

    def process_response(self, first_dim, packing_key) -> np.ndarray:
        keep, views = self._views([(first_dim, packing_key)])
        p = self.params
        out = np.zeros((p.entry_size_multiple, p.degree), dtype=np.uint64)
        err = _err()
        rc = lib().pirg_vpir_process(self._h, views, out.ctypes.data_as(_u64p), err,
                                     ctypes.c_size_t(len(err)))
        _check(rc, err)
        return out

    def batch_stage(self, requests) -> None:
        keep, views = self._views(requests)
        err = _err()
        _check(lib().pirg_vpir_batch_stage(self._h, len(requests), views, err,
                                           ctypes.c_size_t(len(err))), err)

    def batch_run(self, n: int) -> None:
        err = _err()
        _check(lib().pirg_vpir_batch_run(self._h, n, err, ctypes.c_size_t(len(err))), err)

    def batch_fetch(self, n: int):
        p = self.params
        outs = [np.zeros((p.entry_size_multiple, p.degree), dtype=np.uint64) for _ in range(n)]
        ptrs = (_u64p * max(n, 1))(*[o.ctypes.data_as(_u64p) for o in outs])
        err = _err()
        _check(lib().pirg_vpir_batch_fetch(self._h, n, ptrs, err, ctypes.c_size_t(len(err))),
               err)
        return outs

    def use_graphs(self, enable: bool = True) -> None:
        lib().pirg_vpir_use_graphs(self._h, int(enable))

    def time_graph_scans(self, enable: bool = True) -> None:
        _check(lib().pirg_vpir_time_graph_scans(self._h, int(enable)), _err())

    def scan_times_ms(self, reset: bool = False):
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        lib().pirg_vpir_scan_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n),
                                   int(reset))
        return a.value, b.value, n.value

    def autotune(self, widths: Sequence[int]) -> str:
        w = (ctypes.c_int * len(widths))(*widths)
        rep = ctypes.create_string_buffer(1 << 16)
        err = _err()
        _check(lib().pirg_vpir_autotune(self._h, w, len(widths), rep, ctypes.c_size_t(1 << 16),
                                        err, ctypes.c_size_t(len(err))), err)
        return rep.value.decode()


def poly_to_proto(coeffs, log_modulus: int) -> List[bytes]:
    """Polynomial::ToProto(log_modulus).encoded_coeffs (polynomial.cc:562-590)
    for each row of coeffs (npolys x len), on the GPU (pirg_poly_to_proto)."""
    c = _u64_2d(coeffs)
    npolys, length = c.shape
    nb = (length * log_modulus + 7) // 8 if 0 < log_modulus <= 64 else 0
    out = np.zeros(max(npolys * nb, 1), dtype=np.uint8)
    err = _err()
    rc = lib().pirg_poly_to_proto(c.ctypes.data_as(_u64p), npolys, length, log_modulus,
                                  out.ctypes.data_as(ctypes.c_void_p), err,
                                  ctypes.c_size_t(len(err)))
    _check(rc, err)
    return [out[i * nb:(i + 1) * nb].tobytes() for i in range(npolys)]


def poly_from_proto(encoded: Sequence[bytes], length: int, log_modulus: int) -> np.ndarray:
    """Polynomial::CreateFromProto (polynomial.cc:208-252) for byte strings of
    one size, on the GPU (pirg_poly_from_proto): an npolys x length array."""
    npolys = len(encoded)
    nb = len(encoded[0]) if npolys else 0
    if any(len(e) != nb for e in encoded):
        raise PirError(INVALID_ARGUMENT, "Unexpected serialized size.")
    blob = np.frombuffer(b"".join(encoded), dtype=np.uint8) if nb else np.zeros(1, np.uint8)
    out = np.zeros((max(npolys, 1), max(length, 1)), dtype=np.uint64)
    err = _err()
    rc = lib().pirg_poly_from_proto(blob.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(nb),
                                    npolys, length, log_modulus, out.ctypes.data_as(_u64p), err,
                                    ctypes.c_size_t(len(err)))
    _check(rc, err)
    return out[:npolys, :length] if npolys else np.zeros((0, max(length, 0)), np.uint64)


def poly_to_proto_device(d_coeffs: int, npolys: int, length: int, log_modulus: int,
                         d_words: int, pitch_words: int, stream: int = 0) -> None:
    """pirg_poly_to_proto_device on device addresses (e.g. torch data_ptr())."""
    err = _err()
    rc = lib().pirg_poly_to_proto_device(ctypes.c_void_p(d_coeffs), npolys, length, log_modulus,
                                         ctypes.c_void_p(d_words), ctypes.c_longlong(pitch_words),
                                         ctypes.c_void_p(stream), err, ctypes.c_size_t(len(err)))
    _check(rc, err)


def poly_from_proto_device(d_words: int, pitch_words: int, npolys: int, length: int,
                           log_modulus: int, d_coeffs: int, stream: int = 0) -> None:
    """pirg_poly_from_proto_device on device addresses."""
    err = _err()
    rc = lib().pirg_poly_from_proto_device(ctypes.c_void_p(d_words),
                                           ctypes.c_longlong(pitch_words), npolys, length,
                                           log_modulus, ctypes.c_void_p(d_coeffs),
                                           ctypes.c_void_p(stream), err,
                                           ctypes.c_size_t(len(err)))
    _check(rc, err)


def matrix_multiply_u16(cdata: np.ndarray, cols: int, vec: np.ndarray) -> np.ndarray:
    """Matrix<uint16_t>::Multiply<uint64_t> on condensed words (crypto/matrix.cc:261-318)."""
    cdata = _u64(cdata)
    vec = _u64(vec)
    rows = cdata.size // (cols // 4)
    out = np.zeros(rows, dtype=np.uint64)
    err = _err()
    rc = lib().pirg_matrix_multiply_u16(cdata.ctypes.data_as(_u64p), rows, cols,
                                        vec.ctypes.data_as(_u64p), out.ctypes.data_as(_u64p),
                                        err, ctypes.c_size_t(len(err)))
    _check(rc, err)
    return out


def matrix_multiply_i32(cdata: np.ndarray, cols: int, vec: np.ndarray) -> np.ndarray:
    """Matrix<int32_t>::Multiply<uint64_t> on condensed words (crypto/matrix.cc:320-374)."""
    cdata = _u64(cdata)
    vec = _u64(vec)
    rows = cdata.size // (cols // 2)
    out = np.zeros(rows, dtype=np.uint64)
    err = _err()
    rc = lib().pirg_matrix_multiply_i32(cdata.ctypes.data_as(_u64p), rows, cols,
                                        vec.ctypes.data_as(_u64p), out.ctypes.data_as(_u64p),
                                        err, ctypes.c_size_t(len(err)))
    _check(rc, err)
    return out


def inner_product_fft(u, v, u_bits=64, v_bits=64, chunk_bits=20, u_signed=False) -> np.ndarray:
    """Polynomial::InnerProductFft (crypto/polynomial_fft.cc:68-197) on the GPU."""
    u = _u64_2d(u)
    v = _u64_2d(v)
    n, d = u.shape
    out = np.zeros(d, dtype=np.uint64)
    err = _err()
    rc = lib().pirg_inner_product_fft(u.ctypes.data_as(_u64p), v.ctypes.data_as(_u64p), n, d,
                                      u_bits, v_bits, chunk_bits, int(u_signed),
                                      out.ctypes.data_as(_u64p), err, ctypes.c_size_t(len(err)))
    _check(rc, err)
    return out


def inner_product_fft_noalloc(d_u: int, d_v: int, n: int, d: int, u_bits: int, v_bits: int,
                              chunk_bits: int, u_signed: bool, d_out: int,
                              stream: int = 0) -> None:
    """pirg_inner_product_fft_noalloc on device addresses: queued on `stream`,
    no device allocation (the working memory is the library's static array)."""
    err = _err()
    rc = lib().pirg_inner_product_fft_noalloc(ctypes.c_void_p(d_u), ctypes.c_void_p(d_v), n, d,
                                              u_bits, v_bits, chunk_bits, int(u_signed),
                                              ctypes.c_void_p(d_out), ctypes.c_void_p(stream),
                                              err, ctypes.c_size_t(len(err)))
    _check(rc, err)


# ---- Offline preprocessing (crypto/pir/server/preprocessing_server.h) ----

def interpolate(rows, d: int, t: int, plaintext_modulus: int, db_bytes: int = 2) -> np.ndarray:
    """Interpolate (crypto/db_row_interpolation.cc:115-228) of every row (t*d
    values) on the GPU."""
    dt = np.uint16 if db_bytes == 2 else np.uint8
    rows = np.ascontiguousarray(rows, dtype=dt)
    nrows = rows.size // (t * d)
    out = np.empty_like(rows)
    err = _err()
    rc = lib().pirg_interpolate(rows.ctypes.data_as(ctypes.c_void_p), db_bytes,
                                ctypes.c_longlong(nrows), d, t,
                                ctypes.c_uint32(plaintext_modulus),
                                out.ctypes.data_as(ctypes.c_void_p), err,
                                ctypes.c_size_t(len(err)))
    _check(rc, err)
    return out


def compute_hint_matrix(db, a_components, db_bytes: int = 2) -> np.ndarray:
    """ComputeHintMatrix (crypto/hint_computation.cc:78-160) on the GPU: db is
    rows x cols, a_components n x d; returns rows x d."""
    dt = np.uint16 if db_bytes == 2 else np.uint8
    db = np.ascontiguousarray(db, dtype=dt)
    a = _u64_2d(a_components)
    rows = db.shape[0] if db.ndim == 2 else 0
    cols = db.shape[1] if db.ndim == 2 else 0
    n, d = a.shape
    out = np.empty((max(rows, 1), d), dtype=np.uint64)
    err = _err()
    rc = lib().pirg_compute_hint_matrix(db.ctypes.data_as(ctypes.c_void_p), db_bytes, rows, cols,
                                        a.ctypes.data_as(_u64p), n, d,
                                        out.ctypes.data_as(_u64p), err,
                                        ctypes.c_size_t(len(err)))
    _check(rc, err)
    return out[:rows]


def preprocess_matrix_pack(params: PirParams, a_is, w):
    """PreprocessMatrixPack (crypto/lwes_to_rlwe.cc:421-446) on the GPU for
    nblocks blocks: a_is is (nblocks*d) x d, w is 2k x d. Returns (pack,
    a_tilde, t_vec_h): pack is (nblocks*d) x k*d/2 condensed words (mat_bytes
    4) or (nblocks*d) x k*d int16."""
    d, k = params.degree, params.pack_num_digits
    a_is = _u64_2d(a_is)
    w = _u64_2d(w)
    nb = a_is.shape[0] // d
    if params.mat_bytes == 4:
        pack = np.empty((nb * d, k * d // 2), dtype=np.uint64)
    else:
        pack = np.empty((nb * d, k * d), dtype=np.int16)
    a_tilde = np.empty((nb, d), dtype=np.uint64)
    t_vec_h = np.empty((nb, k, d), dtype=np.uint64)
    err = _err()
    cp = params.c()
    rc = lib().pirg_preprocess_matrix_pack(
        ctypes.byref(cp), nb, a_is.ctypes.data_as(_u64p), w.ctypes.data_as(_u64p),
        pack.ctypes.data_as(ctypes.c_void_p), a_tilde.ctypes.data_as(_u64p),
        t_vec_h.ctypes.data_as(_u64p), err, ctypes.c_size_t(len(err)))
    _check(rc, err)
    return pack, a_tilde, t_vec_h


def preprocess(params: PirParams, db_raw, first_dim_a, pack_a) -> dict:
    """PreprocessingServer::Preprocess on the GPU (pirg_preprocess) with the
    given a-components. Returns the material in the pirg_server_create_flat
    layouts (db, db_cols, pack, a_tilde, t_vec_h) and the stage times."""
    d, t, ell, k = params.degree, params.interp_degree, params.entry_size_multiple, \
        params.pack_num_digits
    r = params.num_entries // t
    simd = 64 if params.db_bytes == 1 else 32
    K = (r + simd - 1) // simd * simd
    M = ell * t * d
    raw = np.ascontiguousarray(db_raw, dtype=np.uint16 if params.db_bytes == 2 else np.uint8)
    fa, pa = _u64_2d(first_dim_a), _u64_2d(pack_a)
    db = np.empty((M, K * params.db_bytes // 8), dtype=np.uint64)
    if params.mat_bytes == 4:
        pack = np.empty((M, k * d // 2), dtype=np.uint64)
    else:
        pack = np.empty((M, k * d), dtype=np.int16)
    a_tilde = np.empty((ell * t, d), dtype=np.uint64)
    t_vec_h = np.empty((ell * t, k, d), dtype=np.uint64)
    ms = (ctypes.c_double * 4)()
    cols = ctypes.c_int()
    err = _err()
    cp = params.c()
    rc = lib().pirg_preprocess(
        ctypes.byref(cp), raw.ctypes.data_as(ctypes.c_void_p), fa.ctypes.data_as(_u64p),
        pa.ctypes.data_as(_u64p), db.ctypes.data_as(_u64p), ctypes.byref(cols),
        pack.ctypes.data_as(ctypes.c_void_p), a_tilde.ctypes.data_as(_u64p),
        t_vec_h.ctypes.data_as(_u64p), ms, err, ctypes.c_size_t(len(err)))
    _check(rc, err)
    assert cols.value == K
    return dict(db=db, db_cols=K, pack=pack, a_tilde=a_tilde, t_vec_h=t_vec_h,
                stage_ms=dict(zip(("interpolate", "hints", "preprocess_pack", "total"),
                                  list(ms))))


def version() -> str:
    return lib().pirg_version().decode()
