"""cuspfft_amd -- MI355X-native binary-sparse 2-D R2C FFT.

Python host mirror of the reference's entry points for this path
(reference include/sparse_fft.h:7-13, 57, 70) over the C ABI in
``include/spfft_c.h``, implemented by hand-written gfx950 HIP kernels in
``cuspfft_amd/_lib/libspfft_hip.so``.  PyTorch is used only for device memory
and streams.  There is no CPU fallback: if the HIP library is missing every
call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libspfft_hip.so")
CLI_PATH = os.path.join(_HERE, "_lib", "spfft_cli")

SMOOTH_3 = 3
SMOOTH_7 = 7
PREC_FP32 = 0
PREC_BF16X3 = 1
# row-pass kernels (spfft_opts.row_kernel)
ROW_AUTO, ROW_FOUR_STEP, ROW_FUSED = 0, 1, 2
# column-pass half-period row pairing (spfft_opts.colpass_pair)
PAIR_AUTO, PAIR_OFF = 0, 1
# row transform (spfft_opts.row_transform): Bluestein, a direct length-cols FFT,
# Rader's convolutions for cols = 2p (p an odd prime), or the split cols = n1 * n2 (an
# on-chip row pass of width n2, then length-n1 DFTs across)
TRANSFORM_AUTO, TRANSFORM_BLUESTEIN, TRANSFORM_DIRECT, TRANSFORM_RADER, TRANSFORM_SPLIT = 0, 1, 2, 3, 4
# batch execution (spfft_batch_set_mode): shared descriptor-driven launches, or per matrix
BATCH_AUTO, BATCH_PER_MATRIX = 0, 1

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "HIP", 3: "ALLOC", 4: "UNSUPPORTED", 5: "BAD_INPUT"}


class SpfftError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class Opts(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_int), ("length_rule", ctypes.c_int), ("u_tile", ctypes.c_int),
                ("precision", ctypes.c_int), ("fft_len", ctypes.c_int), ("m1", ctypes.c_int),
                ("conv_rows", ctypes.c_int), ("colpass_rows", ctypes.c_int),
                ("colpass_table", ctypes.c_int), ("row_limit", ctypes.c_int),
                ("row_kernel", ctypes.c_int), ("colpass_pair", ctypes.c_int),
                ("row_transform", ctypes.c_int), ("row_persist", ctypes.c_int),
                ("t_rows", ctypes.c_int), ("row_threads", ctypes.c_int)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_int),
                ("rows", ctypes.c_int), ("cols", ctypes.c_int), ("out_cols", ctypes.c_int),
                ("pitch", ctypes.c_int), ("fft_len", ctypes.c_int), ("m1", ctypes.c_int),
                ("m2", ctypes.c_int), ("packed", ctypes.c_int), ("u_tile", ctypes.c_int),
                ("conv_rows", ctypes.c_int), ("conv_fast", ctypes.c_int),
                ("conv_stages", ctypes.c_int), ("conv_radix", ctypes.c_int * 8),
                ("colpass_rows", ctypes.c_int), ("colpass_full_table", ctypes.c_int),
                ("precision", ctypes.c_int), ("workspace_bytes", ctypes.c_size_t),
                ("row_kernel", ctypes.c_int), ("colpass_paired", ctypes.c_int),
                ("direct", ctypes.c_int), ("rader", ctypes.c_int), ("row_transform", ctypes.c_int),
                ("t_rows", ctypes.c_int), ("row_threads", ctypes.c_int)]

    def __init__(self):
        super().__init__()
        self.struct_size = ctypes.sizeof(PlanInfo)


class Result(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_int), ("d_output", ctypes.c_void_p),
                ("output_cols", ctypes.c_int), ("ms", ctypes.c_float), ("ms_e2e", ctypes.c_float),
                ("mem_bytes", ctypes.c_size_t), ("fft_len", ctypes.c_int),
                ("row_transform", ctypes.c_int)]

    def __init__(self):
        super().__init__()
        self.struct_size = ctypes.sizeof(Result)


class CscView(ctypes.Structure):
    _fields_ = [("n_sparse_cols", ctypes.c_int), ("nnz", ctypes.c_int), ("packed", ctypes.c_int),
                ("d_sparse_cols", ctypes.c_void_p), ("d_col_ptr", ctypes.c_void_p),
                ("d_row_idx", ctypes.c_void_p), ("d_col_start", ctypes.c_void_p)]


class Profile(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 5), ("launches", ctypes.c_int * 5)]


PROFILE_STAGES = ("csc_build", "colpass", "fft1", "conv", "ifft1_finalize")


class HostCoo(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int), ("cols", ctypes.c_int), ("nnz", ctypes.c_int),
                ("row_idx", ctypes.POINTER(ctypes.c_int)),
                ("col_idx", ctypes.POINTER(ctypes.c_int))]


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"cuspfft_amd: HIP library not built ({LIB_PATH}); "
                           "run `python -m cuspfft_amd.build`")
    # torch ships its own HIP runtime with the same SONAME (libamdhip64.so.7).
    # Importing it first makes the library bind to that one runtime; loading the
    # library first would put a second HIP runtime in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "spfft_opts_default": (None, [ctypes.POINTER(Opts)]),
        "spfft_reference_opts": (None, [_i, _i, ctypes.POINTER(Opts)]),
        "spfft_last_error": (ctypes.c_char_p, []),
        "spfft_version": (ctypes.c_char_p, []),
        "spfft_plan_create": (_i, [_i, _i, ctypes.POINTER(Opts), ctypes.POINTER(_vp)]),
        "spfft_plan_destroy": (_i, [_vp]),
        "spfft_plan_get_info": (_i, [_vp, ctypes.POINTER(PlanInfo)]),
        "spfft_plan_describe": (_i, [_i, _i, ctypes.POINTER(Opts), ctypes.POINTER(PlanInfo)]),
        "spfft_csc_build": (_i, [_vp, _vp, _vp, _i, _vp]),
        "spfft_csc_build_noalloc": (_i, [_i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
        "spfft_csc_get": (_i, [_vp, _vp, ctypes.POINTER(CscView)]),
        "spfft_rowpass": (_i, [_vp, _vp, ctypes.c_int64, _vp]),
        "spfft_execute": (_i, [_vp, _vp, _vp, _i, _vp, ctypes.c_int64, _vp]),
        "spfft_execute_to_host": (_i, [_vp, _vp, _vp, _i, _vp, ctypes.c_int64, _vp]),
        "spfft_rowpass_to_host": (_i, [_vp, _vp, ctypes.c_int64, _vp]),
        "spfft_run_host_coo_to_host": (_i, [_i, _i, _i, _vp, _vp, ctypes.POINTER(Opts), _vp,
                                            ctypes.c_int64, ctypes.POINTER(Result)]),
        "spfft_check_input": (_i, [_vp, _vp]),
        "spfft_csc_bluestein_mixed_radix": (_i, [_i, _i, _i, _vp, _vp, _i, ctypes.POINTER(Result)]),
        "spfft_csc_bluestein_cufft_smooth": (_i, [_i, _i, _i, _vp, _vp, _i, ctypes.POINTER(Result)]),
        "spfft_run_host_coo": (_i, [_i, _i, _i, _vp, _vp, ctypes.POINTER(Opts),
                                    ctypes.POINTER(Result)]),
        "spfft_free_output": (_i, [_vp]),
        "spfft_profile_enable": (_i, [_vp, _i]),
        "spfft_graph_capture": (_i, [_vp, _vp, _vp, _i, _vp, ctypes.c_int64, _vp, ctypes.POINTER(_vp)]),
        "spfft_graph_launch": (_i, [_vp, _vp]),
        "spfft_graph_destroy": (_i, [_vp]),
        "spfft_csc_bluestein_mixed_radix_graph": (_i, [_i, _i, _i, _vp, _vp, _i, ctypes.POINTER(Result)]),
        "spfft_batch_create": (_i, [_i, _vp, _vp, _vp, _vp, _i, ctypes.POINTER(_vp)]),
        "spfft_batch_destroy": (_i, [_vp]),
        "spfft_batch_layout": (_i, [_vp, _vp, _vp, _vp, _vp]),
        "spfft_batch_execute": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "spfft_batch_capture": (_i, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
        "spfft_batch_check_input": (_i, [_vp, _vp]),
        "spfft_batch_set_mode": (_i, [_vp, _i]),
        "spfft_batch_get_mode": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
        "spfft_profile_read": (_i, [_vp, _vp, ctypes.POINTER(Profile)]),
        "spfft_next_3_smooth": (_i, [_i]),
        "spfft_next_7_smooth": (_i, [_i]),
        "spfft_factor_3_smooth": (_i, [_i, _vp, _i]),
        "spfft_fused_lengths": (_i, [_vp, _i]),
        "spfft_direct_lengths": (_i, [_vp, _i]),
        "spfft_rader_widths": (_i, [_vp, _i]),
        "spfft_coo_generate": (_i, [ctypes.c_char_p, _i, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                    ctypes.POINTER(HostCoo)]),
        "spfft_coo_free": (None, [ctypes.POINTER(HostCoo)]),
        "spfft_read_mtx": (_i, [ctypes.c_char_p, ctypes.POINTER(HostCoo)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(status: int):
    if status != 0:
        msg = load().spfft_last_error().decode()
        raise SpfftError(status, f"spfft {STATUS.get(status, status)}: {msg}")


def next_3_smooth(n: int) -> int:
    return load().spfft_next_3_smooth(int(n))


def next_7_smooth(n: int) -> int:
    return load().spfft_next_7_smooth(int(n))


def factor_3_smooth(n: int) -> list[int]:
    buf = (ctypes.c_int * 128)()
    k = load().spfft_factor_3_smooth(int(n), ctypes.cast(buf, _vp), 128)
    if k < 0:
        raise RuntimeError("factor_3_smooth: not 3-smooth")
    return list(buf[:k])


def fused_lengths() -> list[int]:
    """Bluestein lengths M with a fused row-kernel shape (row_kernel=ROW_FUSED)."""
    buf = (ctypes.c_int * 256)()
    k = load().spfft_fused_lengths(ctypes.cast(buf, _vp), 256)
    return list(buf[:min(k, 256)])


def direct_lengths() -> list[int]:
    """Widths cols with a direct length-cols transform (row_transform=TRANSFORM_DIRECT)."""
    buf = (ctypes.c_int * 256)()
    k = load().spfft_direct_lengths(ctypes.cast(buf, _vp), 256)
    return list(buf[:min(k, 256)])


def rader_widths() -> list[int]:
    """Widths with a Rader row transform (row_transform=TRANSFORM_RADER): cols = 2p or
    cols = p, p an odd prime with a kernel shape for p - 1."""
    k = load().spfft_rader_widths(None, 0)
    buf = (ctypes.c_int * max(k, 1))()
    k = load().spfft_rader_widths(ctypes.cast(buf, _vp), k)
    return list(buf[:k])


# ---------------------------------------------------------------------------
# COO input (the reference's COOMatrix, include/mtx_reader.h:7-13)
# ---------------------------------------------------------------------------
@dataclass
class COOMatrix:
    rows: int
    cols: int
    row_idx: np.ndarray
    col_idx: np.ndarray

    @property
    def nnz(self) -> int:
        return int(len(self.row_idx))


def _from_host_coo(h: HostCoo) -> COOMatrix:
    n = h.nnz
    r = np.ctypeslib.as_array(h.row_idx, shape=(max(n, 1),))[:n].astype(np.int32).copy()
    c = np.ctypeslib.as_array(h.col_idx, shape=(max(n, 1),))[:n].astype(np.int32).copy()
    return COOMatrix(h.rows, h.cols, r, c)


def generate(kind: str, n: int, p0: float, p1: float = 0.0, p2: float = 0.0, p3: float = 0.0,
             seed: int = 0) -> COOMatrix:
    """Seeded synthetic binary symmetric matrix (see spfft_coo_generate)."""
    L = load()
    h = HostCoo()
    st = L.spfft_coo_generate(kind.encode(), int(n), float(p0), float(p1), float(p2),
                              float(p3), int(seed), ctypes.byref(h))
    try:
        if st != 0:
            raise SpfftError(st, f"generator {kind} failed")
        return _from_host_coo(h)
    finally:
        L.spfft_coo_free(ctypes.byref(h))


def read_mtx(path: str) -> COOMatrix:
    L = load()
    h = HostCoo()
    st = L.spfft_read_mtx(path.encode(), ctypes.byref(h))
    try:
        if st != 0:
            raise RuntimeError(f"Cannot read Matrix Market file: {path}")
        return _from_host_coo(h)
    finally:
        L.spfft_coo_free(ctypes.byref(h))


# Named workloads (sizes fixed by the benchmark configurations).
CONFIGS = {
    "c1": ("uniform", 2310, 8000, 0, 0, 0),
    "c2": ("banded", 8219, 242700, 300, 1200, 0.02),
    "c3": ("chunglu", 16706, 121250, 2.1, 360, 0),
    "c5": ("chunglu", 40421, 351400, 2.2, 0, 0),
    # the shape of the reference's delaunay_n14 benchmark (README.md:659-660), synthetic
    "p14": ("uniform", 16384, 49122, 0, 0, 0),
}

CONFIG_TEXT = {
    "c1": "uniform symmetric 2310x2310, 8000 random i<j pairs mirrored (16000 nnz), seed 0",
    "c2": "banded symmetric 8219x8219, p ~ exp(-|i-j|/300), cutoff 1200, +2% off-band, ~242700 nnz, seed 0",
    "c3": "Chung-Lu power-law 16706x16706, exponent 2.1, max degree 360, ~121250 nnz, seed 0",
    "c5": "Chung-Lu power-law 40421x40421, exponent 2.2, ~351400 nnz, seed 0",
    "p14": "uniform symmetric 16384x16384, 49122 random i<j pairs mirrored (98244 nnz), seed 0 "
           "(the size of delaunay_n14; a power-of-two width)",
}


def config_matrix(name: str, seed: int = 0) -> COOMatrix:
    kind, n, p0, p1, p2, p3 = CONFIGS[name]
    return generate(kind, n, p0, p1, p2, p3, seed)


# ---------------------------------------------------------------------------
# Plans and device execution (torch supplies device memory and the stream)
# ---------------------------------------------------------------------------
def _stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _opts(length_rule=SMOOTH_3, u_tile=0, fft_len=0, m1=0, conv_rows=0, precision=PREC_FP32,
          colpass_rows=0, colpass_table=0, row_limit=0, row_kernel=ROW_AUTO, colpass_pair=PAIR_AUTO,
          row_transform=TRANSFORM_AUTO, row_persist=0, t_rows=0, row_threads=0):
    o = Opts()
    load().spfft_opts_default(ctypes.byref(o))
    o.length_rule, o.u_tile, o.precision = length_rule, u_tile, precision
    o.fft_len, o.m1, o.conv_rows = fft_len, m1, conv_rows
    o.colpass_rows, o.colpass_table, o.row_limit = colpass_rows, colpass_table, row_limit
    o.row_kernel = row_kernel
    o.colpass_pair = colpass_pair
    o.row_transform = row_transform
    o.row_persist = row_persist
    o.t_rows = t_rows
    o.row_threads = row_threads
    return o


def _info_dict(inf: PlanInfo) -> dict:
    d = {k: getattr(inf, k) for k, _ in PlanInfo._fields_ if k != "struct_size"}
    d["conv_radix"] = list(inf.conv_radix)[: inf.conv_stages]
    return d


def describe_plan(rows: int, cols: int, **kw) -> dict:
    """Host-only shape plan (M, M1 x M2 split, convolution kernel) -- no GPU needed."""
    inf = PlanInfo()
    _check(load().spfft_plan_describe(int(rows), int(cols), ctypes.byref(_opts(**kw)),
                                      ctypes.byref(inf)))
    return _info_dict(inf)


class Plan:
    """A planned transform for one (rows, cols) shape.  Mirrors the set-up half
    of the reference drivers (src/sparse_fft_bluestein.cu:2256-2320)."""

    def __init__(self, rows: int, cols: int, length_rule: int = SMOOTH_3, u_tile: int = 0,
                 fft_len: int = 0, m1: int = 0, conv_rows: int = 0, precision: int = PREC_FP32,
                 colpass_rows: int = 0, colpass_table: int = 0, row_limit: int = 0,
                 row_kernel: int = ROW_AUTO, colpass_pair: int = PAIR_AUTO,
                 row_transform: int = TRANSFORM_AUTO, row_persist: int = 0, t_rows: int = 0,
                 row_threads: int = 0):
        L = load()
        o = _opts(length_rule, u_tile, fft_len, m1, conv_rows, precision, colpass_rows,
                  colpass_table, row_limit, row_kernel, colpass_pair, row_transform, row_persist,
                  t_rows, row_threads)
        self._h = _vp()
        _check(L.spfft_plan_create(int(rows), int(cols), ctypes.byref(o), ctypes.byref(self._h)))
        inf = PlanInfo()
        _check(L.spfft_plan_get_info(self._h, ctypes.byref(inf)))
        self.info = _info_dict(inf)
        self.rows, self.cols = rows, cols

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.spfft_plan_destroy(h)
            self._h = None

    @property
    def pitch(self) -> int:
        return self.info["pitch"]

    def alloc_output(self):
        import torch
        return torch.empty((self.rows, self.pitch), dtype=torch.complex64, device="cuda")

    def execute(self, d_row, d_col, d_out, stream=None):
        """All three stages on device COO (int32 torch tensors) into d_out [rows, pitch]."""
        self._check_io(d_row, d_col, d_out)
        _check(load().spfft_execute(self._h, _vp(d_row.data_ptr()), _vp(d_col.data_ptr()),
                                    int(d_row.numel()), _vp(d_out.data_ptr()), int(d_out.shape[1]),
                                    _stream_handle(stream)))

    def alloc_host_output(self):
        """Pinned host memory [rows, pitch] complex64 for execute_to_host."""
        import torch
        return torch.empty((self.rows, self.pitch), dtype=torch.complex64, pin_memory=True)

    def execute_to_host(self, d_row, d_col, h_out, stream=None):
        """All three stages on device COO with the output streamed to host memory h_out
        (a contiguous complex64 CPU tensor [rows, >= cols/2+1], pinned for overlap):
        two device chunks of 2 * u_tile rows, copied on the plan's copy stream while the
        next tile computes (the reference's streaming drivers,
        src/sparse_fft_bluestein.cu:912-1145).  Asynchronous on the stream: synchronise
        before reading h_out."""
        self._check_coo(d_row, d_col)
        self._check_host_out(h_out)
        _check(load().spfft_execute_to_host(self._h, _vp(d_row.data_ptr()), _vp(d_col.data_ptr()),
                                            int(d_row.numel()), _vp(h_out.data_ptr()),
                                            int(h_out.shape[1]), _stream_handle(stream)))

    def rowpass_to_host(self, h_out, stream=None):
        """Stages 2 + 3 of execute_to_host on the CSC of the last csc_build."""
        self._check_host_out(h_out)
        _check(load().spfft_rowpass_to_host(self._h, _vp(h_out.data_ptr()), int(h_out.shape[1]),
                                            _stream_handle(stream)))

    def csc_build(self, d_row, d_col, stream=None):
        self._check_coo(d_row, d_col)
        _check(load().spfft_csc_build(self._h, _vp(d_row.data_ptr()), _vp(d_col.data_ptr()),
                                      int(d_row.numel()), _stream_handle(stream)))

    def rowpass(self, d_out, stream=None):
        _check(load().spfft_rowpass(self._h, _vp(d_out.data_ptr()), int(d_out.shape[1]),
                                    _stream_handle(stream)))

    def check_input(self, stream=None):
        _check(load().spfft_check_input(self._h, _stream_handle(stream)))

    def profile(self, on: bool = True):
        """Start (or stop) per-kernel event timing on the launch stream."""
        _check(load().spfft_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, stream=None) -> dict:
        """{stage: (total_ms, launches)} since the last read."""
        pr = Profile()
        _check(load().spfft_profile_read(self._h, _stream_handle(stream), ctypes.byref(pr)))
        return {name: (pr.ms[i], pr.launches[i]) for i, name in enumerate(PROFILE_STAGES)}

    def csc_get(self, stream=None) -> dict:
        """Copies of the device CSC arrays as numpy (sparse_cols, col_ptr, row_idx, col_start)."""
        import torch
        v = CscView()
        _check(load().spfft_csc_get(self._h, _stream_handle(stream), ctypes.byref(v)))
        it = np.uint16 if v.packed else np.int32
        return {
            "n_sparse_cols": v.n_sparse_cols,
            "packed": bool(v.packed),
            "sparse_cols": _d2h(v.d_sparse_cols, v.n_sparse_cols, it),
            "col_ptr": _d2h(v.d_col_ptr, v.n_sparse_cols + 1, np.int32),
            "row_idx": _d2h(v.d_row_idx, v.nnz, it),
            "col_start": _d2h(v.d_col_start, self.cols + 1, np.int32),
        }

    def _check_coo(self, d_row, d_col):
        import torch
        for t in (d_row, d_col):
            if t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("COO arrays must be contiguous int32 CUDA tensors")
        if d_row.numel() != d_col.numel():
            raise ValueError("row and col arrays differ in length")

    def _check_host_out(self, h_out):
        import torch
        if (h_out.dtype != torch.complex64 or h_out.is_cuda or h_out.dim() != 2
                or h_out.shape[0] != self.rows or h_out.shape[1] < self.info["out_cols"]
                or not h_out.is_contiguous()):
            raise ValueError("host output must be a contiguous complex64 CPU tensor [rows, >= cols/2+1]")

    def _check_io(self, d_row, d_col, d_out):
        import torch
        self._check_coo(d_row, d_col)
        if (d_out.dtype != torch.complex64 or not d_out.is_cuda or d_out.dim() != 2
                or d_out.shape[0] != self.rows or d_out.shape[1] < self.info["out_cols"]
                or not d_out.is_contiguous()):
            raise ValueError("output must be a contiguous complex64 CUDA tensor [rows, >= cols/2+1]")


class Graph:
    """A captured hipGraph of a whole transform (or a whole batch); launch() replays it."""

    def __init__(self, handle, keep=()):
        self._h = handle
        self._keep = keep  # device buffers the graph reads or writes

    def launch(self, stream=None):
        _check(load().spfft_graph_launch(self._h, _stream_handle(stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.spfft_graph_destroy(h)
            self._h = None


def capture_plan(plan: "Plan", d_row, d_col, d_out) -> Graph:
    """Capture CSC build + row pass of `plan` on these device buffers into one graph."""
    plan._check_io(d_row, d_col, d_out)
    h = _vp()
    _check(load().spfft_graph_capture(plan._h, _vp(d_row.data_ptr()), _vp(d_col.data_ptr()),
                                      int(d_row.numel()), _vp(d_out.data_ptr()), int(d_out.shape[1]),
                                      None, ctypes.byref(h)))
    return Graph(h, (plan, d_row, d_col, d_out))


class Batch:
    """Many matrices of uneven shape run as one launch sequence on `streams`
    worker streams (the batched form); each matrix has its own plan."""

    def __init__(self, shapes, opts=None, nnz_hint=None, streams: int = 4):
        L = load()
        n = len(shapes)
        self.count = n
        self.rows = np.ascontiguousarray([s[0] for s in shapes], np.int32)
        self.cols = np.ascontiguousarray([s[1] for s in shapes], np.int32)
        hint = None if nnz_hint is None else np.ascontiguousarray(nnz_hint, np.int32)
        oarr = None
        if opts is not None:
            oarr = (Opts * n)()
            for i, kw in enumerate(opts):
                oarr[i] = _opts(**kw)
        self._h = _vp()
        _check(L.spfft_batch_create(n, self.rows.ctypes.data, self.cols.ctypes.data,
                                    None if hint is None else hint.ctypes.data,
                                    None if oarr is None else ctypes.cast(oarr, _vp),
                                    int(streams), ctypes.byref(self._h)))
        total = ctypes.c_int64()
        self.offsets = np.zeros(n, np.int64)
        self.pitches = np.zeros(n, np.int32)
        infos = (PlanInfo * n)()
        for k in range(n):
            infos[k].struct_size = ctypes.sizeof(PlanInfo)
        _check(L.spfft_batch_layout(self._h, ctypes.byref(total), self.offsets.ctypes.data,
                                    self.pitches.ctypes.data, ctypes.cast(infos, _vp)))
        self.total = int(total.value)
        self.infos = [_info_dict(infos[i]) for i in range(n)]

    def alloc_output(self):
        import torch
        return torch.empty(self.total, dtype=torch.complex64, device="cuda")

    def output_of(self, arena, i: int):
        """View of matrix i's [rows, pitch] output inside the arena."""
        r, p = int(self.rows[i]), int(self.pitches[i])
        o = int(self.offsets[i])
        return arena[o: o + r * p].view(r, p)

    @staticmethod
    def pack(mats):
        """Concatenate COO pairs: (d_row, d_col, coo_off) ready for execute()."""
        import torch
        off = np.zeros(len(mats) + 1, np.int64)
        for i, m in enumerate(mats):
            off[i + 1] = off[i] + m.nnz
        r = np.concatenate([m.row_idx for m in mats]).astype(np.int32) if off[-1] else np.zeros(1, np.int32)
        c = np.concatenate([m.col_idx for m in mats]).astype(np.int32) if off[-1] else np.zeros(1, np.int32)
        return torch.as_tensor(r).cuda(), torch.as_tensor(c).cuda(), off

    def execute(self, d_row, d_col, coo_off, arena, stream=None):
        off = np.ascontiguousarray(coo_off, np.int64)
        _check(load().spfft_batch_execute(self._h, _vp(d_row.data_ptr()), _vp(d_col.data_ptr()),
                                          off.ctypes.data, _vp(arena.data_ptr()),
                                          _stream_handle(stream)))

    def capture(self, d_row, d_col, coo_off, arena) -> Graph:
        off = np.ascontiguousarray(coo_off, np.int64)
        h = _vp()
        _check(load().spfft_batch_capture(self._h, _vp(d_row.data_ptr()), _vp(d_col.data_ptr()),
                                          off.ctypes.data, _vp(arena.data_ptr()), ctypes.byref(h)))
        return Graph(h, (self, d_row, d_col, arena, off))

    def check_input(self, stream=None):
        _check(load().spfft_batch_check_input(self._h, _stream_handle(stream)))

    def set_mode(self, mode: int):
        """BATCH_AUTO: the members (one-tile on-chip plans) run as shared launches driven
        by a device table of their arguments; BATCH_PER_MATRIX: every matrix on its own
        launches.  Graphs captured before a mode change must not be replayed after it."""
        _check(load().spfft_batch_set_mode(self._h, int(mode)))

    def mode(self):
        """(mode, members): the mode and how many matrices run through the shared launches."""
        m, n = ctypes.c_int(), ctypes.c_int()
        _check(load().spfft_batch_get_mode(self._h, ctypes.byref(m), ctypes.byref(n)))
        return m.value, n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.spfft_batch_destroy(h)
            self._h = None


_hip = None


def _hip_lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    return _hip


def _d2h(ptr, n: int, dtype) -> np.ndarray:
    out = np.empty(max(n, 0), dtype=dtype)
    if n > 0:
        rc = _hip_lib().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # DeviceToHost
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out


def device_ptr_to_host(ptr, rows: int, pitch: int) -> np.ndarray:
    """Copy a [rows, pitch] complex64 device array (raw pointer) to host."""
    return _d2h(ptr, rows * pitch * 2, np.float32).view(np.complex64).reshape(rows, pitch)


@dataclass
class SparseFFTResult:
    """Mirror of the reference's SparseFFTResult (include/sparse_fft.h:7-13);
    d_output is a torch tensor [rows, output_cols] (pitch P), or None for the
    streamed variants, whose output is h_output (pinned host tensor, same shape)."""
    d_output: object
    output_cols: int
    ms: float
    mem_bytes: int
    ms_e2e: float = 0.0
    h_output: object = None


def reference_opts(length_rule: int, u_tile: int) -> dict:
    """Plan options of the reference-shaped entry points: the reference's algorithm,
    the Bluestein row pass at the rule's smooth length, u_tile rows per tile."""
    return dict(length_rule=length_rule, u_tile=u_tile, row_transform=TRANSFORM_BLUESTEIN)


def _run_reference_shaped(coo: COOMatrix, u_tile: int, length_rule: int) -> SparseFFTResult:
    import torch
    if u_tile <= 0:
        raise RuntimeError("CSC tile size must be positive")
    free0, _ = torch.cuda.mem_get_info()
    plan = Plan(coo.rows, coo.cols, **reference_opts(length_rule, u_tile))
    d_row = torch.as_tensor(np.ascontiguousarray(coo.row_idx, np.int32)).cuda()
    d_col = torch.as_tensor(np.ascontiguousarray(coo.col_idx, np.int32)).cuda()
    out = plan.alloc_output()
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    plan.csc_build(d_row, d_col)
    e1.record()
    plan.rowpass(out)
    e2.record()
    e2.synchronize()
    plan.check_input()
    free1, _ = torch.cuda.mem_get_info()
    return SparseFFTResult(out, plan.pitch, e1.elapsed_time(e2), max(free0 - free1, 0),
                           e0.elapsed_time(e2))


def sparse_fft_csc_bluestein_mixed_radix(coo: COOMatrix, u_tile: int = 128) -> SparseFFTResult:
    """Bluestein at next_3_smooth(2*cols-1); reference src/sparse_fft_bluestein.cu:2256-2392."""
    return _run_reference_shaped(coo, u_tile, SMOOTH_3)


def sparse_fft_csc_bluestein_cufft_smooth(coo: COOMatrix, u_tile: int = 128) -> SparseFFTResult:
    """Bluestein at next_7_smooth(2*cols-1); reference src/sparse_fft_bluestein.cu:2557-2719."""
    return _run_reference_shaped(coo, u_tile, SMOOTH_7)


def sparse_fft_csc_bluestein_mixed_radix_graph(coo: COOMatrix, u_tile: int = 128) -> SparseFFTResult:
    """The mixed-radix transform with the CSC build and the tile loop captured
    into one hipGraph and launched once; ms is the GPU time of that launch.
    Reference src/sparse_fft_bluestein.cu:2402-2549."""
    import torch
    if u_tile <= 0:
        raise RuntimeError("CSC tile size must be positive")
    free0, _ = torch.cuda.mem_get_info()
    plan = Plan(coo.rows, coo.cols, **reference_opts(SMOOTH_3, u_tile))
    d_row = torch.as_tensor(np.ascontiguousarray(coo.row_idx, np.int32)).cuda()
    d_col = torch.as_tensor(np.ascontiguousarray(coo.col_idx, np.int32)).cuda()
    out = plan.alloc_output()
    torch.cuda.synchronize()
    graph = capture_plan(plan, d_row, d_col, out)
    e0, e1 = (torch.cuda.Event(enable_timing=True) for _ in range(2))
    e0.record()
    graph.launch()
    e1.record()
    e1.synchronize()
    plan.check_input()
    free1, _ = torch.cuda.mem_get_info()
    ms = e0.elapsed_time(e1)
    return SparseFFTResult(out, plan.pitch, ms, max(free0 - free1, 0), ms)


def sparse_fft_csc_stockham_streaming(coo: COOMatrix, u_tile: int = 128,
                                      fft_backend: int = 0) -> SparseFFTResult:
    """The transform with the output streamed to pinned host memory: two device
    chunks of 2 * u_tile rows; ms covers the tile loop with its device-to-host
    copies.  fft_backend 3: the 7-smooth length (the reference's cuFFT backend),
    else the 3-smooth one.  Reference include/sparse_fft.h:35-37,
    src/sparse_fft_bluestein.cu:912-1149."""
    import torch
    if u_tile <= 0:
        raise RuntimeError("CSC tile size must be positive")
    free0, _ = torch.cuda.mem_get_info()
    plan = Plan(coo.rows, coo.cols, **reference_opts(SMOOTH_7 if fft_backend == 3 else SMOOTH_3, u_tile))
    d_row = torch.as_tensor(np.ascontiguousarray(coo.row_idx, np.int32)).cuda()
    d_col = torch.as_tensor(np.ascontiguousarray(coo.col_idx, np.int32)).cuda()
    host = plan.alloc_host_output()
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    plan.csc_build(d_row, d_col)
    e1.record()
    plan.rowpass_to_host(host)
    e2.record()
    e2.synchronize()
    plan.check_input()
    free1, _ = torch.cuda.mem_get_info()
    return SparseFFTResult(None, plan.pitch, e1.elapsed_time(e2), max(free0 - free1, 0),
                           e0.elapsed_time(e2), host)


def sparse_fft_csc_bluestein_cufft_streaming(coo: COOMatrix, u_tile: int = 128) -> SparseFFTResult:
    """sparse_fft_csc_stockham_streaming at the 7-smooth length (reference
    include/sparse_fft.h:37)."""
    return sparse_fft_csc_stockham_streaming(coo, u_tile, 3)
