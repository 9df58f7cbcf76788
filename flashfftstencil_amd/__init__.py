"""flashfftstencil_amd -- Python binding of libffstencil, the MI355X tiled FFT-stencil engine.

The product is the C ABI declared in ``include/ffstencil.h`` and implemented by
``flashfftstencil_amd/libffstencil.so`` (HIP kernels for gfx950 + host plan/run code).  This module is a
thin ctypes layer over it for the test suite and the benchmark; it computes nothing itself and raises
when the library is missing.  Device buffers may be passed as raw pointers or as torch CUDA tensors
(float64, contiguous); torch is only used for device memory and streams.

Mirrors the reference's user-facing pieces (KevinWu2017/FlashFFTStencil):
  * shape table + weights       -> :func:`named_stencil`  (src/1D/1d_main.cu:44-65, src/2D/2d_main.cu:49-52)
  * ``CreatePlan(k, K, print)``  -> :class:`Plan`           (src/2D/create_fft_pfa_plan.cu:17)
  * the launch loop over T       -> :meth:`Plan.run`        (src/2D/2d_main.cu:150-170)
  * the CSV line                 -> :func:`format_csv`      (src/2D/2d_main.cu:185)
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffstencil.so")
BIN_DIR = os.path.join(_HERE, "bin")

FFS_OK = 0
STATUS_NAMES = {
    0: "FFS_OK",
    1: "FFS_ERR_INVALID_ARG",
    2: "FFS_ERR_UNSUPPORTED",
    3: "FFS_ERR_OUT_OF_MEMORY",
    4: "FFS_ERR_HIP",
    5: "FFS_ERR_NO_DEVICE",
    6: "FFS_ERR_INTERNAL",
}
PRECISIONS = {"fp64": 0, "bf16": 1, "mixed": 1, "f32": 2, "bf16x2": 3, "f16": 4}
CONVENTIONS = {"ref": 0, "centred": 1, "centered": 1}
ALGORITHMS = {"fft": 0, "direct": 1, "auto": 2}
MAPPINGS = {"auto": 0, "xcd": 1, "linear": 2}


class FFSError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Problem(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32),
        ("radius", ctypes.c_int32),
        ("sizes", ctypes.c_int64 * 3),
        ("weights", ctypes.POINTER(ctypes.c_double)),
        ("steps", ctypes.c_int64),
        ("precision", ctypes.c_int32),
        ("convention", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("fuse", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("mapping", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


FLAG_FULL_STAGES = 0x1


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32),
        ("radius", ctypes.c_int32),
        ("sizes", ctypes.c_int64 * 3),
        ("steps", ctypes.c_int64),
        ("precision", ctypes.c_int32),
        ("convention", ctypes.c_int32),
        ("tile", ctypes.c_int32 * 3),
        ("valid", ctypes.c_int32 * 3),
        ("fuse", ctypes.c_int32),
        ("fuse_tail", ctypes.c_int32),
        ("blocks", ctypes.c_int64),
        ("tiles", ctypes.c_int64),
        ("launches", ctypes.c_int64),
        ("flops_per_block", ctypes.c_double),
        ("bytes_per_block", ctypes.c_double),
        ("algorithm", ctypes.c_int32),
        ("mapping", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library() -> ctypes.CDLL:
    """Load libffstencil.so (built in-tree by ``__graft_entry__.build()``); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FFSError(5, f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    # torch-ROCm ships its own libamdhip64.so.7; load it first so that this library binds to the same
    # HIP runtime instance (same soname) instead of pulling a second copy into the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sigs = {
        "ffs_problem_init": [P(Problem)],
        "ffs_stencil_named": [ctypes.c_char_p, ctypes.c_uint64, P(ctypes.c_int32), P(ctypes.c_int32),
                              P(ctypes.c_double), ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32],
        "ffs_fill_uniform": [ctypes.c_uint64, P(ctypes.c_double), ctypes.c_int64],
        "ffs_fill_uniform_device": [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p],
        "ffs_plan_create": [P(Problem), P(ctypes.c_void_p)],
        "ffs_plan_get_info": [ctypes.c_void_p, P(PlanInfo)],
        "ffs_plan_destroy": [ctypes.c_void_p],
        "ffs_run": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
        "ffs_run_host": [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double)],
        "ffs_run_block": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p],
        "ffs_run_static": [P(Problem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
        "ffs_format_csv": [ctypes.c_char_p, ctypes.c_int32, P(ctypes.c_int64), ctypes.c_int64, ctypes.c_double,
                           ctypes.c_char_p, ctypes.c_int32],
        "ffs_autotune": [P(Problem), ctypes.c_char_p, P(ctypes.c_double)],
        "ffs_plan_read_table": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                P(ctypes.c_int64)],
        "ffs_batch_create": [P(Problem), ctypes.c_int32, P(ctypes.c_void_p)],
        "ffs_batch_run": [ctypes.c_void_p, P(ctypes.c_void_p), P(ctypes.c_void_p), ctypes.c_void_p],
        "ffs_batch_get_info": [ctypes.c_void_p, ctypes.c_int32, P(PlanInfo)],
        "ffs_batch_destroy": [ctypes.c_void_p],
        "ffs_batch_autotune": [P(Problem), ctypes.c_int32, ctypes.c_char_p, P(ctypes.c_double)],
    }
    for name, argtypes in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.ffs_last_error.restype = ctypes.c_char_p
    lib.ffs_last_error.argtypes = []
    lib.ffs_version.restype = ctypes.c_char_p
    lib.ffs_version.argtypes = []
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != FFS_OK:
        raise FFSError(status, load_library().ffs_last_error().decode())


def version() -> str:
    return load_library().ffs_version().decode()


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def named_stencil(shape: str, seed: int = 0x5EED0000) -> Tuple[int, int, np.ndarray, str]:
    """(dims, radius, weights[(2r+1)^dims], csv tag) of a named stencil."""
    lib = load_library()
    dims, radius = ctypes.c_int32(), ctypes.c_int32()
    w = np.zeros(7 ** 3, dtype=np.float64)
    tag = ctypes.create_string_buffer(32)
    _check(lib.ffs_stencil_named(shape.encode(), ctypes.c_uint64(seed), ctypes.byref(dims), ctypes.byref(radius),
                                 _dptr(w), w.size, tag, 32))
    k = 2 * radius.value + 1
    return dims.value, radius.value, w[: k ** dims.value].copy(), tag.value.decode()


def fill_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float64)
    _check(load_library().ffs_fill_uniform(ctypes.c_uint64(seed), _dptr(out), ctypes.c_int64(out.size)))
    return out


def fill_uniform_device(seed: int, d_out, stream=None) -> None:
    """Write the ffs_fill_uniform sequence into a float64 CUDA tensor (or device pointer with .numel)."""
    n = d_out.numel()
    _check(load_library().ffs_fill_uniform_device(ctypes.c_uint64(seed), ctypes.c_void_p(_ptr(d_out)),
                                                   ctypes.c_int64(n), ctypes.c_void_p(_stream_handle(stream))))


def format_csv(tag: str, sizes: Sequence[int], steps: int, ms: float) -> str:
    buf = ctypes.create_string_buffer(256)
    s = (ctypes.c_int64 * 3)(*([int(x) for x in sizes] + [1] * (3 - len(sizes))))
    _check(load_library().ffs_format_csv(tag.encode(), len(sizes), s, steps, ms, buf, 256))
    return buf.value.decode()


def _ptr(x, numel: Optional[int] = None, device: Optional[int] = None) -> int:
    """Device address of x.  For a torch tensor, also require `numel` elements and the plan's `device`
    (a raw integer pointer cannot be checked: the caller vouches for its size)."""
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):  # torch tensor
        if not x.is_cuda or x.dtype.__str__() != "torch.float64" or not x.is_contiguous():
            raise FFSError(1, "device buffers must be contiguous float64 CUDA tensors")
        if numel is not None and x.numel() != numel:
            raise FFSError(1, f"tensor has {x.numel()} elements, the grid has {numel}")
        if device is not None and x.device.index != device:
            raise FFSError(1, f"tensor is on cuda:{x.device.index}, the plan on device {device}")
        return x.data_ptr()
    raise TypeError(f"expected a device pointer or torch tensor, got {type(x)!r}")


def _problem(dims: int, sizes: Sequence[int], w: np.ndarray, radius: int, steps: int, precision="fp64",
             convention="ref", tile: int = 0, fuse: int = 0, use_graph: bool = True, device: int = 0,
             algorithm="fft", mapping="auto", full_stages: bool = False) -> Problem:
    """An ffs_problem for the given arguments (w must stay alive while the struct is used)."""
    p = Problem()
    _check(load_library().ffs_problem_init(ctypes.byref(p)))
    p.dims = dims
    p.radius = radius
    sz = [int(s) for s in sizes] + [1] * (3 - len(sizes))
    for i in range(3):
        p.sizes[i] = sz[i]
    p.weights = _dptr(w)
    p.steps = int(steps)
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.convention = CONVENTIONS[convention] if isinstance(convention, str) else int(convention)
    p.tile = tile
    p.fuse = fuse
    p.use_graph = 1 if use_graph else 0
    p.device = device
    p.algorithm = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    p.mapping = MAPPINGS[mapping] if isinstance(mapping, str) else int(mapping)
    p.flags = FLAG_FULL_STAGES if full_stages else 0
    return p


class Plan:
    """A planned stencil problem (owns device tables, the ping-pong buffer and the cached hipGraph)."""

    def __init__(self, dims: int, sizes: Sequence[int], weights, radius: int, steps: int,
                 precision: str = "fp64", convention: str = "ref", tile: int = 0, fuse: int = 0,
                 use_graph: bool = True, device: int = 0, algorithm: str = "fft", mapping: str = "auto",
                 full_stages: bool = False):
        lib = load_library()
        self._w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).ravel())
        p = _problem(dims, sizes, self._w, radius, steps, precision, convention, tile, fuse, use_graph, device,
                     algorithm, mapping, full_stages)
        sz = [int(s) for s in sizes]
        self.problem = p
        self.sizes = tuple(sz[:dims])
        self.dims = dims
        self.device = device
        self.numel = int(np.prod(self.sizes))
        h = ctypes.c_void_p()
        _check(lib.ffs_plan_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h

    @property
    def info(self) -> dict:
        inf = PlanInfo()
        _check(load_library().ffs_plan_get_info(self._h, ctypes.byref(inf)))
        return inf.as_dict()

    def read_table(self, which: int) -> bytes:
        """Raw bytes of one of the plan's device tables (0 ops, 1 mixed ops, 2 spectrum, 3 out16 inverse)."""
        lib = load_library()
        n = ctypes.c_int64()
        _check(lib.ffs_plan_read_table(self._h, which, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib.ffs_plan_read_table(self._h, which, buf, n.value, ctypes.byref(n)))
        return buf.raw

    def run(self, d_in, d_out, stream=None) -> None:
        """d_out = S^T(d_in); pointers or float64 CUDA tensors.  stream: int handle, torch stream or None
        (None = torch's current stream when torch is loaded, else the null stream)."""
        pi, po = _ptr(d_in, self.numel, self.device), _ptr(d_out, self.numel, self.device)
        _check(load_library().ffs_run(self._h, ctypes.c_void_p(pi), ctypes.c_void_p(po),
                                      ctypes.c_void_p(_stream_handle(stream))))

    def run_block(self, d_in, d_out, tail: bool = False, stream=None) -> None:
        pi, po = _ptr(d_in, self.numel, self.device), _ptr(d_out, self.numel, self.device)
        _check(load_library().ffs_run_block(self._h, ctypes.c_void_p(pi), ctypes.c_void_p(po),
                                            1 if tail else 0, ctypes.c_void_p(_stream_handle(stream))))

    def run_host(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != int(np.prod(self.sizes)):
            raise FFSError(1, f"host array has {x.size} points, plan expects {self.sizes}")
        out = np.empty_like(x)
        _check(load_library().ffs_run_host(self._h, _dptr(x), _dptr(out)))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().ffs_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


STATIC_SCRATCH_BYTES = 64 << 20  # FFS_STATIC_SCRATCH_BYTES
STATIC_TABLE_BYTES = 1 << 20     # FFS_STATIC_TABLE_BYTES


def run_static(dims: int, sizes: Sequence[int], weights, radius: int, steps: int, d_in, d_out, stream=None,
               **kw) -> None:
    """ffs_run_static: d_out = S^T(d_in) without allocating device memory (plan tables and ping-pong grid
    in the library's static per-GPU array).  Keyword arguments as for :class:`Plan`.  Calls that overlap
    on one GPU race on that array."""
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).ravel())
    p = _problem(dims, sizes, w, radius, steps, **kw)
    n = int(np.prod([int(x) for x in sizes][:dims]))
    dev = kw.get("device", 0)
    pi, po = _ptr(d_in, n, dev), _ptr(d_out, n, dev)
    _check(load_library().ffs_run_static(ctypes.byref(p), ctypes.c_void_p(pi), ctypes.c_void_p(po),
                                         ctypes.c_void_p(_stream_handle(stream))))


def autotune(dims: int, sizes: Sequence[int], weights, radius: int, steps: int, precision: str = "fp64",
             convention: str = "ref", cache: Optional[str] = None, algorithm: str = "fft") -> Tuple[int, int, float]:
    """Measured best (tile, fuse, ms per run) for this problem on device 0 (cached in `cache`).  With
    algorithm "auto" the winner's algorithm is in :func:`autotune_full`."""
    r = autotune_full(dims, sizes, weights, radius, steps, precision, convention, cache, algorithm)
    return r[1], r[2], r[4]


def autotune_full(dims: int, sizes: Sequence[int], weights, radius: int, steps: int, precision: str = "fp64",
                  convention: str = "ref", cache: Optional[str] = None,
                  algorithm: str = "fft") -> Tuple[str, int, int, str, float]:
    """Measured best (algorithm, tile, fuse, mapping, ms per run)."""
    lib = load_library()
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).ravel())
    p = Problem()
    _check(lib.ffs_problem_init(ctypes.byref(p)))
    p.dims, p.radius, p.steps = dims, radius, int(steps)
    for i, s in enumerate(list(sizes) + [1] * (3 - len(sizes))):
        p.sizes[i] = int(s)
    p.weights = _dptr(w)
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.convention = CONVENTIONS[convention] if isinstance(convention, str) else int(convention)
    p.algorithm = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    ms = ctypes.c_double()
    _check(lib.ffs_autotune(ctypes.byref(p), cache.encode() if cache else None, ctypes.byref(ms)))
    name = {v: k for k, v in ALGORITHMS.items()}[p.algorithm]
    mname = {0: "auto", 1: "xcd", 2: "linear"}[p.mapping]
    return name, p.tile, p.fuse, mname, ms.value


def _batch_problems(problems, precision):
    """ctypes array of ffs_problem for a batch (2D, tile 32), the weight arrays it points to, and the sizes."""
    lib = load_library()
    n = len(problems)
    arr = (Problem * n)()
    ws, sizes = [], []
    for i, pr in enumerate(problems):
        p = arr[i]
        _check(lib.ffs_problem_init(ctypes.byref(p)))
        w = np.ascontiguousarray(np.asarray(pr["weights"], dtype=np.float64).ravel())
        ws.append(w)
        p.dims = 2
        p.radius = int(pr["radius"])
        p.sizes[0], p.sizes[1] = int(pr["sizes"][0]), int(pr["sizes"][1])
        p.weights = _dptr(w)
        p.steps = int(pr["steps"])
        prec = pr.get("precision", precision)
        p.precision = PRECISIONS[prec] if isinstance(prec, str) else int(prec)
        conv = pr.get("convention", "ref")
        p.convention = CONVENTIONS[conv] if isinstance(conv, str) else int(conv)
        p.fuse = int(pr.get("fuse", 0))
        mp = pr.get("mapping", "auto")
        p.mapping = MAPPINGS[mp] if isinstance(mp, str) else int(mp)
        sizes.append((int(pr["sizes"][0]), int(pr["sizes"][1])))
    return arr, ws, sizes


class Batch:
    """Independent 2D grids advanced together (one persistent launch per fused block; FP64, FP16 or BF16).

    `problems` is a list of dicts with keys sizes, weights, radius, steps and optionally precision,
    convention, fuse (tile is always 32)."""

    def __init__(self, problems, precision: str = "bf16"):
        lib = load_library()
        arr, self._w, self.sizes = _batch_problems(problems, precision)
        n = len(problems)
        h = ctypes.c_void_p()
        _check(lib.ffs_batch_create(arr, n, ctypes.byref(h)))
        self._h = h
        self.count = n
        self.device = 0

    @staticmethod
    def autotune(problems, precision: str = "bf16", cache: Optional[str] = None):
        """ffs_batch_autotune: copies of `problems` with the tuned fuse (and mapping) of every grid, and the
        tuned batch's ms per run."""
        arr, keep, _ = _batch_problems(problems, precision)
        ms = ctypes.c_double()
        _check(load_library().ffs_batch_autotune(arr, len(problems), cache.encode() if cache else None,
                                                 ctypes.byref(ms)))
        out = []
        for pr, p in zip(problems, arr):
            q = dict(pr)
            q["fuse"] = int(p.fuse)
            q["mapping"] = {v: k for k, v in MAPPINGS.items()}[int(p.mapping)]
            out.append(q)
        del keep
        return out, ms.value

    def info(self, i: int) -> dict:
        inf = PlanInfo()
        _check(load_library().ffs_batch_get_info(self._h, i, ctypes.byref(inf)))
        return inf.as_dict()

    def run(self, d_ins, d_outs, stream=None) -> None:
        n = self.count
        if len(d_ins) != n or len(d_outs) != n:
            raise FFSError(1, f"batch of {n} grids got {len(d_ins)} inputs and {len(d_outs)} outputs")
        ins = (ctypes.c_void_p * n)(*[_ptr(x, s[0] * s[1], self.device) for x, s in zip(d_ins, self.sizes)])
        outs = (ctypes.c_void_p * n)(*[_ptr(x, s[0] * s[1], self.device) for x, s in zip(d_outs, self.sizes)])
        _check(load_library().ffs_batch_run(self._h, ins, outs, ctypes.c_void_p(_stream_handle(stream))))

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().ffs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        import sys
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_initialized():
            return torch.cuda.current_stream().cuda_stream
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream
