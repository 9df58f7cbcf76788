"""The C ABI on the CPU: the library loads, exports every entry point include/ffstencil.h declares, and
its host-only functions (named stencils, seeded fill, CSV line, argument validation) behave as
documented.  No compute calls: plan creation is only exercised on paths that fail validation, which
happens before any HIP call.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from tests import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ffstencil.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\*?\s+\*?(ffs_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for expected in ("ffs_plan_create", "ffs_run", "ffs_run_host", "ffs_run_block", "ffs_plan_destroy",
                     "ffs_plan_get_info", "ffs_stencil_named", "ffs_fill_uniform", "ffs_format_csv",
                     "ffs_last_error", "ffs_version", "ffs_problem_init"):
        assert expected in names


def test_library_exports_every_declared_symbol(ffs):
    lib = ffs.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ffs.library_path()], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(ffs_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layouts_match_header(ffs):
    # ffs_problem: 2*i32, 3*i64, ptr, i64, 8*i32 (algorithm, mapping last), flags, reserved
    assert ctypes.sizeof(ffs.Problem) == 4 * 2 + 8 * 3 + 8 + 8 + 4 * 8 + 4 * 2
    lib = ffs.load_library()
    p = ffs.Problem()
    assert lib.ffs_problem_init(ctypes.byref(p)) == 0
    assert (p.dims, p.radius, p.steps, p.precision, p.convention, p.use_graph) == (2, 1, 1, 0, 0, 1)
    assert list(p.sizes) == [1, 1, 1]


def test_version(ffs):
    assert re.match(r"^\d+\.\d+\.\d+$", ffs.version())


NAMED = {
    "Heat-1D": (1, 1, "1d1r", False), "1D5P": (1, 2, "1d2r", False), "1D7P": (1, 3, "1d3r", False),
    "Heat-2D": (2, 1, "star_2d1r", True), "Box2D9P": (2, 1, "box_2d1r", False),
    "Star2D9P": (2, 2, "star_2d2r", True), "Box2D25P": (2, 2, "box_2d2r", False),
    "Heat-3D": (3, 1, "star_3d1r", True), "Box3D27P": (3, 1, "box_3d1r", False),
}


@pytest.mark.parametrize("shape", sorted(NAMED))
def test_named_stencils(ffs, shape):
    dims, radius, tag, star = NAMED[shape]
    d, r, w, t = ffs.named_stencil(shape, seed=123)
    assert (d, r, t) == (dims, radius, tag)
    K = 2 * r + 1
    assert w.size == K ** d
    assert abs(w.sum() - 1.0) < 1e-15
    W = w.reshape((K,) * d)
    for idx in np.ndindex(*W.shape):
        off_axis = sum(1 for i in idx if i != r)
        if star and off_axis > 1:
            assert W[idx] == 0.0
        else:
            assert W[idx] > 0.0
    if shape == "Heat-1D":
        assert list(w) == [0.25, 0.5, 0.25]
    else:  # seeded: same seed same weights, other seed other weights
        assert np.array_equal(w, ffs.named_stencil(shape, seed=123)[2])
        assert not np.array_equal(w, ffs.named_stencil(shape, seed=124)[2])


def test_named_stencil_alias_and_errors(ffs):
    assert ffs.named_stencil("Box3d27P")[3] == "box_3d1r"
    with pytest.raises(ffs.FFSError) as e:
        ffs.named_stencil("Box4D81P")
    assert e.value.status == 1 and "Box4D81P" in str(e.value)


def test_fill_uniform_matches_independent_splitmix(ffs):
    for seed in (0, 1, 0x5EED0002):
        assert np.array_equal(ffs.fill_uniform(seed, 1000), O.uniform(seed, (1000,)))


def test_format_csv(ffs):
    assert ffs.format_csv("box_2d1r", [8192, 8192], 128, 2.0) == \
        "FlashFFTStencil, box_2d1r, 1, 8192, 8192, 128, 2, 4294.97"
    assert ffs.format_csv("1d1r", [1048576], 64, 0.5).startswith("FlashFFTStencil, 1d1r, 1, 1, 1048576, 64, 0.5, ")
    assert ffs.format_csv("box_3d1r", [512, 512, 512], 32, 10.0).startswith("FlashFFTStencil, box_3d1r, 512, 512, 512, 32, 10, ")


def _status(ffs, **kw):
    lib = ffs.load_library()
    p = ffs.Problem()
    lib.ffs_problem_init(ctypes.byref(p))
    w = kw.pop("weights", np.full(9, 1 / 9.0))
    for k, v in kw.items():
        if k == "sizes":
            for i, s in enumerate(v):
                p.sizes[i] = s
        else:
            setattr(p, k, v)
    p.weights = w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if w is not None else None
    h = ctypes.c_void_p()
    st = lib.ffs_plan_create(ctypes.byref(p), ctypes.byref(h))
    return st, lib.ffs_last_error().decode()


@pytest.mark.parametrize("kw,needle", [
    (dict(dims=0), "dims"), (dict(dims=4), "dims"), (dict(radius=-1), "radius"), (dict(steps=-3), "steps"),
    (dict(sizes=[0, 5]), "size"), (dict(weights=None), "weights"), (dict(precision=9), "precision"),
    (dict(convention=5), "convention"), (dict(fuse=-1), "fuse"), (dict(algorithm=3), "algorithm"),
    (dict(algorithm=-1), "algorithm"), (dict(mapping=3), "mapping"),
    (dict(weights=np.array([1.0, np.nan] + [0.0] * 7)), "finite"),
])
def test_plan_create_rejects_invalid(ffs, kw, needle):
    st, msg = _status(ffs, **{"sizes": [64, 64], **kw})
    assert st == 1, (st, msg)
    assert needle in msg


@pytest.mark.parametrize("kw,needle", [
    (dict(tile=24), "tile"), (dict(tile=32, precision=2), "tile"), (dict(fuse=16), "fuse"),
    (dict(fuse=8, tile=16), "fuse"),
    (dict(radius=16, weights=np.full(33 * 33, 1 / 1089.0)), "radius"),
    (dict(radius=8, tile=16, weights=np.full(17 * 17, 1 / 289.0)), "radius"),
    (dict(precision=4, tile=16), "tile"), (dict(precision=4, dims=1, sizes=[640]), "FP16"),
    (dict(precision=4, dims=3, sizes=[16, 16, 16], weights=np.full(27, 1 / 27.0)), "FP16"),
    (dict(algorithm=1, precision=1), "FP64"),
    (dict(algorithm=1, radius=2, fuse=5, weights=np.full(25, 1 / 25.0)), "fuse"),
    (dict(algorithm=1, dims=3, radius=2, sizes=[16, 16, 16], weights=np.full(125, 1 / 125.0)), "radius"),
])
def test_plan_create_unsupported(ffs, kw, needle):
    st, msg = _status(ffs, **{"sizes": [64, 64], **kw})
    assert st == 2, (st, msg)
    assert needle in msg


def test_null_arguments(ffs):
    lib = ffs.load_library()
    assert lib.ffs_run(None, None, None, None) == 1
    assert lib.ffs_run_block(None, None, None, 0, None) == 1
    assert lib.ffs_plan_get_info(None, None) == 1
    assert lib.ffs_plan_destroy(None) == 0
    assert lib.ffs_plan_create(None, None) == 1


def test_c_struct_layout_matches_binding(ffs, tmp_path):
    # sizes and field offsets of the C structs as the C compiler lays them out, against the ctypes mirror
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "ffstencil.h"\n'
        'int main(void) {\n'
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %d\\n", sizeof(ffs_problem), offsetof(ffs_problem, algorithm),\n'
        '         offsetof(ffs_problem, weights), sizeof(ffs_plan_info), offsetof(ffs_plan_info, algorithm),\n'
        '         offsetof(ffs_plan_info, flops_per_block), offsetof(ffs_problem, mapping),\n'
        '         offsetof(ffs_plan_info, mapping), offsetof(ffs_problem, flags), FFS_FLAG_FULL_STAGES);\n'
        '  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(ffs.Problem), ffs.Problem.algorithm.offset, ffs.Problem.weights.offset,
            ctypes.sizeof(ffs.PlanInfo), ffs.PlanInfo.algorithm.offset, ffs.PlanInfo.flops_per_block.offset,
            ffs.Problem.mapping.offset, ffs.PlanInfo.mapping.offset, ffs.Problem.flags.offset,
            ffs.FLAG_FULL_STAGES]
    assert got == want


def test_plan_create_rejects_unknown_flags(ffs):
    st, msg = _status(ffs, sizes=[64, 64], flags=0x6)
    assert st == 1 and "flags" in msg


def test_bf16x2_withdrawn(ffs):
    st, msg = _status(ffs, sizes=[64, 64], precision=3)
    assert st == 2 and "withdrawn" in msg


class _FakeTensor:
    """Stands in for a CUDA float64 tensor: what the binding inspects before passing a pointer on."""

    def __init__(self, n, device=0, ptr=0x1000):
        self._n, self._ptr = n, ptr
        self.is_cuda = True
        self.dtype = "torch.float64"
        self.device = type("dev", (), {"index": device})()

    def is_contiguous(self):
        return True

    def numel(self):
        return self._n

    def data_ptr(self):
        return self._ptr


def test_binding_checks_tensor_size_and_device(ffs):
    # an undersized tensor or one on another device would let the kernels read or write out of bounds
    assert ffs._ptr(_FakeTensor(100), 100, 0) == 0x1000
    with pytest.raises(ffs.FFSError, match="elements"):
        ffs._ptr(_FakeTensor(99), 100, 0)
    with pytest.raises(ffs.FFSError, match="cuda:1"):
        ffs._ptr(_FakeTensor(100, device=1), 100, 0)
    assert ffs._ptr(0x2000, 100, 0) == 0x2000  # raw pointers are the caller's responsibility


def test_batch_run_checks_counts(ffs):
    b = object.__new__(ffs.Batch)  # no device: only the argument checks of run() are exercised
    b._h, b.count, b.device, b.sizes = None, 2, 0, [(4, 5), (6, 7)]
    with pytest.raises(ffs.FFSError, match="2 grids"):
        b.run([_FakeTensor(20)], [_FakeTensor(20)])
    with pytest.raises(ffs.FFSError, match="elements"):
        b.run([_FakeTensor(20), _FakeTensor(41)], [_FakeTensor(20), _FakeTensor(42)])
