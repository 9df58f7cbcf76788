"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only, no GPU needed).

Three builds, each from the committed sources into a temporary directory:
  * the oracle (oracle/oracle.c and oracle/sweep_fast.c) with gcc, driven by tests/sanitize/oracle_check.c over 1D/2D/3D ragged
    and tiny grids, both conventions, T = 0, aliasing and rejected arguments;
  * libffstencil with the sanitizers on its host code only (device code cannot take them: on every hipcc
    line each -fsanitize= follows -Xarch_host), driven by tests/sanitize/host_abi_check.cpp through every
    entry point that needs no GPU (validation, named stencils, fill, CSV, NULL handling), with leak
    detection on;
  * the 2d driver linked against that library, run through its CPU-side paths (usage errors, --help,
    unknown shapes, --custom files and stdin, rejected option values).
The reference's error behaviour these paths replace is exit() from checkCudaErrors
(helper_cuda.h:170-182) and `return 1` on a bad shape (src/2D/2d_main.cu:43-47).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
HIP_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer"]


def _need(tool):
    if not (os.path.exists(tool) or shutil.which(tool)):
        pytest.skip(f"{tool} not available")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    return r


def _clean(r):
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "LeakSanitizer" not in text, text[-3000:]
    assert "runtime error:" not in text, text[-3000:]
    assert r.returncode not in (98, 99), text[-3000:]


@pytest.fixture(scope="module")
def san_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("san")


@pytest.fixture(scope="module")
def san_lib(san_dir):
    _need(HIPCC)
    src = os.path.join(ROOT, "flashfftstencil_amd", "csrc")
    out = str(san_dir / "libffstencil_san.so")
    cmd = [HIPCC, "--offload-arch=gfx950", "-mcode-object-version=5", "-mllvm", "-amdgpu-mfma-vgpr-form=1", "-O1", "-g",
           "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *HIP_SAN, "-shared",
           "-o", out, os.path.join(src, "ffs_api.cpp"), os.path.join(src, "ffs_kernels.hip"),
           os.path.join(src, "ffs_direct.hip")]
    r = _run(cmd)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_oracle_under_sanitizers(san_dir):
    _need("gcc")
    exe = str(san_dir / "oracle_check")
    r = _run(["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
              "-fno-sanitize-recover=undefined", "-fopenmp", "-std=c11", "-Wall", "-march=x86-64-v3",
              os.path.join(ROOT, "tests", "sanitize", "oracle_check.c"), os.path.join(ROOT, "oracle", "oracle.c"),
              os.path.join(ROOT, "oracle", "sweep_fast.c"), "-lm", "-o", exe])
    assert r.returncode == 0, r.stderr
    r = _run([exe], env=SAN_ENV)
    _clean(r)
    assert r.returncode == 0 and "oracle checks passed" in r.stdout, r.stdout + r.stderr


def test_host_abi_under_sanitizers(san_dir, san_lib):
    _need(CLANGXX)
    exe = str(san_dir / "host_abi_check")
    r = _run([CLANGXX, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
              "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "sanitize", "host_abi_check.cpp"),
              san_lib, f"-Wl,-rpath,{san_dir}", "-o", exe])
    assert r.returncode == 0, r.stderr
    r = _run([exe], env=SAN_ENV)
    _clean(r)
    assert r.returncode == 0 and "host ABI checks passed" in r.stdout, r.stdout + r.stderr


def test_driver_cli_under_sanitizers(san_dir, san_lib, tmp_path):
    exe = str(san_dir / "2d")
    r = _run([HIPCC, "-O1", "-g", "-std=c++17", *HIP_SAN, "-DFFS_DRIVER_DIMS=2",
              os.path.join(ROOT, "flashfftstencil_amd", "drivers", "ffs_driver.cpp"), f"-L{san_dir}", "-lffstencil_san",
              f"-Wl,-rpath,{san_dir}", "-o", exe])
    assert r.returncode == 0, r.stderr[-3000:]
    wrong_count = tmp_path / "w4.txt"
    wrong_count.write_text("0.1 0.2 0.3 0.4\n")
    not_numbers = tmp_path / "junk.txt"
    not_numbers.write_text("a b c\n")
    cases = [
        ([], None, 1),
        (["--help"], None, 0),
        (["Box2D9P", "64"], None, 1),
        (["NoSuchShape", "64", "4"], None, 1),
        (["Heat-1D", "64", "4"], None, 1),
        (["Box2D9P", "abc", "4"], None, 1),
        (["Box2D9P", "64", "4", "--bogus"], None, 1),
        (["Box2D9P", "64", "4", "--custom"], None, 1),
        (["Box2D9P", "64", "4", "--custom", str(wrong_count)], None, 1),
        (["Box2D9P", "64", "4", "--custom", str(not_numbers)], None, 1),
        (["Box2D9P", "64", "4", "--custom", str(tmp_path / "missing.txt")], None, 1),
        (["Box2D9P", "64", "4", "--custom", "-"], "1 2 3\n", 1),
        (["Box2D9P", "64", "4", "--precision", "fp8"], None, 1),
        (["Box2D9P", "64", "4", "--convention", "left"], None, 1),
        (["Box2D9P", "64", "4", "--algorithm", "fourier"], None, 1),
        (["Box2D9P", "0", "4"], None, 1),
        (["Box2D9P", "64", "-3"], None, 1),
    ]
    for args, stdin, code in cases:
        r = _run([exe, *args], input=stdin, env=SAN_ENV)
        _clean(r)
        assert r.returncode == code, (args, r.returncode, r.stdout[-500:], r.stderr[-500:])
