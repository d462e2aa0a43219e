"""Builds the native parts of ktt_amd in-tree.

* ``ktt_amd/libktt_hip.so``: the ktt C++ API with the HIP compute engine and the C ABI of
  ``include/ktt_hip.h`` (g++, linked against the ROCm HIP runtime and hipRTC).
* ``ktt_amd/libktt_hip_counters.so``: hardware counters through rocprofiler-sdk, loaded by the
  main library only when counters are enabled (KTT_HIP_COUNTERS=1 or ktt::EnableHipCounters()).
* ``ktt_amd/bin/``: the example programs (CoulombSum2d, CoulombSum3d, CoulombSum3dIterative)
  written against the ktt C++ API.
* every kernel source under ``ktt_amd/kernels`` is compiled once for gfx950 through the
  library's own hipRTC path (``ktt_hip_compile_check``), which needs no GPU.

Run ``python -m ktt_amd.build`` or call :func:`build`.
"""

from __future__ import annotations

import concurrent.futures
import ctypes
import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
REPO = PACKAGE.parent
CSRC = PACKAGE / "csrc"
KERNELS = PACKAGE / "kernels"
OBJ_DIR = PACKAGE / "_build"
LIBRARY = PACKAGE / "libktt_hip.so"
COUNTERS_LIBRARY = PACKAGE / "libktt_hip_counters.so"
NOALLOC_LIBRARY = PACKAGE / "libktt_hip_noalloc.so"
BIN_DIR = PACKAGE / "bin"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"
# Device-code flags of every hipcc build: packed FP32 instructions off (as for hipRTC).
HIP_DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

CXX_FLAGS = [
    "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
    f"-I{CSRC / 'ktt'}", f"-I{CSRC / 'src'}", f"-I{ROCM / 'include'}",
]
LINK_FLAGS = [f"-L{ROCM / 'lib'}", "-lamdhip64", "-lhiprtc", f"-Wl,-rpath,{ROCM / 'lib'}", "-ldl", "-pthread"]

# (source file, kernel) pairs compiled as the build check of the device code.
KERNEL_CHECKS = [
    ("CoulombSum3d.hip", "directCoulombSum", None),
    ("CoulombSum3d.hip", "directCoulombSumCompensated", None),
    ("CoulombSum3dReference.hip", "directCoulombSumReference", None),
    ("CoulombSum2d.hip", "directCoulombSum", "POINTS_X=2,ROWS_PER_THREAD=8"),
    ("CoulombSum2dReference.hip", "directCoulombSumReference", None),
    ("CoulombSum3dIterative.hip", "directCoulombSum", "POINTS_X=2,ROWS_PER_THREAD=8"),
    ("CoulombSum3dIterative.hip", "directCoulombSumSlice", "POINTS_X=2,ROWS_PER_THREAD=8"),
    ("CoulombSum2d.hip", "directCoulombSum", "POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=128,INNER_UNROLL_FACTOR=16"),
    ("CoulombSum3dIterative.hip", "directCoulombSum", "POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=128,INNER_UNROLL_FACTOR=16"),
    ("CoulombSum3dIterative.hip", "directCoulombSumSlice", "POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=128,INNER_UNROLL_FACTOR=16"),
    ("CoulombSum3dIterativeReference.hip", "directCoulombSumReference", None),
    ("coulomb_util.hip", "kttZeroFill16", None),
    ("coulomb_util.hip", "kttCopy16", None),
    ("coulomb_util.hip", "kttZeroFill4", None),
    ("coulomb_util.hip", "kttJitterAtoms", None),
    ("coulomb_util.hip", "kttRateFma", None),
    ("coulomb_util.hip", "kttRateFmaScalar", None),
    ("coulomb_util.hip", "kttRateRsq", None),
    ("coulomb_util.hip", "kttRateMix", None),
    ("CoulombSum3d.hip", "directCoulombSumCompensated",
     "WORK_GROUP_SIZE_X=32,WORK_GROUP_SIZE_Y=8,POINTS_X=6,ROWS_PER_THREAD=4,ATOM_TILE=128"),
    ("CoulombSum3d.hip", "directCoulombSum",
     "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=8,ATOM_TILE=128,INNER_UNROLL_FACTOR=16"),
    ("CoulombSum3d.hip", "directCoulombSumCompensated",
     "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,POINTS_X=2,ROWS_PER_THREAD=8,ATOM_TILE=256,INNER_UNROLL_FACTOR=16"),
    ("CoulombSum3dZ.hip", "directCoulombSum", None),
    ("CoulombSum3dZ.hip", "directCoulombSum", "Z_ITERATIONS=32,INNER_UNROLL_FACTOR=16,USE_SOA=1"),
    ("CoulombSum3dZ.hip", "directCoulombSum", "WORK_GROUP_SIZE_X=16,WORK_GROUP_SIZE_Y=4,Z_ITERATIONS=1,USE_CONSTANT_MEMORY=1"),
    ("CoulombSum3dZ.hip", "directCoulombSum", "Z_ITERATIONS=4,USE_SOA=1,VECTOR_SIZE=4"),
    ("Nbody.hip", "nbody_kernel", None),
    ("Nbody.hip", "nbody_kernel", "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,BODIES_PER_THREAD=8,ATOM_TILE=512,INNER_UNROLL_FACTOR=8,USE_SOA=1"),
    ("NbodyReference.hip", "nbody_kernel_reference", None),
    ("CoulombBatched.hip", "coulombBatchedPersistent", None),
    ("CoulombBatched.hip", "coulombBatchedSystem", None),
    ("CoulombBatched.hip", "coulombBatchedReset", None),
]

EXAMPLES = ["CoulombSum2d", "CoulombSum3d", "CoulombSum3dIterative", "KttTuningLauncher", "ProfiledCoulombSum3d", "Nbody"]


def _run(command: list[str]) -> None:
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"command failed ({result.returncode}): {' '.join(command)}\n{result.stdout}\n{result.stderr}")


def _compile_object(source: Path) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    target = OBJ_DIR / (source.stem + ".o")
    _run(["g++", *CXX_FLAGS, "-c", str(source), "-o", str(target)])
    return target


def build_library(jobs: int = 8) -> Path:
    sources = sorted((CSRC / "src").glob("*.cpp"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile_object, sources))
    temporary = LIBRARY.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-o", str(temporary), *map(str, objects), *LINK_FLAGS])
    os.replace(temporary, LIBRARY)
    return LIBRARY


def build_counters() -> Path:
    temporary = COUNTERS_LIBRARY.with_suffix(".so.tmp")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", "-shared",
          str(CSRC / "counters" / "HipCounters.cpp"), "-o", str(temporary), f"-L{ROCM / 'lib'}", "-lrocprofiler-sdk",
          f"-Wl,-rpath,{ROCM / 'lib'}", "-pthread"])
    os.replace(temporary, COUNTERS_LIBRARY)
    return COUNTERS_LIBRARY


def build_noalloc() -> Path:
    """libktt_hip_noalloc.so: the precompiled compensated kernels behind
    ktt_coulomb3d_compensated_noalloc, with their scratch in a static device array (hipcc for
    gfx950; libktt_hip.so loads it on first use)."""
    temporary = NOALLOC_LIBRARY.with_suffix(".so.tmp")
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", *HIP_DEVICE_FLAGS,
          f"-I{KERNELS}", str(CSRC / "noalloc" / "CoulombNoAlloc.hip"), "-o", str(temporary)])
    os.replace(temporary, NOALLOC_LIBRARY)
    return NOALLOC_LIBRARY


def build_examples() -> list[Path]:
    BIN_DIR.mkdir(exist_ok=True)
    built = []
    for name in EXAMPLES:
        source = CSRC / "examples" / f"{name}.cpp"
        if not source.exists():
            continue
        target = BIN_DIR / name
        _run(["g++", *CXX_FLAGS, f"-I{REPO / 'include'}", str(source), "-o", str(target), f"-L{PACKAGE}", "-lktt_hip",
              f"-Wl,-rpath,{PACKAGE}", *LINK_FLAGS])
        built.append(target)
    return built


def build_core_tests() -> Path:
    """Host-only test program of the ktt core (tests/test_core.py runs it), and the GPU test
    program of symbol arguments (tests/test_gpu_symbol.py runs it)."""
    BIN_DIR.mkdir(exist_ok=True)
    target = BIN_DIR / "ktt_core_tests"
    _run(["g++", *CXX_FLAGS, str(CSRC / "tests" / "CoreTests.cpp"), "-o", str(target), f"-L{PACKAGE}", "-lktt_hip",
          f"-Wl,-rpath,{PACKAGE}", *LINK_FLAGS])
    _run(["g++", *CXX_FLAGS, str(CSRC / "tests" / "SymbolArgumentTest.cpp"), "-o", str(BIN_DIR / "ktt_symbol_test"),
          f"-L{PACKAGE}", "-lktt_hip", f"-Wl,-rpath,{PACKAGE}", *LINK_FLAGS])
    return target


SANITIZE_DIR = OBJ_DIR / "sanitize"
SANITIZE_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g", "-O1"]


def build_sanitized(jobs: int = 8) -> tuple[Path, Path]:
    """Host-only AddressSanitizer + UBSan build of libktt_hip.so and ktt_core_tests into
    ktt_amd/_build/sanitize/ (g++; the device code is compiled at run time by hipRTC and is not
    instrumented). Run the core tests as ``ktt_amd/_build/sanitize/ktt_core_tests
    tests/golden/coulomb3d_2080_space.json``; for the Python CPU tests, preload the sanitizer
    runtime (``LD_PRELOAD=$(g++ -print-file-name=libasan.so)``, ``ASAN_OPTIONS=detect_leaks=0``)
    and load this library in place of the product one."""
    SANITIZE_DIR.mkdir(parents=True, exist_ok=True)
    flags = [flag for flag in CXX_FLAGS if flag != "-O2"] + SANITIZE_FLAGS

    def compile_one(source: Path) -> Path:
        target = SANITIZE_DIR / (source.stem + ".o")
        _run(["g++", *flags, "-c", str(source), "-o", str(target)])
        return target

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(compile_one, sorted((CSRC / "src").glob("*.cpp"))))
    library = SANITIZE_DIR / "libktt_hip.so"
    _run(["g++", "-shared", *SANITIZE_FLAGS, "-o", str(library), *map(str, objects), *LINK_FLAGS])
    program = SANITIZE_DIR / "ktt_core_tests"
    _run(["g++", *flags, str(CSRC / "tests" / "CoreTests.cpp"), "-o", str(program), f"-L{SANITIZE_DIR}", "-lktt_hip",
          f"-Wl,-rpath,{SANITIZE_DIR}", *LINK_FLAGS])
    return library, program


def check_kernels() -> None:
    library = ctypes.CDLL(str(LIBRARY))
    library.ktt_hip_compile_check.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    library.ktt_hip_last_error.restype = ctypes.c_char_p
    log = ctypes.create_string_buffer(65536)
    failures = []
    for source, kernel, config in KERNEL_CHECKS:
        path = KERNELS / source
        if not path.exists():
            failures.append(f"{source}: missing")
            continue
        status = library.ktt_hip_compile_check(str(path).encode(), kernel.encode(),
                                               config.encode() if config else None, log, len(log))
        if status != 0:
            failures.append(f"{source}:{kernel}: {library.ktt_hip_last_error().decode()}")
    if failures:
        raise RuntimeError("kernel compilation failed:\n" + "\n".join(failures))


def build(verbose: bool = True) -> None:
    if verbose:
        print("[ktt_amd.build] compiling libktt_hip.so", flush=True)
    build_library()
    build_counters()
    if verbose:
        print("[ktt_amd.build] compiling libktt_hip_noalloc.so (hipcc, gfx950)", flush=True)
    build_noalloc()
    if verbose:
        print("[ktt_amd.build] compiling examples", flush=True)
    build_examples()
    build_core_tests()
    if verbose:
        print("[ktt_amd.build] compiling kernels for gfx950 through hipRTC", flush=True)
    check_kernels()


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "sanitize":
        for path in build_sanitized():
            print(path)
    else:
        build(verbose=True)
    sys.exit(0)
