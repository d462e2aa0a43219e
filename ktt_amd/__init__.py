"""ktt_amd: MI355X (gfx950) direct Coulomb summation behind a ktt::Tuner-compatible API.

Layout:
  csrc/ktt/      public C++ headers of the ktt API (Ktt.h, Tuner.h, ...), ComputeApi::HIP
  csrc/src/      the library: kernel/argument model, tuning runner, searchers, validation,
                 the HIP compute engine (hipRTC, hipModule, hipEvent) and the C ABI
  csrc/examples/ the CoulombSum2d / 3d / 3dIterative programs against the C++ API
  kernels/       hand-written HIP kernels for gfx950, compiled at run time by hipRTC
  capi.py        ctypes binding of include/ktt_hip.h (loads libktt_hip.so from this directory)
  workloads.py   seeded synthetic inputs of the benchmark configurations
"""

from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "libktt_hip.so"
KERNEL_DIR = PACKAGE_DIR / "kernels"

__all__ = ["PACKAGE_DIR", "LIBRARY_PATH", "KERNEL_DIR"]
