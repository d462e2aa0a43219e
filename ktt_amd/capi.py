"""ctypes binding of the C ABI in include/ktt_hip.h (libktt_hip.so).

The library is loaded from this package directory only. If it is missing the import fails
with an error naming the build command: there is no fallback of any kind.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE / "libktt_hip.so"

PLAIN = 0
COMPENSATED = 1
SEARCHER_DETERMINISTIC = 0
SEARCHER_RANDOM = 1
SEARCHER_MCMC = 2

# Every function declared in include/ktt_hip.h: name -> (restype, argtypes).
_c = ctypes
_vp = _c.c_void_p
_fp = _c.POINTER(_c.c_float)
_uint3 = _c.c_uint * 3
SIGNATURES = {
    "ktt_hip_abi_version": (_c.c_int, []),
    "ktt_hip_last_error": (_c.c_char_p, []),
    "ktt_hip_device_count": (_c.c_int, [_c.POINTER(_c.c_int)]),
    "ktt_hip_device_info": (_c.c_int, [_c.c_char_p, _c.c_size_t, _c.c_char_p, _c.c_size_t, _c.POINTER(_c.c_int),
                                       _c.POINTER(_c.c_size_t)]),
    "ktt_hip_kernel_directory": (_c.c_char_p, []),
    "ktt_hip_compile_check": (_c.c_int, [_c.c_char_p, _c.c_char_p, _c.c_char_p, _c.c_char_p, _c.c_size_t]),
    "ktt_hip_malloc": (_c.c_int, [_c.POINTER(_vp), _c.c_size_t]),
    "ktt_hip_free": (_c.c_int, [_vp]),
    "ktt_hip_memcpy_htod": (_c.c_int, [_vp, _vp, _c.c_size_t]),
    "ktt_hip_memcpy_dtoh": (_c.c_int, [_vp, _vp, _c.c_size_t]),
    "ktt_hip_zero": (_c.c_int, [_vp, _c.c_size_t, _vp]),
    "ktt_hip_copy": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp]),
    "ktt_hip_issue_rates": (_c.c_int, [_c.c_int] + [_c.POINTER(_c.c_double)] * 4),
    "ktt_hip_synchronize": (_c.c_int, []),
    "ktt_hip_stream_create": (_c.c_int, [_c.POINTER(_vp)]),
    "ktt_hip_stream_destroy": (_c.c_int, [_vp]),
    "ktt_hip_event_create": (_c.c_int, [_c.POINTER(_vp)]),
    "ktt_hip_event_destroy": (_c.c_int, [_vp]),
    "ktt_hip_event_record": (_c.c_int, [_vp, _vp]),
    "ktt_hip_event_elapsed_ms": (_c.c_int, [_vp, _vp, _fp]),
    "ktt_coulomb3d": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _vp, _c.c_int, _c.c_char_p, _vp, _fp]),
    "ktt_coulomb3d_reference": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _vp, _vp, _fp]),
    "ktt_coulomb3d_launch_plan": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_char_p, _uint3, _uint3, _c.POINTER(_c.c_int),
                                             _c.POINTER(_c.c_size_t)]),
    "ktt_coulomb3d_compensated_noalloc": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _vp, _vp]),
    "ktt_coulomb2d": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _c.c_int, _vp, _c.c_char_p, _vp, _fp]),
    "ktt_coulomb2d_reference": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _c.c_int, _vp, _vp, _fp]),
    "ktt_coulomb3d_iterative": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _vp, _c.c_char_p, _vp, _fp, _fp]),
    "ktt_coulomb3d_slices": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _vp, _c.c_char_p, _vp, _fp]),
    "ktt_coulomb3d_iterative_reference": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _vp, _vp, _fp]),
    "ktt_coulomb_launch_shape": (_c.c_int, [_c.c_int, _c.c_int, _c.c_char_p, _uint3, _uint3]),
    "ktt_coulomb3d_tune": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _c.c_char_p,
                                      _c.c_char_p, _c.c_size_t, _c.POINTER(_c.c_double), _c.POINTER(_c.c_int),
                                      _c.POINTER(_c.c_int), _vp]),
    "ktt_run_tuning_script": (_c.c_int, [_c.c_char_p, _c.c_char_p, _c.c_longlong, _c.c_char_p, _c.c_size_t,
                                         _c.POINTER(_c.c_longlong), _c.POINTER(_c.c_longlong), _c.POINTER(_c.c_longlong),
                                         _c.POINTER(_c.c_double)]),
    "ktt_tuning_script_space": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_longlong)]),
    "ktt_script_generate": (_c.c_int, [_c.c_char_p, _c.c_longlong, _vp]),
    "ktt_simulate_search": (_c.c_int, [_c.c_char_p, _c.c_char_p, _c.c_char_p, _c.c_ulonglong, _c.c_longlong, _c.c_uint, _vp,
                                       _c.c_longlong, _c.POINTER(_c.c_longlong)]),
    "ktt_hip_enable_counters": (_c.c_int, []),
    "ktt_nbody_step": (_c.c_int, [_vp, _vp, _c.c_int, _c.c_float, _c.c_float, _c.c_float, _vp, _vp, _c.c_char_p, _vp,
                                  _c.POINTER(_c.c_float)]),
    "ktt_nbody_step_reference": (_c.c_int, [_vp, _vp, _c.c_int, _c.c_float, _c.c_float, _c.c_float, _vp, _vp, _vp,
                                            _c.POINTER(_c.c_float)]),
    "ktt_nbody_tune": (_c.c_int, [_vp, _vp, _c.c_int, _c.c_float, _c.c_float, _c.c_float, _c.c_int, _c.c_int, _c.c_double,
                                  _c.c_char_p, _c.c_char_p, _c.c_size_t, _c.POINTER(_c.c_double), _c.POINTER(_c.c_int),
                                  _c.POINTER(_c.c_int)]),
    "ktt_coulomb3d_dynamic": (_c.c_int, [_vp, _c.c_int, _c.c_float, _c.c_int, _c.c_int, _c.c_float, _c.c_ulonglong, _c.c_int,
                                         _c.c_int, _c.c_char_p, _vp, _vp, _vp, _c.c_char_p, _c.c_size_t,
                                         _c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _vp, _vp]),
    "ktt_coulomb_batch_create": (_c.c_int, [_vp, _vp, _vp, _vp, _c.c_int, _c.c_char_p, _c.POINTER(_vp)]),
    "ktt_coulomb_batch_destroy": (_c.c_int, [_vp]),
    "ktt_coulomb_batch_info": (_c.c_int, [_vp, _c.POINTER(_vp), _c.POINTER(_c.c_size_t), _vp, _c.POINTER(_c.c_int),
                                          _c.POINTER(_c.c_longlong), _c.c_char_p, _c.c_size_t]),
    "ktt_coulomb_batch_zero": (_c.c_int, [_vp, _vp]),
    "ktt_coulomb_batch_run": (_c.c_int, [_vp, _c.c_int, _vp, _fp]),
    "ktt_coulomb_batch_run_per_system": (_c.c_int, [_vp, _c.c_int, _vp, _fp]),
    "ktt_coulomb_batch_tune": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_char_p, _c.c_char_p, _c.c_size_t]),
}


class KttHipError(RuntimeError):
    """A C entry point returned an error."""


_library = None


def library() -> ctypes.CDLL:
    global _library
    if _library is None:
        if not LIBRARY_PATH.exists():
            raise ImportError(f"{LIBRARY_PATH} is not built; run `python -m ktt_amd.build` (or __graft_entry__.build())")
        handle = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            function = getattr(handle, name)
            function.restype = restype
            function.argtypes = argtypes
        _library = handle
    return _library


def _check(status: int, what: str) -> None:
    if status != 0:
        message = library().ktt_hip_last_error().decode(errors="replace")
        raise KttHipError(f"{what} failed: {message}")


def _config(config) -> bytes | None:
    if config is None:
        return None
    if isinstance(config, dict):
        config = ",".join(f"{key}={value}" for key, value in config.items())
    return config.encode()


def device_count() -> int:
    count = ctypes.c_int(0)
    _check(library().ktt_hip_device_count(ctypes.byref(count)), "ktt_hip_device_count")
    return count.value


def device_info() -> dict:
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    units = ctypes.c_int(0)
    memory = ctypes.c_size_t(0)
    _check(library().ktt_hip_device_info(name, 256, arch, 64, ctypes.byref(units), ctypes.byref(memory)), "ktt_hip_device_info")
    return {"name": name.value.decode(), "arch": arch.value.decode(), "compute_units": units.value, "memory": memory.value}


def compile_check(source: str, kernel: str, config=None) -> str:
    log = ctypes.create_string_buffer(65536)
    _check(library().ktt_hip_compile_check(source.encode(), kernel.encode(), _config(config), log, len(log)),
           f"compile {source}:{kernel}")
    return log.value.decode(errors="replace")


class DeviceBuffer:
    """Device allocation owned by Python; freed on close() or garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        pointer = ctypes.c_void_p()
        _check(library().ktt_hip_malloc(ctypes.byref(pointer), max(self.nbytes, 4)), "ktt_hip_malloc")
        self.pointer = pointer

    @classmethod
    def from_array(cls, array: np.ndarray) -> "DeviceBuffer":
        array = np.ascontiguousarray(array)
        buffer = cls(array.nbytes)
        buffer.upload(array)
        return buffer

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self.pointer

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array)
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if array.nbytes:
            _check(library().ktt_hip_memcpy_htod(self.pointer, array.ctypes.data_as(ctypes.c_void_p), array.nbytes), "upload")

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the buffer holds")
        if out.nbytes:
            _check(library().ktt_hip_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.pointer, out.nbytes), "download")
        return out

    def zero(self, stream=None) -> None:
        _check(library().ktt_hip_zero(self.pointer, self.nbytes - self.nbytes % 4, stream), "ktt_hip_zero")

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None, stream=None) -> None:
        count = min(self.nbytes, other.nbytes) if nbytes is None else nbytes
        _check(library().ktt_hip_copy(self.pointer, other.pointer, count - count % 16, stream), "ktt_hip_copy")

    def close(self) -> None:
        if self.pointer is not None and self.pointer.value:
            library().ktt_hip_free(self.pointer)
        self.pointer = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def issue_rates(iterations: int = 2048) -> dict:
    """Marginal vector-ALU issue rates of the device (lane operations per second): FMA with VGPR
    operands (dual issue across waves), FMA with an SGPR operand (single issue), v_rsq_f32 alone,
    and the Coulomb mix of 2 FMA + 1 v_rsq_f32 inside each wave."""
    rates = [ctypes.c_double(0.0) for _ in range(4)]
    _check(library().ktt_hip_issue_rates(iterations, *(ctypes.byref(rate) for rate in rates)), "ktt_hip_issue_rates")
    return {"fma": rates[0].value, "fma_scalar_operand": rates[1].value, "rsq": rates[2].value, "mix": rates[3].value}


def synchronize() -> None:
    _check(library().ktt_hip_synchronize(), "ktt_hip_synchronize")


class Event:
    def __init__(self):
        handle = ctypes.c_void_p()
        _check(library().ktt_hip_event_create(ctypes.byref(handle)), "ktt_hip_event_create")
        self.handle = handle

    def record(self, stream=None) -> None:
        _check(library().ktt_hip_event_record(self.handle, stream), "ktt_hip_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        value = ctypes.c_float(0.0)
        _check(library().ktt_hip_event_elapsed_ms(self.handle, end.handle, ctypes.byref(value)), "ktt_hip_event_elapsed_ms")
        return value.value

    def __del__(self):
        try:
            library().ktt_hip_event_destroy(self.handle)
        except Exception:
            pass


def _timing(timed: bool):
    return (ctypes.c_float(0.0), True) if timed else (None, False)


def coulomb3d(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid: int, energy: DeviceBuffer, form: int = PLAIN,
              config=None, stream=None, timed: bool = True) -> float | None:
    ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb3d(atoms.ptr, n_atoms, spacing, grid, energy.ptr, form, _config(config), stream,
                                   ctypes.byref(ms) if timed else None), "ktt_coulomb3d")
    return ms.value if timed else None


def coulomb3d_compensated_noalloc(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid: int, energy: DeviceBuffer,
                                  stream=None) -> None:
    """The compensated 3D sum through the entry point that allocates nothing (asynchronous)."""
    _check(library().ktt_coulomb3d_compensated_noalloc(atoms.ptr, n_atoms, spacing, grid, energy.ptr, stream),
           "ktt_coulomb3d_compensated_noalloc")


def coulomb3d_reference(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid: int, energy: DeviceBuffer, stream=None,
                        timed: bool = True) -> float | None:
    ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb3d_reference(atoms.ptr, n_atoms, spacing, grid, energy.ptr, stream,
                                             ctypes.byref(ms) if timed else None), "ktt_coulomb3d_reference")
    return ms.value if timed else None


def coulomb2d(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid_x: int, grid_y: int, energy: DeviceBuffer, config=None,
              stream=None, timed: bool = True) -> float | None:
    ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb2d(atoms.ptr, n_atoms, spacing, grid_x, grid_y, energy.ptr, _config(config), stream,
                                   ctypes.byref(ms) if timed else None), "ktt_coulomb2d")
    return ms.value if timed else None


def coulomb2d_reference(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid_x: int, grid_y: int, energy: DeviceBuffer,
                        stream=None, timed: bool = True) -> float | None:
    ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb2d_reference(atoms.ptr, n_atoms, spacing, grid_x, grid_y, energy.ptr, stream,
                                             ctypes.byref(ms) if timed else None), "ktt_coulomb2d_reference")
    return ms.value if timed else None


def coulomb3d_iterative(host_atoms: np.ndarray, spacing: float, grid: int, energy: DeviceBuffer, config=None,
                        stream=None) -> tuple[float, float]:
    host_atoms = np.ascontiguousarray(host_atoms, dtype=np.float32)
    kernel_ms = ctypes.c_float(0.0)
    wall_ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb3d_iterative(host_atoms.ctypes.data_as(ctypes.c_void_p), host_atoms.shape[0], spacing, grid,
                                             energy.ptr, _config(config), stream, ctypes.byref(kernel_ms),
                                             ctypes.byref(wall_ms)), "ktt_coulomb3d_iterative")
    return kernel_ms.value, wall_ms.value


def coulomb3d_slices(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid: int, energy: DeviceBuffer, config=None,
                     stream=None, timed: bool = True) -> float | None:
    ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb3d_slices(atoms.ptr, n_atoms, spacing, grid, energy.ptr, _config(config), stream,
                                          ctypes.byref(ms) if timed else None), "ktt_coulomb3d_slices")
    return ms.value if timed else None


def coulomb3d_iterative_reference(atoms: DeviceBuffer, n_atoms: int, spacing: float, grid: int, energy: DeviceBuffer,
                                  stream=None, timed: bool = True) -> float | None:
    ms = ctypes.c_float(0.0)
    _check(library().ktt_coulomb3d_iterative_reference(atoms.ptr, n_atoms, spacing, grid, energy.ptr, stream,
                                                       ctypes.byref(ms) if timed else None), "ktt_coulomb3d_iterative_reference")
    return ms.value if timed else None


def coulomb3d_launch_plan(n_atoms: int, grid: int, form: int = PLAIN, config=None) -> dict:
    """The launch ktt_coulomb3d makes: workgroup grid (z = atom blocks), workgroup shape, first split
    tile and the split's scratch bytes. "groups" is grid x times grid y, the tile count unless the
    row groups fold into x (then the last run of row groups holds workgroups that exit)."""
    g, b = _uint3(), _uint3()
    split_from = ctypes.c_int(0)
    scratch = ctypes.c_size_t(0)
    _check(library().ktt_coulomb3d_launch_plan(n_atoms, grid, form, _config(config), g, b, ctypes.byref(split_from),
                                               ctypes.byref(scratch)), "ktt_coulomb3d_launch_plan")
    return {"grid": tuple(g), "block": tuple(b), "split_from": split_from.value, "scratch_bytes": scratch.value,
            "groups": g[0] * g[1]}


def launch_shape(kind: int, grid: int, config=None) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
    g = _uint3()
    b = _uint3()
    _check(library().ktt_coulomb_launch_shape(kind, grid, _config(config), g, b), "ktt_coulomb_launch_shape")
    return tuple(g), tuple(b)


def coulomb3d_tune(host_atoms: np.ndarray, spacing: float, grid: int, searcher: int = SEARCHER_DETERMINISTIC,
                   max_configurations: int = 0, tolerance: float = 0.01, results_file: str | None = None,
                   want_grid: bool = False) -> dict:
    host_atoms = np.ascontiguousarray(host_atoms, dtype=np.float32)
    best = ctypes.create_string_buffer(1024)
    best_ms = ctypes.c_double(0.0)
    valid = ctypes.c_int(0)
    tested = ctypes.c_int(0)
    output = np.zeros((grid, grid, grid), dtype=np.float32) if want_grid else None
    _check(library().ktt_coulomb3d_tune(host_atoms.ctypes.data_as(ctypes.c_void_p), host_atoms.shape[0], spacing, grid, searcher,
                                        max_configurations, tolerance, results_file.encode() if results_file else None, best,
                                        len(best), ctypes.byref(best_ms), ctypes.byref(valid), ctypes.byref(tested),
                                        output.ctypes.data_as(ctypes.c_void_p) if output is not None else None),
           "ktt_coulomb3d_tune")
    return {"best_config": best.value.decode(), "best_ms": best_ms.value, "valid": valid.value, "tested": tested.value,
            "grid": output}


def run_tuning_script(script_path: str, output_file: str | None = None, configuration_limit: int = 0) -> dict:
    """Runs a KTT JSON tuning script (Language "HIP"); returns its summary."""
    best = ctypes.create_string_buffer(2048)
    space, tested, valid = ctypes.c_longlong(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
    best_ms = ctypes.c_double(0.0)
    _check(library().ktt_run_tuning_script(str(script_path).encode(), output_file.encode() if output_file else None,
                                           configuration_limit, best, len(best), ctypes.byref(space), ctypes.byref(tested),
                                           ctypes.byref(valid), ctypes.byref(best_ms)), "ktt_run_tuning_script")
    return {"best_config": best.value.decode(), "space": space.value, "tested": tested.value, "valid": valid.value,
            "best_ms": best_ms.value}


def tuning_script_space(script_path: str) -> int:
    size = ctypes.c_longlong(0)
    _check(library().ktt_tuning_script_space(str(script_path).encode(), ctypes.byref(size)), "ktt_tuning_script_space")
    return size.value


def script_generate(expression: str, count: int) -> np.ndarray:
    values = np.zeros(count, dtype=np.float64)
    _check(library().ktt_script_generate(expression.encode(), count, values.ctypes.data_as(ctypes.c_void_p)), "ktt_script_generate")
    return values


def simulate_search(results_file: str, searcher: str, knowledge_file: str | None = None, seed: int = 0, max_steps: int = 0,
                    compute_units: int = 0) -> np.ndarray:
    """Simulated tuning over saved KTT JSON results; the total duration (ms) of every step, -1
    for invalid results."""
    capacity = 1 << 20 if max_steps <= 0 else max_steps
    durations = np.zeros(capacity, dtype=np.float64)
    steps = ctypes.c_longlong(0)
    _check(library().ktt_simulate_search(str(results_file).encode(), searcher.encode(),
                                         str(knowledge_file).encode() if knowledge_file else None, seed, max_steps, compute_units,
                                         durations.ctypes.data_as(ctypes.c_void_p), capacity, ctypes.byref(steps)),
           "ktt_simulate_search")
    return durations[:min(steps.value, capacity)].copy()


def nbody_step(bodies: DeviceBuffer, velocity: DeviceBuffer, n: int, dt: float, damping: float, softening_sqr: float,
               new_bodies: DeviceBuffer, new_velocity: DeviceBuffer, config=None, stream=None, timed: bool = False):
    ms = ctypes.c_float(0.0)
    _check(library().ktt_nbody_step(bodies.ptr, velocity.ptr, n, dt, damping, softening_sqr, new_bodies.ptr, new_velocity.ptr,
                                    _config(config), stream, ctypes.byref(ms) if timed else None), "ktt_nbody_step")
    return ms.value if timed else None


def nbody_step_reference(bodies: DeviceBuffer, velocity: DeviceBuffer, n: int, dt: float, damping: float, softening_sqr: float,
                         new_bodies: DeviceBuffer, new_velocity: DeviceBuffer, stream=None, timed: bool = False):
    ms = ctypes.c_float(0.0)
    _check(library().ktt_nbody_step_reference(bodies.ptr, velocity.ptr, n, dt, damping, softening_sqr, new_bodies.ptr,
                                              new_velocity.ptr, stream, ctypes.byref(ms) if timed else None),
           "ktt_nbody_step_reference")
    return ms.value if timed else None


def nbody_tune(host_bodies: np.ndarray, host_velocity: np.ndarray, dt: float, damping: float, softening_sqr: float,
               searcher: int = SEARCHER_DETERMINISTIC, max_configurations: int = 0, tolerance: float = 0.001,
               results_file: str | None = None) -> dict:
    bodies = np.ascontiguousarray(host_bodies, dtype=np.float32)
    velocity = np.ascontiguousarray(host_velocity, dtype=np.float32)
    best = ctypes.create_string_buffer(1024)
    best_ms = ctypes.c_double(0.0)
    valid = ctypes.c_int(0)
    tested = ctypes.c_int(0)
    _check(library().ktt_nbody_tune(bodies.ctypes.data_as(ctypes.c_void_p), velocity.ctypes.data_as(ctypes.c_void_p), bodies.shape[0],
                                    dt, damping, softening_sqr, searcher, max_configurations, tolerance,
                                    results_file.encode() if results_file else None, best, len(best), ctypes.byref(best_ms),
                                    ctypes.byref(valid), ctypes.byref(tested)), "ktt_nbody_tune")
    return {"best_config": best.value.decode(), "best_ms": best_ms.value, "valid": valid.value, "tested": tested.value}


def enable_counters() -> None:
    """Enables hardware counters; effective only before the process's first HIP call."""
    _check(library().ktt_hip_enable_counters(), "ktt_hip_enable_counters")


def coulomb3d_dynamic(host_atoms: np.ndarray, spacing: float, grid: int, steps: int, sigma: float, seed: int,
                      searcher: int = SEARCHER_MCMC, tuning_budget: int = 0, results_file: str | None = None,
                      want_final: bool = False) -> dict:
    """Dynamic tuning run (``ktt_coulomb3d_dynamic``): per-step kernel and wall times, which
    steps tuned, the best configuration, and optionally the final atoms and map."""
    host_atoms = np.ascontiguousarray(host_atoms, dtype=np.float32)
    if host_atoms.ndim != 2 or host_atoms.shape[1] != 4:
        raise ValueError("atoms must have shape (N, 4)")
    kernel_ms = np.zeros(steps, np.float64)
    wall_ms = np.zeros(steps, np.float64)
    tuned = np.zeros(steps, np.int32)
    best = ctypes.create_string_buffer(1024)
    tested, space = ctypes.c_int(0), ctypes.c_int(0)
    final_atoms = np.empty_like(host_atoms) if want_final else None
    final_grid = np.empty((grid, grid, grid), np.float32) if want_final else None

    def pointer(array):
        return array.ctypes.data_as(ctypes.c_void_p) if array is not None else None

    _check(library().ktt_coulomb3d_dynamic(pointer(host_atoms), host_atoms.shape[0], spacing, grid, steps, sigma, seed, searcher,
                                           tuning_budget, results_file.encode() if results_file else None, pointer(kernel_ms),
                                           pointer(wall_ms), pointer(tuned), best, len(best), ctypes.byref(tested),
                                           ctypes.byref(space), pointer(final_atoms), pointer(final_grid)),
           "ktt_coulomb3d_dynamic")
    return {"kernel_ms": kernel_ms, "wall_ms": wall_ms, "tuned": tuned.astype(bool), "best_config": best.value.decode(),
            "tested": tested.value, "space": space.value, "atoms": final_atoms, "grid": final_grid}


class Batch:
    """Batch of independent systems (include/ktt_hip.h section 5)."""

    def __init__(self, systems, config=None):
        """systems: iterable of objects with .atoms (N, 4) float32, .grid (side) and .spacing."""
        systems = list(systems)
        atoms = np.ascontiguousarray(np.concatenate([s.atoms for s in systems], axis=0) if systems else
                                     np.zeros((0, 4), np.float32), dtype=np.float32)
        counts = np.array([s.atoms.shape[0] for s in systems], dtype=np.int32)
        sides = np.array([s.grid for s in systems], dtype=np.int32)
        spacings = np.array([s.spacing for s in systems], dtype=np.float32)
        handle = ctypes.c_void_p()
        _check(library().ktt_coulomb_batch_create(atoms.ctypes.data_as(ctypes.c_void_p), counts.ctypes.data_as(ctypes.c_void_p),
                                                  sides.ctypes.data_as(ctypes.c_void_p),
                                                  spacings.ctypes.data_as(ctypes.c_void_p), len(systems), _config(config),
                                                  ctypes.byref(handle)), "ktt_coulomb_batch_create")
        self.handle = handle
        self.sides = sides
        self.count = len(systems)

    def info(self) -> dict:
        grids = ctypes.c_void_p()
        points = ctypes.c_size_t()
        offsets = np.zeros(self.count, dtype=np.int64)
        items = ctypes.c_int()
        interactions = ctypes.c_longlong()
        shapes = ctypes.create_string_buffer(2048)
        _check(library().ktt_coulomb_batch_info(self.handle, ctypes.byref(grids), ctypes.byref(points),
                                                offsets.ctypes.data_as(ctypes.c_void_p), ctypes.byref(items),
                                                ctypes.byref(interactions), shapes, len(shapes)), "ktt_coulomb_batch_info")
        return {"grids": grids, "points": points.value, "offsets": offsets, "items": items.value,
                "interactions": interactions.value, "shapes": shapes.value.decode()}

    def zero(self, stream=None) -> None:
        _check(library().ktt_coulomb_batch_zero(self.handle, stream), "ktt_coulomb_batch_zero")

    def run(self, use_graph: bool = True, stream=None, timed: bool = True) -> float | None:
        ms = ctypes.c_float(0.0)
        _check(library().ktt_coulomb_batch_run(self.handle, 1 if use_graph else 0, stream, ctypes.byref(ms) if timed else None),
               "ktt_coulomb_batch_run")
        return ms.value if timed else None

    def run_per_system(self, form: int = PLAIN, stream=None, timed: bool = True) -> float | None:
        ms = ctypes.c_float(0.0)
        _check(library().ktt_coulomb_batch_run_per_system(self.handle, form, stream, ctypes.byref(ms) if timed else None),
               "ktt_coulomb_batch_run_per_system")
        return ms.value if timed else None

    def tune(self, searcher: int = SEARCHER_RANDOM, max_configurations: int = 8, results_file: str | None = None) -> str:
        best = ctypes.create_string_buffer(2048)
        _check(library().ktt_coulomb_batch_tune(self.handle, searcher, max_configurations,
                                                results_file.encode() if results_file else None, best, len(best)),
               "ktt_coulomb_batch_tune")
        return best.value.decode()

    def grids(self) -> list[np.ndarray]:
        """Host copies of every system's grid."""
        info = self.info()
        total = info["points"]
        flat = np.empty(total, dtype=np.float32)
        _check(library().ktt_hip_memcpy_dtoh(flat.ctypes.data_as(ctypes.c_void_p), info["grids"], flat.nbytes), "download grids")
        out = []
        for k, side in enumerate(self.sides):
            start = int(info["offsets"][k])
            out.append(flat[start:start + int(side) ** 3].reshape(int(side), int(side), int(side)))
        return out

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            library().ktt_coulomb_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
