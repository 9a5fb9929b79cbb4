"""Python bindings of the MI355X split-operator propagation backend.

The product is two in-tree shared libraries:
  lib/libspo_hip.so   HIP/CDNA4 kernels behind the C ABI of include/spo_hip.h (the propagation backend);
  lib/libspo_host.so  the host driver behind include/spo_host.h (set-up from the JSON input, the time loop,
                      the analysis results), which runs the HIP backend or a backend made by a caller's factory.

This module only loads them with ctypes and converts arrays; it computes nothing itself. A missing library is an
error: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import json
import os
from typing import Optional

import numpy as np

PACKAGE_DIR = os.path.dirname(os.path.abspath(os.path.dirname(__file__)))
LIB_DIR = os.path.join(PACKAGE_DIR, "lib")

SPO_DEVICE_HIP = 1
SPO_DEVICE_EXTERNAL = 2

_FACTORY = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p)


class SpoError(RuntimeError):
    """A failure reported by the C ABI; code is the status (2 input, 3 not supported, 1 other)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[status {code}] {message}")
        self.code = code


class RunInfo(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int) for name in (
        "n_rot", "nx", "ny", "nr", "nz", "nz_specular", "nz_potential",
        "n_theta", "n_phi", "n_representatives", "initial_channel",
        "n_states", "n_v", "n_energies", "n_diffraction",
        "steps_per_analysis", "analysis_steps_total", "analyses_per_average", "iz_analysis")] + [
        ("time_step", ctypes.c_double)]


_host = None


def library_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def host_library() -> ctypes.CDLL:
    """libspo_host.so (loads libspo_hip.so with it). Raises if it has not been built."""
    global _host
    if _host is None:
        path = library_path("libspo_host.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build the libraries first (python cpp-spo-dvr_amd/build.py)")
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        lib.spo_last_error.restype = ctypes.c_char_p
        lib.spo_run_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_void_p)]
        lib.spo_run_create_with_potential.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_size_t,
                                                      ctypes.POINTER(ctypes.c_void_p)]
        lib.spo_run_destroy.argtypes = [ctypes.c_void_p]
        lib.spo_run_get_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(RunInfo)]
        lib.spo_run_backend.argtypes = [ctypes.c_void_p]
        lib.spo_run_backend.restype = ctypes.c_void_p
        lib.spo_run_call_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        for fn in ("spo_run_steps",):
            getattr(lib, fn).argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.spo_run_propagate.argtypes = [ctypes.c_void_p]
        lib.spo_run_steps_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int]
        lib.spo_run_kinetic_and_potential.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.spo_run_analyse.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
        lib.spo_run_get_wave.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_set_wave.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_get_specular.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_block_densities.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_z_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.spo_run_site_column.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.spo_run_bk_record.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.spo_run_bk_count.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        lib.spo_run_averages.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        lib.spo_run_surface.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_double)]
        lib.spo_run_results.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_s_sums.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_write_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.spo_run_energy_resolved_shape.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                      ctypes.POINTER(ctypes.c_int)]
        lib.spo_run_energy_resolved.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_run_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t)]
        lib.spo_pes_value.argtypes = [ctypes.c_char_p] + [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)]
        lib.spo_run_kernel_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.spo_run_kernel_times.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        lib.spo_crp_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        lib.spo_crp_close.argtypes = [ctypes.c_void_p]
        for fn in ("spo_crp_molecule", "spo_crp_atom"):
            getattr(lib, fn).argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p]
        for fn in ("spo_crp_qd", "spo_crp_gas_phase", "spo_crp_device_values"):
            getattr(lib, fn).argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_crp_geometry.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_crp_parameters.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_scm_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        lib.spo_scm_close.argtypes = [ctypes.c_void_p]
        lib.spo_scm_values.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_scm_device_values.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
        lib.spo_scm_parameters.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _host = lib
    return _host


def _check(status: int) -> None:
    if status != 0:
        raise SpoError(status, host_library().spo_last_error().decode(errors="replace"))


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def pes_value(system_id: str, x, y, z, r, theta, phi) -> float:
    out = ctypes.c_double()
    _check(host_library().spo_pes_value(system_id.encode(), x, y, z, r, theta, phi, ctypes.byref(out)))
    return out.value


_PES_IDS = {"HardWallPES": 1, "H2Cu111DaiLight": 2}  # SPO_HIP_PES_* of include/spo_hip.h


def device_pes_values(system_id: str, points: np.ndarray) -> np.ndarray:
    """The surface in eV at points (n, 6) = (x, y, z, r, theta, phi), evaluated on the GPU (spo_hip_pes_values)."""
    host_library()  # loads libspo_hip.so with it
    lib = ctypes.CDLL(library_path("libspo_hip.so"))
    lib.spo_hip_pes_values.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    lib.spo_hip_last_error.restype = ctypes.c_char_p
    if system_id not in _PES_IDS:
        raise SpoError(3, f"SystemID {system_id} has no device surface")
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 6)
    out = np.empty(len(pts))
    status = lib.spo_hip_pes_values(_PES_IDS[system_id], len(pts), _ptr(pts), _ptr(out))
    if status != 0:
        raise SpoError(status, lib.spo_hip_last_error().decode(errors="replace"))
    return out


# Unit codes of the CRP library (src/crplib/units.hpp:11-17)
CRP_BOHR, CRP_ANGSTROM = 0, 1
CRP_HARTREE, CRP_EV, CRP_KCAL_PER_MOL, CRP_KJ_PER_MOL, CRP_JOULE = 2, 3, 4, 5, 6


def _points(a, width: int) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, width)


class CrpSurface:
    """The CRP potential of a C3v or C6v data set directory (SystemID CRPlibrary), evaluated by libspo_host.so.

    molecule(atoms (n, 6) = xa ya za xb yb zb) -> (n, 7) V and its gradient on both atoms; atom(positions (n, 3))
    -> (n, 4); qd(points (n, 6) = u v Z r theta phi) -> V in eV as the propagation uses it; gas_phase(zr (n, 2)) ->
    (n, 5) F, dF/dZ, V2D, dV2D/dZ, dV2D/dr; parameters() -> dict."""

    def __init__(self, directory: str):
        lib = host_library()
        handle = ctypes.c_void_p()
        _check(lib.spo_crp_open(directory.encode(), ctypes.byref(handle)))
        self._handle = handle

    def close(self) -> None:
        if self._handle:
            host_library().spo_crp_close(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def molecule(self, atoms, length_unit: int = CRP_BOHR, energy_unit: int = CRP_HARTREE) -> np.ndarray:
        pts = _points(atoms, 6)
        out = np.empty((len(pts), 7))
        _check(host_library().spo_crp_molecule(self._handle, len(pts), _ptr(pts), length_unit, energy_unit, _ptr(out)))
        return out

    def atom(self, positions, length_unit: int = CRP_BOHR, energy_unit: int = CRP_HARTREE) -> np.ndarray:
        pts = _points(positions, 3)
        out = np.empty((len(pts), 4))
        _check(host_library().spo_crp_atom(self._handle, len(pts), _ptr(pts), length_unit, energy_unit, _ptr(out)))
        return out

    def qd(self, points) -> np.ndarray:
        pts = _points(points, 6)
        out = np.empty(len(pts))
        _check(host_library().spo_crp_qd(self._handle, len(pts), _ptr(pts), _ptr(out)))
        return out

    def device_qd(self, points) -> np.ndarray:
        """qd(points) evaluated on the GPU."""
        pts = _points(points, 6)
        out = np.empty(len(pts))
        _check(host_library().spo_crp_device_values(self._handle, len(pts), _ptr(pts), _ptr(out)))
        return out

    def gas_phase(self, zr) -> np.ndarray:
        pts = _points(zr, 2)
        out = np.empty((len(pts), 5))
        _check(host_library().spo_crp_gas_phase(self._handle, len(pts), _ptr(pts), _ptr(out)))
        return out

    @staticmethod
    def geometry(atoms) -> np.ndarray:
        pts = _points(atoms, 6)
        out = np.empty((len(pts), 6))
        _check(host_library().spo_crp_geometry(len(pts), _ptr(pts), _ptr(out)))
        return out

    def parameters(self) -> dict:
        out = np.empty(9)
        _check(host_library().spo_crp_parameters(self._handle, _ptr(out)))
        keys = ("lattice_constant", "z_asymptotic", "skewing_angle", "zso_min", "zso_max", "r_min", "r_max", "z_min")
        d = dict(zip(keys, (float(x) for x in out[:8])))
        d["point_group"] = {3.0: "C3v", 4.0: "C4v"}.get(float(out[8]), "C6v")
        return d


class ScmSurface:
    """The CRP surface plus the SCM correction of a distorted slab snapshot (SystemID CRPDistrtSurface).

    values(points (n, 6) = u v Z r theta phi) -> (n, 5): Cartesian x, y, dV (hartree), the CRP at (u/s, v/s) (eV)
    and the total (eV)."""

    def __init__(self, crp_data_set: str, parameter_file: str, position_file: str, configuration: int = 1,
                 cutoff: float = 16.0, rescale: bool = True):
        handle = ctypes.c_void_p()
        _check(host_library().spo_scm_open(crp_data_set.encode(), parameter_file.encode(), position_file.encode(),
                                           configuration, cutoff, int(rescale), ctypes.byref(handle)))
        self._handle = handle

    def close(self) -> None:
        if self._handle:
            host_library().spo_scm_close(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def values(self, points) -> np.ndarray:
        pts = _points(points, 6)
        out = np.empty((len(pts), 5))
        _check(host_library().spo_scm_values(self._handle, len(pts), _ptr(pts), _ptr(out)))
        return out

    def device_values(self, points) -> np.ndarray:
        """The total (eV) of values(points)[:, 4], evaluated on the GPU."""
        pts = _points(points, 6)
        out = np.empty(len(pts))
        _check(host_library().spo_scm_device_values(self._handle, len(pts), _ptr(pts), _ptr(out)))
        return out

    def parameters(self) -> dict:
        out = np.empty(5)
        _check(host_library().spo_scm_parameters(self._handle, _ptr(out)))
        return dict(scale=float(out[0]), lattice_constant=float(out[1]), cos_skew=float(out[2]),
                    sin_skew=float(out[3]), atoms=int(out[4]))


def _hip_library() -> ctypes.CDLL:
    host_library()  # loads libspo_hip.so with it
    lib = ctypes.CDLL(library_path("libspo_hip.so"))
    lib.spo_hip_last_error.restype = ctypes.c_char_p
    lib.spo_hip_mlp_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.spo_hip_mlp_destroy.argtypes = [ctypes.c_void_p]
    lib.spo_hip_mlp_set_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.spo_hip_mlp_get_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.spo_hip_mlp_gradient.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
    lib.spo_hip_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_void_p]
    lib.spo_hip_mlp_predict.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    return lib


class Mlp:
    """The potential surrogate on the GPU (spo_hip_mlp_*): 6 features -> 128 -> 128 -> 128 -> 1, SiLU, hidden layers
    on bf16 MFMA. Parameters are a float32 vector of MLP_PARAMS values (include/spo_hip.h)."""

    PARAMS = 34049

    def __init__(self, batch: int):
        self._lib = _hip_library()
        handle = ctypes.c_void_p()
        self._check(self._lib.spo_hip_mlp_create(batch, ctypes.byref(handle)))
        self._handle = handle
        self.batch = batch

    def _check(self, status: int) -> None:
        if status != 0:
            raise SpoError(status, self._lib.spo_hip_last_error().decode(errors="replace"))

    def close(self) -> None:
        if self._handle:
            self._lib.spo_hip_mlp_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.PARAMS, dtype=np.float32)
        self._check(self._lib.spo_hip_mlp_get_params(self._handle, _ptr(out)))
        return out

    @params.setter
    def params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.PARAMS:
            raise ValueError(f"expected {self.PARAMS} parameters")
        self._check(self._lib.spo_hip_mlp_set_params(self._handle, _ptr(a)))

    def gradient(self, x, t):
        xs = np.ascontiguousarray(x, dtype=np.float32).reshape(self.batch, 6)
        ts = np.ascontiguousarray(t, dtype=np.float32).reshape(self.batch)
        loss = np.zeros(1, dtype=np.float32)
        grad = np.empty(self.PARAMS, dtype=np.float32)
        self._check(self._lib.spo_hip_mlp_gradient(self._handle, _ptr(xs), _ptr(ts), _ptr(loss), _ptr(grad)))
        return float(loss[0]), grad

    def train(self, x, t, steps: int, lr: float) -> np.ndarray:
        xs = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        ts = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        losses = np.empty(max(steps, 1), dtype=np.float32)
        self._check(self._lib.spo_hip_mlp_train(self._handle, len(xs), _ptr(xs), _ptr(ts), steps, lr, _ptr(losses)))
        return losses[:steps]

    def predict(self, x) -> np.ndarray:
        xs = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        out = np.empty(len(xs), dtype=np.float32)
        self._check(self._lib.spo_hip_mlp_predict(self._handle, len(xs), _ptr(xs), _ptr(out)))
        return out


def run_batch(runs, n_analysis_steps: int) -> None:
    """Up to n_analysis_steps analysis steps of every run in lockstep (spo_run_steps_batch): on the GPU the runs'
    time steps are enqueued together, each run's results unchanged."""
    handles = (ctypes.c_void_p * len(runs))(*[r._h for r in runs])
    _check(host_library().spo_run_steps_batch(handles, len(runs), n_analysis_steps))


class Run:
    """A propagation run: the set-up from a JSON input (plus overrides) and a backend.

    device="hip" uses the HIP backend; device="external" uses factory, a C function pointer
    (void* (*)(const void*)) that makes a spo::PropagationBackend (e.g. the CPU oracle's, in tests). potential, when
    given, replaces the set-up's potential grid (hartree, in the order of array("potential.values"); see
    spo_run_create_with_potential)."""

    def __init__(self, input_path: str, overrides: Optional[dict] = None, device: str = "hip", factory=None,
                 potential: Optional[np.ndarray] = None):
        lib = host_library()
        handle = ctypes.c_void_p()
        over = json.dumps(overrides).encode() if overrides else None
        if device == "hip":
            code, fptr = SPO_DEVICE_HIP, None
        elif device == "external":
            if factory is None:
                raise ValueError("device='external' needs a factory")
            code, fptr = SPO_DEVICE_EXTERNAL, ctypes.cast(factory, ctypes.c_void_p)
        else:
            raise ValueError(f"unknown device {device!r}")
        if potential is None:
            _check(lib.spo_run_create(input_path.encode(), over, code, fptr, ctypes.byref(handle)))
        else:
            grid = np.ascontiguousarray(potential, dtype=np.float64)
            _check(lib.spo_run_create_with_potential(input_path.encode(), over, code, fptr, _ptr(grid), grid.size,
                                                     ctypes.byref(handle)))
        self._h = handle
        info = RunInfo()
        _check(lib.spo_run_get_info(self._h, ctypes.byref(info)))
        self.info = info

    def close(self) -> None:
        if getattr(self, "_h", None):
            host_library().spo_run_destroy(self._h)
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

    # shapes
    @property
    def n_zr(self) -> int:
        return self.info.nr * self.info.nz

    @property
    def n_xy(self) -> int:
        return self.info.nx * self.info.ny

    @property
    def scattering_size(self) -> int:
        return self.info.n_rot * self.n_xy * self.n_zr

    @property
    def specular_size(self) -> int:
        return self.info.nr * self.info.nz_specular

    @property
    def backend_pointer(self) -> int:
        return host_library().spo_run_backend(self._h)

    SEAM_CALLS = ("propagate_analysis_step", "kinetic_and_potential", "z_plane", "block_densities", "specular",
                  "wave_packet_on_host", "read_wave_packet", "site_column", "analyse", "s_matrix_sums",
                  "add_earlier_analysis_step", "energy_resolved_add")

    def call_counts(self) -> dict:
        """The calls the run has made across the backend seam so far (spo_run_call_counts)."""
        out = np.zeros(len(self.SEAM_CALLS), dtype=np.int64)
        _check(host_library().spo_run_call_counts(self._h, _ptr(out), len(out)))
        return dict(zip(self.SEAM_CALLS, (int(v) for v in out)))

    # time loop
    def steps(self, n: int) -> None:
        _check(host_library().spo_run_steps(self._h, n))

    def propagate(self) -> None:
        _check(host_library().spo_run_propagate(self._h))

    def kinetic_and_potential(self, half_step: bool, kinetic_only: bool) -> None:
        _check(host_library().spo_run_kinetic_and_potential(self._h, int(half_step), int(kinetic_only)))

    def analyse(self, t: float) -> np.ndarray:
        bk = np.zeros(self.info.n_states * self.info.n_v, dtype=np.complex128)
        _check(host_library().spo_run_analyse(self._h, t, _ptr(bk)))
        return bk

    # wave packet
    def get_wave(self):
        s = np.empty(self.scattering_size, dtype=np.complex128)
        p = np.empty(self.specular_size, dtype=np.complex128)
        _check(host_library().spo_run_get_wave(self._h, _ptr(s), _ptr(p)))
        return s, p

    def set_wave(self, scattering: np.ndarray, specular: np.ndarray) -> None:
        s = np.ascontiguousarray(scattering, dtype=np.complex128)
        p = np.ascontiguousarray(specular, dtype=np.complex128)
        if s.size != self.scattering_size or p.size != self.specular_size:
            raise ValueError("wave packet has the wrong size")
        _check(host_library().spo_run_set_wave(self._h, _ptr(s), _ptr(p)))

    def specular(self) -> np.ndarray:
        p = np.empty(self.specular_size, dtype=np.complex128)
        _check(host_library().spo_run_get_specular(self._h, _ptr(p)))
        return p

    def block_densities(self) -> np.ndarray:
        d = np.empty(self.n_zr, dtype=np.float64)
        _check(host_library().spo_run_block_densities(self._h, _ptr(d)))
        return d

    def z_plane(self, iz: int) -> np.ndarray:
        out = np.empty(self.info.n_rot * self.n_xy * self.info.nr, dtype=np.complex128)
        _check(host_library().spo_run_z_plane(self._h, iz, _ptr(out)))
        return out

    def site_column(self, ixy: int) -> np.ndarray:
        out = np.empty(self.n_zr * self.info.n_rot, dtype=np.complex128)
        _check(host_library().spo_run_site_column(self._h, ixy, _ptr(out)))
        return out

    # analysis
    def bk_records(self) -> np.ndarray:
        count = ctypes.c_int()
        _check(host_library().spo_run_bk_count(self._h, ctypes.byref(count)))
        out = np.zeros((count.value, self.info.n_states * self.info.n_v), dtype=np.complex128)
        for k in range(count.value):
            _check(host_library().spo_run_bk_record(self._h, k, _ptr(out[k])))
        return out

    def averages(self) -> np.ndarray:
        count = ctypes.c_int()
        _check(host_library().spo_run_averages(self._h, None, 0, ctypes.byref(count)))
        rows = np.zeros((max(count.value, 1), 5), dtype=np.float64)
        _check(host_library().spo_run_averages(self._h, _ptr(rows), count.value, ctypes.byref(count)))
        return rows[:count.value]

    def surface(self):
        fixed, index, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(host_library().spo_run_surface(self._h, ctypes.byref(fixed), ctypes.byref(index), ctypes.byref(z)))
        return bool(fixed.value), index.value, z.value

    def results(self, zana_convention: bool = False) -> dict:
        ne, nk = self.info.n_energies, self.info.n_states * self.info.n_v
        s = np.zeros((ne, nk), dtype=np.complex128)
        p = np.zeros((ne, nk), dtype=np.float64)
        pr = np.zeros(ne, dtype=np.float64)
        te = np.zeros(ne, dtype=np.float64)
        _check(host_library().spo_run_results(self._h, int(zana_convention), _ptr(s), _ptr(p), _ptr(pr), _ptr(te)))
        return {"s": s, "p": p, "p_reaction": pr, "total_energy": te}

    def energy_resolved(self):
        """(sums[site, energy, Z, r, rot], specular[energy, Z_specular, r]) of the energy-resolved sites, or None
        when EnergyResolvedSites is not set."""
        ns, ne = ctypes.c_int(0), ctypes.c_int(0)
        _check(host_library().spo_run_energy_resolved_shape(self._h, ctypes.byref(ns), ctypes.byref(ne)))
        if ns.value == 0:
            return None
        i = self.info
        sums = np.zeros((ns.value, ne.value, i.nz, i.nr, i.n_rot), dtype=np.complex128)
        spec = np.zeros((ne.value, i.nz_specular, i.nr), dtype=np.complex128)
        _check(host_library().spo_run_energy_resolved(self._h, _ptr(sums), _ptr(spec)))
        return sums, spec

    def s_sums(self) -> np.ndarray:
        out = np.zeros(self.info.n_energies * self.info.n_states * self.info.n_v, dtype=np.complex128)
        _check(host_library().spo_run_s_sums(self._h, _ptr(out)))
        return out

    def write_checkpoint(self, directory: str) -> None:
        """checkpoint.bin and bk_coefficients.bin of the run so far (the reference's native restart files); a run
        with {"NewCalc": False, "RestartDir": directory, "tStart": <this time>} continues from them."""
        _check(host_library().spo_run_write_checkpoint(self._h, os.fspath(directory).encode()))

    def kernel_timing(self, on: bool) -> None:
        """Starts (and clears) or stops per-kernel HIP-event timing on the backend's stream."""
        _check(host_library().spo_run_kernel_timing(self._h, int(on)))

    def kernel_times(self) -> dict:
        """{kernel name: (launches, total milliseconds)} since timing was enabled."""
        lib = host_library()
        count = ctypes.c_int()
        _check(lib.spo_run_kernel_times(self._h, None, None, None, 0, ctypes.byref(count)))
        n = count.value
        names = ctypes.create_string_buffer(48 * max(n, 1))
        launches = np.zeros(max(n, 1), dtype=np.int64)
        ms = np.zeros(max(n, 1), dtype=np.float64)
        _check(lib.spo_run_kernel_times(self._h, names, _ptr(launches), _ptr(ms), n, ctypes.byref(count)))
        raw = names.raw
        out = {}
        for k in range(n):
            name = raw[48 * k:48 * (k + 1)].split(b"\0", 1)[0].decode()
            out[name] = (int(launches[k]), float(ms[k]))
        return out

    def array(self, name: str) -> np.ndarray:
        lib = host_library()
        count = ctypes.c_size_t()
        _check(lib.spo_run_array(self._h, name.encode(), None, 0, ctypes.byref(count)))
        is_complex = name == "initial.specular"
        out = np.zeros(count.value * (2 if is_complex else 1), dtype=np.float64)
        _check(lib.spo_run_array(self._h, name.encode(), _ptr(out), out.size, ctypes.byref(count)))
        return out.view(np.complex128) if is_complex else out
