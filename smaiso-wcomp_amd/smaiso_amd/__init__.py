"""ctypes binding of libsmaiso.so (include/smaiso.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  This
module only marshals arguments; it never computes a result itself, and it
raises if the shared library is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
from dataclasses import dataclass, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsmaiso.so")
RUN_PATH = os.path.join(PKG_ROOT, "bin", "smaiso_run")  # the command-line driver
HEADER_PATH = os.path.join(REPO_ROOT, "include", "smaiso.h")

OK, E_INVALID, E_HIP, E_NOMEM, E_STATE, E_NONFINITE = 0, -1, -2, -3, -4, -5
FP64, FP32_MIXED = 0, 1
REDFT10, REDFT01, RODFT10, RODFT01 = 0, 1, 2, 3

FIELDS = {
    "psi": 0, "rho": 1, "amp": 2, "vx": 3, "vy": 4, "vz": 5,
    "vsolx": 6, "vsoly": 7, "vsolz": 8, "virrx": 9, "virry": 10, "virrz": 11,
    "divv": 12, "lpsi": 13, "mu": 14, "energy": 15, "psiq": 16, "psiq_old": 17,
    "nlq": 18, "f_rho": 19, "divf": 20, "gpx": 21, "gpy": 22, "gpz": 23,
    "rdx": 24, "rdy": 25, "rdz": 26, "lap_rl": 27, "psi_old": 28,
}
TILE_VARIANTS = 4
# stages of the FP32-mixed step that may compute in FP32 (SMAISO_STAGE_*)
STAGES = {"grad_psi": 1 << 0, "amp": 1 << 1, "lpsi": 1 << 2, "laprhol": 1 << 3, "forces": 1 << 4,
          "divv": 1 << 5, "divf": 1 << 6, "velocity": 1 << 7, "nl": 1 << 8, "velsolve": 1 << 9,
          "grad_rho": 1 << 10}


def fp32_stages_default() -> int:
    return int(lib().smaiso_fp32_stages_default())


KERNEL_CLASSES = {"r2r_z": 0, "r2r_y": 1, "r2r_x": 2, "spectral": 3, "pointwise": 4, "reduce": 5, "r2r_axis": 6}


class SmaisoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"smaiso error {code}: {msg}")
        self.code = code


class CParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dt", C.c_double),
        ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
        ("epsilon", C.c_double), ("q0", C.c_double), ("nu", C.c_double),
        ("lambda_", C.c_double), ("zeta", C.c_double), ("rho0", C.c_double),
        ("rho_m", C.c_double),
        ("divv_switch", C.c_int32), ("load_mode", C.c_int32),
        ("precision", C.c_int32), ("slab_n0", C.c_int32),
    ]


class CIC(C.Structure):
    _fields_ = [
        ("amplitude", C.c_double), ("h0", C.c_double), ("width", C.c_double),
        ("rho_iso", C.c_double), ("rho_sm", C.c_double), ("undulation", C.c_double),
        ("noise", C.c_double), ("seed", C.c_uint64),
    ]


_lib = None


def lib():
    """Load libsmaiso.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.smaiso_abi_version.restype = C.c_int
        L.smaiso_params_default.argtypes = [C.POINTER(CParams)]
        L.smaiso_params_default.restype = None
        L.smaiso_create.argtypes = [C.POINTER(CParams), C.POINTER(vp)]
        L.smaiso_destroy.argtypes = [vp]
        L.smaiso_last_error.argtypes = [vp]
        L.smaiso_last_error.restype = C.c_char_p
        L.smaiso_set_state.argtypes = [vp, dp, dp]
        L.smaiso_set_state_device.argtypes = [vp, vp, vp]
        L.smaiso_step.argtypes = [vp, C.c_int32]
        L.smaiso_step_async.argtypes = [vp, C.c_int32]
        L.smaiso_synchronize.argtypes = [vp]
        L.smaiso_step_count.argtypes = [vp]
        L.smaiso_step_count.restype = C.c_int64
        L.smaiso_set_fp32_stages.argtypes = [vp, C.c_uint32]
        L.smaiso_fp32_stages_default.argtypes = []
        L.smaiso_fp32_stages_default.restype = C.c_uint32
        L.smaiso_r2r3d_noalloc.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, dp, dp,
                                           C.c_int32]
        L.smaiso_get_field.argtypes = [vp, C.c_int32, dp]
        L.smaiso_get_field_device.argtypes = [vp, C.c_int32, vp]
        L.smaiso_diagnostics.argtypes = [vp, dp, dp, dp]
        L.smaiso_r2r3d.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, dp, dp, C.c_int32]
        L.smaiso_r2r3d_device.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, vp, vp,
                                          C.c_int32, vp]
        L.smaiso_make_interface_ic.argtypes = [C.POINTER(CParams), C.POINTER(CIC), dp, dp, dp]
        L.smaiso_timing_enable.argtypes = [vp, C.c_int32]
        L.smaiso_timing_get.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int64), dp]
        L.smaiso_set_graph.argtypes = [vp, C.c_int32]
        L.smaiso_set_adjoint_store.argtypes = [vp, C.c_int32]
        L.smaiso_get_slice.argtypes = [vp, C.c_int32, C.c_int32, dp]
        L.smaiso_surface_height.argtypes = [vp, C.c_double, dp]
        L.smaiso_save_checkpoint.argtypes = [vp, C.c_char_p]
        L.smaiso_load_checkpoint.argtypes = [vp, C.c_char_p]
        L.smaiso_read_psi_text.argtypes = [C.POINTER(C.c_char_p), C.c_int32, C.c_int64, dp]
        L.smaiso_autotune.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_int32)]
        L.smaiso_ensemble_autotune.argtypes = [vp]
        L.smaiso_r2r3d_tiles.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, vp, vp,
                                         C.c_int32, C.c_int32, vp]
        L.smaiso_adjoint.argtypes = [vp, dp, dp, dp, C.c_int32, C.c_int32, dp, dp, dp, dp, dp]
        L.smaiso_interface_ic_vjp.argtypes = [C.POINTER(CParams), C.POINTER(CIC), dp, dp, dp, dp]
        L.smaiso_ensemble_create.argtypes = [C.POINTER(CParams), C.c_int32, C.POINTER(vp)]
        L.smaiso_ensemble_destroy.argtypes = [vp]
        L.smaiso_ensemble_last_error.argtypes = [vp]
        L.smaiso_ensemble_last_error.restype = C.c_char_p
        L.smaiso_ensemble_groups.argtypes = [vp]
        L.smaiso_ensemble_set_state.argtypes = [vp, C.c_int32, dp, dp]
        L.smaiso_ensemble_set_graph.argtypes = [vp, C.c_int32]
        L.smaiso_ensemble_step.argtypes = [vp, C.c_int32]
        L.smaiso_ensemble_step_count.argtypes = [vp, C.c_int32]
        L.smaiso_ensemble_step_count.restype = C.c_int64
        L.smaiso_ensemble_get_field.argtypes = [vp, C.c_int32, C.c_int32, dp]
        L.smaiso_ensemble_diagnostics.argtypes = [vp, C.c_int32, dp, dp, dp]
        for name in ("smaiso_create", "smaiso_destroy", "smaiso_set_state", "smaiso_set_state_device",
                     "smaiso_step", "smaiso_step_async", "smaiso_synchronize", "smaiso_get_field",
                     "smaiso_get_field_device", "smaiso_diagnostics", "smaiso_r2r3d",
                     "smaiso_r2r3d_device", "smaiso_make_interface_ic", "smaiso_timing_enable",
                     "smaiso_timing_get", "smaiso_set_graph", "smaiso_ensemble_create",
                     "smaiso_ensemble_destroy", "smaiso_ensemble_groups", "smaiso_ensemble_set_state",
                     "smaiso_ensemble_set_graph", "smaiso_ensemble_step", "smaiso_ensemble_get_field",
                     "smaiso_ensemble_diagnostics", "smaiso_adjoint", "smaiso_set_adjoint_store",
                     "smaiso_interface_ic_vjp",
                     "smaiso_autotune", "smaiso_ensemble_autotune", "smaiso_r2r3d_tiles",
                     "smaiso_get_slice", "smaiso_surface_height", "smaiso_save_checkpoint",
                     "smaiso_load_checkpoint", "smaiso_read_psi_text"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


_hip = None


def _hip_runtime():
    """The HIP runtime libsmaiso.so is linked against (libamdhip64.so.7)."""
    global _hip
    if _hip is None:
        lib()
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipSetDevice.argtypes = [C.c_int]
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
    return _hip


def set_device(dev: int):
    """Select the HIP device for subsequent handles (default 0)."""
    rc = _hip_runtime().hipSetDevice(int(dev))
    if rc != 0:
        raise SmaisoError(E_HIP, f"hipSetDevice({dev}) failed with {rc}")


def device_count() -> int:
    n = C.c_int(0)
    rc = _hip_runtime().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def device_synchronize():
    rc = _hip_runtime().hipDeviceSynchronize()
    if rc != 0:
        raise SmaisoError(E_HIP, f"hipDeviceSynchronize failed with {rc}")


def header_symbols(path: str = HEADER_PATH):
    """Names of the functions declared in include/smaiso.h."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(smaiso_\w+)\s*\(", text, re.M)))


@dataclass
class Params:
    """Mirror of smaiso_params; defaults are the reference's constants (S:206-283)."""
    nx: int = 512
    ny: int = 512
    nz: int = 512
    dx: float = 2.0 * math.pi / 8.0
    dy: float = 2.0 * math.pi / 8.0
    dz: float = 2.0 * math.pi / 8.0
    dt: float = 0.001
    alpha: float = 1.0
    beta: float = 2.0
    gamma: float = 1.0
    epsilon: float = -0.005
    q0: float = 1.0
    nu: float = 1.0
    lam: float = 1.0
    zeta: float = 1.0
    rho0: float = 0.01
    rho_m: float = 0.0          # <= 0 -> rho_s/2
    divv_switch: int = 50
    load_mode: int = 0
    precision: int = FP64
    slab_n0: int = 0

    def to_c(self) -> CParams:
        c = CParams()
        for f in fields(self):
            setattr(c, "lambda_" if f.name == "lam" else f.name, getattr(self, f.name))
        return c

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


def _check(rc: int, h=None):
    if rc != OK:
        msg = lib().smaiso_last_error(h).decode()
        raise SmaisoError(rc, msg)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Stepper:
    """One simulation on the current HIP device (smaiso_create .. smaiso_destroy)."""

    def __init__(self, params: Params):
        self.params = params
        self._cp = params.to_c()
        h = C.c_void_p()
        _check(lib().smaiso_create(C.byref(self._cp), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().smaiso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, psi: np.ndarray, rho: np.ndarray):
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.params.shape)
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(self.params.shape)
        _check(lib().smaiso_set_state(self.h, _dp(psi), _dp(rho)), self.h)

    def set_state_device(self, psi_ptr: int, rho_ptr: int):
        _check(lib().smaiso_set_state_device(self.h, C.c_void_p(psi_ptr), C.c_void_p(rho_ptr)), self.h)

    def step(self, n: int = 1):
        _check(lib().smaiso_step(self.h, int(n)), self.h)

    def step_async(self, n: int = 1):
        _check(lib().smaiso_step_async(self.h, int(n)), self.h)

    def synchronize(self):
        _check(lib().smaiso_synchronize(self.h), self.h)

    @property
    def nloop(self) -> int:
        return int(lib().smaiso_step_count(self.h))

    def set_fp32_stages(self, mask: int):
        """FP32-arithmetic stages of an FP32-mixed handle (STAGES bits)."""
        _check(lib().smaiso_set_fp32_stages(self.h, int(mask)), self.h)

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.params.shape, dtype=np.float64)
        _check(lib().smaiso_get_field(self.h, FIELDS[name], _dp(out)), self.h)
        return out

    def field_device(self, name: str, out_ptr: int):
        _check(lib().smaiso_get_field_device(self.h, FIELDS[name], C.c_void_p(out_ptr)), self.h)

    def diagnostics(self, allow_nonfinite: bool = False):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        rc = lib().smaiso_diagnostics(self.h, C.byref(a), C.byref(b), C.byref(c))
        if not (allow_nonfinite and rc == E_NONFINITE):
            _check(rc, self.h)
        return a.value, b.value, c.value

    def slice(self, name: str, j: int | None = None) -> np.ndarray:
        """x-z slice of a field at y index j (default Ny/2), the saveMid layout."""
        j = self.params.ny // 2 if j is None else j
        out = np.empty(self.params.nx * self.params.nz, dtype=np.float64)
        _check(lib().smaiso_get_slice(self.h, FIELDS[name], int(j), _dp(out)), self.h)
        return out

    def surface_height(self, threshold: float, prev: np.ndarray | None = None) -> np.ndarray:
        """Surface tracker (smaiso_surface_height); prev holds the stale values."""
        surf = np.zeros(self.params.nx * self.params.ny) if prev is None else \
            np.array(prev, dtype=np.float64).reshape(-1).copy()
        _check(lib().smaiso_surface_height(self.h, float(threshold), _dp(surf)), self.h)
        return surf

    def save_checkpoint(self, path: str):
        _check(lib().smaiso_save_checkpoint(self.h, str(path).encode()), self.h)

    def load_checkpoint(self, path: str):
        _check(lib().smaiso_load_checkpoint(self.h, str(path).encode()), self.h)

    def set_graph(self, enable: bool):
        """Replay pairs of steady steps from one captured hipGraph (smaiso_set_graph)."""
        _check(lib().smaiso_set_graph(self.h, 1 if enable else 0), self.h)

    def autotune(self):
        """Pick tile variants for this handle's plain passes (smaiso_autotune).
        Returns (ms[dtype][axis][variant], chosen[dtype][axis]), dtype 0 = FP64, 1 = FP32."""
        ms = (C.c_float * (2 * 3 * TILE_VARIANTS))()
        ch = (C.c_int32 * 6)()
        _check(lib().smaiso_autotune(self.h, ms, ch), self.h)
        return (np.array(ms[:], dtype=np.float64).reshape(2, 3, TILE_VARIANTS),
                np.array(ch[:], dtype=np.int32).reshape(2, 3))

    def adjoint(self, psi0: np.ndarray, rho0: np.ndarray, target: np.ndarray, nsteps: int,
                checkpoint_every: int = 50) -> "AdjointResult":
        """Loss sum (psi(T) - target)^2 after nsteps from set_state(psi0, rho0)
        and its gradients (smaiso_adjoint).  Consumes the state."""
        sh = self.params.shape
        psi0 = np.ascontiguousarray(psi0, dtype=np.float64).reshape(sh)
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64).reshape(sh)
        target = np.ascontiguousarray(target, dtype=np.float64).reshape(sh)
        dpsi, drho = np.empty(sh), np.empty(sh)
        loss, dz, dn = C.c_double(), C.c_double(), C.c_double()
        _check(lib().smaiso_adjoint(self.h, _dp(psi0), _dp(rho0), _dp(target), int(nsteps),
                                    int(checkpoint_every), C.byref(loss), _dp(dpsi), _dp(drho),
                                    C.byref(dz), C.byref(dn)), self.h)
        return AdjointResult(loss.value, dpsi, drho, dz.value, dn.value)

    def set_adjoint_store(self, enable: bool):
        """Keep each recomputed step's intermediates in the adjoint (default)
        or recompute every step again before its transpose
        (smaiso_set_adjoint_store)."""
        _check(lib().smaiso_set_adjoint_store(self.h, 1 if enable else 0), self.h)

    def timing(self, enable: bool):
        _check(lib().smaiso_timing_enable(self.h, 1 if enable else 0), self.h)

    def timing_get(self, kc: str):
        ms, n, by = C.c_double(), C.c_int64(), C.c_double()
        _check(lib().smaiso_timing_get(self.h, KERNEL_CLASSES[kc], C.byref(ms), C.byref(n), C.byref(by)), self.h)
        return ms.value, n.value, by.value


@dataclass
class AdjointResult:
    loss: float
    psi0: np.ndarray   # dL/dpsi0
    rho0: np.ndarray   # dL/drho0
    zeta: float        # dL/dzeta
    nu: float          # dL/dnu


def _echeck(rc: int, e=None):
    if rc != OK:
        raise SmaisoError(rc, lib().smaiso_ensemble_last_error(e).decode())


class Ensemble:
    """Independent simulations stepped together (smaiso_ensemble_*).

    Members with the same grid and shared scalars run as one batch; nu,
    lambda, zeta and rho0 may differ per member.
    """

    def __init__(self, members):
        self.params = list(members)
        arr = (CParams * len(self.params))(*[p.to_c() for p in self.params])
        e = C.c_void_p()
        _echeck(lib().smaiso_ensemble_create(arr, len(self.params), C.byref(e)))
        self.e = e

    def close(self):
        if getattr(self, "e", None):
            lib().smaiso_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.params)

    @property
    def groups(self) -> int:
        return int(lib().smaiso_ensemble_groups(self.e))

    def _shape(self, member: int):
        # out-of-range members are reported by the library itself
        return self.params[member].shape if 0 <= member < len(self.params) else (-1,)

    def set_state(self, member: int, psi: np.ndarray, rho: np.ndarray):
        shape = self._shape(member)
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(shape)
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(shape)
        _echeck(lib().smaiso_ensemble_set_state(self.e, int(member), _dp(psi), _dp(rho)), self.e)

    def set_graph(self, enable: bool):
        _echeck(lib().smaiso_ensemble_set_graph(self.e, 1 if enable else 0), self.e)

    def autotune(self):
        _echeck(lib().smaiso_ensemble_autotune(self.e), self.e)

    def step(self, n: int = 1):
        _echeck(lib().smaiso_ensemble_step(self.e, int(n)), self.e)

    def nloop(self, member: int = 0) -> int:
        return int(lib().smaiso_ensemble_step_count(self.e, int(member)))

    def field(self, member: int, name: str) -> np.ndarray:
        out = np.empty(self._shape(member) if 0 <= member < len(self.params) else 1, dtype=np.float64)
        _echeck(lib().smaiso_ensemble_get_field(self.e, int(member), FIELDS[name], _dp(out)), self.e)
        return out

    def diagnostics(self, member: int, allow_nonfinite: bool = False):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        rc = lib().smaiso_ensemble_diagnostics(self.e, int(member), C.byref(a), C.byref(b), C.byref(c))
        if not (allow_nonfinite and rc == E_NONFINITE):
            _echeck(rc, self.e)
        return a.value, b.value, c.value


def r2r3d(x: np.ndarray, kinds, dtype: int = FP64) -> np.ndarray:
    """FFTW-semantics 3-D r2r on the GPU (smaiso_r2r3d)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.ndim == 3
    out = np.empty_like(x)
    k = (C.c_int32 * 3)(*[int(v) for v in kinds])
    _check(lib().smaiso_r2r3d(k, x.shape[0], x.shape[1], x.shape[2], _dp(x), _dp(out), int(dtype)))
    return out


def r2r3d_device(in_ptr: int, out_ptr: int, shape, kinds, dtype: int = FP64, stream: int = 0):
    k = (C.c_int32 * 3)(*[int(v) for v in kinds])
    _check(lib().smaiso_r2r3d_device(k, shape[0], shape[1], shape[2], C.c_void_p(in_ptr),
                                     C.c_void_p(out_ptr), int(dtype), C.c_void_p(stream)))


def r2r3d_tiles(in_ptr: int, out_ptr: int, shape, kinds, dtype: int, variant: int, stream: int = 0):
    """r2r3d_device with one tile variant forced on every axis (smaiso_r2r3d_tiles)."""
    k = (C.c_int32 * 3)(*[int(v) for v in kinds])
    _check(lib().smaiso_r2r3d_tiles(k, shape[0], shape[1], shape[2], C.c_void_p(in_ptr),
                                    C.c_void_p(out_ptr), int(dtype), int(variant), C.c_void_p(stream)))


@dataclass
class InterfaceIC:
    amplitude: float = 1.0
    h0: float = -1.0            # < 0 -> Lz/2
    width: float = 2.0
    rho_iso: float = 0.9
    rho_sm: float = 1.0
    undulation: float = 0.0
    noise: float = 1e-3
    seed: int = 1

    def to_c(self) -> CIC:
        c = CIC()
        for f in fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c


def make_interface_ic(params: Params, ic: InterfaceIC, height: np.ndarray | None = None):
    """Builder-defined interface IC (smaiso_make_interface_ic, host code)."""
    psi = np.empty(params.shape, dtype=np.float64)
    rho = np.empty(params.shape, dtype=np.float64)
    cp, ci = params.to_c(), ic.to_c()
    hp = None
    if height is not None:
        height = np.ascontiguousarray(height, dtype=np.float64).reshape(params.nx, params.ny)
        hp = _dp(height)
    _check(lib().smaiso_make_interface_ic(C.byref(cp), C.byref(ci), hp, _dp(psi), _dp(rho)))
    return psi, rho


def interface_ic_vjp(params: Params, ic: InterfaceIC, dpsi: np.ndarray, drho: np.ndarray,
                     height: np.ndarray | None = None) -> np.ndarray:
    """dL/dh(x, y) through the interface IC (smaiso_interface_ic_vjp, host code)."""
    cp, ci = params.to_c(), ic.to_c()
    dpsi = np.ascontiguousarray(dpsi, dtype=np.float64).reshape(params.shape)
    drho = np.ascontiguousarray(drho, dtype=np.float64).reshape(params.shape)
    hp = None
    if height is not None:
        height = np.ascontiguousarray(height, dtype=np.float64).reshape(params.nx, params.ny)
        hp = _dp(height)
    out = np.empty((params.nx, params.ny))
    _check(lib().smaiso_interface_ic_vjp(C.byref(cp), C.byref(ci), hp, _dp(dpsi), _dp(drho), _dp(out)))
    return out


def read_psi_text(paths, count: int) -> np.ndarray:
    """The reference's psi<rank>.dat restart files, in order (smaiso_read_psi_text)."""
    arr = (C.c_char_p * len(paths))(*[str(q).encode() for q in paths])
    out = np.empty(int(count), dtype=np.float64)
    _check(lib().smaiso_read_psi_text(arr, len(paths), int(count), _dp(out)))
    return out


def create_error(params: Params) -> tuple[int, str]:
    """Return (status, message) of smaiso_create without keeping a handle."""
    cp = params.to_c()
    h = C.c_void_p()
    rc = lib().smaiso_create(C.byref(cp), C.byref(h))
    if rc == OK:
        lib().smaiso_destroy(h)
        return rc, ""
    return rc, lib().smaiso_last_error(None).decode()
