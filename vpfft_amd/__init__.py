"""vpfft_amd -- MI355X VPFFT strain-step engine, Python host mirror.

The compute path is libvpfft.so (hand-written gfx950 HIP kernels behind a C
ABI, include/vpfft.h).  This module is a thin ctypes binding that mirrors the
reference's host interface, VPFFT::DataStructures::MaterialGrid
(Src/MaterialGrid.h:59-577), with the same method names and argument
meanings, so a driver written against the reference reads the same here.

There is no CPU fallback: if libvpfft.so is missing, or no GPU is present,
the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpfft.so")

NSLIP = 12
FP64, MIXED = 0, 1

FIELD_ORIENTATION = 0
FIELD_CRSS = 1
FIELD_STRESS = 2
FIELD_LAGRANGE = 3
FIELD_POLARIZATION = 4
FIELD_STRAIN_RATE_VARIATION = 5
FIELD_ACC_SHEAR = 6
FIELD_SCHMID = 7
FIELD_STRAIN_RATE = 8

ERRORS = {
    -1: "VPFFT_E_INVALID",
    -2: "VPFFT_E_NOMEM",
    -3: "VPFFT_E_DEVICE",
    -4: "VPFFT_E_DEGENERATE",
    -5: "VPFFT_E_NONFINITE",
    -6: "VPFFT_E_UNSUPPORTED",
    -7: "VPFFT_E_STATE",
}


class VPFFTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class vpfft_grid_desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("precision", C.c_int)]


class vpfft_material(C.Structure):
    _fields_ = [
        ("gamma_dot_base", C.c_double * NSLIP),
        ("rate_sensitivity", C.c_int * NSLIP),
        ("tau0", C.c_double * NSLIP),
        ("tau1", C.c_double * NSLIP),
        ("theta0", C.c_double * NSLIP),
        ("theta1", C.c_double * NSLIP),
        ("hardening", C.c_double * (NSLIP * NSLIP)),
    ]


class vpfft_step_info(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("status", C.c_int),
        ("precision", C.c_int),
        ("err_stress", C.c_double),
        ("err_strain_rate", C.c_double),
        ("avg_stress", C.c_double * 5),
        ("avg_strain_rate", C.c_double * 5),
        ("L0", C.c_double * 25),
        ("max_nr_error", C.c_double),
        ("nr_steps_last", C.c_double),
        ("nr_steps_total", C.c_double),
        ("ms_prologue", C.c_double),
        ("ms_loop", C.c_double),
        ("ms_epilogue", C.c_double),
        ("max_nr_steps", C.c_double),
        ("nr_wave_steps_total", C.c_double),
        ("taylor_steps_total", C.c_double),
        ("err_ratio_5", C.c_double),
    ]

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class vpfft_deck(C.Structure):
    """include/vpfft_io.h: Utilities::ConfigFile, Src/Config.h:21-52."""
    _fields_ = [
        ("num_x", C.c_int), ("num_y", C.c_int), ("num_z", C.c_int),
        ("input_strain_rate", (C.c_double * 3) * 3),
        ("max_nr_iter", C.c_int),
        ("nr_tolerance", C.c_double),
        ("time_step", C.c_double),
        ("num_time_iter", C.c_int),
        ("init_method", C.c_int),
        ("microstructure_file", C.c_char * 1024),
        ("crss", C.c_double * NSLIP),
        ("gamma_dot_base", C.c_double * NSLIP),
        ("rate_sensitivity", C.c_int * NSLIP),
    ]


_lib = None

# every symbol include/vpfft.h declares
EXPORTED = [
    "vpfft_grid_create", "vpfft_grid_destroy", "vpfft_grid_set_voxel_length",
    "vpfft_grid_get_voxel_length", "vpfft_grid_set_precision", "vpfft_grid_upload",
    "vpfft_grid_upload_field", "vpfft_grid_download", "vpfft_run_strain_step",
    "vpfft_run_vpfft", "vpfft_strain_rate_evolution", "vpfft_solve_constrained_batch",
    "vpfft_solve_taylor_batch", "vpfft_generate_voronoi", "vpfft_last_error", "vpfft_version",
    "vpfft_grid_set_profiling", "vpfft_grid_kernel_times", "vpfft_kernel_name",
    "vpfft_batch_create", "vpfft_batch_destroy", "vpfft_batch_set_tuning", "vpfft_batch_block_shape",
    "vpfft_batch_radix_plan",
    "vpfft_batch_run_strain_step", "vpfft_grid_set_block_shape", "vpfft_grid_bytes", "vpfft_grid_save_state",
    "vpfft_grid_restore_state", "vpfft_strain_rate_evolution_noalloc",
    "vpfft_solvers_taylor", "vpfft_solvers_constrained", "vpfft_solvers_sachs", "vpfft_solvers_apply",
    "vpfft_solvers_shear_rate", "vpfft_solvers_hardening",
    # include/vpfft_io.h
    "vpfft_deck_read", "vpfft_deck_parse", "vpfft_deck_material", "vpfft_deck_strain_rate",
    "vpfft_two_grain_init", "vpfft_euler_deg_to_matrix", "vpfft_read_euler_file", "vpfft_write_euler_file",
    "vpfft_write_material_file", "vpfft_read_material_file", "vpfft_run_deck",
    # include/vpfft_serve.h
    "vpfft_serve",
]


def load_library(path: Optional[str] = None):
    """Load libvpfft.so (raises when it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} not found: build it with `python -m vpfft_amd.build`")
    lib = C.CDLL(p)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.vpfft_grid_create.argtypes = [C.POINTER(vpfft_grid_desc), C.POINTER(vpfft_material), C.POINTER(vp)]
    lib.vpfft_grid_destroy.argtypes = [vp]
    lib.vpfft_grid_set_voxel_length.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    lib.vpfft_grid_get_voxel_length.argtypes = [vp, dp]
    lib.vpfft_grid_set_precision.argtypes = [vp, C.c_int]
    lib.vpfft_grid_upload.argtypes = [vp, dp, dp, C.c_double, dp]
    lib.vpfft_grid_upload_field.argtypes = [vp, C.c_int, dp]
    lib.vpfft_grid_download.argtypes = [vp, C.c_int, dp]
    lib.vpfft_run_strain_step.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, dp, C.POINTER(vpfft_step_info)]
    lib.vpfft_run_vpfft.argtypes = [vp, dp, C.c_int, C.c_int, C.c_double, C.c_double, dp,
                                    C.POINTER(vpfft_step_info), C.c_char_p]
    lib.vpfft_strain_rate_evolution.argtypes = [vp, dp, dp, dp]
    lib.vpfft_strain_rate_evolution_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_void_p,
                                                        C.c_void_p, C.c_void_p]
    lib.vpfft_solve_constrained_batch.argtypes = [C.c_long, C.POINTER(vpfft_material), dp, dp, dp, dp, dp, dp, dp,
                                                  dp, dp, C.POINTER(C.c_int)]
    lib.vpfft_solve_taylor_batch.argtypes = [C.c_long, C.POINTER(vpfft_material), dp, dp, dp, dp, dp]
    mp = C.POINTER(vpfft_material)
    lib.vpfft_solvers_taylor.argtypes = [C.c_long, mp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int, dp]
    lib.vpfft_solvers_constrained.argtypes = [C.c_long, mp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double,
                                              C.c_int, dp, dp]
    lib.vpfft_solvers_sachs.argtypes = [C.c_long, dp, dp, dp, dp]
    lib.vpfft_solvers_apply.argtypes = [C.c_long, mp, dp, dp, dp, dp]
    lib.vpfft_solvers_shear_rate.argtypes = [C.c_long, mp, dp, dp, dp, dp, dp]
    lib.vpfft_solvers_hardening.argtypes = [C.c_long, mp, dp, dp, dp, C.c_double, dp, dp]
    lib.vpfft_generate_voronoi.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, dp,
                                           C.POINTER(C.c_int32), dp]
    lib.vpfft_grid_set_profiling.argtypes = [vp, C.c_int]
    lib.vpfft_grid_kernel_times.argtypes = [vp, dp, C.POINTER(C.c_long), C.c_int]
    lib.vpfft_kernel_name.argtypes = [C.c_int]
    ip = C.POINTER(C.c_int32)
    cp = C.c_char_p
    lib.vpfft_batch_create.argtypes = [C.POINTER(vp)]
    lib.vpfft_batch_destroy.argtypes = [vp]
    lib.vpfft_batch_set_tuning.argtypes = [vp, C.c_int]
    lib.vpfft_batch_block_shape.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    lib.vpfft_batch_radix_plan.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.vpfft_batch_run_strain_step.argtypes = [vp, C.POINTER(vp), C.c_int, dp, C.c_int, dp, C.c_double,
                                                dp, C.POINTER(vpfft_step_info)]
    lib.vpfft_grid_bytes.argtypes = [C.POINTER(vpfft_grid_desc), C.POINTER(C.c_uint64)]
    lib.vpfft_grid_save_state.argtypes = [vp]
    lib.vpfft_grid_set_block_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.vpfft_grid_restore_state.argtypes = [vp]
    lib.vpfft_deck_read.argtypes = [cp, C.POINTER(vpfft_deck)]
    lib.vpfft_deck_parse.argtypes = [cp, C.POINTER(vpfft_deck)]
    lib.vpfft_deck_material.argtypes = [C.POINTER(vpfft_deck), C.POINTER(vpfft_material)]
    lib.vpfft_deck_strain_rate.argtypes = [C.POINTER(vpfft_deck), dp]
    lib.vpfft_two_grain_init.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    lib.vpfft_euler_deg_to_matrix.argtypes = [dp, C.c_long, dp]
    lib.vpfft_read_euler_file.argtypes = [cp, C.c_int, C.c_int, C.c_int, dp, dp, ip, ip]
    lib.vpfft_write_euler_file.argtypes = [cp, C.c_int, C.c_int, C.c_int, dp, ip, ip]
    lib.vpfft_write_material_file.argtypes = [cp, C.c_int, C.c_int, C.c_int, dp, dp]
    lib.vpfft_read_material_file.argtypes = [cp, C.c_int, C.c_int, C.c_int, dp, dp]
    lib.vpfft_run_deck.argtypes = [cp, C.c_int, cp, cp, C.POINTER(vpfft_step_info), C.c_int]
    lib.vpfft_kernel_name.restype = C.c_char_p
    lib.vpfft_last_error.restype = C.c_char_p
    lib.vpfft_version.restype = C.c_char_p
    for name in EXPORTED:
        if name not in ("vpfft_last_error", "vpfft_version", "vpfft_kernel_name"):
            getattr(lib, name).restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = load_library().vpfft_last_error().decode(errors="replace")
        raise VPFFTError(rc, msg)


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


# ---------------------------------------------------------------------------
# material
# ---------------------------------------------------------------------------
@dataclass
class Material:
    """Slip-system constants (the reference hard-codes them in
    MaterialGrid::InitializeMemory, Src/MaterialGrid.cpp:322-336)."""
    gamma_dot_base: Sequence[float] = field(default_factory=lambda: [1.0] * NSLIP)
    rate_sensitivity: Sequence[int] = field(default_factory=lambda: [10] * NSLIP)
    tau0: Sequence[float] = field(default_factory=lambda: [11.0] * NSLIP)
    tau1: Sequence[float] = field(default_factory=lambda: [11.0] * NSLIP)
    theta0: Sequence[float] = field(default_factory=lambda: [430.0] * NSLIP)
    theta1: Sequence[float] = field(default_factory=lambda: [110.0] * NSLIP)
    hardening: np.ndarray = field(default_factory=lambda: np.ones((NSLIP, NSLIP)))

    @classmethod
    def uniform(cls, gamma_dot_base: float = 1.0, rate_sensitivity: int = 10, **kw) -> "Material":
        return cls(gamma_dot_base=[float(gamma_dot_base)] * NSLIP,
                   rate_sensitivity=[int(rate_sensitivity)] * NSLIP, **kw)

    def to_c(self) -> vpfft_material:
        m = vpfft_material()
        for i in range(NSLIP):
            m.gamma_dot_base[i] = float(self.gamma_dot_base[i])
            m.rate_sensitivity[i] = int(self.rate_sensitivity[i])
            m.tau0[i] = float(self.tau0[i])
            m.tau1[i] = float(self.tau1[i])
            m.theta0[i] = float(self.theta0[i])
            m.theta1[i] = float(self.theta1[i])
        h = np.asarray(self.hardening, dtype=np.float64).reshape(NSLIP * NSLIP)
        for i in range(NSLIP * NSLIP):
            m.hardening[i] = float(h[i])
        return m


# ---------------------------------------------------------------------------
# deviatoric basis helpers (Src/LinearAlgebra.h:207-247, 307-360)
# ---------------------------------------------------------------------------
def eigen_basis() -> np.ndarray:
    s2, s6, s3 = 1.0 / np.sqrt(2.0), 1.0 / np.sqrt(6.0), 1.0 / np.sqrt(3.0)
    B = np.zeros((6, 3, 3))
    B[0] = s2 * np.array([[-1, 0, 0], [0, 1, 0], [0, 0, 0]])
    B[1] = s6 * np.array([[-1, 0, 0], [0, -1, 0], [0, 0, 2]])
    B[2] = s2 * np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]])
    B[3] = s2 * np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]])
    B[4] = s2 * np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]])
    B[5] = s3 * np.eye(3)
    return B


def EigenRep(matrix) -> np.ndarray:
    """EigenRep(const SMatrix3x3&): the 5 deviatoric components of a 3x3 tensor."""
    M = np.asarray(matrix, dtype=np.float64).reshape(3, 3)
    B = eigen_basis()
    return np.array([np.sum(M * B[a]) for a in range(5)])


def ToMatrixRep(v) -> np.ndarray:
    """EigenRep::ToMatrixRep: 5-vector -> symmetric traceless 3x3."""
    B = eigen_basis()
    v = np.asarray(v, dtype=np.float64)
    return sum(v[a] * B[a] for a in range(5))


def generate_voronoi(nx: int, ny: int, nz: int, num_grains: int, seed: int):
    """Seeded periodic Voronoi polycrystal: (orientation N x 9, grain N, euler G x 3)."""
    lib = load_library()
    n = nx * ny * nz
    R = np.empty((n, 9))
    grain = np.empty(n, dtype=np.int32)
    eul = np.empty((num_grains, 3))
    _check(lib.vpfft_generate_voronoi(nx, ny, nz, num_grains, C.c_uint64(seed), _dptr(R),
                                      grain.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(eul)))
    return R, grain, eul


# ---------------------------------------------------------------------------
# MaterialGrid mirror
# ---------------------------------------------------------------------------
class MaterialGrid:
    """GPU strain-step engine with the reference's MaterialGrid interface.

    MaterialGrid(DimX, DimY, DimZ)           Src/MaterialGrid.h:72-78
    SetVoxelLength(x, y, z)                  :97-102
    SendDataAndInitialize(R, Stress)         :85-86
    RunSingleStrainStep(Edot, maxIt, dt, eps) -> <sigma>   :221-223
    RunVPFFT(Edot, maxIt, steps, dt, eps)    :571-575
    Stress(), LagrangeMultiplier(), Orientation(), SchmidtTensorList(),
    CRSS(), DisplacementVariation(), PolarizationField()   :129-185, 338-434
    NumX/NumY/NumZ, ToIndex                  :121-124, 235-237
    """

    def __init__(self, DimX: int, DimY: int, DimZ: int, material: Optional[Material] = None,
                 precision: int = FP64, crss: float = 1.0):
        self._lib = load_library()
        self.DimX, self.DimY, self.DimZ = int(DimX), int(DimY), int(DimZ)
        self.material = material or Material.uniform()
        self.crss_value = float(crss)
        desc = vpfft_grid_desc(self.DimX, self.DimY, self.DimZ, int(precision))
        self._mat_c = self.material.to_c()
        h = C.c_void_p()
        _check(self._lib.vpfft_grid_create(C.byref(desc), C.byref(self._mat_c), C.byref(h)))
        self._h = h
        self.last_info: Optional[vpfft_step_info] = None

    # ---- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._lib.vpfft_grid_destroy(self._h)
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

    # ---- geometry
    @property
    def N(self) -> int:
        return self.DimX * self.DimY * self.DimZ

    def NumX(self) -> int:
        return self.DimX

    def NumY(self) -> int:
        return self.DimY

    def NumZ(self) -> int:
        return self.DimZ

    def ToIndex(self, x: int, y: int, z: int) -> int:
        return (x * self.DimY + y) * self.DimZ + z

    def SetVoxelLength(self, x: float, y: float, z: float):
        _check(self._lib.vpfft_grid_set_voxel_length(self._h, x, y, z))

    def VoxelLength(self):
        out = np.empty(3)
        _check(self._lib.vpfft_grid_get_voxel_length(self._h, _dptr(out)))
        return out

    def SetBlockShape(self, precision: int, pass_: int, block: int):
        """FFT pencils per block of a pass (0 z, 1 y, 2 x); 0 = default."""
        _check(self._lib.vpfft_grid_set_block_shape(self._h, precision, pass_, block))

    def SetProfiling(self, enable: bool = True):
        """Per-kernel device timing with HIP events on the grid's stream (resets totals)."""
        _check(self._lib.vpfft_grid_set_profiling(self._h, 1 if enable else 0))

    def KernelTimes(self) -> dict:
        """{kernel name: (total ms, launches)} since profiling was enabled."""
        ms = np.zeros(32)
        cnt = np.zeros(32, dtype=np.int64)
        n = self._lib.vpfft_grid_kernel_times(self._h, _dptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_long)), 32)
        if n < 0:
            _check(n)
        return {self._lib.vpfft_kernel_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    def SetPrecision(self, precision: int):
        _check(self._lib.vpfft_grid_set_precision(self._h, int(precision)))

    def SaveState(self):
        """Device copy of the state a strain step changes (vpfft_grid_save_state)."""
        _check(self._lib.vpfft_grid_save_state(self._h))

    def RestoreState(self):
        """Back to the last SaveState (vpfft_grid_restore_state)."""
        _check(self._lib.vpfft_grid_restore_state(self._h))

    # ---- data
    def SendDataAndInitialize(self, Orientations, Stress=None, CRSS=None):
        R = _f64(Orientations, (self.N, 9))
        S = None if Stress is None else _f64(Stress, (self.N, 5))
        T = None if CRSS is None else _f64(CRSS, (self.N, 12))
        _check(self._lib.vpfft_grid_upload(self._h, _dptr(R), _dptr(T), self.crss_value, _dptr(S)))

    def _download(self, fid: int, shape) -> np.ndarray:
        out = np.empty(shape)
        _check(self._lib.vpfft_grid_download(self._h, fid, _dptr(out)))
        return out

    def _upload(self, fid: int, arr, shape):
        a = _f64(arr, shape)
        _check(self._lib.vpfft_grid_upload_field(self._h, fid, _dptr(a)))

    def Stress(self) -> np.ndarray:
        return self._download(FIELD_STRESS, (self.N, 5))

    def LagrangeMultiplier(self) -> np.ndarray:
        return self._download(FIELD_LAGRANGE, (self.N, 5))

    def Orientation(self) -> np.ndarray:
        return self._download(FIELD_ORIENTATION, (self.N, 9))

    def CRSS(self) -> np.ndarray:
        return self._download(FIELD_CRSS, (self.N, 12))

    def AccumulatedShear(self) -> np.ndarray:
        return self._download(FIELD_ACC_SHEAR, (self.N,))

    def SchmidtTensorList(self) -> np.ndarray:
        return self._download(FIELD_SCHMID, (self.N, 12, 5))

    def StrainRate(self) -> np.ndarray:
        """Solvers::ApplyConstitutiveEquations (Solvers.cpp:255-274) of the
        current stress at every voxel, N x 5."""
        return self._download(FIELD_STRAIN_RATE, (self.N, 5))

    def PolarizationField(self) -> np.ndarray:
        return self._download(FIELD_POLARIZATION, (5, self.N))

    def DisplacementVariation(self) -> np.ndarray:
        return self._download(FIELD_STRAIN_RATE_VARIATION, (5, self.N))

    def SetStress(self, s):
        self._upload(FIELD_STRESS, s, (self.N, 5))

    def SetLagrangeMultiplier(self, s):
        self._upload(FIELD_LAGRANGE, s, (self.N, 5))

    def SetCRSS(self, t):
        self._upload(FIELD_CRSS, t, (self.N, 12))

    def SetOrientation(self, R):
        self._upload(FIELD_ORIENTATION, R, (self.N, 9))

    def SetAccumulatedShear(self, a):
        self._upload(FIELD_ACC_SHEAR, a, (self.N,))

    # ---- the hot path
    def RunSingleStrainStep(self, MacroscopicStrainRate, NumMaxIterations: int, TimeStep: float,
                            ConvergenceEpsilon: float) -> np.ndarray:
        e = _f64(MacroscopicStrainRate, (5,))
        avg = np.empty(5)
        info = vpfft_step_info()
        rc = self._lib.vpfft_run_strain_step(self._h, _dptr(e), int(NumMaxIterations), float(TimeStep),
                                             float(ConvergenceEpsilon), _dptr(avg), C.byref(info))
        self.last_info = info
        _check(rc)
        return avg

    def RunVPFFT(self, MacroscopicStrainRate, NumMaxIterations: int, NumTimeSteps: int, TimeStep: float,
                 ConvergenceEpsilon: float, stress_strain_path: Optional[str] = None):
        e = _f64(MacroscopicStrainRate, (5,))
        avg = np.empty((max(NumTimeSteps, 1), 5))
        infos = (vpfft_step_info * max(NumTimeSteps, 1))()
        path = stress_strain_path.encode() if stress_strain_path else None
        _check(self._lib.vpfft_run_vpfft(self._h, _dptr(e), int(NumMaxIterations), int(NumTimeSteps),
                                         float(TimeStep), float(ConvergenceEpsilon), _dptr(avg), infos, path))
        return avg[:NumTimeSteps], [infos[i] for i in range(NumTimeSteps)]

    def StrainRateEvolutionWithFFT(self, L0, phi) -> np.ndarray:
        """d~ = Re IFFT(-S5 FFT(phi)) / N for a given L0 (MaterialGrid.cpp:624-761)."""
        L = _f64(L0, (25,))
        p = _f64(phi, (5, self.N))
        out = np.empty((5, self.N))
        _check(self._lib.vpfft_strain_rate_evolution(self._h, _dptr(L), _dptr(p), _dptr(out)))
        return out


def solve_constrained_batch(material: Material, R, crss, L0, edot, stress, lam, dtilde, nr_error=None):
    """Solvers::SolveConstrainedConstitutiveEquations over n voxels on the GPU."""
    lib = load_library()
    R = _f64(R)
    n = R.reshape(-1, 9).shape[0]
    crss, L0, edot = _f64(crss, (n, 12)), _f64(L0, (25,)), _f64(edot, (5,))
    stress, lam, dtilde = _f64(stress, (n, 5)), _f64(lam, (n, 5)), _f64(dtilde, (n, 5))
    err = np.zeros(n) if nr_error is None else _f64(nr_error, (n,)).copy()
    out = np.empty((n, 5))
    steps = np.empty(n, dtype=np.int32)
    m = material.to_c()
    _check(lib.vpfft_solve_constrained_batch(n, C.byref(m), _dptr(R), _dptr(crss), _dptr(L0), _dptr(edot),
                                             _dptr(stress), _dptr(lam), _dptr(dtilde), _dptr(out), _dptr(err),
                                             steps.ctypes.data_as(C.POINTER(C.c_int))))
    return out, err, steps


def solve_taylor_batch(material: Material, R, crss, edot, stress):
    """Solvers::SolveConstitutiveEquations (Taylor) over n voxels on the GPU."""
    lib = load_library()
    R = _f64(R)
    n = R.reshape(-1, 9).shape[0]
    crss, edot, stress = _f64(crss, (n, 12)), _f64(edot, (5,)), _f64(stress, (n, 5))
    out = np.empty((n, 5))
    m = material.to_c()
    _check(lib.vpfft_solve_taylor_batch(n, C.byref(m), _dptr(R), _dptr(crss), _dptr(edot), _dptr(stress),
                                        _dptr(out)))
    return out


class Solvers:
    """VPFFT::Solvers (Src/Solvers.h:51-153) over batches of voxels on the GPU,
    with the reference's function names and argument order.  Every per-voxel
    argument takes a leading batch axis: SchmidtTensors (n, 12, 5), CRSS
    (n, 12), stresses (n, 5), AntiSchmidtTensors (n, 12, 3, 3).  The strain
    rate and the material vectors (GammaDotBase, RateSensitivity, ...) are
    one per call.  A single voxel is a batch of one."""

    @staticmethod
    def _mat(gdot0, nexp, hmat=None, tau0=None, tau1=None, theta0=None, theta1=None) -> vpfft_material:
        m = Material(gamma_dot_base=list(np.asarray(gdot0, dtype=np.float64).reshape(NSLIP)),
                     rate_sensitivity=[int(x) for x in np.asarray(nexp).reshape(NSLIP)])
        if hmat is not None:
            m.hardening = np.asarray(hmat, dtype=np.float64).reshape(NSLIP, NSLIP)
            m.tau0, m.tau1, m.theta0, m.theta1 = (list(np.asarray(v, dtype=np.float64).reshape(NSLIP))
                                                  for v in (tau0, tau1, theta0, theta1))
        return m.to_c()

    @staticmethod
    def _batch(SchmidtTensors, CRSS):
        m = _f64(SchmidtTensors)
        n = m.reshape(-1, 60).shape[0]
        return n, _f64(m, (n, 60)), _f64(CRSS, (n, 12))

    @staticmethod
    def SolveConstitutiveEquations(InitialStressState, SchmidtTensors, CRSS, GammaDotBase, RateSensitivity,
                                   StrainRate, EpsilonConvergence, MaxResolvedStress, MaxNumIterations):
        """Taylor Newton-Raphson (Solvers.cpp:19-101): stress per voxel (n, 5)."""
        n, m, t = Solvers._batch(SchmidtTensors, CRSS)
        s0, e = _f64(InitialStressState, (n, 5)), _f64(StrainRate, (5,))
        out = np.empty((n, 5))
        mat = Solvers._mat(GammaDotBase, RateSensitivity)
        _check(load_library().vpfft_solvers_taylor(n, C.byref(mat), _dptr(m), _dptr(t), _dptr(e), _dptr(s0),
                                                   float(EpsilonConvergence), float(MaxResolvedStress),
                                                   int(MaxNumIterations), _dptr(out)))
        return out

    @staticmethod
    def SolveConstrainedConstitutiveEquations(InitialStressState, SchmidtTensors, CRSS, GammaDotBase,
                                              RateSensitivity, MacroscopicStrainRate, LagrangeMultiplier,
                                              LocalDisplacementVariation, HomogeonousReference, EpsilonConvergence,
                                              MaxResolvedStress, MaxNumIterations, NR_Error_Out):
        """Constrained Newton-Raphson (Solvers.cpp:111-222).  NR_Error_Out is an
        (n,) float64 array updated in place, as the reference writes through
        its pointer (and leaves it alone when the budget runs out in the
        admissibility back-track).  Returns the stresses (n, 5)."""
        n, m, t = Solvers._batch(SchmidtTensors, CRSS)
        s0, e = _f64(InitialStressState, (n, 5)), _f64(MacroscopicStrainRate, (5,))
        lam, dtl = _f64(LagrangeMultiplier, (n, 5)), _f64(LocalDisplacementVariation, (n, 5))
        L0 = _f64(HomogeonousReference, (25,))
        err = NR_Error_Out
        if not (isinstance(err, np.ndarray) and err.dtype == np.float64 and err.shape == (n,)
                and err.flags.c_contiguous):
            raise ValueError("NR_Error_Out must be a contiguous float64 array of shape (n,)")
        out = np.empty((n, 5))
        mat = Solvers._mat(GammaDotBase, RateSensitivity)
        _check(load_library().vpfft_solvers_constrained(n, C.byref(mat), _dptr(m), _dptr(t), _dptr(e), _dptr(lam),
                                                        _dptr(dtl), _dptr(L0), _dptr(s0), float(EpsilonConvergence),
                                                        float(MaxResolvedStress), int(MaxNumIterations),
                                                        _dptr(out), _dptr(err)))
        return out

    @staticmethod
    def SolveSachEquation(SchmidtTensors, CRSS, StrainRate):
        """Sachs guess (Solvers.cpp:230-243): (n, 5)."""
        n, m, t = Solvers._batch(SchmidtTensors, CRSS)
        e = _f64(StrainRate, (5,))
        out = np.empty((n, 5))
        _check(load_library().vpfft_solvers_sachs(n, _dptr(m), _dptr(t), _dptr(e), _dptr(out)))
        return out

    @staticmethod
    def ApplyConstitutiveEquations(StressState, SchmidtTensors, CRSS, GammaDotBase, RateSensitivity):
        """Strain rate of given stresses (Solvers.cpp:255-274): (n, 5)."""
        n, m, t = Solvers._batch(SchmidtTensors, CRSS)
        s = _f64(StressState, (n, 5))
        out = np.empty((n, 5))
        mat = Solvers._mat(GammaDotBase, RateSensitivity)
        _check(load_library().vpfft_solvers_apply(n, C.byref(mat), _dptr(m), _dptr(t), _dptr(s), _dptr(out)))
        return out

    @staticmethod
    def CalculateShearRate(StressState, SchmidtTensors, AntiSchmidtTensors, CRSS, GammaDotBase, RateSensitivity):
        """Shear rate (Solvers.cpp:282-302): (n, 3, 3)."""
        n, m, t = Solvers._batch(SchmidtTensors, CRSS)
        s, a = _f64(StressState, (n, 5)), _f64(AntiSchmidtTensors, (n, 108))
        out = np.empty((n, 3, 3))
        mat = Solvers._mat(GammaDotBase, RateSensitivity)
        _check(load_library().vpfft_solvers_shear_rate(n, C.byref(mat), _dptr(m), _dptr(a), _dptr(t), _dptr(s),
                                                       _dptr(out)))
        return out

    @staticmethod
    def CalculateHardening(AccumulatedShear, TimeStep, StressState, SchmidtTensors, CRSS, GammaDotBase,
                           RateSensitivity, HardeningMatrix, Tau0, Tau1, Theta0, Theta1):
        """Voce hardening as written (Solvers.cpp:314-373).  AccumulatedShear
        is an (n,) float64 array updated in place; returns the CRSS (n, 12)."""
        n, m, t = Solvers._batch(SchmidtTensors, CRSS)
        acc = AccumulatedShear
        if not (isinstance(acc, np.ndarray) and acc.dtype == np.float64 and acc.shape == (n,)
                and acc.flags.c_contiguous):
            raise ValueError("AccumulatedShear must be a contiguous float64 array of shape (n,)")
        s = _f64(StressState, (n, 5))
        out = np.empty((n, 12))
        mat = Solvers._mat(GammaDotBase, RateSensitivity, HardeningMatrix, Tau0, Tau1, Theta0, Theta1)
        _check(load_library().vpfft_solvers_hardening(n, C.byref(mat), _dptr(m), _dptr(t), _dptr(s),
                                                      float(TimeStep), _dptr(acc), _dptr(out)))
        return out


# ---------------------------------------------------------------------------
# batched strain steps
# ---------------------------------------------------------------------------
class Batch:
    """Advance several MaterialGrids (any sizes) through one strain step each,
    together (include/vpfft.h, vpfft_batch_*)."""

    def __init__(self, tune=False):
        """tune: False / 0 default schedules; True / 1 tuned block shapes (same
        bits); 2 or "radix": tuned block shapes and radix orders."""
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.vpfft_batch_create(C.byref(h)))
        self._h = h
        level = 2 if tune == "radix" else int(tune)
        if level:
            _check(self._lib.vpfft_batch_set_tuning(self._h, level))
        self.last_infos: list = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vpfft_batch_destroy(self._h)
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

    def block_shape(self, grid: "MaterialGrid", pass_: int) -> int:
        b = C.c_int()
        _check(self._lib.vpfft_batch_block_shape(self._h, grid._h, pass_, C.byref(b)))
        return b.value

    def radix_plan(self, grid: "MaterialGrid", pass_: int) -> list:
        """The radix stages of grid's pass (0 = z, 1 = y, 2 = x), in order."""
        r = (C.c_int * 12)()
        k = C.c_int()
        _check(self._lib.vpfft_batch_radix_plan(self._h, grid._h, pass_, r, C.byref(k)))
        return [r[i] for i in range(k.value)]

    def RunSingleStrainStep(self, grids, edots, NumMaxIterations: int, TimeStep: float,
                            ConvergenceEpsilon: float) -> np.ndarray:
        """One strain step of every grid; returns <sigma> per grid (count x 5)."""
        count = len(grids)
        e = _f64(edots, (count, 5))
        dts = _f64(np.broadcast_to(np.asarray(TimeStep, dtype=np.float64), (count,)))
        arr = (C.c_void_p * count)(*[g._h for g in grids])
        avg = np.empty((count, 5))
        infos = (vpfft_step_info * count)()
        rc = self._lib.vpfft_batch_run_strain_step(self._h, arr, count, _dptr(e), int(NumMaxIterations),
                                                    _dptr(dts), float(ConvergenceEpsilon), _dptr(avg), infos)
        self.last_infos = [infos[j] for j in range(count)]
        for g, inf in zip(grids, self.last_infos):
            g.last_info = inf
        _check(rc)
        return avg
