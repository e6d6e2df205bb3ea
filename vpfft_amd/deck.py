"""Input deck and microstructure files (include/vpfft_io.h) with the
reference's names.

ConfigFile                 Utilities::ConfigFile, Src/Config.h:21-52, Src/Config.cpp:11-261
two_grain_init             UnitTester::SetMaterialGrid, Src/UnitTester.cpp:46-78
read_euler_file            UnitTester::ReadRicardoData, Src/UnitTester.cpp:237-265
write_euler_file           UnitTester::PrintGrid, Src/UnitTester.cpp:17-39
write_material_file /
read_material_file         UnitTester::WriteRandomMaterial / ReadRandomMaterial, Src/UnitTester.cpp:116-168
run_deck                   a deck-driven RunVPFFT (Src/MaterialGrid.cpp:1202-1228)

All of these run on the host; only run_deck uses the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import (FP64, Material, VPFFTError, _check, _dptr, load_library, vpfft_deck, vpfft_material,
               vpfft_step_info)

INIT_TWO_GRAIN = 0  # ConfigFile::eInitTwoGrain
INIT_FILE = 1       # ConfigFile::eInitFile


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def _path(p) -> bytes:
    return os.fsencode(p)


class ConfigFile:
    """The reference's input deck.  Read() returns True, or raises VPFFTError
    with the reference's diagnostic where the reference returns false or
    exits."""

    def __init__(self):
        self._d = vpfft_deck()

    def Read(self, Filename) -> bool:
        _check(load_library().vpfft_deck_read(_path(Filename), C.byref(self._d)))
        return True

    def Parse(self, text: str) -> bool:
        _check(load_library().vpfft_deck_parse(text.encode(), C.byref(self._d)))
        return True

    # members, Src/Config.h:32-50
    @property
    def NumX(self) -> int:
        return self._d.num_x

    @property
    def NumY(self) -> int:
        return self._d.num_y

    @property
    def NumZ(self) -> int:
        return self._d.num_z

    @property
    def InputStrainRate(self) -> np.ndarray:
        return np.array([[self._d.input_strain_rate[i][j] for j in range(3)] for i in range(3)])

    @property
    def MaxNRIter(self) -> int:
        return self._d.max_nr_iter

    @property
    def NRTolerence(self) -> float:
        return self._d.nr_tolerance

    @property
    def TimeStep(self) -> float:
        return self._d.time_step

    @property
    def NumTimeIter(self) -> int:
        return self._d.num_time_iter

    @property
    def InitMethod(self) -> int:
        return self._d.init_method

    @property
    def InputMicrostructureFilename(self) -> str:
        return os.fsdecode(self._d.microstructure_file)

    @property
    def CRSS(self) -> np.ndarray:
        return np.array(self._d.crss[:])

    @property
    def GammaDotBase(self) -> np.ndarray:
        return np.array(self._d.gamma_dot_base[:])

    @property
    def RateSensitivity(self) -> np.ndarray:
        return np.array(self._d.rate_sensitivity[:], dtype=np.int64)

    def material(self) -> Material:
        """The deck's GammaDotBase / RateSensitivity with the reference's Voce defaults."""
        m = vpfft_material()
        _check(load_library().vpfft_deck_material(C.byref(self._d), C.byref(m)))
        return Material(gamma_dot_base=list(m.gamma_dot_base), rate_sensitivity=list(m.rate_sensitivity),
                        tau0=list(m.tau0), tau1=list(m.tau1), theta0=list(m.theta0), theta1=list(m.theta1),
                        hardening=np.array(m.hardening[:]).reshape(12, 12))

    def strain_rate(self) -> np.ndarray:
        """EigenRep(InputStrainRate)."""
        e = np.empty(5)
        _check(load_library().vpfft_deck_strain_rate(C.byref(self._d), _dptr(e)))
        return e


def two_grain_init(nx: int, ny: int, nz: int):
    """(orientation N x 9, stress placeholders N x 5) of SetMaterialGrid."""
    n = nx * ny * nz
    R, S = np.empty((n, 9)), np.empty((n, 5))
    _check(load_library().vpfft_two_grain_init(nx, ny, nz, _dptr(R), _dptr(S)))
    return R, S


def euler_deg_to_matrix(euler_deg) -> np.ndarray:
    """BuildActiveEulerMatrix of angles in degrees (float cosf / sinf, as the reference)."""
    e = np.ascontiguousarray(euler_deg, dtype=np.float64).reshape(-1, 3)
    R = np.empty((e.shape[0], 9))
    _check(load_library().vpfft_euler_deg_to_matrix(_dptr(e), e.shape[0], _dptr(R)))
    return R


def read_euler_file(path, nx: int, ny: int, nz: int):
    """(orientation N x 9, euler N x 3 degrees, grain N, phase N) in ToIndex order."""
    n = nx * ny * nz
    R, E = np.empty((n, 9)), np.empty((n, 3))
    g, p = np.empty(n, dtype=np.int32), np.empty(n, dtype=np.int32)
    _check(load_library().vpfft_read_euler_file(_path(path), nx, ny, nz, _dptr(R), _dptr(E), _ip(g), _ip(p)))
    return R, E, g, p


def write_euler_file(path, nx: int, ny: int, nz: int, euler_deg, grain=None, phase=None):
    e = np.ascontiguousarray(euler_deg, dtype=np.float64).reshape(nx * ny * nz, 3)
    g = None if grain is None else np.ascontiguousarray(grain, dtype=np.int32)
    p = None if phase is None else np.ascontiguousarray(phase, dtype=np.int32)
    _check(load_library().vpfft_write_euler_file(_path(path), nx, ny, nz, _dptr(e), _ip(g), _ip(p)))


def write_material_file(path, nx: int, ny: int, nz: int, orientation, stress=None):
    R = np.ascontiguousarray(orientation, dtype=np.float64).reshape(nx * ny * nz, 9)
    S = None if stress is None else np.ascontiguousarray(stress, dtype=np.float64).reshape(nx * ny * nz, 5)
    _check(load_library().vpfft_write_material_file(_path(path), nx, ny, nz, _dptr(R),
                                                     None if S is None else _dptr(S)))


def read_material_file(path, nx: int, ny: int, nz: int):
    n = nx * ny * nz
    R, S = np.empty((n, 9)), np.empty((n, 5))
    _check(load_library().vpfft_read_material_file(_path(path), nx, ny, nz, _dptr(R), _dptr(S)))
    return R, S


def run_deck(deck_path, precision: int = FP64, stress_strain_path: Optional[str] = "Stress_Strain.txt",
             material_out_path: Optional[str] = None):
    """Run the steps a deck describes on the GPU; returns the per-step records."""
    cfg = ConfigFile()
    cfg.Read(deck_path)
    infos = (vpfft_step_info * max(1, cfg.NumTimeIter))()
    _check(load_library().vpfft_run_deck(
        _path(deck_path), precision, None if stress_strain_path is None else _path(stress_strain_path),
        None if material_out_path is None else _path(material_out_path), infos, max(1, cfg.NumTimeIter)))
    return [infos[i].as_dict() for i in range(cfg.NumTimeIter)]


__all__ = ["ConfigFile", "INIT_TWO_GRAIN", "INIT_FILE", "two_grain_init", "euler_deg_to_matrix", "read_euler_file",
           "write_euler_file", "write_material_file", "read_material_file", "run_deck", "VPFFTError"]
