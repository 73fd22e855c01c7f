"""ctypes mirrors of the SPK reader's C interface (include/spk.h, the reference's src/spk.h:15-122).

The reference reads these structs directly in its unit tests (unit_tests/spk_init,
spk_join_masses, spk_load_constants, spk_planets_calc); the mirrors let Python do the same with
libassist: ``SpkFile`` (struct spk_s), ``SpkTarget``, ``ConstantsAndMasses`` and ``Mpos``, with
the entry points' argument and result types set on the shared library.
"""
from __future__ import annotations

import ctypes as C

from ._libassist import clibassist

_dp = C.POINTER(C.c_double)


class Mpos(C.Structure):
    _fields_ = [("u", C.c_double * 3), ("v", C.c_double * 3), ("w", C.c_double * 3)]


class MassData(C.Structure):
    _fields_ = [("names", C.POINTER(C.c_char_p)), ("values", _dp), ("count", C.c_size_t)]


class ConstantsAndMasses(C.Structure):
    _fields_ = [("AU", C.c_double), ("EMRAT", C.c_double), ("J2E", C.c_double), ("J3E", C.c_double),
                ("J4E", C.c_double), ("J2SUN", C.c_double), ("RE", C.c_double), ("CLIGHT", C.c_double),
                ("ASUN", C.c_double), ("masses", MassData)]


class SpkTarget(C.Structure):
    _fields_ = [("code", C.c_int), ("cen", C.c_int), ("mass", C.c_double), ("beg", C.c_double),
                ("end", C.c_double), ("res", C.c_double), ("one", C.POINTER(C.c_int)), ("two", C.POINTER(C.c_int)),
                ("ind", C.c_int), ("allocated_ind", C.c_int)]


class SpkFile(C.Structure):
    _fields_ = [("targets", C.POINTER(SpkTarget)), ("num", C.c_int), ("allocated_num", C.c_int),
                ("map", C.c_void_p), ("len", C.c_size_t), ("device_cache", C.c_void_p)]


_SP = C.POINTER(SpkFile)
_lib = clibassist
_lib.assist_spk_init.restype = _SP
_lib.assist_spk_init.argtypes = [C.c_char_p]
_lib.assist_spk_free.argtypes = [_SP]
_lib.assist_spk_find_target.restype = C.POINTER(SpkTarget)
_lib.assist_spk_find_target.argtypes = [_SP, C.c_int]
_lib.assist_spk_target_pos.restype = Mpos
_lib.assist_spk_target_pos.argtypes = [_SP, C.POINTER(SpkTarget), C.c_double, C.c_double]
_lib.assist_load_spk_constants_and_masses.restype = ConstantsAndMasses
_lib.assist_load_spk_constants_and_masses.argtypes = [C.c_char_p]
_lib.assist_apply_spk_constants.argtypes = [C.c_void_p, C.POINTER(ConstantsAndMasses)]
_lib.assist_free_spk_constants_and_masses.argtypes = [C.POINTER(ConstantsAndMasses)]
_lib.assist_spk_join_masses.argtypes = [_SP, C.POINTER(MassData), C.c_double]
_lib.assist_spk_calc.argtypes = [_SP, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
_lib.assist_spk_calc_planets.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int] + [_dp] * 10
_lib.assist_spk_calc_planets_by_assist.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int] + [_dp] * 10


def spk_init(path: str):
    """assist_spk_init: a pointer to the file's struct spk_s (NULL pointer on failure)."""
    return _lib.assist_spk_init(path.encode())


def spk_free(sp) -> int:
    return _lib.assist_spk_free(sp)


def load_constants_and_masses(path: str) -> ConstantsAndMasses:
    return _lib.assist_load_spk_constants_and_masses(path.encode())
