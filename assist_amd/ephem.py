"""``Ephem``: the ephemeris object (reference assist/ephem.py:38-120).

Same ctypes field layout as ``struct assist_ephem`` and the same methods,
body-name table and errors as the reference wrapper.
"""
from ctypes import Structure, byref, c_char_p, c_double, c_int, c_void_p
from typing import Optional, Union

from ._libassist import assist_error_messages, clibassist
from .simulation import Particle

ASSIST_BODY_IDS = {
    0: "Sun", 1: "Mercury", 2: "Venus", 3: "Earth", 4: "Moon", 5: "Mars", 6: "Jupiter",
    7: "Saturn", 8: "Uranus", 9: "Neptune", 10: "Pluto",
    11: "Camilla", 12: "Ceres", 13: "Cybele", 14: "Davida", 15: "Eunomia", 16: "Euphrosyne",
    17: "Europa", 18: "Hygiea", 19: "Interamnia", 20: "Iris", 21: "Juno", 22: "Pallas",
    23: "Psyche", 24: "Sylvia", 25: "Thisbe", 26: "Vesta",
}


class Ephem(Structure):
    """Main object for ASSIST ephemeris operations; not tied to a simulation."""

    def __init__(self, planets_path: Optional[str] = None, asteroids_path: Optional[str] = None):
        c_planets = c_char_p(str(planets_path).encode("ascii")) if planets_path is not None else c_char_p(None)
        c_ast = c_char_p(str(asteroids_path).encode("ascii")) if asteroids_path is not None else c_char_p(None)
        clibassist.assist_ephem_init.restype = c_int
        ret = clibassist.assist_ephem_init(byref(self), c_planets, c_ast)
        if ret != 0:
            raise RuntimeError(assist_error_messages(ret))
        self._initialised = True

    def get_particle(self, body: Union[int, str], t: float) -> Particle:
        if isinstance(body, str):
            body_str = body.lower()
            body = -1
            for k in ASSIST_BODY_IDS:
                if body_str == ASSIST_BODY_IDS[k].lower():
                    body = k
            if body < 0:
                raise ValueError("Cannot find body '" + body_str + "'. Needs to be one of: " +
                                 ", ".join([ASSIST_BODY_IDS[k] for k in ASSIST_BODY_IDS]) + ".")
        if not isinstance(body, int):
            raise ValueError("Expecting integer for body id.")
        clibassist.assist_get_particle_with_error.restype = Particle
        e = c_int(0)
        p = clibassist.assist_get_particle_with_error(byref(self), c_int(body), c_double(t), byref(e))
        if e.value:
            raise RuntimeError(assist_error_messages(e.value))
        return p

    def time_bounds(self) -> tuple:
        """The (t_beg, t_end) coverage interval, relative to ``jd_ref``."""
        t_beg = c_double(0.0)
        t_end = c_double(0.0)
        clibassist.assist_ephem_time_bounds(byref(self), byref(t_beg), byref(t_end))
        return (t_beg.value, t_end.value)

    def __del__(self) -> None:
        if getattr(self, "_initialised", False):
            clibassist.assist_ephem_free_pointers(byref(self))
            self._initialised = False

    _fields_ = [
        ("jd_ref", c_double),
        ("spk_planets", c_void_p),
        ("spk_asteroids", c_void_p),
        ("ascii_planets", c_void_p),
        ("planets_source", c_int),
        ("planets_calc", c_void_p),
        ("spk_target_index", c_int * 11),
        ("spk_emb_index", c_int),
        ("AU", c_double),
        ("EMRAT", c_double),
        ("J2E", c_double),
        ("J3E", c_double),
        ("J4E", c_double),
        ("J2SUN", c_double),
        ("RE", c_double),
        ("CLIGHT", c_double),
        ("ASUN", c_double),
        ("Re_eq", c_double),
        ("Rs_eq", c_double),
        ("c_AU_per_day", c_double),
        ("c_squared", c_double),
        ("over_c_squared", c_double),
    ]
