"""Minimal simulation objects standing in for REBOUND's (not part of this build).

``Particle`` and ``Simulation`` mirror ``struct reb_particle`` and
``struct reb_simulation`` of include/assist.h field by field; ``Simulation.integrate``
runs IAS15 on the GPU through libassist.
"""
from __future__ import annotations

from ctypes import POINTER, Structure, byref, c_char, c_double, c_int, c_void_p

from ._libassist import clibassist


class Particle(Structure):
    _fields_ = [("x", c_double), ("y", c_double), ("z", c_double),
                ("vx", c_double), ("vy", c_double), ("vz", c_double),
                ("ax", c_double), ("ay", c_double), ("az", c_double),
                ("m", c_double)]

    def __sub__(self, other):
        p = Particle()
        for f, _ in self._fields_:
            setattr(p, f, getattr(self, f) - getattr(other, f))
        return p

    def __repr__(self):
        return "<Particle x=%.16g y=%.16g z=%.16g vx=%.16g vy=%.16g vz=%.16g>" % (
            self.x, self.y, self.z, self.vx, self.vy, self.vz)


class VariationalConfiguration(Structure):
    _fields_ = [("index", c_int), ("testparticle", c_int)]


class IntegratorIAS15(Structure):
    _fields_ = [("epsilon", c_double), ("min_dt", c_double), ("adaptive_mode", c_int)]


class Simulation(Structure):
    _fields_ = [("t", c_double), ("dt", c_double), ("dt_last_done", c_double),
                ("N", c_int), ("N_var", c_int), ("N_var_config", c_int), ("allocated_N", c_int),
                ("_particles", POINTER(Particle)), ("var_config", POINTER(VariationalConfiguration)),
                ("status", c_int), ("exact_finish_time", c_int),
                ("ri_ias15", IntegratorIAS15),
                ("extras", c_void_p), ("additional_forces", c_void_p),
                ("error_message", c_char * 256)]

    def __init__(self):
        super().__init__()
        clibassist.reb_simulation_init(byref(self))
        self._initialised = True

    def __del__(self):
        if getattr(self, "_initialised", False):
            clibassist.reb_simulation_free_pointers(byref(self))
            self._initialised = False

    @property
    def particles(self):
        return [self._particles[i] for i in range(self.N)]

    def add(self, particle: Particle | None = None, **kw):
        p = Particle()
        if particle is not None:
            for f, _ in Particle._fields_:
                setattr(p, f, getattr(particle, f))
        for k, v in kw.items():
            setattr(p, k, float(v))
        clibassist.reb_simulation_add(byref(self), p)

    def add_variation(self, testparticle: int = 0) -> int:
        """First-order variational particle of real particle ``testparticle``; returns its index."""
        return clibassist.reb_simulation_add_variation_1st_order(byref(self), c_int(testparticle))

    def integrate(self, tmax: float):
        clibassist.reb_simulation_integrate(byref(self), c_double(tmax))
        if self.status:
            raise RuntimeError(self.error_message.decode("ascii", "replace"))

    def update_acceleration(self):
        """Evaluates the attached ASSIST force model (on the GPU) at ``t``."""
        clibassist.assist_additional_forces(byref(self))
        if self.status:
            raise RuntimeError(self.error_message.decode("ascii", "replace"))
