"""CPU oracle of the AlgoGas path -- TEST INFRASTRUCTURE ONLY.

Loads liboracle.so (built from oracle.hpp / oracle_capi.cpp by oracle/Makefile)
for tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
product library never imports this.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    if not os.path.exists(LIBPATH):
        build()
    lib = ctypes.CDLL(LIBPATH)
    vp, d, i, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
    sigs = {
        "orc_create": ([vp, vp, i], vp), "orc_destroy": ([vp], None),
        "orc_step": ([vp, i], ctypes.c_longlong),
        "orc_get_field": ([vp, i, vp, sz], i), "orc_set_field": ([vp, i, vp, sz], i),
        "orc_get_ring": ([vp, i, vp, sz], i),
        "orc_get_planets": ([vp, vp], None), "orc_set_planets": ([vp, vp], None),
        "orc_info": ([vp, vp], None), "orc_force": ([vp, d, d, d, d, vp], None),
        "orc_condition_cfl": ([vp, d], i), "orc_stage": ([vp, ctypes.c_char_p, d], i),
        "orc_get_int_velocities": ([vp, vp, vp, sz], i),
        "orc_fargo_shift_f32": ([ctypes.c_float, ctypes.c_float, d, i], i),
        "orc_fargo_shift_f64": ([d, d, d, i], i),
        "orc_fft2d": ([vp, vp, i, i, i], None), "orc_sg_accel": ([vp, vp, vp], i),
        "orc_get_array": ([vp, ctypes.c_char_p, vp, sz], i), "orc_set_array": ([vp, ctypes.c_char_p, vp, sz], i),
        "orc_call": ([vp, ctypes.c_char_p, d], i),
        "orc_sizeof_params": ([], sz), "orc_sizeof_planets": ([], sz), "orc_num_threads": ([], i),
    }
    for n, (a, r) in sigs.items():
        f = getattr(lib, n)
        f.argtypes, f.restype = a, r
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


FIELDS = {"dens": 0, "vrad": 1, "vtheta": 2, "energy": 3, "temperature": 4,
          "soundspeed": 5, "divv": 6, "qplus": 7, "label": 8, "potential": 9,
          "sg_accr": 10, "sg_acct": 11}
RINGS = {"radii": 0, "rmed": 1, "surf": 2, "sigmamed": 3, "energymed": 4,
         "viscosity": 5, "vmed": 6, "nshift": 7, "cs": 8, "newdt": 9,
         "invrmed": 10}


FP32, FP64, EXT80 = 0, 1, 2   # arithmetic of an oracle instance (EXT80: x87 long double)


class Oracle:
    """One disk on the CPU; params/planets are the product's fgx structs
    (only their bytes are used).  precision: FP32, FP64 (the restatement
    proper and the CPU baseline) or EXT80 (the same code in 80-bit long
    double, the checker of the FP64 engine)."""

    def __init__(self, params, planets, precision=1):
        L = lib()
        self.nrad, self.nsec = int(params.get("NRAD")), int(params.get("NSEC"))
        self.h = L.orc_create(ctypes.cast(params.buf, ctypes.c_void_p), ctypes.addressof(planets), precision)
        if not self.h:
            raise RuntimeError("oracle set-up failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def step(self, n=1):
        return lib().orc_step(self.h, n)

    def field(self, name, ghost=False):
        n = ((self.nrad + 1) if ghost else self.nrad) * self.nsec
        out = np.empty(n)
        if lib().orc_get_field(self.h, FIELDS[name], out.ctypes.data, n):
            raise ValueError(name)
        return out.reshape(-1, self.nsec)

    def set_field(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if lib().orc_set_field(self.h, FIELDS[name], v.ctypes.data, v.size):
            raise ValueError(name)

    def ring(self, name):
        out = np.empty(self.nrad + 1)
        lib().orc_get_ring(self.h, RINGS[name], out.ctypes.data, self.nrad + 1)
        return out

    def info(self):
        out = np.zeros(11)
        lib().orc_info(self.h, out.ctypes.data)
        keys = ["time", "omega_frame", "lost_mass", "mdcp1", "exces_mdcp", "last_dt", "last_ncfl",
                "substeps", "masstaper", "ind_x", "ind_y"]
        return dict(zip(keys, out.tolist()))

    def planets(self, cls):
        pl = cls()
        lib().orc_get_planets(self.h, ctypes.addressof(pl))
        return pl

    def set_planets(self, pl):
        lib().orc_set_planets(self.h, ctypes.addressof(pl))

    def force(self, x, y, sm, m):
        out = np.zeros(44)
        lib().orc_force(self.h, x, y, sm, m, out.ctypes.data)
        return out

    def condition_cfl(self, delta_t):
        return lib().orc_condition_cfl(self.h, delta_t)

    def array(self, name):
        """a named array of the restatement (fields, work arrays, stresses),
        shaped (nrad+1, nsec)"""
        n = (self.nrad + 1) * self.nsec
        out = np.empty(n)
        if lib().orc_get_array(self.h, name.encode(), out.ctypes.data, n):
            raise ValueError(name)
        return out.reshape(-1, self.nsec)

    def set_array(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if lib().orc_set_array(self.h, name.encode(), v.ctypes.data, v.size):
            raise ValueError(name)

    def call(self, name, dt=0.0):
        """one function of the restatement on the current arrays"""
        if lib().orc_call(self.h, name.encode(), dt):
            raise ValueError(name)

    def stage(self, name, dt):
        r = lib().orc_stage(self.h, name.encode(), dt)
        if r:
            raise ValueError(name)
