"""Python mirror of the fgx C ABI (include/fgx.h) over ctypes.

The reference drives its hot path from C++ main() (src/Main.cu:87-401); this
module exposes the same run surface to Python callers and tests: parameter
files (.par / .cfg), one disk or a batch of disks stepped by AlgoGas calls,
field and ring-profile read-back, planets, forces and per-kernel timing.
Every call goes to the native library libfgx.so built next to this file; an
import fails loudly when it is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libfgx.so")

FGX_OK = 0
F32, F64 = 0, 1
MAX_PLANETS = 8

FIELDS = {"dens": 0, "vrad": 1, "vtheta": 2, "energy": 3, "temperature": 4,
          "soundspeed": 5, "divv": 6, "qplus": 7, "label": 8, "potential": 9,
          "sg_accr": 10, "sg_acct": 11}
RINGS = {"radii": 0, "rmed": 1, "surf": 2, "sigmamed": 3, "energymed": 4,
         "viscosity": 5, "vmed": 6, "nshift": 7, "cs": 8, "newdt": 9,
         "invrmed": 10}


class FgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fgx error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(_LIBPATH):
        raise ImportError(
            f"native library {_LIBPATH} is missing: build it with "
            "`python -m cudaprotoplanetarydisksfloat_amd.build`")
    # One HIP runtime per process: PyTorch-ROCm brings its own copy of the
    # HIP and HSA runtime libraries.  Loaded after them, libfgx.so shares
    # them; loaded before them (this module imported ahead of torch), the
    # first GPU use by torch leaves libfgx.so's runtime with no device
    # (hipGetDeviceCount fails).  So torch, where installed, is loaded first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return ctypes.CDLL(_LIBPATH)


lib = _load()


class Planets(ctypes.Structure):
    """fgx_planets: the reference's PlanetarySystem (include/Types.cuh:2-15)."""
    _fields_ = [("nb", ctypes.c_int),
                ("mass", ctypes.c_double * MAX_PLANETS),
                ("x", ctypes.c_double * MAX_PLANETS),
                ("y", ctypes.c_double * MAX_PLANETS),
                ("vx", ctypes.c_double * MAX_PLANETS),
                ("vy", ctypes.c_double * MAX_PLANETS),
                ("acc", ctypes.c_double * MAX_PLANETS),
                ("feel_disk", ctypes.c_int * MAX_PLANETS),
                ("feel_others", ctypes.c_int * MAX_PLANETS),
                ("name", (ctypes.c_char * 64) * MAX_PLANETS)]

    def as_dict(self):
        n = self.nb
        return {"nb": n, "mass": list(self.mass[:n]), "x": list(self.x[:n]),
                "y": list(self.y[:n]), "vx": list(self.vx[:n]), "vy": list(self.vy[:n]),
                "feel_disk": list(self.feel_disk[:n]), "feel_others": list(self.feel_others[:n])}


class DiskInfo(ctypes.Structure):
    _fields_ = [("nrad", ctypes.c_int), ("nsec", ctypes.c_int), ("precision", ctypes.c_int),
                ("physical_time", ctypes.c_double), ("omega_frame", ctypes.c_double),
                ("lost_mass", ctypes.c_double), ("mdcp1", ctypes.c_double),
                ("exces_mdcp", ctypes.c_double), ("last_dt", ctypes.c_double),
                ("last_ncfl", ctypes.c_int), ("substeps", ctypes.c_longlong),
                ("dt_steps", ctypes.c_longlong), ("health", ctypes.c_int)]


class Tiling(ctypes.Structure):
    """fgx_tiling: strip heights of the batch kernels and the short-strip
    threshold (0 = the engine's choice)."""
    _fields_ = [("strip_rings", ctypes.c_int), ("source_rings", ctypes.c_int),
                ("fine_source_rings", ctypes.c_int), ("fine_strip_rings", ctypes.c_int),
                ("fine_cells", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneReport(ctypes.Structure):
    """fgx_tune_report: what the autotuner timed and chose."""
    _fields_ = [("from_cache", ctypes.c_int), ("candidates", ctypes.c_int),
                ("default_rate", ctypes.c_double), ("best_rate", ctypes.c_double),
                ("seconds", ctypes.c_double), ("signature", ctypes.c_char * 128)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["signature"] = self.signature.decode()
        return d


class ServerConfig(ctypes.Structure):
    """fgx_server_config (include/fgx.h)."""
    _fields_ = [("precision", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("max_batch_disks", ctypes.c_int), ("snapshot_every", ctypes.c_int),
                ("batch_window", ctypes.c_double), ("keep_final", ctypes.c_int),
                ("max_live_batches", ctypes.c_int), ("spare_slots", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    """fgx_request_info: timestamps (s from the start of the run) and work."""
    _fields_ = [("arrival", ctypes.c_double), ("admit", ctypes.c_double), ("finish", ctypes.c_double),
                ("substeps", ctypes.c_longlong), ("cell_updates", ctypes.c_longlong),
                ("n_dt", ctypes.c_int), ("snapshots", ctypes.c_int), ("batch", ctypes.c_int),
                ("batch_size", ctypes.c_int), ("status", ctypes.c_int), ("slot", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """fgx_server_stats: batches created, slot refills, peak live batches and bytes."""
    _fields_ = [("batches", ctypes.c_int), ("refills", ctypes.c_int), ("peak_live", ctypes.c_int),
                ("peak_bytes", ctypes.c_size_t)]


_vp, _cp, _sz, _i, _d = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
_sigs = {
    "fgx_version": ([], _i), "fgx_last_error": ([], _cp),
    "fgx_sizeof_params": ([], _sz), "fgx_sizeof_planets": ([], _sz),
    "fgx_device_count": ([ctypes.POINTER(_i)], _i),
    "fgx_params_default": ([_vp], _i), "fgx_params_load": ([_cp, _vp], _i),
    "fgx_params_parse": ([_cp, _vp], _i), "fgx_params_set": ([_vp, _cp, _cp], _i),
    "fgx_params_get": ([_vp, _cp, _cp, _sz], _i),
    "fgx_planets_load": ([_cp, _vp, ctypes.POINTER(Planets)], _i),
    "fgx_planets_parse": ([_cp, _vp, ctypes.POINTER(Planets)], _i),
    "fgx_disk_create": ([_vp, ctypes.POINTER(Planets), _i, ctypes.POINTER(_vp)], _i),
    "fgx_batch_create": ([_i, _vp, ctypes.POINTER(Planets), _i, ctypes.POINTER(_vp)], _i),
    "fgx_batch_create_tiled": ([_i, _vp, ctypes.POINTER(Planets), _i, ctypes.POINTER(Tiling),
                                ctypes.POINTER(_vp)], _i),
    "fgx_batch_tiling": ([_vp, ctypes.POINTER(Tiling)], _i),
    "fgx_autotune": ([_i, _vp, ctypes.POINTER(Planets), _i, _i, _cp, ctypes.POINTER(Tiling),
                      ctypes.POINTER(TuneReport)], _i),
    "fgx_disk_step": ([_vp, _i], _i), "fgx_batch_step": ([_vp, _i], _i),
    "fgx_disk_destroy": ([_vp], _i), "fgx_batch_destroy": ([_vp], _i),
    "fgx_batch_size": ([_vp, ctypes.POINTER(_i)], _i),
    "fgx_batch_info": ([_vp, _i, ctypes.POINTER(DiskInfo)], _i),
    "fgx_batch_planets": ([_vp, _i, ctypes.POINTER(Planets)], _i),
    "fgx_batch_read_field": ([_vp, _i, _i, _i, _vp, _sz], _i),
    "fgx_batch_write_field": ([_vp, _i, _i, _i, _vp, _sz], _i),
    "fgx_disk_read_ring": ([_vp, _i, _i, _vp, _sz], _i),
    "fgx_disk_planets": ([_vp, ctypes.POINTER(Planets)], _i),
    "fgx_disk_set_planets": ([_vp, ctypes.POINTER(Planets)], _i),
    "fgx_disk_force": ([_vp, _d, _d, _d, _d, ctypes.POINTER(_d)], _i),
    "fgx_disk_force_noalloc": ([_vp, _d, _d, _d, _d, ctypes.POINTER(_d)], _i),
    "fgx_disk_write_outputs": ([_vp, _i], _i), "fgx_disk_write_logs": ([_vp, _i], _i),
    "fgx_disk_write_setup_files": ([_vp], _i),
    "fgx_footprint_bytes": ([_vp, _i, ctypes.POINTER(_sz)], _i),
    "fgx_profile_enable": ([_vp, _i], _i),
    "fgx_profile_get": ([_vp, _i, ctypes.POINTER(_cp), ctypes.POINTER(_d),
                         ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_d)], _i),
    "fgx_profile_reset": ([_vp], _i),
    "fgx_run_main": ([_i, ctypes.POINTER(_cp)], _i),
    "fgx_batch_freeze": ([_vp, _i, _i], _i),
    "fgx_batch_checkpoint": ([_vp, _i, _cp], _i), "fgx_batch_restore": ([_vp, _i, _cp], _i),
    "fgx_batch_reset_disk": ([_vp, _i, _vp, ctypes.POINTER(Planets)], _i),
    "fgx_server_create": ([ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)], _i),
    "fgx_server_submit": ([_vp, _vp, ctypes.POINTER(Planets), _i, _d, ctypes.POINTER(_i)], _i),
    "fgx_server_run": ([_vp], _i),
    "fgx_server_request": ([_vp, _i, ctypes.POINTER(RequestInfo)], _i),
    "fgx_server_read_field": ([_vp, _i, _i, _i, _vp, _sz], _i),
    "fgx_server_stats_get": ([_vp, ctypes.POINTER(ServerStats)], _i),
    "fgx_server_destroy": ([_vp], _i),
    "fgx_group_create": ([_i, _vp, ctypes.POINTER(Planets), _i, _i, ctypes.POINTER(Tiling), ctypes.POINTER(_vp)], _i),
    "fgx_group_step": ([_vp, _i], _i), "fgx_group_parts": ([_vp, ctypes.POINTER(_i)], _i),
    "fgx_group_batch": ([_vp, _i, ctypes.POINTER(_vp)], _i),
    "fgx_group_disk": ([_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_i)], _i),
    "fgx_group_destroy": ([_vp], _i),
}
for _n, (_a, _r) in _sigs.items():
    _f = getattr(lib, _n)
    _f.argtypes, _f.restype = _a, _r


def _check(code):
    if code != FGX_OK:
        raise FgxError(code, lib.fgx_last_error().decode(errors="replace"))


def exported_symbols():
    """Names the C ABI header declares (checked by the tests)."""
    return list(_sigs)


class Params:
    """fgx_params: the reference's .par keys (include/Param.cuh) and flags."""

    def __init__(self, text=None, path=None, **keys):
        self.buf = ctypes.create_string_buffer(lib.fgx_sizeof_params())
        if path is not None:
            _check(lib.fgx_params_load(os.fsencode(path), self.buf))
        elif text is not None:
            _check(lib.fgx_params_parse(text.encode(), self.buf))
        else:
            _check(lib.fgx_params_default(self.buf))
        for k, v in keys.items():
            self.set(k, v)

    def set(self, key, value):
        if isinstance(value, bool):
            value = int(value)
        if isinstance(value, float):
            value = repr(value)
        _check(lib.fgx_params_set(self.buf, key.encode(), str(value).encode()))
        return self

    def get(self, key):
        out = ctypes.create_string_buffer(600)
        _check(lib.fgx_params_get(self.buf, key.encode(), out, 600))
        s = out.value.decode()
        try:
            return float(s)
        except ValueError:
            return s

    def __getitem__(self, key):
        return self.get(key)

    def copy(self):
        p = Params.__new__(Params)
        p.buf = ctypes.create_string_buffer(self.buf.raw, len(self.buf))
        return p

    @property
    def ptr(self):
        return ctypes.cast(self.buf, ctypes.c_void_p)

    def footprint(self, precision=F32):
        n = ctypes.c_size_t()
        _check(lib.fgx_footprint_bytes(self.buf, precision, ctypes.byref(n)))
        return n.value


def planets_from_text(text, params):
    pl = Planets()
    _check(lib.fgx_planets_parse(text.encode(), params.buf, ctypes.byref(pl)))
    return pl


def planets_from_file(path, params):
    pl = Planets()
    _check(lib.fgx_planets_load(os.fsencode(path), params.buf, ctypes.byref(pl)))
    return pl


def _disk_arrays(params, planets):
    n = len(params)
    sz = lib.fgx_sizeof_params()
    arr = ctypes.create_string_buffer(sz * n)
    for k, p in enumerate(params):
        ctypes.memmove(ctypes.addressof(arr) + k * sz, p.buf, sz)
    return n, arr, (Planets * n)(*planets)


def autotune(params, planets, precision=F32, n_dt=2, cache=None):
    """fgx_autotune: the tiling under which these disks advance fastest, and
    the tuner's report (a cache file hit times nothing)."""
    if isinstance(params, Params):
        params, planets = [params], [planets]
    n, arr, pls = _disk_arrays(params, planets)
    t, rep = Tiling(), TuneReport()
    _check(lib.fgx_autotune(n, arr, pls, precision, n_dt, cache.encode() if cache else None,
                            ctypes.byref(t), ctypes.byref(rep)))
    return t, rep


class Batch:
    """N disks advanced together; a single disk is a batch of one.
    tiling: a Tiling (or dict of its fields) for the batch kernels, None for
    the engine's own choice."""

    def __init__(self, params, planets, precision=F32, tiling=None):
        if isinstance(params, Params):
            params, planets = [params], [planets]
        n, arr, pls = _disk_arrays(params, planets)
        h = ctypes.c_void_p()
        if tiling is None:
            _check(lib.fgx_batch_create(n, arr, pls, precision, ctypes.byref(h)))
        else:
            if isinstance(tiling, dict):
                tiling = Tiling(**tiling)
            _check(lib.fgx_batch_create_tiled(n, arr, pls, precision, ctypes.byref(tiling), ctypes.byref(h)))
        self.h = h
        self.n = n
        self.precision = precision
        self.params = [p.copy() for p in params]
        self.shape = [(int(p.get("NRAD")), int(p.get("NSEC"))) for p in params]

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):   # a group's batch goes with its group
                lib.fgx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n_dt=1):
        _check(lib.fgx_batch_step(self.h, n_dt))

    def tiling(self):
        """The tiling the batch runs with, every field resolved."""
        t = Tiling()
        _check(lib.fgx_batch_tiling(self.h, ctypes.byref(t)))
        return t

    def freeze(self, disk, frozen=True):
        """Skip (or resume) one disk in every later DT."""
        _check(lib.fgx_batch_freeze(self.h, disk, int(bool(frozen))))

    def info(self, disk=0):
        out = DiskInfo()
        _check(lib.fgx_batch_info(self.h, disk, ctypes.byref(out)))
        return out

    def field(self, name, disk=0, ghost=False, dtype=np.float64):
        nr, ns = self.shape[disk]
        n = (nr + 1 if ghost else nr) * ns
        out = np.empty(n, dtype=dtype)
        _check(lib.fgx_batch_read_field(self.h, disk, FIELDS[name], F64 if dtype == np.float64 else F32,
                                        out.ctypes.data, n))
        return out.reshape(-1, ns)

    def set_field(self, name, values, disk=0):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        _check(lib.fgx_batch_write_field(self.h, disk, FIELDS[name], F64, v.ctypes.data, v.size))

    def checkpoint(self, path, disk=0):
        """Complete state of a disk, restorable bit-identically (fgx_batch_checkpoint)."""
        _check(lib.fgx_batch_checkpoint(self.h, disk, os.fsencode(path)))

    def restore(self, path, disk=0):
        _check(lib.fgx_batch_restore(self.h, disk, os.fsencode(path)))

    def reset_disk(self, disk, params, planets):
        """Slot `disk` takes a new disk of the same grid and modes (fgx_batch_reset_disk)."""
        _check(lib.fgx_batch_reset_disk(self.h, disk, ctypes.cast(params.buf, ctypes.c_void_p), ctypes.byref(planets)))
        self.params[disk] = params.copy()

    # single-disk conveniences (disk 0)
    def ring(self, name):
        nr, _ = self.shape[0]
        out = np.empty(nr + 1, dtype=np.float64)
        _check(lib.fgx_disk_read_ring(self.h, RINGS[name], F64, out.ctypes.data, nr + 1))
        return out

    def planets(self, disk=0):
        pl = Planets()
        _check(lib.fgx_batch_planets(self.h, disk, ctypes.byref(pl)))
        return pl

    def set_planets(self, pl):
        _check(lib.fgx_disk_set_planets(self.h, ctypes.byref(pl)))

    def force(self, x, y, smoothing, mass, noalloc=False):
        """ComputeForce's 44 values (Force.cu:31-67) of disk 0 at (x, y);
        noalloc=True uses the allocation-free entry point."""
        out = (ctypes.c_double * 44)()
        fn = lib.fgx_disk_force_noalloc if noalloc else lib.fgx_disk_force
        _check(fn(self.h, x, y, smoothing, mass, out))
        return np.array(out[:])

    def write_setup_files(self):
        _check(lib.fgx_disk_write_setup_files(self.h))

    def write_outputs(self, index):
        _check(lib.fgx_disk_write_outputs(self.h, index))

    def write_logs(self, timestep):
        _check(lib.fgx_disk_write_logs(self.h, timestep))

    # per-kernel device time (HIP events on the library's stream)
    def profile(self, on=True):
        _check(lib.fgx_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        _check(lib.fgx_profile_reset(self.h))

    def profile_stats(self):
        out = {}
        k = 0
        while True:
            name, ms, n, b = ctypes.c_char_p(), ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
            if lib.fgx_profile_get(self.h, k, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(n),
                                   ctypes.byref(b)) != 0:
                break
            out[name.value.decode()] = {"ms": ms.value, "launches": n.value, "bytes_per_launch": b.value}
            k += 1
        return out


Disk = Batch


class Group:
    """One set of disks advanced as `parts` concurrent batches (fgx_group_*):
    each batch on its own stream and host thread, a disk's fields those it
    has in one batch of all the disks.  `parts` is the list of the batches
    (views that the group owns); disk(d) is (batch, index) of disk d."""

    def __init__(self, params, planets, precision=F32, parts=2, tiling=None):
        n, arr, pls = _disk_arrays(params, planets)
        if isinstance(tiling, dict):
            tiling = Tiling(**tiling)
        h = ctypes.c_void_p()
        _check(lib.fgx_group_create(n, arr, pls, precision, parts, ctypes.byref(tiling) if tiling else None,
                                    ctypes.byref(h)))
        self.h = h
        self.n = n
        self.precision = precision
        np_ = ctypes.c_int()
        _check(lib.fgx_group_parts(h, ctypes.byref(np_)))
        self.parts = []
        for k in range(np_.value):
            bh = ctypes.c_void_p()
            _check(lib.fgx_group_batch(h, k, ctypes.byref(bh)))
            b = Batch.__new__(Batch)   # a view: the group destroys the batch
            b.h, b.precision, b.params, b.shape, b._owned = bh, precision, [], [], False
            self.parts.append(b)
        self._where = []
        for d in range(n):
            bh, ix = ctypes.c_void_p(), ctypes.c_int()
            _check(lib.fgx_group_disk(h, d, ctypes.byref(bh), ctypes.byref(ix)))
            k = next(q for q, b in enumerate(self.parts) if b.h.value == bh.value)
            self._where.append((k, ix.value))
            self.parts[k].params.append(params[d].copy())
            self.parts[k].shape.append((int(params[d].get("NRAD")), int(params[d].get("NSEC"))))
        for b in self.parts:
            b.n = len(b.shape)
        self.shape = [(int(p.get("NRAD")), int(p.get("NSEC"))) for p in params]

    def disk(self, d):
        """(batch view, index in it) of disk d."""
        k, ix = self._where[d]
        return self.parts[k], ix

    def step(self, n_dt=1):
        _check(lib.fgx_group_step(self.h, n_dt))

    def info(self, disk=0):
        b, ix = self.disk(disk)
        return b.info(ix)

    def field(self, name, disk=0, ghost=False, dtype=np.float64):
        b, ix = self.disk(disk)
        return b.field(name, ix, ghost, dtype)

    def planets(self, disk=0):
        b, ix = self.disk(disk)
        return b.planets(ix)

    def close(self):
        if getattr(self, "h", None):
            for b in self.parts:
                b.h = None   # the views' handles go with the group
            lib.fgx_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    n = ctypes.c_int()
    _check(lib.fgx_device_count(ctypes.byref(n)))
    return n.value


def run_main(argv):
    """The reference executable's main() (src/Main.cu:87-401) in-process."""
    args = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    return lib.fgx_run_main(len(argv), args)


class Server:
    """Serving loop (fgx_server_*): requests arrive at given times (seconds
    from the start of run()), are admitted under an HBM budget and served in
    concurrent batches; per-request admit/finish times come back."""

    def __init__(self, precision=F32, hbm_budget=256 << 30, max_batch_disks=64, snapshot_every=100,
                 batch_window=0.0, keep_final=False, max_live_batches=4, spare_slots=2):
        cfg = ServerConfig(precision, hbm_budget, max_batch_disks, snapshot_every, batch_window, int(keep_final),
                           max_live_batches, spare_slots)
        self.h = ctypes.c_void_p()
        _check(lib.fgx_server_create(ctypes.byref(cfg), ctypes.byref(self.h)))
        self.shapes = []

    def submit(self, params, planets, n_dt, arrival):
        rid = ctypes.c_int()
        _check(lib.fgx_server_submit(self.h, ctypes.cast(params.buf, ctypes.c_void_p), ctypes.byref(planets),
                                     int(n_dt), float(arrival), ctypes.byref(rid)))
        self.shapes.append((int(params["NRAD"]), int(params["NSEC"])))
        return rid.value

    def run(self):
        _check(lib.fgx_server_run(self.h))

    def request(self, rid):
        info = RequestInfo()
        _check(lib.fgx_server_request(self.h, rid, ctypes.byref(info)))
        return info

    def stats(self):
        out = ServerStats()
        _check(lib.fgx_server_stats_get(self.h, ctypes.byref(out)))
        return out

    def field(self, rid, name):
        nr, ns = self.shapes[rid]
        out = np.empty((nr + 1) * ns, dtype=np.float64)
        _check(lib.fgx_server_read_field(self.h, rid, FIELDS[name], F64, out.ctypes.data, out.size))
        return out.reshape(nr + 1, ns)[:nr]

    def close(self):
        if self.h:
            lib.fgx_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
