"""MI355X-native FARGO-style polar-grid disk hydrodynamics (fgx).

The engine lives in the native library libfgx.so (HIP kernels for gfx950 plus
C++ host code behind the C ABI of include/fgx.h); this package is its Python
face.  See fgx.py for the bindings.
"""
from .fgx import (F32, F64, FIELDS, RINGS, Batch, Disk, FgxError, Group, Params, Planets, Server, Tiling,  # noqa: F401
                  TuneReport, autotune, device_count, planets_from_file, planets_from_text, run_main)
