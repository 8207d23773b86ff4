"""Shared helpers: a GPU disk and the CPU oracle started from one state."""
import numpy as np

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
import oracle

STATE = ("dens", "vrad", "vtheta")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# The oracle instance each engine precision is checked against: the FP64
# one for FP32 fields, the 80-bit one for FP64 fields.  At 512 x 1536 the
# FP64 oracle's own rounding already moves vr by 2.8e-12 against the 80-bit
# instance, more than the 1e-12 bar, so only the wider one can judge FP64.
ORACLE_FOR = {F32: oracle.FP64, F64: oracle.EXT80}


def pair(par, cfg, precision=F32, **keys):
    """GPU disk + oracle; the oracle is loaded with the GPU's initial fields
    so both start from bit-identical values."""
    p, pl = configs.load(par, cfg, **keys)
    disk = Batch(p, pl, precision)
    orc = oracle.Oracle(p, pl, ORACLE_FOR[precision])
    names = STATE + (("energy",) if int(p.get("ADIABATIC_F")) else ())
    for name in names:
        orc.set_field(name, disk.field(name, ghost=True))
    return disk, orc, names


def compare(disk, orc, names):
    return {n: rel_l2(disk.field(n), orc.field(n)) for n in names}
