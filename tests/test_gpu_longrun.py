"""Long runs against the oracle: a disk whose planet feels the disk, 200
CFL-driven DTs on the GPU and in the FP64 oracle from the same initial
state.  Every 20 DTs the two are compared on integral quantities, each
formed on the host in FP64 by one formula from either side's fields:
total disk mass, total disk angular momentum, the planet's semi-major axis
(its drift since the start), and the disk's torque on the planet (averaged
over the run).  The bounds are stated per precision below.
"""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, configs
from helpers import pair

pytestmark = pytest.mark.gpu

DTS, EVERY = 200, 20
# bound on |gpu - oracle| for each measure, relative to the oracle's value
# (mass, angular momentum, mean torque) or to the oracle's drift (a).  The
# net torque is a small difference of the inner and outer disks' torques, so
# its relative error is the largest (measured: 1.5e-3 in FP32, 2e-11 in FP64;
# the other FP32 measures stay near 1e-8 and 1e-5)
BOUND = {
    F32: {"mass": 3e-7, "ang_mom": 3e-7, "a_drift": 3e-4, "torque": 1e-2},
    F64: {"mass": 1e-14, "ang_mom": 3e-14, "a_drift": 1e-10, "torque": 1e-9},
}


class Geometry:
    def __init__(self, disk):
        nr, ns = disk.shape[0]
        self.rmed = disk.ring("rmed")[:nr]
        self.surf = disk.ring("surf")[:nr]
        ang = 2 * np.pi * np.arange(ns) / ns
        self.x = self.rmed[:, None] * np.cos(ang)[None, :]
        self.y = self.rmed[:, None] * np.sin(ang)[None, :]


def measures(g, dens, vtheta, pl, smoothing):
    """Mass, angular momentum (vtheta at the cell centre, fixed frame), the
    planet's semi-major axis (G = M_star = 1) and the disk's torque on it."""
    dm = dens * g.surf[:, None]
    vc = 0.5 * (vtheta + np.roll(vtheta, -1, axis=1))
    xp, yp, vx, vy, mp = pl.x[0], pl.y[0], pl.vx[0], pl.vy[0], pl.mass[0]
    r = np.hypot(xp, yp)
    a = 1.0 / (2.0 / r - (vx * vx + vy * vy) / (1.0 + mp))
    dx, dy = g.x - xp, g.y - yp
    inv3 = (dx * dx + dy * dy + smoothing ** 2) ** -1.5
    torque = mp * np.sum(dm * (xp * dy - yp * dx) * inv3)
    return {"mass": dm.sum(), "ang_mom": np.sum(dm * g.rmed[:, None] * vc), "a": a, "torque": torque}


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
def test_long_run_integrals_match_oracle(gpu, precision):
    par, _ = configs.plain_disk(nrad=48, nsec=128)
    cfg = "Jupiter\t1.0\t1e-3\t0.0\tYES\tNO\n"     # feels the disk, migrates
    disk, orc, _ = pair(par, cfg, precision, FORCEDCIRCULAR="NO")
    g = Geometry(disk)
    sm = 0.6 * 0.05 * 1.0                            # 0.6 H at the planet
    P = type(disk.planets())
    m0 = measures(g, disk.field("dens"), disk.field("vtheta"), disk.planets(), sm)
    gpu_t, orc_t, worst = [], [], {}
    for k in range(DTS // EVERY):
        disk.step(EVERY)
        orc.step(EVERY)
        assert disk.info().substeps == orc.info()["substeps"], k
        mg = measures(g, disk.field("dens"), disk.field("vtheta"), disk.planets(), sm)
        mo = measures(g, orc.field("dens"), orc.field("vtheta"), orc.planets(P), sm)
        gpu_t.append(mg["torque"])
        orc_t.append(mo["torque"])
        for q in ("mass", "ang_mom"):
            worst[q] = max(worst.get(q, 0.0), abs(mg[q] / mo[q] - 1))
        drift = mo["a"] - m0["a"]
        assert drift != 0      # the planet does migrate
        worst["a_drift"] = max(worst.get("a_drift", 0.0), abs(mg["a"] - mo["a"]) / abs(drift))
    worst["torque"] = abs(np.mean(gpu_t) / np.mean(orc_t) - 1)
    print(precision, worst)
    for q, e in worst.items():
        assert e < BOUND[precision][q], (q, e)
