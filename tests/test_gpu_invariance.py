"""A disk's fields do not depend on what else its batch holds or on how the
batch is tiled, at the workloads' full grid sizes: frozen spare disks beside
it, the radial strip height, the short-strip tiling and narrowed tile lists
all give the same bits as the disk alone with the default tiling.  (The
serving loop relies on this: a request in a slot of a shared batch ends as
it would alone.)"""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359
THREE = ("P0\t0.8\t1e-3\t0.0\tYES\tYES\n"
         "P1\t1.15\t3e-4\t0.0\tYES\tYES\n"
         "P2\t1.55\t1e-4\t0.0\tYES\tYES\n")


def _case(name):
    if name == "plain_1024x3072":
        par, _ = configs.plain_disk(nrad=1024, nsec=3072, disk_id=31)
        return configs.load(par, THREE, DT=DT_SMALL, FORCEDCIRCULAR="NO", VISCOSITY=0, ALPHAVISCOSITY=1e-3)
    if name == "adiabatic_512x1536":
        par, cfg = configs.large_disk(nrad=512, nsec=1536, disk_id=3)
        return configs.load(par, cfg, DT=DT_SMALL)
    par, cfg = configs.sg_disk(nrad=512, nsec=1536, disk_id=32)
    return configs.load(par, cfg, DT=DT_SMALL)


def _names(p):
    return ("dens", "vrad", "vtheta") + (("energy",) if int(p.get("ADIABATIC_F")) else ())


def _run(p, pl, precision, n_dt, spares=0):
    b = Batch([p] * (1 + spares), [pl] * (1 + spares), precision)
    for k in range(1, 1 + spares):
        b.freeze(k)
    b.step(n_dt)
    out = {n: b.field(n, ghost=True) for n in _names(p)}
    out["substeps"] = b.info(0).substeps
    b.close()
    return out


def _same(a, b, what):
    assert a["substeps"] == b["substeps"], what
    for n in a:
        if n != "substeps":
            assert np.array_equal(a[n], b[n]), (what, n, int(np.sum(a[n] != b[n])))


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["plain_1024x3072", "adiabatic_512x1536", "sg_512x1536"])
def test_frozen_spares_leave_a_disk_unchanged(gpu, case, precision):
    p, pl = _case(case)
    _same(_run(p, pl, precision, 2), _run(p, pl, precision, 2, spares=2), case)


@pytest.mark.parametrize("env", [("FGX_STRIP_RINGS", "2"), ("FGX_STRIP_RINGS", "16"), ("FGX_STRIP_RINGS", "64"),
                                 ("FGX_FINE_CELLS", "1e12"), ("FGX_SOURCE_RINGS", "5"), ("FGX_COMPACT", "0")],
                         ids=lambda e: "%s=%s" % e)
def test_tilings_at_full_size(gpu, monkeypatch, env):
    p, pl = _case("plain_1024x3072")
    ref = _run(p, pl, F32, 1)
    monkeypatch.setenv(*env)
    _same(ref, _run(p, pl, F32, 1), env)
