"""Slot refill (fgx_batch_reset_disk): a slot of a live batch takes a new
disk of the same grid and modes between DTs.  The new disk then evolves
bit-identically to the same disk run alone from its own set-up, and the
other disks of the batch continue bit-identically to their uninterrupted
solo runs.  A wrong grid or mode is refused before anything changes."""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
from cudaprotoplanetarydisksfloat_amd.fgx import FgxError

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359


def _plain(k, **keys):
    par, cfg = configs.plain_disk(nrad=24 + 8 * k, nsec=96, disk_id=k)
    return configs.load(par, cfg, DT=DT_SMALL, **keys)


def _adiabatic(k, **keys):
    par, cfg = configs.large_disk(nrad=24 + 8 * k, nsec=96, disk_id=k)
    return configs.load(par, cfg, DT=DT_SMALL, **keys)


def _sg(k, **keys):
    par, cfg = configs.sg_disk(nrad=24 + 6 * k, nsec=96, disk_id=k)
    return configs.load(par, cfg, DT=DT_SMALL, **keys)


# a different disk for slot 1: same grid, other surface density, slope,
# aspect ratio, noise draw and planet mass
NEW = dict(SIGMA0=9e-4, SIGMASLOPE=0.8, ASPECTRATIO=0.06, NOISESEED=77, DISKID=5)
CASES = {"isothermal": _plain, "adiabatic": _adiabatic, "self_gravity": _sg}


def _names(p):
    return ("dens", "vrad", "vtheta") + (("energy",) if int(p.get("ADIABATIC_F")) else ())


def _solo(p, pl, precision, n_dt):
    b = Batch(p, pl, precision)
    b.step(n_dt)
    out = {n: b.field(n, ghost=True) for n in _names(p)}
    out["info"] = b.info()
    pls = b.planets()
    out["planets"] = [pls.x[0], pls.y[0], pls.vx[0], pls.vy[0], pls.mass[0]]
    b.close()
    return out


def _check(b, k, ref):
    for n in _names(b.params[k]):
        assert np.array_equal(b.field(n, disk=k, ghost=True), ref[n]), (k, n)
    pls = b.planets(k)
    assert [pls.x[0], pls.y[0], pls.vx[0], pls.vy[0], pls.mass[0]] == ref["planets"], k
    assert b.info(k).substeps == ref["info"].substeps
    assert b.info(k).physical_time == ref["info"].physical_time


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", list(CASES))
def test_refilled_slot_equals_its_solo_run(gpu, case, precision):
    make = CASES[case]
    loaded = [make(k) for k in range(3)]
    b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], precision)
    b.step(2)
    new_p, new_pl = make(1, **NEW)
    new_pl.mass[0] *= 1.7
    b.freeze(1)
    b.step(1)                       # slot 1 idle for a DT
    b.reset_disk(1, new_p, new_pl)
    assert b.info(1).dt_steps == 0 and b.info(1).substeps == 0 and b.info(1).health == 0
    b.step(3)
    _check(b, 1, _solo(new_p, new_pl, precision, 3))
    for k in (0, 2):
        _check(b, k, _solo(*loaded[k], precision, 6))
    b.close()


def test_refill_refuses_a_different_grid_or_mode(gpu):
    loaded = [_plain(k) for k in range(3)]
    b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], F32)
    b.step(1)
    before = b.field("dens", disk=1, ghost=True)
    wrong = [_plain(2),                                    # 40 rings into a 32-ring slot
             _plain(1, ADIABATIC="YES"),                   # another mode
             _plain(1, TRANSPORT="STANDARD")]
    for p, pl in wrong:
        with pytest.raises(FgxError) as e:
            b.reset_disk(1, p, pl)
        assert e.value.code == -1
    assert np.array_equal(b.field("dens", disk=1, ghost=True), before)
    b.step(2)
    for k in range(3):
        ref = _solo(*loaded[k], F32, 3)
        assert np.array_equal(b.field("dens", disk=k, ghost=True), ref["dens"]), k
    b.close()
