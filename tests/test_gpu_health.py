"""Per-disk numeric health: a disk whose density goes non-positive or not a
number, whose CFL time step is not positive, or whose CFL count exceeds
MAXSUBSTEPS is frozen where its substep starts and flagged; the other disks
of its batch finish bit-identical to their solo runs, and the step reports
FGX_ENUMERIC.  The bad states are written from the host; every kernel still
runs to its end."""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, Server, configs
from cudaprotoplanetarydisksfloat_amd.fgx import FgxError

pytestmark = pytest.mark.gpu

FGX_ENUMERIC = -3
H_DENSITY, H_NAN, H_TIMESTEP, H_SUBSTEPS = 1, 2, 4, 8
DT_SMALL = 0.0314159265359
DT_LONG = 1.5   # several CFL substeps per DT on these grids
NAMES = ("dens", "vrad", "vtheta")


def _three(dt=None, **bad_keys):
    out = []
    for k in range(3):
        par, cfg = configs.plain_disk(nrad=24 + 8 * k, nsec=96, disk_id=k)
        keys = dict(bad_keys) if k == 1 else {}
        if dt is not None:
            keys["DT"] = dt
        out.append(configs.load(par, cfg, **keys))
    return out


def _batch(loaded, precision):
    return Batch([p for p, _ in loaded], [pl for _, pl in loaded], precision)


def _solo(p, pl, precision, n_dt):
    b = Batch(p, pl, precision)
    b.step(n_dt)
    out = {n: b.field(n) for n in NAMES}
    out["info"] = b.info()
    b.close()
    return out


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("poison", ["negative", "nan"])
def test_a_bad_density_stops_only_its_disk(gpu, poison, precision):
    loaded = _three(DT_SMALL)
    b = _batch(loaded, precision)
    sig = b.field("dens", disk=1, ghost=True)
    before = sig.copy()
    sig[5, 7] = -sig[5, 7] if poison == "negative" else np.nan
    b.set_field("dens", sig, disk=1)
    written = {n: b.field(n, disk=1, ghost=True) for n in NAMES}
    assert np.isnan(written["dens"][5, 7]) == (poison == "nan")
    with pytest.raises(FgxError) as e:
        b.step(3)
    assert e.value.code == FGX_ENUMERIC
    info = b.info(1)
    assert info.health == H_DENSITY, info.health
    assert info.dt_steps == 0 and info.substeps == 0   # stopped before its first substep
    # the flagged disk keeps the fields it was stopped with
    for n in NAMES:
        got = b.field(n, disk=1, ghost=True)
        assert np.array_equal(got, written[n], equal_nan=True), n
    for k in (0, 2):
        assert b.info(k).health == 0 and b.info(k).dt_steps == 3
        solo = _solo(*loaded[k], precision, 3)
        for n in NAMES:
            assert np.array_equal(b.field(n, disk=k), solo[n]), (k, n)
    # a later step leaves the frozen disk alone and reports nothing new
    b.step(1)
    assert b.info(1).dt_steps == 0 and b.info(0).dt_steps == 4
    # thawed with its density restored, it runs again
    b.set_field("dens", before, disk=1)
    b.freeze(1, False)
    assert b.info(1).health == 0
    b.step(1)
    assert b.info(1).health == 0 and b.info(1).dt_steps == 1
    b.close()


def _capped():
    """The three disks with disk 1 allowed one substep fewer than its CFL count."""
    par, cfg = configs.plain_disk(nrad=32, nsec=96, disk_id=1)
    # the first substep's CFL count is the number of substeps the DT takes
    ncfl = _solo(*configs.load(par, cfg, DT=DT_LONG), F32, 1)["info"].substeps
    assert ncfl > 1, "the disk must need more than one substep per DT"
    return _three(DT_LONG, MAXSUBSTEPS=ncfl - 1)


def test_a_substep_count_above_the_cap_stops_only_its_disk(gpu):
    loaded = _capped()
    b = _batch(loaded, F32)
    with pytest.raises(FgxError) as e:
        b.step(2)
    assert e.value.code == FGX_ENUMERIC
    assert b.info(1).health == H_SUBSTEPS
    assert b.info(1).dt_steps == 0
    for k in (0, 2):
        solo = _solo(*loaded[k], F32, 2)
        for n in NAMES:
            assert np.array_equal(b.field(n, disk=k), solo[n]), (k, n)
    b.close()


def test_the_server_retires_only_the_unhealthy_request(gpu):
    srv = Server(precision=F32, hbm_budget=1 << 30, max_batch_disks=4, snapshot_every=0, keep_final=True)
    loaded = _capped()
    for p, pl in loaded:
        srv.submit(p, pl, 3, 0.0)
    with pytest.raises(FgxError) as e:
        srv.run()
    assert e.value.code == FGX_ENUMERIC
    assert srv.request(1).status == FGX_ENUMERIC
    assert len({srv.request(k).batch for k in range(3)}) == 1   # one batch held all three
    for k in (0, 2):
        assert srv.request(k).status == 0
        solo = _solo(*loaded[k], F32, 3)
        for n in NAMES:
            assert np.array_equal(srv.field(k, n), solo[n]), (k, n)
    srv.close()
