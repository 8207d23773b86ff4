"""Disk groups (fgx_group_*): the disks of one set dealt over several batches
stepped concurrently, each on its own stream and host thread.  Every disk's
fields, substep counts and planets equal, bit for bit, those it has in one
batch of all the disks, on disks of the uneven 64-disk workload at their full
grid sizes; a disk frozen for its numeric health stops only itself."""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, FgxError, Group, configs

pytestmark = pytest.mark.gpu

FGX_ENUMERIC = -3
NAMES = ("dens", "vrad", "vtheta")


def _disks(n):
    return [configs.load(p, c) for p, c in configs.batch_disks(n=n)]


def _state(b, d):
    out = {n: b.field(n, disk=d, ghost=True) for n in NAMES}
    info = b.info(d)
    pl = b.planets(d)
    out["counts"] = (info.substeps, info.dt_steps, info.health)
    out["planets"] = np.array([pl.x[0], pl.y[0], pl.vx[0], pl.vy[0]])
    return out


def _same(a, b, what):
    assert a["counts"] == b["counts"], what
    for n in a:
        if n != "counts":
            assert np.array_equal(a[n], b[n], equal_nan=True), (what, n, int(np.sum(a[n] != b[n])))


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("parts", [2, 3])
def test_group_equals_one_batch(gpu, parts, precision):
    loaded = _disks(10)
    ps, pls = [p for p, _ in loaded], [pl for _, pl in loaded]
    one = Batch(ps, pls, precision)
    one.step(2)
    g = Group(ps, pls, precision, parts=parts)
    assert len(g.parts) == parts
    # every disk is in exactly one batch, each batch keeps the given order
    seen = sorted(g._where)
    assert len(set(seen)) == len(loaded)
    for b in g.parts:
        idx = [d for d in range(len(loaded)) if g.disk(d)[0] is b]
        assert [g.disk(d)[1] for d in idx] == list(range(len(idx)))
        assert [g.shape[d] for d in idx] == b.shape
    g.step(1)
    g.step(1)
    for d in range(len(loaded)):
        _same(_state(one, d), _state(g, d), d)
    g.close()
    one.close()


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
def test_group_freezes_only_the_unhealthy_disk(gpu, precision):
    loaded = _disks(6)
    ps, pls = [p for p, _ in loaded], [pl for _, pl in loaded]
    one = Batch(ps, pls, precision)
    g = Group(ps, pls, precision, parts=2)
    bad = 4
    for h, set_dens in ((one, lambda s: one.set_field("dens", s, disk=bad)),
                        (g, lambda s: g.disk(bad)[0].set_field("dens", s, disk=g.disk(bad)[1]))):
        sig = h.field("dens", disk=bad, ghost=True)
        sig[3, 5] = np.nan
        set_dens(sig)
    with pytest.raises(FgxError) as e1:
        one.step(2)
    with pytest.raises(FgxError) as e2:
        g.step(2)
    assert e1.value.code == FGX_ENUMERIC and e2.value.code == FGX_ENUMERIC
    assert g.info(bad).health != 0 and g.info(bad).dt_steps == 0
    for d in range(len(loaded)):
        if d != bad:
            assert g.info(d).dt_steps == 2
        _same(_state(one, d), _state(g, d), d)
    g.close()
    one.close()
