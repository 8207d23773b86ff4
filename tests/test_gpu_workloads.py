"""The batched and serving workloads at their own sizes against the oracle,
one oracle per disk (the reference runs one disk per process,
src/Main.cu:87-98).

- The 64-disk uneven isothermal batch (NRAD in {128,192,256} x NSEC in
  {384,512,768}, seeded q/h/alpha/slope) stepped as one batch for 10 DTs of
  0.0314 with CFL-driven substeps: every disk's substep count equals its
  oracle's, every field is within 1e-5 (FP32, against the FP64 oracle) or
  1e-12 (FP64, against the 80-bit oracle).
- A serving mix with the largest grid (1024 x 3072, three planets that feel
  the disk and each other), a 512 x 1536 self-gravitating request and a
  small plain one: each request's fields equal the same request run alone,
  bit for bit, and are within 1e-5 of its own oracle.
"""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, Server, configs
import oracle
from helpers import ORACLE_FOR, rel_l2

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359
TOL = {F32: 1e-5, F64: 1e-12}
NAMES = ("dens", "vrad", "vtheta")


def _oracle_run(p, pl, precision, init, n_dt):
    orc = oracle.Oracle(p, pl, precision)
    for n, v in init.items():
        orc.set_field(n, v)
    orc.step(n_dt)
    out = {n: orc.field(n) for n in init}, orc.info()["substeps"]
    del orc
    return out


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
def test_batch64_every_disk_against_its_oracle(gpu, precision):
    loaded = [configs.load(par, cfg, DT=DT_SMALL) for par, cfg in configs.batch_disks(64)]
    b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], precision)
    init = [{n: b.field(n, disk=k, ghost=True) for n in NAMES} for k in range(len(loaded))]
    b.step(10)
    worst = {n: 0.0 for n in NAMES}
    shapes = set()
    for k, (p, pl) in enumerate(loaded):
        ref, subs = _oracle_run(p, pl, ORACLE_FOR[precision], init[k], 10)
        assert b.info(k).substeps == subs, k
        shapes.add((int(p["NRAD"]), int(p["NSEC"])))
        for n in NAMES:
            e = rel_l2(b.field(n, disk=k), ref[n])
            worst[n] = max(worst[n], e)
            assert e < TOL[precision], (k, p["NRAD"], p["NSEC"], n, e)
    print(precision, "worst", worst, "grids", sorted(shapes))
    assert len(shapes) == 9   # every NRAD x NSEC pair of the workload is present
    b.close()


def _mix():
    big, _ = configs.plain_disk(nrad=1024, nsec=3072, disk_id=31)
    three = ("P0\t0.8\t1e-3\t0.0\tYES\tYES\n"
             "P1\t1.15\t3e-4\t0.0\tYES\tYES\n"
             "P2\t1.55\t1e-4\t0.0\tYES\tYES\n")
    sg, sg_cfg = configs.sg_disk(nrad=512, nsec=1536, disk_id=32)
    small, small_cfg = configs.plain_disk(nrad=128, nsec=384, disk_id=33)
    return [(big, three, {"FORCEDCIRCULAR": "NO", "VISCOSITY": 0, "ALPHAVISCOSITY": 1e-3}, 3, 0.0),
            (sg, sg_cfg, {}, 2, 0.01),
            (small, small_cfg, {}, 4, 0.02)]


def test_serving_mix_at_full_grid_sizes(gpu):
    srv = Server(precision=F32, hbm_budget=64 << 30, max_batch_disks=8, snapshot_every=1, keep_final=True)
    loaded = []
    for par, cfg, keys, n_dt, t in _mix():
        p, pl = configs.load(par, cfg, DT=DT_SMALL, **keys)
        assert int(pl.nb) == (3 if int(p["NRAD"]) == 1024 else 1)
        loaded.append((p, pl, n_dt))
        srv.submit(p, pl, n_dt, t)
    srv.run()
    for rid, (p, pl, n_dt) in enumerate(loaded):
        info = srv.request(rid)
        assert info.status == 0 and info.n_dt == n_dt
        alone = Batch(p, pl, F32)
        init = {n: alone.field(n, ghost=True) for n in NAMES}
        alone.step(n_dt)
        assert alone.info().substeps == info.substeps
        ref, subs = _oracle_run(p, pl, ORACLE_FOR[F32], init, n_dt)
        assert subs == info.substeps
        for n in NAMES:
            got = srv.field(rid, n)
            assert np.array_equal(got, alone.field(n)), (rid, n)
            e = rel_l2(got, ref[n])
            print(rid, p["NRAD"], p["NSEC"], n, e)
            assert e < TOL[F32], (rid, n, e)
        alone.close()
    srv.close()
