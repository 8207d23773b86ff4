"""The tile-shape autotuner (fgx_autotune): it returns a tiling at least as
fast as the engine's own, records it under the workload's signature and reads
it back without timing; and a batch gives the same fields bit for bit under
the tuned tiling, the engine's own and an extreme one (tall source strips,
short radial strips, the short-strip tiling never or always)."""
import os

import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, Tiling, autotune, configs

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359


def _disks(n=6):
    out = []
    for k in range(n):
        par, cfg = configs.plain_disk(nrad=40 + 12 * k, nsec=96 + 32 * (k % 3), disk_id=k)
        p, pl = configs.load(par, cfg, DT=DT_SMALL, ASPECTRATIO=0.04 + 0.005 * k)
        out.append((p, pl))
    return [p for p, _ in out], [pl for _, pl in out]


def test_tuner_records_and_reuses_its_choice(tmp_path):
    ps, pls = _disks()
    cache = str(tmp_path / "tiling.txt")
    t, rep = autotune(ps, pls, F32, n_dt=1, cache=cache)
    print("tuned", t.as_dict(), rep.as_dict(), flush=True)
    assert rep.from_cache == 0
    assert rep.candidates >= 10
    assert rep.best_rate >= rep.default_rate > 0
    assert 2 <= t.strip_rings <= 64 and 1 <= t.source_rings <= 24
    assert 1 <= t.fine_source_rings <= t.source_rings and 2 <= t.fine_strip_rings <= t.strip_rings
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].startswith(rep.signature.decode())
    # the same workload again: read back, nothing timed
    t2, rep2 = autotune(ps, pls, F32, n_dt=1, cache=cache)
    assert rep2.from_cache == 1 and rep2.candidates == 0
    assert t2.as_dict() == t.as_dict()
    # another workload (a disk fewer) gets a record of its own
    t3, rep3 = autotune(ps[:-1], pls[:-1], F32, n_dt=1, cache=cache)
    assert rep3.from_cache == 0 and rep3.signature != rep.signature
    assert len(open(cache).read().splitlines()) == 2
    # a batch created with the tuned tiling reports it
    b = Batch(ps, pls, F32, tiling=t)
    assert b.tiling().as_dict() == t.as_dict()
    b.close()


@pytest.mark.parametrize("precision", [F32, F64])
def test_fields_do_not_depend_on_the_tiling(precision):
    ps, pls = _disks()
    tilings = [None,
               Tiling(strip_rings=4, source_rings=24, fine_source_rings=1, fine_strip_rings=2, fine_cells=-1.0),
               Tiling(strip_rings=16, source_rings=5, fine_source_rings=3, fine_strip_rings=4, fine_cells=1e12)]
    results = []
    for t in tilings:
        b = Batch(ps, pls, precision, tiling=t)
        b.step(3)
        results.append([np.concatenate([b.field(n, disk=d) for n in ("dens", "vrad", "vtheta")])
                        for d in range(len(ps))])
        b.close()
    for other in results[1:]:
        for d in range(len(ps)):
            assert np.array_equal(results[0][d], other[d]), d


def test_tiling_fields_are_resolved_and_clamped():
    ps, pls = _disks(2)
    b = Batch(ps, pls, F32, tiling=Tiling(strip_rings=1, source_rings=99, fine_source_rings=50,
                                          fine_strip_rings=9, fine_cells=0.0))
    t = b.tiling()
    assert t.strip_rings == 2 and t.source_rings == 24
    assert t.fine_source_rings == 24 and t.fine_strip_rings == 2
    assert t.fine_cells == 2.0e6 or os.environ.get("FGX_FINE_CELLS")
    b.close()
