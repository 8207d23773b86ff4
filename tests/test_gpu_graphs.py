"""Substeps replayed as instantiated hipGraphs (FGX_GRAPH=1) do the same work
as launched one kernel at a time: every field and substep count is
bit-identical after DTs in which the tile lists narrow, the short-strip tiling
switches on and graphs are both built and replayed from the cache, with the
host's own summary reads and synchronisations between replays."""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359


def _uneven(n=6):
    out = []
    for k in range(n):
        par, cfg = configs.plain_disk(nrad=32 + 16 * k, nsec=96 + 32 * (k % 3), disk_id=k)
        out.append(configs.load(par, cfg, DT=DT_SMALL, ASPECTRATIO=0.035 + 0.006 * k))
    return [p for p, _ in out], [pl for _, pl in out]


def _run(monkeypatch, graph, ps, pls, precision, n_dt):
    monkeypatch.setenv("FGX_GRAPH", "1" if graph else "0")
    monkeypatch.setenv("FGX_FINE_CELLS", "5e4")   # the short-strip tiling in the tail of each DT
    b = Batch(ps, pls, precision)
    out = []
    for _ in range(n_dt):   # one DT per call: summaries, narrowing and syncs between replays
        b.step(1)
        out.append([np.concatenate([b.field(n, disk=d) for n in ("dens", "vrad", "vtheta")]) for d in range(len(ps))])
    subs = [b.info(d).substeps for d in range(len(ps))]
    b.close()
    return out, subs


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
def test_graph_replay_equals_launches(monkeypatch, precision):
    ps, pls = _uneven()
    a, sa = _run(monkeypatch, False, ps, pls, precision, 4)
    g, sg = _run(monkeypatch, True, ps, pls, precision, 4)
    assert sa == sg
    for t in range(len(a)):
        for d in range(len(ps)):
            assert np.array_equal(a[t][d], g[t][d]), (t, d)


@pytest.mark.parametrize("case", ["adiabatic", "self_gravity"])
def test_graph_replay_other_modes(monkeypatch, case):
    if case == "adiabatic":
        par, cfg = configs.large_disk(nrad=48, nsec=144)
    else:
        par, cfg = configs.sg_disk(nrad=24, nsec=96)
    p, pl = configs.load(par, cfg, DT=DT_SMALL)
    a, sa = _run(monkeypatch, False, [p], [pl], F32, 3)
    g, sg = _run(monkeypatch, True, [p], [pl], F32, 3)
    assert sa == sg
    for t in range(len(a)):
        assert np.array_equal(a[t][0], g[t][0]), t
