"""GPU engine versus the oracle with StubbedPhysics: FARGO's planet
accretion, outer mass source and exact centrifugal balance, which the
reference parses but does not run (src/Planet.cu:7-47; src/SideEuler.cu:143-146,
181; src/Pframeforce.cu:202-237).  Same bars as the main parity tests: FP32
within 1e-5 relative L2 of the FP64 oracle after 10 fixed-dt steps, FP64
within 1e-12.
"""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
import oracle
from helpers import compare, pair, rel_l2

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359
TOL = {F32: 1e-5, F64: 1e-12}
ON = {"STUBBEDPHYSICS": "YES"}


def _case(name):
    if name == "outer_source":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return par, cfg, {"OUTERSOURCEMASS": "YES", "IMPOSEDDISKDRIFT": -1e-4, **ON}
    if name == "outer_source_adiabatic":
        par, cfg = configs.large_disk(nrad=40, nsec=120)
        return par, cfg, {"OUTERSOURCEMASS": "YES", "IMPOSEDDISKDRIFT": 2e-5, **ON}
    if name == "accretion":
        par, _ = configs.plain_disk(nrad=48, nsec=160)
        cfg = "P1\t1.0\t1e-3\t2.0\tYES\tNO\n"
        return par, cfg, {"FORCEDCIRCULAR": "NO", "RMIN": 0.5, "RMAX": 1.8, **ON}
    if name == "accretion_two":
        par, _ = configs.plain_disk(nrad=48, nsec=144)
        cfg = "P1\t1.0\t1e-3\t1.0\tYES\tYES\nP2\t1.5\t5e-4\t3.0\tNO\tYES\n"
        return par, cfg, {"FORCEDCIRCULAR": "NO", "EXCLUDEHILL": "YES", **ON}
    if name == "centrifugal":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return par, cfg, {"CENTRIFUGALBALANCE": 1, "FLARINGINDEX": 0.25, **ON}
    if name == "centrifugal_sg":
        par, cfg = configs.sg_disk(nrad=24, nsec=96)
        return par, cfg, {"CENTRIFUGALBALANCE": 1, **ON}
    raise KeyError(name)


CASES = ["outer_source", "outer_source_adiabatic", "accretion", "accretion_two", "centrifugal", "centrifugal_sg"]


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", CASES)
def test_ten_fixed_steps_stubbed(gpu, case, precision):
    par, cfg, keys = _case(case)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, SUBSTEPS=1, **keys)
    m0 = [disk.planets().mass[k] for k in range(disk.planets().nb)]
    disk.step(10)
    orc.step(10)
    err = compare(disk, orc, names)
    print(case, precision, err)
    for n, e in err.items():
        assert e < TOL[precision], (case, n, e)
    gp, op = disk.planets(), orc.planets(type(disk.planets()))
    for k in range(gp.nb):
        assert abs(gp.x[k] - op.x[k]) < 1e-9 and abs(gp.y[k] - op.y[k]) < 1e-9
        dm = op.mass[k] - m0[k]
        if dm:   # accreted mass to the bar of the density it came from
            assert abs(gp.mass[k] - op.mass[k]) < (1e-4 if precision == F32 else 1e-9) * dm, (k, gp.mass[k], op.mass[k])
        else:
            assert gp.mass[k] == op.mass[k]
    if case.startswith("accretion"):
        assert gp.mass[0] > m0[0]


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["centrifugal", "centrifugal_sg"])
def test_centrifugal_balance_initial_velocity(gpu, case, precision):
    """The -b rotation curve, formed on the host, against the oracle's own
    (not loaded from the GPU)."""
    par, cfg, keys = _case(case)
    p, pl = configs.load(par, cfg, **keys)
    disk = Batch(p, pl, precision)
    orc = oracle.Oracle(p, pl, 1)
    e = rel_l2(disk.field("vtheta"), orc.field("vtheta"))
    assert e < (1e-7 if precision == F32 else 1e-14), e
    # and it is not the plain initial curve
    p0, pl0 = configs.load(par, cfg, **{k: v for k, v in keys.items() if k != "STUBBEDPHYSICS"})
    assert rel_l2(Batch(p0, pl0, precision).field("vtheta"), orc.field("vtheta")) > 1e-6


def test_batch_with_accretion_equals_alone(gpu):
    """An accreting disk and two that do not, in one batch: each ends
    bit-identical to its solo run (the accreting disk's split substep start
    must leave the others untouched)."""
    par, _, keys = _case("accretion")
    acc = configs.load(par, "P1\t1.0\t1e-3\t2.0\tYES\tNO\n", **keys)
    # a zero rate, and a rate without StubbedPhysics: neither accretes
    quiet = configs.load(par, "P1\t1.2\t5e-4\t0.0\tYES\tNO\n", NRAD=40, NSEC=128, **keys)
    off = configs.load(par, "P1\t0.9\t1e-3\t2.0\tYES\tNO\n", NRAD=56, NSEC=192,
                       **{k: v for k, v in keys.items() if k != "STUBBEDPHYSICS"})
    plain = [quiet, off]
    loaded = [quiet, acc, off]
    b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], F32)
    b.step(2)
    for k, (p, pl) in enumerate(loaded):
        solo = Batch(p, pl, F32)
        solo.step(2)
        for name in ("dens", "vrad", "vtheta"):
            assert np.array_equal(b.field(name, disk=k), solo.field(name)), (k, name)
        assert b.planets(k).mass[0] == solo.planets().mass[0]
    assert b.planets(1).mass[0] > 1e-3
    assert b.planets(0).mass[0] == plain[0][1].mass[0] and b.planets(2).mass[0] == plain[1][1].mass[0]


def test_accretion_conserves_mass_on_the_gpu(gpu):
    """Closed boundaries, no outflow: gas plus planet mass is constant."""
    par, _, keys = _case("accretion")
    p, pl = configs.load(par, "P1\t1.0\t1e-3\t2.0\tYES\tNO\n", **keys)
    disk = Batch(p, pl, F64)
    surf = disk.ring("surf")[:-1]
    total = lambda: (disk.field("dens") * surf[:, None]).sum() + disk.planets().mass[0]
    t0 = total()
    disk.step(3)
    assert disk.planets().mass[0] > 1e-3
    assert abs(total() / t0 - 1) < 1e-13


# ------------------------------------------------------------- label advection
LAB = {"LABELADVECTION": "YES", **ON}


def _label_case(name):
    if name == "label":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return par, cfg, dict(LAB)
    if name == "label_adiabatic":
        par, cfg = configs.large_disk(nrad=40, nsec=120)
        return par, cfg, dict(LAB)
    if name == "label_multiarc":      # two arcs of one wave each
        par, cfg = configs.plain_disk(nrad=16, nsec=384)
        return par, cfg, dict(LAB)
    if name == "label_lds":           # no arc split fits: shared-memory kernel
        par, cfg = configs.plain_disk(nrad=16, nsec=1030)
        return par, cfg, dict(LAB)
    raise KeyError(name)


def _smooth_label(nr, ns, seed=3):
    rng = np.random.RandomState(seed)
    i = np.arange(nr + 1)[:, None] / nr
    j = np.arange(ns)[None, :] * 2 * np.pi / ns
    lab = 1.0 + 0.5 * np.sin(3 * j + 5 * i) + 0.2 * rng.rand(nr + 1, ns)
    lab[nr] = 0
    return lab


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["label", "label_adiabatic", "label_multiarc", "label_lds"])
def test_label_advection_ten_steps(gpu, case, precision):
    par, cfg, keys = _label_case(case)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, SUBSTEPS=1, **keys)
    nr, ns = disk.shape[0]
    lab = _smooth_label(nr, ns)
    disk.set_field("label", lab)
    orc.set_field("label", disk.field("label", ghost=True))
    l0 = disk.field("label")
    disk.step(10)
    orc.step(10)
    err = compare(disk, orc, names + ("label",))
    print(case, precision, err)
    for n, e in err.items():
        assert e < TOL[precision], (case, n, e)
    assert np.abs(disk.field("label") - l0).max() > 1e-3   # it moved


def test_label_without_stubbed_physics_stays_put(gpu):
    """The reference's LABELADVECTION YES leaves the label as initialised."""
    par, cfg = configs.plain_disk(nrad=32, nsec=96)
    p, pl = configs.load(par, cfg, LABELADVECTION="YES")
    disk = Batch(p, pl, F32)
    l0 = disk.field("label")
    disk.step(2)
    assert np.array_equal(disk.field("label"), l0)
    assert l0.max() == 1 and l0.min() == 0          # InitLabel: 1 inside the Hill sphere


def test_label_checkpoint_restores_bit_identically(gpu, tmp_path):
    par, cfg, keys = _label_case("label")
    p, pl = configs.load(par, cfg, **keys)
    a = Batch(p, pl, F32)
    nr, ns = a.shape[0]
    a.set_field("label", _smooth_label(nr, ns))
    a.step(2)
    ck = str(tmp_path / "lab.bin")
    a.checkpoint(ck)
    a.step(2)
    b = Batch(p, pl, F32)
    b.restore(ck)
    b.step(2)
    for n in ("dens", "vtheta", "label"):
        assert np.array_equal(a.field(n, ghost=True), b.field(n, ghost=True)), n
