"""Monitoring outputs and initial conditions of the product against the
oracle: the eleven-variant disk force of UpdateLog (ComputeForce,
src/Force.cu:31-67), CircumPlanetaryMass (src/SourceEuler.cu:539-553), the
temperature field (ComputeTemperatureFieldKernel, src/Kernels.cu:211-223),
and the product's own initial fields against the oracle's own set-up
(InitGasDensity / InitGasEnergy / InitGasVelocities, src/Pframeforce.cu:173-272,
src/Theo.cu:13-51), with no field copied from one side to the other."""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
import oracle
from helpers import ORACLE_FOR, pair, rel_l2

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359
TOL = {F32: 1e-5, F64: 1e-12}
PREC = [pytest.param(F32, id="fp32"), pytest.param(F64, id="fp64")]


def _stepped(precision, n_dt=3, **keys):
    par, cfg = configs.plain_disk(nrad=64, nsec=192, noise=1e-2)
    cfg = "Jupiter\t1.0\t0.001\t0.0\tYES\tNO\n"
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, FORCEDCIRCULAR="NO", **keys)
    disk.step(n_dt)
    orc.step(n_dt)
    return disk, orc


@pytest.mark.parametrize("precision", PREC)
def test_disk_force_all_variants(gpu, precision):
    disk, orc = _stepped(precision)
    pl = disk.planets()
    x, y, m = pl.x[0], pl.y[0], pl.mass[0]
    sm = 0.6 * 0.05 * np.hypot(x, y)    # ThicknessSmoothing * h * r (Force.cu:69-74)
    for pt in ((x, y, sm, m), (0.0, 0.0, 0.0, 0.0), (0.7, -0.9, 0.02, 2e-3)):
        got = disk.force(*pt)
        ref = orc.force(*pt)
        scale = np.max(np.abs(ref))
        assert scale > 0
        err = np.max(np.abs(got - ref)) / scale
        print(precision, pt, err)
        assert err < TOL[precision], (pt, err)
        # the split is the reference's: inner and outer parts per variant
        assert np.all((ref[:11] == 0) == (got[:11] == 0))


@pytest.mark.parametrize("precision", PREC)
def test_circumplanetary_mass(gpu, precision):
    disk, orc = _stepped(precision)
    got, ref = disk.info().mdcp1, orc.info()["mdcp1"]
    print(precision, got, ref)
    assert ref > 0
    assert abs(got - ref) / ref < TOL[precision]


@pytest.mark.parametrize("precision", PREC)
@pytest.mark.parametrize("case", ["isothermal", "adiabatic"])
def test_temperature_field(gpu, case, precision):
    if case == "isothermal":
        disk, orc = _stepped(precision, WRITETEMPERATURE="YES")
    else:
        par, cfg = configs.large_disk(nrad=48, nsec=144)
        disk, orc, _ = pair(par, cfg, precision, DT=DT_SMALL)
        disk.step(3)
        orc.step(3)
    e = rel_l2(disk.field("temperature"), orc.field("temperature"))
    print(case, precision, e)
    assert e < TOL[precision]


def test_isothermal_temperature_needs_the_output_switch(gpu):
    disk, _ = _stepped(F32, n_dt=1)
    with pytest.raises(Exception, match="WRITETEMPERATURE"):
        disk.field("temperature")


IC_CASES = {
    "plain_noise": lambda: (*configs.plain_disk(nrad=64, nsec=192, noise=1e-2), {}),
    "log_flared_adiabatic": lambda: (*configs.large_disk(nrad=48, nsec=144, noise=1e-3), {}),
    "alpha_viscosity_drift": lambda: (*configs.plain_disk(nrad=40, nsec=128), {"VISCOSITY": 0, "ALPHAVISCOSITY": 3e-3,
                                                                                "IMPOSEDDISKDRIFT": 1e-4}),
    "self_gravity": lambda: (*configs.sg_disk(nrad=24, nsec=96), {}),
    "batch_disk_7": lambda: (*configs.batch_disks(8)[7], {}),
}


@pytest.mark.parametrize("precision", PREC)
@pytest.mark.parametrize("case", list(IC_CASES))
def test_initial_fields_against_the_oracle_setup(gpu, case, precision):
    """The product's set-up (seeded noise included) against the oracle's own:
    nothing is loaded from one into the other."""
    par, cfg, keys = IC_CASES[case]()
    p, pl = configs.load(par, cfg, **keys)
    disk = Batch(p, pl, precision)
    orc = oracle.Oracle(p, pl, ORACLE_FOR[precision])
    names = ["dens", "vrad", "vtheta"] + (["energy"] if int(p.get("ADIABATIC_F")) else [])
    tol = {F32: 1e-7, F64: 1e-14}[precision]
    for n in names:
        a, b = disk.field(n, ghost=True), orc.field(n, ghost=True)
        e = rel_l2(a, b)
        print(case, precision, n, e)
        assert e < tol, (n, e)
