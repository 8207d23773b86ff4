"""GPU engine versus the CPU oracle through the C ABI.

FP32 fields must stay within 1e-5 relative L2 of the FP64 oracle after 10
fixed-dt steps; FP64 fields within 1e-12.  Integer results (FARGO shifts,
CFL counts) must agree exactly on identical inputs.
"""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
import oracle
from helpers import compare, pair, rel_l2

pytestmark = pytest.mark.gpu

DT_SMALL = 0.0314159265359   # one fixed substep per DT stays below the CFL limit

TOL = {F32: 1e-5, F64: 1e-12}


def _case(name):
    """(par, cfg, extra keys) of the small parity cases."""
    if name == "plain":
        par, cfg = configs.plain_disk(nrad=48, nsec=144)
        return par, cfg, {}
    if name == "plain_log_alpha":
        par, cfg = configs.plain_disk(nrad=48, nsec=160)
        return par, cfg, {"RADIALSPACING": "LOGARITHM", "VISCOSITY": 0, "ALPHAVISCOSITY": 1e-3,
                          "SIGMASLOPE": 1.0, "FLARINGINDEX": 0.25}
    if name == "adiabatic_evanescent":
        par, cfg = configs.large_disk(nrad=48, nsec=144)
        return par, cfg, {}
    if name == "adiabatic_cooling_open":
        par, cfg = configs.large_disk(nrad=40, nsec=128)
        return par, cfg, {"COOLING": "YES", "COOLINGTIME0": 6.28, "INNERBOUNDARY": "OPEN"}
    if name == "nonreflecting":
        par, cfg = configs.plain_disk(nrad=40, nsec=120)
        return par, cfg, {"INNERBOUNDARY": "NONREFLECTING"}
    if name == "nonreflecting_adiabatic":
        par, cfg = configs.large_disk(nrad=40, nsec=120)
        return par, cfg, {"INNERBOUNDARY": "NONREFLECTING"}
    if name == "open":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return par, cfg, {"INNERBOUNDARY": "OPEN"}
    if name == "standard_transport":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return par, cfg, {"TRANSPORT": "STANDARD"}
    if name == "guiding_center":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return par, cfg, {"FRAME": "GUIDING-CENTER", "FORCEDCIRCULAR": "NO"}
    if name == "two_planets":
        par, cfg = configs.plain_disk(nrad=48, nsec=144)
        cfg = "P1\t1.0\t1e-3\t0.0\tYES\tYES\nP2\t1.6\t3e-4\t0.0\tYES\tYES\n"
        return par, cfg, {"FORCEDCIRCULAR": "NO", "EXCLUDEHILL": "YES"}
    if name == "no_planets":
        # the star's potential and the indirect term still act (quirk Q5).
        # Without a planet the disk's vr is a small residual of O(1) terms
        # that cancel (rms 1e-4 at 1% noise against 1.5e-3 with the planet),
        # so its relative norm would measure the rounding of that
        # cancellation; a 10% density perturbation gives it real flow
        par, _ = configs.plain_disk(nrad=40, nsec=128, noise=1e-1)
        return par, "", {}
    if name == "ragged":
        # 37 rings: partial last strips of 8 and 12 rings; 100 sectors: one
        # partial wave of 50 lanes x 2 sectors, two source chunks of 57
        par, cfg = configs.plain_disk(nrad=37, nsec=100)
        return par, cfg, {}
    if name == "tiny":
        par, cfg = configs.plain_disk(nrad=8, nsec=32)
        return par, cfg, {}
    if name == "self_gravity":
        par, cfg = configs.sg_disk(nrad=24, nsec=96)
        return par, cfg, {}
    if name == "self_gravity_wide":
        # 3072 sectors: the row passes run 1024-thread blocks
        par, cfg = configs.sg_disk(nrad=32, nsec=3072)
        return par, cfg, {}
    if name == "self_gravity_tall":
        # 1024 rings: the column pass runs 1024-thread blocks (transforms of 2048)
        par, cfg = configs.sg_disk(nrad=1024, nsec=96)
        return par, cfg, {}
    if name == "self_gravity_radix5":
        par, cfg = configs.sg_disk(nrad=20, nsec=80, sigma0=4e-3)
        return par, cfg, {"ADIABATIC": "YES", "ALPHAVISCOSITY": 0, "VISCOSITY": 1e-5}
    raise KeyError(name)


CASES = ["plain", "plain_log_alpha", "adiabatic_evanescent", "adiabatic_cooling_open", "nonreflecting",
         "nonreflecting_adiabatic", "open", "standard_transport", "guiding_center", "two_planets",
         "self_gravity", "self_gravity_radix5", "no_planets", "ragged", "tiny"]


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", CASES)
def test_ten_fixed_steps(gpu, case, precision):
    par, cfg, keys = _case(case)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, SUBSTEPS=1, **keys)
    disk.step(10)
    orc.step(10)
    err = compare(disk, orc, names)
    print(case, precision, err)
    for n, e in err.items():
        assert e < TOL[precision], (case, n, e)
    # planets: FP64 state on both sides
    gp, op = disk.planets(), orc.planets(type(disk.planets()))
    for k in range(gp.nb):
        assert abs(gp.x[k] - op.x[k]) < 1e-9 and abs(gp.y[k] - op.y[k]) < 1e-9
    assert abs(disk.info().physical_time - orc.info()["time"]) < 1e-12


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["self_gravity_wide", "self_gravity_tall"])
def test_self_gravity_large_transforms_steps(gpu, case, precision):
    """Self-gravitating disks whose transforms run 1024-thread blocks (rows of
    3072 sectors; columns of 2048 for 1024 rings) stepped with CFL-driven
    substeps (their cells are too thin for one fixed substep per DT): fields
    and substep counts against the oracle after two DTs."""
    par, cfg, keys = _case(case)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, **keys)
    disk.step(2)
    orc.step(2)
    assert disk.info().substeps == orc.info()["substeps"]
    err = compare(disk, orc, names)
    for n, e in err.items():
        assert e < TOL[precision], (case, n, e)


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("nsec", [384, 512, 768, 1536, 1030])
def test_wide_rings(gpu, nsec, precision):
    """Rings split over several waves (384: 2 arcs, 512 and 768: 4, 1536: 8)
    and a sector count no split fits (1030, shared-memory kernel); CFL-driven
    substeps, which must agree in number."""
    par, cfg = configs.plain_disk(nrad=16, nsec=nsec)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL)
    for _ in range(4):
        disk.step(1)
        orc.step(1)
        assert disk.info().substeps == orc.info()["substeps"]
    err = compare(disk, orc, names)
    print(nsec, precision, err)
    for n, e in err.items():
        assert e < TOL[precision], (nsec, n, e)


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["self_gravity", "self_gravity_radix5", "self_gravity_wide", "self_gravity_tall"])
def test_initial_self_gravity_field(gpu, case, precision):
    """SG accelerations of the initial density (the three transform passes)
    against the oracle's FFT solution, itself pinned to a direct sum."""
    par, cfg, keys = _case(case)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, SUBSTEPS=1, **keys)
    nr, ns = disk.shape[0]
    want = {"sg_accr": np.zeros((nr, ns)), "sg_acct": np.zeros((nr, ns))}
    # the oracle's field of the density it was just given (the GPU's)
    assert oracle.lib().orc_sg_accel(orc.h, want["sg_accr"].ctypes.data, want["sg_acct"].ctypes.data) == 0
    tol = 1e-7 if precision == F32 else 1e-12
    got = {name: disk.field(name) for name in ("sg_accr", "sg_acct")}
    if case in ("self_gravity_wide", "self_gravity_tall"):
        # On these grids the azimuthal component is about 1% of the field of
        # a near-axisymmetric disk, so its own relative error carries the
        # rounding of transforms of 2048 and 3072 points magnified a hundred
        # times (3e-12 in FP64); the bar holds for the acceleration as a
        # vector field, and each component within a hundred times it
        d2 = sum(float(np.sum((got[n] - want[n]) ** 2)) for n in got)
        w2 = sum(float(np.sum(want[n] ** 2)) for n in got)
        assert np.sqrt(d2 / w2) < tol, np.sqrt(d2 / w2)
        assert float(np.linalg.norm(want["sg_acct"])) < 0.05 * float(np.linalg.norm(want["sg_accr"]))
        for name in got:
            assert rel_l2(got[name], want[name]) < 100 * tol, (name, rel_l2(got[name], want[name]))
    else:
        for name in got:
            e = rel_l2(got[name], want[name])
            assert e < tol, (name, e)
    # planets start on SG-corrected circular orbits
    gp, op = disk.planets(), orc.planets(type(disk.planets()))
    assert abs(gp.vy[0] - op.vy[0]) < 1e-14


def test_cfl_driven_substep_counts(gpu):
    """Without a forced substep count, both sides must pick the same CFL
    counts (integer) and stay within tolerance."""
    par, cfg = configs.plain_disk(nrad=48, nsec=144)
    disk, orc, names = pair(par, cfg, F64)
    for _ in range(3):
        disk.step(1)
        orc.step(1)
        assert disk.info().substeps == orc.info()["substeps"]
    err = compare(disk, orc, names)
    assert max(err.values()) < 1e-12, err


def test_fargo_shift_bit_exact(gpu):
    """Nshift of every ring equals the reference's FP32 formula applied to the
    GPU's own ring-mean velocity and dt (identical inputs)."""
    par, cfg = configs.plain_disk(nrad=64, nsec=192)
    p, pl = configs.load(par, cfg, DT=DT_SMALL, SUBSTEPS=1)
    disk = Batch(p, pl, F32)
    disk.step(1)
    vmed = disk.ring("vmed")
    nshift = disk.ring("nshift")
    invrmed = disk.ring("invrmed")
    dt = disk.info().last_dt
    L = oracle.lib()
    for i in range(64):
        want = L.orc_fargo_shift_f32(float(vmed[i]), float(invrmed[i]), dt, 192)
        assert int(nshift[i]) == want, (i, nshift[i], want)


def test_batch_equals_alone(gpu):
    """Each disk of an uneven batch ends bit-identical to its solo run."""
    disks = configs.batch_disks(n=3, substeps=0)
    loaded = [configs.load(par, cfg, NRAD=48 + 16 * k, NSEC=96 + 32 * k) for k, (par, cfg) in enumerate(disks)]
    b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], F32)
    b.step(2)
    for k, (p, pl) in enumerate(loaded):
        solo = Batch(p, pl, F32)
        solo.step(2)
        for name in ("dens", "vrad", "vtheta"):
            assert np.array_equal(b.field(name, disk=k), solo.field(name)), (k, name)
        assert b.info(k).substeps == solo.info().substeps


def test_mass_conserved_closed(gpu):
    par, cfg = configs.plain_disk(nrad=64, nsec=192)
    p, pl = configs.load(par, cfg)
    disk = Batch(p, pl, F64)
    surf = disk.ring("surf")[:-1]
    m0 = (disk.field("dens") * surf[:, None]).sum()
    disk.step(3)
    m1 = (disk.field("dens") * surf[:, None]).sum()
    assert abs(m1 / m0 - 1) < 1e-13


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["plain", "plain_log_alpha", "adiabatic_evanescent", "adiabatic_cooling_open",
                                  "nonreflecting", "two_planets", "guiding_center", "self_gravity",
                                  "self_gravity_radix5"])
def test_fused_source_equals_separate_kernels(gpu, monkeypatch, case, precision):
    """The one-pass source kernel (potential, Substep1, the self-gravity
    update, stresses, viscous update, Substep2) does the separate kernels'
    arithmetic operation for operation: after CFL-driven steps every field
    is bit-identical."""
    par, cfg, keys = _case(case)
    keys = dict(keys)
    # several strips and a ragged last one; self-gravity needs 2*NRAD to
    # factor into 2, 3 and 5 (150 = 2*3*5*5)
    keys.setdefault("NRAD", 75 if case.startswith("self_gravity") else 70)
    p, pl = configs.load(par, cfg, DT=DT_SMALL, **keys)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("FGX_FUSED_SOURCE", fused)
        monkeypatch.setenv("FGX_SOURCE_RINGS", "8")
        disk = Batch(p, pl, precision)
        disk.step(3)
        names = ("dens", "vrad", "vtheta") + (("energy",) if int(p.get("ADIABATIC_F")) else ())
        out[fused] = {n: disk.field(n) for n in names}
        out[fused]["substeps"] = disk.info().substeps
    assert out["1"]["substeps"] == out["0"]["substeps"]
    for n in out["0"]:
        if n == "substeps":
            continue
        a, b = out["1"][n], out["0"][n]
        assert np.array_equal(a, b), (case, n, float(np.abs(a - b).max()))


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["adiabatic_evanescent", "adiabatic_cooling_open", "nonreflecting_adiabatic",
                                  "lambda_doubling"])
@pytest.mark.parametrize("fused", ["1", "0"])
def test_viscosity_profile_reuse(gpu, monkeypatch, case, precision, fused):
    """The alpha viscosity profile of an adiabatic disk is re-formed only
    where the sound speed it averages has changed (each DT's start, a
    non-reflecting boundary) or where it depends on time (LambdaDoubling);
    every field ends bit-identical to re-forming it at each of its places in
    the substep, over several CFL-driven substeps per DT and two step calls."""
    if case == "lambda_doubling":
        par, cfg, keys = _case("adiabatic_evanescent")
        keys = dict(keys, LAMBDADOUBLING=0.1, CAVITYRATIO=0.5, CAVITYRADIUS=1.0, CAVITYWIDTH=1.0)
    else:
        par, cfg, keys = _case(case)
    p, pl = configs.load(par, cfg, **keys)
    monkeypatch.setenv("FGX_FUSED_SOURCE", fused)
    out = {}
    for skip in ("1", "0"):
        monkeypatch.setenv("FGX_VISC_SKIP", skip)
        b = Batch(p, pl, precision)
        b.step(2)
        b.step(1)
        out[skip] = ({n: b.field(n) for n in ("dens", "vrad", "vtheta", "energy")}, b.info().substeps,
                     b.ring("viscosity"))
        b.close()
    assert out["1"][1] == out["0"][1] and out["1"][1] > 3, (out["1"][1], out["0"][1])
    for n in out["0"][0]:
        assert np.array_equal(out["1"][0][n], out["0"][0][n]), (case, n)
    assert np.array_equal(out["1"][2], out["0"][2])


@pytest.mark.parametrize("case", ["plain", "adiabatic_evanescent", "open"])
def test_fine_tiling_equals_coarse(gpu, monkeypatch, case):
    """The short-strip tiling used when few disks are still active gives the
    same fields, bit for bit, as the default tiling (every substep forced one
    way or the other)."""
    par, cfg, keys = _case(case)
    keys = dict(keys)
    keys.setdefault("NRAD", 66)
    p, pl = configs.load(par, cfg, DT=DT_SMALL, **keys)
    out = {}
    for frac in ("0", "1e12"):
        monkeypatch.setenv("FGX_FINE_CELLS", frac)
        disk = Batch(p, pl, F32)
        disk.step(3)
        names = ("dens", "vrad", "vtheta") + (("energy",) if int(p.get("ADIABATIC_F")) else ())
        out[frac] = {n: disk.field(n) for n in names}
    for n in out["0"]:
        assert np.array_equal(out["0"][n], out["1e12"][n]), (case, n)


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("nsec", [384, 512, 768])
def test_one_wave_per_ring_layout(gpu, monkeypatch, nsec, precision):
    """FGX_THETA_ARCS=0 runs the azimuthal pass with one wave per ring (6, 8
    or 12 sectors per lane) where the default splits the ring into arcs; both
    layouts stay within the oracle's bar."""
    monkeypatch.setenv("FGX_THETA_ARCS", "0")
    par, cfg = configs.plain_disk(nrad=16, nsec=nsec)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, SUBSTEPS=1)
    disk.step(4)
    orc.step(4)
    err = compare(disk, orc, names)
    for n, e in err.items():
        assert e < TOL[precision], (nsec, n, e)


def test_narrowed_tile_lists_equal_whole_lists(gpu, monkeypatch):
    """Within a DT the tile lists are narrowed to the disks still active;
    an uneven batch (disks taking different CFL substep counts) ends bit
    for bit as with every disk's tiles launched throughout."""
    disks = configs.batch_disks(n=6, substeps=0)
    loaded = [configs.load(par, cfg, NRAD=32 + 8 * (k % 3), NSEC=96 + 32 * (k % 2)) for k, (par, cfg) in enumerate(disks)]
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("FGX_COMPACT", flag)
        b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], F32)
        b.step(3)
        out[flag] = [[b.field(n, disk=k) for n in ("dens", "vrad", "vtheta")] for k in range(len(loaded))]
        counts = [b.info(k).substeps for k in range(len(loaded))]
    assert len(set(counts)) > 1, counts          # the disks do take different counts
    for k in range(len(loaded)):
        for a, c in zip(out["0"][k], out["1"][k]):
            assert np.array_equal(a, c), k


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("nsec", [384, 1030])
def test_wide_rings_adiabatic(gpu, nsec, precision):
    """The energy, stored as a residual of its ring profile, through the
    multi-arc azimuthal kernel (384) and the shared-memory one (1030), with
    evanescent boundaries and CFL-driven substeps."""
    par, cfg = configs.large_disk(nrad=16, nsec=nsec)
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL)
    for _ in range(3):
        disk.step(1)
        orc.step(1)
        assert disk.info().substeps == orc.info()["substeps"]
    err = compare(disk, orc, names)
    print(nsec, precision, err)
    for n, e in err.items():
        assert e < TOL[precision], (nsec, n, e)


def test_frozen_disk_is_skipped(gpu):
    """A frozen disk keeps its fields and DT count while the batch steps;
    the others end as without it."""
    disks = configs.batch_disks(n=3, substeps=0)
    loaded = [configs.load(par, cfg, NRAD=24, NSEC=64) for par, cfg in disks]
    b = Batch([p for p, _ in loaded], [pl for _, pl in loaded], F32)
    b.step(1)
    b.freeze(1, True)
    d1 = b.field("dens", disk=1)
    b.step(2)
    assert np.array_equal(b.field("dens", disk=1), d1)
    assert [b.info(k).dt_steps for k in range(3)] == [3, 1, 3]
    solo = Batch(*loaded[2], F32)
    solo.step(3)
    assert np.array_equal(b.field("dens", disk=2), solo.field("dens"))


FULL = {   # the workloads' own grid sizes
    "plain_128x384": lambda: configs.plain_disk(nrad=128, nsec=384),
    "adiabatic_evanescent_512x1536": lambda: configs.large_disk(nrad=512, nsec=1536),
    "self_gravity_256x768": lambda: configs.sg_disk(nrad=256, nsec=768),
}


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", list(FULL))
def test_full_size_ten_steps(gpu, case, precision):
    """The single-disk workloads at their full grid sizes: 10 DTs of 0.0314
    with CFL-driven substeps; the same substep count as the oracle, and every
    field within the bar (FP32 against the FP64 oracle, FP64 against the
    80-bit one)."""
    par, cfg = FULL[case]()
    disk, orc, names = pair(par, cfg, precision, DT=DT_SMALL, SUBSTEPS=0)
    disk.step(10)
    orc.step(10)
    assert disk.info().substeps == orc.info()["substeps"]
    err = compare(disk, orc, names)
    print(case, precision, disk.info().substeps, err)
    for n, e in err.items():
        assert e < TOL[precision], (case, n, e)
