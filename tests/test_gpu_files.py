"""Values of the text files the product writes, against the oracle run from
the same state: bigplanet<k>.dat (WriteBigPlanetFile, src/Output.cu:70-102),
tqwk<k>.dat and tqwk<k>_<v>.dat (UpdateLog, src/Force.cu:76-138),
used_rad.dat (src/SourceEuler.cu:159-171), dims.dat (WriteDim,
src/Output.cu:137-153), and the GASNAMES layouts of the field files."""
import os
import subprocess

import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
from cudaprotoplanetarydisksfloat_amd.fgx import Planets
import oracle
from helpers import pair

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FGX_RUN = os.path.join(ROOT, "cudaprotoplanetarydisksfloat_amd", "fgx-run")
DT_SMALL = 0.0314159265359
TOL = {F32: 1e-5, F64: 1e-11}
PREC = [pytest.param(F32, id="fp32"), pytest.param(F64, id="fp64")]
# two planets feeling the disk, the second on an eccentric orbit
CFG = "Jupiter\t1.0\t0.001\t0.0\tYES\tYES\nSaturn\t1.6\t0.0003\t0.0\tYES\tYES\n"


def _rows(path):
    return [ln.split("\t") for ln in open(path).read().splitlines()]


def _f32(v):
    return float(np.float32(v))


@pytest.mark.parametrize("precision", PREC)
def test_log_files_against_the_oracle(gpu, tmp_path, precision):
    par, _ = configs.plain_disk(nrad=64, nsec=192, noise=1e-2)
    disk, orc, _ = pair(par, CFG, precision, DT=DT_SMALL, FORCEDCIRCULAR="NO", OUTPUTDIR=str(tmp_path) + "/")
    disk.step(4)
    orc.step(4)
    disk.write_logs(4)
    pl = orc.planets(Planets)
    info = orc.info()
    tol = TOL[precision]
    for k in range(2):
        x, y, vx, vy, m = pl.x[k], pl.y[k], pl.vx[k], pl.vy[k], pl.mass[k]
        # bigplanet: step, the planet as floats, then LostMass, PhysicalTime,
        # OmegaFrame, mdcp1 and exces_mdcp as floats
        (row,) = _rows(tmp_path / ("bigplanet%d.dat" % k))
        assert row[0] == "4"
        want = [x, y, vx, vy, m, info["lost_mass"], info["time"], info["omega_frame"], info["mdcp1"],
                info["exces_mdcp"]]
        got = [float(v) for v in row[1:]]
        for c, (g, w) in enumerate(zip(got, want)):
            assert g == _f32(g), (k, c)                  # printed from a float
            assert abs(g - w) <= max(tol, 6e-8) * max(abs(w), 1e-12) + 1e-30, (k, c, g, w)
        # tqwk: torques and powers of the inner/outer disk, with and without
        # the Hill sphere, from the Force struct's floats
        r = np.hypot(x, y)
        sm = 0.6 * 0.05 * r                              # ThicknessSmoothing * h * r, FlaringIndex 0
        gf = orc.force(x, y, sm, m)
        f = [_f32(v) for v in gf]
        fxi, fxei, fxo, fxeo = f[0], f[10], f[11], f[21]
        fyi, fyei, fyo, fyeo = f[22], f[32], f[33], f[43]
        want = [x * fyi - y * fxi, x * fyo - y * fxo, x * fyei - y * fxei, x * fyeo - y * fxeo,
                vx * fxi + vy * fyi, vx * fxo + vy * fyo, vx * fxei + vy * fyei, vx * fxeo + vy * fyeo]
        scale = [abs(x * fyi) + abs(y * fxi), abs(x * fyo) + abs(y * fxo), abs(x * fyei) + abs(y * fxei),
                 abs(x * fyeo) + abs(y * fxeo), abs(vx * fxi) + abs(vy * fyi), abs(vx * fxo) + abs(vy * fyo),
                 abs(vx * fxei) + abs(vy * fyei), abs(vx * fxeo) + abs(vy * fyeo)]
        (row,) = _rows(tmp_path / ("tqwk%d.dat" % k))
        assert row[0] == "4" and len(row) == 10
        for c in range(8):
            g = float(row[1 + c])
            assert abs(g - want[c]) <= 4 * max(tol, 6e-8) * scale[c], (k, c, g, want[c])
        assert abs(float(row[9]) - info["time"]) <= 6e-8 * info["time"]
        # tqwk_<v>: the eleven Hill-cut variants of GlobalForce
        for v in range(11):
            (row,) = _rows(tmp_path / ("tqwk%d_%d.dat" % (k, v)))
            a, b, c_, e = f[v], f[11 + v], f[22 + v], f[33 + v]
            want = [x * c_ - y * a, x * e - y * b, vx * a + vy * c_, vx * b + vy * e]
            scale = [abs(x * c_) + abs(y * a), abs(x * e) + abs(y * b), abs(vx * a) + abs(vy * c_),
                     abs(vx * b) + abs(vy * e)]
            assert row[0] == "4" and len(row) == 6
            for c in range(4):
                g = float(row[1 + c])
                assert abs(g - want[c]) <= 4 * max(tol, 6e-8) * scale[c], (k, v, c, g, want[c])
    disk.close()


def test_setup_files_against_the_oracle(gpu, tmp_path):
    par, cfg = configs.plain_disk(nrad=48, nsec=144)
    for spacing in ("ARITHMETIC", "LOGARITHM"):   # the value Interpret.cu:132 compares
        out = tmp_path / spacing
        out.mkdir()
        p, pl = configs.load(par, cfg, RADIALSPACING=spacing, OUTPUTDIR=str(out) + "/", NTOT=300, NINTERM=25)
        d = Batch(p, pl, F32)
        d.write_setup_files()
        radii = np.array([float(v) for v in open(out / "used_rad.dat").read().split()])
        orc = oracle.Oracle(p, pl, oracle.FP64)
        ref = orc.ring("radii")
        assert radii.shape == (49,)
        assert np.array_equal(radii, ref), np.max(np.abs(radii - ref))
        # the radii as a log grid or an even one, from the edges
        # the edges are float keys; the even grid steps by a float step
        # (drrsep, src/SourceEuler.cu:83), the log grid is formed in double
        lo, hi = np.float32(0.4), np.float32(2.5)
        step = float((hi - lo) / np.float32(48))
        want = (np.geomspace(float(lo), float(hi), 49) if spacing == "LOGARITHM"
                else float(lo) + step * np.arange(49))
        assert np.allclose(radii, want, rtol=1e-14, atol=0)
        dims = open(out / "dims.dat").read()
        assert dims == "0\t0\t\t0\t0\t%f\t%d\t%d\t%d\n" % (np.float32(2.5), 300 // 25, 48, 144)
        # planet files are emptied at set-up
        assert os.path.getsize(out / "planet0.dat") == 0
        d.close()


def _cli_case(d, extra):
    par, cfg = configs.plain_disk(nrad=32, nsec=96)
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    open(os.path.join(d, "planet.cfg"), "w").write(cfg)
    lines = [ln for ln in par.splitlines() if ln.split()[0].upper() not in ("PLANETCONFIG", "OUTPUTDIR", "NTOT",
                                                                            "NINTERM")]
    lines += ["PlanetConfig %s" % os.path.join(d, "planet.cfg"), "OutputDir %s/" % os.path.join(d, "out"),
              "Ntot 2", "Ninterm 1"] + extra
    path = os.path.join(d, "run.par")
    open(path, "w").write("\n".join(lines) + "\n")
    r = subprocess.run([FGX_RUN, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return os.path.join(d, "out")


def test_gas_names_layouts(gpu, tmp_path):
    """GASNAMES 0: dens/dens<N>.dat as the reference writes; 1: FARGO's
    gasdens<N>.dat in the output directory; 2: both, byte-identical."""
    outs = {g: _cli_case(str(tmp_path / ("g%d" % g)), ["GasNames %d" % g, "WriteTemperature YES"]) for g in (0, 1, 2)}
    pairs = (("dens", "gasdens"), ("vrad", "gasvrad"), ("vtheta", "gasvtheta"), ("temperature", "gasTemperature"))
    for n in range(3):
        for sub, gas in pairs:
            a = os.path.join(outs[0], sub, "%s%d.dat" % (sub, n))
            b = os.path.join(outs[1], "%s%d.dat" % (gas, n))
            assert os.path.getsize(a) == 32 * 96 * 4
            assert not os.path.exists(os.path.join(outs[0], "%s%d.dat" % (gas, n)))
            assert not os.path.exists(os.path.join(outs[1], sub))
            both = [os.path.join(outs[2], sub, "%s%d.dat" % (sub, n)), os.path.join(outs[2], "%s%d.dat" % (gas, n))]
            data = [open(p, "rb").read() for p in (a, b) + tuple(both)]
            assert all(x == data[0] for x in data), (sub, n)
