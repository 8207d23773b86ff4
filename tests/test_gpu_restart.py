"""Checkpoint and restore, and the executable's restart switch.

A disk checkpointed between DTs and restored into a fresh disk of the same
parameters continues bit-identically.  fgx-run writes restart<N>.bin next to
output N when CHECKPOINT is YES, and `fgx-run -s N` resumes from it: the output
files it writes afterwards equal, byte for byte, those of the run that was
never interrupted.  (The reference parses -s but restores nothing,
src/Main.cu:147-155, 175-180.)
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, FgxError, configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FGX_RUN = os.path.join(ROOT, "cudaprotoplanetarydisksfloat_amd", "fgx-run")


def _case(name):
    if name == "plain":
        par, cfg = configs.plain_disk(nrad=40, nsec=128)
        return configs.load(par, cfg)
    if name == "adiabatic":
        par, cfg = configs.large_disk(nrad=36, nsec=96)
        return configs.load(par, cfg)
    if name == "planets":
        par, _ = configs.plain_disk(nrad=40, nsec=120)
        cfg = "P1\t1.0\t1e-3\t0.0\tYES\tYES\nP2\t1.6\t3e-4\t0.0\tYES\tYES\n"
        return configs.load(par, cfg, FORCEDCIRCULAR="NO")
    raise KeyError(name)


def _state_names(p):
    return ("dens", "vrad", "vtheta") + (("energy",) if int(p.get("ADIABATIC_F")) else ())


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["plain", "adiabatic", "planets"])
def test_restore_continues_bit_identically(gpu, tmp_path, case, precision):
    p, pl = _case(case)
    ck = str(tmp_path / "state.bin")
    a = Batch(p, pl, precision)
    a.step(3)
    a.checkpoint(ck)
    a.step(2)
    b = Batch(p, pl, precision)
    b.restore(ck)
    b.step(2)
    for n in _state_names(p):
        fa, fb = a.field(n, ghost=True), b.field(n, ghost=True)
        assert np.array_equal(fa, fb), (case, n, float(np.abs(fa - fb).max()))
    ia, ib = a.info(), b.info()
    assert ia.physical_time == ib.physical_time
    assert ia.substeps == ib.substeps and ia.dt_steps == ib.dt_steps
    ga, gb = a.planets(), b.planets()
    for k in range(ga.nb):
        assert (ga.x[k], ga.y[k], ga.vx[k], ga.vy[k]) == (gb.x[k], gb.y[k], gb.vx[k], gb.vy[k])


def test_batch_disk_restores_into_a_solo_disk(gpu, tmp_path):
    """Disk 1 of an uneven batch, checkpointed after a DT in which the other
    disks took a different number of substeps, continues alone exactly as it
    does inside the batch."""
    disks = configs.batch_disks(n=3, substeps=0)
    loaded = [configs.load(par, cfg, NRAD=40 + 8 * k, NSEC=96 + 32 * k) for k, (par, cfg) in enumerate(disks)]
    batch = Batch([p for p, _ in loaded], [pl for _, pl in loaded], F32)
    batch.step(2)
    ck = str(tmp_path / "disk1.bin")
    batch.checkpoint(ck, disk=1)
    batch.step(1)
    solo = Batch(loaded[1][0], loaded[1][1], F32)
    solo.restore(ck)
    solo.step(1)
    for n in ("dens", "vrad", "vtheta"):
        assert np.array_equal(batch.field(n, disk=1), solo.field(n)), n
    assert batch.info(1).substeps == solo.info().substeps


def test_restore_rejects_a_mismatched_or_damaged_file(gpu, tmp_path):
    par, cfg = configs.plain_disk(nrad=32, nsec=64)
    p, pl = configs.load(par, cfg)
    a = Batch(p, pl, F32)
    ck = str(tmp_path / "a.bin")
    a.checkpoint(ck)
    other = Batch(*configs.load(par, cfg, NRAD=48), F32)
    with pytest.raises(FgxError) as e:
        other.restore(ck)
    assert e.value.code == -1          # FGX_EPARAM: another grid
    wide = Batch(p, pl, F64)
    with pytest.raises(FgxError) as e:
        wide.restore(ck)
    assert e.value.code == -1          # FP32 state into an FP64 disk
    raw = bytearray(open(ck, "rb").read())
    raw[len(raw) // 2] ^= 0x40
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(FgxError) as e:
        a.restore(bad)
    assert e.value.code == -4          # FGX_EIO: checksum
    with pytest.raises(FgxError) as e:
        a.restore(str(tmp_path / "missing.bin"))
    assert e.value.code == -4


def _write_case(d, extra=""):
    """A small run of the executable: 32 x 96 disk, NTOT 6, NINTERM 2.  The
    output directory must exist, as for the reference (it only creates the
    dens/, vrad/, ... subdirectories, Output.cu:122-124)."""
    par, cfg = configs.plain_disk(nrad=32, nsec=96)
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    open(os.path.join(d, "planet.cfg"), "w").write(cfg)
    lines = []
    for line in par.splitlines():
        k = line.split()[0].upper() if line.split() else ""
        if k in ("PLANETCONFIG", "OUTPUTDIR", "NTOT", "NINTERM"):
            continue
        lines.append(line)
    lines += ["PlanetConfig %s" % os.path.join(d, "planet.cfg"), "OutputDir %s/" % os.path.join(d, "out"),
              "Ntot 6", "Ninterm 2", extra]
    path = os.path.join(d, "run.par")
    open(path, "w").write("\n".join(lines) + "\n")
    return path


def _run(args):
    r = subprocess.run([FGX_RUN] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def test_cli_writes_the_reference_file_set(gpu, tmp_path):
    par = _write_case(str(tmp_path))
    _run([par])
    out = str(tmp_path / "out")
    for n in range(4):   # outputs 0..3 at iterations 0, 2, 4, 6
        for f in ("dens", "vrad", "vtheta"):
            path = os.path.join(out, f, "%s%d.dat" % (f, n))
            assert os.path.getsize(path) == 32 * 96 * 4, path
    assert os.path.getsize(os.path.join(out, "planet0.dat")) == 0        # truncated, never written
    big = open(os.path.join(out, "bigplanet0.dat")).read().splitlines()
    assert len(big) == 7                                                 # one line per iteration 0..6
    assert len(open(os.path.join(out, "used_rad.dat")).read().split()) == 33
    # output 0 is the initial density of the same parameters
    p, pl = configs.load(open(par).read(), open(str(tmp_path / "planet.cfg")).read())
    d = Batch(p, pl, F32)
    dens0 = np.fromfile(os.path.join(out, "dens", "dens0.dat"), dtype="<f4").reshape(32, 96)
    assert np.array_equal(dens0, d.field("dens", dtype=np.float32))


def test_cli_restart_resumes_bit_identically(gpu, tmp_path):
    full = str(tmp_path / "full")
    par = _write_case(full, "Checkpoint YES")
    _run([par])
    out = os.path.join(full, "out")
    for n in range(4):
        assert os.path.exists(os.path.join(out, "restart%d.bin" % n))
    # a second directory holding what the first run had written up to output 1
    part = str(tmp_path / "part")
    par2 = _write_case(part, "Checkpoint YES")
    out2 = os.path.join(part, "out")
    shutil.copy(os.path.join(out, "restart1.bin"), out2)
    stdout = _run(["-s", "1", par2])
    assert "Restarting from output 1" in stdout
    for n in (2, 3):
        for f in ("dens", "vrad", "vtheta"):
            a = open(os.path.join(out, f, "%s%d.dat" % (f, n)), "rb").read()
            b = open(os.path.join(out2, f, "%s%d.dat" % (f, n)), "rb").read()
            assert a == b, (f, n)
        assert open(os.path.join(out, "restart%d.bin" % n), "rb").read() == \
            open(os.path.join(out2, "restart%d.bin" % n), "rb").read()
    # the resumed run's log lines are the last ones of the full run
    big = open(os.path.join(out, "bigplanet0.dat")).read().splitlines()
    big2 = open(os.path.join(out2, "bigplanet0.dat")).read().splitlines()
    assert big2 == big[len(big) - len(big2):]
    with pytest.raises(AssertionError):
        _run(["-s", "7", par2])          # no checkpoint for output 7
