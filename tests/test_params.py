"""The .par / .cfg parser against the reference's parsing rules.

Known answers follow src/Interpret.cu:35-208 (ReadVariables, SearchVariable)
and src/Psys.cu:85-134 (InitPlanetarySystem); the expected values below are
written from those rules, not produced by the code under test.
"""
import math
import os

import pytest

from cudaprotoplanetarydisksfloat_amd import FgxError, Params, planets_from_text

REF_IN = "/root/reference/in"

TEMPLATE_LIKE = """### Disk parameters

AspectRatio     \t0.05            Thickness over Radius in the disc
Sigma0\t\t\t6.3661977237e-4\tSurface Density at r=1
Viscosity\t\t1e-5\t\tUniform kinematic viscosity
SigmaSlope\t\t0.0\t\tSlope of surface density profile.
\t\t\t\t\t#here constant.
PlanetConfig\t\tin/Jup.cfg
ThicknessSmoothing \t0.6\t\tSmoothing parameters in disk thickness
Transport\t\tFARGO
InnerBoundary\t\tRIGID\tchoose : OPEN or RIGID or NONREFLECTING
Disk\t\t\tYES
OmegaFrame     \t\t1.0
Frame\t\t\tGUIDING-CENTER
IndirectTerm\t\tYES
Nrad\t\t\t128\t\tRadial number of zones
Nsec\t\t\t384\t\tAzimuthal number of zones (sectors)
Rmin\t\t\t0.4\t\tInner boundary radius
Rmax\t\t\t2.5\t\tOuter boundary radius
RadialSpacing \t\tARITHMETIC      Zone interfaces evenly spaced
Ntot\t\t\t10001\t\tTotal number of time steps
Ninterm\t \t\t20\t\tTime steps between outputs
DT\t\t\t0.314159265359\tTime step length. 2PI = 1 orbit
OutputDir\t\tout/
SelfGravity\tNO     #choose: Yes, Z or No
Adiabatic\tNO        #choose No for an isothermal eq. of state
AdiabaticIndex\t1.4 #Default value: 1.4
"""


def f32(x):
    import numpy as np
    return float(np.float32(x))


def test_defaults_match_param_header():
    p = Params()
    # include/Param.cuh:9-62
    assert p["DT"] == 1.0 and p["SIGMA0"] == 173.0 and p["NINTERM"] == 10 and p["NTOT"] == 1501
    assert p["NRAD"] == 64 and p["NSEC"] == 64 and p["RMIN"] == 1.0 and p["RMAX"] == 1.0
    assert p["OUTPUTDIR"] == "~masset" and p["INNERBOUNDARY"] == "WALL" and p["TRANSPORT"] == "FAST"
    assert p["ASPECTRATIO"] == f32(0.05) and p["ADIABATICINDEX"] == f32(1.4) and p["COOLINGTIME0"] == f32(6.28)
    assert p["MASSTAPER"] == f32(1e-7) and p["CAVITYRATIO"] == 1.0 and p["TRANSITIONRATIO"] == 1.0
    # src/Interpret.cu:25-31
    assert p["FASTTRANSPORT"] == 1 and p["INDIRECT_TERM"] == 1 and p["ISDISK"] == 1
    assert p["WRITE_DENSITY"] == 1 and p["WRITE_VELOCITY"] == 1 and p["WRITE_ENERGY"] == 0


def test_template_like_file():
    p = Params(text=TEMPLATE_LIKE)
    assert p["NRAD"] == 128 and p["NSEC"] == 384 and p["NINTERM"] == 20 and p["NTOT"] == 10001
    assert p["DT"] == f32(0.314159265359) and p["SIGMA0"] == f32(6.3661977237e-4)
    assert p["RMIN"] == f32(0.4) and p["RMAX"] == f32(2.5) and p["VISCOSITY"] == f32(1e-5)
    assert p["PLANETCONFIG"] == "in/Jup.cfg" and p["OUTPUTDIR"] == "out/"
    assert p["COROTATING"] == 1 and p["GUIDINGCENTER"] == 1       # GUIDING-CENTER sets both
    assert p["OPENINNER"] == 0 and p["NONREFLECTING"] == 0 and p["EVANESCENT"] == 0
    assert p["LOGGRID"] == 0 and p["ADIABATIC_F"] == 0 and p["SELFGRAVITY_F"] == 0
    assert p["OMEGAFRAME"] == 1.0


def test_separators_and_case():
    # value starts after any run of "\t :=>_"; keys are case-insensitive
    p = Params(text="nrad :=> 96\nnSeC\t_ 48\nrmin 0.5\nrmax 2\nthicknesssmoothing 0.6\n")
    assert p["NRAD"] == 96 and p["NSEC"] == 48
    # the key is the first whitespace-delimited token, so "nrad:=>" is one
    # unknown key and the line is ignored
    assert Params(text="nrad:=> 96\nthicknesssmoothing 0.6\n")["NRAD"] == 64
    # values are case-sensitive: "Open" is not OPEN
    p = Params(text="InnerBoundary Open\nThicknessSmoothing 0.6\n")
    assert p["OPENINNER"] == 0
    p = Params(text="innerboundary OPEN\nThicknessSmoothing 0.6\n")
    assert p["OPENINNER"] == 1
    # "key=value" without a separator is one unknown token
    p = Params(text="Nrad=32\nThicknessSmoothing 0.6\n")
    assert p["NRAD"] == 64
    # comments only when the first token starts with '#'
    p = Params(text="#Nrad 32\nNsec 32 # trailing text is ignored\nThicknessSmoothing 0.6\n")
    assert p["NRAD"] == 64 and p["NSEC"] == 32


def test_flag_rules():
    base = "ThicknessSmoothing 0.6\n"
    assert Params(text=base + "RadialSpacing L\n")["LOGGRID"] == 0            # only LOGARITHM
    assert Params(text=base + "RadialSpacing LOGARITHM\n")["LOGGRID"] == 1
    assert Params(text=base + "Transport STANDARD\n")["FASTTRANSPORT"] == 0
    p = Params(text=base + "Adiabatic YES\n")
    assert p["ADIABATIC_F"] == 1 and p["WRITE_TEMPERATURE"] == 1              # Interpret.cu:159-162
    assert Params(text=base + "Adiabatic YES\nAdiabaticIndex 1.0\n")["ADIABATIC_F"] == 0
    assert Params(text=base + "WriteDivV YES\n")["WRITE_DIVV"] == 0           # Interpret.cu:73
    p = Params(text=base + "SelfGravity Z\n")
    assert p["SELFGRAVITY_F"] == 1 and p["SGZEROMODE"] == 1
    assert Params(text=base + "Frame COROTATING\n")["GUIDINGCENTER"] == 0
    assert Params(text=base + "AlphaViscosity 1e-3\n")["VISCOSITYALPHA"] == 1
    assert Params(text="RocheSmoothing 0.5\n")["ROCHESMOOTHING_F"] == 1
    assert Params(text=base + "IndirectTerm NO\n")["INDIRECT_TERM"] == 0


@pytest.mark.parametrize("text", [
    "Viscosity 1e-5\nAlphaViscosity 1e-3\nThicknessSmoothing 0.6\n",   # Interpret.cu:173-179
    "ThicknessSmoothing 0.6\nRocheSmoothing 0.5\n",                   # Interpret.cu:184-190
    "Nrad 32\n",                                                      # Interpret.cu:191-197
])
def test_invalid_combinations_are_errors(text):
    with pytest.raises(FgxError) as e:
        Params(text=text)
    assert e.value.code == -1


def test_extension_keys():
    p = Params(text="ThicknessSmoothing 0.6\nNoiseAmplitude 1e-4\nNoiseSeed 7\nDiskId 3\nSubsteps 5\n")
    assert p["NOISEAMPLITUDE"] == 1e-4 and p["NOISESEED"] == 7 and p["DISKID"] == 3 and p["SUBSTEPS"] == 5


def test_planet_config():
    p = Params(text="ThicknessSmoothing 0.6\n")
    pl = planets_from_text("# Planet Name Distance Mass Accretion FeelsDisk FeelsOthers\n"
                           "Jupiter\t\t1.0\t\t0.001\t 0.0\t\tNO\t\tNO\n"
                           "Neptune 2.2 0.0001 0.0 Yes no\n", p)
    assert pl.nb == 2
    # Psys.cu:120-128: x = d(1+e), vy = sqrt((1+m)/d) sqrt((1-e)/(1+e)), vx = -1e-10 vy
    assert pl.x[0] == 1.0 and pl.y[0] == 0.0 and pl.mass[0] == 0.001
    assert pl.vy[0] == pytest.approx(math.sqrt(1.001), rel=1e-15)
    assert pl.vx[0] == pytest.approx(-1e-10 * math.sqrt(1.001), rel=1e-15)
    assert pl.feel_disk[0] == 0 and pl.feel_others[0] == 0
    assert pl.feel_disk[1] == 1 and pl.feel_others[1] == 0       # tolower(first char) == 'n'
    assert pl.vy[1] == pytest.approx(math.sqrt(1.0001 / 2.2), rel=1e-15)
    pe = Params(text="ThicknessSmoothing 0.6\nEccentricity 0.1\n")
    ple = planets_from_text("P 1.0 0.001 0 NO NO\n", pe)
    e = f32(0.1)
    assert ple.x[0] == pytest.approx(1.0 + e, rel=1e-15)
    assert ple.vy[0] == pytest.approx(math.sqrt(1.001) * math.sqrt((1 - e) / (1 + e)), rel=1e-15)


def test_cic_planet_sits_on_a_ring_edge():
    p = Params(text="ThicknessSmoothing 0.6\nNrad 10\nNsec 16\nRmin 0.5\nRmax 1.5\nCICPlanet YES\n")
    pl = planets_from_text("P 1.03 0.001 0 NO NO\n", p)
    # Psys.cu:104-109: first ring with Rmed >= d, planet placed at its outer edge
    edges = [0.5 + 0.1 * i for i in range(11)]
    rmed = [2 / 3 * (edges[i + 1] ** 3 - edges[i] ** 3) / (edges[i + 1] ** 2 - edges[i] ** 2) for i in range(10)]
    j = next(k for k in range(10) if rmed[k] >= 1.03)
    assert pl.x[0] == pytest.approx(edges[j + 1], rel=1e-6)


@pytest.mark.skipif(not os.path.isdir(REF_IN), reason="reference inputs not mounted")
def test_reference_input_files_parse():
    seen = 0
    for name in sorted(os.listdir(REF_IN)):
        if not name.endswith(".par"):
            continue
        text = open(os.path.join(REF_IN, name)).read()
        try:
            p = Params(text=text)
        except FgxError:
            continue
        assert p["NRAD"] > 0 and p["NSEC"] > 0
        seen += 1
    assert seen >= 5
    p = Params(text=open(os.path.join(REF_IN, "hd100.par")).read())
    assert p["NONREFLECTING"] == 1 and p["ADIABATIC_F"] == 1 and p["COOLING_F"] == 1
    assert p["VISCOSITYALPHA"] == 1 and p["RADIALSPACING"] == "ARITHMETHIC" and p["LOGGRID"] == 0
    p = Params(text=open(os.path.join(REF_IN, "adsg.par")).read())
    assert p["EVANESCENT"] == 1 and p["COROTATING"] == 1 and p["FORCEDCIRCULAR_F"] == 1


def test_serving_workload_sampler():
    """Poisson arrivals, Hill-stable planets, FFT-friendly self-gravity grids."""
    import numpy as np
    from cudaprotoplanetarydisksfloat_amd import configs
    reqs = configs.serve_requests(n=200, rate=4.0)
    t = np.array([r[3] for r in reqs])
    assert np.all(np.diff(t) > 0)
    # mean inter-arrival 1/rate (200 draws: within 25%)
    assert abs(np.diff(np.concatenate([[0.0], t])).mean() - 0.25) < 0.0625
    nsg = 0
    for par, cfg, n_dt, _ in reqs:
        assert 500 <= n_dt <= 5000
        p = Params(text=par)
        a = [float(l.split()[1]) for l in cfg.strip().splitlines()]
        q = [float(l.split()[2]) for l in cfg.strip().splitlines()]
        assert 1 <= len(a) <= 3 and all(0.7 <= x <= 1.6 for x in a)
        for k in range(len(a) - 1):
            rh = ((q[k] + q[k + 1]) / 3) ** (1 / 3) * 0.5 * (a[k] + a[k + 1])
            assert a[k + 1] - a[k] >= 2 * 3 ** 0.5 * rh
        if p["SELFGRAVITY_F"]:
            nsg += 1
            assert p["LOGGRID"] == 1
            for n in (int(p["NSEC"]), 2 * int(p["NRAD"])):
                while n % 2 == 0 or n % 3 == 0 or n % 5 == 0:
                    n //= 2 if n % 2 == 0 else (3 if n % 3 == 0 else 5)
                assert n == 1
    assert 20 <= nsg <= 60


def test_checkpoint_key():
    """CHECKPOINT (unknown to the reference, so .par files stay interchangeable)
    turns on restart<N>.bin at each output of fgx-run."""
    from cudaprotoplanetarydisksfloat_amd import configs
    par, _ = configs.plain_disk(nrad=16, nsec=32)
    assert int(Params(text=par).get("CHECKPOINT")) == 0
    assert int(Params(text=par + "Checkpoint YES\n").get("CHECKPOINT")) == 1
    assert int(Params(text=par + "Checkpoint 1\n").get("CHECKPOINT")) == 1
    assert int(Params(text=par + "Checkpoint NO\n").get("CHECKPOINT")) == 0
