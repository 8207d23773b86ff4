"""Workload definitions as reference-format .par / .cfg text.

Each function returns (par_text, cfg_text) in the reference's parameter
format (src/Interpret.cu:97-208, src/Psys.cu:85-134).  Keys the reference
does not know (NOISEAMPLITUDE, NOISESEED, DISKID, SUBSTEPS) are ignored by
it, so every file also runs under the reference.

- plain_disk:    128 x 384 isothermal disk, arithmetic r in [0.4, 2.5],
                 Sigma0 6.3662e-4 with slope 0.5, h = 0.05, nu = 1e-5, one
                 q = 1e-3 planet on a fixed circular orbit at a = 1, closed
                 boundaries, +-1e-4 density noise
- large_disk:    512 x 1536 adiabatic disk (gamma 1.4), log r in [0.3, 3.0],
                 alpha = 1e-3, flaring 0.25, Sigma ~ r^-1, migrating
                 q = 3e-4 planet (RK5), evanescent boundaries
- sg_disk:       256 x 768 self-gravitating disk, log r in [0.5, 4.0], Sigma ~ r^-1.5,
                 alpha = 1e-4, q = 1e-3 planet, +-1e-3 noise
- serve_requests: Poisson-arriving disk runs of the serving workload: grids
                 128x384 .. 1024x3072, 1-3 Hill-stable planets, 20% self-gravity
- batch_disks:   64 isothermal disks with NRAD in {128,192,256}, NSEC in
                 {384,512,768} and parameters drawn from a seed
"""
import numpy as np

DT_ORBIT_20 = 0.314159265359   # DT of the reference's example runs (2pi/20)
BASE_SEED = 20261203


def _par(keys):
    return "".join(f"{k}\t{v}\n" for k, v in keys.items())


def plain_disk(nrad=128, nsec=384, noise=1e-4, disk_id=0, substeps=0):
    keys = {
        "AspectRatio": 0.05, "Sigma0": 6.3662e-4, "Viscosity": 1e-5, "SigmaSlope": 0.5,
        "FlaringIndex": 0.0, "PlanetConfig": "in/Jup.cfg", "ThicknessSmoothing": 0.6,
        "ForcedCircular": "YES", "Transport": "FARGO", "InnerBoundary": "RIGID", "Disk": "YES",
        "OmegaFrame": 0.0, "Frame": "FIXED", "IndirectTerm": "YES", "Nrad": nrad, "Nsec": nsec,
        "Rmin": 0.4, "Rmax": 2.5, "RadialSpacing": "ARITHMETIC", "Ntot": 100, "Ninterm": 20,
        "DT": DT_ORBIT_20, "OutputDir": "out/", "Adiabatic": "NO",
        "NoiseAmplitude": noise, "NoiseSeed": BASE_SEED, "DiskId": disk_id,
    }
    if substeps:
        keys["Substeps"] = substeps
    cfg = "Jupiter\t1.0\t0.001\t0.0\tNO\tNO\n"
    return _par(keys), cfg


def large_disk(nrad=512, nsec=1536, noise=1e-4, disk_id=0, substeps=0):
    keys = {
        "AspectRatio": 0.05, "Sigma0": 6.3662e-4, "AlphaViscosity": 1e-3, "SigmaSlope": 1.0,
        "FlaringIndex": 0.25, "PlanetConfig": "in/planet.cfg", "ThicknessSmoothing": 0.6,
        "ForcedCircular": "NO", "Transport": "FARGO", "InnerBoundary": "EVANESCENT",
        "Disk": "YES", "OmegaFrame": 0.0, "Frame": "FIXED", "IndirectTerm": "YES",
        "Nrad": nrad, "Nsec": nsec, "Rmin": 0.3, "Rmax": 3.0, "RadialSpacing": "LOGARITHM",
        "Ntot": 2000, "Ninterm": 100, "DT": DT_ORBIT_20, "OutputDir": "out/",
        "Adiabatic": "YES", "AdiabaticIndex": 1.4, "Cooling": "NO",
        "NoiseAmplitude": noise, "NoiseSeed": BASE_SEED, "DiskId": disk_id,
    }
    if substeps:
        keys["Substeps"] = substeps
    cfg = "Planet\t1.0\t3e-4\t0.0\tYES\tYES\n"
    return _par(keys), cfg


def sg_disk(nrad=256, nsec=768, noise=1e-3, disk_id=0, substeps=0, sigma0=6.155e-4):
    """Self-gravitating disk: log r in [0.5, 4.0], Sigma ~ r^-1.5, h = 0.05,
    alpha = 1e-4, one q = 1e-3 planet at a = 1.  The default Sigma0 gives a
    disk mass of 0.01 inside [0.5, 4] (Toomre Q ~ 26 at r = 1)."""
    keys = {
        "AspectRatio": 0.05, "Sigma0": sigma0, "AlphaViscosity": 1e-4, "SigmaSlope": 1.5,
        "FlaringIndex": 0.0, "PlanetConfig": "in/Jup.cfg", "ThicknessSmoothing": 0.6,
        "ForcedCircular": "YES", "Transport": "FARGO", "InnerBoundary": "RIGID", "Disk": "YES",
        "OmegaFrame": 0.0, "Frame": "FIXED", "IndirectTerm": "YES", "Nrad": nrad, "Nsec": nsec,
        "Rmin": 0.5, "Rmax": 4.0, "RadialSpacing": "LOGARITHM", "Ntot": 1000, "Ninterm": 100,
        "DT": DT_ORBIT_20, "OutputDir": "out/", "Adiabatic": "NO", "SelfGravity": "YES",
        "NoiseAmplitude": noise, "NoiseSeed": BASE_SEED, "DiskId": disk_id,
    }
    if substeps:
        keys["Substeps"] = substeps
    cfg = "Jupiter\t1.0\t0.001\t0.0\tNO\tNO\n"
    return _par(keys), cfg


def batch_disks(n=64, seed=BASE_SEED, noise=1e-4, substeps=0):
    """Uneven isothermal batch: per-disk grid and physics drawn from a seed."""
    rng = np.random.RandomState(seed % (2 ** 32))
    out = []
    for d in range(n):
        nrad = int(rng.choice([128, 192, 256]))
        nsec = int(rng.choice([384, 512, 768]))
        q = float(10 ** rng.uniform(np.log10(1e-5), np.log10(3e-3)))
        h = float(rng.uniform(0.03, 0.08))
        alpha = float(10 ** rng.uniform(-4, -2))
        slope = float(rng.uniform(-1.5, 0.0))
        keys = {
            "AspectRatio": f"{h:.6f}", "Sigma0": 6.3662e-4, "AlphaViscosity": f"{alpha:.6e}",
            "SigmaSlope": f"{-slope:.6f}", "FlaringIndex": 0.0, "PlanetConfig": "in/planet.cfg",
            "ThicknessSmoothing": 0.6, "ForcedCircular": "YES", "Transport": "FARGO",
            "InnerBoundary": "RIGID", "Disk": "YES", "OmegaFrame": 0.0, "Frame": "FIXED",
            "IndirectTerm": "YES", "Nrad": nrad, "Nsec": nsec, "Rmin": 0.4, "Rmax": 2.5,
            "RadialSpacing": "ARITHMETIC", "Ntot": 500, "Ninterm": 100, "DT": DT_ORBIT_20,
            "OutputDir": "out/", "Adiabatic": "NO", "NoiseAmplitude": noise,
            "NoiseSeed": seed, "DiskId": d,
        }
        if substeps:
            keys["Substeps"] = substeps
        cfg = f"Planet\t1.0\t{q:.6e}\t0.0\tNO\tNO\n"
        out.append((_par(keys), cfg))
    return out


def load(par_text, cfg_text, **overrides):
    """Params and Planets of one workload (with optional key overrides)."""
    from .fgx import Params, planets_from_text
    p = Params(text=par_text)
    for k, v in overrides.items():
        p.set(k, v)
    return p, planets_from_text(cfg_text, p)


SERVE_GRIDS = ((128, 384), (256, 768), (512, 1536), (1024, 3072))
SERVE_GRID_WEIGHTS = (0.4, 0.3, 0.2, 0.1)


def _hill_stable_orbits(rng, nplanets, lo=0.7, hi=1.6):
    """Semi-major axes in [lo, hi] and masses log-U[1e-5, 3e-3] with every
    neighbouring pair at least 2*sqrt(3) mutual Hill radii apart."""
    for _ in range(1000):
        q = 10 ** rng.uniform(np.log10(1e-5), np.log10(3e-3), nplanets)
        a = np.sort(rng.uniform(lo, hi, nplanets))
        ok = True
        for k in range(nplanets - 1):
            rh = ((q[k] + q[k + 1]) / 3.0) ** (1.0 / 3.0) * 0.5 * (a[k] + a[k + 1])
            ok &= (a[k + 1] - a[k]) >= 2 * np.sqrt(3.0) * rh
        if ok:
            return a, q
    raise RuntimeError("no Hill-stable configuration found")


def serve_requests(n=40, rate=4.0, seed=BASE_SEED, dt_range=(500, 5000), sg_frac=0.2, noise=1e-4):
    """Serving workload: n requests with Poisson arrivals at `rate` per second
    (seeded), each a disk run of U[dt_range] DTs on one of SERVE_GRIDS, with
    the batch workload's physics draws, 1-3 Hill-stable planets at
    a in U[0.7, 1.6] and self-gravity (log grid, r in [0.5, 4]) for a
    fraction sg_frac.  Returns [(par_text, cfg_text, n_dt, arrival_s)]."""
    rng = np.random.RandomState((seed + 5) % (2 ** 32))
    t = 0.0
    out = []
    for d in range(n):
        t += rng.exponential(1.0 / rate)
        nrad, nsec = SERVE_GRIDS[rng.choice(len(SERVE_GRIDS), p=SERVE_GRID_WEIGHTS)]
        sg = rng.uniform() < sg_frac
        h = float(rng.uniform(0.03, 0.08))
        alpha = float(10 ** rng.uniform(-4, -2))
        slope = float(rng.uniform(-1.5, 0.0))
        npl = int(rng.randint(1, 4))
        a, q = _hill_stable_orbits(rng, npl)
        n_dt = int(rng.randint(dt_range[0], dt_range[1] + 1))
        keys = {
            "AspectRatio": f"{h:.6f}", "Sigma0": 6.3662e-4, "AlphaViscosity": f"{alpha:.6e}",
            "SigmaSlope": f"{-slope:.6f}", "FlaringIndex": 0.0, "PlanetConfig": "in/planets.cfg",
            "ThicknessSmoothing": 0.6, "ForcedCircular": "YES", "Transport": "FARGO",
            "InnerBoundary": "RIGID", "Disk": "YES", "OmegaFrame": 0.0, "Frame": "FIXED",
            "IndirectTerm": "YES", "Nrad": nrad, "Nsec": nsec,
            "Rmin": 0.5 if sg else 0.4, "Rmax": 4.0 if sg else 2.5,
            "RadialSpacing": "LOGARITHM" if sg else "ARITHMETIC",
            "Ntot": n_dt, "Ninterm": 100, "DT": DT_ORBIT_20, "OutputDir": "out/", "Adiabatic": "NO",
            "SelfGravity": "YES" if sg else "NO",
            "NoiseAmplitude": noise, "NoiseSeed": seed, "DiskId": 1000 + d,
        }
        cfg = "".join(f"P{k}\t{a[k]:.6f}\t{q[k]:.6e}\t0.0\tNO\tNO\n" for k in range(npl))
        out.append((_par(keys), cfg, n_dt, t))
    return out
