"""Benchmark: aggregate cell-updates/s of the 64-disk uneven batch.

One step = one AlgoGas call (one DT, with per-disk CFL substeps) for every
disk of the batch.  value = total cell-updates (cells x substeps, summed over
disks) / wall time of the timed steps, inputs resident in HBM, with the disks
run as a group of concurrent batches (`group`: how many, picked by timing 1..3
before the timed region, and the throughput of the same disks as one batch).
The same disks advanced one at a time through the single-disk path give
`plain`.  `roofline` is the dominant kernel's algorithmic bytes per launch
over its average device time (HIP events on the library's stream, recorded
over as many steps of the disks as one batch, so that each kernel has the
device to itself); `cpu_baseline` times the FP64 OpenMP oracle on a bounded
sample of the same disks.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch|large|plain|sg]
                    [--serve-requests N]

A serving phase follows (fgx_server_*): Poisson-arriving disk runs of mixed
sizes served concurrently; its p50/p99 completion times go into "serving".
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s batched vs plain at 1e-5 rel L2; %HBM roofline; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_GINSTR = 256 * 2.4   # wave64 VALU instructions per ns: one per CU per clock


def _cells(batch):
    return [nr * ns for nr, ns in batch.shape]


def _substeps(batch):
    return [batch.info(d).substeps for d in range(batch.n)]


def _updates(batch, before):
    after = _substeps(batch)
    return sum(c * (a - b) for c, a, b in zip(_cells(batch), after, before))


def _workload(name):
    from cudaprotoplanetarydisksfloat_amd import configs
    if name == "batch":
        disks = configs.batch_disks(n=64)
        desc = {"workload": "64 isothermal disks, NRAD in {128,192,256} x NSEC in {384,512,768}, "
                            "seeded q/h/alpha/slope, per-disk CFL, FARGO, closed boundaries",
                "disks": 64, "seed": configs.BASE_SEED}
    elif name == "large":
        disks = [configs.large_disk()]
        desc = {"workload": "512x1536 adiabatic disk, log grid, alpha 1e-3, evanescent boundaries, "
                            "migrating planet (RK5)", "disks": 1}
    elif name == "sg":
        disks = [configs.sg_disk()]
        desc = {"workload": "256x768 self-gravitating disk, log r in [0.5, 4], Sigma ~ r^-1.5, Toomre Q ~ 26, "
                            "one planet", "disks": 1}
    else:
        disks = [configs.plain_disk()]
        desc = {"workload": "128x384 isothermal disk, one fixed-orbit planet, closed boundaries", "disks": 1}
    loaded = [configs.load(p, c) for p, c in disks]
    return loaded, desc


def serving(n_req, precision, dt_range=(4, 24), snapshot_every=4):
    """The serving workload (Poisson arrivals at 4/s, grids 128x384..1024x3072,
    1-3 planets, 20% self-gravity) with request lengths scaled down from
    U[500, 5000] DTs to U[dt_range] so the run takes seconds; completion time =
    finish - arrival of each request."""
    import numpy as np
    from cudaprotoplanetarydisksfloat_amd import Server, configs
    import torch
    reqs = configs.serve_requests(n=n_req, rate=4.0, dt_range=dt_range)
    budget = int(torch.cuda.get_device_properties(0).total_memory * 0.9)
    srv = Server(precision=precision, hbm_budget=budget, max_batch_disks=64, snapshot_every=snapshot_every,
                 batch_window=0.0, max_live_batches=4, spare_slots=2)
    for par, cfg, n_dt, t in reqs:
        p, pl = configs.load(par, cfg)
        srv.submit(p, pl, n_dt, t)
    t0 = time.perf_counter()
    srv.run()
    wall = time.perf_counter() - t0
    infos = [srv.request(k) for k in range(len(reqs))]
    st = srv.stats()
    srv.close()
    lat = np.array([i.finish - i.arrival for i in infos])
    cu = sum(i.cell_updates for i in infos)
    return {"requests": len(reqs), "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4), "max_s": round(float(lat.max()), 4),
            "cell_updates_per_s": cu / wall, "makespan_s": round(wall, 3),
            "last_arrival_s": round(reqs[-1][3], 3), "batches": st.batches, "slot_refills": st.refills,
            "peak_live_batches": st.peak_live, "peak_bytes": int(st.peak_bytes),
            "snapshots": int(sum(i.snapshots for i in infos)), "self_gravity_requests":
                sum(1 for par, _, _, _ in reqs if "SelfGravity\tYES" in par),
            "workload": "Poisson 4/s, grids 128x384..1024x3072, 1-3 planets, 20% SG, "
                        f"U[{dt_range[0]},{dt_range[1]}] DTs per request (scaled from U[500,5000]), "
                        f"snapshot every {snapshot_every} DTs"}


def cpu_baseline(loaded, budget_s=20.0):
    """FP64 OpenMP oracle on the first disks of the workload, one DT each,
    until about budget_s of CPU work."""
    import oracle
    total, t0, ndisk, ndt = 0, time.perf_counter(), 0, 0
    orcs = []
    while time.perf_counter() - t0 < budget_s:
        # one more DT of every disk, disks created on first use
        for k, (p, pl) in enumerate(loaded):
            if k == len(orcs):
                orcs.append(oracle.Oracle(p, pl, 1))
                ndisk += 1
            o = orcs[k]
            total += o.step(1) * o.nrad * o.nsec
            ndt += 1
            if time.perf_counter() - t0 > budget_s:
                break
    dt = time.perf_counter() - t0
    return {"value": total / dt, "unit": "cell-updates/s", "cores": oracle.lib().orc_num_threads(),
            "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"{ndisk} disk(s) of the workload, {ndt} DTs in all, {total} cell-updates in {dt:.1f} s"}


def _cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's processor string."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or None


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic_batch64_fp32.txt")


def measured_traffic(kernel, args):
    """HBM bytes and VALU wave-instructions per launch of a kernel class:
    FETCH_SIZE + WRITE_SIZE from the two rocprofv3 --pmc runs of this workload
    recorded in profiles/ (scaled by a 256 MiB copy kernel's counts), and
    SQ_INSTS_VALU from the SQ counter run; Nones for other workloads."""
    if args.workload != "batch" or args.precision != "f32" or not os.path.exists(TRAFFIC_FILE):
        return None, None, None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) >= 6 and f[0] == kernel:
            src = os.path.relpath(TRAFFIC_FILE, os.path.dirname(TRAFFIC_FILE) + "/..")
            return int(float(f[3]) + float(f[4])), int(float(f[5])), src
    return None, None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch", choices=["batch", "large", "plain", "sg"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--serve-requests", type=int, default=40,
                    help="requests of the serving phase (p50/p99 completion time); 0 skips it")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--parts", type=int, default=0,
                    help="concurrent batches the disks run as (fgx_group_*); 0: the fastest of 1..3")
    ap.add_argument("--no-tune", action="store_true",
                    help="run with the engine's own tiling instead of the autotuner's choice")
    ap.add_argument("--tiling-cache", default=None,
                    help="the autotuner's record file: a matching record is used without timing")
    ap.add_argument("--tune-only", action="store_true", help="run the autotuner, print its report and stop")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); disks shard across GPUs "
              f"by disk, which this build does not launch", file=sys.stderr)
        sys.exit(2)
    import torch
    from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, Group, autotune

    prec = F32 if args.precision == "f32" else F64
    if args.copy_mib:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE
        # counters calibrate the HBM-traffic counters of the workload kernels
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    loaded, desc = _workload(args.workload)
    # the tile shapes: the autotuner times candidate tilings on scratch
    # batches of the same disks (outside the timed region; any tiling gives
    # the same fields), or reads its record from --tiling-cache
    tiling, tuner = None, None
    if not args.no_tune:
        tiling, rep = autotune([p for p, _ in loaded], [pl for _, pl in loaded], prec, n_dt=2,
                               cache=args.tiling_cache)
        tuner = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rep.as_dict().items()}
        if args.tune_only:
            print(json.dumps({"tiling": tiling.as_dict(), "tuner": tuner}))
            return
    ps, pls = [p for p, _ in loaded], [pl for _, pl in loaded]
    # the disks as `parts` concurrent batches (fgx_group_*, each on its own
    # stream and host thread): the count is the fastest of 1..3 over two DTs
    # after one, outside the timed region (every count gives the same fields)
    cands = [args.parts] if args.parts else ([1, 2, 3] if len(loaded) > 1 else [1])
    runner, rates = None, {}
    for parts in cands:
        g = Group(ps, pls, prec, parts=parts, tiling=tiling)
        g.step(1)
        torch.cuda.synchronize()
        bf = _substeps(g)
        t0 = time.perf_counter()
        g.step(2)
        torch.cuda.synchronize()
        rates[parts] = _updates(g, bf) / (time.perf_counter() - t0)
        if runner is None or rates[parts] > rates[len(runner.parts)]:
            if runner is not None:
                runner.close()
            runner = g
        else:
            g.close()
    group = {"parts": len(runner.parts), "candidates": {str(k): round(v) for k, v in rates.items()},
             "disks_per_batch": [b.n for b in runner.parts]}
    # the same disks as one batch on one stream, brought to the same DT, so
    # that both are timed over the same DTs
    batch = Batch(ps, pls, prec, tiling=tiling)
    desc = dict(desc, tiling=batch.tiling().as_dict())
    runner.step(args.warmup)
    batch.step(3 + args.warmup)
    torch.cuda.synchronize()
    before = _substeps(runner)
    t0 = time.perf_counter()
    runner.step(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    updates = _updates(runner, before)
    value = updates / wall
    runner.close()
    before = _substeps(batch)
    t0 = time.perf_counter()
    batch.step(args.steps)
    torch.cuda.synchronize()
    group["one_batch_value"] = _updates(batch, before) / (time.perf_counter() - t0)
    group["over_one_batch"] = round(value / group["one_batch_value"], 3)

    # a second pass of the one batch over as many steps with per-kernel HIP
    # events, each kernel alone on the device (in a group, launches of
    # different batches overlap and share the CUs, so a launch's duration
    # would not be its own)
    # (the substep counts are read outside it: each disk query forms the
    # disk's circumplanetary mass with a launch of its own)
    before = _substeps(batch)
    batch.profile(True)
    batch.step(args.steps)
    torch.cuda.synchronize()
    batch.profile(False)
    prof_updates = _updates(batch, before)
    stats = batch.profile_stats()
    ncell = sum(_cells(batch))
    kern = {}
    for name, s in stats.items():
        if s["launches"] == 0:
            continue
        bpc = s["bytes_per_launch"] / ncell  # algorithmic bytes per cell of one launch
        kern[name] = {"ms": s["ms"], "launches": s["launches"], "avg_us": 1e3 * s["ms"] / s["launches"],
                      "alg_bytes_per_launch": bpc * prof_updates / s["launches"]}
    dom = max((k for k in kern if kern[k]["alg_bytes_per_launch"] > 0), key=lambda k: kern[k]["ms"])
    d = kern[dom]
    achieved = d["alg_bytes_per_launch"] / (d["avg_us"] * 1e-6) / 1e9
    sub_ms = sum(v["ms"] for v in kern.values())
    traffic, valu, traffic_src = measured_traffic(dom, args)
    hbm_frac = achieved / HBM_PEAK_GBS
    # the other roof of these kernels: VALU issue, one wave64 instruction per
    # CU per clock (4 SIMDs, 4 clocks each), 256 CUs at 2.4 GHz
    valu_issue = None
    if valu:
        rate = valu / (d["avg_us"] * 1e-6)
        valu_issue = {"instr_per_launch": valu, "achieved": round(rate / 1e9, 1), "peak": VALU_PEAK_GINSTR,
                      "unit": "G wave-instr/s", "frac": round(rate / 1e9 / VALU_PEAK_GINSTR, 4),
                      "source": traffic_src}
    binding = "valu_issue" if valu_issue and valu_issue["frac"] > hbm_frac else "hbm"
    roofline = {"bound": "hbm", "kernel": dom, "pass": "the disks as one batch on one stream, kernels alone on the device", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(hbm_frac, 4), "traffic": traffic,
                "traffic_source": traffic_src, "valu_issue": valu_issue, "binding": binding,
                "alg_bytes_per_launch": round(d["alg_bytes_per_launch"]),
                "kernel_ms_per_step": round(sub_ms / args.steps, 3),
                "avg_us": round(d["avg_us"], 2),
                "kernels": {k: {"avg_us": round(v["avg_us"], 2), "share": round(v["ms"] / sub_ms, 3),
                                "GBs": round(v["alg_bytes_per_launch"] / (v["avg_us"] * 1e-6) / 1e9, 1)}
                            for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}}

    plain = None
    if not args.no_plain and len(loaded) > 1:
        # the same disks one at a time through the single-disk path, each with
        # the tiling the autotuner picks for one disk of its grid
        solo_tiling = {}
        if not args.no_tune:
            for p, pl in loaded:
                shape = (int(p.get("NRAD")), int(p.get("NSEC")))
                if shape not in solo_tiling:
                    solo_tiling[shape] = autotune([p], [pl], prec, n_dt=2, cache=args.tiling_cache)[0]
        solo = [Batch(p, pl, prec, tiling=solo_tiling.get((int(p.get("NRAD")), int(p.get("NSEC")))))
                for p, pl in loaded]
        for b in solo:
            b.step(1)
        torch.cuda.synchronize()
        befores = [_substeps(b) for b in solo]
        t0 = time.perf_counter()
        for b in solo:
            b.step(args.steps)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        pup = sum(_updates(b, bf) for b, bf in zip(solo, befores))
        plain = {"value": pup / pwall, "unit": "cell-updates/s", "wall_s": round(pwall, 3),
                 "batched_over_plain": round(value / (pup / pwall), 3),
                 "tiling": "tuned per grid" if solo_tiling else "engine's own"}
        for b in solo:
            b.close()

    serve = serving(args.serve_requests, prec) if args.serve_requests > 0 else None
    cpu = None if args.no_cpu else cpu_baseline(loaded)
    line = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
            "config": dict(desc, cell_updates=updates, cells=ncell), "roofline": roofline,
            "cpu_baseline": cpu, "plain": plain, "serving": serve, "tuner": tuner, "group": group}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
