"""Serving loop (fgx_server_*): admission, concurrent batches, retirement,
snapshots, and results identical to the same disks run alone."""
import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, Batch, Server, configs

pytestmark = pytest.mark.gpu


def _requests():
    out = []
    for k in range(5):
        par, cfg = configs.plain_disk(nrad=32 + 8 * k, nsec=96, disk_id=k)
        out.append((par, cfg, 3 + k, 0.05 * k, {"VISCOSITY": 0, "ALPHAVISCOSITY": 1e-3 * (k + 1)}))
    par, cfg = configs.sg_disk(nrad=24, nsec=96, disk_id=9)
    out.append((par, cfg, 4, 0.12, {}))
    return out


def test_serving_timestamps_and_alone_equality(gpu):
    reqs = _requests()
    srv = Server(precision=F32, hbm_budget=1 << 30, max_batch_disks=3, snapshot_every=2,
                 batch_window=0.02, keep_final=True)
    loaded = []
    for par, cfg, n_dt, t, keys in reqs:
        p, pl = configs.load(par, cfg, **keys)
        loaded.append((p, pl, n_dt))
        srv.submit(p, pl, n_dt, t)
    srv.run()
    batches = set()
    for rid, (p, pl, n_dt) in enumerate(loaded):
        info = srv.request(rid)
        print(rid, info.arrival, info.admit, info.finish, info.batch, info.batch_size, info.substeps, info.snapshots)
        assert info.status == 0
        assert info.admit >= info.arrival - 1e-6 and info.finish > info.admit
        assert info.n_dt == n_dt and info.substeps >= n_dt
        assert info.cell_updates == info.substeps * int(p["NRAD"]) * int(p["NSEC"])
        assert info.snapshots == (n_dt - 1) // 2
        batches.add(info.batch)
        # the same disk alone gives bit-identical fields
        alone = Batch(p, pl, F32)
        alone.step(n_dt)
        for name in ("dens", "vrad", "vtheta"):
            assert np.array_equal(srv.field(rid, name), alone.field(name)), (rid, name)
        assert alone.info().substeps == info.substeps
        alone.close()
    assert len(batches) >= 2   # the self-gravity request cannot share a batch with the others
    srv.close()


def test_serving_budget_serialises(gpu):
    """With a budget for one request at a time every request still completes,
    each admitted only after the previous one finished."""
    reqs = _requests()[:3]
    loaded = [configs.load(par, cfg, **keys) for par, cfg, _, _, keys in reqs]
    from cudaprotoplanetarydisksfloat_amd.fgx import lib
    import ctypes
    sizes = []
    for p, _ in loaded:
        b = ctypes.c_size_t()
        assert lib.fgx_footprint_bytes(ctypes.cast(p.buf, ctypes.c_void_p), F32, ctypes.byref(b)) == 0
        sizes.append(b.value)
    srv = Server(precision=F32, hbm_budget=max(sizes) + 1, max_batch_disks=8, snapshot_every=0)
    for (p, pl), (_, _, n_dt, t, _) in zip(loaded, reqs):
        srv.submit(p, pl, n_dt, 0.0)
    srv.run()
    infos = [srv.request(k) for k in range(3)]
    for a, b in zip(infos, infos[1:]):
        assert b.admit >= a.finish - 1e-6
    srv.close()


def test_continuous_batching_refills_slots(gpu):
    """Eight staggered requests of two grids, at most two live batches with
    a spare slot each: later requests take the slots that earlier ones
    leave, so the requests run in fewer batches than there are requests, and
    each still ends bit-identical to its solo run."""
    shapes = [(32, 96), (48, 144), (32, 96), (32, 96), (48, 144), (32, 96), (48, 144), (32, 96)]
    srv = Server(precision=F32, hbm_budget=1 << 30, max_batch_disks=2, snapshot_every=2, keep_final=True,
                 max_live_batches=2, spare_slots=1)
    loaded = []
    for k, (nr, ns) in enumerate(shapes):
        par, cfg = configs.plain_disk(nrad=nr, nsec=ns, disk_id=40 + k)
        p, pl = configs.load(par, cfg, SIGMA0=6e-4 * (1 + 0.1 * k), ASPECTRATIO=0.05 + 0.002 * k)
        n_dt = 2 + k % 3
        loaded.append((p, pl, n_dt))
        srv.submit(p, pl, n_dt, 0.04 * k)
    srv.run()
    st = srv.stats()
    print("batches", st.batches, "refills", st.refills, "peak live", st.peak_live, "peak bytes", st.peak_bytes)
    assert st.peak_live <= 2
    assert st.batches < len(shapes) and st.refills >= 1
    for rid, (p, pl, n_dt) in enumerate(loaded):
        info = srv.request(rid)
        assert info.status == 0 and info.admit >= info.arrival - 1e-6 and info.finish > info.admit
        assert 0 <= info.batch < st.batches and 0 <= info.slot < 2
        assert info.snapshots == (n_dt - 1) // 2
        alone = Batch(p, pl, F32)
        alone.step(n_dt)
        assert alone.info().substeps == info.substeps
        for name in ("dens", "vrad", "vtheta"):
            assert np.array_equal(srv.field(rid, name), alone.field(name)), (rid, name)
        alone.close()
    srv.close()


def test_package_imported_before_torch(gpu):
    """The package loads PyTorch's HIP runtime before its own library, so
    that a program importing the package first and initialising the GPU
    through torch afterwards still reaches the device (a batch and a served
    request), in a fresh process."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from cudaprotoplanetarydisksfloat_amd import Batch, Server, F32, configs\n"
        "import torch\n"
        "torch.cuda.get_device_properties(0)\n"
        "par, cfg = configs.plain_disk(nrad=32, nsec=96)\n"
        "p, pl = configs.load(par, cfg)\n"
        "Batch(p, pl, F32).close()\n"
        "srv = Server(precision=F32, hbm_budget=1 << 30, max_batch_disks=2, snapshot_every=0)\n"
        "srv.submit(p, pl, 2, 0.0)\n"
        "srv.run()\n"
        "assert srv.request(0).status == 0\n"
        "srv.close()\n"
        "print('ok')\n" % root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]


def test_batches_of_different_lds_needs(gpu):
    """Two batches whose self-gravity column pass needs different amounts of
    LDS (128 KB at NRAD=1024, 64 KB at NRAD=512), created one after the other
    as the serving loop's workers may: the second creation does not lower the
    first one's limit, and the first batch still steps, bit-identical to a
    batch of the same disk created alone."""
    big = configs.load(*configs.sg_disk(nrad=1024, nsec=96))
    small = configs.load(*configs.sg_disk(nrad=512, nsec=96))
    a = Batch(*big, F32)
    b = Batch(*small, F32)
    a.step(1)
    b.step(1)
    fa = a.field("dens")
    a.close()
    b.close()
    solo = Batch(*big, F32)
    solo.step(1)
    assert np.array_equal(fa, solo.field("dens"))
    solo.close()
