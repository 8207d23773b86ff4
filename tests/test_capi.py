"""The C-ABI library loads and exports every entry point include/fgx.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fgx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fgx_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "cudaprotoplanetarydisksfloat_amd", "libfgx.so"))
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    from cudaprotoplanetarydisksfloat_amd import fgx
    bound = set(fgx.exported_symbols())
    assert set(_declared()) - bound <= {"fgx_disk_info_get", "fgx_disk_read_field", "fgx_disk_write_field"}


def test_layout_sizes_agree():
    from cudaprotoplanetarydisksfloat_amd import fgx
    import oracle
    assert fgx.lib.fgx_sizeof_params() == oracle.lib().orc_sizeof_params()
    assert fgx.lib.fgx_sizeof_planets() == ctypes.sizeof(fgx.Planets) == oracle.lib().orc_sizeof_planets()


def test_create_without_device_fails_cleanly():
    from cudaprotoplanetarydisksfloat_amd import Batch, FgxError, configs, device_count
    if device_count() > 0:
        pytest.skip("a device is present")
    p, pl = configs.load(*configs.plain_disk(nrad=16, nsec=32))
    with pytest.raises(FgxError) as e:
        Batch(p, pl)
    assert e.value.code == -2


def test_group_arguments_and_missing_device():
    from cudaprotoplanetarydisksfloat_amd import FgxError, Group, configs, device_count
    loaded = [configs.load(*configs.plain_disk(nrad=16 + 8 * k, nsec=32, disk_id=k)) for k in range(3)]
    ps, pls = [p for p, _ in loaded], [pl for _, pl in loaded]
    for parts in (0, 9):   # outside 1..FGX_GROUP_MAX_PARTS, refused before any device call
        with pytest.raises(FgxError) as e:
            Group(ps, pls, parts=parts)
        assert e.value.code == -1 and "batches" in str(e.value)
    if device_count() == 0:
        # the first batch's creation fails; its message is the caller's last error
        with pytest.raises(FgxError) as e:
            Group(ps, pls, parts=2)
        assert e.value.code == -2


def test_autotune_without_device_fails_cleanly(tmp_path):
    from cudaprotoplanetarydisksfloat_amd import FgxError, autotune, configs, device_count
    if device_count() > 0:
        pytest.skip("a device is present")
    p, pl = configs.load(*configs.plain_disk(nrad=16, nsec=32))
    cache = tmp_path / "tiling.txt"
    with pytest.raises(FgxError) as e:
        autotune([p], [pl], cache=str(cache))
    assert e.value.code == -2
    assert not cache.exists()   # nothing recorded for a failed tuning
    with pytest.raises(FgxError):
        autotune([p], [pl], n_dt=0)


def test_server_without_device_reports_the_workers_message():
    """A request that fails in the server's worker thread ends with that
    thread's status and message: fgx_server_run returns the first failed
    request's status with the library's text for it, not an empty one."""
    from cudaprotoplanetarydisksfloat_amd import F32, FgxError, Server, configs, device_count
    if device_count() > 0:
        pytest.skip("a device is present")
    p, pl = configs.load(*configs.plain_disk(nrad=16, nsec=32))
    srv = Server(precision=F32, hbm_budget=1 << 30, max_batch_disks=2, snapshot_every=0)
    srv.submit(p, pl, 1, 0.0)
    with pytest.raises(FgxError) as e:
        srv.run()
    assert e.value.code == -2
    assert "request 0: no HIP device available" in str(e.value)
    assert srv.request(0).status == -2
    srv.close()
