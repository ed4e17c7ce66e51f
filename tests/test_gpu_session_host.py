"""ROCm device discovery and the session's refusal rules on the host, over a
fake sysfs tree laid out like the amdgpu driver's (the KFD topology node and
render-node files read on an MI355X box: node 1, gpu_id 36622,
gfx_target_version 90500, simd_count 1024, num_xcc 8, unique_id
11138430765219937217 = 0x9a93a7c45c377bc1, renderD128)."""
import fcntl
import os

import pytest

from gplsaco_amd import gpu_session as G

UNIQUE = 11138430765219937217


def fake_sysfs(root, *, gfx=90500, busy=0, used=None, users=()):
    nodes = root / "class" / "kfd" / "kfd" / "topology" / "nodes"
    (nodes / "0").mkdir(parents=True)
    (nodes / "0" / "gpu_id").write_text("0\n")
    (nodes / "0" / "properties").write_text("cpu_cores_count 16\nsimd_count 0\ngfx_target_version 0\n")
    (nodes / "1").mkdir()
    (nodes / "1" / "gpu_id").write_text("36622\n")
    (nodes / "1" / "properties").write_text(
        f"simd_count 1024\narray_count 32\ngfx_target_version {gfx}\nlocation_id 9728\ndrm_render_minor 128\n"
        f"max_engine_clk_fcompute 2400\nunique_id {UNIQUE}\nnum_xcc 8\nsimd_per_cu 4\n")
    device = root / "class" / "drm" / "renderD128" / "device"
    device.mkdir(parents=True)
    (device / "product_name").write_text("AMD Instinct MI355 OAM\n")
    (device / "mem_info_vram_total").write_text("309220868096\n")
    (device / "gpu_busy_percent").write_text(f"{busy}\n")
    if used is not None:
        (device / "mem_info_vram_used").write_text(f"{used}\n")
    for pid in users:
        proc = root / "class" / "kfd" / "kfd" / "proc" / str(pid)
        proc.mkdir(parents=True)
        (proc / "vram_36622").write_text("1048576\n")
    return root


def test_list_gpus_reads_the_kfd_topology(tmp_path):
    gpus = G.list_gpus(fake_sysfs(tmp_path))
    assert len(gpus) == 1
    g = gpus[0]
    assert g["uuid"] == "GPU-9a93a7c45c377bc1" and g["gfx"] == "gfx950" and g["compute_units"] == 256
    assert g["xcc"] == 8 and g["render_minor"] == 128 and g["product"] == "AMD Instinct MI355 OAM"
    assert g["vram_total"] == 309220868096 and g["vram_used"] is None and g["busy_percent"] == 0
    assert g["processes"] == []
    assert G.list_gpus(tmp_path / "absent") == []


def test_gfx_names_and_ants_rule():
    assert G.gfx_name(90500) == "gfx950" and G.gfx_name(90402) == "gfx942" and G.gfx_name(90010) == "gfx90a"
    # 64 * ceil(sqrt(n) / 16): 64 up to 256 cities, 128 up to 1024, 192 up to 2304
    assert [G.faco_ants(n) for n in (3, 256, 257, 1000, 1024, 1025, 2304, 2305)] == [64, 64, 128, 128, 128, 192,
                                                                                       192, 256]
    with pytest.raises(ValueError):
        G.faco_ants(2)


@pytest.mark.parametrize("case", ["gfx", "busy", "memory", "process", "missing"])
def test_session_refuses_a_device_that_is_not_an_idle_mi355x(tmp_path, case):
    kw = {"gfx": {"gfx": 90402}, "busy": {"busy": 40}, "memory": {"used": 4 << 30},
          "process": {"users": (os.getpid() + 1,)}, "missing": {}}[case]
    root = fake_sysfs(tmp_path / "sys", **kw)
    uuid = "GPU-0000000000000001" if case == "missing" else "GPU-9a93a7c45c377bc1"
    with pytest.raises(RuntimeError):
        G.GpuSession(uuid, lock_directory=tmp_path / "locks", sysfs=root)
    # the lock is released after a refusal
    lock = (tmp_path / "locks" / f"worker-{uuid}.lock").open("a")
    fcntl.flock(lock, fcntl.LOCK_EX | fcntl.LOCK_NB)
    lock.close()


def test_session_lock_is_exclusive(tmp_path):
    root = fake_sysfs(tmp_path / "sys")
    (tmp_path / "locks").mkdir()
    held = (tmp_path / "locks" / "worker-GPU-9a93a7c45c377bc1.lock").open("a")
    fcntl.flock(held, fcntl.LOCK_EX | fcntl.LOCK_NB)
    with pytest.raises(BlockingIOError):
        G.GpuSession("GPU-9a93a7c45c377bc1", lock_directory=tmp_path / "locks", sysfs=root)
    held.close()
    with pytest.raises(ValueError):
        G.GpuSession("GPU-9a93a7c45c377bc1", colonies=0, lock_directory=tmp_path / "locks", sysfs=root)
