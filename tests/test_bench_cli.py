"""bench.py on the host: more than one GPU is refused with a message (the
library runs on one GPU) before anything touches a device, and the
measurement helpers (the traffic row of the run's kernel and block size,
the kernel names, the loop clock's rate) behave as the bench line needs."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_traffic_row_matches_the_kernel_and_block_size():
    """roofline.traffic comes from the summary row that measured the kernel
    and block size of the run, and is null with the reason otherwise."""
    b = _bench()
    k64 = b.chain_kernel(("lane+lds-z", 64), 15)
    k128 = b.chain_kernel(("lane+lds-z", 128), 15)
    v64, note = b.traffic("updateBF_chain_n1e6_m15", k64)
    v128, _ = b.traffic("updateBF_chain_n1e6_m15", k128)
    assert v64 and v128 and v64 != v128
    assert "hbm_traffic_per_launch.md" in note
    v, why = b.traffic("updateBF_chain_n1e6_m15",
                       b.chain_kernel(("lane", 256), 15))
    assert v is None and "this run used" in why
    v, why = b.traffic("no_such_workload", k64)
    assert v is None and "no summary row" in why


def test_chain_kernel_names_follow_the_launch_shapes():
    b = _bench()
    assert b.chain_kernel(("lane+lds-z", 64), 15) == \
        "nngp::k_bf_lane_lds<15, 0, 2, false> (64 threads)"
    # the mixed register copy stages rows (zs = 4) while the block's stage
    # fits 64 KiB: 64 and 128 threads, not 256
    assert b.chain_kernel(("lane+reg-copy", 128), 15, True, 4) == \
        "nngp::k_bf_mixed_reg<15, 0, 2> (128 threads)"
    assert b.chain_kernel(("lane+reg-copy", 256), 15, True, 4) == \
        "nngp::k_bf_mixed_reg<15, 0, 0> (256 threads)"
    assert b.chain_kernel(("lane+lds-copy", 64), 15, True) == \
        "nngp::k_bf_mixed<15, 0, true> (64 threads)"


def test_loop_rate_needs_the_loop_clock_of_every_step():
    b = _bench()
    it_s, dt = b.loop_rate({"loop_ms": 500.0, "loop_iters": 100}, 100)
    assert it_s == 200.0 and dt == 0.5
    import pytest
    with pytest.raises(RuntimeError):
        b.loop_rate({"loop_ms": 500.0, "loop_iters": 99}, 100)
    with pytest.raises(RuntimeError):
        b.loop_rate({"loop_ms": 0.0, "loop_iters": 100}, 100)
