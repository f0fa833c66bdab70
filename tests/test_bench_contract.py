"""bench.py's command-line contract and its byte table, on the CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_a_message(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_every_timed_kernel_has_a_known_role():
    """The kernels the bench times carry algorithmic bytes (the roofline's
    numerator), or belong to an operation measured as a whole (the paints:
    sort plus deposit), or are bookkeeping of the comparison modes."""
    in_ops = {k for members, _, _ in bench.op_bytes(512, 4).values() for k in members}
    other = {"paint_deposit", "paint_convert", "loss"}
    for k in bench.KERNELS:
        by = bench.kernel_bytes(k, 512, 4)
        assert (by is not None and by > 0) or k in in_ops or k in other, k
    # no kernel of an operation carries the operation's bytes by itself
    for k in in_ops:
        assert bench.kernel_bytes(k, 512, 4) is None or k == "paint_absmax", k
    ops = bench.op_bytes(512, 4)
    assert ops["paint1"][2] == (3 * 4 + 4) * 512 ** 3       # positions in, mesh out
    assert ops["paint3"][2] == (6 * 4 + 3 * 4) * 512 ** 3   # positions and 3 masses in, 3 meshes out
    # the dominant kernel's bytes: positions and g in (12 + 12 B), g out (12 B), 4 fields (16 B per cell)
    assert bench.kernel_bytes("sm_bwd_gather", 512, 4) == (9 * 4 + 4 * 4) * 512 ** 3
