"""The benchmark's host-side contract, checked without a GPU: the kernel it
names for the roofline (the one the solves launched), the algorithmic bytes
it divides by, and the refusal of more than one GPU."""
import json
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parent.parent


def test_roofline_kernel_names():
    sell = {"lanes": 0, "nt": 1, "slices": 1, "order": 0}
    assert bench.spmv_kernel_name(sell, "float") == "k_spmv_sell<float, true, true, false>"
    assert bench.spmv_kernel_name(dict(sell, slices=2), "float") == "k_spmv_sell2<float, true, true, false>"
    assert bench.spmv_kernel_name(dict(sell, slices=2, order=1), "float") == "k_spmv_sell2<float, true, true, true>"
    assert (bench.spmv_kernel_name(dict(sell, nt=0), "double", jacobi=False)
            == "k_spmv_sell<double, false, false, false>")
    assert bench.spmv_kernel_name({"lanes": 4, "nt": 0, "slices": 0}, "double") == "k_spmv_pc<double, 4, false>"


def test_measured_traffic_is_keyed_by_launched_kernel():
    # every entry names a kernel the bench can report, for the channel's size
    for (name, cells), traffic in bench.MEASURED_TRAFFIC.items():
        assert name.startswith(("k_spmv_sell<", "k_spmv_sell2<", "k_spmv_pc<"))
        assert cells == 5_995_990 and traffic > 0


def test_spmv_algorithmic_bytes():
    # channel 200x100x200: 39,146,960 entries on 5,995,990 rows, FP32 + Jacobi
    assert bench.spmv_bytes(39_146_960, 5_995_990, 4, 4) == 409_111_524
    # without the Jacobi diagonal one vector less
    assert bench.spmv_bytes(10, 3, 8, 8, jacobi=False) == 12 * 10 + 4 * 4 + 2 * 8 * 3


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
