"""bench.py pieces that run without a GPU: the HBM traffic it reports comes
from the committed rocprofv3 summary, the algorithmic bytes of one heatbath
half-sweep launch, and the refusal of more than one GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_from_committed_profile():
    t = bench.measured_traffic()
    # FETCH_SIZE x 2 + WRITE_SIZE of k_hb_half<float> at 32^4: between the
    # two-row compulsory bytes (226.5 MB) and the 18-real algorithmic bytes
    assert t is not None and 226.5e6 < t < 339.8e6
    assert bench.measured_traffic("no_such_kernel") is None


def test_algorithmic_bytes_per_half_sweep():
    V = 32 ** 4
    # R+W of U_mu(p), R of U_mu(1-p) and R of the three U_nu fields: 81 reals
    # per site for V/2 links, i.e. 162 reals per link-update
    assert bench.HB_BYTES_PER_LINK == 162
    assert bench.HB_BYTES_PER_LINK * 4 * (V // 2) == 339738624
    assert bench.WF_BYTES_PER_SITE == 720   # 216 + 288 + 216 reals over the three stages


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "one GPU" in r.stderr and r.stdout == ""
