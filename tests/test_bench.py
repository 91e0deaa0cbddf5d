"""bench.py's host-side contract: the one-GPU refusal and the algorithmic
byte model (no GPU needed)."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["n_gpus"] == 2 and "one GPU" in d["error"]


def test_algorithmic_bytes_model():
    """Structural properties of the layout byte model: moving a particle's
    integration from the node kick into the force pass saves bytes (its
    forces never leave the pass), the inverse masses cost 4 bytes per
    integrated particle, and only list mode has a list build."""
    n, ncells, ntri, nb = 1000, 100, 50, 10
    for L in (None, 20.0):
        allnode = bench.algorithmic_bytes(n, ncells, ntri, nb, L, n_corner=n)
        fused = bench.algorithmic_bytes(n, ncells, ntri, nb, L, n_corner=100)
        tot = lambda d: d["kick"] + d["forces"]
        assert tot(fused) < tot(allnode)
        # per particle moved into the pass: the node kick's read of fx,fy
        # and the pass's write of them are gone
        assert (tot(allnode) - tot(fused)) % 900 == 0 and tot(allnode) - tot(fused) >= 16 * 900
        uni = bench.algorithmic_bytes(n, ncells, ntri, nb, L, uniform_mass=True, n_corner=100)
        assert tot(fused) - tot(uni) == 4 * n
        assert {k: uni[k] for k in ("bin", "sort", "list_build")} == \
            {k: fused[k] for k in ("bin", "sort", "list_build")}
        assert (fused["list_build"] > 0) == (L is not None)
    assert set(bench.REBUILD_KERNELS) == {"bin", "sort", "list_build"}
    # the survey's compulsory figures: the fused pass adds 16 per kicked particle
    cb = bench.compulsory_bytes(n, ncells, ntri, nb, 300, 20, n_corner=300)
    assert cb["forces_fused"] - cb["forces"] == 16 * (n - 300)
    assert cb["forces"] == 24 * n + 4 * ncells + 84 * ntri + 40 * nb


def test_committed_counter_summaries_are_read():
    """roofline.traffic and roofline.valu_us come from the committed
    rocprofv3 summaries: both k_forces lines parse, and the measured HBM
    bytes sit within 1.5 x of the algorithmic model at the c2 size."""
    t = bench.measured_traffic("k_forces", "list")
    v = bench.measured_valu_us("k_forces", "list")
    assert t is not None and v is not None
    ab = bench.algorithmic_bytes(1141760, 380 * 380, 153600, 0, 21.66, uniform_mass=True,
                                 n_corner=93184)
    assert 0.9 <= t / ab["forces"] <= 1.5
    assert 5.0 < v < 100.0
    assert bench.measured_valu_us("k_none", "list") is None
