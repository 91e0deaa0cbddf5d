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
    n, ncells, ntri, nb = 1000, 100, 50, 10
    cells = bench.algorithmic_bytes(n, ncells, ntri, nb)
    lists = bench.algorithmic_bytes(n, ncells, ntri, nb, 20.0)
    corners = 36 * ntri + 24 * nb
    elems = 100 * ntri + 48 * nb
    # cell walk: the binning is part of the kick, no list build
    assert cells["kick"] == 52 * n + 8 * ncells + corners
    assert cells["bin"] == 0 and cells["list_build"] == 0
    assert cells["sort"] == 36 * n + 12 * ncells
    assert cells["forces"] == 28 * n + 4 * ncells + elems
    # lists: 2 bytes per slot, L + 1.5 slots per particle
    assert lists["kick"] == 60 * n + corners
    assert lists["list_build"] == int((20 + 2 * 21.5) * n) + 4 * ncells
    assert lists["forces"] == int((24 + 2 * 21.5) * n) + 4 * ncells + elems
    assert set(bench.REBUILD_KERNELS) == {"bin", "sort", "list_build"}
    # one mass for every particle: the kick reads no inverse masses
    uni = bench.algorithmic_bytes(n, ncells, ntri, nb, 20.0, uniform_mass=True)
    assert uni["kick"] == lists["kick"] - 4 * n
    assert {k: v for k, v in uni.items() if k != "kick"} == \
        {k: v for k, v in lists.items() if k != "kick"}


def test_committed_counter_summaries_are_read():
    """roofline.traffic and roofline.valu_us come from the committed
    rocprofv3 summaries: both k_forces lines parse, and the measured HBM
    bytes sit within 1.5 x of the algorithmic model at the c2 size."""
    t = bench.measured_traffic("k_forces", "list")
    v = bench.measured_valu_us("k_forces", "list")
    assert t is not None and v is not None
    ab = bench.algorithmic_bytes(1141760, 380 * 380, 153600, 0, 21.66)
    assert 0.9 <= t / ab["forces"] <= 1.5
    assert 5.0 < v < 100.0
    assert bench.measured_valu_us("k_none", "list") is None
