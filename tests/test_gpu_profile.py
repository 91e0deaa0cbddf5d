"""The profile pass (fm_profile_step): per-kernel dispatch durations, taken
from the start and end timestamps of each dispatch.

  * list mode times all five kernels; cell-walk mode launches no separate
    binning or list build, and reports 0 for them;
  * every kernel's time is positive and no longer than the whole step body;
  * the steps it advances are ordinary steps: the state after it equals the
    state after the same number of steps taken by run() (bitwise).
"""
import numpy as np
import pytest

from femmd_old_amd import configs

pytestmark = pytest.mark.gpu

KEYS = ("kick", "bin", "sort", "list_build", "forces", "step")


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_profile_pass_times_each_kernel(skin):
    w = configs.c1(seed=7)
    s = configs.build_system(w, skin=skin)
    s.run(20)
    p = s.profile(8)
    print(f"  skin {skin}: " + ", ".join(f"{k} {p[k] * 1e3:.2f} us" for k in KEYS))
    assert tuple(p) == KEYS
    for k in ("kick", "sort", "forces", "step"):
        assert p[k] > 0.0
    if skin > 0:
        assert p["bin"] > 0.0 and p["list_build"] > 0.0
    else:
        assert p["bin"] == 0.0 and p["list_build"] == 0.0
    kernels = [p[k] for k in KEYS[:-1]]
    assert max(kernels) <= p["step"]
    assert sum(kernels) <= p["step"] * 1.001
    assert s.device_errors() == 0


def test_profile_pass_steps_like_run():
    w = configs.c1(seed=8)
    a = configs.build_system(w)
    b = configs.build_system(w)
    a.run(10)
    a.profile(5)  # advances 5 + 1 steps
    b.run(16)
    for u, v in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(u, v)
