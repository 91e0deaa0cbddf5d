"""fm_set_temperature_noalloc (the entry point that allocates no device
memory) against fm_set_temperature through the C ABI: two calls in a row
(all particles, then a smaller group of ids), each giving the same velocity
bits as the allocating entry point on an identical system; sizes beyond the
scratch limit are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from femmd_old_amd import _lib, configs

pytestmark = pytest.mark.gpu


def _pair(w, steps):
    a = configs.build_system(w)
    b = configs.build_system(w)
    a.run(steps)
    b.run(steps)
    return a, b


def _rescale(lib, h, temp, ids, noalloc):
    fn = lib.fm_set_temperature_noalloc if noalloc else lib.fm_set_temperature
    if ids is None:
        return fn(h, float(temp), 0, None)
    a = np.ascontiguousarray(ids, np.int32)
    return fn(h, float(temp), len(a), a.ctypes.data_as(_lib._i))


@pytest.mark.parametrize("cfg", ["c1", "c2s"])
def test_noalloc_rescale_matches_allocating_entry_point(cfg):
    lib = _lib.load()
    w = {"c1": lambda: configs.c1(), "c2s": lambda: configs.c2(scale=0.01)}[cfg]()
    a, b = _pair(w, 12)
    rng = np.random.default_rng(7)
    calls = [(1.7, None),  # every particle
             (0.6, np.sort(rng.choice(w.n_total, 100, replace=False)))]  # then 100 ids
    for temp, ids in calls:
        assert _rescale(lib, a.handle, temp, ids, False) == _lib.FM_OK, lib.fm_last_error(a.handle)
        assert _rescale(lib, b.handle, temp, ids, True) == _lib.FM_OK, lib.fm_last_error(b.handle)
        for u, v in zip(a.get_state(), b.get_state()):
            np.testing.assert_array_equal(u, v)
        ea, eb = a.energies(), b.energies()
        assert ea["ke"] == eb["ke"] and ea["temperature"] == eb["temperature"]
        if ids is None:
            assert eb["temperature"] == pytest.approx(temp, rel=1e-5)
        # the system goes on stepping the same
        a.run(5)
        b.run(5)
        for u, v in zip(a.get_state(), b.get_state()):
            np.testing.assert_array_equal(u, v)
    assert a.device_errors() == 0 and b.device_errors() == 0


def test_noalloc_rejects_requests_beyond_its_scratch():
    lib = _lib.load()
    s = configs.build_system(configs.c1())
    x0 = s.get_state()
    big = np.zeros((1 << 20) + 1, np.int32)
    rc = lib.fm_set_temperature_noalloc(s.handle, 1.0, len(big), big.ctypes.data_as(_lib._i))
    assert rc == _lib.FM_ERR_ARG
    assert b"at most 1048576 ids" in lib.fm_last_error(s.handle)
    bad = np.array([0, s.N], np.int32)  # an id out of range
    rc = lib.fm_set_temperature_noalloc(s.handle, 1.0, 2, bad.ctypes.data_as(_lib._i))
    assert rc == _lib.FM_ERR_ARG
    for u, v in zip(x0, s.get_state()):
        np.testing.assert_array_equal(u, v)
