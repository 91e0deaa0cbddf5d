"""The C ABI library loads, exports every entry point include/femmd.h
declares, and its host-side validation behaves (no GPU work is issued by
any call before fm_commit)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from femmd_old_amd import _lib, System, FemmdError

ROOT = Path(__file__).resolve().parent.parent


def declared_entry_points():
    text = (ROOT / "include" / "femmd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_entry_points()
    assert len(names) >= 35
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fm_[a-z_]+)\b", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)
    # the ctypes table covers the whole header
    assert set(_lib.PROTOTYPES) == set(names)


def test_library_is_gfx950_only():
    # the offload bundle ids embedded in our own build name only gfx950
    blob = _lib.LIB_PATH.read_bytes()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa-+(gfx[0-9a-z]+)", blob))
    assert targets == {b"gfx950"}, targets


def test_create_rejects_bad_box():
    lib = _lib.load()
    h = C.c_void_p()
    p = _lib.FmParams(Lx=-1.0, Ly=10.0, dt=0.005, device=0)
    assert lib.fm_create(C.byref(p), C.byref(h)) == _lib.FM_ERR_ARG
    p = _lib.FmParams(Lx=10.0, Ly=10.0, dt=0.0, device=0)
    assert lib.fm_create(C.byref(p), C.byref(h)) == _lib.FM_ERR_ARG


def test_host_side_construction_and_validation():
    s = System(20.0, 20.0)
    first = s.add_particles([1.0, 2.0], [3.0, 4.0], m=1.0, sig=1.0)
    assert first == 0 and s.N == 2
    v, t = np.array([0, 0, 1, 0, 0, 1], np.float32), np.array([1, 2, 3], np.int32)
    node0 = s.add_elastomer(v, t, 10.0, 10.0, mass=3.0)
    assert node0 == 2 and s.N == 5
    x, y, vx, vy = s.get_state()  # before commit: host copy, caller order
    np.testing.assert_allclose(x, [1, 2, 10, 11, 10])
    with pytest.raises(FemmdError) as e:
        s.add_elastomer(v, np.array([1, 2, 4], np.int32), 0, 0, mass=1.0)
    assert e.value.code == _lib.FM_ERR_ARG
    with pytest.raises(FemmdError) as e:
        s.add_bonds([0], [7], [1.0])
    assert e.value.code == _lib.FM_ERR_ARG
    with pytest.raises(ValueError):
        s.set_pair_potential([1.0, 1.0], [2.5, 2.5], [0.0, 0.0])  # not square
    with pytest.raises(FemmdError) as e:
        s.set_pair_potential([1.0], [-2.5], [0.0])  # negative cutoff
    assert e.value.code == _lib.FM_ERR_ARG
    with pytest.raises(ValueError):
        s.add_particles([1.0], [1.0, 2.0])
    with pytest.raises(FemmdError) as e:
        s.run(10)  # before commit
    assert e.value.code == _lib.FM_ERR_STATE
    with pytest.raises(FemmdError) as e:
        s.add_particles([np.nan], [1.0])
    assert e.value.code == _lib.FM_ERR_ARG
    with pytest.raises(FemmdError) as e:
        s.set_youngs_modulus(10.0, 0.5)  # nu = 0.5 is incompressible
    assert e.value.code == _lib.FM_ERR_ARG
    s.close()


def test_set_mass_and_tuning_validation():
    s = System(20.0, 20.0)
    s.add_particles([1.0, 2.0], [3.0, 4.0], vx=[1.0, 1.0], vy=[0.5, 0.5])
    s.set_mass([1], np.inf)  # pinned: its velocity is cleared
    _, _, vx, vy = s.get_state()
    assert vx[1] == 0 and vy[1] == 0 and vx[0] == 1.0
    with pytest.raises(FemmdError):
        s.set_mass([5], 1.0)
    with pytest.raises(FemmdError):
        s.set_tuning(tile_cells=100)
    with pytest.raises(FemmdError):
        s.set_tuning(block=96)
    s.set_tuning(tile_cells=16, block=128, use_graph=0)
    with pytest.raises(FemmdError):
        s.set_reorder_interval(-1)
    s.set_reorder_interval(0)
    assert s.reorder_info() == {"interval": 0, "count": 0}
    s.close()


def test_ingestion_is_all_or_nothing():
    """A rejected fm_add_particles / fm_add_elastomer / fm_add_bonds call
    leaves the system as it was: every input is validated before anything is
    appended, an exclusion group is taken or an id is written."""
    lib = _lib.load()
    s = System(30.0, 30.0)
    s.add_particles([1.0, 2.0, 3.0], [1.0, 1.0, 1.0])
    first = C.c_int32(-7)
    bad = np.array([4.0, 5.0, np.nan], np.float32)
    ones = np.ones(3, np.float32)
    rc = lib.fm_add_particles(s.handle, 3, bad.ctypes.data_as(_lib._f), ones.ctypes.data_as(_lib._f),
                              None, None, ones.ctypes.data_as(_lib._f),
                              ones.ctypes.data_as(_lib._f), None, C.byref(first))
    assert rc == _lib.FM_ERR_ARG and first.value == -7 and s.N == 3
    # a bad sigma, velocity or type is refused before any particle is appended
    for kw in ({"sig": [1.0, 0.0]}, {"vx": [0.0, np.inf]}, {"typ": [0, -1]}):
        with pytest.raises(FemmdError):
            s.add_particles([4.0, 5.0], [1.0, 1.0], **kw)
        assert s.N == 3
    v = np.array([0, 0, 1, 0, 0, 1, 1, 1], np.float32)
    t = np.array([1, 2, 3, 2, 4, 3], np.int32)
    # per-element nu outside (-1, 1/2), E negative or NaN: nothing appended
    for kw in ({"nu": [0.3, 0.6]}, {"E": [10.0, -1.0]}, {"E": [np.nan, 1.0]}):
        with pytest.raises(FemmdError):
            s.add_elastomer(v, t, 10.0, 10.0, mass=4.0, **kw)
        assert s.N == 3
        assert s.n_elements == 0
    with pytest.raises(FemmdError):
        s.add_elastomer(np.array([0, 0, 1, 0, 0, np.inf], np.float32), t[:3], 10.0, 10.0, mass=1.0)
    assert s.N == 3
    node0 = s.add_elastomer(v, t, 10.0, 10.0, mass=4.0, E=[20.0, 30.0], nu=[0.3, 0.45])
    assert node0 == 3 and s.N == 7
    ne = C.c_int32()
    assert lib.fm_num_elements(s.handle, C.byref(ne)) == _lib.FM_OK and ne.value == 2
    # bonds: the second bond is bad, so neither is added
    with pytest.raises(FemmdError):
        s.add_bonds([0, 1], [1, 99], [10.0, 10.0])
    assert s.n_bonds() == 0
    with pytest.raises(FemmdError):
        s.add_bonds([0, 1], [1, 2], [10.0, np.nan])
    assert s.n_bonds() == 0
    with pytest.raises(FemmdError):
        s.add_bonds([0, 1], [1, 2], [10.0, 10.0], [1.0, np.nan])
    assert s.n_bonds() == 0
    s.add_bonds([0, 1], [1, 2], [10.0, 10.0])
    np.testing.assert_array_equal(s.bond_pairs(), [[0, 1], [1, 2]])
    s.close()
