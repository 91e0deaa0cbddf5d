"""A C99 program compiled against include/femmd.h and linked with
libfemmd.so (tests/c_caller.c): the header is valid C, every call it makes
links, and on the GPU the C program's results are the same bits as the same
system built and stepped through the Python binding."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from femmd_old_amd import System, _lib

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "c_caller.c"


def build_c_caller(out_dir: Path, cxx: bool = False) -> Path:
    exe = out_dir / ("c_caller_cxx" if cxx else "c_caller")
    lib_dir = _lib.LIB_PATH.parent
    lang = ["g++", "-x", "c++", "-std=c++17"] if cxx else ["gcc", "-std=c99", "-pedantic"]
    cmd = [*lang, "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(SRC),
           f"-L{lib_dir}", "-lfemmd", f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath-link,/opt/rocm/lib",
           "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_compiles_links_and_fails_cleanly_without_a_device(tmp_path):
    exe_c = build_c_caller(tmp_path)
    build_c_caller(tmp_path, cxx=True)  # the header under a C++ compiler too
    if _lib.device_count() > 0:
        return
    # no device: fm_commit reports the HIP failure as a status, the process
    # is not killed (exit code 1 from the caller, message on stderr)
    r = subprocess.run([str(exe_c), "5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "fm_commit failed" in r.stderr


def parse(out: str):
    d = {"s": [], "p": []}
    for line in out.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] in ("s", "p"):
            d[t[0]].append([float(v) for v in t[1 + (t[0] == "p"):]])
        elif t[0] in ("n", "tris", "bonds"):
            d[t[0]] = [int(v) for v in t[1:]]
        else:
            d[t[0]] = [float(v) for v in t[1:]]
    return d


@pytest.mark.gpu
def test_c_caller_matches_python_binding(tmp_path):
    exe = build_c_caller(tmp_path)
    nsteps = 50
    r = subprocess.run([str(exe), str(nsteps)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    d = parse(r.stdout)
    ntot, first, node0, nb, nt = d["n"]
    assert (first, nb, nt) == (0, 2, 6) and ntot == node0 + 7
    sv = np.array(d["s"], np.float32)
    s = System(30.0, 30.0, dt=0.005)
    s.add_particles(sv[:, 0], sv[:, 1], sv[:, 2], sv[:, 3], 1.0, 1.0)
    s.set_youngs_modulus(50.0, 0.3)
    s.add_elastomer(np.array(d["verts"], np.float32), np.array(d["tris"], np.int32), 15.3, 15.2,
                    0, 0, 1.0, 0.0, True, node_mass=1.0)
    b = d["bonds"]
    s.add_bonds([b[0], b[2]], [b[1], b[3]], [100.0, 100.0], [-1.0, -1.0])
    s.set_pair_potential([1.0], [2.5], [0.0])
    s.allocate_cuda()
    s.run(nsteps)
    x, y, vx, vy = s.get_state()
    pc = np.array(d["p"], np.float32)
    for k, a in enumerate((x, y, vx, vy)):
        np.testing.assert_array_equal(a, pc[:, k])
    e = s.energies()
    keys = ("ke", "u_lj", "w_lj", "u_fea", "w_fea", "u_bond", "w_bond", "pressure", "temperature")
    for k, v in zip(keys, d["energy"]):
        assert e[k] == v, (k, e[k], v)
    assert int(d["energy"][-1]) == nsteps
    avg = s.averages()
    assert int(d["avg"][0]) == avg["count"] == nsteps
    for k, name in enumerate(_lib.AVG_NAMES):
        assert d["avg"][1 + 2 * k] == avg[name]["avg"], name
    assert e["u_fea"] > 0 and e["u_bond"] > 0
