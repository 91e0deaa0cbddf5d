"""Opt-in elastomer edge bonds on the device: the rest length of each is the
edge's length when it was created (Bonds::createBond, femmd2D/src/
bonds.cpp:17-20) and its stiffness the given k, shown through the bond
energy and forces of a stretched mesh against the FP64 oracle
(femmd2D/src/interactions.cpp:143-173)."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from femmd_old_amd import System
from femmd_old_amd.io import load_mesh

pytestmark = pytest.mark.gpu
MESH = Path(__file__).resolve().parent / "golden" / "meshes"


@pytest.mark.parametrize("mode", ["all", "dedup"])
def test_edge_bonds_rest_length_and_stiffness(mode):
    v, t = load_mesh(str(MESH / "diskVertices.txt"), str(MESH / "diskTriangles.txt"))
    v = v * np.float32(8.0)  # edges of about one sigma (the unit disk mesh is 1 across)
    L, k = 40.0, 12.5
    s = System(L, L, dt=0.002)
    s.set_lj(False)
    s.set_youngs_modulus(0.0, 0.3)  # only the bonds act
    first = s.add_elastomer(v.ravel(), t.ravel(), 20.0, 20.0, node_mass=1.0, edge_bonds=mode,
                            bond_k=k)
    assert first == 0
    s.allocate_cuda()
    x0, y0, _, _ = s.get_state()
    e0 = s.energies()
    assert abs(e0["u_bond"]) < 1e-9  # every bond at its rest length
    pairs = s.bond_pairs()
    # the rest lengths the oracle is given: the creation-time lengths
    r0 = np.hypot(x0[pairs[:, 0]].astype(np.float64) - x0[pairs[:, 1]],
                  y0[pairs[:, 0]].astype(np.float64) - y0[pairs[:, 1]]).astype(np.float32)
    # stretch the mesh by 6 % about its centre and compare
    xs = (20.0 + 1.06 * (x0 - 20.0)).astype(np.float32)
    ys = (20.0 + 1.06 * (y0 - 20.0)).astype(np.float32)
    s.set_state(xs, ys, None, None)
    s.calc_forces()
    fx, fy = s.get_forces()
    e = s.energies()
    Lf = O.box_f32(L)
    osys = O.OracleSystem(Lf, Lf, xs, ys, lj_on=False,
                          bonds=(pairs[:, 0], pairs[:, 1], np.full(len(pairs), k), r0))
    rx, ry, en, fs = O.forces(osys, pred_f32=True, scale=True)
    err = O.force_error(fx, fy, rx, ry, fs)
    print(f"  {mode}: {len(pairs)} bonds, U_bond {e['u_bond']:.8e} / {en[O.OR_U_BOND]:.8e}, "
          f"max rel force err {err.max():.3e}")
    assert e["u_bond"] == pytest.approx(en[O.OR_U_BOND], rel=1e-5)
    assert err.max() <= 1e-5
    # the duplicated bonds of "all" double the shared edges' share
    n_int = len({tuple(sorted(p)) for p in pairs.tolist()})
    assert len(pairs) == (3 * t.shape[0] if mode == "all" else n_int)
    assert s.device_errors() == 0
