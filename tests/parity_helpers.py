"""Shared checks of the GPU parity tests (imported by them; not collected)."""
import numpy as np
import pytest

import oracle as O


def check_energies(s, en, w=None, rtol=1e-5):
    """Energies, virials, kinetic energy and pressure of the engine's current
    state against the FP64 oracle at the same positions.  Each bar is rtol
    times the scale of the cancellation in its sum: the energies' magnitudes,
    the virial terms' magnitudes (oracle OR_WS_*).  The engine forms W_fea as
    sum_k f_k . (r_k - r_2) per element, the oracle as the reference's
    sum_k f_k . r_k over unfolded corners (femmd2D/src/interactions.cpp:681):
    equal since sum_k f_k = 0.  Pressure: 0.5 sum m v^2 / V + 0.5 (W_lj +
    W_fea) / V (femmd2D/src/properties.cpp:34) from the host's FP64 kinetic
    energy of the engine's velocities and the oracle's virials."""
    e = s.energies()
    scale_u = abs(en[O.OR_U_LJ]) + abs(en[O.OR_U_FEA]) + abs(en[O.OR_U_BOND]) + 1.0
    assert abs(e["u_lj"] - en[O.OR_U_LJ]) <= rtol * scale_u, (e["u_lj"], en[O.OR_U_LJ])
    assert abs(e["w_lj"] - en[O.OR_W_LJ]) <= rtol * (abs(en[O.OR_W_LJ]) + 1.0)
    assert abs(e["u_fea"] - en[O.OR_U_FEA]) <= rtol * scale_u
    assert abs(e["u_bond"] - en[O.OR_U_BOND]) <= rtol * scale_u
    ws_fea = en[O.OR_WS_FEA] + abs(en[O.OR_W_FEA])
    assert abs(e["w_fea"] - en[O.OR_W_FEA]) <= rtol * ws_fea + 1e-9, (e["w_fea"], en[O.OR_W_FEA])
    assert abs(e["w_bond"] - en[O.OR_W_BOND]) <= rtol * en[O.OR_WS_BOND] + 1e-9, \
        (e["w_bond"], en[O.OR_W_BOND])
    if w is not None:
        _, _, vx, vy = s.get_state()
        m = w.global_arrays()["mass"].astype(np.float64)
        m = np.where(np.isfinite(m) & (m > 0), m, 0.0)
        vx = vx.astype(np.float64)
        vy = vy.astype(np.float64)
        ke2 = float(np.sum(m * (vx * vx + vy * vy)))
        assert e["ke"] == pytest.approx(0.5 * ke2, rel=rtol, abs=1e-12)
        V = float(np.float32(w.Lx)) * float(np.float32(w.Ly))
        p_ref = 0.5 * ke2 / V + 0.5 * (en[O.OR_W_LJ] + en[O.OR_W_FEA]) / V
        p_scale = (0.5 * ke2 + 0.5 * (abs(en[O.OR_W_LJ]) + 1.0 + ws_fea)) / V
        assert abs(e["pressure"] - p_ref) <= rtol * p_scale, (e["pressure"], p_ref)
        assert e["temperature"] == pytest.approx(0.5 * ke2 / w.n_total, rel=rtol)
    return e
