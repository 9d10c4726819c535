"""Pins of the oracle's joint-limit rows (solve_tree_limits,
newt/src/solver.rs:2337-2628).

* One active row on a single hinge has a closed form: PGS converges in one
  sweep to f = max(0, a_ref·dt·d·M), where M is the hinge's mass,
  a_ref = -b·v_row + k·violation the reference acceleration
  (solver.rs:353-384) and d the impedance.  The Euler step's qacc is then
  (c·f/dt + gravity torque)/M.  Both sides of the range are checked, and a
  case where the row is inactive because the joint moves back in.
* The reference's own anchors (newt/tests/solver_joint_limits.rs), on the
  oracle's float instantiation (newt's f32 arithmetic): a hinge released
  above its range under an outward torque stays within pi/3 + 0.2; started at
  the limit it creeps less than 0.05 rad over 10 000 steps; a soft SolRef
  override lets more through than the default.
"""
import math

import numpy as np
import pytest

from tooling_amd import abi, scenes


def _hinge(l, m, ixx, rng, integrator=abi.INTEGRATOR_EULER, dt=0.001, iterations=5, damping=0.0,
           limit_solref=None):
    links = [scenes.link(-1, abi.JOINT_FIXED, mass=1.0, inertia=(1, 0, 0, 0, 1, 0, 0, 0, 1)),
             scenes.link(0, abi.JOINT_HINGE, axis=(1.0, 0.0, 0.0), child_pos=(0.0, 0.0, l), mass=m,
                         inertia=(ixx, 0, 0, 0, ixx, 0, 0, 0, ixx), damping=damping, range=rng,
                         limit_solref=limit_solref)]
    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=integrator, solver=abi.SOLVER_PGS,
                iterations=iterations)
    return scenes.Scene("hinge_limit", opts, links, [], 0, None)


def _impedance(viol, dmin=0.9, dmax=0.95, width=0.001, mid=0.5, power=2):
    x = abs(-viol / width)
    if x <= 0.0:
        y = 0.0
    elif x >= 1.0:
        y = 1.0
    elif x < mid:
        y = mid * (x / mid) ** power
    else:
        y = 1.0 - (1.0 - mid) * ((1.0 - x) / (1.0 - mid)) ** power
    return dmin + (dmax - dmin) * y


@pytest.mark.parametrize("q,qd", [(0.95, 0.3), (0.95, -0.02), (-1.3, -0.4), (-1.3, 5.0), (0.2, 1.0)])
def test_single_limit_row_closed_form(oracle_lib, q, qd):
    l, m, ixx, dt = 0.6, 1.4, 0.03, 0.001
    lo, hi = -1.2, 0.9
    sc = _hinge(l, m, ixx, (lo, hi), dt=dt)
    ow = oracle_lib.OracleWorlds(sc, 1, 0)
    ow.set_state(np.array([[q]]), np.array([[qd]]))
    ow.step(1)
    a = ow.qacc()[0, 0]
    L = oracle_lib.lib()
    M = m * l * l + ixx
    sin_q = 2.0 * L.oracle_nsin64(q * 0.5) * L.oracle_ncos64(q * 0.5)
    tau_g = -m * 9.81 * l * sin_q
    if q > hi:
        c, viol = -1.0, q - hi
    elif q < lo:
        c, viol = 1.0, lo - q
    else:
        c, viol = 0.0, 0.0
    f = 0.0
    if c != 0.0:
        dmax, tc, dr = 0.95, max(0.02, 2 * dt), 1.0
        d = _impedance(viol)
        b = 2.0 / (dmax * tc)
        k = d / (dmax * dmax * tc * tc * dr * dr)
        a_ref = -b * (c * qd) - k * (-viol)
        f = max(0.0, a_ref * dt * d * M)   # -bias / diag with bias = -a_ref dt, diag = A/d, A = 1/M
    expected = (c * f / dt + tau_g) / M
    assert a == pytest.approx(expected, rel=1e-12, abs=1e-12)
    if q == 0.95 and qd == -0.02:
        assert f > 0.0          # penetrating but moving in slowly: the spring term wins
    if q == -1.3 and qd == 5.0:
        assert f == 0.0         # moving back in fast: row inactive after the projection


def _anchor(oracle_lib, q0, torque, solref=None):
    """newt/tests/solver_joint_limits.rs:43-72: World::new (dt 5 ms, RK4),
    PGS 30; hinge about X, COM 1 m from the joint, range ±pi/4, damping 0.5."""
    sc = _hinge(1.0, 1.0, 0.01, (-math.pi / 4, math.pi / 4), integrator=abi.INTEGRATOR_RK4, dt=0.005,
                iterations=30, damping=0.5, limit_solref=solref)
    ow = oracle_lib.OracleWorlds(sc, 1, 1)   # float: newt's arithmetic
    ow.set_state(np.array([[q0]]), np.zeros((1, 1)))
    ow.set_qfrc(np.array([[torque]]))
    return ow


def test_reference_anchor_released_above_range(oracle_lib):
    ow = _anchor(oracle_lib, math.pi / 3, 12.0)
    peak = math.pi / 3
    for _ in range(60):
        ow.step(100)
        peak = max(peak, ow.state()[0][0, 0])
    assert peak < math.pi / 3 + 0.2
    # and it settles at the upper limit
    assert abs(ow.state()[0][0, 0] - math.pi / 4) < 0.05


def test_reference_anchor_no_creep(oracle_lib):
    ow = _anchor(oracle_lib, math.pi / 4, 12.0)
    worst = 0.0
    for _ in range(100):
        ow.step(100)
        worst = max(worst, ow.state()[0][0, 0] - math.pi / 4)
    assert worst < 0.05


def test_reference_anchor_solref_override(oracle_lib):
    stiff = _anchor(oracle_lib, math.pi / 4, 12.0)
    soft = _anchor(oracle_lib, math.pi / 4, 12.0, solref=(0.2, 1.0))
    stiff.step(3000)
    soft.step(3000)
    ps = stiff.state()[0][0, 0] - math.pi / 4
    pf = soft.state()[0][0, 0] - math.pi / 4
    assert ps < 0.05 and pf < 0.5 and pf > ps + 0.001


def test_limits_without_pgs_are_rejected():
    """Penalty limit torques are not built: such a model fails at finalize
    instead of running with the limit silently missing."""
    from tooling_amd.newt import Model
    sc = _hinge(1.0, 1.0, 0.01, (-1.0, 1.0))
    sc = scenes.with_options(sc, "_penalty", solver=abi.SOLVER_PENALTY)
    try:
        abi.load_library()
    except RuntimeError:
        pytest.skip("HIP library not built")
    with pytest.raises(abi.NewtError):
        Model(sc)
