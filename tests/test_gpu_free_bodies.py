"""Free bodies on the GPU vs the CPU oracle, same seeded inputs.

The free-body step (World::add_body models: step_bodies / step_bodies_euler,
newt/src/world.rs:1705-1817, and the free-body PGS solve_free_bodies,
newt/src/solver.rs:703-1257) runs in FP64 with the reference's operation
order in every precision mode, so state and qacc must equal the FP64 oracle
bit for bit.  The oracle's float instantiation reproduces newt's own golden
files for these paths byte for byte (test_oracle_golden.py: tumbling_3_body,
integrator_*, solver_stack, solver_incline).

FP64_EXACT runs the lane-per-world kernel, FP64 and MIXED the
wavefront-per-world kernel.  The Newton solve runs one wavefront per world in
every mode; under FP64_EXACT it keeps newt's order throughout (bitwise),
under FP64 and MIXED its back substitution runs column by column, held to
the drop-in's 1e-9.

Scenes: three tumbling boxes with no contacts (newt/tests/golden.rs) with
speed caps and a gravity scale; newt's golden stack of three boxes under
RK4 and Euler; 64 boxes dropped on a plane (box-box and box-plane
contacts, hundreds of rows per world).
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def tumbling(caps=False):
    """newt/tests/golden.rs:18-57; with caps, speed limits and a gravity scale."""
    q_x = scenes.quat_axis_angle((1.0, 0.0, 0.0), 0.3)
    q_yz = scenes.quat_axis_angle((0.0, 1.0, 1.0), 0.7)
    specs = [(1.0, (0.4, 0.3, 0.2), (0.0, 0.0, 5.0), scenes.IDENTITY_Q, (1.0, 0.5, 3.0), (1.0, 2.0, 0.5)),
             (2.0, (0.3, 0.3, 0.3), (1.5, -1.0, 4.0), q_x, (-0.5, 1.5, 2.0), (0.2, 3.5, 0.1)),
             (0.5, (0.5, 0.2, 0.1), (-1.0, 2.0, 6.0), q_yz, (2.0, -1.0, 0.5), (0.05, 0.1, 4.0))]
    bodies = []
    for i, (m, h, p, q, v, w) in enumerate(specs):
        extra = {}
        if caps:
            extra = dict(max_linear_velocity=[2.5, None, 0.0][i], max_angular_velocity=[None, 3.0, 4.5][i],
                         gravity_scale=[0.5, 1.0, 2.0][i])
        bodies.append(scenes.body(m, scenes.solid_box_inertia(m, h), p, q, v, w, **extra))
    q0 = np.array([x for b in bodies for x in (*b["position"], *b["orientation"])])
    v0 = np.array([x for b in bodies for x in (*b["linear_velocity"], *b["angular_velocity_body"])])

    def sampler(n, seed):
        r = np.random.default_rng(seed)
        return np.tile(q0, (n, 1)), np.tile(v0, (n, 1)) + r.normal(0.0, 0.3, size=(n, v0.size))

    opts = dict(dt=0.005, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_RK4, solver=abi.SOLVER_PENALTY,
                iterations=20)
    return scenes.Scene("tumbling3" + ("_caps" if caps else ""), opts, [], [], 0, sampler, "", bodies=bodies)


def run_pair(oracle, sc, n, steps, precision, seed=1, chunks=1):
    q0, v0 = sc.initial_state(n, seed)
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    for _ in range(chunks):
        step(m, d, steps // chunks)
        ow.step(steps // chunks, threads=8)
    out = d.state() + (d.qacc(), d.faults())
    ref = ow.state() + (ow.qacc(),)
    d.close()
    m.close()
    ow.close()
    return out, ref


def check_exact(out, ref, what):
    (q, v, a, f), (qr, vr, ar) = out, ref
    assert not f.any(), f"{what}: faults {np.unique(f)}"
    for name, x, y in (("qpos", q, qr), ("qvel", v, vr), ("qacc", a, ar)):
        if not np.array_equal(x, y):
            bad = np.argwhere(x != y)
            w, k = bad[0]
            pytest.fail(f"{what}: {name} differs in {len(bad)} entries, first world {w} index {k}: "
                        f"{x[w, k]!r} vs {y[w, k]!r} (max abs {np.abs(x - y).max():.3e})")


def check_close(out, ref, what, tol=1e-9):
    (q, v, a, f), (qr, vr, ar) = out, ref
    assert not f.any(), f"{what}: faults {np.unique(f)}"
    for name, x, y in (("qpos", q, qr), ("qvel", v, vr), ("qacc", a, ar)):
        err = (np.abs(x - y).max(axis=1) / np.maximum(np.abs(y).max(axis=1), 1.0)).max()
        assert err <= tol, f"{what}: {name} relative error {err:.3e}"


@pytest.mark.parametrize("caps", [False, True])
@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_RK4, abi.INTEGRATOR_EULER])
def test_tumbling_bodies(oracle, caps, integrator):
    sc = scenes.with_options(tumbling(caps), f"_i{integrator}", integrator=integrator)
    say(f"{sc.name}: 128 worlds x 400 steps")
    out, ref = run_pair(oracle, sc, 128, 400, abi.PRECISION_FP64_EXACT)
    check_exact(out, ref, sc.name)


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_RK4, abi.INTEGRATOR_EULER])
@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_golden_stack(oracle, integrator, precision):
    sc = scenes.box_stack(integrator=integrator)
    say(f"{sc.name} integrator {integrator} precision {precision}: 64 worlds x 600 steps")
    out, ref = run_pair(oracle, sc, 64, 600, precision, seed=3, chunks=3)
    check_exact(out, ref, sc.name)
    # the stack stands: the top box stays on the two below (start height 1.25 m)
    assert np.all(out[0][:, 7 * 2 + 2] > 1.1)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_box_pile(oracle, precision):
    """FP64_EXACT runs the lane-per-world step, FP64 the wavefront-per-world
    step (rows in LDS, or in global memory for worlds with more active
    contacts than fit): both bit for bit."""
    sc = scenes.box_pile()
    say(f"{sc.name} precision {precision}: 32 worlds x 400 steps (drop, impacts, piling)")
    out, ref = run_pair(oracle, sc, 32, 400, precision, seed=5, chunks=4)
    check_exact(out, ref, sc.name)


def test_box_pile_contacts_match(oracle):
    """Contacts of the last step (geoms, slot, point, normal, depth), world by
    world, bit for bit."""
    sc = scenes.box_pile()
    n = 16
    q0, v0 = sc.initial_state(n, 9)
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    step(m, d, 250)
    ow.step(250, threads=8)
    nc, ids_g, geo_g = d.contacts()
    for w in range(n):
        ids, geo = ow.contacts(w)
        assert nc[w] == len(ids), (w, nc[w], len(ids))
        assert np.array_equal(ids_g[w, :nc[w]], ids), w
        assert np.array_equal(geo_g[w, :nc[w]], geo), w
    say(f"box pile contacts per world after 250 steps: {nc.min()}..{nc.max()}")
    assert nc.max() > 20
    d.close()
    m.close()
    ow.close()


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_newton_golden_stack(oracle, precision):
    """The Newton solver (solve_free_bodies_newton, solver.rs:1295-1351;
    NewtonSystem::solve and the exact line search, newton.rs:275-556) on the
    golden stack: dense response matrix, Cholesky, Newton directions, line
    search, bit for bit."""
    sc = scenes.with_options(scenes.box_stack(), "_newton", solver=abi.SOLVER_NEWTON)
    # The drop-in's contract: within 1e-9 over 100 steps (the reordered back
    # substitution's rounding then compounds, about 1e-9 by 400 steps).
    steps = 400 if precision == abi.PRECISION_FP64_EXACT else 100
    say(f"{sc.name} precision {precision}: 32 worlds x {steps} steps")
    out, ref = run_pair(oracle, sc, 32, steps, precision, seed=7, chunks=2)
    if precision == abi.PRECISION_FP64_EXACT:
        check_exact(out, ref, sc.name)
    else:
        check_close(out, ref, sc.name)


@pytest.mark.parametrize("precision,steps", [(abi.PRECISION_FP64_EXACT, 300), (abi.PRECISION_FP64, 100)])
def test_newton_box_pile(oracle, precision, steps):
    """64 boxes under the Newton solver: a few hundred rows per world.
    FP64_EXACT bit for bit over 300 steps; FP64 within 1e-9 over 100 steps
    (16 worlds)."""
    sc = scenes.with_options(scenes.box_pile(), "_newton", solver=abi.SOLVER_NEWTON)
    say(f"{sc.name} precision {precision}: 16 worlds x {steps} steps")
    out, ref = run_pair(oracle, sc, 16, steps, precision, seed=11, chunks=3)
    if precision == abi.PRECISION_FP64_EXACT:
        check_exact(out, ref, sc.name)
    else:
        check_close(out, ref, sc.name)


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_RK4, abi.INTEGRATOR_EULER])
@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_penalty_stack(oracle, integrator, precision):
    """Penalty contacts (compute_penalty_wrenches / apply_contact_wrench,
    newt/src/world.rs:2281-2301, 2868-2949; the legacy vertex-face box-box
    manifold, contact.rs:534-714) on newt's penalty golden scene: contacts
    and wrenches re-sampled at every RK4 stage, bit for bit in every
    precision mode (the lane-per-world step runs them all)."""
    sc = scenes.penalty_stack(integrator=integrator)
    say(f"{sc.name} integrator {integrator} precision {precision}: 64 worlds x 600 steps")
    out, ref = run_pair(oracle, sc, 64, 600, precision, seed=13, chunks=3)
    check_exact(out, ref, sc.name)


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_RK4, abi.INTEGRATOR_EULER])
def test_penalty_box_pile(oracle, integrator):
    """64 boxes under penalty contacts: box-plane and box-box (vertex-face
    with the SAT fallback), over a hundred contacts per world, bit for bit;
    the last stage's contacts match world by world."""
    sc = scenes.with_options(scenes.box_pile(), f"_pen_i{integrator}", solver=abi.SOLVER_PENALTY,
                             integrator=integrator)
    n = 16
    say(f"{sc.name}: {n} worlds x 300 steps")
    q0, v0 = sc.initial_state(n, 17)
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    for _ in range(3):
        step(m, d, 100)
        ow.step(100, threads=8)
    check_exact(d.state() + (d.qacc(), d.faults()), ow.state() + (ow.qacc(),), sc.name)
    nc, ids_g, geo_g = d.contacts()
    for w in range(n):
        ids, geo = ow.contacts(w)
        assert nc[w] == len(ids), (w, nc[w], len(ids))
        assert np.array_equal(ids_g[w, :nc[w]], ids), w
        assert np.array_equal(geo_g[w, :nc[w]], geo), w
    say(f"penalty pile contacts per world: {nc.min()}..{nc.max()}")
    assert nc.max() > 50
    d.close()
    m.close()
    ow.close()
