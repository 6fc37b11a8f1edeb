"""hipGraph replay and batched ensembles against single eager handles and the oracle.

A graph replays the same kernels with the same arguments as the eager steps,
and a batch member runs the same per-element arithmetic as a single handle, so
both are required to reproduce the eager single-handle results bit for bit."""
import numpy as np
import pytest

import smaiso_amd as S
import smaiso_oracle as O
from helpers import params_pair, rel_err

pytestmark = pytest.mark.gpu

NAMES = ("psi", "rho", "amp", "vx", "vy", "vz", "mu", "lpsi", "divv", "psiq", "psiq_old", "nlq",
         "f_rho", "psi_old", "lap_rl")


def assert_same(got: dict, want: dict, what: str):
    for k in want:
        if not np.array_equal(got[k], want[k]):
            mx, l2 = rel_err(got[k], want[k])
            raise AssertionError(f"{what}: field {k} differs (max {mx:.3e}, l2 {l2:.3e})")


def handle_fields(st):
    return {k: st.field(k) for k in NAMES}


def member_fields(en, m):
    return {k: en.field(m, k) for k in NAMES}


def ic_for(p: S.Params, seed: int, undulation: float = 0.5, rho_iso: float = 0.9):
    return S.make_interface_ic(p, S.InterfaceIC(undulation=undulation, seed=seed, rho_iso=rho_iso))


@pytest.mark.parametrize("precision", [S.FP64, S.FP32_MIXED])
def test_graph_replay_equals_eager(precision):
    """Chunks of odd and even length make replayed pairs start at both step
    parities; field readouts and diagnostics run eagerly between replays; a
    second set_state after an odd number of steps must keep the captured
    buffer rotation straight."""
    p = S.Params(nx=32, ny=24, nz=48, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=3,
                 precision=precision)
    psi, rho = ic_for(p, 7)
    eager, graph = S.Stepper(p), S.Stepper(p)
    graph.set_graph(True)
    for st in (eager, graph):
        st.set_state(psi, rho)
    for n in (1, 4, 3, 2, 5, 6, 1, 8):
        eager.step(n)
        graph.step(n)
        assert graph.nloop == eager.nloop
        assert_same(handle_fields(graph), handle_fields(eager), f"after {eager.nloop} steps")
        assert graph.diagnostics() == eager.diagnostics()
    # restart from another state (17 steps taken: odd rotation)
    psi2, rho2 = ic_for(p, 11, undulation=0.2)
    fresh = S.Stepper(p)
    fresh.set_state(psi2, rho2)
    graph.set_state(psi2, rho2)
    for n in (2, 5, 4):
        fresh.step(n)
        graph.step(n)
        assert_same(handle_fields(graph), handle_fields(fresh), f"restart, {fresh.nloop} steps")
    # async steps replay too
    fresh.step(6)
    graph.step_async(6)
    graph.synchronize()
    assert_same(handle_fields(graph), handle_fields(fresh), "async")


def test_graph_replay_matches_oracle():
    sp, op = params_pair(nx=32, ny=32, nz=32, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(undulation=0.5, seed=3))
    o = O.Oracle(op)
    o.set_state(psi, rho)
    o.step(12)
    st = S.Stepper(sp)
    st.set_graph(True)
    st.set_state(psi, rho)
    st.step(12)
    for name, ref in (("psi", o.f.psi), ("rho", o.f.rho), ("psiq", o.f.psiq), ("vx", o.f.extra["vx"])):
        mx, l2 = rel_err(st.field(name), ref)
        assert mx < 1e-10 and l2 < 1e-10, (name, mx, l2)


def test_graph_inactive_while_timing():
    p = S.Params(nx=32, ny=32, nz=32, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    psi, rho = ic_for(p, 2)
    st = S.Stepper(p)
    st.set_graph(True)
    st.timing(True)
    st.set_state(psi, rho)
    st.step(4)
    ms, n, _ = st.timing_get("r2r_z")
    assert n > 0 and ms > 0  # every launch was timed, so none came from a graph


def _mixed_members():
    base = dict(dt=0.01, rho_m=1.0, divv_switch=2)
    return [
        S.Params(nx=32, ny=32, nz=32, nu=0.5, zeta=1.5, rho0=0.8, **base),
        S.Params(nx=24, ny=20, nz=36, nu=1.0, **base),
        S.Params(nx=32, ny=32, nz=32, nu=1.7, zeta=0.6, rho0=0.9, lam=0.5, **base),
        S.Params(nx=32, ny=32, nz=32, nu=1.0, precision=S.FP32_MIXED, **base),
        S.Params(nx=24, ny=20, nz=36, nu=0.3, zeta=2.0, rho0=0.7, **base),
        S.Params(nx=32, ny=32, nz=32, nu=0.9, zeta=1.1, rho0=0.75, precision=S.FP32_MIXED, **base),
    ]


@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_members_equal_single_handles(use_graph):
    members = _mixed_members()
    ics = [ic_for(p, 20 + m, undulation=0.1 * m, rho_iso=0.7 + 0.04 * m) for m, p in enumerate(members)]
    singles = []
    for p, (psi, rho) in zip(members, ics):
        st = S.Stepper(p)
        st.set_state(psi, rho)
        singles.append(st)
    with S.Ensemble(members) as en:
        assert en.groups == 3  # 32^3 FP64, 24x20x36 FP64, 32^3 FP32-mixed
        en.set_graph(use_graph)
        for m, (psi, rho) in enumerate(ics):
            en.set_state(m, psi, rho)
        for n in (1, 4, 5):
            en.step(n)
            for st in singles:
                st.step(n)
            for m, st in enumerate(singles):
                assert en.nloop(m) == st.nloop
                assert_same(member_fields(en, m), handle_fields(st), f"member {m}, {st.nloop} steps")
                assert en.diagnostics(m) == st.diagnostics()


def test_ensemble_restart_of_one_batch():
    """set_state of one member restarts its own batch only."""
    base = dict(dt=0.01, rho_m=1.0, divv_switch=1)
    members = [S.Params(nx=32, ny=32, nz=32, nu=1.0, **base),
               S.Params(nx=32, ny=32, nz=32, nu=0.4, **base),
               S.Params(nx=24, ny=24, nz=32, nu=1.0, **base)]
    ics = [ic_for(p, 40 + m) for m, p in enumerate(members)]
    with S.Ensemble(members) as en:
        for m, (psi, rho) in enumerate(ics):
            en.set_state(m, psi, rho)
        en.step(5)
        other = S.Stepper(members[2])
        other.set_state(*ics[2])
        other.step(5)
        # restart batch 0 (members 0 and 1) from new states
        new = [ic_for(members[m], 60 + m, undulation=0.3) for m in (0, 1)]
        for m in (0, 1):
            en.set_state(m, *new[m])
        assert en.nloop(0) == 0 and en.nloop(2) == 5
        en.step(6)
        other.step(6)
        assert_same(member_fields(en, 2), handle_fields(other), "untouched batch")
        for m in (0, 1):
            st = S.Stepper(members[m])
            st.set_state(*new[m])
            st.step(6)
            assert_same(member_fields(en, m), handle_fields(st), f"restarted member {m}")


def test_ensemble_uneven_grids_fp32_mixed():
    """The batched-ensemble workload at reduced count: 64^3, 96^3 and
    128x128x64 grids, FP32-mixed, per-member viscosity, coupling and density
    drawn from a seeded generator; members equal their single-handle runs."""
    rng = np.random.default_rng(5)
    shapes = [(64, 64, 64), (96, 96, 96), (128, 128, 64)]
    members, ics = [], []
    for shape in shapes:
        for _ in range(2):
            rho_iso = float(rng.uniform(0.7, 0.95))
            p = S.Params(nx=shape[0], ny=shape[1], nz=shape[2], dt=0.01, rho_m=1.0, divv_switch=0,
                         rho0=rho_iso, nu=float(rng.uniform(0.1, 2.0)),
                         zeta=float(rng.uniform(0.5, 2.0)), precision=S.FP32_MIXED)
            members.append(p)
            ics.append(ic_for(p, int(rng.integers(1 << 30)), undulation=float(rng.uniform(0, 1)),
                              rho_iso=rho_iso))
    with S.Ensemble(members) as en:
        assert en.groups == 3
        for m, (psi, rho) in enumerate(ics):
            en.set_state(m, psi, rho)
        en.step(20)
        for m, p in enumerate(members):
            st = S.Stepper(p)
            st.set_state(*ics[m])
            st.step(20)
            for k in ("psi", "rho", "vx", "vz"):
                assert np.array_equal(en.field(m, k), st.field(k)), (m, k)
            l1, mass, en_ = en.diagnostics(m)
            assert np.isfinite(l1) and mass > 0


def test_ensemble_errors():
    p = S.Params(nx=16, ny=16, nz=16, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    with S.Ensemble([p, p]) as en:
        psi, rho = ic_for(p, 1)
        en.set_state(0, psi, rho)
        with pytest.raises(S.SmaisoError) as e:
            en.step(1)  # member 1 not set
        assert e.value.code == S.E_STATE
        with pytest.raises(S.SmaisoError) as e:
            en.set_state(2, psi, rho)
        assert e.value.code == S.E_INVALID
        with pytest.raises(S.SmaisoError) as e:
            en.field(0, "psi")
        assert e.value.code == S.E_STATE
        en.set_state(1, psi, rho)
        en.step(2)
        assert en.diagnostics(0) == en.diagnostics(1)



def test_ensemble_setting_one_member_of_a_stepped_batch_unsets_the_others():
    """Setting one member restarts its batch's step count (and with it the
    first-step history of every member), so the batch's other members must be
    set again before it steps."""
    p = S.Params(nx=16, ny=16, nz=16, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    q = S.Params(nx=16, ny=16, nz=32, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    with S.Ensemble([p, p, q]) as en:
        ics = [ic_for(m_p, 70 + m) for m, m_p in enumerate((p, p, q))]
        for m, (psi, rho) in enumerate(ics):
            en.set_state(m, psi, rho)
        en.step(3)
        en.set_state(0, *ics[0])
        with pytest.raises(S.SmaisoError) as e:
            en.step(1)
        assert e.value.code == S.E_STATE
        with pytest.raises(S.SmaisoError) as e:
            en.field(1, "psi")
        assert e.value.code == S.E_STATE
        assert en.nloop(2) == 3  # the other batch keeps its state
        en.set_state(1, *ics[1])
        en.step(2)
        st = S.Stepper(p)
        st.set_state(*ics[1])
        st.step(2)
        assert np.array_equal(en.field(1, "psi"), st.field("psi"))

@pytest.mark.parametrize("precision", [S.FP64, S.FP32_MIXED])
def test_autotune_picks_fastest_and_keeps_parity(precision):
    sp, op = params_pair(nx=64, ny=48, nz=32, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0,
                         precision=precision)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(undulation=0.5, seed=8))
    st = S.Stepper(sp)
    ms, chosen = st.autotune()
    assert ms.shape == (2, 3, S.TILE_VARIANTS) and chosen.shape == (2, 3)
    for d in range(2 if precision == S.FP32_MIXED else 1):
        for a in range(3):
            timed = ms[d, a] >= 0
            assert timed[0] and timed.sum() >= 2
            assert ms[d, a, chosen[d, a]] == ms[d, a][timed].min()
    if precision == S.FP64:
        assert np.all(ms[1] == -1)
    st.set_graph(True)
    st.set_state(psi, rho)
    st.step(10)
    o = O.Oracle(op)
    o.set_state(psi, rho)
    o.step(10)
    tol = 1e-10 if precision == S.FP64 else 1e-4
    for name, ref in (("psi", o.f.psi), ("rho", o.f.rho), ("vx", o.f.extra["vx"])):
        mx, l2 = rel_err(st.field(name), ref)
        assert mx < tol and l2 < tol, (name, mx, l2)


def test_ensemble_autotune():
    members = _mixed_members()
    ics = [ic_for(p, 30 + m) for m, p in enumerate(members)]
    with S.Ensemble(members) as en:
        for m, (psi, rho) in enumerate(ics):
            en.set_state(m, psi, rho)
        en.step(3)
        en.autotune()  # mid-run: the captured graph is dropped and recaptured
        en.step(5)
        for m, p in enumerate(members):
            st = S.Stepper(p)
            st.set_state(*ics[m])
            st.step(8)
            tol = 1e-12 if p.precision == S.FP64 else 1e-5
            for k in ("psi", "rho", "vz"):
                mx, l2 = rel_err(en.field(m, k), st.field(k))
                assert mx < tol and l2 < tol, (m, k, mx, l2)


def test_ensemble_48_members_configuration_4():
    """Configuration 4 at its real size: 48 members (16 each of 64^3, 96^3 and
    128x128x64), FP32-mixed, seeded draws of rho0, nu, zeta and undulation
    (the bench's workload, dt = 0.0025), graph replay, 10 steps: one member
    of each grid equals its single handle bit for bit and all 48 stay
    finite."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    members, ics = bench.ensemble_members(S, 48)
    assert len(members) == 48
    assert sorted({(p.nx, p.ny, p.nz) for p in members}) == [(64, 64, 64), (96, 96, 96), (128, 128, 64)]
    states = [S.make_interface_ic(p, ic) for p, ic in zip(members, ics)]
    with S.Ensemble(members) as en:
        assert en.groups == 3
        for m, (psi, rho) in enumerate(states):
            en.set_state(m, psi, rho)
        en.step(10)
        finite = sum(all(np.isfinite(v) for v in en.diagnostics(m, allow_nonfinite=True)) for m in range(48))
        assert finite == 48
        for m in (0, 16, 47):
            st = S.Stepper(members[m])
            st.set_state(*states[m])
            st.step(10)
            for k in ("psi", "rho", "vx", "vz", "amp"):
                assert np.array_equal(en.field(m, k), st.field(k)), (m, k)
            assert en.diagnostics(m) == st.diagnostics()
            st.close()
