"""The data of the reference's output writers (slices, surface tracker), full-
state checkpoints, the reference's text restart files and the command-line
driver that writes the reference's output set."""
import os
import subprocess

import numpy as np
import pytest

import smaiso_amd as S
import smaiso_oracle as O

pytestmark = pytest.mark.gpu


def _params(n=(32, 20, 48), **kw):
    base = dict(nx=n[0], ny=n[1], nz=n[2], dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0)
    base.update(kw)
    return S.Params(**base)


def _ic(p, seed=3):
    return S.make_interface_ic(p, S.InterfaceIC(rho_iso=p.rho0, undulation=0.6, seed=seed))


def test_slices_equal_field_planes():
    p = _params()
    st = S.Stepper(p)
    st.set_state(*_ic(p))
    st.step(3)
    for name in ("psi", "rho", "vx", "vz", "amp", "divv", "lpsi", "vsoly"):
        f = st.field(name)
        for j in (0, p.ny // 2, p.ny - 1):
            assert np.array_equal(st.slice(name, j), O.mid_slice(f, j)), (name, j)
    with pytest.raises(S.SmaisoError) as e:
        st.slice("psi", p.ny)
    assert e.value.code == S.E_INVALID


def test_surface_height_matches_oracle_with_stale_columns():
    p = _params((24, 16, 64))
    psi, rho = _ic(p)
    rho[3, 5, :] = 0.5          # never above the threshold: keeps the previous value
    st = S.Stepper(p)
    st.set_state(psi, rho)
    prev = np.arange(p.nx * p.ny, dtype=float) + 1000.0
    got = st.surface_height(0.85, prev)
    assert np.array_equal(got, O.surface_track(st.field("rho"), 0.85, prev))
    assert got[3 * p.ny + 5] == prev[3 * p.ny + 5]
    st.step(10)  # and on an evolved density
    got = st.surface_height(0.85, prev)
    assert np.array_equal(got, O.surface_track(st.field("rho"), 0.85, prev))


@pytest.mark.parametrize("precision", [S.FP64, S.FP32_MIXED])
def test_checkpoint_continuation_is_bit_identical(tmp_path, precision):
    p = _params(precision=precision)
    a = S.Stepper(p)
    a.set_state(*_ic(p))
    a.step(7)
    ck = tmp_path / "state.ck"
    a.save_checkpoint(ck)
    b = S.Stepper(p)
    b.set_graph(True)
    b.load_checkpoint(ck)
    assert b.nloop == 7
    a.step(8)
    b.step(8)
    for name in ("psi", "rho", "psiq", "psiq_old", "f_rho", "nlq", "vx", "mu"):
        assert np.array_equal(a.field(name), b.field(name)), name
    assert a.diagnostics() == b.diagnostics()
    # grid or precision mismatch and foreign files are refused
    other = S.Stepper(_params((32, 20, 32), precision=precision))
    with pytest.raises(S.SmaisoError) as e:
        other.load_checkpoint(ck)
    assert e.value.code == S.E_INVALID and "does not match" in str(e.value)
    bad = tmp_path / "bad.ck"
    bad.write_bytes(b"not a checkpoint" * 8)
    with pytest.raises(S.SmaisoError) as e:
        b.load_checkpoint(bad)
    assert e.value.code == S.E_INVALID



@pytest.mark.parametrize("precision", [S.FP64, S.FP32_MIXED])
def test_checkpoint_continuation_with_carried_lines(tmp_path, precision):
    """With 256 points along x each step leaves d_x psi and the z lines of
    the next step's L psi for the next step, and a checkpoint load rebuilds
    them from psiq: continuing from the checkpoint (graph replay on that
    side) stays bit-identical."""
    p = _params((256, 16, 48), precision=precision)
    a = S.Stepper(p)
    a.set_state(*_ic(p))
    a.step(5)
    ck = tmp_path / "state.ck"
    a.save_checkpoint(ck)
    b = S.Stepper(p)
    b.set_graph(True)
    b.load_checkpoint(ck)
    a.step(6)
    b.step(6)
    for name in ("psi", "rho", "psiq", "psiq_old", "nlq", "lpsi", "gpx", "amp", "mu", "vx"):
        assert np.array_equal(a.field(name), b.field(name)), name
    assert a.diagnostics() == b.diagnostics()


def test_derived_fields_need_a_step_after_set_or_load(tmp_path):
    """Fields a step produces (amp, L psi, v, psi_old, ...) and the diagnostics
    are refused until a step has run since set_state or a checkpoint load; the
    state fields are readable at once."""
    p = _params()
    psi, rho = _ic(p)
    a = S.Stepper(p)
    a.set_state(psi, rho)
    for name in ("amp", "lpsi", "vx", "psi_old", "energy", "nlq", "mu"):
        with pytest.raises(S.SmaisoError) as e:
            a.field(name)
        assert e.value.code == S.E_STATE, name
    with pytest.raises(S.SmaisoError) as e:
        a.diagnostics()
    assert e.value.code == S.E_STATE
    assert np.array_equal(a.field("rho"), rho)
    a.step(4)
    ck = tmp_path / "state.ck"
    a.save_checkpoint(ck)
    b = S.Stepper(p)
    b.load_checkpoint(ck)
    with pytest.raises(S.SmaisoError) as e:
        b.diagnostics()
    assert e.value.code == S.E_STATE
    with pytest.raises(S.SmaisoError) as e:
        b.field("amp")
    assert e.value.code == S.E_STATE
    assert np.array_equal(b.field("nlq"), a.field("nlq"))  # restored by the load
    assert np.array_equal(b.field("psi"), a.field("psi"))
    a.step(1)
    b.step(1)
    assert a.diagnostics() == b.diagnostics()
    assert np.array_equal(a.field("amp"), b.field("amp"))
    # a truncated checkpoint leaves the handle without a state
    data = ck.read_bytes()
    cut = tmp_path / "cut.ck"
    cut.write_bytes(data[: len(data) // 2])
    with pytest.raises(S.SmaisoError) as e:
        b.load_checkpoint(cut)
    assert e.value.code == S.E_INVALID and "truncated" in str(e.value)
    with pytest.raises(S.SmaisoError) as e:
        b.field("psi")
    assert e.value.code == S.E_STATE
    with pytest.raises(S.SmaisoError) as e:
        b.step(1)
    assert e.value.code == S.E_STATE

def _lines(path):
    with open(path) as f:
        return [float(x) for x in f.read().split()]


def _run(args, cwd):
    r = subprocess.run([S.RUN_PATH, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def test_driver_writes_the_reference_output_set(tmp_path):
    n = 32
    out = tmp_path / "run"
    out.mkdir()
    common = ["1", "5", "--n", n, n, n, "--dt", "0.01", "--rho0", "0.8", "--rho-m", "1",
              "--divv-switch", "0"]
    _run(common + ["--steps", 100, "--save-every", 50, "--out", out, "--checkpoint", out / "ck.bin"], tmp_path)
    names = ["info", "L1.dat", "mass.dat", "energy.dat", "psiMid.dat", "energyMid.dat", "vx.dat", "vy.dat",
             "vz.dat", "vsolx.dat", "vsoly.dat", "vsolz.dat", "sx.dat", "sy.dat", "sz.dat", "divv.dat",
             "rho.dat", "amp.dat", "surfPsi.dat", "psi0.dat"]
    for nm in names:
        assert (out / nm).exists(), nm
    assert len(_lines(out / "L1.dat")) == 2
    for nm in ("psiMid.dat", "vx.dat", "vsolz.dat", "divv.dat", "rho.dat", "amp.dat", "energyMid.dat"):
        assert len(_lines(out / nm)) == 2 * n * n, nm
    assert len(_lines(out / "surfPsi.dat")) == 2 * n * n
    assert len(_lines(out / "psi0.dat")) == n ** 3
    assert os.path.getsize(out / "sx.dat") == 0  # created and never written, as in the reference

    # the same run through the library: the files hold its values to 6 digits
    p = S.Params(nx=n, ny=n, nz=n, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=1.0, epsilon=-0.005)
    psi, rho = S.make_interface_ic(p, S.InterfaceIC(amplitude=1.0, h0=-1.0, width=2.0, rho_iso=0.8,
                                                    rho_sm=1.0, undulation=0.0, noise=1e-3, seed=1))
    st = S.Stepper(p)
    st.set_state(psi, rho)
    st.step(50)
    surf = st.surface_height(0.85, np.zeros(n * n))
    st.step(50)
    surf = st.surface_height(0.85, surf)
    close = lambda a, b: np.allclose(a, b, rtol=1e-5, atol=1e-12 * max(1.0, np.abs(b).max()))
    assert close(_lines(out / "psiMid.dat")[n * n:], st.slice("psi"))
    assert close(_lines(out / "vx.dat")[n * n:], st.slice("vx"))
    assert close(_lines(out / "rho.dat")[n * n:], st.slice("rho"))
    assert _lines(out / "surfPsi.dat")[n * n:] == list(surf)
    assert close(_lines(out / "psi0.dat"), st.field("psi").reshape(-1))
    ps, lp = st.slice("psi"), st.slice("lpsi")
    e = -0.5 * p.epsilon * ps ** 2 + 0.5 * lp ** 2 - 0.25 * p.beta * ps ** 4 + (p.gamma / 6) * ps ** 6
    assert close(_lines(out / "energyMid.dat")[n * n:], e)
    l1, mass, en = st.diagnostics()
    assert close(_lines(out / "L1.dat")[-1:], [l1]) and close(_lines(out / "mass.dat")[-1:], [mass])

    # restart from the checkpoint: 50 more steps equal an uninterrupted 150-step run
    out2, out3 = tmp_path / "restart", tmp_path / "straight"
    out2.mkdir()
    out3.mkdir()
    _run(common + ["--restart", out / "ck.bin", "--steps", 50, "--save-every", 50, "--out", out2], tmp_path)
    _run(common + ["--steps", 150, "--save-every", 50, "--out", out3], tmp_path)
    assert (out2 / "psi0.dat").read_text() == (out3 / "psi0.dat").read_text()

    # the reference's load mode from its psi0.dat (rho = 0, no relaxation)
    out4 = tmp_path / "load"
    out4.mkdir()
    stdout = _run(common + ["--load", out, "--steps", 50, "--out", out4], tmp_path)
    assert "step 50" in stdout
    assert np.isfinite(_lines(out4 / "L1.dat")).all()


def test_driver_runs_until_stationary_reference_ic(tmp_path):
    """The run-until-stationary loop (S:1081, S:2246) without --steps, from the
    reference's own IC (Amp cos(q0 z) layers, rho = 0, S:571-590) with the
    relaxation branch for the first divvSwitch = 50 steps (S:206-214,
    S:1312-1380): the driver checks L1 every 50 steps and stops at the first
    check where L1 <= the threshold.  Its L1.dat equals the library's
    diagnostics of the same run, and it stops where they say it must."""
    n = (16, 16, 32)
    p = S.Params(nx=n[0], ny=n[1], nz=n[2], dt=0.001, rho0=0.01, divv_switch=50, nu=1.0, epsilon=-0.005)
    op = O.Params(nx=n[0], ny=n[1], nz=n[2], dt=0.001, rho0=0.01, divv_switch=50)
    psi, rho = O.reference_layers_ic(op)
    st = S.Stepper(p)
    st.set_state(psi, rho)
    trace = []
    for _ in range(12):
        st.step(50)
        trace.append(st.diagnostics()[0])
    st.close()
    # a threshold the trace crosses: the 7th check's L1; the run must stop at
    # the first check at or below it
    lim = trace[6]
    stop = next(i for i, v in enumerate(trace) if v <= lim)
    assert stop >= 1, trace
    out = tmp_path / "run"
    out.mkdir()
    stdout = _run(["1", "5", "--n", *n, "--dt", "0.001", "--rho0", "0.01", "--divv-switch", "50",
                   "--ic", "reference", "--lim-l1", repr(lim), "--out", out], tmp_path)
    l1 = _lines(out / "L1.dat")
    assert len(l1) == stop + 1, (len(l1), stop, trace)
    assert f"{50 * (stop + 1)} steps in" in stdout
    assert np.allclose(l1, trace[: stop + 1], rtol=1e-5, atol=0), (l1, trace)
    assert l1[-1] <= lim * (1 + 1e-5) and all(v > lim for v in trace[:stop])
    # the default threshold is the reference's 1e-6 (S:1029); a bad one is refused
    r = subprocess.run([S.RUN_PATH, "1", "5", "--lim-l1", "-1"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "lim-l1" in r.stderr
