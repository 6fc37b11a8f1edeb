"""GPU stepper vs the FP64 oracle (per field, per step) and error paths."""
import numpy as np
import pytest

import smaiso_amd as S
import smaiso_oracle as O
from helpers import config1, group_scales, params_pair, rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-10  # per-field max-abs and L2 error, relative to the field's scale


def oracle_fields(o: O.Oracle):
    f, x = o.f, o.f.extra
    d = dict(psi=f.psi, rho=f.rho, psiq=f.psiq, psiq_old=f.psiq_old, nlq=f.nlq, f_rho=f.f_rho)
    for k in ("amp", "vx", "vy", "vz", "vsolx", "vsoly", "vsolz", "virrx", "virry", "virrz",
              "divv", "divf", "lpsi", "mu", "gpx", "gpy", "gpz", "rdx", "rdy", "rdz", "lap_rl",
              "psi_old"):
        d[k] = x[k]
    d["energy"] = O.energy_density(o.p, f.psi, x["lpsi"], f.rho, x["amp"])
    return d


def run_pair(sp, op, psi, rho, steps):
    o = O.Oracle(op)
    o.set_state(psi, rho)
    st = S.Stepper(sp)
    st.set_state(psi, rho)
    o.step(steps)
    st.step(steps)
    return st, o


def compare_all(st, o, tol, names=None):
    ref = oracle_fields(o)
    sc = group_scales(ref)
    bad = {}
    for name in names or S.FIELDS:
        got = st.field(name)
        mx, l2 = rel_err(got, ref[name], *sc[name])
        if not (mx < tol and l2 < tol):
            bad[name] = (mx, l2)
    assert not bad, bad


def test_first_step_every_field_odd_grid():
    """One step on a non-cubic grid mixing the FFT (32, 48) and direct (20) paths."""
    sp, op = params_pair(nx=32, ny=20, nz=48, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(undulation=0.5, seed=5))
    st, o = run_pair(sp, op, psi, rho, 1)
    compare_all(st, o, TOL64)


def test_carried_lines_every_field_nx256():
    """Three steps on 256 x 16 x 48, where each step's psi transform also
    leaves d_x psi and the z lines of the next step's L psi (the 256-point x
    kernel; z length 48 on the FFT path): every field against the oracle."""
    sp, op = params_pair(nx=256, ny=16, nz=48, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(undulation=0.5, seed=6))
    st, o = run_pair(sp, op, psi, rho, 3)
    compare_all(st, o, TOL64)


def test_config1_20_steps_every_field():
    sp, op = config1(64)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(seed=1))
    st, o = run_pair(sp, op, psi, rho, 20)
    compare_all(st, o, TOL64)
    l1, mass, en = st.diagnostics()
    rl1, rmass, ren = o.diagnostics()
    assert abs(l1 - rl1) <= 1e-10 * abs(rl1)
    assert abs(mass - rmass) <= 1e-12 * abs(rmass)
    assert abs(en - ren) <= 1e-10 * abs(ren)


def test_config1_200_steps_trajectory():
    """Config 1 as specified (64^3, dt=0.01, 200 steps, FP64), with the mass
    and free-energy trajectories every 50 steps."""
    sp, op = config1(64)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(seed=1))
    o = O.Oracle(op); o.set_state(psi, rho)
    st = S.Stepper(sp); st.set_state(psi, rho)
    for _ in range(4):
        o.step(50); st.step(50)
        l1, mass, en = st.diagnostics()
        rl1, rmass, ren = o.diagnostics()
        assert abs(mass - rmass) <= 1e-12 * abs(rmass), (mass, rmass)
        assert abs(en - ren) <= 1e-10 * abs(ren), (en, ren)
        assert abs(l1 - rl1) <= 1e-9 * abs(rl1), (l1, rl1)
    compare_all(st, o, TOL64, ["psi", "rho", "vx", "vy", "vz", "amp", "psiq"])


def test_relaxation_branch_reference_ic():
    """divvSwitch > 0 with the reference's own IC (S:571-590) and rho = 0:
    rho is slaved to the amplitude for the first steps (quirk Q4)."""
    sp, op = params_pair(nx=32, ny=16, nz=64, dt=0.001, divv_switch=3)
    psi, rho = O.reference_layers_ic(op)
    st, o = run_pair(sp, op, psi, rho, 6)
    compare_all(st, o, TOL64)


@pytest.mark.parametrize("slab", [1, 8, 5])
def test_slab_quirk_q3(slab):
    """Quirk Q3 follows the emulated x-slab height."""
    sp, op = params_pair(nx=16, ny=16, nz=16, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, slab_n0=slab)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(rho_iso=0.8, undulation=0.5, seed=9))
    st, o = run_pair(sp, op, psi, rho, 2)
    compare_all(st, o, TOL64, ["virrz", "vz", "psi", "rho"])


def test_load_mode_skips_relaxation():
    sp, op = params_pair(nx=16, ny=16, nz=32, dt=0.005, rho0=0.9, rho_m=1.0, divv_switch=50, load_mode=1)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(seed=2))
    st, o = run_pair(sp, op, psi, rho, 3)
    compare_all(st, o, TOL64, ["psi", "rho", "lap_rl"])


def test_uneven_parameters():
    """Non-default nu, lambda, zeta, alpha, beta, gamma, q0, anisotropic dx."""
    kw = dict(nx=24, ny=32, nz=40, dx=0.7, dy=0.8, dz=0.75, dt=0.004, alpha=1.3, beta=1.7,
              gamma=0.9, epsilon=-0.02, q0=1.1, nu=0.3, lam=0.5, zeta=1.7, rho0=0.85, rho_m=1.1,
              divv_switch=1)
    sp, op = params_pair(**kw)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(rho_iso=0.85, undulation=0.7, seed=4))
    st, o = run_pair(sp, op, psi, rho, 3)
    compare_all(st, o, TOL64)


def test_mixed_precision_close_to_fp64():
    """FP32 storage with FP64 arithmetic: after 20 steps at 64^3 the fields
    sit within the rounding of their FP32 copies of the FP64 oracle."""
    sp, op = config1(64, precision=S.FP32_MIXED)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(seed=1))
    st, o = run_pair(sp, op, psi, rho, 20)
    ref = oracle_fields(o)
    sc = group_scales(ref)
    for name, tol in (("psi", 1e-6), ("rho", 1e-7), ("vx", 1e-4), ("vz", 1e-4)):
        mx, l2 = rel_err(st.field(name), ref[name], *sc[name])
        assert l2 < tol, (name, mx, l2)


def test_mixed_precision_psi_within_1e6_of_fp64():
    """The FP32-mixed accuracy target (psi within 1e-6 relative L2 of the
    FP64 stepper) on the perturbed-interface state at 128^3 after 300 steps
    (the bench checks 256^3 after 2000 steps)."""
    n = 128
    ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.8, rho_sm=1.0, undulation=0.5,
                       noise=1e-3, seed=2)

    def params(prec):
        return S.Params(nx=n, ny=n, nz=n, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0,
                        precision=prec)

    psi, rho = S.make_interface_ic(params(S.FP64), ic)
    a, b = S.Stepper(params(S.FP64)), S.Stepper(params(S.FP32_MIXED))
    a.set_state(psi, rho)
    b.set_state(psi, rho)
    a.step(300)
    b.step(300)
    for name, tol in (("psi", 1e-6), ("rho", 1e-7)):
        fa, fb = a.field(name), b.field(name)
        err = np.linalg.norm(fb - fa) / np.linalg.norm(fa)
        assert err < tol, (name, err)


def test_mixed_precision_flat_interface_256():
    """FP32-mixed on the flat-interface state (rho 1.0/0.9, rho0 0.9) at 256^3
    after 500 steps: psi within 1e-7 of the FP64 stepper with the FP32 stages
    that stay at the FP32 storage level (the default without the amplitude
    filter), and within 3e-7 with the default.  This near-stationary state is
    the one where FP32 rounding of partly transformed values between the
    passes of psi's inverse transform used to dominate (3.6e-7 here); with
    FP64 intermediates the error is the FP32 rounding of psi itself (about
    3e-8)."""
    shape = (256, 256, 256)
    ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.9, rho_sm=1.0, undulation=0.0,
                       noise=1e-3, seed=1)

    def params(prec):
        return S.Params(nx=shape[0], ny=shape[1], nz=shape[2], dt=0.01, rho0=0.9, rho_m=1.0,
                        divv_switch=0, precision=prec)

    psi, rho = S.make_interface_ic(params(S.FP64), ic)
    a, b, d = S.Stepper(params(S.FP64)), S.Stepper(params(S.FP32_MIXED)), S.Stepper(params(S.FP32_MIXED))
    b.set_fp32_stages(S.fp32_stages_default() & ~S.STAGES["amp"])
    for st in (a, b, d):
        st.set_state(psi, rho)
        st.step(500)
    for st, ptol in ((b, 1e-7), (d, 3e-7)):
        for name, tol in (("psi", ptol), ("rho", 1e-8 if st is b else 3e-8)):
            fa, fb = a.field(name), st.field(name)
            err = np.linalg.norm(fb - fa) / np.linalg.norm(fa)
            assert err < tol, (name, err, "default" if st is d else "storage-level stages")
    a.close()
    b.close()



def test_mixed_precision_flat_interface_512x512x256():
    """The configuration-3 check at its large grid: FP32-mixed against the
    FP64 stepper from the flat-interface state on 512x512x256, psi within
    1e-6 relative L2 at steps 100 and 500 (and rho, v reported)."""
    shape = (512, 512, 256)
    ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.9, rho_sm=1.0, undulation=0.0,
                       noise=1e-3, seed=1)

    def params(prec):
        return S.Params(nx=shape[0], ny=shape[1], nz=shape[2], dt=0.01, rho0=0.9, rho_m=1.0,
                        divv_switch=0, precision=prec)

    psi, rho = S.make_interface_ic(params(S.FP64), ic)
    a, b = S.Stepper(params(S.FP64)), S.Stepper(params(S.FP32_MIXED))
    a.set_state(psi, rho)
    b.set_state(psi, rho)
    del psi, rho
    done = 0
    for c in (100, 500):
        a.step(c - done)
        b.step(c - done)
        done = c
        errs = {}
        for name in ("psi", "rho", "vx"):
            fa, fb = a.field(name), b.field(name)
            errs[name] = float(np.linalg.norm(fb - fa) / np.linalg.norm(fa))
        print(f"step {c}: {errs}")
        assert errs["psi"] < 1e-6, (c, errs)
        assert errs["rho"] < 1e-7, (c, errs)
    a.close()
    b.close()

def test_mixed_odd_lengths_direct_cast_path():
    """FP32-mixed on a grid whose x (20) and z (40) lengths run on the direct
    path, so the FP32 -> FP64 z pass and the FP64 -> FP32 x pass of psi's
    transform run the direct kernel's casting instances: 3 steps against the
    FP64 oracle at FP32 storage accuracy."""
    sp, op = params_pair(nx=20, ny=32, nz=40, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0,
                         precision=S.FP32_MIXED)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(undulation=0.5, seed=5))
    st, o = run_pair(sp, op, psi, rho, 3)
    ref = oracle_fields(o)
    sc = group_scales(ref)
    for name, tol in (("psi", 1e-6), ("rho", 1e-7), ("psiq", 1e-6)):
        mx, l2 = rel_err(st.field(name), ref[name], *sc[name])
        assert l2 < tol, (name, mx, l2)


def test_uniform_state_is_stationary():
    sp, op = params_pair(nx=16, ny=16, nz=16, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    psi = np.zeros(sp.shape)
    rho = np.full(sp.shape, 0.9)
    st = S.Stepper(sp)
    st.set_state(psi, rho)
    st.step(5)
    assert np.abs(st.field("psi")).max() == 0.0
    assert np.abs(st.field("rho") - 0.9).max() < 1e-15


def test_error_paths():
    sp, _ = config1(16)
    st = S.Stepper(sp)
    with pytest.raises(S.SmaisoError) as e:
        st.step(1)
    assert e.value.code == S.E_STATE
    psi = np.zeros(sp.shape); rho = np.full(sp.shape, 0.9)
    st.set_state(psi, rho)
    with pytest.raises(S.SmaisoError) as e:
        st.diagnostics()
    assert e.value.code == S.E_STATE
    with pytest.raises(KeyError):
        st.field("nope")
    psi[3, 4, 5] = np.nan
    st.set_state(psi, rho)
    st.step(1)
    with pytest.raises(S.SmaisoError) as e:
        st.diagnostics()
    assert e.value.code == S.E_NONFINITE


def test_nloop_and_async():
    sp, op = config1(32)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(seed=8))
    st = S.Stepper(sp)
    st.set_state(psi, rho)
    st.step_async(3)
    st.synchronize()
    assert st.nloop == 3
    a = st.field("psi")
    st.set_state(psi, rho)
    assert st.nloop == 0
    st.step(3)
    assert np.array_equal(a, st.field("psi"))


def test_config2_full_size_two_steps():
    """Config 2 at its full size (256^3, perturbed interface, rho 1.0/0.8,
    dt=0.01): every field after 2 steps within 1e-10 of the FP64 oracle."""
    kw = dict(nx=256, ny=256, nz=256, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0)
    sp, op = params_pair(**kw)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(rho_iso=0.8, undulation=0.5, seed=2))
    st, o = run_pair(sp, op, psi, rho, 2)
    compare_all(st, o, TOL64, ["psi", "rho", "vx", "vy", "vz", "amp", "mu", "divv", "psiq", "lap_rl"])
    l1, mass, en = st.diagnostics()
    rl1, rmass, ren = o.diagnostics()
    assert abs(mass - rmass) <= 1e-12 * abs(rmass)
    assert abs(en - ren) <= 1e-10 * abs(ren)


def test_config2_trace_against_oracle_fixture():
    """Config 2's diagnostics trace (256^3, perturbed interface, 1.0/0.8,
    dt = 0.01): L1, mass and free energy every 50 steps, the reference's
    logging cadence, against the FP64 oracle's trace committed in
    tests/golden/config2_trace.json (made by make_config2_trace.py)."""
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "config2_trace.json")
    with open(path) as f:
        doc = json.load(f)
    cfg = doc["config"]
    trace = doc["trace"]
    assert trace and trace[-1]["step"] >= 50
    n = cfg["n"]
    p = S.Params(nx=n, ny=n, nz=n, dt=cfg["dt"], rho0=cfg["rho0"], rho_m=cfg["rho_m"],
                 divv_switch=cfg["divv_switch"])
    psi, rho = S.make_interface_ic(p, S.InterfaceIC(**cfg["ic"]))
    st = S.Stepper(p)
    st.set_state(psi, rho)
    done = 0
    worst = dict(l1=0.0, mass=0.0, energy=0.0)
    for row in trace:
        st.step(row["step"] - done)
        done = row["step"]
        l1, mass, en = st.diagnostics()
        errs = dict(l1=abs(l1 - row["l1"]) / abs(row["l1"]), mass=abs(mass - row["mass"]) / abs(row["mass"]),
                    energy=abs(en - row["energy"]) / abs(row["energy"]))
        worst = {k: max(worst[k], errs[k]) for k in worst}
        assert errs["mass"] <= 1e-12, (done, errs)
        assert errs["energy"] <= 1e-10 and errs["l1"] <= 1e-10, (done, errs)
    print(f"{done} steps against the oracle trace, worst relative errors {worst}")
    st.close()


@pytest.mark.parametrize("mask", ["default", "all"])
def test_mixed_fp32_stages_one_step_every_field(mask):
    """FP32 arithmetic in the transform passes (the 256-point y/x kernels and
    the 64-point z, one-axis and fused-load kernels of this grid) against the
    FP64 oracle after one step: each FP32 pass rounds relative to its line's
    norm, so every field stays within 2e-5 of its scale (a wrong kernel
    would be off by O(1)); with the default stages psi, rho and psiq stay
    at FP32 storage accuracy."""
    sp, op = params_pair(nx=256, ny=256, nz=64, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0,
                         precision=S.FP32_MIXED)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(undulation=0.5, seed=7))
    o = O.Oracle(op)
    o.set_state(psi, rho)
    o.step(1)
    st = S.Stepper(sp)
    if mask == "all":
        st.set_fp32_stages(sum(S.STAGES.values()))
    st.set_state(psi, rho)
    st.step(1)
    ref = oracle_fields(o)
    sc = group_scales(ref)
    worst = {}
    for name in S.FIELDS:
        mx, l2 = rel_err(st.field(name), ref[name], *sc[name])
        worst[name] = mx
        assert mx < 2e-5 and l2 < 2e-5, (name, mx, l2)
    if mask == "default":
        for name in ("psi", "rho", "psiq"):
            assert worst[name] < 1e-6, (name, worst[name])
    st.close()


def test_fp32_stages_api():
    assert S.fp32_stages_default() == (S.STAGES["amp"] | S.STAGES["forces"] | S.STAGES["divv"]
                                       | S.STAGES["divf"] | S.STAGES["velocity"] | S.STAGES["nl"]
                                       | S.STAGES["velsolve"] | S.STAGES["grad_rho"])
    sp, op = config1(32)
    psi, rho = O.make_interface_ic(op, O.InterfaceIC(seed=3))
    a, b = S.Stepper(sp), S.Stepper(sp)
    b.set_fp32_stages(sum(S.STAGES.values()))  # an FP64 handle ignores the stages
    for st in (a, b):
        st.set_state(psi, rho)
        st.step(3)
    assert np.array_equal(a.field("psi"), b.field("psi"))
    with pytest.raises(S.SmaisoError) as e:
        a.set_fp32_stages(1 << 12)
    assert e.value.code == S.E_INVALID
