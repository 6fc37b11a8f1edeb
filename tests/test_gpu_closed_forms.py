"""One GPU step from the closed-form states of test_oracle_pins: the forces,
density update, nonlinear term and psi update of the library against the
closed forms themselves (not against the oracle), on grids that run the
direct and the FFT transform paths, FP64 and FP32-mixed."""
import numpy as np
import pytest

import smaiso_amd as S
import test_oracle_pins as P

pytestmark = pytest.mark.gpu

GPU_FIELDS = ("lpsi", "amp", "vx", "vy", "vz", "divv", "rho", "f_rho", "mu", "divf", "nlq", "psiq")


def to_lib(p, precision):
    return S.Params(nx=p.nx, ny=p.ny, nz=p.nz, dx=p.dx, dy=p.dy, dz=p.dz, dt=p.dt, alpha=p.alpha,
                    beta=p.beta, gamma=p.gamma, epsilon=p.epsilon, q0=p.q0, nu=p.nu, lam=p.lam,
                    zeta=p.zeta, rho0=p.rho0, rho_m=p.rho_m, divv_switch=p.divv_switch,
                    precision=precision)


def gpu_case_errors(p, case, precision):
    psi, rho, want = case(p)
    st = S.Stepper(to_lib(p, precision))
    st.set_state(psi, rho)
    st.step(1)
    got = {k: st.field(k) for k in GPU_FIELDS if k in want}
    st.close()
    return P.case_errors(got, {k: want[k] for k in got})


CASES = [
    ("z-layers 4x6x16", lambda: P.layered_z_params(4, 6), P.layered_z_case),
    ("z-layers 16x32x16", lambda: P.layered_z_params(16, 32), P.layered_z_case),
    ("x-layers 16x4x6", lambda: P.layered_x_params(4, 6), P.layered_x_case),
    ("x-layers 16x32x16", lambda: P.layered_x_params(32, 16), P.layered_x_case),
]


@pytest.mark.parametrize("name,params,case", CASES, ids=[c[0] for c in CASES])
def test_gpu_step_matches_closed_form_fp64(name, params, case):
    err = gpu_case_errors(params(), case, S.FP64)
    assert max(err.values()) < 1e-11, err


@pytest.mark.parametrize("name,params,case", CASES, ids=[c[0] for c in CASES])
def test_gpu_step_matches_closed_form_mixed(name, params, case):
    err = gpu_case_errors(params(), case, S.FP32_MIXED)
    # FP32 storage of the work fields: a few FP32 ulps of each field's scale
    assert max(err.values()) < 2e-6, err
