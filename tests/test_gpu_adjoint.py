"""GPU adjoint (smaiso_adjoint) against the adjoint oracle and central finite
differences of the GPU forward stepper."""
import dataclasses

import numpy as np
import pytest

import smaiso_adjoint as A
import smaiso_amd as S
import smaiso_oracle as O
from helpers import params_pair, rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _height(p, amp=0.5, shift=0.0):
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    return (p.nz * p.dz / 2 + shift + amp * np.cos(2 * np.pi * x / (p.nx * p.dx))[:, None]
            * np.cos(2 * np.pi * y / (p.ny * p.dy))[None, :])


def _problem(sp, seed=4):
    """IC on an explicit height map and a target from a perturbed height."""
    h = _height(sp)
    ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=sp.rho0, rho_sm=1.0, noise=1e-3, seed=seed)
    psi0, rho0 = S.make_interface_ic(sp, ic, height=h)
    target, _ = S.make_interface_ic(sp, dataclasses.replace(ic, seed=seed + 7), height=_height(sp, 0.8, 0.6))
    return ic, h, psi0, rho0, target


@pytest.mark.parametrize("shape,nsteps,every", [
    ((16, 12, 20), 6, 4),    # FFT and direct lengths, a partial last segment
    ((16, 16, 16), 1, 50),   # the first step alone (quirk Q5)
    ((24, 20, 32), 5, 1),
    ((16, 16, 16), 0, 3),    # loss of the projected initial state
])
def test_adjoint_matches_oracle(shape, nsteps, every):
    sp, op = params_pair(nx=shape[0], ny=shape[1], nz=shape[2], dt=0.01, rho0=0.85, rho_m=1.0,
                         divv_switch=0, nu=0.7, zeta=1.3, lam=0.9)
    _, _, psi0, rho0, target = _problem(sp)
    st = S.Stepper(sp)
    g = st.adjoint(psi0, rho0, target, nsteps, every)
    r = A.loss_and_gradients(op, psi0, rho0, target, nsteps)
    assert abs(g.loss - r.loss) <= 1e-12 * abs(r.loss)
    for name, got, want in (("psi0", g.psi0, r.psi0), ("rho0", g.rho0, r.rho0)):
        mx, l2 = rel_err(got, want)
        assert mx < TOL64 and l2 < TOL64, (name, mx, l2)
    if nsteps > 0:
        assert abs(g.zeta - r.zeta) <= TOL64 * abs(r.zeta), (g.zeta, r.zeta)
        assert abs(g.nu - r.nu) <= TOL64 * abs(r.nu), (g.nu, r.nu)
    else:
        assert g.zeta == 0 and g.nu == 0


def test_adjoint_independent_of_checkpoint_interval():
    sp = S.Params(nx=16, ny=16, nz=24, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0)
    _, _, psi0, rho0, target = _problem(sp)
    ref = None
    for every in (1, 3, 7, 100):
        g = S.Stepper(sp).adjoint(psi0, rho0, target, 7, every)
        if ref is None:
            ref = g
            continue
        assert g.loss == ref.loss and g.zeta == ref.zeta and g.nu == ref.nu
        assert np.array_equal(g.psi0, ref.psi0) and np.array_equal(g.rho0, ref.rho0)


def _fd_loss(sp, ic, h, target, nsteps):
    psi0, rho0 = S.make_interface_ic(sp, ic, height=h)
    st = S.Stepper(sp)
    st.set_state(psi0, rho0)
    st.step(nsteps)
    r = st.field("psi") - target
    return float(np.sum(r * r))


def _check_fd(sp, nsteps, every, tol):
    """Central differences (relative step 1e-6) in zeta, nu and one smooth
    direction of the interface height, against the adjoint."""
    ic, h, psi0, rho0, target = _problem(sp)
    st = S.Stepper(sp)
    g = st.adjoint(psi0, rho0, target, nsteps, every)
    dh = S.interface_ic_vjp(sp, ic, g.psi0, g.rho0, height=h)
    out = {}
    for name, got in (("zeta", g.zeta), ("nu", g.nu)):
        v = getattr(sp, name)
        e = 1e-6 * v
        lp = _fd_loss(dataclasses.replace(sp, **{name: v + e}), ic, h, target, nsteps)
        lm = _fd_loss(dataclasses.replace(sp, **{name: v - e}), ic, h, target, nsteps)
        out[name] = ((lp - lm) / (2 * e), got)
    x = np.arange(sp.nx) * sp.dx
    y = np.arange(sp.ny) * sp.dy
    dirn = np.sin(2 * np.pi * x / (sp.nx * sp.dx))[:, None] * np.cos(4 * np.pi * y / (sp.ny * sp.dy))[None, :] + 0.3
    e = 1e-6
    lp = _fd_loss(sp, ic, h + e * dirn, target, nsteps)
    lm = _fd_loss(sp, ic, h - e * dirn, target, nsteps)
    out["h"] = ((lp - lm) / (2 * e), float(np.sum(dh * dirn)))
    for name, (fd, got) in out.items():
        assert abs(fd - got) <= tol * abs(fd), (name, fd, got, abs(fd - got) / abs(fd))
    return out


def test_adjoint_vs_finite_differences_64():
    sp = S.Params(nx=64, ny=64, nz=64, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=0.6, zeta=1.4)
    _check_fd(sp, 40, 10, 1e-5)


def test_adjoint_vs_finite_differences_128_500_steps():
    """128^3 FP64, 500 steps, a checkpoint every 50: dL/dzeta, dL/dnu and
    dL/dh against central differences to 1e-5 relative."""
    sp = S.Params(nx=128, ny=128, nz=128, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=0.6, zeta=1.4)
    res = _check_fd(sp, 500, 50, 1e-5)
    print({k: (fd, got, abs(fd - got) / abs(fd)) for k, (fd, got) in res.items()})


def test_adjoint_errors_and_state():
    sp = S.Params(nx=16, ny=16, nz=16, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=3)
    _, _, psi0, rho0, target = _problem(sp)
    st = S.Stepper(sp)
    with pytest.raises(S.SmaisoError) as e:
        st.adjoint(psi0, rho0, target, 4)
    assert e.value.code == S.E_INVALID and "steady" in str(e.value)
    sp2 = dataclasses.replace(sp, divv_switch=0)
    st2 = S.Stepper(sp2)
    with pytest.raises(S.SmaisoError) as e:
        st2.adjoint(psi0, rho0, target, 4, 0)
    assert e.value.code == S.E_INVALID
    st2.adjoint(psi0, rho0, target, 4, 2)
    with pytest.raises(S.SmaisoError) as e:
        st2.step(1)  # the adjoint consumed the state
    assert e.value.code == S.E_STATE
    st2.set_state(psi0, rho0)
    st2.step(1)


@pytest.mark.parametrize("precision,tol", [(S.FP64, 1e-11), (S.FP32_MIXED, 1e-6)])
def test_adjoint_step_store_matches_recompute(precision, tol):
    """Intermediates kept while a segment is recomputed (the default) against
    recomputing every step again before its transpose: the same gradients,
    up to the rounding by which the two recomputations differ (the first
    runs the forward sweep's steps and solves virr^ afterwards, the second
    runs the steps that keep virr^).  nz = 32 puts lap(rho L psi)'s last z
    pass in the forces pass of the FP64 step."""
    sp = S.Params(nx=24, ny=16, nz=32, dt=0.01, rho0=0.85, rho_m=1.0, divv_switch=0, nu=0.7, zeta=1.3,
                  lam=0.9, precision=precision)
    _, _, psi0, rho0, target = _problem(sp)
    a = S.Stepper(sp).adjoint(psi0, rho0, target, 9, 4)
    st = S.Stepper(sp)
    st.set_adjoint_store(False)
    b = st.adjoint(psi0, rho0, target, 9, 4)
    assert a.loss == b.loss  # the forward sweep is the same
    for name, got, want in (("psi0", a.psi0, b.psi0), ("rho0", a.rho0, b.rho0)):
        mx, l2 = rel_err(got, want)
        assert mx < tol and l2 < tol, (name, mx, l2)
    assert abs(a.zeta - b.zeta) <= tol * abs(b.zeta), (a.zeta, b.zeta)
    assert abs(a.nu - b.nu) <= tol * abs(b.nu), (a.nu, b.nu)


def test_adjoint_fp32_mixed_close_to_fp64():
    """The FP32-mixed trajectory's gradients against the FP64 ones (the
    adjoint fields are FP64 in both)."""
    sp = S.Params(nx=64, ny=64, nz=64, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=0.6, zeta=1.4)
    spm = dataclasses.replace(sp, precision=S.FP32_MIXED)
    _, _, psi0, rho0, target = _problem(sp)
    g64 = S.Stepper(sp).adjoint(psi0, rho0, target, 40, 10)
    gm = S.Stepper(spm).adjoint(psi0, rho0, target, 40, 10)
    errs = {"loss": abs(gm.loss - g64.loss) / abs(g64.loss),
            "zeta": abs(gm.zeta - g64.zeta) / abs(g64.zeta),
            "nu": abs(gm.nu - g64.nu) / abs(g64.nu),
            "psi0": rel_err(gm.psi0, g64.psi0)[1], "rho0": rel_err(gm.rho0, g64.rho0)[1]}
    print(errs)
    assert all(v < 1e-4 for v in errs.values()), errs


def test_adjoint_fp32_mixed_128_500_steps_vs_fp64_finite_differences():
    """Configuration 5's FP32-mixed leg: 128^3, 500 steps, a checkpoint every
    50; the mixed adjoint's dL/dzeta, dL/dnu and dL/dh (one smooth direction)
    against central differences (relative step 1e-6) of the FP64 forward
    run, i.e. the FP64 tangent, to 1e-5 relative.  Differences of FP32-mixed
    runs themselves would be dominated by their rounding at this step."""
    sp = S.Params(nx=128, ny=128, nz=128, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=0.6, zeta=1.4)
    spm = dataclasses.replace(sp, precision=S.FP32_MIXED)
    ic, h, psi0, rho0, target = _problem(sp)
    g = S.Stepper(spm).adjoint(psi0, rho0, target, 500, 50)
    dh = S.interface_ic_vjp(sp, ic, g.psi0, g.rho0, height=h)
    out = {}
    for name, got in (("zeta", g.zeta), ("nu", g.nu)):
        v = getattr(sp, name)
        e = 1e-6 * v
        lp = _fd_loss(dataclasses.replace(sp, **{name: v + e}), ic, h, target, 500)
        lm = _fd_loss(dataclasses.replace(sp, **{name: v - e}), ic, h, target, 500)
        out[name] = ((lp - lm) / (2 * e), got)
    x = np.arange(sp.nx) * sp.dx
    y = np.arange(sp.ny) * sp.dy
    dirn = np.sin(2 * np.pi * x / (sp.nx * sp.dx))[:, None] * np.cos(4 * np.pi * y / (sp.ny * sp.dy))[None, :] + 0.3
    e = 1e-6
    lp = _fd_loss(sp, ic, h + e * dirn, target, 500)
    lm = _fd_loss(sp, ic, h - e * dirn, target, 500)
    out["h"] = ((lp - lm) / (2 * e), float(np.sum(dh * dirn)))
    rel = {k: abs(fd - got) / abs(fd) for k, (fd, got) in out.items()}
    print({k: (fd, got, rel[k]) for k, (fd, got) in out.items()})
    assert all(v <= 1e-5 for v in rel.values()), rel
