"""Pins of the adjoint oracle (CPU): transposes, dot-product identities against
complex-step tangents of the forward oracle, and loss gradients against
complex-step and central finite differences."""
import dataclasses

import numpy as np
import pytest

import smaiso_adjoint as A
import smaiso_oracle as O

H = 1e-20  # complex-step size: J v = Im F(x + i H v) / H, exact to rounding


def small_params(**kw):
    base = dict(nx=8, ny=6, nz=10, dt=0.01, rho0=0.85, rho_m=1.0, divv_switch=0, nu=0.7,
                zeta=1.3, lam=0.9)
    base.update(kw)
    return O.Params(**base)


def dense(fn, shape):
    n = int(np.prod(shape))
    cols = [fn(np.eye(n)[k].reshape(shape)).ravel() for k in range(n)]
    return np.array(cols).T


@pytest.mark.parametrize("which", ["ct", "ict", "st0", "st1", "st2", "ist0", "ist1", "ist2"])
def test_transposed_transforms_are_matrix_transposes(which):
    shape = (4, 5, 3)
    W = A.Weights(shape)
    fwd = {"ct": lambda x: O.r2r3d(x, O.CT), "ict": lambda x: O.r2r3d(x, O.ICT)}
    adj = {"ct": lambda y: A.ct_T(W, y), "ict": lambda y: A.ict_T(W, y)}
    for a in range(3):
        fwd[f"st{a}"] = lambda x, a=a: O.r2r3d(x, O.ST[a])
        fwd[f"ist{a}"] = lambda x, a=a: O.r2r3d(x, O.IST[a])
        adj[f"st{a}"] = lambda y, a=a: A.st_T(W, y, a)
        adj[f"ist{a}"] = lambda y, a=a: A.ist_T(W, y, a)
    M = dense(fwd[which], shape)
    MT = dense(adj[which], shape)
    assert np.allclose(MT, M.T, rtol=0, atol=1e-12 * np.abs(M).max())


def test_shift_transposes():
    p = small_params()
    t = O.Tables(p)
    rng = np.random.default_rng(0)
    for ax in range(3):
        x, y = rng.standard_normal(p.shape), rng.standard_normal(p.shape)
        assert np.isclose(np.sum(y * O._shift_m(x, ax)), np.sum(A.shift_p(y, ax) * x), rtol=1e-13)
        assert np.isclose(np.sum(y * A.grad_shift(t, x, ax)), np.sum(A.grad_shift_T(t, y, ax) * x),
                          rtol=1e-13)
    # grad_shift is the oracle's own spectral gradient step (S:1166-1221)
    psiq = rng.standard_normal(p.shape)
    gpx, gpy, gpz, *_ = O.gradients(p, t, psiq, np.ones(p.shape))
    for ax, g in enumerate((gpx, gpy, gpz)):
        assert np.allclose(g, O.r2r3d(A.grad_shift(t, psiq, ax), O.IST[ax]), rtol=0, atol=1e-13)


def _state(p, rng, first):
    psi0, rho0 = O.make_interface_ic(p, O.InterfaceIC(undulation=0.6, seed=3, rho_iso=p.rho0))
    o = O.Oracle(p)
    o.set_state(psi0 + 0.05 * rng.standard_normal(p.shape), rho0)
    if not first:
        o.step(2)
    return A.snapshot(o)


def _perturb(X: A.State, V: dict, h: float) -> A.State:
    psiq = X.psiq + 1j * h * V["psiq"]
    return A.State(psiq, X.psiq_old + 1j * h * V["psiq_old"], X.rho + 1j * h * V["rho"],
                   X.f_rho + 1j * h * V["f_rho"],
                   None if X.nlq_old is None else X.nlq_old + 1j * h * V["nlq_old"],
                   O.r2r3d(psiq, O.ICT))


def _outputs(p, t, X, first):
    F = A.forward(p, t, X, first)
    return [F["psiq_new"], X.psiq, F["rho_new"], F["s_rho"], F["nlq"]]


@pytest.mark.parametrize("first", [True, False])
def test_step_vjp_dot_product_vs_complex_step(first):
    """<w, J v> = <J^T w, v> for one step; J v by complex step of the forward
    oracle.  Covers every state input and both parameters."""
    p = small_params()
    t = O.Tables(p)
    W = A.Weights(p.shape)
    rng = np.random.default_rng(1)
    X = _state(p, rng, first)
    names = ("psiq", "psiq_old", "rho", "f_rho", "nlq_old")
    # scale directions like the fields they perturb
    V = {k: rng.standard_normal(p.shape) * (np.abs(getattr(X, k)).max() if getattr(X, k) is not None else 1.0)
         for k in names}
    Wt = [rng.standard_normal(p.shape) for _ in range(5)]
    bar = A.Bar(*Wt)
    adj, dz, dn = A.step_vjp(p, t, W, X, first, bar)
    # state directions, one at a time so that a wrong term is named
    for k in names:
        if first and k == "nlq_old":
            assert np.all(adj.nlq_old == 0)
            continue
        Vk = {q: (V[q] if q == k else np.zeros(p.shape)) for q in names}
        out = _outputs(p, t, _perturb(X, Vk, H), first)
        jv = [np.imag(o) / H for o in out]
        lhs = sum(np.sum(w * j) for w, j in zip(Wt, jv))
        rhs = np.sum(getattr(adj, k) * V[k])
        assert np.isclose(lhs, rhs, rtol=1e-10, atol=1e-12 * abs(lhs)), (k, lhs, rhs)
    # parameters
    for name, got in (("zeta", dz), ("nu", dn)):
        pc = dataclasses.replace(p, **{name: getattr(p, name) + 1j * H})
        out = _outputs(pc, O.Tables(pc), X, first)
        lhs = sum(np.sum(w * np.imag(o) / H) for w, o in zip(Wt, out))
        assert np.isclose(lhs, got, rtol=1e-10), (name, lhs, got)


def _ic(p, seed=5):
    nx, ny = p.nx, p.ny
    x = np.arange(nx) * p.dx
    y = np.arange(ny) * p.dy
    h = p.nz * p.dz / 2 + 0.4 * np.cos(2 * np.pi * x / (nx * p.dx))[:, None] * np.cos(
        2 * np.pi * y / (ny * p.dy))[None, :]
    ic = O.InterfaceIC(undulation=0.0, seed=seed, rho_iso=p.rho0, rho_sm=1.0, height=h)
    return ic, h


@pytest.mark.parametrize("nsteps", [1, 5])
def test_loss_gradients_vs_complex_step(nsteps):
    p = small_params()
    ic, h = _ic(p)
    psi0, rho0 = O.make_interface_ic(p, ic)
    rng = np.random.default_rng(2)
    target = O.make_interface_ic(p, dataclasses.replace(ic, seed=9, height=h + 0.3))[0]
    g = A.loss_and_gradients(p, psi0, rho0, target, nsteps)
    assert np.isclose(g.loss, A.loss(p, psi0, rho0, target, nsteps), rtol=1e-14)
    # parameters
    for name, got in (("zeta", g.zeta), ("nu", g.nu)):
        pc = dataclasses.replace(p, **{name: getattr(p, name) + 1j * H})
        cs = np.imag(A.loss(pc, psi0, rho0, target, nsteps)) / H
        assert np.isclose(got, cs, rtol=1e-9), (name, got, cs)
    # initial state directions
    for arr, grad in ((0, g.psi0), (1, g.rho0)):
        v = rng.standard_normal(p.shape)
        x0 = [psi0.astype(complex), rho0.astype(complex)]
        x0[arr] = x0[arr] + 1j * H * v
        cs = np.imag(A.loss(p, x0[0], x0[1], target, nsteps)) / H
        assert np.isclose(np.sum(grad * v), cs, rtol=1e-9), (arr, np.sum(grad * v), cs)
    # interface height through the IC
    dh = A.interface_ic_vjp(p, ic, h, g.psi0, g.rho0)
    v = rng.standard_normal(h.shape)
    icc = dataclasses.replace(ic, height=h + 1j * H * v)
    pc_psi, pc_rho = O.make_interface_ic(p, icc)
    cs = np.imag(A.loss(p, pc_psi, pc_rho, target, nsteps)) / H
    assert np.isclose(np.sum(dh * v), cs, rtol=1e-9), (np.sum(dh * v), cs)


def test_loss_gradient_vs_central_differences():
    p = small_params()
    ic, h = _ic(p)
    psi0, rho0 = O.make_interface_ic(p, ic)
    target = O.make_interface_ic(p, dataclasses.replace(ic, seed=9, height=h - 0.2))[0]
    g = A.loss_and_gradients(p, psi0, rho0, target, 3)
    eps = 1e-6
    for name, got in (("zeta", g.zeta), ("nu", g.nu)):
        lp = A.loss(dataclasses.replace(p, **{name: getattr(p, name) + eps}), psi0, rho0, target, 3)
        lm = A.loss(dataclasses.replace(p, **{name: getattr(p, name) - eps}), psi0, rho0, target, 3)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - got) <= 1e-6 * max(abs(got), 1e-12 * g.loss), (name, fd, got)
