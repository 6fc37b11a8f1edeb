"""smaiso_adjoint wrapped in a torch.autograd.Function (CPU tensors in and
out; the GPU work happens inside libsmaiso), composed with a differentiable
torch restatement of the interface initial condition, so that autograd
carries dL/dpsi0 and dL/drho0 back to the interface height h(x, y).

torch is imported inside the tests, after libsmaiso has loaded its HIP
runtime: the torch wheel carries its own copy, and only CPU tensors are used
here."""
import dataclasses

import numpy as np
import pytest

import smaiso_amd as S

pytestmark = pytest.mark.gpu


def _torch():
    S.lib()
    return pytest.importorskip("torch")


def make_loss_fn(torch):
    class SmaisoLoss(torch.autograd.Function):
        """L(psi0, rho0, zeta, nu) = sum (psi(T) - target)^2; one smaiso_adjoint
        call gives L and all four gradients."""

        @staticmethod
        def forward(ctx, psi0, rho0, zeta, nu, params, target, nsteps, every):
            p = dataclasses.replace(params, zeta=float(zeta), nu=float(nu))
            with S.Stepper(p) as st:
                g = st.adjoint(psi0.detach().numpy(), rho0.detach().numpy(), target, nsteps, every)
            ctx.save_for_backward(torch.from_numpy(g.psi0), torch.from_numpy(g.rho0),
                                  torch.tensor(g.zeta, dtype=torch.float64),
                                  torch.tensor(g.nu, dtype=torch.float64))
            return torch.tensor(g.loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            dpsi, drho, dz, dn = ctx.saved_tensors
            return gout * dpsi, gout * drho, gout * dz, gout * dn, None, None, None, None

    return SmaisoLoss


def interface_ic_torch(torch, p: S.Params, ic: S.InterfaceIC, h, noise):
    """psi0 = A cos(q0 z) profile + noise, rho0 = rho_iso + (rho_sm - rho_iso)
    profile, profile = (1 - tanh((z - h)/w))/2, as smaiso_make_interface_ic."""
    z = torch.arange(p.nz, dtype=torch.float64) * p.dz
    prof = 0.5 * (1.0 - torch.tanh((z[None, None, :] - h[:, :, None]) / ic.width))
    psi0 = ic.amplitude * torch.cos(p.q0 * z)[None, None, :] * prof + noise
    rho0 = ic.rho_iso + (ic.rho_sm - ic.rho_iso) * prof
    return psi0, rho0


def _setup(n=(16, 16, 24)):
    p = S.Params(nx=n[0], ny=n[1], nz=n[2], dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=0.6, zeta=1.4)
    ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.8, rho_sm=1.0, noise=1e-3, seed=3)
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    h = p.nz * p.dz / 2 + 0.5 * np.cos(2 * np.pi * x / (p.nx * p.dx))[:, None] * np.cos(
        2 * np.pi * y / (p.ny * p.dy))[None, :]
    noise = S.make_interface_ic(p, dataclasses.replace(ic, amplitude=0.0), height=h)[0]
    target = S.make_interface_ic(p, dataclasses.replace(ic, seed=11), height=h + 0.7)[0]
    return p, ic, h, noise, target


def test_torch_ic_matches_library_ic():
    torch = _torch()
    p, ic, h, noise, _ = _setup()
    psi_t, rho_t = interface_ic_torch(torch, p, ic, torch.from_numpy(h), torch.from_numpy(noise))
    psi_c, rho_c = S.make_interface_ic(p, ic, height=h)
    assert np.allclose(psi_t.numpy(), psi_c, rtol=0, atol=1e-14)
    assert np.allclose(rho_t.numpy(), rho_c, rtol=0, atol=1e-14)


def test_autograd_gradients_equal_library_gradients():
    torch = _torch()
    p, ic, h, noise, target = _setup()
    Loss = make_loss_fn(torch)
    ht = torch.tensor(h, requires_grad=True)
    zt = torch.tensor(p.zeta, dtype=torch.float64, requires_grad=True)
    nt = torch.tensor(p.nu, dtype=torch.float64, requires_grad=True)
    psi0, rho0 = interface_ic_torch(torch, p, ic, ht, torch.from_numpy(noise))
    L = Loss.apply(psi0, rho0, zt, nt, p, target, 8, 3)
    L.backward()
    g = S.Stepper(p).adjoint(psi0.detach().numpy(), rho0.detach().numpy(), target, 8, 3)
    dh = S.interface_ic_vjp(p, ic, g.psi0, g.rho0, height=h)
    assert float(L.detach()) == g.loss
    assert float(zt.grad) == g.zeta and float(nt.grad) == g.nu
    assert np.allclose(ht.grad.numpy(), dh, rtol=1e-12, atol=1e-12 * np.abs(dh).max())


def test_autograd_gradcheck_height_zeta_nu():
    """torch.autograd.gradcheck (central differences of the whole GPU
    forward run) through h(x, y), zeta and nu."""
    torch = _torch()
    p, ic, h, noise, target = _setup((8, 8, 16))
    Loss = make_loss_fn(torch)
    nz = torch.from_numpy(noise)

    def f(ht, zt, nt):
        psi0, rho0 = interface_ic_torch(torch, p, ic, ht, nz)
        return Loss.apply(psi0, rho0, zt, nt, p, target, 5, 2)

    inputs = (torch.tensor(h, requires_grad=True),
              torch.tensor(p.zeta, dtype=torch.float64, requires_grad=True),
              torch.tensor(p.nu, dtype=torch.float64, requires_grad=True))
    assert torch.autograd.gradcheck(f, inputs, eps=1e-6, atol=1e-7, rtol=1e-5)
