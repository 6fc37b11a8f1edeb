"""Closed-form pins of the oracle's forces, density update, nonlinear term and
free-energy density (CPU only).

The four pointwise stages of the step, smaiso-wcomp.cpp:1385-1425 (forces,
chemical potential), :2121-2139 (density AB2), :2144-2162 (nonlinear term)
and :2227-2233 (free-energy summand), are checked through whole oracle steps
from states whose every intermediate field has a closed form:

* layers along z, psi = A cos(q0 z) with q0 on the DCT lattice and rho
  uniform: L psi = 0, so S = 0 and v = 0; amp = A q0; div f = -f_c q0^2 psi;
  mu and Nl are polynomials in psi;
* layers along x, psi = A cos(q0 x), with a single-mode density
  rho = rho_c + d cos(k x): L psi = 0 again, but fgrad has modes k and 2k, so
  the irrotational velocity v_x is a sum of sines with known amplitudes, the
  density source is -d(rho v_x)/dx (the continuity equation, expanded by
  trigonometric identities), and Nl carries v . grad psi and the coupling
  term with the updated density;
* mu = dE/dpsi / dV (the chemical potential is the variational derivative of
  the free energy) at alpha = q0 = 1, zeta = 0, uniform rho, by complex-step
  differentiation of the oracle's energy sum on a random state;
* the density AB2 step is exact for a source linear in time.

Each planted mistake in MUTANTS (one sign or term each) makes at least one
pin fail; test_planted_mistakes_fail checks that.
"""
import math

import numpy as np
import pytest

import smaiso_oracle as O

A = 1.3          # layer amplitude
RHO_C = 0.93     # mean density
DEN = 0.05       # density mode amplitude


def centres(n, d):
    return (np.arange(n) + 0.5) * d


# ---------------------------------------------------------------------------
# state 1: layers along z, uniform density
# ---------------------------------------------------------------------------
def layered_z_params(nx=4, ny=6):
    # Nz = 16, dz = pi/4: Vq_z[4] = 4 pi/(16 pi/4) = 1 = q0
    return O.Params(nx=nx, ny=ny, nz=16, dx=0.7, dy=0.9, dz=math.pi / 4, dt=0.01, rho0=0.8,
                    rho_m=1.0, divv_switch=0, beta=2.0, gamma=1.0, epsilon=-0.005, zeta=1.4,
                    nu=0.7, lam=1.1)


def layered_z_case(p):
    """psi = A cos(q0 z), rho = RHO_C: the state and the closed forms of one
    step's fields (real space; nlq spectral, unnormalised)."""
    z = centres(p.nz, p.dz)
    psi = A * np.cos(p.q0 * z)[None, None, :] * np.ones(p.shape)
    rho = np.full(p.shape, RHO_C)
    ampc = A * p.q0
    fc = p.kp * p.zeta * (RHO_C - p.rho0 - p.kp * ampc) / (ampc + 0.001)
    mu = RHO_C * (-p.epsilon * psi - p.beta * psi ** 3 + p.gamma * psi ** 5)
    nl = RHO_C * (p.epsilon * psi + p.beta * psi ** 3 - p.gamma * psi ** 5) + fc * (1 + p.q0 ** 2) * psi
    # first step (quirk Q5): psiq_old = psiq, Nl_old = Nl
    t = O.Tables(p)
    psiq = p.scale * O.r2r3d(psi, O.CT)
    nlq = O.r2r3d(nl, O.CT)
    psiq_new = (t.C1 * psiq - 0.5 * psiq + p.dt * p.scale * nlq) / t.C2
    zero = np.zeros(p.shape)
    want = dict(lpsi=zero, amp=np.full(p.shape, ampc), vx=zero, vy=zero, vz=zero, divv=zero,
                rho=rho, f_rho=zero, mu=mu, divf=-fc * p.q0 ** 2 * psi, nl=nl, nlq=nlq, psiq=psiq_new)
    return psi, rho, want


# ---------------------------------------------------------------------------
# state 2: layers along x, single-mode density along x
# ---------------------------------------------------------------------------
def layered_x_params(ny=4, nz=6):
    # Nx = 16, dx = pi/4: Vq_x[n] = n/4, q0 = Vq_x[4]; the density mode is
    # k = Vq_x[1], so k, 2k, 3k, q0 +- k are all on the lattice
    return O.Params(nx=16, ny=ny, nz=nz, dx=math.pi / 4, dy=0.9, dz=0.8, dt=0.01, rho0=0.8,
                    rho_m=1.0, divv_switch=0, beta=2.0, gamma=1.0, epsilon=-0.005, zeta=1.4,
                    nu=0.7, lam=1.1)


def layered_x_case(p):
    """psi = A cos(q0 x), rho = RHO_C + DEN cos(k x): the state and the closed
    forms of one step's fields."""
    q0 = p.q0
    k = math.pi / (p.dx * p.nx)  # Vq_x[1]
    x = centres(p.nx, p.dx)[:, None, None]
    one = np.ones(p.shape)
    C = lambda m: np.cos(m * k * x) * one
    Sn = lambda m: np.sin(m * k * x) * one
    psi = A * np.cos(q0 * x) * one
    dpsi = -A * q0 * np.sin(q0 * x) * one
    rho = RHO_C + DEN * C(1)
    ampc = A * q0
    R = p.rho0 + p.kp * ampc
    L = p.lam + 2 * p.nu
    # fgrad = zeta/2 (rho^2 - R^2): modes F1 cos(kx) and F2 cos(2kx) besides the mean
    F1 = p.zeta * RHO_C * DEN
    F2 = 0.25 * p.zeta * DEN ** 2
    # irrotational velocity of a cosine mode F cos(mkx): F sin(mkx)/(mk L)
    a, b = F1 / (k * L), F2 / (2 * k * L)
    vx = a * Sn(1) + b * Sn(2)
    divv = a * k * C(1) + 2 * b * k * C(2)
    # rho vx = (RHO_C a + DEN b/2) sin kx + (RHO_C b + DEN a/2) sin 2kx + (DEN b/2) sin 3kx
    s_rho = -(k * (RHO_C * a + 0.5 * DEN * b) * C(1) + 2 * k * (RHO_C * b + 0.5 * DEN * a) * C(2)
              + 3 * k * 0.5 * DEN * b * C(3))
    rho_new = rho + 1.5 * p.dt * s_rho  # first step: f_rho = 0
    # f = f_c(x) grad psi, f_c = c0 + c1 cos kx
    c0 = p.kp * p.zeta * (RHO_C - R) / (ampc + 0.001)
    c1 = p.kp * p.zeta * DEN / (ampc + 0.001)
    # f_x = -A q0 (c0 sin q0x + c1/2 (sin (q0+k)x + sin (q0-k)x)); div f = d f_x/dx
    divf = -A * q0 * (c0 * q0 * np.cos(q0 * x) + 0.5 * c1 * ((q0 + k) * np.cos((q0 + k) * x)
                                                            + (q0 - k) * np.cos((q0 - k) * x))) * one
    mu = rho * (-p.epsilon * psi - p.beta * psi ** 3 + p.gamma * psi ** 5)
    coup = p.kp * p.zeta / (ampc + 0.001) * (rho_new - p.rho0 - p.kp * ampc) * psi
    nl = -mu + coup - divf - vx * dpsi
    zero = np.zeros(p.shape)
    want = dict(lpsi=zero, amp=np.full(p.shape, ampc), vx=vx, vy=zero, vz=zero, divv=divv,
                rho=rho_new, f_rho=s_rho, mu=mu, divf=divf, nl=nl, nlq=O.r2r3d(nl, O.CT))
    return psi, rho, want


def case_errors(got, want):
    """Max-abs error per field relative to the field's max (absolute for a
    field that vanishes)."""
    out = {}
    for k, w in want.items():
        s = np.abs(w).max()
        out[k] = float(np.abs(got[k] - w).max() / (s if s > 0 else 1.0))
    return out


def oracle_case_errors(p, case):
    psi, rho, want = case(p)
    o = O.Oracle(p)
    o.set_state(psi, rho)
    o.step(1)
    x = o.f.extra
    got = dict(x, rho=o.f.rho, f_rho=o.f.f_rho, nlq=o.f.nlq, psiq=o.f.psiq)
    return case_errors(got, want)


def layered_z_errors():
    return oracle_case_errors(layered_z_params(), layered_z_case)


def layered_x_errors():
    return oracle_case_errors(layered_x_params(), layered_x_case)


# ---------------------------------------------------------------------------
# state 3: mu is the variational derivative of the free energy
# ---------------------------------------------------------------------------
def variational_error(seed=3):
    """max |mu - (dE/dpsi)/dV| / max|mu| on a random psi (uniform rho, zeta = 0,
    alpha = q0 = 1), dE/dpsi by complex steps along every unit vector."""
    p = O.Params(nx=4, ny=3, nz=5, dx=0.7, dy=0.8, dz=0.6, dt=0.01, rho0=0.8, rho_m=1.0,
                 divv_switch=0, zeta=0.0, alpha=1.0, q0=1.0, beta=2.0, gamma=1.0, epsilon=-0.3)
    t = O.Tables(p)
    rng = np.random.default_rng(seed)
    psi = 0.8 * rng.standard_normal(p.shape)
    rho = np.full(p.shape, RHO_C)
    dv = p.dx * p.dy * p.dz

    def energy(ps):
        lpsi = O.sh_operator(p, t, p.scale * O.r2r3d(ps, O.CT))
        amp = np.ones(p.shape)  # enters only with zeta = 0
        return np.sum(O.energy_density(p, ps, lpsi, rho, amp)) * dv

    psiq = p.scale * O.r2r3d(psi, O.CT)
    lpsi = O.sh_operator(p, t, psiq)
    lap_rl, _ = O.lap_rho_l(p, t, psiq, rho, lpsi, np.ones(p.shape), relax=False)
    z = np.zeros(p.shape)
    mu = O.forces(p, lpsi, lap_rl, rho, np.ones(p.shape), psi, z, z, z, z, z, z)[2]
    h = 1e-30
    grad = np.zeros(p.shape)
    for idx in np.ndindex(p.shape):
        e = np.zeros(p.shape, dtype=complex)
        e[idx] = 1j * h
        grad[idx] = energy(psi + e).imag / h
    return np.abs(mu - grad / dv).max() / np.abs(mu).max()


def layered_energy_error():
    """Energy sum of psi = A cos(q0 z), L psi = B cos(q0 z), uniform rho and
    amp: the discrete means of cos^2, cos^4, cos^6 over whole periods are
    1/2, 3/8, 5/16."""
    p = layered_z_params()
    z = centres(p.nz, p.dz)
    B = 0.37
    c = np.cos(p.q0 * z)[None, None, :] * np.ones(p.shape)
    psi, lpsi = A * c, B * c
    amp = np.full(p.shape, 0.6)
    rho = np.full(p.shape, RHO_C)
    got = np.sum(O.energy_density(p, psi, lpsi, rho, amp)) / psi.size
    want = (RHO_C * (-0.5 * p.epsilon * A ** 2 / 2 + 0.5 * B ** 2 / 2 - 0.25 * p.beta * A ** 4 * 3 / 8
                     + p.gamma / 6 * A ** 6 * 5 / 16)
            + 0.5 * p.zeta * (RHO_C - p.rho0 - p.kp * 0.6) ** 2)
    return abs(got - want) / abs(want)


def ab2_error():
    """rho_new - rho = integral of a source linear in time, s(t) = s0 + s1 t,
    over [t, t + dt] when f_rho holds s(t - dt)."""
    p = O.Params(nx=3, ny=3, nz=3, dt=0.013)
    rng = np.random.default_rng(5)
    rho = 0.9 + 0.05 * rng.standard_normal(p.shape)
    s_now = rng.standard_normal(p.shape)
    s1 = rng.standard_normal(p.shape)
    s_prev = s_now - s1 * p.dt
    # a source of s_now: v = 0, div v = -s_now / rho
    z = np.zeros(p.shape)
    rho_new, f = O.density_update(p, z, z, z, z, z, z, rho, -s_now / rho, s_prev)
    want = rho + p.dt * s_now + 0.5 * s1 * p.dt ** 2
    return max(np.abs(rho_new - want).max(), np.abs(f - s_now).max())


PIN_TOL = 1e-11


def all_pin_errors():
    out = {}
    for k, v in layered_z_errors().items():
        out["z:" + k] = v
    for k, v in layered_x_errors().items():
        out["x:" + k] = v
    out["variational"] = variational_error()
    out["energy"] = layered_energy_error()
    out["ab2"] = ab2_error()
    return out


def test_layered_z_closed_form():
    err = layered_z_errors()
    assert max(err.values()) < PIN_TOL, err


def test_layered_x_single_mode_density_closed_form():
    err = layered_x_errors()
    assert max(err.values()) < PIN_TOL, err


def test_mu_is_variational_derivative_of_energy():
    assert variational_error() < 1e-11


def test_layered_energy_discrete_means():
    assert layered_energy_error() < 1e-13


def test_density_ab2_exact_for_linear_source():
    assert ab2_error() < 1e-14


def test_layered_x_state_exercises_every_term():
    """The x-layered state has non-zero v . grad psi, rho div v, div f and
    coupling contributions, so a sign error in any of them shows."""
    p = layered_x_params()
    psi, rho, cf = layered_x_case(p)
    dpsi = -A * p.q0 * np.sin(p.q0 * centres(p.nx, p.dx))[:, None, None]
    assert np.abs(cf["vx"] * dpsi).max() > 1e-4
    assert np.abs(rho * cf["divv"]).max() > 1e-4
    assert np.abs(cf["divf"]).max() > 1e-3
    assert np.abs(cf["nl"] - (-cf["mu"] - cf["divf"] - cf["vx"] * dpsi)).max() > 1e-3


# ---------------------------------------------------------------------------
# planted mistakes: each must make some pin fail
# ---------------------------------------------------------------------------
_forces, _nonlinear, _density, _energy = O.forces, O.nonlinear, O.density_update, O.energy_density


def _forces_mut(sign_eps=1, sign_beta=1, sign_gamma=1):
    def f(p, lpsi, lap_rl, rho, amp, psi, gpx, gpy, gpz, rdx, rdy, rdz):
        r = list(_forces(p, lpsi, lap_rl, rho, amp, psi, gpx, gpy, gpz, rdx, rdy, rdz))
        poly = -p.epsilon * psi - p.beta * psi ** 3 + p.gamma * psi ** 5
        bad = (-sign_eps * p.epsilon * psi - sign_beta * p.beta * psi ** 3 + sign_gamma * p.gamma * psi ** 5)
        r[2] = r[2] + rho * (bad - poly)
        return tuple(r)
    return f


def _nl_mut(coupling=1, vgrad=1, divf=1):
    def f(p, mu, amp, rho_new, psi, dvf, vx, vy, vz, gpx, gpy, gpz):
        good = _nonlinear(p, mu, amp, rho_new, psi, dvf, vx, vy, vz, gpx, gpy, gpz)
        c = (p.kp * p.zeta / (amp + 0.001)) * (rho_new - p.rho0 - p.kp * amp) * psi
        vg = vx * gpx + vy * gpy + vz * gpz
        return good + (coupling - 1) * c - (vgrad - 1) * vg - (divf - 1) * dvf
    return f


def _density_plus_rho_divv(p, vx, vy, vz, rdx, rdy, rdz, rho, divv, f_rho_old):
    s = -vx * rdx - vy * rdy - vz * rdz + rho * divv
    return rho + p.dt * (1.5 * s - 0.5 * f_rho_old), s


def _density_vgrad_mut(p, vx, vy, vz, rdx, rdy, rdz, rho, divv, f_rho_old):
    s = vx * rdx + vy * rdy + vz * rdz - rho * divv
    return rho + p.dt * (1.5 * s - 0.5 * f_rho_old), s


def _density_ab_mut(p, vx, vy, vz, rdx, rdy, rdz, rho, divv, f_rho_old):
    s = -vx * rdx - vy * rdy - vz * rdz - rho * divv
    return rho + p.dt * (1.5 * s + 0.5 * f_rho_old), s


def _energy_mut(term):
    def f(p, psi, lpsi, rho, amp):
        good = _energy(p, psi, lpsi, rho, amp)
        if term == "beta":
            return good + 0.5 * p.beta * psi ** 4 * rho
        if term == "lpsi":
            return good - lpsi ** 2 * rho
        if term == "coupling":
            return good - p.zeta * (rho - p.rho0 - p.kp * amp) ** 2
        return good
    return f


MUTANTS = {
    "forces: +epsilon psi in mu": ("forces", _forces_mut(sign_eps=-1)),
    "forces: +beta psi^3 in mu": ("forces", _forces_mut(sign_beta=-1)),
    "forces: -gamma psi^5 in mu": ("forces", _forces_mut(sign_gamma=-1)),
    "nonlinear: coupling term negated": ("nonlinear", _nl_mut(coupling=-1)),
    "nonlinear: +v.grad psi": ("nonlinear", _nl_mut(vgrad=-1)),
    "nonlinear: +div f": ("nonlinear", _nl_mut(divf=-1)),
    "density: +rho div v": ("density_update", _density_plus_rho_divv),
    "density: +v.grad rho": ("density_update", _density_vgrad_mut),
    "density: +0.5 f_rho_old": ("density_update", _density_ab_mut),
    "energy: +beta/4 psi^4": ("energy_density", _energy_mut("beta")),
    "energy: -1/2 (L psi)^2": ("energy_density", _energy_mut("lpsi")),
    "energy: coupling negated": ("energy_density", _energy_mut("coupling")),
}


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistakes_fail(name, monkeypatch):
    attr, fn = MUTANTS[name]
    monkeypatch.setattr(O, attr, fn)
    err = all_pin_errors()
    worst = max(err, key=err.get)
    assert err[worst] > 1e-6, (name, err)
