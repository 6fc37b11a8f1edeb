"""Pins of the FP64 oracle (CPU only).

Each test checks the oracle against something it does not compute itself:
the FFTW manual's defining sums, closed forms of single modes, invariants,
a linear recurrence, exact summation, or committed fixtures.
"""
import math
import os

import numpy as np
import pytest

import smaiso_oracle as O

PLANS = {"CT": O.CT, "ICT": O.ICT, "STx": O.ST[0], "STy": O.ST[1], "STz": O.ST[2],
         "ISTx": O.IST[0], "ISTy": O.IST[1], "ISTz": O.IST[2]}


# --------------------------------------------------------------------------
# transforms (S:499-545)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 12, 16, 17])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_r2r_1d_equals_fftw_definition(n, kind):
    x = np.random.default_rng(n + 10 * kind).standard_normal(n)
    got = O.r2r_1d(x, kind, 0)
    want = O.direct_r2r_1d(x, kind)
    assert np.abs(got - want).max() < 1e-12 * max(1, np.abs(want).max())


def test_fftw_definition_known_values():
    """Hand-evaluated N=2 cases of the FFTW formulas: REDFT10 [1,0] ->
    [2, 2cos(pi/4)], REDFT01 [1,0] -> [1,1], RODFT10 [1,0] -> [2 sin(pi/4), 2],
    RODFT01 [0,1] -> [1,-1]."""
    r = math.sqrt(2.0)
    assert np.allclose(O.direct_r2r_1d(np.array([1.0, 0.0]), 0), [2, r], atol=1e-15)
    assert np.allclose(O.direct_r2r_1d(np.array([1.0, 0.0]), 1), [1, 1], atol=1e-15)
    assert np.allclose(O.direct_r2r_1d(np.array([1.0, 0.0]), 2), [r, 2], atol=1e-15)
    assert np.allclose(O.direct_r2r_1d(np.array([0.0, 1.0]), 3), [1, -1], atol=1e-15)


@pytest.mark.parametrize("shape", [(4, 5, 3), (6, 2, 7), (8, 8, 8)])
@pytest.mark.parametrize("plan", list(PLANS))
def test_r2r3d_equals_direct(shape, plan):
    x = np.random.default_rng(1).standard_normal(shape)
    got = O.r2r3d(x, PLANS[plan])
    want = O.direct_r2r3d(x, PLANS[plan])
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_golden_fixture():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "r2r_golden.npz"))
    keys = [k for k in g.files if k.startswith("in_")]
    assert len(keys) == 40
    for key in keys:
        name = key[3:]
        plan = name.split("@")[0]
        got = O.r2r3d(g[key], PLANS[plan])
        assert np.abs(got - g["out_" + name]).max() < 1e-12, name


@pytest.mark.parametrize("fwd,inv", [("CT", "ICT"), ("STx", "ISTx"), ("STy", "ISTy"), ("STz", "ISTz")])
def test_round_trip_is_8N3(fwd, inv):
    """FFTW: REDFT01(REDFT10(x)) = 2N x and RODFT01(RODFT10(x)) = 2N x per
    axis, so scale = 1/(8 NxNyNz) undoes a 3-D round trip (S:278, S:982)."""
    x = np.random.default_rng(2).standard_normal((6, 5, 4))
    y = O.r2r3d(O.r2r3d(x, PLANS[fwd]), PLANS[inv]) * (0.125 / x.size)
    assert np.abs(y - x).max() < 1e-13


# --------------------------------------------------------------------------
# single-mode closed forms of the stages
# --------------------------------------------------------------------------
def grid_params(nx=8, ny=6, nz=10, **kw):
    return O.Params(nx=nx, ny=ny, nz=nz, dx=0.7, dy=0.9, dz=0.8, **kw)


def cell_centres(p):
    """Real-space points of the DCT-II/DST-II representation: (n + 1/2) d."""
    x = (np.arange(p.nx) + 0.5) * p.dx
    y = (np.arange(p.ny) + 0.5) * p.dy
    z = (np.arange(p.nz) + 0.5) * p.dz
    return x[:, None, None], y[None, :, None], z[None, None, :]


def mode(p, a, b, c):
    vq, _ = O.wavenumbers(p)
    X, Y, Z = cell_centres(p)
    kx, ky, kz = vq[0][a], vq[1][b], vq[2][c]
    cx, cy, cz = np.cos(kx * X), np.cos(ky * Y), np.cos(kz * Z)
    sx, sy, sz = np.sin(kx * X), np.sin(ky * Y), np.sin(kz * Z)
    return (kx, ky, kz), (cx, cy, cz), (sx, sy, sz)


@pytest.mark.parametrize("abc", [(1, 2, 3), (3, 1, 7), (0, 4, 2), (7, 5, 9)])
def test_spectral_gradient_single_mode(abc):
    """grad cos(kx x)cos(ky y)cos(kz z) = -k sin(...) (S:1166-1250); the DCT
    index n of psi maps to DST index n-1 of the derivative, with a minus."""
    p = grid_params()
    t = O.Tables(p)
    (kx, ky, kz), (cx, cy, cz), (sx, sy, sz) = mode(p, *abc)
    psi = cx * cy * cz
    rho = 0.3 + 0.2 * psi
    psiq = p.scale * O.r2r3d(psi, O.CT)
    gpx, gpy, gpz, rdx, rdy, rdz = O.gradients(p, t, psiq, rho)
    assert np.abs(gpx - (-kx * sx * cy * cz)).max() < 1e-12
    assert np.abs(gpy - (-ky * cx * sy * cz)).max() < 1e-12
    assert np.abs(gpz - (-kz * cx * cy * sz)).max() < 1e-12
    assert np.abs(rdx - 0.2 * gpx).max() < 1e-12
    assert np.abs(rdy - 0.2 * gpy).max() < 1e-12
    assert np.abs(rdz - 0.2 * gpz).max() < 1e-12


def test_amplitude_of_layers_is_A_q0():
    """psi = A cos(q0 z) has |A| = sqrt(q0^2 psi^2 + psi_z^2) = A q0 exactly,
    and the Gaussian filter keeps a constant (S:1254-1289)."""
    p = O.Params(nx=4, ny=4, nz=16, dz=math.pi / 4)  # Vq_z[4] = 1 = q0
    t = O.Tables(p)
    _, _, Z = cell_centres(p)
    psi = 1.3 * np.cos(p.q0 * Z) * np.ones(p.shape)
    psiq = p.scale * O.r2r3d(psi, O.CT)
    gpx, gpy, gpz, *_ = O.gradients(p, t, psiq, np.zeros(p.shape))
    amp = O.amplitude(p, t, psi, gpx, gpy, gpz)
    assert np.abs(amp - 1.3 * p.q0).max() < 1e-12


def test_gaussian_filter_damps_each_mode_by_G():
    p = grid_params()
    t = O.Tables(p)
    (kx, ky, kz), (cx, cy, cz), _ = mode(p, 2, 1, 3)
    # sqrt(q0^2 psi^2 + 0) with psi >= 0 is psi itself: feed a positive field
    f = 2.0 + cx * cy * cz
    q2 = kx * kx + ky * ky + kz * kz
    want = 2.0 + np.exp(-1.57 * 1.57 * 2 * q2) * cx * cy * cz
    zero = np.zeros(p.shape)
    got = O.amplitude(p, t, f / p.q0, zero, zero, zero)
    assert np.abs(got - want).max() < 1e-12


def test_sh_operator_and_laplacian_eigenvalues():
    p = grid_params(alpha=1.3, q0=1.1)
    t = O.Tables(p)
    (kx, ky, kz), (cx, cy, cz), _ = mode(p, 3, 2, 5)
    psi = cx * cy * cz
    q2 = kx * kx + ky * ky + kz * kz
    psiq = p.scale * O.r2r3d(psi, O.CT)
    L = O.sh_operator(p, t, psiq)
    assert np.abs(L - 1.3 * (1.21 - q2) * psi).max() < 1e-12
    # steady branch: lap(rho L) with rho = r uniform is -q2 r L (S:1312-1342)
    lap, rho = O.lap_rho_l(p, t, psiq, np.full(p.shape, 0.7), L, None, relax=False)
    assert np.abs(lap - (-q2) * 0.7 * L).max() < 1e-12
    # relaxation branch: rho <- rho0 + (1-rho0)/Amp amp; lap = rho * lap(L) (S:1348-1378)
    amp = np.full(p.shape, 0.4)
    lap2, rho2 = O.lap_rho_l(p, t, psiq, None, L, amp, relax=True)
    r = p.rho0 + (1 - p.rho0) / p.amp * 0.4
    assert np.abs(rho2 - r).max() < 1e-15
    assert np.abs(lap2 - r * (-q2) * L).max() < 1e-12


def test_divergence_of_gradient_is_laplacian():
    """div(grad phi) = -|q|^2 phi through the forward DST transforms and the
    DST -> DCT shift of S:2042-2080 (plus sign), for both div v (already
    normalised input) and div f (raw transforms times scale)."""
    p = grid_params()
    t = O.Tables(p)
    (kx, ky, kz), (cx, cy, cz), (sx, sy, sz) = mode(p, 2, 3, 4)
    phi = cx * cy * cz
    q2 = kx * kx + ky * ky + kz * kz
    gx, gy, gz = -kx * sx * cy * cz, -ky * cx * sy * cz, -kz * cx * cy * sz
    fx, fy, fz = O.r2r3d(gx, O.ST[0]), O.r2r3d(gy, O.ST[1]), O.r2r3d(gz, O.ST[2])
    divv, divf = O.divergences(p, t, p.scale * fx, p.scale * fy, p.scale * fz, fx, fy, fz)
    assert np.abs(O.r2r3d(divv, O.ICT) - (-q2) * phi).max() < 1e-12
    assert np.abs(O.r2r3d(divf, O.ICT) - (-q2) * phi).max() < 1e-12


def _stokes_inputs(p, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(p.shape) for _ in range(4)]


def test_stokes_solenoidal_part_is_orthogonal_to_q():
    """On the staggered lattice, component a of physical mode M=(I,J,K) lives
    at index M - e_a.  vsol = scale/(nu |q|^2) (S - q (q.S)/|q|^2) must satisfy
    q(M).vsol(M) = 0 wherever all three components are computed (S:1583-1974)."""
    p = grid_params(nx=7, ny=6, nz=9, nu=0.7)
    t = O.Tables(p)
    Sx, Sy, Sz, fg = _stokes_inputs(p, 3)
    vsx, vix, vsy, viy, vsz, viz = O.velocity_solve(p, t, Sx, Sy, Sz, fg)
    vq = t.vq
    worst, scale = 0.0, 0.0
    for I in range(1, p.nx):
        for J in range(1, p.ny - 1):
            for K in range(1, p.nz - 1):
                q = np.array([vq[0][I], vq[1][J], vq[2][K]])
                v = np.array([vsx[I - 1, J, K], vsy[I, J - 1, K], vsz[I, J, K - 1]])
                S = np.array([Sx[I - 1, J, K], Sy[I, J - 1, K], Sz[I, J, K - 1]])
                want = p.scale / (p.nu * q @ q) * (S - q * (q @ S) / (q @ q))
                assert np.allclose(v, want, rtol=1e-12, atol=1e-300)
                worst = max(worst, abs(q @ v))
                scale = max(scale, np.linalg.norm(q) * np.linalg.norm(v))
    assert worst < 1e-13 * scale


def test_stokes_irrotational_part_is_parallel_to_q_and_quirk_q3():
    """virr(M) = scale/((lam+2nu)|q|^2) q (fgrad(M) + (q.S)/|q|^2 G(|q|^2)), except
    that virrz at DST index k=0, j>=1 takes G from the y lattice (quirk Q3,
    S:1932) on every plane but the first of a slab."""
    for slab in (0, 3):
        p = grid_params(nx=7, ny=6, nz=9, nu=0.7, lam=0.4, slab_n0=slab)
        t = O.Tables(p)
        Sx, Sy, Sz, fg = _stokes_inputs(p, 4)
        vsx, vix, vsy, viy, vsz, viz = O.velocity_solve(p, t, Sx, Sy, Sz, fg)
        vq, vs = t.vq, t.vs
        for I in range(1, p.nx):
            for J in range(1, p.ny - 1):
                for K in range(1, p.nz - 1):
                    q = np.array([vq[0][I], vq[1][J], vq[2][K]])
                    q2 = q @ q
                    S = np.array([Sx[I - 1, J, K], Sy[I, J - 1, K], Sz[I, J, K - 1]])
                    G = math.exp(-1.57 * 1.57 * 2 * q2)
                    want = p.scale / ((p.lam + 2 * p.nu) * q2) * q * (fg[I, J, K] + (q @ S) / q2 * G)
                    got = np.array([vix[I - 1, J, K], viy[I, J - 1, K], viz[I, J, K - 1]])
                    if K == 1 and J >= 1 and (I % (slab or p.nx)) != 0:
                        # z component at k=0 uses the y-lattice |q|^2 = Vq_x^2 + Vs_y[J]^2 + Vq_z[0]^2
                        qy2 = vq[0][I] ** 2 + vs[1][J] ** 2 + vq[2][0] ** 2
                        Gq = math.exp(-1.57 * 1.57 * 2 * qy2)
                        want[2] = p.scale / ((p.lam + 2 * p.nu) * q2) * q[2] * (fg[I, J, K] + (q @ S) / q2 * Gq)
                    assert np.allclose(got, want, rtol=1e-12, atol=1e-300), (slab, I, J, K)


def test_velocity_truncations_q1_q2():
    p = grid_params(nx=5, ny=4, nz=6)
    t = O.Tables(p)
    vsx, vix, vsy, viy, vsz, viz = O.velocity_solve(p, t, *_stokes_inputs(p, 5))
    assert not vsx[-1].any() and not vix[-1].any()                     # Q2
    for f in (vsy, viy, vsz, viz):                                      # Q1
        assert not f[:, -1, :].any() and not f[:, :, -1].any()
    # and nonzero elsewhere (away from modes parallel to one axis, where vsol = 0)
    assert vix[:-1].all() and viy[:, :-1, :-1].all() and viz[:, :-1, :-1].all()
    assert vsx[:-1, 1:, 1:].all() and vsy[1:, :-1, 1:-1].all() and vsz[1:, 1:-1, :-1].all()


# --------------------------------------------------------------------------
# whole steps
# --------------------------------------------------------------------------
def test_uniform_state_is_stationary():
    p = O.Params(nx=8, ny=8, nz=8, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    o = O.Oracle(p)
    o.set_state(np.zeros(p.shape), np.full(p.shape, 0.9))
    o.step(3)
    assert not o.f.psi.any()
    assert np.abs(o.f.rho - 0.9).max() < 1e-15


def test_linear_recurrence_of_psi_update():
    """With rho = rho0 uniform and |psi| ~ 1e-10, every explicit term but the
    linear -mu = rho0 aLin psi^ is O(1e3 |psi|^2), so each mode follows
    psi1 = (C1 - 1/2 + dt rho0 aLin) psi0 / C2 (first step: psiq_old = psiq,
    Nl_old = Nl, quirk Q5) and then
    psi_{n+1} = (C1 psi_n - psi_{n-1}/2 + dt/2 rho0 aLin (3 psi_n - psi_{n-1})) / C2
    (S:2177-2190 with C1 = 2 - dt rho_m aLin, C2 = 1.5 - dt rho_m aLin)."""
    p = O.Params(nx=8, ny=6, nz=10, dt=0.003, rho0=0.85, rho_m=0.95, divv_switch=0)
    t = O.Tables(p)
    rng = np.random.default_rng(6)
    psi0 = 1e-10 * rng.standard_normal(p.shape)
    o = O.Oracle(p)
    o.set_state(psi0, np.full(p.shape, p.rho0))
    a = t.aLin
    c1 = 2 - p.dt * p.rho_m * a
    c2 = 1.5 - p.dt * p.rho_m * a
    lin = p.dt / 2 * p.rho0 * a
    prev = o.f.psiq.copy()
    cur = (c1 * prev - 0.5 * prev + lin * 2 * prev) / c2
    o.step(1)
    assert np.abs(o.f.psiq - cur).max() < 1e-6 * np.abs(cur).max()
    for _ in range(3):
        nxt = (c1 * cur - 0.5 * prev + lin * (3 * cur - prev)) / c2
        o.step(1)
        assert np.abs(o.f.psiq - nxt).max() < 1e-6 * np.abs(nxt).max()
        prev, cur = cur, nxt


def test_dt_0p01_with_half_rho_s_is_unstable_and_rho_m_1_is_not():
    """Stability of the stabilised scheme: per mode the growth factor root of
    (1.5 - s l) z^2 + (-2 + s l - 1.5 r l) z + (0.5 + 0.5 r l), l = dt aLin,
    s = rho_m, r = rho.  At dt = 0.01, dx = pi/4 the top mode is unstable for
    rho_m = rho_s/2 and stable for rho_m = 1 when rho <= 1."""
    p = O.Params(nx=16, ny=16, nz=16, dt=0.01, rho0=0.9, divv_switch=0)
    l = p.dt * O.Tables(p).aLin.min()
    for s, r, unstable in ((p.rho_s / 2, 1.0, True), (1.0, 1.0, False), (1.0, 0.9, False)):
        z = np.roots([1.5 - s * l, -2 + s * l - 1.5 * r * l, 0.5 + 0.5 * r * l])
        assert (np.abs(z).max() > 1) == unstable, (s, r, z)


def test_diagnostics_equal_exact_sums():
    p = O.Params(nx=6, ny=5, nz=8, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0)
    psi, rho = O.make_interface_ic(p, O.InterfaceIC(seed=3, undulation=0.5))
    o = O.Oracle(p)
    o.set_state(psi, rho)
    o.step(2)
    l1, mass, en = o.diagnostics()
    f, x = o.f, o.f.extra
    vol = p.dx * p.dy * p.dz
    sa = math.fsum(np.abs(f.psi - x["psi_old"]).ravel())
    sb = math.fsum(np.abs(f.psi).ravel())
    e = (-0.5 * p.epsilon * f.psi ** 2 + 0.5 * x["lpsi"] ** 2 - 0.25 * p.beta * f.psi ** 4
         + p.gamma / 6 * f.psi ** 6)
    es = math.fsum((f.rho * e + 0.5 * p.zeta * (f.rho - p.rho0 - p.kp * x["amp"]) ** 2).ravel())
    assert math.isclose(l1, sa / (p.dt * sb), rel_tol=1e-13)
    assert math.isclose(mass, math.fsum(f.rho.ravel()) * vol, rel_tol=1e-13)
    assert math.isclose(en, es * vol, rel_tol=1e-13)


def test_reference_constants():
    """Derived parameters of S:250-255 at the defaults: Amp = sqrt((6 +
    sqrt(36 - 0.2))/5), rho_s = 1, rho_m = 1/2, scale = 1/(8N^3)."""
    p = O.Params(nx=4, ny=4, nz=4)
    assert math.isclose(p.amp, math.sqrt((6 + math.sqrt(35.8)) / 5), rel_tol=1e-15)
    assert math.isclose(p.rho_s, 1.0, rel_tol=1e-15)
    assert math.isclose(p.rho_m_eff, 0.5, rel_tol=1e-15)
    assert p.scale == 0.125 / 64


def test_mid_slice_layout():
    f = np.arange(3 * 4 * 5, dtype=float).reshape(3, 4, 5)
    s = O.mid_slice(f, 2)
    # element (i, k) of the slice is field index (i*Ny + j)*Nz + k
    for i in range(3):
        for k in range(5):
            assert s[i * 5 + k] == (i * 4 + 2) * 5 + k


def test_surface_track_closed_form_and_stale_columns():
    """For the tanh interface rho = rho_iso + (1 - rho_iso)/2 (1 - tanh((z - h)/w))
    the tracked height is the largest k with k dz < h + w atanh(1 - 2 c),
    c = (thr - rho_iso)/(1 - rho_iso); columns never above thr keep prev."""
    p = O.Params(nx=6, ny=5, nz=40, dx=0.5, dy=0.5, dz=0.5)
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    h = 10.0 + 2.3 * np.cos(2 * np.pi * x / (p.nx * p.dx))[:, None] * np.sin(2 * np.pi * y / (p.ny * p.dy))[None, :]
    ic = O.InterfaceIC(width=1.7, rho_iso=0.8, rho_sm=1.0, noise=0.0, height=h)
    _, rho = O.make_interface_ic(p, ic)
    thr = 0.85
    c = (thr - 0.8) / 0.2
    zstar = h + ic.width * np.arctanh(1 - 2 * c)
    want = np.floor(zstar / p.dz - 1e-12).astype(float).reshape(-1)  # k dz < z*
    prev = np.full(p.nx * p.ny, -7.0)
    got = O.surface_track(rho, thr, prev)
    assert np.array_equal(got, want)
    # a column with no value above the threshold keeps its previous entry
    rho2 = rho.copy()
    rho2[2, 3, :] = 0.5
    got2 = O.surface_track(rho2, thr, prev)
    assert got2[2 * p.ny + 3] == -7.0
    assert np.array_equal(np.delete(got2, 2 * p.ny + 3), np.delete(want, 2 * p.ny + 3))
