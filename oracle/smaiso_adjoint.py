"""FP64 CPU oracle of the reverse-mode derivative (adjoint) of the time step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module; the
product path (``smaiso_adjoint`` in ``libsmaiso.so``) never calls into it.

The reference has no adjoint: this is the transpose, written by hand, of the
oracle's step (``smaiso_oracle.Oracle._step``, which restates
smaiso-wcomp.cpp:1081-2198) on its steady branch (``nLoop > divvSwitch`` or
load mode, S:1312), including the first-step quirk Q5 (``Nl_old = N^l``,
S:2170-2172) and the velocity quirks Q1-Q3.

Loss: L = sum (psi(T) - psi_target)^2 over the grid after T steps from
``set_state(psi0, rho0)``.  Returned gradients: dL/dpsi0, dL/drho0 (the arrays
given to set_state), dL/dzeta, dL/dnu, and, through the interface initial
condition, dL/dh for the height map h(x, y).

Transposes of FFTW's unnormalised r2r kinds (1-D, length N):
  REDFT10^T = REDFT01 . diag(2, 1, ..., 1)     RODFT10^T = RODFT01 . diag(1, ..., 1, 2)
so for the 3-D plans CT^T y = ICT(wc * y) and ST_a^T y = IST_a(ws_a * y), and
ICT^T x = CT(x) / wc, IST_a^T x = ST_a(x) / ws_a, with wc = 2^([i=0]+[j=0]+[k=0])
and ws_a the same with the DST axis weighted at index N_a - 1 instead of 0.

Parity pins (tests/test_oracle_adjoint.py): the transposes against dense
matrices; dot-product identities <w, J v> = <J^T w, v> for one step (first and
later steps, every state input and the parameters zeta and nu) where J v is
the complex-step tangent of the forward oracle (exact to rounding); the
multi-step loss gradient against complex-step derivatives in zeta, nu and
random directions of psi0, rho0 and h; and central finite differences.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import smaiso_oracle as O
from smaiso_oracle import CT, ICT, IST, ST, r2r3d

# ---------------------------------------------------------------------------
# transposed transforms
# ---------------------------------------------------------------------------


def _w1(n: int, at: int) -> np.ndarray:
    w = np.ones(n)
    w[at] = 2.0
    return w


class Weights:
    """wc and ws_a of the module docstring, broadcast to the grid."""

    def __init__(self, shape):
        nx, ny, nz = shape
        c = [_w1(nx, 0), _w1(ny, 0), _w1(nz, 0)]
        s = [_w1(nx, nx - 1), _w1(ny, ny - 1), _w1(nz, nz - 1)]
        self.wc = c[0][:, None, None] * c[1][None, :, None] * c[2][None, None, :]
        self.ws = []
        for a in range(3):
            f = [s[b] if b == a else c[b] for b in range(3)]
            self.ws.append(f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :])


def ct_T(W: Weights, y):
    return r2r3d(W.wc * y, ICT)


def ict_T(W: Weights, x):
    return r2r3d(x, CT) / W.wc


def st_T(W: Weights, y, a):
    return r2r3d(W.ws[a] * y, IST[a])


def ist_T(W: Weights, x, a):
    return r2r3d(x, ST[a]) / W.ws[a]


# ---------------------------------------------------------------------------
# index shifts and their transposes
# ---------------------------------------------------------------------------


def _sl(ax, s):
    out = [slice(None)] * 3
    out[ax] = s
    return tuple(out)


def shift_p(a, ax):
    """out[n] = a[n+1] along ax, 0 at n = N-1: the transpose of O._shift_m."""
    out = np.zeros_like(a)
    out[_sl(ax, slice(0, -1))] = a[_sl(ax, slice(1, None))]
    return out


def grad_shift(t: O.Tables, x, ax):
    """Sh_a of the spectral gradient (S:1166-1221): out[n] = -vs_a[n] x[n+e_a]
    for n_a <= N_a-2, 0 on the last plane."""
    vs = O._b(t.vs[ax], ax)
    out = np.zeros_like(x)
    out[_sl(ax, slice(0, -1))] = -(vs * shift_p(x, ax))[_sl(ax, slice(0, -1))]
    return out


def grad_shift_T(t: O.Tables, y, ax):
    """Transpose of grad_shift: out[m] = -vs_a[m-1] y[m-e_a] for m_a >= 1."""
    vs = O._b(t.vs[ax], ax)
    return O._shift_m(-vs * y, ax)


# ---------------------------------------------------------------------------
# forward intermediates of one step
# ---------------------------------------------------------------------------


@dataclass
class State:
    """Inputs of one step: psiq, psiq_old, rho, f_rho, the previous step's
    N^l (None before the first step) and psi = ICT(psiq)."""
    psiq: np.ndarray
    psiq_old: np.ndarray
    rho: np.ndarray
    f_rho: np.ndarray
    nlq_old: np.ndarray | None
    psi: np.ndarray


def snapshot(o: O.Oracle) -> State:
    f = o.f
    return State(f.psiq.copy(), f.psiq_old.copy(), f.rho.copy(), f.f_rho.copy(),
                 None if f.nlq is None else f.nlq.copy(), f.psi.copy())


def forward(p: O.Params, t: O.Tables, X: State, first: bool) -> dict:
    """One steady step from X, keeping every intermediate the transpose reads
    (the same stage functions as Oracle._step)."""
    psiq, rho, psi = X.psiq, X.rho, X.psi
    gp = list(O.gradients(p, t, psiq, rho))
    gp, rd = gp[:3], gp[3:]
    a0 = np.sqrt(p.q0 * p.q0 * psi ** 2 + gp[0] ** 2 + gp[1] ** 2 + gp[2] ** 2)
    amp = O.amplitude(p, t, psi, *gp)
    lpsi = O.sh_operator(p, t, psiq)
    lap_rl, _ = O.lap_rho_l(p, t, psiq, rho, lpsi, amp, relax=False)
    fgrad, erho, mu, Sx, Sy, Sz, fx, fy, fz = O.forces(p, lpsi, lap_rl, rho, amp, psi, *gp, *rd)
    S = (Sx, Sy, Sz)
    fv = (fx, fy, fz)
    fgq = r2r3d(fgrad, CT) + r2r3d(erho, CT) * t.G
    Sq = [r2r3d(S[a], ST[a]) for a in range(3)]
    fq = [r2r3d(fv[a], ST[a]) for a in range(3)]
    vsx, vix, vsy, viy, vsz, viz = O.velocity_solve(p, t, *Sq, fgq)
    vsol, virr = [vsx, vsy, vsz], [vix, viy, viz]
    divv_q, divf_q = O.divergences(p, t, *virr, *fq)
    divv = r2r3d(divv_q, ICT)
    divf = r2r3d(divf_q, ICT)
    v = [r2r3d(vsol[a] + virr[a], IST[a]) for a in range(3)]
    rho_new, s_rho = O.density_update(p, *v, *rd, rho, divv, X.f_rho)
    nl = O.nonlinear(p, mu, amp, rho_new, psi, divf, *v, *gp)
    nlq = r2r3d(nl, CT)
    no = nlq if (first or X.nlq_old is None) else X.nlq_old
    psiq_new = O.psi_update(p, t, psiq, X.psiq_old, nlq, no)
    return dict(gp=gp, rd=rd, a0=a0, amp=amp, lpsi=lpsi, lap_rl=lap_rl, vsol=vsol, virr=virr,
                divv=divv, v=v, rho_new=rho_new, s_rho=s_rho, nlq=nlq, psiq_new=psiq_new)


def next_state(p: O.Params, X: State, F: dict) -> State:
    return State(F["psiq_new"], X.psiq, F["rho_new"], F["s_rho"], F["nlq"],
                 r2r3d(F["psiq_new"], ICT))


# ---------------------------------------------------------------------------
# transpose of the velocity solve (S:1583-1974)
# ---------------------------------------------------------------------------


def velocity_solve_T(p: O.Params, t: O.Tables, bvsol, bvirr, vsol, virr):
    """Given the adjoints of (vsol_a, virr_a) return those of (Sq_a, fgq) and
    dL/dnu (vsol scales with 1/nu, virr with 1/(lambda + 2 nu))."""
    nx, ny, nz = p.shape
    sc = p.scale
    qx, qy, qz = (O._b(t.vq[a], a) for a in range(3))
    sx, sy, sz = (O._b(t.vs[a], a) for a in range(3))
    nu, lam = p.nu, p.lam
    gfac = -1.57 * 1.57 * 2 * sc
    bvsol = [b.copy() for b in bvsol]
    bvirr = [b.copy() for b in bvirr]
    # output masks (Q1, Q2)
    bvsol[0][-1] = 0; bvirr[0][-1] = 0
    for a in (1, 2):
        bvsol[a][:, -1, :] = 0; bvsol[a][:, :, -1] = 0
        bvirr[a][:, -1, :] = 0; bvirr[a][:, :, -1] = 0
    dnu = 0.0
    for a in range(3):
        dnu += np.sum(bvsol[a] * (-vsol[a] / nu)) + np.sum(bvirr[a] * (-2.0 * virr[a] / (lam + 2 * nu)))
    bS = [np.zeros(p.shape, dtype=bvsol[0].dtype) for _ in range(3)]
    bfg = np.zeros(p.shape, dtype=bvsol[0].dtype)

    # x component
    e = np.exp(gfac / t.CM1x)
    bdot = -(t.CM1x / nu) * t.CM2x * bvsol[0] + (t.CM1x / (lam + 2 * nu)) * t.CM2x * e * bvirr[0]
    bS[0] += (t.CM1x / nu) * bvsol[0] + sx * bdot
    w = qy * bdot; bS[1][1:, :-1, :] += w[:-1, 1:, :]
    w = qz * bdot; bS[2][1:, :, :-1] += w[:-1, :, 1:]
    w = (t.CM1x / (lam + 2 * nu)) * sx * bvirr[0]; bfg[1:] += w[:-1]

    # y component
    e = np.exp(gfac / t.CM1y)
    bdot = -(t.CM1y / nu) * t.CM2y * bvsol[1] + (t.CM1y / (lam + 2 * nu)) * t.CM2y * e * bvirr[1]
    bS[1] += (t.CM1y / nu) * bvsol[1] + sy * bdot
    w = qx * bdot; bS[0][:-1, 1:, :] += w[1:, :-1, :]
    w = qz * bdot; bS[2][:, 1:, :-1] += w[:, :-1, 1:]
    w = (t.CM1y / (lam + 2 * nu)) * sy * bvirr[1]; bfg[:, 1:] += w[:, :-1]

    # z component (Q3 exponent)
    cm_exp = t.CM1z.copy()
    slab = p.slab_n0 if p.slab_n0 > 0 else nx
    q3 = (np.arange(nx) % slab) != 0
    cm_exp[q3, 1:, 0] = t.CM1y[q3, 1:, 0]
    e = np.exp(gfac / cm_exp)
    bdot = -(t.CM1z / nu) * t.CM2z * bvsol[2] + (t.CM1z / (lam + 2 * nu)) * t.CM2z * e * bvirr[2]
    bS[2] += (t.CM1z / nu) * bvsol[2] + sz * bdot
    w = qx * bdot; bS[0][:-1, :, 1:] += w[1:, :, :-1]
    w = qy * bdot; bS[1][:, :-1, 1:] += w[:, 1:, :-1]
    w = (t.CM1z / (lam + 2 * nu)) * sz * bvirr[2]; bfg[:, :, 1:] += w[:, :, :-1]
    return bS, bfg, dnu


# ---------------------------------------------------------------------------
# transpose of one step
# ---------------------------------------------------------------------------


@dataclass
class Bar:
    """Adjoints of the step's state: psiq, psiq_old, rho, f_rho, nlq_old."""
    psiq: np.ndarray
    psiq_old: np.ndarray
    rho: np.ndarray
    f_rho: np.ndarray
    nlq_old: np.ndarray


def step_vjp(p: O.Params, t: O.Tables, W: Weights, X: State, first: bool, bar: Bar, F: dict | None = None):
    """Adjoint of one step: the adjoints of the outputs (psiq_new, psiq_old_new
    = psiq, rho_new, f_rho_new = s_rho, nlq_old_new = nlq) -> those of the
    inputs, plus the step's dL/dzeta and dL/dnu."""
    if F is None:
        F = forward(p, t, X, first)
    sc, dt = p.scale, p.dt
    kp, zeta, rho0 = p.kp, p.zeta, p.rho0
    psi, rho = X.psi, X.rho
    gp, rd, v = F["gp"], F["rd"], F["v"]
    amp, lpsi, lap_rl = F["amp"], F["lpsi"], F["lap_rl"]

    # a17 (S:2177-2190)
    u = bar.psiq / t.C2
    bP = t.C1 * u + bar.psiq_old
    bPo = -0.5 * u
    bnlq = 1.5 * dt * sc * u + bar.nlq_old
    bno = -0.5 * dt * sc * u
    if first:
        bnlq = bnlq + bno
        bNo = np.zeros_like(bno)
    else:
        bNo = bno
    bnl = ct_T(W, bnlq)

    # a16 (S:2144-2162)
    ia = 1.0 / (amp + 0.001)
    c = kp * zeta * ia
    D = F["rho_new"] - rho0 - kp * amp
    bmu = -bnl
    bRn = bar.rho + bnl * c * psi
    bamp = bnl * psi * (-kp * zeta * D * ia * ia - kp * c)
    bpsi = bnl * c * D
    dzeta = np.sum(bnl * kp * ia * D * psi)
    bdivf = -bnl
    bv = [-bnl * gp[a] for a in range(3)]
    bgp = [-bnl * v[a] for a in range(3)]

    # a15 (S:2121-2139)
    bs = 1.5 * dt * bRn + bar.f_rho
    bR = bRn.copy()
    bFr = -0.5 * dt * bRn
    for a in range(3):
        bv[a] = bv[a] - bs * rd[a]
    brd = [-bs * v[a] for a in range(3)]
    bR = bR - bs * F["divv"]
    bdivv = -bs * rho

    # a14 inverse transforms and a13 divergences (S:1979-2117)
    bdivv_q = ict_T(W, bdivv)
    bdivf_q = ict_T(W, bdivf)
    qv = [O._b(t.vq[a], a) for a in range(3)]
    bvsol = [ist_T(W, bv[a], a) for a in range(3)]
    bvirr = [bvsol[a] + shift_p(qv[a] * bdivv_q, a) for a in range(3)]
    bfq = [sc * shift_p(qv[a] * bdivf_q, a) for a in range(3)]

    # a12 velocity solve
    bSq, bfgq, dnu = velocity_solve_T(p, t, bvsol, bvirr, F["vsol"], F["virr"])

    # a11 forward transforms (S:1427-1488)
    bfgrad = ct_T(W, bfgq)
    berho = ct_T(W, t.G * bfgq)
    bS = [st_T(W, bSq[a], a) for a in range(3)]
    bf = [st_T(W, bfq[a], a) for a in range(3)]

    # a10 stresses, forces, chemical potential (S:1385-1425)
    d0 = rho - rho0 - kp * amp
    fc = kp * zeta * d0 * ia
    bfc = bf[0] * gp[0] + bf[1] * gp[1] + bf[2] * gp[2]
    for a in range(3):
        bgp[a] = bgp[a] + fc * bf[a]
    bR = bR + bfc * kp * zeta * ia
    bamp = bamp + bfc * kp * zeta * (-kp * ia - d0 * ia * ia)
    dzeta += np.sum(bfc * kp * d0 * ia)
    e = 0.5 * lpsi ** 2
    bR = bR + berho * e
    be = berho * rho
    poly = lpsi - p.epsilon * psi - p.beta * psi ** 3 + p.gamma * psi ** 5
    blap = bmu.copy()
    bR = bR + bmu * poly
    blpsi = bmu * rho
    bpsi = bpsi + bmu * rho * (-p.epsilon - 3 * p.beta * psi ** 2 + 5 * p.gamma * psi ** 4)
    mu1 = lap_rl + rho * lpsi
    bmu1 = bS[0] * gp[0] + bS[1] * gp[1] + bS[2] * gp[2]
    for a in range(3):
        bgp[a] = bgp[a] + bS[a] * mu1
        be = be + bS[a] * rd[a]
        brd[a] = brd[a] + bS[a] * e
    blap = blap + bmu1
    bR = bR + bmu1 * lpsi
    blpsi = blpsi + bmu1 * rho
    rs = rho0 + kp * amp
    bR = bR + bfgrad * zeta * rho
    bamp = bamp - bfgrad * zeta * rs * kp
    dzeta += np.sum(bfgrad * 0.5 * (rho ** 2 - rs ** 2))
    blpsi = blpsi + be * lpsi

    # a9 (S:1312-1380): lap_rl = ICT(-scale |q|^2 CT(rho lpsi))
    by = ct_T(W, -sc * t.mq2c * ict_T(W, blap))
    bR = bR + by * lpsi
    blpsi = blpsi + by * rho

    # a8 (S:1293-1308): lpsi = ICT(alpha (q0^2 - |q|^2) psiq)
    bP = bP + p.alpha * (p.q0 * p.q0 - t.mq2c) * ict_T(W, blpsi)

    # a7 (S:1254-1289): amp = ICT(scale G CT(a0))
    ba0 = ct_T(W, sc * t.G * ict_T(W, bamp))
    a0 = F["a0"]
    inv = np.where(a0 > 0, 1.0 / np.where(a0 > 0, a0, 1.0), 0.0)
    bpsi = bpsi + ba0 * p.q0 * p.q0 * psi * inv
    for a in range(3):
        bgp[a] = bgp[a] + ba0 * gp[a] * inv

    # a6 (S:1104-1250): gp_a = IST_a(Sh_a psiq), rd_a = IST_a(scale Sh_a CT(rho))
    for a in range(3):
        bP = bP + grad_shift_T(t, ist_T(W, bgp[a], a), a)
    bRq = sum(grad_shift_T(t, sc * ist_T(W, brd[a], a), a) for a in range(3))
    bR = bR + ct_T(W, bRq)

    # psi = ICT(psiq)
    bP = bP + ict_T(W, bpsi)
    return Bar(bP, bPo, bR, bFr, bNo), float(np.real(dzeta)), float(np.real(dnu))


# ---------------------------------------------------------------------------
# loss and gradients over T steps
# ---------------------------------------------------------------------------


@dataclass
class Gradients:
    loss: float
    psi0: np.ndarray
    rho0: np.ndarray
    zeta: float
    nu: float


def loss(p: O.Params, psi0, rho0, target, nsteps: int):
    o = O.Oracle(p)
    o.set_state(psi0, rho0)
    o.step(nsteps)
    r = o.f.psi - target
    return np.sum(r * r)


def loss_and_gradients(p: O.Params, psi0, rho0, target, nsteps: int) -> Gradients:
    """L and its gradients (every intermediate state kept: small grids only)."""
    t = O.Tables(p)
    W = Weights(p.shape)
    o = O.Oracle(p)
    o.set_state(psi0, rho0)
    states = []
    for _ in range(nsteps):
        states.append(snapshot(o))
        o.step(1)
    r = o.f.psi - target
    L = float(np.sum(r * r))
    z = np.zeros(p.shape)
    bar = Bar(ict_T(W, 2.0 * r), z, z, z, z)
    dz = dn = 0.0
    for n in reversed(range(nsteps)):
        bar, a, b = step_vjp(p, t, W, states[n], n == 0, bar)
        dz += a
        dn += b
    # set_state: psiq = scale CT(psi0), psiq_old = psiq, rho = rho0, f_rho = 0
    dpsi0 = p.scale * ct_T(W, bar.psiq + bar.psiq_old)
    return Gradients(L, dpsi0, bar.rho.copy(), dz, dn)


def interface_ic_vjp(p: O.Params, ic: O.InterfaceIC, height, dpsi, drho):
    """dL/dh(x, y) from dL/dpsi0 and dL/drho0 through the interface IC:
    d profile/dh = 1/(2w) sech^2((z - h)/w)."""
    nx, ny, nz = p.shape
    h = np.asarray(height, dtype=np.float64).reshape(nx, ny)
    z = np.arange(nz, dtype=np.float64) * p.dz
    s = (z[None, None, :] - h[:, :, None]) / ic.width
    dprof = 0.5 / ic.width / np.cosh(s) ** 2
    g = ic.amplitude * np.cos(p.q0 * z)[None, None, :] * dpsi + (ic.rho_sm - ic.rho_iso) * drho
    return np.sum(g * dprof, axis=2)
