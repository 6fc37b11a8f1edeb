"""FP64 CPU oracle for the smaiso-wcomp pseudo-spectral time stepper.

TEST INFRASTRUCTURE ONLY.  This module is the checker, never the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``libsmaiso.so`` behind
``include/smaiso.h``) never calls into it.

It restates, in numpy, the reference's algorithm from
``smaiso-wcomp.cpp`` (abbreviated ``S:`` below):

* parameters ................ S:239-283, S:206-214
* wavenumbers ............... S:447-481
* r2r transform kinds ....... S:499-545 (FFTW REDFT10/REDFT01/RODFT10/RODFT01)
* operator tables, pre-loop . S:936-1010
* one time step ............. S:1081-2198
* diagnostics ............... S:2202-2263

The x-slab MPI decomposition of the reference is replaced by the global
array; every halo exchange becomes a +-1 index shift with the out-of-range
value 0 (S:1108-1164, S:1492-1579, S:1713-1753, S:1979-2035).  The only
result that depends on the decomposition is quirk Q3 (``CM1y`` in the
``virrz`` exponent at k=0, S:1932), which the reference applies on every
plane except the first plane of each rank's slab; ``slab_n0`` selects the
slab height (default: one slab, so only plane i=0 is exempt).

The 1-D transforms use ``scipy.fft.dct``/``dst`` types 2 and 3, whose
definitions are FFTW's REDFT10/REDFT01/RODFT10/RODFT01 (unnormalised).  That
identity is pinned in the tests against direct O(N^2) sums written from the
FFTW manual's formulas (``direct_r2r_1d`` below).

Parity pins (all in ``tests/test_oracle.py``): transforms vs direct sums and
round trips; the spectral gradient vs a single-mode closed form; amplitude of
A cos(q0 z) equal to A q0; the SH and Laplacian eigenvalues; the uniform
state psi=0, rho=rho0 being stationary; the single-mode Stokes solve
(solenoidal part orthogonal to q on the staggered lattice); the per-mode
linear recurrence of the psi update; diagnostics vs math.fsum; and an
independent pure-Python loop restatement of the whole step on tiny grids.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft

REDFT10, REDFT01, RODFT10, RODFT01 = 0, 1, 2, 3
KIND_NAMES = {REDFT10: "REDFT10", REDFT01: "REDFT01", RODFT10: "RODFT10", RODFT01: "RODFT01"}

_WORKERS = max(1, int(os.environ.get("SMAISO_ORACLE_THREADS", os.cpu_count() or 1)))


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
def r2r_1d(x: np.ndarray, kind: int, axis: int) -> np.ndarray:
    """FFTW r2r kind along one axis (unnormalised), S:499-545."""
    if kind == REDFT10:
        return sfft.dct(x, type=2, axis=axis, workers=_WORKERS)
    if kind == REDFT01:
        return sfft.dct(x, type=3, axis=axis, workers=_WORKERS)
    if kind == RODFT10:
        return sfft.dst(x, type=2, axis=axis, workers=_WORKERS)
    if kind == RODFT01:
        return sfft.dst(x, type=3, axis=axis, workers=_WORKERS)
    raise ValueError(f"unknown r2r kind {kind}")


def r2r3d(x: np.ndarray, kinds) -> np.ndarray:
    """3-D r2r with one kind per axis, like ``fftw_mpi_plan_r2r_3d`` (S:499-545).
    A complex array is transformed as real and imaginary parts (the transforms
    are real-linear), which is what complex-step differentiation needs."""
    if np.iscomplexobj(x):
        return r2r3d(np.real(x), kinds) + 1j * r2r3d(np.imag(x), kinds)
    y = np.asarray(x, dtype=np.float64)
    for ax in (2, 1, 0):
        y = r2r_1d(y, kinds[ax], ax)
    return y


def direct_r2r_1d(x: np.ndarray, kind: int) -> np.ndarray:
    """O(N^2) sums written from the FFTW manual's definitions of the four kinds."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    idx = np.arange(n, dtype=np.float64)
    kk = idx[:, None]  # output index
    nn = idx[None, :]  # input index
    if kind == REDFT10:   # Y_k = 2 sum_j X_j cos(pi (j+1/2) k / n)
        m = 2.0 * np.cos(np.pi * (nn + 0.5) * kk / n)
        return m @ x
    if kind == REDFT01:   # Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2) / n)
        m = 2.0 * np.cos(np.pi * nn * (kk + 0.5) / n)
        m[:, 0] = 1.0
        return m @ x
    if kind == RODFT10:   # Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1) / n)
        m = 2.0 * np.sin(np.pi * (nn + 0.5) * (kk + 1.0) / n)
        return m @ x
    if kind == RODFT01:   # Y_k = (-1)^k X_{n-1} + 2 sum_{j<n-1} X_j sin(pi (j+1)(k+1/2)/n)
        m = 2.0 * np.sin(np.pi * (nn + 1.0) * (kk + 0.5) / n)
        m[:, n - 1] = (-1.0) ** idx
        return m @ x
    raise ValueError(kind)


def direct_r2r3d(x: np.ndarray, kinds) -> np.ndarray:
    y = np.asarray(x, dtype=np.float64)
    for ax in (2, 1, 0):
        y = np.moveaxis(np.tensordot(_direct_matrix(y.shape[ax], kinds[ax]), np.moveaxis(y, ax, 0), axes=(1, 0)), 0, ax)
    return y


def _direct_matrix(n: int, kind: int) -> np.ndarray:
    return direct_r2r_1d(np.eye(n), kind)


CT = (REDFT10, REDFT10, REDFT10)      # planCT   S:499-503
ICT = (REDFT01, REDFT01, REDFT01)     # iPlanCT  S:505-509
ST = {0: (RODFT10, REDFT10, REDFT10),  # planSTx  S:511-515
      1: (REDFT10, RODFT10, REDFT10),  # planSTy  S:517-521
      2: (REDFT10, REDFT10, RODFT10)}  # planSTz  S:523-527
IST = {0: (RODFT01, REDFT01, REDFT01),  # iPlanSTx S:529-533
       1: (REDFT01, RODFT01, REDFT01),  # iPlanSTy S:535-539
       2: (REDFT01, REDFT01, RODFT01)}  # iPlanSTz S:541-545


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------
@dataclass
class Params:
    """Inputs of the stepper; defaults are the reference's constants (S:206-283)."""
    nx: int = 64
    ny: int = 64
    nz: int = 64
    dx: float = 2.0 * math.pi / 8.0     # L/N = 2 pi/(q0 Nw), Nw=8  (S:261, S:267-272)
    dy: float = 2.0 * math.pi / 8.0
    dz: float = 2.0 * math.pi / 8.0
    dt: float = 0.001                   # S:262
    alpha: float = 1.0                  # S:241
    beta: float = 2.0                   # S:240
    gamma: float = 1.0                  # S:239
    epsilon: float = -0.005             # ep = -0.001*ep_arg, ep_arg=5 (S:242-243)
    q0: float = 1.0                     # S:244
    nu: float = 1.0                     # argv[1] (S:249)
    lam: float = 1.0                    # lambda, S:256
    zeta: float = 1.0                   # S:257
    rho0: float = 0.01                  # rho_0, S:251
    rho_m: float | None = None          # None -> rho_s/2 (S:255)
    divv_switch: int = 50               # S:214
    load_mode: int = 0                  # argv[4] (S:176)
    slab_n0: int = 0                    # reference slab height for quirk Q3; 0 -> nx

    # derived (S:250-255, S:278)
    @property
    def amp(self) -> float:
        b, g, e = self.beta, self.gamma, self.epsilon
        return math.sqrt((3 * b + math.sqrt(9 * b * b + 40 * e * g)) / 5)

    @property
    def kp(self) -> float:
        return (1 - self.rho0) / self.amp

    @property
    def rho_s(self) -> float:
        return self.kp * self.amp + self.rho0

    @property
    def rho_m_eff(self) -> float:
        return self.rho_s / 2 if self.rho_m is None else self.rho_m

    @property
    def scale(self) -> float:
        return 0.125 / (self.nx * self.ny * self.nz)

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


def wavenumbers(p: Params):
    """Vq (DCT lattice) and Vs (DST lattice) per axis, S:447-481."""
    vq, vs = [], []
    for n, d in ((p.nx, p.dx), (p.ny, p.dy), (p.nz, p.dz)):
        idx = np.arange(n, dtype=np.float64)
        vq.append(1.0 * math.pi * idx / (d * n))
        vs.append(1.0 * math.pi * (idx + 1) / (d * n))
    return vq, vs


def _b(v, ax):
    s = [None, None, None]
    s[ax] = slice(None)
    return v[tuple(s)]


def gauss(mq2):
    """exp(-1.57*1.57*|q|^2*2), S:1281 / S:1486 (quirk Q7: the constant)."""
    return np.exp(-1.57 * 1.57 * mq2 * 2)


class Tables:
    """Operator tables of S:936-963 (mq2c, aLin, C1, C2, CM1a, CM2a)."""

    def __init__(self, p: Params):
        vq, vs = wavenumbers(p)
        self.vq, self.vs = vq, vs
        qx, qy, qz = (_b(vq[a], a) for a in range(3))
        sx, sy, sz = (_b(vs[a], a) for a in range(3))
        self.mq2c = qx ** 2 + qy ** 2 + qz ** 2
        op_sh = p.alpha * (p.q0 * p.q0 - self.mq2c) ** 2
        self.aLin = p.epsilon - op_sh
        self.C1 = 2.0 - p.dt * p.rho_m_eff * self.aLin
        self.C2 = 1.5 - p.dt * p.rho_m_eff * self.aLin
        mq2 = sx ** 2 + qy ** 2 + qz ** 2
        self.CM1x = p.scale / mq2
        self.CM2x = sx / mq2
        mq2 = qx ** 2 + sy ** 2 + qz ** 2
        self.CM1y = p.scale / mq2
        self.CM2y = sy / mq2
        mq2 = qx ** 2 + qy ** 2 + sz ** 2
        self.CM1z = p.scale / mq2
        self.CM2z = sz / mq2
        self.G = gauss(self.mq2c)


# ---------------------------------------------------------------------------
# initial conditions (builder-defined; the reference's own IC, S:571-590,
# is available as ``reference_layers_ic``)
# ---------------------------------------------------------------------------
_SM_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_SM_M1 = np.uint64(0xBF58476D1CE4E5B9)
_SM_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """U[0,1) from the splitmix64 stream (Steele, Lea, Flood 2014), element n
    uses state seed + (n+1)*0x9E3779B97F4A7C15, top 53 bits."""
    with np.errstate(over="ignore"):
        n = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + n * _SM_GAMMA
        z = (z ^ (z >> np.uint64(30))) * _SM_M1
        z = (z ^ (z >> np.uint64(27))) * _SM_M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class InterfaceIC:
    """Layered smectic below an interface at height h(x,y), isotropic above.

    psi = A cos(q0 z) * 1/2 (1 - tanh((z - h)/w)) + noise
    rho = rho_iso + (rho_sm - rho_iso) * 1/2 (1 - tanh((z - h)/w))
    h(x,y) = h0 + u cos(2 pi x/Lx) cos(2 pi y/Ly)  (or an explicit height map)
    with x = i dx, z = k dz (grid points, as S:588), noise U(-a, a) from
    splitmix64 in row-major order.
    """
    amplitude: float = 1.0
    h0: float | None = None          # None -> Lz/2
    width: float = 2.0
    rho_iso: float = 0.9
    rho_sm: float = 1.0
    undulation: float = 0.0
    noise: float = 1e-3
    seed: int = 1
    height: np.ndarray | None = None  # (nx, ny) explicit h(x,y); overrides h0/undulation


def make_interface_ic(p: Params, ic: InterfaceIC):
    nx, ny, nz = p.shape
    lx, ly, lz = nx * p.dx, ny * p.dy, nz * p.dz
    if ic.height is not None:
        h = np.asarray(ic.height).reshape(nx, ny)  # complex allowed (complex-step)
    else:
        h0 = lz / 2 if ic.h0 is None else ic.h0
        x = np.arange(nx, dtype=np.float64) * p.dx
        y = np.arange(ny, dtype=np.float64) * p.dy
        h = h0 + ic.undulation * np.cos(2 * math.pi * x / lx)[:, None] * np.cos(2 * math.pi * y / ly)[None, :]
    z = np.arange(nz, dtype=np.float64) * p.dz
    prof = 0.5 * (1.0 - np.tanh((z[None, None, :] - h[:, :, None]) / ic.width))
    psi = ic.amplitude * np.cos(p.q0 * z)[None, None, :] * prof
    u = splitmix64_uniform(ic.seed, nx * ny * nz).reshape(nx, ny, nz)
    psi = psi + ic.noise * (2.0 * u - 1.0)
    rho = ic.rho_iso + (ic.rho_sm - ic.rho_iso) * prof
    return np.ascontiguousarray(psi), np.ascontiguousarray(rho)


def reference_layers_ic(p: Params):
    """The reference's active IC, S:571-590: Amp cos(q0 k dz) for Nz/2-8<k<Nz/2+8."""
    psi = np.zeros(p.shape)
    k = np.arange(p.nz)
    sel = (k > p.nz // 2 - 8) & (k < p.nz // 2 + 8)
    psi[:, :, sel] = p.amp * np.cos(p.q0 * k[sel] * p.dz)
    return psi, np.zeros(p.shape)


# ---------------------------------------------------------------------------
# the stepper
# ---------------------------------------------------------------------------
@dataclass
class Fields:
    psi: np.ndarray
    psiq: np.ndarray
    psiq_old: np.ndarray
    rho: np.ndarray
    f_rho: np.ndarray
    nlq: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


class Oracle:
    """FP64 restatement of the reference time loop body (S:1081-2198)."""

    def __init__(self, p: Params):
        self.p = p
        self.t = Tables(p)
        self.nloop = 0
        self.f: Fields | None = None

    # S:965-1010: psiq = scale*DCT(psi); psi = iDCT(psiq); psiq_old = psiq.
    # The pre-loop Nl of S:996-1008 is overwritten before use (S:2170-2172).
    def set_state(self, psi: np.ndarray, rho: np.ndarray):
        p = self.p
        # complex states are allowed for complex-step derivatives
        dt = np.result_type(psi, rho, np.float64)
        psi = np.asarray(psi, dtype=dt).reshape(p.shape)
        rho = np.asarray(rho, dtype=dt).reshape(p.shape)
        psiq = r2r3d(psi, CT)
        psiq = p.scale * psiq
        psi = r2r3d(psiq, ICT)
        self.f = Fields(psi=psi, psiq=psiq, psiq_old=psiq.copy(), rho=rho.copy(),
                        f_rho=np.zeros(p.shape, dtype=dt))
        self.nloop = 0

    def step(self, nsteps: int = 1):
        for _ in range(nsteps):
            self._step()

    def _step(self):
        p, t, f = self.p, self.t, self.f
        self.nloop += 1
        nl_old = f.nlq                                   # S:1100
        psi, psiq, rho = f.psi, f.psiq, f.rho

        gpx, gpy, gpz, rdx, rdy, rdz = gradients(p, t, psiq, rho)          # a6
        amp = amplitude(p, t, psi, gpx, gpy, gpz)                          # a7
        lpsi = sh_operator(p, t, psiq)                                     # a8
        relax = not (self.nloop > p.divv_switch or p.load_mode == 1)
        lap_rl, rho = lap_rho_l(p, t, psiq, rho, lpsi, amp, relax)         # a9
        fgrad, erho, mu, Sx, Sy, Sz, fx, fy, fz = forces(
            p, lpsi, lap_rl, rho, amp, psi, gpx, gpy, gpz, rdx, rdy, rdz)  # a10

        # -- a11: forward transforms, S:1427-1488 -------------------------
        fgq = r2r3d(fgrad, CT)
        erq = r2r3d(erho, CT)
        Sxq = r2r3d(Sx, ST[0]); Syq = r2r3d(Sy, ST[1]); Szq = r2r3d(Sz, ST[2])
        fxq = r2r3d(fx, ST[0]); fyq = r2r3d(fy, ST[1]); fzq = r2r3d(fz, ST[2])
        fgq = fgq + erq * np.exp(-1.57 * 1.57 * t.mq2c * 2)

        vsolx, virrx, vsoly, virry, vsolz, virrz = velocity_solve(p, t, Sxq, Syq, Szq, fgq)  # a12
        divv, divf = divergences(p, t, virrx, virry, virrz, fxq, fyq, fzq)                # a13

        # -- a14: inverse transforms, S:2085-2117 -------------------------
        divv = r2r3d(divv, ICT)
        divf = r2r3d(divf, ICT)
        vsolx = r2r3d(vsolx, IST[0]); vsoly = r2r3d(vsoly, IST[1]); vsolz = r2r3d(vsolz, IST[2])
        virrx = r2r3d(virrx, IST[0]); virry = r2r3d(virry, IST[1]); virrz = r2r3d(virrz, IST[2])
        vx, vy, vz = vsolx + virrx, vsoly + virry, vsolz + virrz

        rho, f_rho = density_update(p, vx, vy, vz, rdx, rdy, rdz, rho, divv, f.f_rho)  # a15
        nl = nonlinear(p, mu, amp, rho, psi, divf, vx, vy, vz, gpx, gpy, gpz)          # a16
        nlq = r2r3d(nl, CT)                              # S:2166-2168
        if self.nloop == 1 or nl_old is None:           # quirk Q5, S:2170-2172
            nl_old = nlq
        psiq_new = psi_update(p, t, psiq, f.psiq_old, nlq, nl_old)                     # a17
        psi_new = r2r3d(psiq_new, ICT)                   # S:2194-2198

        f.extra = dict(psi_old=psi, lpsi=lpsi, amp=amp, mu=mu, divv=divv, divf=divf,
                       vsolx=vsolx, vsoly=vsoly, vsolz=vsolz,
                       virrx=virrx, virry=virry, virrz=virrz, vx=vx, vy=vy, vz=vz,
                       gpx=gpx, gpy=gpy, gpz=gpz, rdx=rdx, rdy=rdy, rdz=rdz,
                       lap_rl=lap_rl, nl=nl)
        f.psiq_old = psiq
        f.psiq = psiq_new
        f.psi = psi_new
        f.rho = rho
        f.f_rho = f_rho
        f.nlq = nlq

    # -- a18: diagnostics, S:2202-2263 --------------------------------------
    def diagnostics(self):
        """L1, mass, energy of the last completed step (S:2215-2258)."""
        p, f = self.p, self.f
        x = f.extra
        psi, psi_old, rho, lpsi, amp = f.psi, x["psi_old"], f.rho, x["lpsi"], x["amp"]
        sum_a = np.sum(np.abs(psi - psi_old))
        sum_b = np.sum(np.abs(psi))
        l1 = sum_a / (p.dt * sum_b)
        mass = np.sum(rho) * p.dx * p.dy * p.dz
        en = energy_density(p, psi, lpsi, rho, amp)
        energy = np.sum(en) * p.dx * p.dy * p.dz
        return float(l1), float(mass), float(energy)


# ---------------------------------------------------------------------------
# stages of one step
# ---------------------------------------------------------------------------
def gradients(p: Params, t: Tables, psiq, rho):
    """a6: rho^ = DCT(rho); DCT -> DST index shift with the minus sign,
    S:1104-1250.  Returns real-space grad psi and grad rho."""
    sc = p.scale
    sx, sy, sz = (_b(t.vs[a], a) for a in range(3))
    rhoq = r2r3d(rho, CT)                            # S:1104-1106
    dt = np.result_type(psiq, rhoq)
    gpx = np.zeros(p.shape, dt); gpy = np.zeros(p.shape, dt); gpz = np.zeros(p.shape, dt)
    rdx = np.zeros(p.shape, dt); rdy = np.zeros(p.shape, dt); rdz = np.zeros(p.shape, dt)
    gpx[:-1] = -sx[:-1] * psiq[1:]                   # S:1166-1192
    rdx[:-1] = -sc * sx[:-1] * rhoq[1:]
    gpy[:, :-1] = -sy[:, :-1] * psiq[:, 1:]          # S:1196-1206
    rdy[:, :-1] = -sc * sy[:, :-1] * rhoq[:, 1:]
    gpz[:, :, :-1] = -sz[:, :, :-1] * psiq[:, :, 1:]  # S:1210-1221
    rdz[:, :, :-1] = -sc * sz[:, :, :-1] * rhoq[:, :, 1:]
    gpx = r2r3d(gpx, IST[0]); gpy = r2r3d(gpy, IST[1]); gpz = r2r3d(gpz, IST[2])
    rdx = r2r3d(rdx, IST[0]); rdy = r2r3d(rdy, IST[1]); rdz = r2r3d(rdz, IST[2])
    return gpx, gpy, gpz, rdx, rdy, rdz


def amplitude(p: Params, t: Tables, psi, gpx, gpy, gpz):
    """a7: amp = sqrt(q0^2 psi^2 + |grad psi|^2), Gaussian-filtered, S:1254-1289."""
    amp = np.sqrt(p.q0 * p.q0 * psi ** 2 + gpx ** 2 + gpy ** 2 + gpz ** 2)
    amp = r2r3d(amp, CT)
    amp = p.scale * amp * np.exp(-1.57 * 1.57 * t.mq2c * 2)
    return r2r3d(amp, ICT)


def sh_operator(p: Params, t: Tables, psiq):
    """a8: L psi = iDCT(alpha (q0^2 - |q|^2) psi^), S:1293-1308."""
    return r2r3d(p.alpha * (p.q0 * p.q0 - t.mq2c) * psiq, ICT)


def lap_rho_l(p: Params, t: Tables, psiq, rho, lpsi, amp, relax: bool):
    """a9: lap(rho L psi), S:1312-1380.  In the relaxation branch rho is
    replaced by rho0 + (1-rho0)/Amp amp (quirk Q4).  Returns (lap_rl, rho)."""
    if not relax:
        w = rho * lpsi
        w = r2r3d(w, CT)
        w = -p.scale * t.mq2c * w
        return r2r3d(w, ICT), rho
    w = -t.mq2c * p.alpha * (p.q0 * p.q0 - t.mq2c) * psiq
    lap_rl = r2r3d(w, ICT)
    rho = p.rho0 + ((1 - p.rho0) / p.amp) * amp       # S:1373-1374
    return rho * lap_rl, rho


def forces(p: Params, lpsi, lap_rl, rho, amp, psi, gpx, gpy, gpz, rdx, rdy, rdz):
    """a10: stresses, forces and chemical potential, S:1385-1425."""
    e = 0.5 * lpsi ** 2
    fgrad = 0.5 * p.zeta * (rho ** 2 - (p.rho0 + p.kp * amp) ** 2)
    mu1 = lap_rl + rho * lpsi
    Sx = mu1 * gpx + e * rdx
    Sy = mu1 * gpy + e * rdy
    Sz = mu1 * gpz + e * rdz
    mu = lap_rl + rho * (lpsi - p.epsilon * psi - p.beta * psi ** 3 + p.gamma * psi ** 5)
    erho = rho * e
    fc = p.kp * p.zeta * (rho - p.rho0 - p.kp * amp) / (amp + 0.001)
    return fgrad, erho, mu, Sx, Sy, Sz, fc * gpx, fc * gpy, fc * gpz


def divergences(p: Params, t: Tables, virrx, virry, virrz, fxq, fyq, fzq):
    """a13: div v^ and div f^ on the DCT lattice (DST -> DCT shift, plus sign),
    S:2042-2080."""
    qx, qy, qz = (_b(t.vq[a], a) for a in range(3))
    divv = qx * _shift_m(virrx, 0) + qy * _shift_m(virry, 1) + qz * _shift_m(virrz, 2)
    divf = p.scale * (qx * _shift_m(fxq, 0) + qy * _shift_m(fyq, 1) + qz * _shift_m(fzq, 2))
    return divv, divf


def density_update(p: Params, vx, vy, vz, rdx, rdy, rdz, rho, divv, f_rho_old):
    """a15: AB2 density update with the step-start grad rho, S:2121-2139."""
    s_rho = -vx * rdx - vy * rdy - vz * rdz - rho * divv
    rho_new = rho + p.dt * (1.5 * s_rho - 0.5 * f_rho_old)
    return rho_new, s_rho


def nonlinear(p: Params, mu, amp, rho_new, psi, divf, vx, vy, vz, gpx, gpy, gpz):
    """a16: explicit terms of the psi equation, S:2144-2162."""
    return (-mu + (p.kp * p.zeta / (amp + 0.001)) * (rho_new - p.rho0 - p.kp * amp) * psi
            - divf - vx * gpx - vy * gpy - vz * gpz)


def psi_update(p: Params, t: Tables, psiq, psiq_old, nlq, nlq_old):
    """a17: stabilised BDF2 / AB2 update, S:2177-2190."""
    return (t.C1 * psiq - 0.5 * psiq_old + (p.dt / 2) * p.scale * (3.0 * nlq - nlq_old)) / t.C2


def energy_density(p: Params, psi, lpsi, rho, amp):
    """Summand of the free energy at S:2227-2233."""
    e = (-0.5 * p.epsilon * psi ** 2 + 0.5 * lpsi ** 2 - 0.25 * p.beta * psi ** 4
         + (p.gamma / 6) * psi ** 6)
    return rho * e + 0.5 * p.zeta * (rho - p.rho0 - p.kp * amp) ** 2


def _shift_m(a, ax):
    """out[n] = a[n-1] along ax, 0 at n=0 (the 'index-1' reads of S:2042-2080)."""
    out = np.zeros_like(a)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    src[ax] = slice(0, -1)
    dst[ax] = slice(1, None)
    out[tuple(dst)] = a[tuple(src)]
    return out


def velocity_solve(p: Params, t: Tables, Sxq, Syq, Szq, fgq):
    """Per-mode solenoidal/irrotational velocity on the staggered lattices,
    S:1583-1974, including quirks Q1 (v_y, v_z only for j<=Ny-2 and k<=Nz-2),
    Q2 (v_x = 0 at i = Nx-1) and Q3 (CM1y in the virrz exponent at k = 0,
    j >= 1, except on the first plane of each slab)."""
    nx, ny, nz = p.shape
    sc = p.scale
    vq, vs = t.vq, t.vs
    qx, qy, qz = (_b(vq[a], a) for a in range(3))
    sx, sy, sz = (_b(vs[a], a) for a in range(3))
    nu, lam = p.nu, p.lam
    shape = p.shape
    gfac = -1.57 * 1.57 * 2 * sc
    dt = np.result_type(Sxq, Syq, Szq, fgq, nu, lam)

    def zeros(sh):
        return np.zeros(sh, dt)

    # x component: i <= Nx-2 (S:1583-1709)
    Sy_s = zeros(shape); Sy_s[:-1, 1:, :] = Syq[1:, :-1, :]
    Sz_s = zeros(shape); Sz_s[:-1, :, 1:] = Szq[1:, :, :-1]
    fg_s = zeros(shape); fg_s[:-1] = fgq[1:]
    dot = sx * Sxq + qy * Sy_s + qz * Sz_s
    vsolx = (t.CM1x / nu) * (Sxq - t.CM2x * dot)
    virrx = (t.CM1x / (lam + 2 * nu)) * (sx * fg_s + t.CM2x * dot * np.exp(gfac / t.CM1x))
    vsolx[-1] = 0.0
    virrx[-1] = 0.0

    # y component: j <= Ny-2, k <= Nz-2 (S:1759-1974)
    Sx_s = zeros(shape); Sx_s[1:, :-1, :] = Sxq[:-1, 1:, :]
    Sz_s = zeros(shape); Sz_s[:, :-1, 1:] = Szq[:, 1:, :-1]
    fg_s = zeros(shape); fg_s[:, :-1] = fgq[:, 1:]
    dot = qx * Sx_s + sy * Syq + qz * Sz_s
    vsoly = (t.CM1y / nu) * (Syq - t.CM2y * dot)
    virry = (t.CM1y / (lam + 2 * nu)) * (sy * fg_s + t.CM2y * dot * np.exp(gfac / t.CM1y))
    vsoly[:, -1, :] = 0.0; vsoly[:, :, -1] = 0.0
    virry[:, -1, :] = 0.0; virry[:, :, -1] = 0.0

    # z component: j <= Ny-2, k <= Nz-2
    Sx_s = zeros(shape); Sx_s[1:, :, :-1] = Sxq[:-1, :, 1:]
    Sy_s = zeros(shape); Sy_s[:, 1:, :-1] = Syq[:, :-1, 1:]
    fg_s = zeros(shape); fg_s[:, :, :-1] = fgq[:, :, 1:]
    dot = qx * Sx_s + qy * Sy_s + sz * Szq
    vsolz = (t.CM1z / nu) * (Szq - t.CM2z * dot)
    cm_exp = t.CM1z.copy()
    slab = p.slab_n0 if p.slab_n0 > 0 else nx
    i_idx = np.arange(nx)
    q3 = (i_idx % slab) != 0                           # i_local >= 1
    cm_exp[q3, 1:, 0] = t.CM1y[q3, 1:, 0]              # quirk Q3 (S:1932)
    virrz = (t.CM1z / (lam + 2 * nu)) * (sz * fg_s + t.CM2z * dot * np.exp(gfac / cm_exp))
    vsolz[:, -1, :] = 0.0; vsolz[:, :, -1] = 0.0
    virrz[:, -1, :] = 0.0; virrz[:, :, -1] = 0.0
    return vsolx, virrx, vsoly, virry, vsolz, virrz


# ---------------------------------------------------------------------------
# data of the output writers (S:94-129, S:2273-2446)
# ---------------------------------------------------------------------------
def mid_slice(field: np.ndarray, j: int) -> np.ndarray:
    """saveMid (S:94-129): the x-z slice at y index j, i-major then k."""
    return np.ascontiguousarray(np.asarray(field)[:, j, :]).reshape(-1)


def surface_track(rho: np.ndarray, threshold: float, prev: np.ndarray) -> np.ndarray:
    """Surface tracker (S:2278-2318): for each column (i, j) the largest k with
    rho > threshold, scanning down from k = Nz-1; a column without one keeps
    its previous value (the reference leaves surfZ_local untouched, quirk Q9).
    Returns the Nx*Ny values in (i, j) order."""
    rho = np.asarray(rho)
    nx, ny, nz = rho.shape
    out = np.array(prev, dtype=np.float64).reshape(nx * ny).copy()
    for i in range(nx):
        for j in range(ny):
            for k in range(nz - 1, -1, -1):
                if rho[i, j, k] > threshold:
                    out[i * ny + j] = k
                    break
    return out
