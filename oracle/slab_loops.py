"""Loop-level restatement of the reference's x-slab MPI step, for tiny grids.

TEST INFRASTRUCTURE ONLY (the checker of the checker).  It follows the
structure of smaiso-wcomp.cpp (S:) literally: P emulated ranks each own
local_n0 x-planes; the four halo exchanges are explicit copies of the
per-rank plane buffers psi_front/psi_back(2,3), which persist between
exchanges as in the reference (S:385-391); the velocity solve keeps the
reference's case split (S:1583-1974) and the divergence its i_local = 0
branch (S:2042-2080).  The FFTW-MPI transforms act on the gathered global
array through the direct O(N^2) sums of the FFTW manual, so nothing here
shares code with the vectorised oracle except the parameter class.

Comparing the two pins the vectorised oracle's index shifts, halo
semantics, truncations (Q1, Q2), the slab-dependent quirk Q3 and the stale
plane of Q8.
"""
from __future__ import annotations

import math

import numpy as np

from smaiso_oracle import Params, direct_r2r3d, REDFT10, REDFT01, RODFT10, RODFT01

CT = (REDFT10,) * 3
ICT = (REDFT01,) * 3
ST = {0: (RODFT10, REDFT10, REDFT10), 1: (REDFT10, RODFT10, REDFT10), 2: (REDFT10, REDFT10, RODFT10)}
IST = {0: (RODFT01, REDFT01, REDFT01), 1: (REDFT01, RODFT01, REDFT01), 2: (REDFT01, REDFT01, RODFT01)}
G = -1.57 * 1.57


class SlabStep:
    def __init__(self, p: Params, nranks: int):
        self.p = p
        nx, ny, nz = p.shape
        block = -(-nx // nranks)
        self.starts = [r * block for r in range(nranks)]
        self.n0 = [min(block, nx - s) for s in self.starts]
        assert all(n > 0 for n in self.n0), "every rank needs at least one plane"
        self.P = nranks
        pi = math.pi
        self.Vq = [[pi * i / (d * n) for i in range(n)] for n, d in ((nx, p.dx), (ny, p.dy), (nz, p.dz))]
        self.Vs = [[pi * (i + 1) / (d * n) for i in range(n)] for n, d in ((nx, p.dx), (ny, p.dy), (nz, p.dz))]
        # persistent halo planes per rank (S:385-391)
        z = lambda: np.zeros((ny, nz))
        self.front = [[z(), z(), z()] for _ in range(nranks)]
        self.back = [[z(), z(), z()] for _ in range(nranks)]
        self.nloop = 0

    # --- helpers ------------------------------------------------------------
    def gather(self, loc):
        return np.concatenate(loc, axis=0)

    def scatter(self, glob):
        return [glob[s:s + n].copy() for s, n in zip(self.starts, self.n0)]

    def xform(self, loc, kinds):
        return self.scatter(direct_r2r3d(self.gather(loc), kinds))

    def mq2(self, i, j, k, lat):
        vx = (self.Vs if lat == 0 else self.Vq)[0][i]
        vy = (self.Vs if lat == 1 else self.Vq)[1][j]
        vz = (self.Vs if lat == 2 else self.Vq)[2][k]
        return vx ** 2 + vy ** 2 + vz ** 2

    # --- state --------------------------------------------------------------
    def set_state(self, psi, rho):
        p = self.p
        psiq = direct_r2r3d(psi, CT) * p.scale
        self.psiq = self.scatter(psiq)
        self.psiq_old = self.scatter(psiq)
        self.psi = self.scatter(direct_r2r3d(psiq, ICT))
        self.rho = self.scatter(np.asarray(rho, dtype=np.float64))
        self.f_rho = [np.zeros_like(a) for a in self.rho]
        self.nl = None
        self.nloop = 0

    def step(self):
        p = self.p
        P, n0, st = self.P, self.n0, self.starts
        nx, ny, nz = p.shape
        Vq, Vs = self.Vq, self.Vs
        sc = p.scale
        self.nloop += 1
        nl_old = self.nl
        rhoq = self.xform(self.rho, CT)

        # halo #1: i_local = 0 planes of psiq, rhoq to rank-1 (S:1108-1164)
        for r in range(P):
            self.back[r][0] = self.psiq[r][0].copy()
            self.back[r][1] = rhoq[r][0].copy()
        for r in range(P - 1):
            self.front[r][0] = self.back[r + 1][0].copy()
            self.front[r][1] = self.back[r + 1][1].copy()
        gp = [[np.zeros((n0[r], ny, nz)) for r in range(P)] for _ in range(3)]
        rd = [[np.zeros((n0[r], ny, nz)) for r in range(P)] for _ in range(3)]
        for r in range(P):
            for il in range(n0[r]):
                i = il + st[r]
                for j in range(ny):
                    for k in range(nz):
                        if il < n0[r] - 1:
                            gp[0][r][il, j, k] = -Vs[0][i] * self.psiq[r][il + 1, j, k]
                            rd[0][r][il, j, k] = -sc * Vs[0][i] * rhoq[r][il + 1, j, k]
                        elif r != P - 1:
                            gp[0][r][il, j, k] = -Vs[0][i] * self.front[r][0][j, k]
                            rd[0][r][il, j, k] = -sc * Vs[0][i] * self.front[r][1][j, k]
                        if j < ny - 1:
                            gp[1][r][il, j, k] = -Vs[1][j] * self.psiq[r][il, j + 1, k]
                            rd[1][r][il, j, k] = -sc * Vs[1][j] * rhoq[r][il, j + 1, k]
                        if k < nz - 1:
                            gp[2][r][il, j, k] = -Vs[2][k] * self.psiq[r][il, j, k + 1]
                            rd[2][r][il, j, k] = -sc * Vs[2][k] * rhoq[r][il, j, k + 1]
        for a in range(3):
            gp[a] = self.xform(gp[a], IST[a])
            rd[a] = self.xform(rd[a], IST[a])

        # amplitude and smoothing (S:1254-1289)
        amp = [np.sqrt(p.q0 * p.q0 * self.psi[r] ** 2 + gp[0][r] ** 2 + gp[1][r] ** 2 + gp[2][r] ** 2)
               for r in range(P)]
        amp = self.xform(amp, CT)
        for r in range(P):
            for il in range(n0[r]):
                i = il + st[r]
                for j in range(ny):
                    for k in range(nz):
                        amp[r][il, j, k] = sc * amp[r][il, j, k] * math.exp(G * self.mq2(i, j, k, -1) * 2)
        amp = self.xform(amp, ICT)

        # L psi (S:1293-1308)
        L = [np.zeros((n0[r], ny, nz)) for r in range(P)]
        for r in range(P):
            for il in range(n0[r]):
                i = il + st[r]
                for j in range(ny):
                    for k in range(nz):
                        L[r][il, j, k] = p.alpha * (p.q0 * p.q0 - self.mq2(i, j, k, -1)) * self.psiq[r][il, j, k]
        L = self.xform(L, ICT)

        # lap(rho L) (S:1312-1380)
        rho = [a.copy() for a in self.rho]
        lap = [np.zeros((n0[r], ny, nz)) for r in range(P)]
        if self.nloop > p.divv_switch or p.load_mode == 1:
            lap = self.xform([rho[r] * L[r] for r in range(P)], CT)
            for r in range(P):
                for il in range(n0[r]):
                    i = il + st[r]
                    for j in range(ny):
                        for k in range(nz):
                            lap[r][il, j, k] = -sc * self.mq2(i, j, k, -1) * lap[r][il, j, k]
            lap = self.xform(lap, ICT)
        else:
            for r in range(P):
                for il in range(n0[r]):
                    i = il + st[r]
                    for j in range(ny):
                        for k in range(nz):
                            m = self.mq2(i, j, k, -1)
                            lap[r][il, j, k] = -m * p.alpha * (p.q0 * p.q0 - m) * self.psiq[r][il, j, k]
            lap = self.xform(lap, ICT)
            for r in range(P):
                rho[r] = p.rho0 + ((1 - p.rho0) / p.amp) * amp[r]
                lap[r] = rho[r] * lap[r]

        # stresses and forces (S:1385-1425)
        S = [[None] * P for _ in range(3)]
        F = [[None] * P for _ in range(3)]
        fgrad, erho, mu = [None] * P, [None] * P, [None] * P
        for r in range(P):
            e = 0.5 * L[r] ** 2
            fgrad[r] = 0.5 * p.zeta * (rho[r] ** 2 - (p.rho0 + p.kp * amp[r]) ** 2)
            m1 = lap[r] + rho[r] * L[r]
            for a in range(3):
                S[a][r] = m1 * gp[a][r] + e * rd[a][r]
            ps = self.psi[r]
            mu[r] = lap[r] + rho[r] * (L[r] - p.epsilon * ps - p.beta * ps ** 3 + p.gamma * ps ** 5)
            erho[r] = rho[r] * e
            fc = p.kp * p.zeta * (rho[r] - p.rho0 - p.kp * amp[r]) / (amp[r] + 0.001)
            for a in range(3):
                F[a][r] = fc * gp[a][r]
        fgrad = self.xform(fgrad, CT)
        erho = self.xform(erho, CT)
        for a in range(3):
            S[a] = self.xform(S[a], ST[a])
            F[a] = self.xform(F[a], ST[a])
        for r in range(P):
            for il in range(n0[r]):
                i = il + st[r]
                for j in range(ny):
                    for k in range(nz):
                        fgrad[r][il, j, k] += erho[r][il, j, k] * math.exp(G * self.mq2(i, j, k, -1) * 2)

        vsol = [[np.zeros((n0[r], ny, nz)) for r in range(P)] for _ in range(3)]
        virr = [[np.zeros((n0[r], ny, nz)) for r in range(P)] for _ in range(3)]
        nu, lam = p.nu, p.lam

        # halo #2 (S:1492-1579)
        for r in range(P):
            for j in range(ny):
                for k in range(nz):
                    self.back[r][0][j, k] = S[1][r][0, j - 1, k] if j != 0 else 0.0
                    self.back[r][1][j, k] = S[2][r][0, j, k - 1] if k != 0 else 0.0
                    self.back[r][2][j, k] = fgrad[r][0, j, k]
        for r in range(P - 1):
            for q in range(3):
                self.front[r][q] = self.back[r + 1][q].copy()

        def vx_at(r, il, j, k, Sy, Sz, fgn):
            i = il + st[r]
            m = self.mq2(i, j, k, 0)
            cm1, cm2 = sc / m, Vs[0][i] / m
            sx = S[0][r][il, j, k]
            dot = Vs[0][i] * sx + Vq[1][j] * Sy + Vq[2][k] * Sz
            vsol[0][r][il, j, k] = (cm1 / nu) * (sx - cm2 * dot)
            virr[0][r][il, j, k] = (cm1 / (lam + 2 * nu)) * (Vs[0][i] * fgn + cm2 * dot * math.exp(G * 2 * sc / cm1))

        for r in range(P):
            if r != P - 1:
                il = n0[r] - 1
                for j in range(ny):
                    for k in range(nz):
                        vx_at(r, il, j, k, self.front[r][0][j, k], self.front[r][1][j, k], self.front[r][2][j, k])
            for il in range(n0[r] - 1):
                for j in range(ny):
                    for k in range(nz):
                        Sy = S[1][r][il + 1, j - 1, k] if j > 0 else 0.0
                        Sz = S[2][r][il + 1, j, k - 1] if k > 0 else 0.0
                        vx_at(r, il, j, k, Sy, Sz, fgrad[r][il + 1, j, k])

        # halo #3 (S:1713-1753): i_local = local_n0-1 plane of Sx to rank+1
        for r in range(P):
            self.back[r][0] = np.zeros((ny, nz))
            self.front[r][0] = S[0][r][n0[r] - 1].copy()
        for r in range(1, P):
            self.back[r][0] = self.front[r - 1][0].copy()

        def vyz_at(r, il, j, k, Sx_y, Sz_y, Sx_z, Sy_z, quirk):
            i = il + st[r]
            m = self.mq2(i, j, k, 1)
            cm1, cm2 = sc / m, Vs[1][j] / m
            sy = S[1][r][il, j, k]
            dot = Vq[0][i] * Sx_y + Vs[1][j] * sy + Vq[2][k] * Sz_y
            vsol[1][r][il, j, k] = (cm1 / nu) * (sy - cm2 * dot)
            virr[1][r][il, j, k] = (cm1 / (lam + 2 * nu)) * (Vs[1][j] * fgrad[r][il, j + 1, k] + cm2 * dot * math.exp(G * 2 * sc / cm1))
            cm1y = cm1
            m = self.mq2(i, j, k, 2)
            cm1, cm2 = sc / m, Vs[2][k] / m
            sz = S[2][r][il, j, k]
            dot = Vq[0][i] * Sx_z + Vq[1][j] * Sy_z + Vs[2][k] * sz
            vsol[2][r][il, j, k] = (cm1 / nu) * (sz - cm2 * dot)
            cme = cm1y if quirk else cm1   # S:1932 uses CM1y in case 3 of i_local >= 1
            virr[2][r][il, j, k] = (cm1 / (lam + 2 * nu)) * (Vs[2][k] * fgrad[r][il, j, k + 1] + cm2 * dot * math.exp(G * 2 * sc / cme))

        for r in range(P):
            for il in range(n0[r]):
                for j in range(ny - 1):
                    for k in range(nz - 1):
                        if il == 0:
                            Sx_y = self.back[r][0][j + 1, k]
                            Sx_z = self.back[r][0][j, k + 1]
                        else:
                            Sx_y = S[0][r][il - 1, j + 1, k]
                            Sx_z = S[0][r][il - 1, j, k + 1]
                        Sz_y = S[2][r][il, j + 1, k - 1] if k != 0 else 0.0
                        Sy_z = S[1][r][il, j - 1, k + 1] if j != 0 else 0.0
                        quirk = (il >= 1 and k == 0 and j >= 1)
                        vyz_at(r, il, j, k, Sx_y, Sz_y, Sx_z, Sy_z, quirk)

        # halo #4 (S:1979-2035); rank 0 zeroes psi_back only: psi_back2 keeps
        # its stale plane, multiplied by Vqx[0] = 0 (quirk Q8)
        for r in range(P):
            self.front[r][0] = virr[0][r][n0[r] - 1].copy()
            self.front[r][1] = F[0][r][n0[r] - 1].copy()
        for r in range(1, P):
            self.back[r][0] = self.front[r - 1][0].copy()
            self.back[r][1] = self.front[r - 1][1].copy()
        self.back[0][0] = np.zeros((ny, nz))
        divv = [np.zeros((n0[r], ny, nz)) for r in range(P)]
        divf = [np.zeros((n0[r], ny, nz)) for r in range(P)]
        for r in range(P):
            for il in range(n0[r]):
                i = il + st[r]
                for j in range(ny):
                    for k in range(nz):
                        if il > 0:
                            sx, fx = virr[0][r][il - 1, j, k], F[0][r][il - 1, j, k]
                        else:
                            sx, fx = self.back[r][0][j, k], self.back[r][1][j, k]
                        sy, fy = (virr[1][r][il, j - 1, k], F[1][r][il, j - 1, k]) if j > 0 else (0.0, 0.0)
                        sz, fz = (virr[2][r][il, j, k - 1], F[2][r][il, j, k - 1]) if k > 0 else (0.0, 0.0)
                        divv[r][il, j, k] = Vq[0][i] * sx + Vq[1][j] * sy + Vq[2][k] * sz
                        divf[r][il, j, k] = sc * (Vq[0][i] * fx + Vq[1][j] * fy + Vq[2][k] * fz)
        divv = self.xform(divv, ICT)
        divf = self.xform(divf, ICT)
        for a in range(3):
            vsol[a] = self.xform(vsol[a], IST[a])
            virr[a] = self.xform(virr[a], IST[a])

        # density AB2 and Nl (S:2121-2162)
        nl = [None] * P
        for r in range(P):
            v = [vsol[a][r] + virr[a][r] for a in range(3)]
            src = -v[0] * rd[0][r] - v[1] * rd[1][r] - v[2] * rd[2][r] - rho[r] * divv[r]
            rho[r] = rho[r] + p.dt * (1.5 * src - 0.5 * self.f_rho[r])
            self.f_rho[r] = src
            nl[r] = (-mu[r] + (p.kp * p.zeta / (amp[r] + 0.001)) * (rho[r] - p.rho0 - p.kp * amp[r]) * self.psi[r]
                     - divf[r] - v[0] * gp[0][r] - v[1] * gp[1][r] - v[2] * gp[2][r])
        nl = self.xform(nl, CT)
        if self.nloop == 1:
            nl_old = nl
        new = [None] * P
        for r in range(P):
            new[r] = np.zeros_like(nl[r])
            for il in range(n0[r]):
                i = il + st[r]
                for j in range(ny):
                    for k in range(nz):
                        a_lin = p.epsilon - p.alpha * (p.q0 * p.q0 - self.mq2(i, j, k, -1)) ** 2
                        c1 = 2.0 - p.dt * p.rho_m_eff * a_lin
                        c2 = 1.5 - p.dt * p.rho_m_eff * a_lin
                        new[r][il, j, k] = (c1 * self.psiq[r][il, j, k] - 0.5 * self.psiq_old[r][il, j, k]
                                            + (p.dt / 2) * sc * (3.0 * nl[r][il, j, k] - nl_old[r][il, j, k])) / c2
        self.psiq_old = self.psiq
        self.psiq = new
        self.psi = self.xform(new, ICT)
        self.rho = rho
        self.nl = nl
        self.out = dict(amp=amp, lpsi=L, lap_rl=lap, mu=mu, divv=divv, divf=divf,
                        vsolx=vsol[0], vsoly=vsol[1], vsolz=vsol[2],
                        virrx=virr[0], virry=virr[1], virrz=virr[2],
                        gpx=gp[0], gpy=gp[1], gpz=gp[2], rdx=rd[0], rdy=rd[1], rdz=rd[2])

    def fields(self):
        d = {k: self.gather(v) for k, v in self.out.items()}
        d.update(psi=self.gather(self.psi), rho=self.gather(self.rho), psiq=self.gather(self.psiq),
                 f_rho=self.gather(self.f_rho), nlq=self.gather(self.nl))
        return d
