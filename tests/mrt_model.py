"""Independent matrix-form model of the reference time step, used to pin the C oracle.

The C oracle (oracle/rb_oracle.c) restates the reference's unrolled Fortran: 19 hand-written
moment rows, 19 hand-written inverse rows, hand-expanded forcing terms.  This model rebuilds the
same scheme from its published definitions instead, so that a transcription slip in the oracle
(a dropped term, a wrong sign, a swapped index or coefficient) shows up as a mismatch:

* moment bases from their polynomial definitions: D3Q19 after d'Humieres et al. (2002),
  D2Q9 after Lallemand & Luo (2000), D3Q7 / D2Q5 after the rows the reference names;
* the inverse transforms by numerical matrix inversion;
* equilibrium moments as M @ f_eq where the second-order equilibrium fixes them (rho, e, j,
  the stress tensor), with the model constants (w_eps = 3, w_epsj = -11/2, w_xx = -1/2,
  q = -2/3 j; 2D: eps = rho - 3 j^2/rho, q = -j) for the rest;
* the Guo forcing term in moment space as (1 - S/2) (d m_eq / d j) . F, evaluated exactly by a
  complex step (m_eq is quadratic in j);
* the walls as vectorised masks applied in the reference's sweep order.

Only the time-step semantics (sweep order, which rho/T the half force uses, M1G history order)
are taken from the reference text: 2DRB/3DRB.F90:225-243 and 2DRB/2DRB.F90:236-254.
"""
from __future__ import annotations

import numpy as np

# Direction order of the reference (2DRB/3DRB.F90:656-658, 2DRB/2DRB.F90:188-190).
E3 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
               [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1],
               [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
E2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [1, 1, 0], [-1, 1, 0],
               [-1, -1, 0], [1, -1, 0]])


def opposite(E):
    return np.array([int(np.where((E == -e).all(axis=1))[0][0]) for e in E])


def weights(E, dims):
    e2 = (E ** 2).sum(axis=1)
    if dims == 3:
        table = {0: 1 / 3, 1: 1 / 18, 2: 1 / 36} if len(E) == 19 else {0: 1 / 4, 1: 1 / 8}
    else:
        table = {0: 4 / 9, 1: 1 / 9, 2: 1 / 36} if len(E) == 9 else {0: 1 / 3, 1: 1 / 6}
    return np.array([table[int(x)] for x in e2])


def basis_d3q19(E):
    ex, ey, ez = (E[:, i].astype(float) for i in range(3))
    e2 = ex ** 2 + ey ** 2 + ez ** 2
    return np.array([
        np.ones_like(ex), 19 * e2 - 30, (21 * e2 ** 2 - 53 * e2 + 24) / 2,
        ex, (5 * e2 - 9) * ex, ey, (5 * e2 - 9) * ey, ez, (5 * e2 - 9) * ez,
        3 * ex ** 2 - e2, (3 * e2 - 5) * (3 * ex ** 2 - e2),
        ey ** 2 - ez ** 2, (3 * e2 - 5) * (ey ** 2 - ez ** 2),
        ex * ey, ey * ez, ex * ez,
        (ey ** 2 - ez ** 2) * ex, (ez ** 2 - ex ** 2) * ey, (ex ** 2 - ey ** 2) * ez])


def basis_d3q7(E):
    ex, ey, ez = (E[:, i].astype(float) for i in range(3))
    e2 = ex ** 2 + ey ** 2 + ez ** 2
    return np.array([np.ones_like(ex), ex, ey, ez, 7 * e2 - 6, 3 * ex ** 2 - e2, ey ** 2 - ez ** 2])


def basis_d2q9(E):
    ex, ey = E[:, 0].astype(float), E[:, 1].astype(float)
    e2 = ex ** 2 + ey ** 2
    return np.array([np.ones_like(ex), -4 + 3 * e2, 4 - 21 / 2 * e2 + 9 / 2 * e2 ** 2,
                     ex, (-5 + 3 * e2) * ex, ey, (-5 + 3 * e2) * ey, ex ** 2 - ey ** 2, ex * ey])


def basis_d2q5(E):
    ex, ey = E[:, 0].astype(float), E[:, 1].astype(float)
    e2 = ex ** 2 + ey ** 2
    return np.array([np.ones_like(ex), ex, ey, 5 * e2 - 4, ex ** 2 - ey ** 2])


def params(dims, nx, ny, Ra, Pr, Ma, thot, tcold, side_heated=False, legacy=False):
    """Dimensionless groups of the reference, rebuilt from their physical meaning:
    the free-fall velocity sqrt(g beta dT L) equals Ma times the lattice sound speed.
    Legacy thermal scheme: fixed rates Qk = 3 - sqrt3, Qnu = 4 sqrt3 - 6; the diffusivity
    kappa = cT^2 (1/Qk - 1/2) then fixes the moving-direction weight, with the thermal sound
    speed cT^2 = sum_a wT_a e_ax^2 = 2 wT_1 and wT_0 = 1 - (q - 1) wT_1."""
    L = float(nx if side_heated else ny)
    cs = 1 / np.sqrt(3.0)
    U = Ma * cs
    nu = U * L * np.sqrt(Pr / Ra)
    kappa = nu / Pr
    p = dict(L=L, Tref=0.5 * (thot + tcold), viscosity=nu, diffusivity=kappa,
             gbeta=U * U / L, tauf=0.5 + 3 * nu)
    p["taug"] = 0.5 + 4 * kappa if dims == 3 else 0.5 + 3 * kappa
    p["Qk"] = 1 / p["taug"]
    p["Qnu"] = 1.0
    if legacy:
        p["Qk"] = 3 - np.sqrt(3.0)
        p["Qnu"] = 4 * np.sqrt(3.0) - 6
        cT2 = kappa / (1 / p["Qk"] - 0.5)
        p["wT1"] = cT2 / 2
        p["wT0"] = 1 - (6 if dims == 3 else 4) * p["wT1"]
    # "magic" combination (1/s_q - 1/2)(1/s_nu - 1/2) = 3/16 fixes the energy-flux rate.
    p["Sq"] = 1.0 / (0.5 + (3.0 / 16.0) / (p["tauf"] - 0.5))
    return p


class Model:
    """Vectorised reference step on arrays shaped (nz, ny, nx, q)."""

    def __init__(self, dims, nx, ny, nz=1, Ra=1e6, Pr=0.71, Ma=0.1, thot=0.5, tcold=-0.5,
                 vel_x=0, vel_z=0, temp_x=1, temp_y=0, temp_z=1, use_g=True,
                 side_heated=False, legacy=False):
        self.dims, self.nx, self.ny, self.nz = dims, nx, ny, (nz if dims == 3 else 1)
        self.thot, self.tcold = thot, tcold
        self.vel_x, self.vel_z = vel_x, vel_z
        self.temp = (temp_x, temp_y, temp_z)
        self.use_g = use_g
        self.p = params(dims, nx, ny, Ra, Pr, Ma, thot, tcold, side_heated, legacy)
        if dims == 3:
            self.Ef, self.Eg = E3, E3[:7]
            self.M, self.N = basis_d3q19(self.Ef), basis_d3q7(self.Eg)
            s_e = s_nu = 1 / self.p["tauf"]
            sq = self.p["Sq"]
            self.S = np.array([0, s_e, s_e, 0, sq, 0, sq, 0, sq, s_nu, s_nu, s_nu, s_nu,
                               s_nu, s_nu, s_nu, sq, sq, sq])
            self.Q = np.array([0] + [self.p["Qk"]] * 3 + [self.p["Qnu"]] * 3)
        else:
            self.Ef, self.Eg = E2, E2[:5]
            self.M, self.N = basis_d2q9(self.Ef), basis_d2q5(self.Eg)
            s_nu, sq = 1 / self.p["tauf"], self.p["Sq"]
            self.S = np.array([0, s_nu, s_nu, 0, sq, 0, sq, s_nu, s_nu])
            self.Q = np.array([0, self.p["Qk"], self.p["Qk"], self.p["Qnu"], self.p["Qnu"]])
        self.Minv, self.Ninv = np.linalg.inv(self.M), np.linalg.inv(self.N)
        self.wf, self.wg = weights(self.Ef, dims), weights(self.Eg, dims)
        if legacy:
            self.wg = np.array([self.p["wT0"]] + [self.p["wT1"]] * (len(self.Eg) - 1))
        self.opp_f, self.opp_g = opposite(self.Ef), opposite(self.Eg)
        self.SG = 1 - 0.5 * self.p["Qk"]

    # -- equilibria ------------------------------------------------------------------------
    def feq(self, rho, u):
        eu = np.einsum("...d,qd->...q", u, self.Ef.astype(float))
        uu = (u * u).sum(axis=-1)[..., None]
        return self.wf * rho[..., None] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu)

    def meq(self, rho, u):
        m = np.einsum("...q,mq->...m", self.feq(rho, u), self.M)
        j2 = (u * u).sum(axis=-1) * rho
        if self.dims == 3:
            m[..., 2] = 3 * rho - 11 / 2 * j2
            for a, c in ((4, 0), (6, 1), (8, 2)):
                m[..., a] = -2 / 3 * rho * u[..., c]
            m[..., 10] = -0.5 * m[..., 9]
            m[..., 12] = -0.5 * m[..., 11]
            m[..., 16:19] = 0
        else:
            m[..., 2] = rho - 3 * j2
            m[..., 4] = -rho * u[..., 0]
            m[..., 6] = -rho * u[..., 1]
        return m

    def geq(self, T, u):
        # first moment of g_eq = u T: c sum_a wT_a e_ax^2 = 1
        c = 1.0 / (2.0 * self.wg[1])
        eu = np.einsum("...d,qd->...q", u, self.Eg.astype(float))
        return self.wg * T[..., None] * (1 + c * eu)

    # -- sweeps ----------------------------------------------------------------------------
    def _pull(self, post, E):
        pad = np.pad(post, ((1, 1), (1, 1), (1, 1), (0, 0))) if self.dims == 3 else \
            np.pad(post, ((0, 0), (1, 1), (1, 1), (0, 0)))
        out = np.empty_like(post)
        nz, ny, nx = post.shape[:3]
        oz = 1 if self.dims == 3 else 0
        for a, (ex, ey, ez) in enumerate(E):
            out[..., a] = pad[oz - ez:oz - ez + nz, 1 - ey:1 - ey + ny, 1 - ex:1 - ex + nx, a]
        return out

    def _walls_flow(self, f, P):
        E, opp = self.Ef, self.opp_f
        # x walls (periodic keeps the destination row), then y walls, then z walls.
        for a, e in enumerate(E):
            if e[0] == 1:
                f[:, :, 0, a] = P[:, :, -1, a] if self.vel_x else P[:, :, 0, opp[a]]
            if e[0] == -1:
                f[:, :, -1, a] = P[:, :, 0, a] if self.vel_x else P[:, :, -1, opp[a]]
        for a, e in enumerate(E):
            if e[1] == 1:
                f[:, 0, :, a] = P[:, 0, :, opp[a]]
            if e[1] == -1:
                f[:, -1, :, a] = P[:, -1, :, opp[a]]
        if self.dims == 3:
            for a, e in enumerate(E):
                if e[2] == 1:
                    f[0, :, :, a] = P[-1, :, :, a] if self.vel_z else P[0, :, :, opp[a]]
                if e[2] == -1:
                    f[-1, :, :, a] = P[0, :, :, a] if self.vel_z else P[-1, :, :, opp[a]]

    def _walls_heat(self, g, P):
        E, opp, w = self.Eg, self.opp_g, self.wg
        axes = 3 if self.dims == 3 else 2
        for ax in range(axes):
            kind = self.temp[ax]
            sl = [slice(None)] * 3
            dim = 2 - ax  # array axis of lattice axis ax
            for a, e in enumerate(E):
                for sign, wall, src_periodic, Tw in ((1, 0, -1, self.thot), (-1, -1, 0, self.tcold)):
                    if e[ax] != sign:
                        continue
                    dst = list(sl)
                    dst[dim] = wall
                    if kind == 2:
                        srcs = list(sl)
                        srcs[dim] = src_periodic
                        g[tuple(dst) + (a,)] = P[tuple(srcs) + (a,)]
                    elif kind == 0:
                        g[tuple(dst) + (a,)] = -P[tuple(dst) + (opp[a],)] + 2 * w[a] * Tw
                    else:
                        g[tuple(dst) + (a,)] = P[tuple(dst) + (opp[a],)]

    # -- state -----------------------------------------------------------------------------
    def init(self, T0):
        sh = (self.nz, self.ny, self.nx)
        self.T = np.array(T0, dtype=float).reshape(sh)
        self.rho = np.ones(sh)
        self.u = np.zeros(sh + (3,))
        self.f = self.feq(self.rho, self.u)
        self.g = self.geq(self.T, self.u)
        self.B = np.zeros(sh + (3,))
        self.Fy = np.zeros(sh)

    def step(self, n=1):
        p = self.p
        for _ in range(n):
            # collision (reference: 3DRB.F90:689-949, 2DRB.F90:675-766)
            m = np.einsum("...q,mq->...m", self.f, self.M)
            F = np.zeros(self.u.shape)
            F[..., 1] = self.rho * p["gbeta"] * (self.T - p["Tref"])
            self.Fy = F[..., 1].copy()
            meq = self.meq(self.rho, self.u)
            dmeq = self.meq(self.rho.astype(complex), self.u + 1j * F / self.rho[..., None]).imag
            mp = m - self.S * (m - meq) + (1 - self.S / 2) * dmeq
            fpost = np.einsum("...m,qm->...q", mp, self.Minv)
            # streaming + walls
            self.f = self._pull(fpost, self.Ef)
            self._walls_flow(self.f, fpost)
            # macro: 3D half force uses the new rho, 2D the force stored by collision
            self.rho = self.f.sum(axis=-1)
            j = np.einsum("...q,qd->...d", self.f, self.Ef.astype(float))
            if self.dims == 3:
                half = 0.5 * self.rho * p["gbeta"] * (self.T - p["Tref"])
            else:
                half = 0.5 * self.Fy
            j[..., 1] += half
            self.u = j / self.rho[..., None]
            if self.dims == 2:
                self.u[..., 2] = 0.0
            # collisionT with the M1G correction (3DRB.F90:1106-1198, 2DRB.F90:906-980)
            B = self.u * self.T[..., None]
            dB = B - self.B if self.use_g else np.zeros_like(B)
            if self.use_g:
                self.B = B
            nmom = np.einsum("...q,mq->...m", self.g, self.N)
            neq = np.einsum("...q,mq->...m", self.geq(self.T, self.u), self.N)
            npost = nmom - self.Q * (nmom - neq)
            nd = 3 if self.dims == 3 else 2
            npost[..., 1:1 + nd] += self.SG * dB[..., :nd]
            gpost = np.einsum("...m,qm->...q", npost, self.Ninv)
            self.g = self._pull(gpost, self.Eg)
            self._walls_heat(self.g, gpost)
            self.T = self.g.sum(axis=-1)
