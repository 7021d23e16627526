"""Numpy restatements of the reference's DBA kernel sequences on tiny grids.

Test infrastructure, written from the reference's CUDA kernels (not from the
oracle) to pin the oracle's DBA data flow. Each class keeps the reference's
global arrays (fields, new fields, derived arrays) and runs one step kernel by
kernel, with the awake tiles of every step given by the test. A derived array is
written only where its kernel runs and read everywhere, so what an awake tile
reads from a sleeping neighbour is the reference's by construction.

  CHSeq  Calc_mu_wakeup, Update_wakeup, Swap_wakeup
         (example_Reactive_ModelB/kernel.cu:469-654, sequence :1094-1131) with the
         single-field Cahn-Hilliard chemical potential of the 2D config;
         conserved: change capture (example_Reactive_ModelB/Conserved/kernel.cu:630-702),
         reduce + Conservation + Swap_regular (:526-553, :1152-1173).
  MBSeq  the same kernels with the reference's ternary Flory-Huggins potentials,
         reaction and 27-point Laplacian (example_Reactive_ModelB/kernel.cu:290-347,
         :519-654).
  KHSeq  getPrimitive_wakeup, extrap_in_time_wakeup, calc_flux_wakeup,
         apply_flux_wakeup (example_Kelvin_Helmholtz/KH_DBA.cu:418-923, sequence
         :1436-1500), periodic in x and y, fixed dt; conserved: the mean shift of
         example_Kelvin_Helmholtz/Conserved/KH_DBA.cu:1553-1575.

The arithmetic is numpy's vector form, not the oracle's operation order, so the
tests compare with a tolerance near 1e-12 and use states where a stale and a
fresh read differ by many orders of magnitude more. `fresh=True` recomputes the
derived arrays on every cell every step: the tests use it to show that their
schedules tell the two apart.
"""
import numpy as np


def cell_mask(tiles, shape, tile):
    """Boolean cell mask of the awake tiles (tiles: flat, x fastest)."""
    if len(shape) == 2:
        ny, nx = shape
        t = np.asarray(tiles, dtype=bool).reshape(ny // tile[1], nx // tile[0])
        return np.repeat(np.repeat(t, tile[1], axis=0), tile[0], axis=1)
    nz, ny, nx = shape
    t = np.asarray(tiles, dtype=bool).reshape(nz // tile[2], ny // tile[1], nx // tile[0])
    return np.repeat(np.repeat(np.repeat(t, tile[2], axis=0), tile[1], axis=1), tile[0], axis=2)


def _sh(f, *d):
    """f at (index + d) with periodic wrap; d is given in (x, y[, z]) order."""
    out = f
    for axis, k in zip(range(f.ndim - 1, -1, -1), d):
        if k:
            out = np.roll(out, -k, axis=axis)
    return out


def lap2d(f, stencil, dx):
    """5-point, or the 9-point isotropic stencil (the 27-point one on a single z-plane)."""
    s1 = _sh(f, 1, 0) + _sh(f, -1, 0) + _sh(f, 0, 1) + _sh(f, 0, -1)
    if stencil == 5:
        return (s1 - 4.0 * f) / (dx * dx)
    s2 = _sh(f, 1, 1) + _sh(f, -1, 1) + _sh(f, 1, -1) + _sh(f, -1, -1)
    return (4.0 * s1 + s2 - 20.0 * f) / (6.0 * dx * dx)


def lap27(f, dx):
    """Laplacian of example_Reactive_ModelB/kernel.cu:290-347 (weights -64/15, 7/15, 1/10, 1/30)."""
    faces = edges = corners = 0.0
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                k = abs(a) + abs(b) + abs(c)
                if k == 1:
                    faces = faces + _sh(f, a, b, c)
                elif k == 2:
                    edges = edges + _sh(f, a, b, c)
                elif k == 3:
                    corners = corners + _sh(f, a, b, c)
    return (-64.0 / 15.0 * f + 7.0 / 15.0 * faces + 0.1 * edges + corners / 30.0) / (dx * dx)


class CHSeq:
    """dc/dt = M lap(c^3 - c - kappa lap c) - Gamma (c - cbar) in the reference's Model B sequence."""

    def __init__(self, p, c0, conserved=False, fresh=False):
        self.p, self.conserved, self.fresh = p, conserved, fresh
        self.tile = (p.tile_x, p.tile_y)
        self.c = np.array(c0, dtype=np.float64)
        self.c_new = self.c.copy()
        self.mu = np.zeros_like(self.c)      # dfdx of the reference: stored per cell
        self.metric = np.zeros_like(self.c)
        self.change = np.zeros_like(self.c)
        self.last_avg = 0.0

    def step(self, awake_tiles):
        p = self.p
        A = cell_mask(awake_tiles, self.c.shape, self.tile)
        W = np.ones_like(A) if self.fresh else A
        # Calc_mu_wakeup: awake cells only; sleeping cells keep their stored mu
        lc = lap2d(self.c, p.stencil, p.dx)
        mu = self.c * self.c * self.c - self.c - p.kappa * lc
        self.mu[W] = mu[W]
        self.metric[A] = np.abs(lc)[A]
        # Update_wakeup: awake cells; the Laplacian of mu reads every stored value
        chg = p.dt * (p.mobility * lap2d(self.mu, p.stencil, p.dx) - p.react_gamma * (self.c - p.cbar))
        self.c_new[A] = self.c[A] + chg[A]
        if self.conserved:
            self.change = np.where(A, chg, 0.0)          # 0 in inactive blocks
            self.last_avg = self.change.sum() / self.c.size
            self.c_new -= self.last_avg                   # Conservation: every cell
            self.c = self.c_new.copy()                    # Swap_regular: every cell
        else:
            self.c[A] = self.c_new[A]                     # Swap_wakeup: awake cells


class MBSeq:
    """Ternary Flory-Huggins Model B with reaction (reference parameter names)."""

    def __init__(self, p, x0, y0, z0, phi0, conserved=False, fresh=False):
        self.p, self.conserved, self.fresh = p, conserved, fresh
        self.tile = (p.tile_x, p.tile_y, p.tile_z)
        self.X = [np.array(a, dtype=np.float64) for a in (x0, y0, z0)]
        self.phi = np.array(phi0, dtype=np.float64)
        self.Xn = [a.copy() for a in self.X]
        self.df = [np.zeros_like(self.phi) for _ in range(3)]
        self.metric = np.zeros_like(self.phi)

    def step(self, awake_tiles):
        p = self.p
        x, y, z = self.X
        A = cell_mask(awake_tiles, x.shape, self.tile)
        W = np.ones_like(A) if self.fresh else A
        # Calc_mu_wakeup (example_Reactive_ModelB/kernel.cu:519-591)
        b = 1.0 - x - y - z
        common = p.chi_xb * b - p.chi_xb * x - p.chi_yb * y - p.chi_zb * z
        lap_z = lap27(z, p.dx)
        dfx = (-1.0 + 1.0 / p.r_x + p.chi_xy * y + p.chi_xz * z + common + np.log(x) / p.r_x - np.log(b)
               - p.epsx_sq * lap27(x, p.dx))
        dfy = (-1.0 + 1.0 / p.r_y + p.chi_xy * x + p.chi_yz * z + p.chi_yb * b - p.chi_xb * x - p.chi_yb * y
               - p.chi_zb * z + np.log(y) / p.r_y - np.log(b) - p.epsy_sq * lap27(y, p.dx))
        dfz = (-1.0 + 1.0 / p.r_z + p.chi_xz * x + p.chi_yz * y + p.chi_zb * b - p.chi_xb * x - p.chi_yb * y
               - p.chi_zb * z + np.log(z) / p.r_z - np.log(b) - p.p_gel * self.phi * self.phi / 2.0 / (1.0 - p.z_crit)
               - p.epsz_sq * lap_z - np.log(p.k_eq) / p.r_z)
        for arr, val in zip(self.df, (dfx, dfy, dfz)):
            arr[W] = val[W]
        self.metric[A] = np.abs(lap_z)[A]
        # Update_wakeup (:594-654): the reaction uses the cell's own potentials,
        # the Laplacians read the stored potentials of every neighbour
        dx_, dy_, dz_ = self.df
        R = p.k_0 * (np.exp(p.n_exp * p.v_x * dx_ + p.m_exp * p.v_y * dy_) - np.exp(p.v_z * dz_))
        chg = [p.dt * p.v_x * (p.mob_x * lap27(dx_, p.dx) - p.n_exp * R),
               p.dt * p.v_y * (p.mob_y * lap27(dy_, p.dx) - p.m_exp * R),
               p.dt * p.v_z * (p.mob_z * lap27(dz_, p.dx) + R)]
        for k in range(3):
            self.Xn[k][A] = self.X[k][A] + chg[k][A]
        if self.conserved:
            for k in range(3):
                avg = np.where(A, chg[k], 0.0).sum() / x.size
                self.Xn[k] -= avg
            self.X = [a.copy() for a in self.Xn]
        else:
            for k in range(3):
                self.X[k][A] = self.Xn[k][A]


class KHSeq:
    """Compressible Euler, MUSCL-Hancock + Rusanov, periodic x and y, fixed dt."""

    def __init__(self, p, U0, conserved=False, fresh_prims=False, fresh_flux=False):
        self.p, self.conserved = p, conserved
        self.fresh_prims, self.fresh_flux = fresh_prims, fresh_flux
        self.tile = (p.tile_x, p.tile_y)
        self.U = [np.array(u, dtype=np.float64) for u in U0]   # Mass, Momx, Momy, Energy
        z = np.zeros_like(self.U[0])
        self.P = [z.copy() for _ in range(4)]    # rho, vx, vy, p
        self.G = [z.copy() for _ in range(8)]    # rho_dx rho_dy p_dx p_dy vx_dx vx_dy vy_dx vy_dy
        self.PR = [z.copy() for _ in range(4)]   # rho', vx', vy', p'
        self.FX = [z.copy() for _ in range(4)]   # x-face fluxes of Mass, Momx, Momy, Energy
        self.FY = [z.copy() for _ in range(4)]   # y-face fluxes
        self.metric = z.copy()

    def _flux(self, rL, rR, pL, pR, uL, uR, vL, vR):
        """getFlux (example_Kelvin_Helmholtz/KH_DBA.cu:662-697): (mass, energy, mom-normal, mom-tangential)."""
        g = self.p.gas_gamma
        en_L = pL / (g - 1.0) + 0.5 * rL * (uL * uL + vL * vL)
        en_R = pR / (g - 1.0) + 0.5 * rR * (uR * uR + vR * vR)
        rho_s = 0.5 * (rL + rR)
        mu_s = 0.5 * (rL * uL + rR * uR)
        mv_s = 0.5 * (rL * vL + rR * vR)
        en_s = 0.5 * (en_L + en_R)
        p_s = (g - 1.0) * (en_s - 0.5 * (mu_s * mu_s + mv_s * mv_s) / rho_s)
        C = np.maximum(np.sqrt(g * pL / rL) + np.abs(uL), np.sqrt(g * pR / rR) + np.abs(uR))
        fm = mu_s - C * 0.5 * (rL - rR)
        fu = mu_s * mu_s / rho_s + p_s - C * 0.5 * (rL * uL - rR * uR)
        fv = mu_s * mv_s / rho_s - C * 0.5 * (rL * vL - rR * vR)
        fe = (en_s + p_s) * mu_s / rho_s - C * 0.5 * (en_L - en_R)
        return fm, fe, fu, fv

    def step(self, awake_tiles):
        p = self.p
        g, dx, dy, dt = p.gas_gamma, p.dx, p.dy, p.dt
        A = cell_mask(awake_tiles, self.U[0].shape, self.tile)
        all_cells = np.ones_like(A)
        WP = all_cells if self.fresh_prims else A
        WF = all_cells if self.fresh_flux else A
        # getPrimitive_wakeup
        vol = dx * dy
        mass, momx, momy, energy = self.U
        rho = mass / vol
        vx = momx / rho / vol
        vy = momy / rho / vol
        pr = (energy / vol - 0.5 * rho * (vx * vx + vy * vy)) * (g - 1.0)
        for arr, val in zip(self.P, (rho, vx, vy, pr)):
            arr[WP] = val[WP]
        # extrap_in_time_wakeup: central gradients of the stored primitives
        r, u, v, pp = self.P
        gx = lambda f: (_sh(f, 1, 0) - _sh(f, -1, 0)) / 2.0 / dx
        gy = lambda f: (_sh(f, 0, 1) - _sh(f, 0, -1)) / 2.0 / dy
        grads = (gx(r), gy(r), gx(pp), gy(pp), gx(u), gy(u), gx(v), gy(v))
        for arr, val in zip(self.G, grads):
            arr[WP] = val[WP]
        r_x, r_y, p_x, p_y, u_x, u_y, v_x, v_y = grads
        primes = (r - 0.5 * dt * (u * r_x + r * u_x + v * r_y + r * v_y),
                  u - 0.5 * dt * (u * u_x + v * u_y + (1.0 / r) * p_x),
                  v - 0.5 * dt * (u * v_x + v * v_y + (1.0 / r) * p_y),
                  pp - 0.5 * dt * (g * pp * (u_x + v_y) + u * p_x + v * p_y))
        for arr, val in zip(self.PR, primes):
            arr[WP] = val[WP]
        if p.metric_kind == 1:
            met = np.abs(r_x) + np.abs(r_y)
        else:
            met = sum(np.abs(a) for a in grads)
        self.metric[A] = met[A]
        # calc_flux_wakeup: faces from the stored primes and gradients at the
        # cell and at its +x / +y neighbour (extrapolateInSpaceToFace)
        G = self.G
        pairs = ((self.PR[0], G[0], G[1]), (self.PR[1], G[4], G[5]), (self.PR[2], G[6], G[7]),
                 (self.PR[3], G[2], G[3]))
        face = []
        for f, fdx, fdy in pairs:
            face.append((_sh(f, 1, 0) - _sh(fdx, 1, 0) * dx / 2.0, f + fdx * dx / 2.0,
                         _sh(f, 0, 1) - _sh(fdy, 0, 1) * dy / 2.0, f + fdy * dy / 2.0))
        fr, fu, fv, fp = face
        fm, fe, fmu, fmv = self._flux(fr[0], fr[1], fp[0], fp[1], fu[0], fu[1], fv[0], fv[1])
        for arr, val in zip(self.FX, (fm, fmu, fmv, fe)):
            arr[WF] = val[WF]
        # y faces: the momentum roles swap (KH_DBA.cu:738-740)
        gm, ge, gmv, gmu = self._flux(fr[2], fr[3], fp[2], fp[3], fv[2], fv[3], fu[2], fu[3])
        for arr, val in zip(self.FY, (gm, gmu, gmv, ge)):
            arr[WF] = val[WF]
        # apply_flux_wakeup: own faces and the stored fluxes of the -x / -y neighbours
        changes = []
        for q in range(4):
            ch = dt * dx * (-1.0 * self.FX[q] + _sh(self.FX[q], -1, 0) - 1.0 * self.FY[q] + _sh(self.FY[q], 0, -1))
            self.U[q][A] += ch[A]
            changes.append(np.where(A, ch, 0.0))
        if self.conserved:
            for q in range(4):
                self.U[q] -= changes[q].sum() / self.U[q].size
