"""FP64 dense-pair torch model of the WCSPH step, for gradient checks (tests only).

Restates the step the engine runs — Tait EOS, continuity, momentum with
Monaghan artificial viscosity (c_ij = c0/2, re-averaged h, Q4/Q5), gravity,
the Lennard-Jones wall force, PEC predict/correct and the Courant/force dt —
with torch autograd on CPU in float64. Pairs are found from the full distance
matrix (no neighbour list): with the cubic spline every pair with r <= 2h lies
inside the 3x3-cell stencil, so the dense set and the engine's set give the
same terms. Uniform h. Used by tests/test_gpu_adjoint.py as the reference
gradient of a final-state loss; its forward is tied to the CPU oracle there.
"""
import math

import torch

G = 9.81


class DenseWCSPH:
    def __init__(self, pA, method, h):
        f = pA['label'] == 0
        t = lambda a: torch.as_tensor(a, dtype=torch.float64)  # noqa: E731
        self.x0 = t(pA['x'][f].copy())
        self.y0 = t(pA['y'][f].copy())
        self.rho_init = t(pA['rho'][f].copy())
        self.m = t(pA['m'][f].copy())
        self.wx = t(pA['x'][~f].copy())
        self.wy = t(pA['y'][~f].copy())
        self.h = float(h)
        self.rho0, self.gamma, self.Pb = method.rho0, method.gamma, method.Pb
        self.beta = method.beta
        self.r0, self.D, self.p1, self.p2 = method.r0, method.D, method.p1, method.p2

    def rates(self, x, y, vx, vy, rho, alpha, co):
        h = self.h
        n = x.shape[0]
        with torch.no_grad():
            d2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
            mask = (d2 <= (2 * h) ** 2) & ~torch.eye(n, dtype=torch.bool)
            ii, jj = torch.nonzero(mask, as_tuple=True)
            w2 = (x[:, None] - self.wx[None, :]) ** 2 + (y[:, None] - self.wy[None, :]) ** 2
            wmask = (w2 <= self.r0 ** 2) & (w2 > 1e-24)
            wi, wj = torch.nonzero(wmask, as_tuple=True)
        B = co * co * self.rho0 / self.gamma
        p = B * ((rho / self.rho0) ** self.gamma - 1.0) + self.Pb
        slf = p / (rho * rho)
        rx, ry = x[ii] - x[jj], y[ii] - y[jj]
        d = torch.sqrt(rx * rx + ry * ry)
        q = d / h
        aC = 10.0 / (7.0 * math.pi) / (h * h)
        g = torch.where(q <= 1.0, -3.0 * q * (1.0 - 0.75 * q), -0.75 * (2.0 - q) ** 2)
        Gf = aC * g / (h * d)
        wgx, wgy = Gf * rx, Gf * ry
        vxr, vyr = vx[ii] - vx[jj], vy[ii] - vy[jj]
        dv = vxr * rx + vyr * ry
        hm = h  # 0.5 (h_i + h_ij) with uniform h
        mu = hm * dv / (d * d + 0.01 * hm * hm)
        rhob = 0.5 * (rho[ii] + rho[jj])
        PI = torch.where(dv < 0, mu * (self.beta * mu - alpha * 0.5 * co) / rhob, torch.zeros_like(dv))
        S = slf[ii] + slf[jj] + PI
        mj = self.m[jj]
        drho = torch.zeros(n, dtype=torch.float64).index_add(0, ii, mj * (vxr * wgx + vyr * wgy))
        ax = torch.zeros(n, dtype=torch.float64).index_add(0, ii, -mj * S * wgx)
        ay = torch.zeros(n, dtype=torch.float64).index_add(0, ii, -mj * S * wgy)
        # Lennard-Jones wall force (src/Equations/BoundaryForce.py:31-41)
        lx, ly = x[wi] - self.wx[wj], y[wi] - self.wy[wj]
        ld2 = lx * lx + ly * ly
        fr = self.r0 / torch.sqrt(ld2)
        fac = self.D * (fr ** self.p1 - fr ** self.p2) / ld2
        ax = ax.index_add(0, wi, fac * lx)
        ay = ay.index_add(0, wi, fac * ly) - G
        return ax, ay, drho

    def run(self, alpha, co, vx, vy, nsteps, gamma_c=0.25, gamma_f=0.25, fixed_dt=None):
        """PEC with dynamic dt from the previous rates (src/Solver.py:366-410), or a fixed dt."""
        x, y, rho = self.x0.clone(), self.y0.clone(), self.rho_init.clone()
        n = x.shape[0]
        ax = torch.zeros(n, dtype=torch.float64)
        ay = torch.zeros(n, dtype=torch.float64)
        drho = torch.zeros(n, dtype=torch.float64)
        for _ in range(nsteps):
            if fixed_dt:
                dt = fixed_dt
            else:
                a2 = torch.max(ax * ax + ay * ay)
                dt_c = gamma_c * self.h / co
                dt = dt_c if float(a2) < 1e-12 else torch.minimum(dt_c, gamma_f * torch.sqrt(self.h / a2))
            xh, yh = x + 0.5 * dt * vx, y + 0.5 * dt * vy
            vxh, vyh = vx + 0.5 * dt * ax, vy + 0.5 * dt * ay
            rhoh = rho + 0.5 * dt * drho
            ax, ay, drho = self.rates(xh, yh, vxh, vyh, rhoh, alpha, co)
            x = 2 * (x + 0.5 * dt * vxh) - x
            y = 2 * (y + 0.5 * dt * vyh) - y
            vx = 2 * (vx + 0.5 * dt * ax) - vx
            vy = 2 * (vy + 0.5 * dt * ay) - vy
            rho = 2 * (rho + 0.5 * dt * drho) - rho
        return x, y, vx, vy, rho
