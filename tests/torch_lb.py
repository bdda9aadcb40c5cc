"""Differentiable FP64 PyTorch restatement of the forward model (test reference).

A-B pull form of the evolve step on a small lattice, written with torch ops so
that torch.autograd provides reference gradients for the hand-written adjoint:
streaming with stationary halfway bounce-back (isothermal.pyx:1112-1248),
periodic / symmetric faces through the halo passes' net effect
(isothermal.pyx:509-837), Zou-He FIXED_VEL at face 1X and FIXED_DEN at face X1
(bcs.pyx:265-296 and 75-110, with _assign_f_1x/_assign_f_x1 at 487-556),
BGK / TRT / regularized collision with forcing (D3Q19.pyx:655-940).
The state is the post-collision population array f[l, site] of the interior.
"""
import numpy as np
import torch

LI = [-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1]
LJ = [-1, 0, 0, 0, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, 1]
LK = [0, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, -1, 0, 1, 0, -1, 0, 1, 0]
W = [1 / 36 if abs(a) + abs(b) + abs(c) == 2 else 2 / 36 if abs(a) + abs(b) + abs(c) == 1 else 12 / 36
     for a, b, c in zip(LI, LJ, LK)]
PERIODIC, SYMMETRIC = 1, 4


def _mirror(l, axis):
    c = [LI[l], LJ[l], LK[l]]
    c[axis] = -c[axis]
    for m in range(19):
        if [LI[m], LJ[m], LK[m]] == c:
            return m
    raise ValueError


def _resolve(pos, l, lat, bc):
    """Site and lattice vector whose post-collision value arrives at a halo
    position, or None for a wall (solid) halo; axes z, y, x in that order."""
    pos = list(pos)
    for a in (2, 1, 0):
        n = lat[a]
        if 0 <= pos[a] < n:
            continue
        lo = pos[a] < 0
        t = bc[2 * a + (0 if lo else 1)]
        if t == PERIODIC:
            pos[a] += n if lo else -n
        elif t == SYMMETRIC:
            pos[a] = 0 if lo else n - 1
            l = _mirror(l, a)
        else:
            return None
    return tuple(pos), l


class TorchLB:
    def __init__(self, phase, bc, rlx, kvisc, device="cpu"):
        self.lat = phase.shape
        nx, ny, nz = self.lat
        self.nc = nx * ny * nz
        self.rlx = rlx
        inv_tau = 1.0 / (0.5 + 3.0 * kvisc)
        self.we = inv_tau
        self.wo = 8.0 * (2.0 - inv_tau) / (8.0 - inv_tau) if rlx == 1 else inv_tau
        self.dev = device
        fluid = np.argwhere(phase != 0)
        self.fluid = torch.tensor([(i * ny + j) * nz + k for i, j, k in fluid], device=device)
        src = np.zeros((19, len(fluid)), dtype=np.int64)
        for n, (i, j, k) in enumerate(fluid):
            cell = (i * ny + j) * nz + k
            for l in range(19):
                r = _resolve((i - LI[l], j - LJ[l], k - LK[l]), l, self.lat, bc)
                if r is None or phase[r[0]] == 0:
                    src[l, n] = (18 - l) * self.nc + cell  # bounce-back
                else:
                    (a, b, c), lp = r
                    src[l, n] = lp * self.nc + (a * ny + b) * nz + c
        self.src = torch.tensor(src, device=device)
        # index into the halo-padded face arrays of faces 1X / X1: (j+1)*(nz+2) + (k+1)
        self.face_idx = torch.tensor((fluid[:, 1] + 1) * (nz + 2) + fluid[:, 2] + 1, device=device)
        self.in_1x = torch.nonzero(torch.tensor(fluid[:, 0] == 0, device=device))[:, 0]
        self.in_x1 = torch.nonzero(torch.tensor(fluid[:, 0] == nx - 1, device=device))[:, 0]
        self.rows = torch.arange(19, device=device).unsqueeze(1)

    # ---- Zou-He -----------------------------------------------------------
    @staticmethod
    def _assign_1x(f, jx, jy, jz):
        t1 = 0.5 * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7])
        t2 = 0.5 * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11])
        new = {14: f[4] + jx / 6 + t1 - 0.5 * jy, 15: f[3] + jx / 6 + t2 - 0.5 * jz, 16: f[2] + jx / 3,
               17: f[1] + jx / 6 - t2 + 0.5 * jz, 18: f[0] + jx / 6 - t1 + 0.5 * jy}
        return [new.get(l, f[l]) for l in range(19)]

    @staticmethod
    def _assign_x1(f, jx, jy, jz):
        t1 = 0.5 * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7])
        t2 = 0.5 * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11])
        new = {0: f[18] - jx / 6 + t1 - 0.5 * jy, 1: f[17] - jx / 6 + t2 - 0.5 * jz, 2: f[16] - jx / 3,
               3: f[15] - jx / 6 - t2 + 0.5 * jz, 4: f[14] - jx / 6 - t1 + 0.5 * jy}
        return [new.get(l, f[l]) for l in range(19)]

    def _zouhe(self, g, v_in, den_out):
        """FIXED_VEL at 1X with face velocity data v_in ((ny+2)*(nz+2), 3) and
        FIXED_DEN at X1 with density den_out and zero tangential velocity."""
        idx = self.in_1x
        if len(idx):
            fs = [g[l, idx] for l in range(19)]
            v = v_in[self.face_idx[idx]]
            den_kn = (fs[9] + fs[2] + fs[6] + fs[12] + fs[8] + fs[5] + fs[11] + fs[1] + fs[0] + fs[4]
                      + fs[10] + fs[7] + fs[13] + fs[3])
            jneg = fs[2] + fs[1] + fs[0] + fs[4] + fs[3]
            den = (den_kn + jneg) / (1.0 - v[:, 0])
            out = torch.stack(self._assign_1x(fs, den * v[:, 0], den * v[:, 1], den * v[:, 2]))
            g = g.index_put((self.rows, idx.unsqueeze(0)), out)
        idx = self.in_x1
        if len(idx):
            fs = [g[l, idx] for l in range(19)]
            den_kn = (fs[9] + fs[16] + fs[6] + fs[12] + fs[8] + fs[5] + fs[11] + fs[15] + fs[14] + fs[18]
                      + fs[10] + fs[7] + fs[13] + fs[17])
            jneg = den_out - den_kn
            jpos = fs[16] + fs[15] + fs[14] + fs[18] + fs[17]
            zero = torch.zeros_like(jpos)
            out = torch.stack(self._assign_x1(fs, jpos - jneg, zero, zero))
            g = g.index_put((self.rows, idx.unsqueeze(0)), out)
        return g

    # ---- collision --------------------------------------------------------
    @staticmethod
    def _project(h):
        h0, hx, hy, hz, xx, xy, xz, yy, yz, zz = h
        f = []
        for l in range(19):
            c = (LI[l], LJ[l], LK[l])
            val = h0 + 3 * (c[0] * hx + c[1] * hy + c[2] * hz)
            val = val + 4.5 * ((c[0] ** 2 - 1 / 3) * xx + (c[1] ** 2 - 1 / 3) * yy + (c[2] ** 2 - 1 / 3) * zz)
            val = val + 9 * (c[0] * c[1] * xy + c[0] * c[2] * xz + c[1] * c[2] * yz)
            f.append(W[l] * val)
        return torch.stack(f)

    def _collide(self, g, a):
        c = torch.tensor([LI, LJ, LK], dtype=g.dtype, device=g.device)  # (3,19)
        rho = g.sum(0)
        j = c @ g
        js = j + 0.5 * rho * a.T
        us = js / rho
        we, wo = self.we, self.wo
        c1, c2 = 1 - 0.5 * wo, 1 - 0.5 * we
        ax, ay, az = a[:, 0], a[:, 1], a[:, 2]
        jx, jy, jz = js
        H = [we * rho, wo * jx + c1 * rho * ax, wo * jy + c1 * rho * ay, wo * jz + c1 * rho * az,
             we * jx * jx / rho + 2 * c2 * ax * jx, we * jx * jy / rho + c2 * (ax * jy + ay * jx),
             we * jx * jz / rho + c2 * (ax * jz + az * jx), we * jy * jy / rho + 2 * c2 * ay * jy,
             we * jy * jz / rho + c2 * (ay * jz + az * jy), we * jz * jz / rho + 2 * c2 * az * jz]
        if self.rlx == 0:
            f = (1 - we) * g + self._project(H)
        elif self.rlx == 1:
            c3, c4 = 1 - 0.5 * (we + wo), 0.5 * (we - wo)
            opp = g.flip(0)
            d = c3 * g - c4 * opp
            d = torch.cat([d[:9], ((1 - we) * g[9]).unsqueeze(0), d[10:]])
            f = d + self._project(H)
        else:
            m2 = lambda p, q: (c[p] * c[q]) @ g  # noqa: E731
            h = [rho, j[0], j[1], j[2], m2(0, 0) - rho / 3, m2(0, 1), m2(0, 2), m2(1, 1) - rho / 3, m2(1, 2),
                 m2(2, 2) - rho / 3]
            H = [(1 - we) * hh + HH for hh, HH in zip(h, H)]
            f = self._project(H)
        return f, rho, us

    def run(self, f0, acc, v_in, den_out, g, steps):
        """f0: (19, nc) post-collision state; acc: (nc, 3) field; v_in:
        ((ny+2)*(nz+2), 3) face-1X velocity data. Returns (rho, u*) of the
        last step at the fluid sites (in the order of self.fluid); the last
        step's collision input and acceleration stay in self.last_input."""
        f = f0
        a = acc[self.fluid] + g.unsqueeze(0)
        rho = us = None
        for _ in range(steps):
            gsite = f.reshape(-1)[self.src]
            gsite = self._zouhe(gsite, v_in, den_out)
            self.last_input = (gsite, a)
            fpost, rho, us = self._collide(gsite, a)
            f = f.index_put((self.rows, self.fluid.unsqueeze(0)), fpost)
        return rho, us

    def dissipation(self):
        """Mean over the fluid sites of Pi_neq : Pi_neq of the last step's
        collision input, Pi_neq = sum_l c_l c_l g_l - rho u* u* - rho/3 I."""
        g, a = self.last_input
        c = torch.tensor([LI, LJ, LK], dtype=g.dtype, device=g.device)
        rho = g.sum(0)
        us = (c @ g) / rho + 0.5 * a.T
        tot = 0
        for p in range(3):
            for q in range(3):
                pn = (c[p] * c[q]) @ g - rho * us[p] * us[q] - (rho / 3 if p == q else 0)
                tot = tot + pn * pn
        return tot.mean()
