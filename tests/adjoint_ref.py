"""Torch expressions of the three pair passes that the adjoint runs as HIP
kernels (SPH force, fluid-particle drag, DEM contacts), differentiated by
torch autograd, and a pair provider backed by the CPU oracle.  Test
infrastructure: these are the checkers of the hand-written backward kernels
and of the differentiable step on the CPU."""
from __future__ import annotations

import torch

import oracle
import sphdem as sd
import sphdem_adjoint as sa


class TorchPasses:
    """The pair passes as torch expressions over a pair set: calcForce
    (sphIncompress.h:899-972 with sph.h:579-666, 692-734), calcOmega
    (sph.h:437-452), calcDddtDudt (sphIncompress.h:783-851), the fluid-side
    coupling sums (sph.h:1303-1362), calcFPI (sph.h:1708-1792 with 1575-1631)
    and calcDEMContacts (sph.h:1676-1706)."""

    def omega(self, C, ps, r, h, dens):
        c, m, n = C.m.P.wcon, C.m.mass, C.m.n
        a, b, dx, rr = sa._dx(ps, r)
        acc = torch.zeros(n, dtype=r.dtype, device=r.device).index_add(
            0, a, m[b] * sa.kPdWh(c, rr / h[a], h[a]))
        acc = acc + m * sa.kPdWh(c, torch.zeros_like(h), h)
        om = 1.0 + h * acc / (dens * 3.0)
        return torch.where(C.m.grp, om, torch.ones_like(om))

    def dddt(self, C, ps, r, vhat, h, omega):
        c, m = C.m.P.wcon, C.m.mass
        a, b, dx, rr = sa._dx(ps, r)
        F = sa.kF(c, rr / h[a], h[a]) / omega[a]
        vdr = ((vhat[a] - vhat[b]) * dx).sum(1) * F
        return torch.zeros_like(h).index_add(0, a, m[b] * vdr)

    def coupling(self, C, ps, r, dens, fdrag):
        P, m = C.m.P, C.m.mass
        Rc = P.coupling_radius
        a, d, dx, rr = sa._dx(ps, r)
        W = sa.kW(P.wcon, rr * (2.0 / Rc), Rc / 2.0)
        z = torch.zeros_like(dens)
        por = z.index_add(0, a, P.porosity_incr * W)
        shep = z.index_add(0, d, (m[a] / dens[a]) * W)
        safe = torch.where(shep[d] > 0, shep[d], torch.ones_like(shep[d]))
        w = torch.where(shep[d] > 0, W / safe, torch.zeros_like(W))
        drag = torch.zeros_like(r).index_add(0, a, -(m[d] / dens[a] * w)[:, None] * fdrag[d])
        return por, drag

    def force(self, C, ps, r, vhat, h, dens, omega, pdr2, visc):
        m, P = C.m.mass, C.m.P
        a, b, dx, rr = sa._dx(ps, r)
        r2 = (dx * dx).sum(1)
        Fa = sa.kF(P.wcon, rr / h[a], h[a])
        Fb = sa.kF(P.wcon, rr / h[b], h[b])
        ia = 1.0 / omega[a]
        vis = 2.0 * visc * P.dens / (dens[a] * dens[b])
        d = r2 * (Fa + Fb) * 0.5 * ia
        hav = 0.5 * (h[a] + h[b])
        den = r2 + 0.01 * hav * hav
        A = torch.where(C.m.sph[a], m[b] * vis * d / den, torch.zeros_like(d))
        fv = torch.zeros_like(r).index_add(0, a, A[:, None] * (vhat[a] - vhat[b]))
        B = -m[b] * (pdr2[a] * Fa + pdr2[b] * Fb) * ia
        fp = torch.zeros_like(r).index_add(0, a, B[:, None] * dx)
        return fv, fp

    def fpi(self, C, ps, r, vhat, h, dens, fpfb, fv, por, visc, drag_a, drag_b):
        m, P = C.m.mass, C.m.P
        p, nb, dx, rr = sa._dx(ps, r)
        W = sa.kW(P.wcon, rr / h[nb], h[nb])
        dvW = (m[nb] / dens[nb]) * W
        mW = m[nb] * W
        z, z3 = torch.zeros_like(h), torch.zeros_like(r)
        S = z.index_add(0, p, dvW)
        U = z3.index_add(0, p, dvW[:, None] * vhat[nb])
        Pn = z.index_add(0, p, dvW * por[nb])
        Gp = z3.index_add(0, p, mW[:, None] * fpfb[nb])
        Gv = z3.index_add(0, p, (mW + dvW)[:, None] * fv[nb])
        on = C.m.dem & (S > 0.5)
        Ss = torch.where(on, S, torch.ones_like(S))
        fvel = U / Ss[:, None]
        pr = Pn / Ss
        dv = pr[:, None] * (fvel - vhat)
        vd2 = (dv * dv).sum(1)
        nz = on & (vd2 != 0)
        vds = torch.sqrt(torch.where(nz, vd2, torch.ones_like(vd2)))
        R = P.dem_radius
        Re = (2.0 * R / visc) * vds
        L10 = torch.log10(Re)
        chi = drag_a - drag_b * torch.exp(-(1.5 - L10) ** 2 / 2.0)
        Cd = (0.63 + 4.8 / torch.sqrt(Re)) ** 2
        s = 0.5 * (P.pi * R * R) * Cd * vds
        pw = torch.where(nz, pr, torch.ones_like(pr)) ** (-chi)
        fd = torch.where(nz[:, None], (s * pw)[:, None] * dv, torch.zeros_like(dv))
        out = (P.dem_vol / m)[:, None] * (Gv + Gp) / Ss[:, None] + (P.dens / m)[:, None] * fd
        return torch.where(on[:, None], out, torch.zeros_like(out))

    def contacts(self, C, ps, r, vhat, dem_k, dem_gamma):
        m, P = C.m.mass, C.m.P
        a, b, dx, rr = sa._dx(ps, r)
        ov = 2.0 * P.dem_radius - rr
        on = ov > 0
        n = dx / rr[:, None]
        od = -((vhat[a] - vhat[b]) * n).sum(1)
        sN = torch.where(on, (dem_k * ov + dem_gamma * od) / m[a], torch.zeros_like(ov))
        return torch.zeros_like(r).index_add(0, a, sN[:, None] * n)


class OraclePairs:
    """Pair sets of the oracle's searches for given positions and h."""

    def __init__(self, case: sd.Case):
        self.o = oracle.Oracle(case)

    def __call__(self, r, h, kinds):
        self.o.set("r", r.detach().cpu().numpy())
        self.o.set("h", h.detach().cpu().numpy())
        self.o.reset()
        return {k: torch.from_numpy(self.o.pairs(k)).to(r.device) for k in kinds}

    def close(self):
        self.o.close()
