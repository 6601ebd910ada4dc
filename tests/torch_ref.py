"""Pure-PyTorch FP64 restatement of the dry preqx step (one
prim_run_subcycle_c: qsplit * max(rsplit, 1) dynamics steps, then the
remap), used as the reference gradient for the HIP adjoint through
``torch.autograd``.

TEST INFRASTRUCTURE ONLY.  It restates, in vectorised torch, the same step
the CPU oracle runs (``oracle/preqx_oracle.c``) for a configuration without
moisture: dp3d from ps (prim_driver.cpp:101-110), the five CAAR stages of
u3_5stage with their DSS (prim_advance_exp.cpp:113-161,
CaarFunctorImpl.hpp:98-729), the subcycled biharmonic hyperviscosity with
the nu_top sponge (HyperviscosityFunctorImpl.cpp, biharmonic_wk_dp3d), the
dynamics steps chained through dp3d (prim_step.cpp:57-80,
prim_driver.cpp:115-130), and the remap: for rsplit = 0 the surface-pressure
update ps_v = hyai(1) ps0 + sum dp3d, for rsplit > 0 the PPM remap of v and
T (PpmRemap.hpp / VerticalRemapManager.cpp).
Tracers do not feed back on the dynamics of a dry run, so they are left out.

Fields use the device layout of the library: scalars ``(nelem, nlev, 4, 4)``
indexed [ie][lev][igp][jgp], vectors ``(nelem, nlev, 2, 4, 4)``, surface
pressure ``(nelem, 4, 4)``.
"""
import numpy as np
import torch

RGAS = 287.04
CP = 1005.0
KAPPA = RGAS / CP
RRE = 1.0 / 6.376e6


class TorchStep:
    """One dry remap interval as a differentiable torch function."""

    def __init__(self, prob, device="cpu"):
        p = prob.params
        if p.moisture or p.hypervis_scaling != 0.0 or p.remap_alg not in (1, 2):
            raise ValueError("the torch reference covers dry runs with constant-coefficient HV and remap alg 1/2")
        self.rsplit = p.rsplit
        self.ndyn = p.qsplit * max(p.rsplit, 1)
        self.remap_alg = p.remap_alg
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
        m = prob.metric
        self.L = prob.nlev
        self.dt = prob.dt
        self.D = t(m["D"]).unsqueeze(3)  # (E, 2, 2, 1, 4, 4): D(a, b)
        self.Dinv = t(m["Dinv"]).unsqueeze(3)
        self.metinv = t(m["metinv"]).unsqueeze(3)
        e2 = lambda a: t(a).unsqueeze(1)  # (E, 1, 4, 4)
        self.metdet = e2(m["metdet"])
        self.sph = e2(m["spheremp"])
        self.rsph = e2(m["rspheremp"])
        self.mp = e2(m["mp"])
        self.fcor = e2(m["fcor"])
        self.phis = e2(prob.state["phis"])
        self.dvv = t(prob.dvv)  # dvv[a][b]
        vc = prob.vcoord
        self.ps0 = vc.ps0
        self.ptop = vc.hyai[0] * vc.ps0
        self.dA = t((vc.hyai[1:] - vc.hyai[:-1]) * vc.ps0).view(1, -1, 1, 1)
        self.dB = t(vc.hybi[1:] - vc.hybi[:-1]).view(1, -1, 1, 1)
        self.hybi = t(vc.hybi[:-1]).view(1, -1, 1, 1)  # hybi(k) at interface k = 0..L-1
        self.nu, self.nu_s, self.nu_p, self.nu_top = p.nu, p.nu_s, p.nu_p, p.nu_top
        self.nsub = p.hypervis_subcycle
        if p.nu != p.nu_div:
            self.nr1 = self.nr2 = p.nu_div / p.nu
        else:
            self.nr1 = self.nr2 = 1.0
        # DSS: every GLL copy of a geometric node sums into one slot
        key = np.round(prob.mesh.pos * 1e9).astype(np.int64).reshape(-1, 3)
        _, inv = np.unique(key, axis=0, return_inverse=True)
        self.node = torch.as_tensor(inv.reshape(-1), device=device)
        self.nnode = int(inv.max()) + 1

    # ------------------------------------------------------ element operators
    def _rowc(self, s):  # sum_k dvv[jg, k] s[ig, k]
        return torch.einsum("...ik,jk->...ij", s, self.dvv)

    def _colc(self, s):  # sum_k dvv[ig, k] s[k, jg]
        return torch.einsum("...kj,ik->...ij", s, self.dvv)

    def grad(self, s):
        Di = self.Dinv
        v0 = self._rowc(s) * RRE
        v1 = self._colc(s) * RRE
        return Di[:, 0, 0] * v0 + Di[:, 0, 1] * v1, Di[:, 1, 0] * v0 + Di[:, 1, 1] * v1

    def div(self, a, b):
        Di, md = self.Dinv, self.metdet
        gv0 = (Di[:, 0, 0] * a + Di[:, 1, 0] * b) * md
        gv1 = (Di[:, 0, 1] * a + Di[:, 1, 1] * b) * md
        return (self._rowc(gv0) + self._colc(gv1)) * (1.0 / md * RRE)

    def vort(self, u, v):
        D, md = self.D, self.metdet
        c0 = D[:, 0, 0] * u + D[:, 0, 1] * v
        c1 = D[:, 1, 0] * u + D[:, 1, 1] * v
        return (self._rowc(c1) - self._colc(c0)) * (1.0 / md * RRE)

    def _wk(self, x):
        """(A, B) with A[n,m] = sum_j x[n,j] dvv[j,m], B[n,m] = sum_j x[j,m] dvv[j,n]"""
        return torch.einsum("...nj,jm->...nm", x, self.dvv), torch.einsum("...jm,jn->...nm", x, self.dvv)

    def div_wk(self, v0, v1):
        Di = self.Dinv
        a = Di[:, 0, 0] * v0 + Di[:, 1, 0] * v1
        b = Di[:, 0, 1] * v0 + Di[:, 1, 1] * v1
        A, _ = self._wk(self.sph * a)
        _, B = self._wk(self.sph * b)
        return -(A + B) * RRE

    def lap(self, f):
        return self.div_wk(*self.grad(f))

    def grad_wk_testcov(self, s):
        mi, md, D = self.metinv, self.metdet, self.D
        A, B = self._wk(self.mp * s)
        b0 = -md * (mi[:, 0, 0] * A + mi[:, 0, 1] * B)
        b1 = -md * (mi[:, 1, 0] * A + mi[:, 1, 1] * B)
        return (D[:, 0, 0] * b0 + D[:, 1, 0] * b1) * RRE, (D[:, 0, 1] * b0 + D[:, 1, 1] * b1) * RRE

    def curl_wk_testcov(self, s):
        D = self.D
        A, B = self._wk(self.mp * s)
        sb0, sb1 = -B, A
        return (D[:, 0, 0] * sb0 + D[:, 1, 0] * sb1) * RRE, (D[:, 0, 1] * sb0 + D[:, 1, 1] * sb1) * RRE

    def vlap(self, u0, u1, nr):
        d = self.div(u0, u1)
        if nr > 0 and nr != 1.0:
            d = d * nr
        g0, g1 = self.grad_wk_testcov(d)
        c0, c1 = self.curl_wk_testcov(self.vort(u0, u1))
        f = 2.0 * self.sph * RRE * RRE
        return f * u0 + (g0 - c0), f * u1 + (g1 - c1)

    # ------------------------------------------------------------------ DSS
    def assemble(self, x):
        """sum of all copies of every node, x: (E, nl, 4, 4)"""
        E, nl = x.shape[0], x.shape[1]
        flat = x.permute(0, 2, 3, 1).reshape(E * 16, nl)
        s = torch.zeros(self.nnode, nl, dtype=x.dtype, device=x.device).index_add(0, self.node, flat)
        return s[self.node].reshape(E, 4, 4, nl).permute(0, 3, 1, 2)

    def dss_r(self, x):
        return self.rsph * self.assemble(x)

    # ----------------------------------------------------------------- CAAR
    def caar(self, am, bn, h):
        """pre-DSS outputs sph*(x(nm1) + h*RHS(x(n0))), rsplit = 0, dry"""
        vm, Tm, dpm = am
        v, T, dp = bn
        u0, u1 = v[:, :, 0], v[:, :, 1]
        L = self.L
        divv = self.div(u0 * dp, u1 * dp)
        excl = lambda x: torch.cumsum(x, 1) - x  # sum_{j<k}
        p = self.ptop + excl(dp) + 0.5 * dp
        r = RGAS * T * (dp * 0.5 / p)
        suffix = torch.flip(torch.cumsum(torch.flip(r, [1]), 1), [1]) - r  # sum_{j>k}
        phi = self.phis + 2.0 * suffix + r
        gp0, gp1 = self.grad(p)
        omb = (u0 * gp0 + u1 * gp1 - (excl(divv) + 0.5 * divv)) / p
        stot = divv.sum(1, keepdim=True)
        eta = self.hybi * stot - excl(divv)
        eta = torch.cat([torch.zeros_like(eta[:, :1]), eta[:, 1:]], 1)  # eta(0) = 0
        if self.rsplit > 0:
            eta = torch.zeros_like(eta)  # Lagrangian levels: no vertical flux
        eta_up = torch.cat([eta[:, 1:], torch.zeros_like(eta[:, :1])], 1)  # eta(k+1), 0 at the bottom
        facp = 0.5 * (1 / dp) * eta_up
        facm = 0.5 * (1 / dp) * eta

        def vadv(f):
            fp = torch.cat([f[:, 1:], f[:, -1:]], 1)
            fm = torch.cat([f[:, :1], f[:, :-1]], 1)
            return facp * (fp - f) + facm * (f - fm)

        gT0, gT1 = self.grad(T)
        ttens = -vadv(T) - (u0 * gT0 + u1 * gT1) + KAPPA * T * omb
        Tout = self.sph * (Tm + h * ttens)
        coef = RGAS * (T / p)
        e0, e1 = self.grad(0.5 * (u0 * u0 + u1 * u1) + phi)
        e0 = coef * gp0 + e0
        e1 = coef * gp1 + e1
        vort = self.vort(u0, u1) + self.fcor
        du0 = -e0 - vadv(u0) + u1 * vort
        du1 = -e1 - vadv(u1) - u0 * vort
        vout = self.sph.unsqueeze(2) * (vm + h * torch.stack([du0, du1], 2))
        dpout = self.sph * (dpm - h * (eta_up + divv - eta))
        return vout, Tout, dpout

    def stage(self, am, bn, h):
        v, T, dp = self.caar(am, bn, h)
        v = torch.stack([self.dss_r(v[:, :, 0]), self.dss_r(v[:, :, 1])], 2)
        return v, self.dss_r(T), self.dss_r(dp)

    # -------------------------------------------------------- hyperviscosity
    def hv_cycle(self, x, dt):
        v, T, dp = x
        u0, u1 = v[:, :, 0], v[:, :, 1]
        l0, l1 = self.vlap(u0, u1, self.nr1)
        lt, ld = self.lap(T), self.lap(dp)
        l0, l1, lt, ld = (self.dss_r(a) for a in (l0, l1, lt, ld))
        l0, l1 = self.vlap(l0, l1, self.nr2)
        lt, ld = self.lap(lt), self.lap(ld)
        v0 = -self.nu * l0
        v1 = -self.nu * l1
        tt = -self.nu_s * lt
        dpt = -self.nu_p * ld
        if self.nu_top > 0:
            nt = min(3, self.L)
            sc = torch.tensor([4.0, 2.0, 1.0][:nt], dtype=torch.float64, device=T.device).view(1, nt, 1, 1) \
                * self.nu_top
            t0, t1 = self.vlap(u0[:, :nt], u1[:, :nt], 1.0)
            top = lambda a, b: torch.cat([a[:, :nt] + sc * b, a[:, nt:]], 1)
            v0, v1 = top(v0, t0), top(v1, t1)
            tt = top(tt, self.lap(T[:, :nt]))
            dpt = top(dpt, self.lap(dp[:, :nt]))
        dpt = dpt * dt + dp * self.sph
        v0, v1, tt, dpt = (self.assemble(a) for a in (v0, v1, tt, dpt))
        r = self.rsph
        vt0, vt1 = dt * v0 * r, dt * v1 * r
        n0, n1 = u0 + vt0, u1 + vt1
        heating = vt0 * n0 + vt1 * n1
        Tn = T + dt * tt * r - heating / CP
        return torch.stack([n0, n1], 2), Tn, dpt * r

    # ----------------------------------------------------------------- step
    def dyn(self, u0):
        """one dynamics step: RK5 stages (u3_5stage) and the hyperviscosity
        subcycles, (v, T, dp3d) -> (v, T, dp3d)"""
        dt = self.dt
        u1 = self.stage(u0, u0, dt / 5.0)
        u2 = self.stage(u0, u1, dt / 5.0)
        u3 = self.stage(u0, u2, dt / 3.0)
        u4 = self.stage(u0, u3, 2.0 * dt / 3.0)
        w = tuple((5.0 * a - b) / 4.0 for a, b in zip(u1, u0))
        x = self.stage(w, u4, 3.0 * dt / 4.0)
        for _ in range(self.nsub):
            x = self.hv_cycle(x, dt / self.nsub)
        return x

    def step(self, v, T, ps):
        x = (v, T, self.dA + self.dB * ps.unsqueeze(1))
        for _ in range(self.ndyn):
            x = self.dyn(x)
        vn, Tn, dpn = x
        if self.rsplit == 0:
            return vn, Tn, self.ptop + dpn.sum(1)
        return self.remap_states(vn, Tn, dpn)

    # ------------------------------------------------------ vertical remap
    def remap_states(self, v, T, src):
        """PPM remap of v and T (as extensive v*dp, T*dp) from the Lagrangian
        thicknesses ``src`` to the hybrid levels of ps = hyai(0) ps0 + sum(src)
        (PpmRemap.hpp:288-673, the oracle's ppm_grids / ppm_remap_column)."""
        L = self.L
        ps = self.ptop + src.sum(1)
        tgt = self.dA + self.dB * ps.unsqueeze(1)
        # columns: (ncol, L)
        col = lambda x: x.permute(0, 2, 3, 1).reshape(-1, L)
        uncol = lambda x: x.reshape(src.shape[0], 4, 4, L).permute(0, 3, 1, 2)
        dx, tg = col(src), col(tgt)
        fields = [col(v[:, :, 0]), col(v[:, :, 1]), col(T)]
        out = [uncol(self._ppm_column(dx, tg, f * dx) / tg) for f in fields]
        return torch.stack([out[0], out[1]], 2), out[2], ps

    def _ppm_column(self, dx, tg, var):
        L = self.L
        z = torch.zeros_like(dx[:, :1])
        pio = torch.cat([z, torch.cumsum(dx, 1)], 1)               # pio(0..L)
        pio = torch.cat([pio, pio[:, -1:] + 1.0], 1)                # pio(L+1)
        pin = torch.cat([z, torch.cumsum(tg, 1)], 1)                # pin(0..L)
        pin = torch.cat([pin[:, :L], pio[:, L:L + 1]], 1)           # pin(L) = pio(L)
        # mirrored ghosts: DX(i), i = 0..L+3, is dx at Fortran cell i-1
        D = torch.cat([dx[:, 1:2], dx[:, 0:1], dx, dx[:, L - 1:L], dx[:, L - 2:L - 1]], 1)
        # target search on detached values (kid is piecewise constant)
        with torch.no_grad():
            pio_d, pin_d = pio.detach(), pin.detach()
            kid = torch.zeros(dx.shape[0], L, dtype=torch.long, device=dx.device)
            for k in range(L):
                # kk = smallest index >= k+1 with pio(kk-1) > pin(k+1), capped
                cond = pio_d[:, k:L + 2] <= pin_d[:, k + 1:k + 2]       # pio(kk-1) for kk-1 = k..L+1
                # count consecutive trues from the start
                run = torch.cumprod(cond.to(torch.int64), 1).sum(1)
                kk = k + 1 + run - 1
                kk = torch.where(kk == L + 1, torch.full_like(kk, L), kk)
                kid[:, k] = kk - 1
        rows = torch.arange(dx.shape[0], device=dx.device)
        DXk = lambda j: D[:, j]
        # ppmdx coefficients, j = 0..L+1 (and 0..L)
        G = [[None] * (L + 2) for _ in range(10)]
        for j in range(L + 2):
            G[0][j] = DXk(j + 1) / (DXk(j) + DXk(j + 1) + DXk(j + 2))
            G[1][j] = (2.0 * DXk(j) + DXk(j + 1)) / (DXk(j + 1) + DXk(j + 2))
            G[2][j] = (DXk(j + 1) + 2.0 * DXk(j + 2)) / (DXk(j) + DXk(j + 1))
        for j in range(L + 1):
            G[3][j] = DXk(j + 1) / (DXk(j + 1) + DXk(j + 2))
            G[4][j] = 1.0 / (DXk(j) + DXk(j + 1) + DXk(j + 2) + DXk(j + 3))
            G[5][j] = (2.0 * DXk(j + 1) * DXk(j + 2)) / (DXk(j + 1) + DXk(j + 2))
            G[6][j] = (DXk(j) + DXk(j + 1)) / (2.0 * DXk(j + 1) + DXk(j + 2))
            G[7][j] = (DXk(j + 3) + DXk(j + 2)) / (2.0 * DXk(j + 2) + DXk(j + 1))
            G[8][j] = DXk(j + 1) * (DXk(j) + DXk(j + 1)) / (2.0 * DXk(j + 1) + DXk(j + 2))
            G[9][j] = DXk(j + 2) * (DXk(j + 2) + DXk(j + 3)) / (DXk(j + 1) + 2.0 * DXk(j + 2))
        ao = var / dx
        if self.remap_alg == 3:
            raise ValueError("remap alg 3 is not accepted by the C-ABI")
        A = torch.cat([ao[:, 1:2], ao[:, 0:1], ao, ao[:, L - 1:L], ao[:, L - 2:L - 1]], 1)  # a(j) = A[:, j+1]
        a = lambda j: A[:, j + 1]
        dma = []
        for j in range(L + 2):
            dp_, dm_ = a(j + 1) - a(j), a(j) - a(j - 1)
            da = G[0][j] * (G[1][j] * dp_ + G[2][j] * dm_)
            m3 = torch.minimum(torch.minimum(torch.abs(da), 2.0 * torch.abs(dm_)), 2.0 * torch.abs(dp_))
            # Homme::min(a, b, c) selection with its tie rule
            bc = torch.where(2.0 * torch.abs(dm_) < 2.0 * torch.abs(dp_), 2.0 * torch.abs(dm_), 2.0 * torch.abs(dp_))
            m3 = torch.where(torch.abs(da) < bc, torch.abs(da), bc)
            dma.append(torch.where(dp_ * dm_ > 0.0, m3 * torch.sign(da), torch.zeros_like(da)))
        ai = []
        for j in range(L + 1):
            d1 = a(j + 1) - a(j)
            ai.append(a(j) + G[3][j] * d1 + G[4][j] * (G[5][j] * (G[6][j] - G[7][j]) * d1 - G[8][j] * dma[j + 1]
                                                      + G[9][j] * dma[j]))
        c0, c1, c2 = [], [], []
        for j in range(1, L + 1):
            al, ar, aa = ai[j - 1], ai[j], a(j)
            flat = (ar - aa) * (aa - al) <= 0.0
            al = torch.where(flat, aa, al)
            ar = torch.where(flat, aa, ar)
            c2a = (ar - al) * (aa - (al + ar) / 2.0) > (ar - al) * (ar - al) / 6.0
            al = torch.where(c2a, 3.0 * aa - 2.0 * ar, al)
            c3a = (ar - al) * (aa - (al + ar) / 2.0) < -(ar - al) * (ar - al) / 6.0
            ar = torch.where(c3a, 3.0 * aa - 2.0 * al, ar)
            x0, x1, x2 = 1.5 * aa - (al + ar) / 4.0, ar - al, 3.0 * (-2.0 * aa + (al + ar))
            if self.remap_alg == 2 and (j <= 2 or j >= L - 1):
                x0, x1, x2 = aa, torch.zeros_like(aa), torch.zeros_like(aa)
            c0.append(x0)
            c1.append(x1)
            c2.append(x2)
        C0, C1, C2 = torch.stack(c0, 1), torch.stack(c1, 1), torch.stack(c2, 1)
        mass = torch.cat([z, torch.cumsum(var, 1)], 1)              # mass(kk+1) = sum_{j<kk} var
        outs = []
        prev = None
        for k in range(L):
            kk = kid[:, k]
            dxk = dx[rows, kk]
            z2 = (pin[:, k + 1] - (pio[rows, kk] + pio[rows, kk + 1]) * 0.5) / dxk
            x1, x2 = -0.5, z2
            integ = (C0[rows, kk] * (x2 - x1) + C1[rows, kk] * (x2 * x2 - x1 * x1) / 2.0) + \
                C2[rows, kk] * (x2 * x2 * x2 - x1 * x1 * x1) / 3.0
            m = mass[rows, kk] + integ * dxk
            outs.append(m if prev is None else m - prev)
            prev = m
        return torch.stack(outs, 1)

    def run(self, v, T, ps, nsteps):
        for _ in range(nsteps):
            v, T, ps = self.step(v, T, ps)
        return v, T, ps


def initial_state(prob, device="cpu"):
    """(v, T, ps) of time level n0 of ``prob`` in the device layout."""
    s = prob.state
    n0 = prob.time_level[1] - 1
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    return t(s["v"][:, n0]), t(s["T"][:, n0]), t(s["ps_v"][:, n0])
