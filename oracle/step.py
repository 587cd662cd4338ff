"""CPU restatement of the library's immersed-boundary timestep, composed from
the oracle's pieces (test infrastructure only).

The reference's own step (ib::solver, test/ib-whole.cc:226-275) is not part
of the reference, so this follows the library's documented scheme:
  y = E^T X (oracle/rbf.py), f = -k (y - y_ref) and / or the membrane
  models of oracle/forces.py (tension against the undeformed shape's
  geometry; bending from H at the data sites, differentiated at the sample
  sites by the sample derivative operators),
  F = spread(y, f) (oracle/orc.py; two cell types spread separately and added),
  H = advection(u) (oracle/fd.py),
  rhs = dt I~ (F / rho - H) + dt nu (b + L u),  (I~ - dt nu / 2 L) du = rhs
  (ins/diffusion.h:46-56, sparse direct solves),
  optionally (pseudo_pressure) rhs += dt nu / 2 dt B_i (G_i p) with
  B_i = ins::boundary(grid_i, i) at order 2 and p the previous pressure
  (the w = u + k grad phi wall data of ins/boundary.h:16-18),
  L_p phi = D u* / dt on the Neumann cell-centred grid (mean-free; pinned
  phi_0 = 0 then the mean removed), u = u* - dt G phi, p = phi,
  X += dt interpolate(u, X).
"""
import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
CENTRE = [0.5, 0.5, 0.5]


class Step:
    """spec: [(length, solid, lower, upper)] x 3 as for orc.grid_axes."""

    def __init__(self, orc, fd, rbf, spec, refinement, dt, nu, rho, delta, types, pseudo_pressure=False,
                 forces=None, solver="direct"):
        """solver "direct": sparse LU of every operator; "iterative": the
        Helmholtz solves by the CSR multigrid (solvers/multigrid.h) to
        1e-14 relative and the pressure Poisson by its exact spectral
        solution (Fourier modes on periodic axes, DCT-II modes on Neumann
        walls: the eigenvectors of the cell-centred second difference), for
        grids too large for sparse LU."""
        self.orc, self.fd, self.rbf = orc, fd, rbf
        self.forces_mod = forces
        self.spec, self.ref = spec, refinement
        self.dt, self.nu, self.rho, self.delta = dt, nu, rho, delta
        self.types = types                  # dicts: E, Et, Ep, nd, ns, nc, w, ref (3, nc, ns), k; membrane
                                            # types add seconds (Euu, Euv, Evv), data_ops (5), reference (3, nd),
                                            # membrane {tension, shear, bulk, bending, preferred}
        self.pseudo_pressure = pseudo_pressure
        pspec = [(L, s, (0.0, 1.0) if s else lo, (0.0, 1.0) if s else up) for (L, s, lo, up) in spec]
        self.pspec = pspec
        self.axes_of = lambda align: self._axes(spec, align)
        self.paxes_of = lambda align: self._axes(pspec, align)
        self.mac = orc.mac_axes(spec, refinement)
        self.uaxes = [self.axes_of(a) for a in MAC]
        self.paxes = self.paxes_of(CENTRE)
        lam = dt * nu / 2
        self.Bw = [fd.ins_boundary(ax, i, 2) for i, ax in enumerate(self.uaxes)]
        self.I = [fd.identity(ax, 2) for ax in self.uaxes]
        self.L = [fd.laplacian(ax) for ax in self.uaxes]
        self.solver = solver
        if solver == "direct":
            self.helm = [spla.splu(sp.csc_matrix(fd.operator(ax, 1.0, -lam))) for ax in self.uaxes]
            Lp = sp.lil_matrix(fd.laplacian(self.paxes))
            Lp[0, :] = 0
            Lp[0, 0] = 1.0
            self.pois = spla.splu(sp.csc_matrix(Lp))
        else:
            self.helm_mg = [fd.Multigrid(lambda r, al=al: self._axes_ref(spec, r, al), refinement, 1.0, -lam,
                                         order="x-fastest") for al in MAC]
            self.helm_A = [fd.operator(ax, 1.0, -lam) for ax in self.uaxes]

    def _axes(self, spec, align):
        a = self.orc.grid_axes(spec, self.ref, align)
        return [a[j] for j in range(3)]

    def _axes_ref(self, spec, refinement, align):
        a = self.orc.grid_axes(spec, refinement, list(align))
        return [a[j] for j in range(3)]

    def _helmholtz(self, i, rhs):
        if self.solver == "direct":
            return self.helm[i].solve(rhs)
        mg, A = self.helm_mg[i], self.helm_A[i]
        x = np.zeros_like(rhs)
        r = rhs.copy()
        for _ in range(100):
            if np.linalg.norm(r) <= 1e-14 * np.linalg.norm(rhs):
                break
            e = mg.vcycle(r)
            x = x + e
            r = rhs - A @ x
        return x

    def _poisson(self, div):
        """Mean-free phi with L_p phi = div (div mean-free)."""
        if self.solver == "direct":
            d = div.copy()
            d[0] = 0.0
            phi = self.pois.solve(d)
            return phi - phi.mean()
        ax = self.paxes
        n = [a.points for a in ax]
        f = div.reshape(n[2], n[1], n[0])                # z, y, x
        lam = np.zeros((n[2], n[1], n[0]))
        coef = f
        for j, a in enumerate(ax):
            axis = 2 - j
            k = np.arange(n[j])
            r2 = a.resolution ** 2
            if a.solid:
                coef = sfft.dct(coef, type=2, norm="ortho", axis=axis)
                ev = -4 * r2 * np.sin(np.pi * k / (2 * n[j])) ** 2
            else:
                coef = sfft.fft(coef, axis=axis)
                ev = -4 * r2 * np.sin(np.pi * k / n[j]) ** 2
            shape = [1, 1, 1]
            shape[axis] = n[j]
            lam = lam + ev.reshape(shape)
        lam[0, 0, 0] = 1.0
        coef = coef / lam
        coef[0, 0, 0] = 0.0
        for j, a in enumerate(ax):
            axis = 2 - j
            if a.solid:
                coef = sfft.idct(coef, type=2, norm="ortho", axis=axis)
            else:
                coef = sfft.ifft(coef, axis=axis)
        phi = np.real(coef).reshape(-1)
        return phi - phi.mean()

    def geometry(self, xs):
        out = []
        for T, X in zip(self.types, xs):
            Xm = X.reshape(3 * T["nc"], T["nd"])
            y, t1, t2 = (self.rbf.evaluate(M, T["nd"], Xm).reshape(3, T["nc"], T["ns"])
                         for M in (T["E"], T["Et"], T["Ep"]))
            out.append((y, t1, t2))
        return out

    def forces(self, T, X, geo):
        """Forces (3, nc, ns) of one cell type at geometry geo = (y, t1, t2)."""
        y = geo[0]
        f = np.zeros_like(y)
        if T.get("k", 0.0):
            f = f - T["k"] * (y - T["ref"])
        m = T.get("membrane")
        if not m:
            return f
        fo = self.forces_mod
        nd, ns, nc = T["nd"], T["ns"], T["nc"]
        P = ns * nc
        ev = lambda M, Z, K=nd: self.rbf.evaluate(M, K, Z)

        def blocks(mats, Z, rows):
            # (5, 3, rows * cells) from operators applied to Z (3 * cells, nd)
            cells = Z.shape[0] // 3
            return np.stack([ev(M, Z).reshape(3, cells * rows) for M in mats])

        sample_ops = [T["Et"], T["Ep"]] + list(T["seconds"])
        Xm = X.reshape(3 * nc, nd)
        g = blocks(sample_ops, Xm, ns)
        t, tt = g[:2], g[2:]
        ref = T["reference"].reshape(3, nd)
        g0 = blocks(sample_ops, ref, ns)
        t0, tt0 = g0[:2], g0[2:]
        c0 = np.cross(t0[0], t0[1], axis=0)
        s0 = T["w"] * np.sqrt((c0 * c0).sum(0))
        out = np.zeros((3, P))
        if m.get("tension", "none") != "none":
            law = getattr(fo, m["tension"])(m["shear"], m["bulk"])
            out += fo.tension(law, t, tt, t0, tt0, s0)
        if m.get("bending", 0.0):
            cr = np.cross(t[0], t[1], axis=0)
            det = np.sqrt((cr * cr).sum(0))
            n, s = cr / det, np.tile(T["w"], nc) * det
            gd = blocks(T["data_ops"], Xm, nd)
            nd_ = np.cross(gd[0], gd[1], axis=0)
            nd_ /= np.sqrt((nd_ * nd_).sum(0))
            if m.get("preferred"):
                gd0 = blocks(T["data_ops"], ref, nd)
                n0 = np.cross(gd0[0], gd0[1], axis=0)
                n0 /= np.sqrt((n0 * n0).sum(0))
                h = fo.mean_curvature_sites(gd[:2], gd[2:], nd_, gd0[:2], gd0[2:], n0)
                n0s = np.cross(t0[0], t0[1], axis=0)
                n0s /= np.sqrt((n0s * n0s).sum(0))
                ref_geo = (t0, tt0, n0s)
            else:
                h = fo.mean_curvature_sites(gd[:2], gd[2:], nd_)
                ref_geo = (None, None, None)
            hm = h.reshape(nc, nd)
            hd = np.stack([ev(M, hm).reshape(P) for M in sample_ops])
            out += fo.bending(m["bending"], t, tt, n, s, hd, *ref_geo)
        return f + out.reshape(3, nc, ns)

    def step(self, u, p, xs, boundary=None):
        fd, orc, dt, nu = self.fd, self.orc, self.dt, self.nu
        geo = self.geometry(xs)
        F = [np.zeros(self.I[i].shape[0]) for i in range(3)]
        for T, X, g in zip(self.types, xs, geo):
            y = g[0]
            f = self.forces(T, X, g)
            Fi = orc.spread(self.mac, self.delta, y.reshape(-1), f.reshape(-1))
            F = [a + b for a, b in zip(F, Fi)]
        H = fd.advection(self.axes_of, MAC, u)
        Gp = fd.gradient(self.paxes_of, CENTRE, p) if self.pseudo_pressure else None
        us = []
        for i in range(3):
            acc = F[i] / self.rho - H[i]
            rhs = dt * (self.I[i] @ acc)
            lu = self.L[i] @ u[i]
            if boundary is not None:
                lu = lu + boundary[i]
            rhs = rhs + dt * nu * lu
            if Gp is not None:
                rhs = rhs + (dt * nu / 2) * dt * (self.Bw[i] @ Gp[i])
            us.append(u[i] + self._helmholtz(i, rhs))
        div = fd.divergence(self.axes_of, MAC, us) / dt
        div -= div.mean()
        phi = self._poisson(div)
        G = fd.gradient(self.paxes_of, CENTRE, phi)
        un = [us[i] - dt * G[i] for i in range(3)]
        xn = []
        for T, X in zip(self.types, xs):
            U = orc.interpolate(self.mac, self.delta, X, un)
            xn.append(X + dt * U)
        return un, phi, xn, geo
