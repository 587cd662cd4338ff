"""CPU restatement of the library's immersed-boundary timestep, composed from
the oracle's pieces (test infrastructure only).

The reference's own step (ib::solver, test/ib-whole.cc:226-275) is not part
of the reference, so this follows the library's documented scheme:
  y = E^T X (oracle/rbf.py), f = -k (y - y_ref),
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
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MAC = [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
CENTRE = [0.5, 0.5, 0.5]


class Step:
    """spec: [(length, solid, lower, upper)] x 3 as for orc.grid_axes."""

    def __init__(self, orc, fd, rbf, spec, refinement, dt, nu, rho, delta, types, pseudo_pressure=False):
        self.orc, self.fd, self.rbf = orc, fd, rbf
        self.spec, self.ref = spec, refinement
        self.dt, self.nu, self.rho, self.delta = dt, nu, rho, delta
        self.types = types                  # dicts: E, Et, Ep, nd, ns, nc, w, ref (3, nc, ns), k
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
        self.helm = [spla.splu(sp.csc_matrix(fd.operator(ax, 1.0, -lam))) for ax in self.uaxes]
        Lp = sp.lil_matrix(fd.laplacian(self.paxes))
        Lp[0, :] = 0
        Lp[0, 0] = 1.0
        self.pois = spla.splu(sp.csc_matrix(Lp))

    def _axes(self, spec, align):
        a = self.orc.grid_axes(spec, self.ref, align)
        return [a[j] for j in range(3)]

    def geometry(self, xs):
        out = []
        for T, X in zip(self.types, xs):
            Xm = X.reshape(3 * T["nc"], T["nd"])
            y, t1, t2 = (self.rbf.evaluate(M, T["nd"], Xm).reshape(3, T["nc"], T["ns"])
                         for M in (T["E"], T["Et"], T["Ep"]))
            out.append((y, t1, t2))
        return out

    def step(self, u, p, xs, boundary=None):
        fd, orc, dt, nu = self.fd, self.orc, self.dt, self.nu
        geo = self.geometry(xs)
        F = [np.zeros(self.I[i].shape[0]) for i in range(3)]
        for T, (y, _, _) in zip(self.types, geo):
            f = -T["k"] * (y - T["ref"])
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
            us.append(u[i] + self.helm[i].solve(rhs))
        div = fd.divergence(self.axes_of, MAC, us) / dt
        div -= div.mean()
        div[0] = 0.0
        phi = self.pois.solve(div)
        phi -= phi.mean()
        G = fd.gradient(self.paxes_of, CENTRE, phi)
        un = [us[i] - dt * G[i] for i in range(3)]
        xn = []
        for T, X in zip(self.types, xs):
            U = orc.interpolate(self.mac, self.delta, X, un)
            xn.append(X + dt * U)
        return un, phi, xn, geo
