"""Small immersed-boundary configurations shared by the timestep tests."""
import math

import numpy as np

from conftest import load_inputs, load_surfaces

I = load_inputs()
S = load_surfaces()

NU = 10.0          # 1e-6 m^2/s in library units
RHO = 1e-3         # 1 g/mL
DT = 1e-4          # 0.1 us


def domain_spec(walls):
    """0.8 x 0.8 x 0.8 units (8 um), periodic in x and z, y walls or periodic."""
    per = (0.8, 0, (0.0, 0.0), (0.0, 0.0))
    wall = (0.8, 1, (1.0, 0.0), (1.0, 0.0))
    return [per, wall if walls else per, per]


def domain(ibm, walls):
    return [ibm.periodic(0.8), ibm.dirichlet(0.8) if walls else ibm.periodic(0.8), ibm.periodic(0.8)]


def cell_types(nd_rbc=150, ns_rbc=400, nd_plt=40, ns_plt=120):
    """A half-size RBC at the centre and two platelets; tethered to a 5%
    larger copy of their current shape (outward forces)."""
    out = []
    rbc_shape = lambda th, ph: 0.5 * I.rbc_shape(th, ph)
    for shape, nd, ns, deg, centres in (
            (rbc_shape, nd_rbc, ns_rbc, 5, [(0.4, 0.41, 0.39)]),
            (I.platelet_shape, nd_plt, ns_plt, 0, [(0.15, 0.2, 0.6), (0.62, 0.63, 0.18)])):
        dp, sp = I.bauer_params(nd), I.bauer_params(ns)
        E, Et, Ep = S.operators(dp, sp, deg)
        w = S.sphere_weights(*sp)
        nc = len(centres)
        X = np.stack([shape(*dp) + np.array(c)[:, None] for c in centres], axis=1)       # (3, nc, nd)
        ref = np.stack([1.05 * shape(*sp) + np.array(c)[:, None] for c in centres], axis=1)
        out.append({"E": E, "Et": Et, "Ep": Ep, "nd": nd, "ns": ns, "nc": nc, "w": w, "ref": ref, "k": 0.01,
                    "X": X.reshape(-1)})
    return out
