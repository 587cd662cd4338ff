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


# Membrane moduli of test/ib-whole.cc:208-215 in library units
# (1 dyn/cm = 1, 1 erg = 1e6): RBCs Skalak (2.5e-3, 2.5e-1) with bending
# 2e-12 erg; platelets neo-Hookean (1e-1, 1) with preferred bending 2e-11 erg.
RBC_MEMBRANE = {"tension": "skalak", "shear": 2.5e-3, "bulk": 2.5e-1, "bending": 2e-6, "preferred": False}
PLT_MEMBRANE = {"tension": "neohookean", "shear": 1e-1, "bulk": 1.0, "bending": 2e-5, "preferred": True}


def membrane_cell_types(nd_rbc=150, ns_rbc=400, nd_plt=40, ns_plt=120, inflate=1.04):
    """The cells of cell_types() with the reference's membrane models instead
    of tethers: each cell starts inflated by `inflate` (and the platelets
    tilted) relative to its undeformed shape, so tension and bending are both
    active."""
    out = []
    rbc_shape = lambda th, ph: 0.5 * I.rbc_shape(th, ph)
    for shape, nd, ns, deg, centres, mem in (
            (rbc_shape, nd_rbc, ns_rbc, 5, [(0.4, 0.41, 0.39)], RBC_MEMBRANE),
            (I.platelet_shape, nd_plt, ns_plt, 0, [(0.15, 0.2, 0.6), (0.62, 0.63, 0.18)], PLT_MEMBRANE)):
        dp, sp = I.bauer_params(nd), I.bauer_params(ns)
        ops = S.operators(dp, sp, deg, seconds=True)
        dops = S.data_operators(dp, deg)
        w = S.sphere_weights(*sp)
        nc = len(centres)
        ref = shape(*dp)                                                           # (3, nd)
        X = np.stack([I.rotate(inflate * ref, 0.1 * k, (1.0, 0.0, 0.0)) + np.array(c)[:, None]
                      for k, c in enumerate(centres)], axis=1)                     # (3, nc, nd)
        out.append({"E": ops[0], "Et": ops[1], "Ep": ops[2], "seconds": ops[3:], "data_ops": dops,
                    "nd": nd, "ns": ns, "nc": nc, "w": w, "k": 0.0, "reference": ref.reshape(-1),
                    "membrane": mem, "X": X.reshape(-1)})
    return out
