"""Seeded synthetic inputs for the benchmark configurations.

All inputs are generated on the host in FP64 by this module, so the GPU path
and the CPU oracle see identical bits:
  * random numbers: SplitMix64 -> 53-bit uniforms -> Box-Muller normals;
  * sphere sample sites: the Bauer spiral of bases/shapes/sphere.h:80-99 (the
    reference falls back to it when its data/*.bin site files are absent,
    include/rbc.h:52-58);
  * cell shapes: RBC biconcave map (include/rbc.h:19-35), platelet ellipsoid
    (include/platelet.h:21-34);
  * poses: bases::rotate (Rodrigues, bases/transforms.h:102-128) then translate.
Lengths are in library units (1 unit = 10 um).
"""
import math

import numpy as np

UM = 0.1
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014) producing 53-bit uniforms."""

    def __init__(self, seed):
        self.state = np.uint64(seed & MASK64)

    def next_u64(self, n):
        # Vectorised: state_k = seed + k * golden, then the mix function.
        golden = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * golden
            self.state = self.state + np.uint64(n) * golden
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n):
        """Uniforms in [0, 1) with 53 random bits."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normal(self, n):
        """N(0,1) by Box-Muller on pairs of uniforms (u1 in (0,1])."""
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1 = 1.0 - u[0::2]
        u2 = u[1::2]
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.empty(2 * m)
        z[0::2] = r * np.cos(2 * math.pi * u2)
        z[1::2] = r * np.sin(2 * math.pi * u2)
        return z[:n]


def bauer_params(n):
    """(theta, phi) of the n-point Bauer spiral (bases/shapes/sphere.h:80-99)."""
    t = np.arange(n, dtype=np.float64)
    l = math.sqrt(math.pi * n)
    z = -1.0 + (2.0 * t + 1.0) / n
    phi = np.arcsin(z)
    theta = np.fmod(l * phi, 2 * math.pi)
    return theta, phi


def sphere_shape(theta, phi, radius=1.0):
    x = np.cos(theta) * np.cos(phi)
    y = np.sin(theta) * np.cos(phi)
    z = np.sin(phi)
    return np.stack([radius * x, radius * y, radius * z])


RBC_RADIUS = 3.91 * UM
PLT_MAJOR = 1.55 * UM
PLT_MINOR = 0.50 * UM


def rbc_shape(theta, phi):
    """Biconcave RBC (include/rbc.h:19-35); returns 3 x n."""
    x = np.cos(theta) * np.cos(phi)
    y = np.sin(theta) * np.cos(phi)
    z0 = np.sin(phi)
    r2 = x * x + y * y
    z = 0.5 * z0 * (0.21 + 2.0 * r2 - 1.12 * r2 * r2)
    return np.stack([RBC_RADIUS * x, RBC_RADIUS * z, RBC_RADIUS * y])


def platelet_shape(theta, phi):
    """Platelet ellipsoid (include/platelet.h:96-108); returns 3 x n."""
    x = np.cos(theta) * np.cos(phi)
    y = np.sin(theta) * np.cos(phi)
    z = np.sin(phi)
    return np.stack([PLT_MAJOR * x, PLT_MINOR * y, PLT_MAJOR * z])


def rotate(points, angle, axis):
    """bases::rotate (Rodrigues) applied to 3 x n points."""
    a = np.asarray(axis, dtype=np.float64)
    l = math.sqrt(float(a @ a))
    c, s = math.cos(angle), math.sin(angle)
    d = (a @ points) / l
    w = np.cross(a[:, None], points, axis=0)
    return c * points + s * w / l + (1 - c) * d[None, :] * a[:, None] / l


def soa(cells):
    """Concatenates a list of 3 x n_k arrays into the reference's SoA
    layout x[n*i + m] (ib/novel.h:83, bases/container.h:54)."""
    pts = np.concatenate(cells, axis=1)
    return np.ascontiguousarray(pts.reshape(-1))


# ----------------------------------------------------------------------
# Configurations
# ----------------------------------------------------------------------
def config_c1(seed=1):
    """One sphere of 2562 Bauer-spiral markers, radius 3.9 um, centred in a
    periodic 16 um box; forces iid N(0,1) per component."""
    n = 2562
    th, ph = bauer_params(n)
    x = sphere_shape(th, ph, 3.9 * UM) + 8.0 * UM
    rng = SplitMix64(seed)
    f = rng.normal(3 * n)
    return {"L": (16 * UM,) * 3, "walls": False, "refinement": 32, "x": soa([x]), "f": f, "n": n}


def rbc_reference(ns):
    th, ph = bauer_params(ns)
    return rbc_shape(th, ph)


def platelet_reference(ns):
    th, ph = bauer_params(ns)
    return platelet_shape(th, ph)


def cell_poses(Lx, Ly, Lz, n_rbc, n_plt, seed):
    """Seeded non-overlapping poses: RBC layers (discs normal to y) on a
    per-layer lattice of 8 um slots with small jitter and tilt, platelets on a
    5 um lattice in the gaps between RBC layers.  Returns (rbc_poses,
    plt_poses), each a list of (rotations, centre) with rotations a list of
    (angle, axis) applied in order."""
    rng = SplitMix64(seed)
    per_layer_x = max(int(Lx / (8.0 * UM)), 1)
    per_layer_z = max(int(Lz / (8.0 * UM)), 1)
    per_layer = per_layer_x * per_layer_z
    layers = (n_rbc + per_layer - 1) // per_layer if n_rbc else 0
    dy = Ly / max(layers, 1)
    rbcs = []
    for k in range(n_rbc):
        layer, slot = divmod(k, per_layer)
        ix, iz = slot % per_layer_x, slot // per_layer_x
        u = rng.uniform(5)
        cx = (ix + 0.5) * Lx / per_layer_x + (u[0] - 0.5) * 0.08 * UM
        cz = (iz + 0.5) * Lz / per_layer_z + (u[1] - 0.5) * 0.08 * UM
        cy = (layer + 0.5) * dy + (u[2] - 0.5) * 0.1 * UM
        spin = 2 * math.pi * u[3]
        tilt = math.radians(2.0) * (u[4] - 0.5)
        rbcs.append(([(spin, (0.0, 1.0, 0.0)), (tilt, (1.0, 0.0, 0.0))], np.array([cx, cy, cz])))
    plts = []
    gaps = max(layers - 1, 1)
    nx = max(int(Lx / (5.0 * UM)), 1)
    nz = max(int(Lz / (5.0 * UM)), 1)
    for k in range(n_plt):
        gap, slot = divmod(k, nx * nz)
        gap = gap % gaps
        ix, iz = slot % nx, slot // nx
        u = rng.uniform(3)
        cx = (ix + 0.5) * Lx / nx + (u[0] - 0.5) * 0.5 * UM
        cz = (iz + 0.5) * Lz / nz + (u[1] - 0.5) * 0.5 * UM
        cy = (gap + 1) * dy if layers > 1 else 0.5 * Ly
        plts.append(([(2 * math.pi * u[2], (0.0, 1.0, 0.0))], np.array([cx, cy, cz])))
    return rbcs, plts


def posed(points, pose, scale=1.0):
    """Applies a pose (rotations, centre) to 3 x n reference points, after
    scaling them about the origin."""
    p = scale * points
    for angle, axis in pose[0]:
        p = rotate(p, angle, axis)
    return p + pose[1][:, None]


def layered_cells(Lx, Ly, Lz, n_rbc, n_plt, ns_rbc, ns_plt, seed):
    """Sample-site positions (lists of 3 x ns arrays) of the cell_poses
    placement."""
    rp, pp = cell_poses(Lx, Ly, Lz, n_rbc, n_plt, seed)
    rbc_ref = rbc_reference(ns_rbc)
    plt_ref = platelet_reference(ns_plt)
    return [posed(rbc_ref, q) for q in rp], [posed(plt_ref, q) for q in pp]


def config_c4(seed=4, ns_rbc=8832, ns_plt=1552):
    """200 RBCs + 400 platelets in a 32 x 64 x 32 um box, walls in y,
    fd::mac{128} over a 32 um base unit -> 128 x 256 x 128 cells."""
    L = (32 * UM, 64 * UM, 32 * UM)
    rbcs, plts = layered_cells(*L, 200, 400, ns_rbc, ns_plt, seed)
    rng = SplitMix64(seed + 1000)
    xr = soa(rbcs)
    xp = soa(plts)
    nr = xr.size // 3
    npl = xp.size // 3
    return {"L": L, "walls": True, "refinement": 128,
            "x_rbc": xr, "x_plt": xp, "n_rbc": nr, "n_plt": npl,
            "f_rbc": rng.normal(3 * nr), "f_plt": rng.normal(3 * npl)}


def random_markers(n, L, seed, margin=0.0):
    """n markers uniform in the box [margin, L - margin] per dimension (SoA)."""
    rng = SplitMix64(seed)
    out = []
    for j in range(3):
        u = rng.uniform(n)
        out.append(margin + u * (L[j] - 2 * margin))
    return np.ascontiguousarray(np.concatenate(out))


def config_c5(seed=5, nboxes=64, L=32 * UM, refinement=32, ns_rbc=546, ns_plt=386):
    """64 independent periodic boxes of 32 um with 32^3 cells; box b holds
    U{1..20} cells (seeded), each an RBC (546 sample sites) or a platelet
    (386) with probability 1/2, at the first slots of a jittered 3 x 3 x 3
    lattice (10.7 um spacing, so cells do not overlap), randomly rotated;
    forces iid N(0,1).  Returns box-local SoA positions of all markers in
    box order, the box offsets and the forces."""
    rng = SplitMix64(seed)
    rbc_ref = rbc_reference(ns_rbc)
    plt_ref = platelet_reference(ns_plt)
    spacing = L / 3
    xs, offsets, kinds = [], [0], []
    for b in range(nboxes):
        ncells = 1 + int(rng.uniform(1)[0] * 20)
        pts = []
        for k in range(ncells):
            u = rng.uniform(6)
            iz, rem = divmod(k, 9)
            iy, ix = divmod(rem, 3)
            centre = np.array([(ix + 0.5 + 0.1 * (u[0] - 0.5)) * spacing,
                               (iy + 0.5 + 0.1 * (u[1] - 0.5)) * spacing,
                               (iz + 0.5 + 0.1 * (u[2] - 0.5)) * spacing])
            ref = rbc_ref if u[3] < 0.5 else plt_ref
            kinds.append(0 if u[3] < 0.5 else 1)
            p = rotate(ref, 2 * math.pi * u[4], (0.0, 1.0, 0.0))
            p = rotate(p, math.pi * u[5], (1.0, 0.0, 0.0))
            pts.append(p + centre[:, None])
        box = np.concatenate(pts, axis=1)
        xs.append(box)
        offsets.append(offsets[-1] + box.shape[1])
    x = np.ascontiguousarray(np.concatenate(xs, axis=1).reshape(-1))
    n = offsets[-1]
    f = rng.normal(3 * n)
    return {"L": L, "refinement": refinement, "nboxes": nboxes, "x": x, "f": f, "offsets": offsets, "n": n,
            "kinds": kinds}


# The eight RBC poses of test/ib-whole.cc:163-170: rotation (angle, axis)
# then translation (um).
C2_POSES = [
    (5.068183442568522, (1.030437666936895, -0.021108450766973392, -0.04630059618399827),
     (4.015214448901787, 6.100138060189724, 1.9880461041978537)),
    (5.007815431576859, (0.9792381457522887, 0.06271914842308295, -0.05642529490935586),
     (11.877848594136138, 5.83714041009, 3.97975684227)),
    (5.074088738134238, (1.0142231005007727, -2.9623760263254767e-05, -0.039362970458103955),
     (4.146345896973395, 5.881458848974533, 5.962170272743408)),
    (5.079676113839819, (0.9890875402932139, -0.031109409809641483, -0.05032173371464803),
     (12.013925005641164, 6.158047729278891, 7.93432970522188)),
    (5.1362464926490805, (1.0348848942058133, 0.01720012363571293, -0.11947361656641693),
     (3.908805848485636, 6.091134893946915, 9.855992001042118)),
    (5.054989211831597, (0.9675754189340313, -0.048388248918717285, 0.03738865256187718),
     (12.083267206719906, 5.80425714837, 12.1799588627)),
    (5.108732820408364, (0.9569821581817692, -0.056252540193673733, -0.07872719696376375),
     (4.110024751201524, 5.91689772487407, 13.976372295071807)),
    (4.931079097988467, (0.9593196951752679, -0.020300578072308676, 0.07572396922527558),
     (12.02165847715475, 6.012459328852655, 15.96571613373596)),
]


def config_c2_cells(points):
    """The eight posed copies (3, 8, n) of reference points (3, n)."""
    return np.stack([rotate(points, a, ax) + np.array(t)[:, None] * UM for a, ax, t in C2_POSES], axis=1)
