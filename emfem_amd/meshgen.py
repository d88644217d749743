"""Synthetic EM models on graded tensor grids split into Kuhn tetrahedra.

TetGen is not available, so every workload is generated here as a deterministic
graded tensor grid whose hexahedra are each split into 6 tetrahedra along the
main diagonal (the Kuhn/Freudenthal split, which is conforming across hexes).
Heights can be warped for topography or bathymetry.  All vertex coordinates are
rounded to whole metres, which keeps the orientation predicate used by point
location exact in double precision (every product of three coordinate
differences stays below 2**53).

A `Model` holds what the reference reads from its `.mesh` + `.emd` inputs
(`src/em_io.cc:155-287`): points, tets with integer attributes, the resistivity
table with its anisotropy form (`src/em_io.cc:9-86`), frequencies, transmitters
(7 numbers each, `src/em_defs.h:25`), receivers and the observation list.

Tetrahedra are emitted with det([1; x_v]) > 0, the orientation TetGen produces
and the reference's `TetAccessor::contains_point` assumes
(`src/em_mesh.h:208-230`).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Observation type codes (src/em_defs.h:49-74).
F_EX_RI, F_EY_RI, F_EZ_RI, F_HX_RI, F_HY_RI, F_HZ_RI = 111, 121, 131, 141, 151, 161
F_EX_AP, F_EY_AP, F_EZ_AP, F_HX_AP, F_HY_AP, F_HZ_AP = 112, 122, 132, 142, 152, 162
R_XY_AP, R_YX_AP = 212, 222
Z_XX_RI, Z_XY_RI, Z_YX_RI, Z_YY_RI, T_ZX_RI, T_ZY_RI = 311, 321, 331, 341, 351, 361
Z_XX_AP, Z_XY_AP, Z_YX_AP, Z_YY_AP = 312, 322, 332, 342

MT_TIDX = -3  # transmitter index the reference uses for MT observations (src/em_fwd.cc:766)


@dataclasses.dataclass
class Model:
    points: np.ndarray          # (V, 3) float64, metres, z positive down
    tets: np.ndarray            # (T, 4) int32
    attr: np.ndarray            # (T,) int32, row of `rho`
    aniso_form: int             # 0 isotropic, 1 vertical, 2 triaxial, 3 arbitrary
    rho: np.ndarray             # (n_rho, {1,2,3,6}) float64
    freqs: np.ndarray           # (F,) float64, Hz
    tx: np.ndarray              # (n_tx, 7) float64: x y z azimuth dip current length
    rx: np.ndarray              # (n_rx, 3) float64
    obs: np.ndarray             # (n_obs, 4) int32: otype fidx tidx ridx
    name: str = ""

    @property
    def n_tets(self):
        return int(self.tets.shape[0])

    @property
    def n_vertices(self):
        return int(self.points.shape[0])


# Local vertex paths of the 6 Kuhn tetrahedra of a unit cube: vertex bits (x=1, y=2, z=4).
def _kuhn_local():
    out = []
    for perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)):
        b0 = 0
        b1 = b0 | (1 << perm[0])
        b2 = b1 | (1 << perm[1])
        b3 = 7
        even = perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1))
        out.append((b0, b1, b2, b3) if even else (b0, b1, b3, b2))
    return out


KUHN = _kuhn_local()


def kuhn_grid(xn, yn, zn):
    """Points and tets of the tensor grid with node coordinates xn, yn, zn."""
    xn = np.asarray(xn, dtype=np.float64)
    yn = np.asarray(yn, dtype=np.float64)
    zn = np.asarray(zn, dtype=np.float64)
    nx, ny, nz = len(xn), len(yn), len(zn)
    Z, Y, X = np.meshgrid(zn, yn, xn, indexing="ij")
    points = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def vid(i, j, k):
        return (k * ny + j) * nx + i

    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    # order cells z-major so tets of one layer are contiguous (better locality)
    order = np.lexsort((ci, cj, ck))
    ci, cj, ck = ci[order], cj[order], ck[order]
    corner = [vid(ci + (b & 1), cj + ((b >> 1) & 1), ck + ((b >> 2) & 1)) for b in range(8)]
    tets = np.empty((ci.size, 6, 4), dtype=np.int64)
    for s, loc in enumerate(KUHN):
        for v in range(4):
            tets[:, s, v] = corner[loc[v]]
    cell = np.repeat(np.arange(ci.size), 6)
    return points, tets.reshape(-1, 4).astype(np.int32), cell, (ci, cj, ck)


def tet_det(points, tets):
    p = points[tets]
    return np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))


def centroids(points, tets):
    return points[tets].mean(axis=1)


def graded(core_half, core_h, outer, n_out):
    """Symmetric node list: uniform core [-core_half, core_half] with spacing core_h,
    then n_out geometrically growing cells out to +-outer."""
    core = np.arange(-core_half, core_half + 0.5 * core_h, core_h)
    grow = _geometric(core_half, outer, core_h, n_out)
    return np.concatenate([-grow[::-1], core, grow])


def _geometric(start, end, h0, n):
    """n nodes after `start` ending exactly at `end`, first step about h0, growing geometrically."""
    L = end - start
    lo, hi = 1.0 + 1e-9, 10.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        s = h0 * (r ** n - 1) / (r - 1)
        if s > L:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    steps = h0 * r ** np.arange(n)
    steps *= L / steps.sum()
    return start + np.cumsum(steps)


def _round(points):
    return np.round(points)


def _warp(points, surf, ztop, zbot):
    """Move the z=0 node layer to z=surf(x,y); blend linearly to no shift at ztop and zbot."""
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    s = surf(x, y)
    w = np.where(z <= 0, 1.0 - z / ztop, 1.0 - z / zbot)
    out = points.copy()
    out[:, 2] = z + s * w
    return out


def _check(points, tets):
    d = tet_det(points, tets)
    if not np.all(d > 0):
        raise ValueError("mesh has non-positive tetrahedra after warping/rounding")


def mt_obs(n_freqs, n_rx, types=(Z_XY_RI, Z_YX_RI, Z_XX_RI, Z_YY_RI, T_ZX_RI, T_ZY_RI)):
    rows = []
    for f in range(n_freqs):
        for r in range(n_rx):
            for t in types:
                rows.append((t, f, MT_TIDX, r))
    return np.asarray(rows, dtype=np.int32).reshape(-1, 4)


def surface_grid(lo, hi, n, z=0.0):
    g = np.linspace(lo, hi, n)
    X, Y = np.meshgrid(g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)


# ----------------------------------------------------------------------------------------
# Input files in the reference's text formats (src/em_io.cc:155-287)
# ----------------------------------------------------------------------------------------

def write_mesh(m: Model, path, bounds=(1.0e-2, 1.0e5)):
    """`.mesh`: points, tets with attribute, then the resistivity block
    (n_rho aniso_form; each row its values and two bounds)."""
    with open(path, "w") as f:
        f.write("# points\n%d\n" % m.n_vertices)
        np.savetxt(f, m.points, fmt="%.17g")
        f.write("# tets: 4 vertex ids and the attribute\n%d\n" % m.n_tets)
        np.savetxt(f, np.concatenate([m.tets, m.attr[:, None]], axis=1), fmt="%d")
        f.write("# resistivity\n%d %d\n" % (m.rho.shape[0], m.aniso_form))
        for row in m.rho:
            f.write(" ".join("%.17g" % v for v in row) + " %g %g\n" % bounds)


def write_emd(m: Model, path, values=None):
    """`.emd`: frequencies, transmitters, receivers, observations (otype fidx tidx ridx
    re im err_re err_im)."""
    with open(path, "w") as f:
        f.write("# frequencies\n%d\n" % len(m.freqs))
        for v in m.freqs:
            f.write("%.17g\n" % v)
        f.write("# transmitters\n%d\n" % len(m.tx))
        for row in m.tx:
            f.write(" ".join("%.17g" % v for v in row) + "\n")
        f.write("# receivers\n%d\n" % len(m.rx))
        for row in m.rx:
            f.write(" ".join("%.17g" % v for v in row) + "\n")
        f.write("# observations\n%d\n" % len(m.obs))
        vals = np.ones((len(m.obs), 4)) if values is None else values
        for o, v in zip(m.obs, vals):
            f.write("%d %d %d %d %.17g %.17g %.17g %.17g\n" % (tuple(o) + tuple(v)))


# ----------------------------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------------------------

AIR_RHO = 1.0e8


def plain_mt(scale=1):
    """Plain MT, FP64: 20x20 km earth under 50 km of air, 100 ohm-m half-space with a
    2x2x1 km 10 ohm-m block at 500 m depth.  scale=1 gives 20x20x21 cells
    (50,400 tets, 9,702 vertices, 62,581 edges).  One frequency (1 Hz), 11x11 receivers."""
    xn = graded(2000.0, 500.0 / scale, 10000.0, 6 * scale)
    air = -np.array([50000, 30000, 16000, 8000, 4000, 2000, 1000, 400], dtype=np.float64)
    earth = np.array([0, 250, 500, 750, 1000, 1250, 1500, 2000, 3000, 4500, 7000, 10000, 14000, 20000],
                     dtype=np.float64)
    zn = np.concatenate([air, earth])
    if scale > 1:
        zn = _subdivide(zn, scale)
        xn = np.unique(np.round(xn))
    pts, tets, _, _ = kuhn_grid(xn, xn, zn)
    pts = _round(pts)
    c = centroids(pts, tets)
    attr = np.where(c[:, 2] < 0, 0, 1).astype(np.int32)
    blk = (np.abs(c[:, 0]) < 1000) & (np.abs(c[:, 1]) < 1000) & (c[:, 2] > 500) & (c[:, 2] < 1500)
    attr[blk] = 2
    rho = np.array([[AIR_RHO], [100.0], [10.0]])
    rx = surface_grid(-5000.0, 5000.0, 11)
    _check(pts, tets)
    return Model(pts, tets, attr, 0, rho, np.array([1.0]), np.zeros((0, 7)), rx,
                 mt_obs(1, rx.shape[0]), name="plain_mt")


def _subdivide(nodes, s):
    out = [nodes[0]]
    for a, b in zip(nodes[:-1], nodes[1:]):
        for i in range(1, s + 1):
            out.append(a + (b - a) * i / s)
    return np.round(np.asarray(out))


def halfspace_mt(n=8, rho=100.0, freq=1.0, extent=20000.0, depth=20000.0, air=50000.0,
                 core=4000.0, n_rx=3, rx_span=None):
    """Homogeneous half-space under air: the closed-form MT test case."""
    xn = graded(core / 2, core / n, extent / 2, max(2, n // 2))
    earth = np.concatenate([np.linspace(0, core, n + 1), _geometric(core, depth, core / n, max(2, n // 2))])
    airn = -_geometric(0.0, air, core / n, max(3, n // 2))[::-1]
    zn = np.concatenate([airn, earth])
    pts, tets, _, _ = kuhn_grid(np.round(xn), np.round(xn), np.round(zn))
    pts = _round(pts)
    c = centroids(pts, tets)
    attr = np.where(c[:, 2] < 0, 0, 1).astype(np.int32)
    span = core / 4 if rx_span is None else rx_span
    rx = surface_grid(-span, span, n_rx)
    return Model(pts, tets, attr, 0, np.array([[AIR_RHO], [rho]]), np.array([freq]),
                 np.zeros((0, 7)), rx, mt_obs(1, rx.shape[0]), name="halfspace_mt")


def tiny_mt(n=3, aniso_form=0, seed=0, freq=1.0):
    """A few hundred tets: the brute-force and golden-fixture case."""
    xn = np.linspace(-3000.0, 3000.0, n + 1)
    zn = np.concatenate([-np.array([20000.0, 6000.0, 2000.0]), np.linspace(0.0, 3000.0, n + 1)])
    zn = np.unique(zn)
    pts, tets, _, _ = kuhn_grid(xn, xn, zn)
    pts = _round(pts)
    c = centroids(pts, tets)
    attr = np.where(c[:, 2] < 0, 0, 1).astype(np.int32)
    blk = (np.abs(c[:, 0]) < 1000) & (np.abs(c[:, 1]) < 1000) & (c[:, 2] > 0) & (c[:, 2] < 1000)
    attr[blk] = 2
    rng = np.random.default_rng(seed)
    rho = _rho_table(aniso_form, [AIR_RHO, 100.0, 10.0], rng)
    rx = np.array([[-500.0, 250.0, 0.0], [0.0, 0.0, 0.0], [700.0, -300.0, 0.0]])
    return Model(pts, tets, attr, aniso_form, rho, np.array([freq]), np.zeros((0, 7)), rx,
                 mt_obs(1, rx.shape[0]), name="tiny_mt")


def _rho_table(form, base, rng):
    rows = []
    for i, r in enumerate(base):
        if form == 0:
            rows.append([r])
        elif form == 1:
            rows.append([r, r if i == 0 else r * rng.uniform(1, 5)])
        elif form == 2:
            rows.append([r] * 3 if i == 0 else [r * rng.uniform(0.5, 2) for _ in range(3)])
        else:
            if i == 0:
                rows.append([r, r, r, 0.0, 0.0, 0.0])
            else:
                rows.append([r * rng.uniform(0.5, 2) for _ in range(3)] + list(rng.uniform(0, 360, 3)))
    return np.asarray(rows, dtype=np.float64)


def adaptive_mt(cells=(40, 40, 40), seed=2, n_freqs=8, n_rx_side=20, amplitude=300.0, length=3000.0):
    """Adaptive-MT model: anisotropic (arbitrary form) blocks, Gaussian-hill topography with
    300 m amplitude, 8 log-spaced frequencies 1e-3..100 Hz and 20x20 surface receivers.
    `cells` sets the grid; (40,40,40) gives ~0.45M edges, (60,60,60) ~1.5M."""
    nxc, nyc, nzc = cells
    rng = np.random.default_rng(seed)
    n_out = max(3, nxc // 6)
    core_cells = nxc - 2 * n_out
    xn = graded(4000.0, 8000.0 / core_cells, 20000.0, n_out)
    yn = xn
    n_air = max(4, nzc // 4)
    n_earth = nzc - n_air
    n_core_z = max(2, (2 * n_earth) // 3)
    earth = np.concatenate([np.linspace(0.0, 4000.0, n_core_z + 1),
                            _geometric(4000.0, 30000.0, 4000.0 / n_core_z, n_earth - n_core_z)])
    airn = -_geometric(0.0, 50000.0, 4000.0 / n_core_z, n_air)[::-1]
    zn = np.concatenate([airn, earth])
    pts, tets, _, _ = kuhn_grid(np.round(xn), np.round(yn), np.round(zn))

    def surf(x, y):
        return -amplitude * np.exp(-(x * x + y * y) / (2 * length * length))

    pts = _round(_warp(pts, surf, zn[0], zn[-1]))
    _check(pts, tets)
    c = centroids(pts, tets)
    below = c[:, 2] > surf(c[:, 0], c[:, 1])
    attr = np.where(below, 1, 0).astype(np.int32)
    blocks = [((-2500, -500), (-1500, 1500), (200, 1400)), ((500, 2500), (-2000, 0), (600, 2000)),
              ((-1000, 1500), (800, 2800), (1500, 3500))]
    for b, ((x0, x1), (y0, y1), (z0, z1)) in enumerate(blocks):
        m = below & (c[:, 0] > x0) & (c[:, 0] < x1) & (c[:, 1] > y0) & (c[:, 1] < y1) & (c[:, 2] > z0) & (c[:, 2] < z1)
        attr[m] = 2 + b
    rho = [[AIR_RHO, AIR_RHO, AIR_RHO, 0.0, 0.0, 0.0]]
    for _ in range(1 + len(blocks)):
        rho.append(list(np.exp(rng.uniform(0.0, math.log(1000.0), 3))) + list(rng.uniform(0.0, 360.0, 3)))
    rho = np.asarray(rho)
    g = np.linspace(-3800.0, 3800.0, n_rx_side)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rx = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    rx[:, 2] = _surface_height_on_mesh(pts, tets, rx, surf)
    freqs = 10.0 ** (-3.0 + 5.0 * np.arange(n_freqs) / max(1, n_freqs - 1))
    return Model(pts, tets, attr, 3, rho, freqs, np.zeros((0, 7)), rx, mt_obs(n_freqs, rx.shape[0]),
                 name="adaptive_mt")


def _surface_height_on_mesh(pts, tets, rx, surf):
    """Receivers sit on the discrete (piecewise-linear) surface: take the analytic height
    slightly below it so they fall inside the top earth layer rather than in the air."""
    return surf(rx[:, 0], rx[:, 1]) + 1.0


def marine_csem(cells=(48, 48, 40), n_tx=64, freqs=(0.25, 0.75, 1.25), with_air=True):
    """Marine CSEM: 0.3 S/m seawater over 1 ohm-m sediments, seafloor at 1000 m with a
    200 m sinusoidal bathymetry, a 100 ohm-m 4x4x0.1 km reservoir 1 km below the seafloor,
    64 x-directed point dipoles 50 m above the seafloor along x in [-5, 5] km."""
    nxc, nyc, nzc = cells
    n_out = max(3, nxc // 6)
    core_cells = nxc - 2 * n_out
    xn = graded(6000.0, 12000.0 / core_cells, 20000.0, n_out)
    yn = graded(3000.0, 6000.0 / max(4, nyc - 2 * n_out), 20000.0, n_out)
    n_air = max(3, nzc // 8) if with_air else 0
    n_sea = max(4, nzc // 5)
    n_rest = nzc - n_air - n_sea
    n_core = max(4, (2 * n_rest) // 3)
    sea = np.linspace(0.0, 1000.0, n_sea + 1)
    sed = np.concatenate([np.linspace(1000.0, 3200.0, n_core + 1)[1:],
                          _geometric(3200.0, 30000.0, 2200.0 / n_core, n_rest - n_core)])
    zn = np.concatenate([sea, sed])
    if with_air:
        zn = np.concatenate([-_geometric(0.0, 30000.0, 200.0, n_air)[::-1], zn])
    pts, tets, _, _ = kuhn_grid(np.round(xn), np.round(yn), np.round(zn))

    def seafloor(x, y):
        return 1000.0 + 200.0 * np.sin(2 * np.pi * x / 5000.0)

    # warp the z=1000 layer onto the seafloor; blend to no shift at 0 and at the bottom
    z = pts[:, 2]
    s = seafloor(pts[:, 0], pts[:, 1]) - 1000.0
    w = np.where(z <= 1000.0, np.clip(z / 1000.0, 0.0, None), (zn[-1] - z) / (zn[-1] - 1000.0))
    w = np.where(z < 0, 0.0, w)
    pts = pts.copy()
    pts[:, 2] = z + s * w
    pts = _round(pts)
    _check(pts, tets)
    c = centroids(pts, tets)
    sf = seafloor(c[:, 0], c[:, 1])
    attr = np.where(c[:, 2] < 0, 0, np.where(c[:, 2] < sf, 1, 2)).astype(np.int32)
    res = (np.abs(c[:, 0]) < 2000) & (np.abs(c[:, 1]) < 2000) & (c[:, 2] > sf + 950) & (c[:, 2] < sf + 1050)
    attr[res] = 3
    rho = np.array([[AIR_RHO], [1.0 / 0.3], [1.0], [100.0]])
    txx = np.linspace(-5000.0, 5000.0, n_tx)
    tx = np.zeros((n_tx, 7))
    tx[:, 0] = txx
    tx[:, 2] = seafloor(txx, 0.0) - 50.0
    tx[:, 5] = 1.0
    rxx = np.linspace(-7000.0, 7000.0, 29)
    rx = np.stack([rxx, np.zeros_like(rxx), seafloor(rxx, 0.0) - 1.0], axis=1)
    obs = []
    for f in range(len(freqs)):
        for t in range(n_tx):
            for r in range(rx.shape[0]):
                for o in (F_EX_RI, F_EY_RI, F_EZ_RI, F_HX_RI, F_HY_RI, F_HZ_RI):
                    obs.append((o, f, t, r))
    return Model(pts, tets, attr, 0, rho, np.asarray(freqs, dtype=np.float64), tx, rx,
                 np.asarray(obs, dtype=np.int32), name="marine_csem")


def uneven_mt_batch(n_models=256, seed=4, e_min=2.0e4, e_max=2.0e5, f_min=1e-2, f_max=10.0):
    """Uneven batch of independent MT models: edge counts log-uniform in
    [e_min, e_max], a half-space of 10-1000 ohm-m (log-uniform) with 1-4 random blocks
    (1-1000 ohm-m), one random frequency per model, log-uniform in [f_min, f_max], and a
    square surface receiver grid of 2-7 per side spanning +-0.5..4 km.  The grid size n per model is the one whose edge count (about
    7 (n+1)^2 (n+1+n_pad) on the padded Kuhn grid) is nearest the drawn target."""
    rng = np.random.default_rng(seed)
    models = []
    for k in range(n_models):
        target = float(np.exp(rng.uniform(np.log(e_min), np.log(e_max))))
        best = None
        for n in range(4, 64):
            nx = n + 2 * max(2, n // 2)
            nz = n + max(2, n // 2) + max(3, n // 2)
            e = 7.0 * (nx + 1) * (nx + 1) * (nz + 1)
            if best is None or abs(np.log(e / target)) < abs(np.log(best[1] / target)):
                best = (n, e)
        n = best[0]
        rho_hs = float(np.exp(rng.uniform(np.log(10.0), np.log(1000.0))))
        freq = float(np.exp(rng.uniform(np.log(f_min), np.log(f_max))))
        m = halfspace_mt(n=n, rho=rho_hs, freq=freq, n_rx=int(rng.integers(2, 8)),
                         rx_span=float(rng.uniform(500.0, 4000.0)))
        c = centroids(m.points, m.tets)
        rho = [[AIR_RHO], [rho_hs]]
        for b in range(int(rng.integers(1, 5))):
            lo = rng.uniform([-1500.0, -1500.0, 100.0], [1000.0, 1000.0, 2500.0])
            size = rng.uniform([300.0, 300.0, 200.0], [1500.0, 1500.0, 1200.0])
            inside = np.all((c > lo) & (c < lo + size), axis=1)
            m.attr[inside] = len(rho)
            rho.append([float(np.exp(rng.uniform(np.log(1.0), np.log(1000.0))))])
        m.rho = np.array(rho)
        m.name = "uneven_mt_%03d" % k
        models.append(m)
    return models


def varied_mt_batch(n_models=32, seed=21, n_min=6, n_max=12, f_min=1e-3, f_max=10.0):
    """Training models for the learned indicator: the uneven batch's half-spaces of
    10-1000 ohm-m with 1-4 random blocks and one random frequency, widened by what the
    adaptive-MT models add: a Gaussian-hill topography of random amplitude (0-400 m) and
    width (1.5-5 km), and anisotropic tensors (arbitrary form, principal resistivities
    within a factor of 1-10 of the block's value, random Euler angles) for about half of
    the blocks and the background.  Receivers: a square grid of 2-7 per side spanning
    +-0.5..4 km, 1 m below the discrete surface."""
    rng = np.random.default_rng(seed)
    models = []
    for k in range(n_models):
        n = int(rng.integers(n_min, n_max + 1))
        n_out = max(3, n // 6 + 2)
        xn = graded(4000.0, 8000.0 / n, 20000.0, n_out)
        n_air = max(4, n // 3)
        n_core_z = max(3, (2 * n) // 3)
        earth = np.concatenate([np.linspace(0.0, 4000.0, n_core_z + 1),
                                _geometric(4000.0, 30000.0, 4000.0 / n_core_z, max(3, n // 3))])
        airn = -_geometric(0.0, 50000.0, 4000.0 / n_core_z, n_air)[::-1]
        zn = np.concatenate([airn, earth])
        pts, tets, _, _ = kuhn_grid(np.round(xn), np.round(xn), np.round(zn))
        amp = float(rng.uniform(0.0, 400.0))
        length = float(rng.uniform(1500.0, 5000.0))
        x0, y0 = rng.uniform(-2000.0, 2000.0, 2)

        def surf(x, y, amp=amp, length=length, x0=x0, y0=y0):
            return -amp * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * length * length))

        pts = _round(_warp(pts, surf, zn[0], zn[-1]))
        _check(pts, tets)
        c = centroids(pts, tets)
        below = c[:, 2] > surf(c[:, 0], c[:, 1])
        attr = np.where(below, 1, 0).astype(np.int32)

        def tensor(rho0):
            if rng.uniform() < 0.5:
                return [rho0, rho0, rho0, 0.0, 0.0, 0.0]
            r = rho0 * np.exp(rng.uniform(-0.5, 0.5, 3) * math.log(10.0))
            return list(r) + list(rng.uniform(0.0, 360.0, 3))

        rho_hs = float(np.exp(rng.uniform(np.log(10.0), np.log(1000.0))))
        rho = [[AIR_RHO, AIR_RHO, AIR_RHO, 0.0, 0.0, 0.0], tensor(rho_hs)]
        for b in range(int(rng.integers(1, 5))):
            lo = rng.uniform([-1500.0, -1500.0, 100.0], [1000.0, 1000.0, 2500.0])
            size = rng.uniform([300.0, 300.0, 200.0], [1500.0, 1500.0, 1200.0])
            inside = below & np.all((c > lo) & (c < lo + size), axis=1)
            attr[inside] = len(rho)
            rho.append(tensor(float(np.exp(rng.uniform(np.log(1.0), np.log(1000.0))))))
        n_rx = int(rng.integers(2, 8))
        span = float(rng.uniform(500.0, 4000.0))
        g = np.linspace(-span, span, n_rx)
        X, Y = np.meshgrid(g, g, indexing="ij")
        rx = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
        rx[:, 2] = _surface_height_on_mesh(pts, tets, rx, surf)
        freq = float(np.exp(rng.uniform(np.log(f_min), np.log(f_max))))
        models.append(Model(pts, tets, attr, 3, np.asarray(rho), np.array([freq]), np.zeros((0, 7)), rx,
                            mt_obs(1, rx.shape[0]), name="varied_mt_%03d" % k))
    return models


MU0 = 4e-7 * np.pi


def layered_impedance(tops, rho, freq, z):
    """Exact MT impedance E_x / H_y at depth z (>= 0, below the surface) of a layered
    earth (layer i from tops[i] to tops[i+1], the last one unbounded; e^{+i w t},
    z down), by the impedance recursion from the bottom half-space up."""
    w = 2 * np.pi * freq
    k = np.sqrt(1j * w * MU0 / np.asarray(rho, dtype=float))
    zi = 1j * w * MU0 / k                      # intrinsic impedances
    Z = zi[-1]
    layer = max(i for i in range(len(tops)) if z >= tops[i])
    for i in range(len(tops) - 2, layer - 1, -1):
        h = (tops[i + 1] - (z if i == layer else tops[i]))
        t = np.tanh(k[i] * h)
        Z = zi[i] * (Z + zi[i] * t) / (zi[i] + Z * t)
    return Z


def layered_mt(freqs, h=1000.0, hz=300.0, tops=(0.0, 600.0, 1500.0), rho=(100.0, 10.0, 300.0)):
    """Flat layered earth under air on a graded Kuhn mesh (default: 100 ohm-m to 600 m,
    10 ohm-m to 1500 m, 300 ohm-m below); 6 x 6 receivers over +-3 km, 1 m below the
    surface.  Returns (model, exact impedances of its observations): the MT boundary
    data are the exact 1-D field here, so the discrete solutions converge to it."""
    xn = graded(4000.0, h, 20000.0, 4)
    earth = np.concatenate([np.arange(0.0, 4200.0 + 1, hz), _geometric(4200.0, 30000.0, hz, 5)])
    air = -_geometric(0.0, 50000.0, hz, 6)[::-1]
    zn = np.concatenate([air, earth])
    pts, tets, _, _ = kuhn_grid(np.round(xn), np.round(xn), np.round(zn))
    c = centroids(pts, tets)
    conds = [c[:, 2] < 0] + [c[:, 2] < tp for tp in tops[1:]]
    attr = np.select(conds, list(range(len(conds))), len(conds)).astype(np.int32)
    rho_t = np.array([[AIR_RHO]] + [[r] for r in rho])
    g = np.linspace(-3000.0, 3000.0, 6)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rx = np.stack([X.ravel(), Y.ravel(), np.ones(X.size)], axis=1)
    obs = mt_obs(len(freqs), len(rx), types=(Z_XY_RI, Z_YX_RI))
    m = Model(pts, tets, attr, 0, rho_t, np.asarray(freqs, dtype=float), np.zeros((0, 7)), rx, obs, name="layered_mt")
    target = np.zeros(len(obs), complex)
    for o, (code, f, _, r) in enumerate(obs):
        Z = layered_impedance(list(tops), list(rho), m.freqs[f], rx[r, 2])
        target[o] = Z if code == Z_XY_RI else -Z
    return m, target
