"""Seeded synthetic meshes and load cases for the five workloads.

The reference reads meshes from files (BasicMesh::ReadMesh, BasicMesh.cpp:
79-190); here the same data (nodal coordinates, element connectivity in the
reference node order, per-element material records, restrained DOFs with
prescribed values, nodal point loads) are generated from a seed with numpy's
PCG64 generator.  Draw order is fixed and documented per workload so that the
CPU oracle and the device path consume identical inputs.

Node numbering is x-fastest: node(i,j,k) = i + (nx+1)*(j + (ny+1)*k).
Hex8 local node order follows the reference (BasicElement.cpp:518-525):
(xi,eta,zeta) = (---),(-+-),(-++),(--+),(+--),(++-),(+++),(+-+) with
xi->x, eta->y, zeta->z.  Quad4 is counter-clockwise from (-1,-1).
"""
from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 20261203  # seed of workload k is SEED_BASE + k


@dataclass
class Problem:
    name: str
    dim: int
    xyz: np.ndarray          # (n_nodes, dim)
    conn: np.ndarray         # (n_el, nnpe) int32
    ndof: int                # DOFs per node (uniform)
    mat_kind: int            # capi.MAT_*
    mat: np.ndarray          # (n_groups, stride)
    restrained: np.ndarray   # (n_dof,) uint8
    prescribed: np.ndarray   # (n_dof,) float64
    load: np.ndarray         # (n_dof,) float64
    order: int = 2
    factor: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    @property
    def n_el(self):
        return self.conn.shape[0]

    @property
    def n_dof(self):
        return self.n_nodes * self.ndof


def structured_hex8(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # element index e = i + nx*(j + ny*k): order by k, then j, then i
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()

    def nid(a, b, c):
        return a + (nx + 1) * (b + (ny + 1) * c)

    conn = np.stack([
        nid(i, j, k), nid(i, j + 1, k), nid(i, j + 1, k + 1), nid(i, j, k + 1),
        nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i + 1, j + 1, k + 1), nid(i + 1, j, k + 1),
    ], axis=1).astype(np.int32)
    return xyz, conn


def structured_quad4(nx, ny, lx=1.0, ly=1.0):
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()

    def nid(a, b):
        return a + (nx + 1) * b

    conn = np.stack([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)], axis=1).astype(np.int32)
    return xy, conn


def boundary_mask_hex(nx, ny, nz):
    i, j, k = _node_ijk(nx, ny, nz)
    return (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)


def _node_ijk(nx, ny, nz):
    n = np.arange((nx + 1) * (ny + 1) * (nz + 1))
    i = n % (nx + 1)
    j = (n // (nx + 1)) % (ny + 1)
    k = n // ((nx + 1) * (ny + 1))
    return i, j, k


def plane_stress_quad(n=64, seed=None):
    """Workload 1: unit square, n x n quad4, plane stress E=1 nu=0.3,
    thickness 1, x=0 edge clamped, unit x-traction on x=1 as consistent
    nodal loads (h/2 at the two corners, h at interior edge nodes)."""
    xy, conn = structured_quad4(n, n)
    nn = xy.shape[0]
    h = 1.0 / n
    E, nu = 1.0, 0.3
    G = E / (2 * (1 + nu))
    mat = np.array([[E, E, nu, G]])
    rest = np.zeros(2 * nn, np.uint8)
    left = np.where(np.isclose(xy[:, 0], 0.0))[0]
    rest[2 * left] = 1
    rest[2 * left + 1] = 1
    load = np.zeros(2 * nn)
    right = np.where(np.isclose(xy[:, 0], 1.0))[0]
    w = np.full(len(right), h)
    w[np.isclose(xy[right, 1], 0.0) | np.isclose(xy[right, 1], 1.0)] = h / 2
    load[2 * right] = w
    from . import capi
    return Problem("plane_stress_quad4_%d" % n, 2, xy, conn, 2, capi.MAT_PLANE_STRESS, mat, rest,
                   np.zeros(2 * nn), load, order=2, factor=1.0)


def elastic_cube(n=96, seed=None, jitter=0.1, e_range=(1.0, 100.0), nu_range=(0.2, 0.35)):
    """Workload 2: unit cube, n^3 hex8.  Draws, in order: E (n_el, log-uniform
    in e_range), nu (n_el, uniform), interior-node jitter (n_nodes x 3,
    uniform in +-jitter*h; boundary nodes are not moved).  z=0 clamped, unit
    pressure on z=1 in -z as consistent bilinear nodal loads (h^2/4 corner,
    h^2/2 edge, h^2 interior)."""
    seed = SEED_BASE + 2 if seed is None else seed
    rng = np.random.default_rng(seed)
    xyz, conn = structured_hex8(n, n, n)
    ne = conn.shape[0]
    nn = xyz.shape[0]
    h = 1.0 / n
    E = np.exp(rng.uniform(np.log(e_range[0]), np.log(e_range[1]), ne))
    nu = rng.uniform(nu_range[0], nu_range[1], ne)
    jit = rng.uniform(-jitter * h, jitter * h, (nn, 3))
    interior = ~boundary_mask_hex(n, n, n)
    xyz[interior] += jit[interior]
    mat = np.stack([E, nu], axis=1)
    rest = np.zeros(3 * nn, np.uint8)
    bottom = np.where(np.isclose(xyz[:, 2], 0.0))[0]
    for c in range(3):
        rest[3 * bottom + c] = 1
    load = np.zeros(3 * nn)
    top = np.where(np.isclose(xyz[:, 2], 1.0))[0]
    ex = np.isclose(xyz[top, 0], 0.0) | np.isclose(xyz[top, 0], 1.0)
    ey = np.isclose(xyz[top, 1], 0.0) | np.isclose(xyz[top, 1], 1.0)
    w = np.full(len(top), h * h)
    w[ex ^ ey] = h * h / 2
    w[ex & ey] = h * h / 4
    load[3 * top + 2] = -w
    from . import capi
    return Problem("elastic_hex8_%d" % n, 3, xyz, conn, 3, capi.MAT_ISO_ELASTIC, mat, rest,
                   np.zeros(3 * nn), load, order=2)


def diffusion_cube(n=128, seed=None, k_range=(0.1, 10.0), sigma=0.1):
    """Workload 3: unit cube, n^3 hex8, 1 DOF/node.  Draws: k per element
    (log-uniform in k_range); capacity 1.  u0 = exp(-|x-c|^2/(2 sigma^2)) at
    interior nodes and 0 on the boundary, where u = 0 is prescribed."""
    seed = SEED_BASE + 3 if seed is None else seed
    rng = np.random.default_rng(seed)
    xyz, conn = structured_hex8(n, n, n)
    ne = conn.shape[0]
    nn = xyz.shape[0]
    k = np.exp(rng.uniform(np.log(k_range[0]), np.log(k_range[1]), ne))
    mat = np.stack([k, k, k, np.ones(ne)], axis=1)
    bnd = boundary_mask_hex(n, n, n)
    rest = bnd.astype(np.uint8)
    c = np.array([0.5, 0.5, 0.5])
    u0 = np.exp(-np.sum((xyz - c) ** 2, axis=1) / (2 * sigma ** 2))
    u0[bnd] = 0.0
    from . import capi
    return Problem("diffusion_hex8_%d" % n, 3, xyz, conn, 1, capi.MAT_HEAT, mat, rest,
                   np.zeros(nn), np.zeros(nn), order=2, extra={"u0": u0, "dt": 1e-3, "steps": 100})


def plastic_bar(n=64, E=200.0, nu=0.3, sigma0=0.25, H=2.0, strain=0.005, increments=20):
    """Workload 4: unit cube, n^3 hex8, J2 with linear isotropic hardening.
    Rollers: u_z=0 on z=0, u_x=0 on x=0, u_y=0 on y=0; u_z prescribed on z=1,
    reaching `strain` in `increments` equal steps."""
    xyz, conn = structured_hex8(n, n, n)
    nn = xyz.shape[0]
    rest = np.zeros(3 * nn, np.uint8)
    rest[3 * np.where(np.isclose(xyz[:, 0], 0.0))[0] + 0] = 1
    rest[3 * np.where(np.isclose(xyz[:, 1], 0.0))[0] + 1] = 1
    rest[3 * np.where(np.isclose(xyz[:, 2], 0.0))[0] + 2] = 1
    top = np.where(np.isclose(xyz[:, 2], 1.0))[0]
    rest[3 * top + 2] = 1
    presc_inc = np.zeros(3 * nn)
    presc_inc[3 * top + 2] = strain / increments
    mat = np.array([[E, nu, sigma0, H]])
    from . import capi
    return Problem("plastic_hex8_%d" % n, 3, xyz, conn, 3, capi.MAT_J2, mat, rest, presc_inc,
                   np.zeros(3 * nn), order=2,
                   extra={"increments": increments, "strain": strain, "E": E, "nu": nu,
                          "sigma0": sigma0, "H": H})


_RVE_BASE = {}


def _rve_base(n):
    """Geometry shared by every RVE request of size n (cached): nodes,
    connectivity, element centroids as an (n, n, n, 3) [k][j][i] view,
    restraints and the six affine load cases."""
    b = _RVE_BASE.get(n)
    if b is not None:
        return b
    xyz, conn = structured_hex8(n, n, n)
    nn = xyz.shape[0]
    cent = xyz[conn].mean(axis=1)
    bnd = boundary_mask_hex(n, n, n)
    rest = np.zeros(3 * nn, np.uint8)
    for c in range(3):
        rest[3 * np.where(bnd)[0] + c] = 1
    rb = rest.astype(bool)
    cases = []
    for k in range(6):
        eps = np.zeros((3, 3))
        if k < 3:
            eps[k, k] = 1.0
        else:
            a, bb = [(0, 1), (1, 2), (0, 2)][k - 3]
            eps[a, bb] = eps[bb, a] = 0.5
        u = xyz @ eps.T
        pv = np.zeros(3 * nn)
        pv[rb] = u.ravel()[rb]
        pv.flags.writeable = False
        cases.append(pv)
    for a in (xyz, conn, rest):
        a.flags.writeable = False
    # every size of the stream (16..96) stays cached: rebuilding an evicted
    # base cost ~90 ms per request; all 81 bases take ~5 GB of host memory
    if len(_RVE_BASE) >= 128:
        _RVE_BASE.clear()
    b = {"xyz": xyz, "conn": conn, "cent": cent.reshape(n, n, n, 3), "rest": rest, "cases": cases}
    _RVE_BASE[n] = b
    return b


def rve_request(n, seed, n_incl=None, r_range=(0.05, 0.15)):
    """One request of workload 5: n^3 unit cube with spherical inclusions
    (E=70, nu=0.22 inside, E=3, nu=0.35 outside, by element centroid).
    Draws: number of inclusions (5..40) unless given, centres (U(0,1)^3),
    radii (uniform in r_range).  Every boundary DOF is prescribed; the six
    load cases are the affine fields u = eps . x of the unit macro-strains
    (xx, yy, zz, xy, yz, xz with engineering shear split symmetrically).
    Geometry, restraints and load cases are shared (read-only) between
    requests of the same n; each inclusion tests only the elements in its
    bounding box (one element of margin), which gives the same mask as
    testing every centroid."""
    rng = np.random.default_rng(seed)
    if n_incl is None:
        n_incl = int(rng.integers(5, 41))
    cen = rng.uniform(0.0, 1.0, (n_incl, 3))
    rad = rng.uniform(r_range[0], r_range[1], n_incl)
    b = _rve_base(n)
    cent = b["cent"]
    inside = np.zeros((n, n, n), bool)  # [k][j][i]
    for c, r in zip(cen, rad):
        lo = np.clip(np.floor((c - r) * n).astype(int) - 1, 0, n)
        hi = np.clip(np.ceil((c + r) * n).astype(int) + 1, 0, n)
        sub = cent[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        inside[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] |= np.sum((sub - c) ** 2, axis=-1) < r * r
    mats = np.array([[3.0, 0.35], [70.0, 0.22]])
    group = inside.ravel().astype(np.int32)
    nn = b["xyz"].shape[0]
    from . import capi
    return Problem("rve_hex8_%d" % n, 3, b["xyz"], b["conn"], 3, capi.MAT_ISO_ELASTIC, mats, b["rest"],
                   b["cases"][0], np.zeros(3 * nn), order=2,
                   extra={"cases": b["cases"], "group": group, "n_incl": n_incl})
