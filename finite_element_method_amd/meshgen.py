"""Seeded synthetic meshes for the assembly workloads.

Every generator returns a ``Mesh``: node numbers and coordinates in insertion
order (node index = position, node number = index + 1, as the crate assigns
indices in add order, src/fem/methods_for_node_data_handle.rs:66-79) and the
elements of one type as SoA arrays in the layout of the batched C-ABI adds.

Workloads:
  * truss_lattice(11)      3D lattice, bars on axis edges + face diagonals
  * plate_grid(512)        jittered Mindlin plate, 4-node quads in z = 0
  * beam_frame(64)         3D frame, 2-node beams on every axis edge
  * batched_meshes(256)    uneven meshes of mixed type
"""

from dataclasses import dataclass, field

import numpy as np

E_STEEL = 210e9
E_ALU = 70e9


@dataclass
class Mesh:
    kind: str                    # "truss" | "beam" | "plate"
    node_numbers: np.ndarray     # (N,) uint32
    xyz: np.ndarray              # (N, 3) float64
    numbers: np.ndarray          # (E,) uint32 element numbers
    conn: np.ndarray             # (E, nn) uint32 node numbers
    props: np.ndarray            # truss: (E, 3) E, A, A2(NaN=None); beam: (E, 11); plate: (E, 4)
    rel_tol: float = 1e-4
    abs_tol: float = 1e-12
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self):
        return len(self.node_numbers)

    @property
    def n_elements(self):
        return len(self.numbers)


def truss_lattice(n=11, seed=0, spacing=1.0):
    """n^3 nodes; bars on all axis edges, one diagonal per unit face square
    (direction by seed) and the crossing diagonal with probability 0.5."""
    rng = np.random.default_rng(seed)
    idx = np.arange(n ** 3).reshape(n, n, n)  # [k, j, i] -> index i + n j + n^2 k
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)  # k, j, i
    xyz = np.stack([g[..., 2], g[..., 1], g[..., 0]], -1).reshape(-1, 3).astype(np.float64) * spacing
    bars = []
    bars.append(np.stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()], 1))  # x edges
    bars.append(np.stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()], 1))  # y edges
    bars.append(np.stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()], 1))  # z edges
    faces = []
    # xy faces: corners (k, j, i), (k, j+1, i+1) and (k, j, i+1), (k, j+1, i)
    faces.append((idx[:, :-1, :-1], idx[:, 1:, 1:], idx[:, :-1, 1:], idx[:, 1:, :-1]))
    faces.append((idx[:-1, :, :-1], idx[1:, :, 1:], idx[:-1, :, 1:], idx[1:, :, :-1]))  # xz faces
    faces.append((idx[:-1, :-1, :], idx[1:, 1:, :], idx[:-1, 1:, :], idx[1:, :-1, :]))  # yz faces
    for a, b, c, d in faces:
        a, b, c, d = a.ravel(), b.ravel(), c.ravel(), d.ravel()
        flip = rng.random(a.size) < 0.5
        first = np.where(flip[:, None], np.stack([c, d], 1), np.stack([a, b], 1))
        second = np.where(flip[:, None], np.stack([a, b], 1), np.stack([c, d], 1))
        bars.append(first)
        both = rng.random(a.size) < 0.5
        bars.append(second[both])
    conn_idx = np.concatenate(bars, 0)
    ne = len(conn_idx)
    A = rng.uniform(1e-4, 5e-4, ne)
    props = np.stack([np.full(ne, E_STEEL), A, np.full(ne, np.nan)], 1)
    return Mesh("truss", np.arange(1, n ** 3 + 1, dtype=np.uint32), xyz,
                np.arange(1, ne + 1, dtype=np.uint32), (conn_idx + 1).astype(np.uint32), props,
                meta={"workload": "truss_lattice", "n": n, "seed": seed})


def plate_grid(nx=512, ny=None, seed=0, jitter=0.1, thickness=(0.005, 0.02), E=E_ALU, nu=0.3,
               shear=5.0 / 6.0, lx=1.0, ly=1.0, rotation=None, shuffle=False):
    """(nx+1)(ny+1) nodes on [0,lx]x[0,ly] in z = 0, interior nodes jittered by
    U[-jitter, jitter] * spacing in x and y; quads ordered 1(+,+) 2(-,+) 3(-,-) 4(+,-).

    rotation: optional 3x3 matrix applied to the coordinates (a tilted plate);
    shuffle: insert nodes and elements in a random order (irregular numbering)."""
    ny = nx if ny is None else ny
    rng = np.random.default_rng(seed)
    hx, hy = lx / nx, ly / ny
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    x = ii * hx
    y = jj * hy
    interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)
    x = x + np.where(interior, rng.uniform(-jitter, jitter, x.shape) * hx, 0.0)
    y = y + np.where(interior, rng.uniform(-jitter, jitter, y.shape) * hy, 0.0)
    xyz = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], 1)
    nid = np.arange((nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
    n3 = nid[:-1, :-1]
    n4 = nid[:-1, 1:]
    n1 = nid[1:, 1:]
    n2 = nid[1:, :-1]
    conn_idx = np.stack([n1.ravel(), n2.ravel(), n3.ravel(), n4.ravel()], 1)
    ne = len(conn_idx)
    t = rng.uniform(thickness[0], thickness[1], ne)
    props = np.stack([np.full(ne, E), np.full(ne, nu), t, np.full(ne, shear)], 1)
    if rotation is not None:
        xyz = xyz @ np.asarray(rotation, dtype=np.float64).T
    if shuffle:
        perm = rng.permutation(xyz.shape[0])          # new position -> old index
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        xyz = xyz[perm]
        conn_idx = inv[conn_idx]
        eperm = rng.permutation(ne)
        conn_idx, props = conn_idx[eperm], props[eperm]
    return Mesh("plate", np.arange(1, xyz.shape[0] + 1, dtype=np.uint32), xyz,
                np.arange(1, ne + 1, dtype=np.uint32), (conn_idx + 1).astype(np.uint32), props,
                meta={"workload": "plate_grid", "nx": nx, "ny": ny, "seed": seed})


def _log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def beam_frame(nx=64, ny=None, nz=None, seed=0, spacing=1.0, E=E_STEEL, nu=0.3, shear=5.0 / 6.0,
               i12_frac=0.0, equal_frac=0.0):
    """nx*ny*nz nodes at unit spacing, 2-node beams on every axis edge; A, I11,
    I22, It log-uniform over two decades, I12 = 0, local axis 1 a random unit
    vector whose component perpendicular to the beam is at least 0.1.

    i12_frac: fraction of beams with a product of inertia I12 ~ U(-0.4, 0.4) *
    sqrt(I11 I22) (rotated principal axes, beam.rs:119-160); equal_frac:
    fraction of beams with I22 = I11 (the crate's rel_tol branch)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    rng = np.random.default_rng(seed)
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    g = np.stack(np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij"), -1)
    xyz = np.stack([g[..., 2], g[..., 1], g[..., 0]], -1).reshape(-1, 3).astype(np.float64) * spacing
    parts, dirs = [], []
    parts.append(np.stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()], 1)); dirs.append(0)
    parts.append(np.stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()], 1)); dirs.append(1)
    parts.append(np.stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()], 1)); dirs.append(2)
    conn_idx = np.concatenate(parts, 0)
    axis = np.concatenate([np.full(len(p), d) for p, d in zip(parts, dirs)])
    ne = len(conn_idx)
    v1 = rng.normal(size=(ne, 3))
    v1 /= np.linalg.norm(v1, axis=1, keepdims=True)
    for _ in range(100):
        perp = np.sqrt(np.maximum(0.0, 1.0 - v1[np.arange(ne), axis] ** 2))
        bad = perp < 0.1
        if not bad.any():
            break
        nv = rng.normal(size=(bad.sum(), 3))
        v1[bad] = nv / np.linalg.norm(nv, axis=1, keepdims=True)
    A = _log_uniform(rng, 1e-4, 1e-2, ne)
    i11 = _log_uniform(rng, 1e-8, 1e-6, ne)
    i22 = _log_uniform(rng, 1e-8, 1e-6, ne)
    it = _log_uniform(rng, 1e-8, 1e-6, ne)
    i12 = np.zeros(ne)
    if equal_frac > 0.0:
        eq = rng.random(ne) < equal_frac
        i22 = np.where(eq, i11, i22)
    if i12_frac > 0.0:
        on = rng.random(ne) < i12_frac
        i12 = np.where(on, rng.uniform(-0.4, 0.4, ne) * np.sqrt(i11 * i22), 0.0)
    props = np.stack([np.full(ne, E), np.full(ne, nu), A, i11, i22, i12, it, np.full(ne, shear),
                      v1[:, 0], v1[:, 1], v1[:, 2]], 1)
    return Mesh("beam", np.arange(1, xyz.shape[0] + 1, dtype=np.uint32), xyz,
                np.arange(1, ne + 1, dtype=np.uint32), (conn_idx + 1).astype(np.uint32), props,
                meta={"workload": "beam_frame", "n": (nx, ny, nz), "seed": seed})


def batched_meshes(count=256, seed=0, min_elems=1_000, max_elems=200_000):
    """`count` independent meshes, type uniform over {truss lattice, beam frame,
    plate}, element counts log-uniform in [min_elems, max_elems]."""
    rng = np.random.default_rng(seed)
    meshes = []
    for m in range(count):
        kind = ["truss", "beam", "plate"][rng.integers(3)]
        target = float(np.exp(rng.uniform(np.log(min_elems), np.log(max_elems))))
        s = int(rng.integers(1 << 30))
        if kind == "plate":
            n = max(2, int(round(np.sqrt(target))))
            meshes.append(plate_grid(n, seed=s))
        elif kind == "beam":
            n = max(2, int(round((target / 3.0) ** (1.0 / 3.0))) + 1)
            meshes.append(beam_frame(n, seed=s))
        else:
            n = max(2, int(round((target / 6.4) ** (1.0 / 3.0))) + 1)
            meshes.append(truss_lattice(n, seed=s))
    return meshes


def load_oracle(mesh, oracle_model_cls, f32=False):
    """Feed a mesh into the oracle model class (tests and the CPU baseline)."""
    m = oracle_model_cls(mesh.rel_tol, mesh.abs_tol, mesh.n_nodes, f32=f32)
    m.add_nodes(mesh.node_numbers, mesh.xyz)
    if mesh.kind == "truss":
        m.add_trusses(mesh.numbers, mesh.conn, mesh.props[:, 0], mesh.props[:, 1], mesh.props[:, 2])
    elif mesh.kind == "beam":
        m.add_beams(mesh.numbers, mesh.conn, mesh.props)
    else:
        m.add_plates(mesh.numbers, mesh.conn, mesh.props)
    return m
