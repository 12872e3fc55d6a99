"""Boundary-condition cases shared by the separation tests: which dofs to
constrain and load on a mesh, applied identically to the GPU model and to the
oracle."""

import numpy as np


def diagonal(rp, ci, v):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    d = np.zeros(n)
    on = ci == rows
    d[rows[on]] = v[on]
    return d


def clamp_face(mesh, diag, axis=0, shift=None):
    """Constrains every active dof (non-zero diagonal) of the nodes on the
    lowest face along `axis`.  shift: imposed rigid translation (3,) or None
    (zero displacements).  Returns (numbers, dofs, values)."""
    x = mesh.xyz[:, axis]
    face = np.nonzero(np.isclose(x, x.min()))[0]
    numbers, dofs, values = [], [], []
    for i in face:
        for k in range(6):
            if diag[6 * i + k] != 0.0:
                numbers.append(int(mesh.node_numbers[i]))
                dofs.append(k)
                values.append(0.0 if shift is None or k >= 3 else float(shift[k]))
    return numbers, dofs, values


def random_bc(mesh, diag, seed, frac_fixed=0.05, frac_loaded=0.2):
    """Random constraints on active dofs (at least one) with random imposed
    values, and random loads (repeated dofs included, so they add up)."""
    rng = np.random.default_rng(seed)
    active = np.nonzero(diag != 0.0)[0]
    k = max(1, int(frac_fixed * len(active)))
    fixed = rng.choice(active, size=k, replace=False)
    disp = ([int(mesh.node_numbers[g // 6]) for g in fixed], [int(g % 6) for g in fixed],
            list(rng.normal(size=k) * 1e-3))
    m = max(1, int(frac_loaded * mesh.n_nodes))
    where = rng.integers(0, 6 * mesh.n_nodes, size=m)
    where = np.concatenate([where, where[: m // 4]])  # some dofs loaded twice
    loads = ([int(mesh.node_numbers[g // 6]) for g in where], [int(g % 6) for g in where],
             list(rng.normal(size=len(where)) * 1e3))
    return disp, loads


def apply(model, disp, loads):
    """Applies displacements then loads, one call per entry (the crate's API)."""
    for n, d, v in zip(*disp):
        model.add_displacement(n, d, v)
    for n, d, v in zip(*loads):
        model.add_concentrated_load(n, d, v)
