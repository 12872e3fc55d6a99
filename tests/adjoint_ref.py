"""Reference gradients for the adjoint tests, from the oracle's element
evaluator (oracle.eval_element, CPU FP64).

For L(K) = sum_ij G_ij K_ij and one element e, dL/dp = <G_e, dK_e/dp> where
G_e is G restricted to the element's global dofs (0 outside the pattern).
  * exact: the stiffness terms are homogeneous in E, thickness, area, It and the
    shear factor, so those derivatives are contractions with the oracle's
    separately evaluated terms (no differencing);
  * finite differences: Richardson-extrapolated central differences of the
    oracle's K_e for everything else (nu, I11, I22, node coordinates), and
    plain central differences with a relative step for the FD-h check.
"""

import numpy as np
import scipy.sparse as sp

from oracle import oracle as orc

TRUSS, BEAM, PLATE = 0, 1, 2
NODES_PER = {TRUSS: 2, BEAM: 2, PLATE: 4}


def element_dofs(etype, idx):
    per = 3 if etype == TRUSS else 6
    return np.array([6 * a + i for a in idx for i in range(per)])


class ElementGrad:
    """G_e blocks from a CSR-shaped dL/dK, or -u_e u_e^T from displacements."""

    def __init__(self, mesh, rp=None, ci=None, g=None, u=None):
        self.mesh = mesh
        if g is not None:
            n = len(rp) - 1
            self.G = sp.csr_matrix((g, ci, rp), shape=(n, n))
        self.u = u

    def block(self, etype, e):
        idx = self.mesh.conn[e].astype(np.int64) - 1
        dofs = element_dofs(etype, idx)
        if self.u is not None:
            ue = self.u[dofs]
            return -np.outer(ue, ue)
        return self.G[dofs][:, dofs].toarray()


def _xyz(mesh, e):
    return mesh.xyz[mesh.conn[e].astype(np.int64) - 1].copy()


def l_elem(mesh, etype, e, Ge, xyz=None, params=None, parts=None):
    X = _xyz(mesh, e) if xyz is None else xyz
    P = mesh.props[e] if params is None else params
    return float(np.sum(Ge * orc.eval_element(etype, X, P, mesh.rel_tol, mesh.abs_tol, parts)))


def richardson(f, h):
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def param_fd(mesh, etype, e, Ge, i, rel=1e-3):
    p0 = mesh.props[e].astype(np.float64)

    def f(h):
        p = p0.copy()
        p[i] += h
        return l_elem(mesh, etype, e, Ge, params=p)

    return richardson(f, rel * abs(p0[i]))


def coord_fd(mesh, etype, e, Ge, a, k, h):
    X0 = _xyz(mesh, e)

    def f(d):
        X = X0.copy()
        X[a, k] += d
        return l_elem(mesh, etype, e, Ge, xyz=X)

    return richardson(f, h)


def plate_param_grads(mesh, e, Ge):
    """dL/d(E, nu, t, k) of plate e: E, t and k exact, nu by differences."""
    E, nu, t, k = mesh.props[e]
    lm = l_elem(mesh, PLATE, e, Ge, parts=orc.PLATE_MEMBRANE)
    lb = l_elem(mesh, PLATE, e, Ge, parts=orc.PLATE_BENDING)
    ls = l_elem(mesh, PLATE, e, Ge, parts=orc.PLATE_SHEAR)
    return np.array([(lm + lb + ls) / E, param_fd(mesh, PLATE, e, Ge, 1), (lm + 3 * lb + ls) / t, ls / k])


def truss_param_grads(mesh, e, Ge):
    E, A, A2 = mesh.props[e]
    le = l_elem(mesh, TRUSS, e, Ge)
    area = A if np.isnan(A2) else 0.5 * (A + A2)
    ga = le / area
    return np.array([le / E, ga if np.isnan(A2) else 0.5 * ga, 0.0 if np.isnan(A2) else 0.5 * ga])


def beam_param_grads(mesh, e, Ge):
    """dL/d(E, nu, A, I11, I22, It, k) of beam e (I12 = 0 meshes): E, A, It, k
    exact from the rank-one terms, nu, I11, I22 by differences."""
    p = mesh.props[e]
    E, A, it, k = p[0], p[2], p[6], p[7]
    lt = [l_elem(mesh, BEAM, e, Ge, parts=1 << t) for t in range(6)]
    le = sum(lt)
    return {0: le / E, 1: param_fd(mesh, BEAM, e, Ge, 1), 2: (lt[0] + lt[1] + lt[2]) / A,
            3: param_fd(mesh, BEAM, e, Ge, 3), 4: param_fd(mesh, BEAM, e, Ge, 4), 6: lt[3] / it,
            7: (lt[1] + lt[2]) / k}


def node_coord_grads(mesh, etype, eg, h, elements=None):
    """dL/dX (N x 3) summed over the given elements (default: all)."""
    out = np.zeros((mesh.n_nodes, 3))
    elements = range(mesh.n_elements) if elements is None else elements
    for e in elements:
        Ge = eg.block(etype, e)
        idx = mesh.conn[e].astype(np.int64) - 1
        for a in range(NODES_PER[etype]):
            for k in range(3):
                out[idx[a], k] += coord_fd(mesh, etype, e, Ge, a, k, h)
    return out


def rel_err(got, ref):
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref))
    return float(np.max(np.abs(np.asarray(got) - ref)) / scale) if scale > 0 else float(np.max(np.abs(got)))
