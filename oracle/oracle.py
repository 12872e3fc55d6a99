"""ctypes wrapper of the CPU oracle (oracle/fem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  It restates the
crate `finite_element_method`'s element-stiffness and assembly path; see the
header of fem_oracle.cpp for what each function follows.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfem_oracle.so")
_lib = None


class OracleError(Exception):
    """Raised with the crate-identical message when the oracle rejects an input."""


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, u32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_double
        L.om_create.restype = vp
        L.om_create.argtypes = [dbl, dbl, u32, ctypes.c_int]
        L.om_destroy.argtypes = [vp]
        L.om_last_error.restype = ctypes.c_char_p
        L.om_last_error.argtypes = [vp]
        for name, nargs in [("om_add_nodes", 3), ("om_add_trusses", 6), ("om_add_beams", 4), ("om_add_plates", 4)]:
            f = getattr(L, name)
            f.restype = i64
            f.argtypes = [vp, i64] + [vp] * (nargs - 1)
        L.om_get_rotation.restype = ctypes.c_int
        L.om_get_rotation.argtypes = [vp, ctypes.c_int, u32, vp]
        L.om_element_matrix.restype = ctypes.c_int
        L.om_element_matrix.argtypes = [vp, ctypes.c_int, u32, ctypes.c_int, vp]
        L.om_principal_moments.argtypes = [dbl, dbl, dbl, dbl, vp]
        L.om_quad_on_line.restype = ctypes.c_int
        L.om_quad_on_line.argtypes = [vp]
        L.om_quad_on_plane.restype = ctypes.c_int
        L.om_quad_on_plane.argtypes = [vp, dbl]
        L.om_convex_hull.restype = ctypes.c_int
        L.om_convex_hull.argtypes = [vp, vp, dbl, vp]
        L.om_rust_debug.restype = ctypes.c_int
        L.om_rust_debug.argtypes = [dbl, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.om_export.restype = i64
        L.om_export.argtypes = [vp, ctypes.c_int]
        L.om_ndof.restype = i64
        L.om_ndof.argtypes = [vp]
        L.om_export_copy.argtypes = [vp, vp, vp, vp]
        L.om_export_magnitude.restype = i64
        L.om_export_magnitude.argtypes = [vp, vp, i64]
        L.om_num_contributions.restype = i64
        L.om_num_contributions.argtypes = [vp]
        L.om_structural_violations.restype = i64
        L.om_structural_violations.argtypes = [vp]
        L.om_nonstructural_max_rel.restype = ctypes.c_double
        L.om_nonstructural_max_rel.argtypes = [vp]
        L.om_add_displacement.restype = ctypes.c_int
        L.om_add_displacement.argtypes = [vp, u32, ctypes.c_int, dbl]
        L.om_add_concentrated_load.restype = ctypes.c_int
        L.om_add_concentrated_load.argtypes = [vp, u32, ctypes.c_int, dbl]
        L.om_add_line_load.restype = ctypes.c_int
        L.om_add_line_load.argtypes = [vp, u32, ctypes.c_int, dbl]
        L.om_add_surface_load.restype = ctypes.c_int
        L.om_add_surface_load.argtypes = [vp, u32, ctypes.c_int, dbl]
        L.om_forces.argtypes = [vp, vp]
        L.om_separate.restype = ctypes.c_int
        L.om_separate.argtypes = [vp, vp]
        L.om_separated_copy.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.om_separated_vectors.argtypes = [vp, vp, vp, vp, vp]
        L.om_num_elements.restype = ctypes.c_int64
        L.om_num_elements.argtypes = [vp, ctypes.c_int]
        L.om_element_forces.restype = ctypes.c_int64
        L.om_element_forces.argtypes = [vp, ctypes.c_int, vp, vp]
        L.om_time_values.restype = ctypes.c_double
        L.om_time_values.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.om_eval_element.restype = ctypes.c_int
        L.om_eval_element.argtypes = [ctypes.c_int, vp, vp, dbl, dbl, ctypes.c_uint, vp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rust_debug(x, f32=False):
    buf = ctypes.create_string_buffer(64)
    lib().om_rust_debug(float(x), int(f32), buf, 64)
    return buf.value.decode()


PLATE_MEMBRANE, PLATE_BENDING, PLATE_SHEAR, PLATE_DRILLING = 1, 2, 4, 8


def eval_element(etype, xyz, params, rel_tol=1e-4, abs_tol=1e-12, parts=None):
    """Global-frame K_e (n x n) of one element from explicit node coordinates
    (nodes_per x 3) and parameters, as add_* computes it (no validation).

    parts: plates, a mask of PLATE_*; beams, bit t keeps rank-one term t
    (u, v, w, theta_u, theta_v, theta_w); None = every term."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    params = np.ascontiguousarray(params, dtype=np.float64)
    if parts is None:
        parts = 15 if etype == 2 else 63
    out = np.zeros(24 * 24)
    n = lib().om_eval_element(int(etype), _ptr(xyz), _ptr(params), float(rel_tol), float(abs_tol), int(parts), _ptr(out))
    if n < 0:
        raise OracleError("no element frame for these coordinates")
    return out[: n * n].reshape(n, n)


def quad_on_line(points):
    """is_points_of_quadrilateral_on_the_same_line (quadrilateral_4n_element_functions.rs:14-86)."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(12)
    return bool(lib().om_quad_on_line(_ptr(p)))


def quad_on_plane(points, abs_tol):
    """is_points_of_quadrilateral_on_the_same_plane (quadrilateral_4n_element_functions.rs:88-129)."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(12)
    return bool(lib().om_quad_on_plane(_ptr(p), float(abs_tol)))


def convex_hull(numbers, points, abs_tol):
    """convex_hull_on_four_points_on_plane (quadrilateral_4n_element_functions.rs:173-250):
    the hull's point numbers, or None when points 2-4 have no frame."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(12)
    nums = np.ascontiguousarray(numbers, dtype=np.uint32).reshape(4)
    out = np.zeros(4, dtype=np.uint32)
    n = lib().om_convex_hull(_ptr(nums), _ptr(p), float(abs_tol), _ptr(out))
    return None if n < 0 else [int(x) for x in out[:n]]


def principal_moments(i11, i22, i12, rel_tol):
    out = np.zeros(3)
    lib().om_principal_moments(i11, i22, i12, rel_tol, _ptr(out))
    return out


DOF_NAMES = ("X", "Y", "Z", "ThX", "ThY", "ThZ")


def separate_csr(row_ptr, col_idx, val, constrained, disp=None, force=None, node_numbers=None):
    """separate_stiffness_matrix_sparse_iterative
    (methods_for_separate_stiffness_matrix.rs:217-320) with compose_r_a_vector /
    compose_u_b_vector (:323-345), restated in numpy over any CSR with sorted
    columns.  Global dof = 6 * node_index + dof; node_numbers[node_index] names
    the node in the error message.  Returns ({"aa","ab","ba","bb": (row_ptr,
    col_idx, val)}, k_aa_indexes, k_bb_indexes, r_a, u_b)."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    ci = np.asarray(col_idx, dtype=np.int64)
    v = np.asarray(val, dtype=np.float64)
    n = len(rp) - 1
    cons = np.asarray(constrained).astype(bool)
    disp = np.zeros(n) if disp is None else np.asarray(disp, dtype=np.float64)
    force = np.zeros(n) if force is None else np.asarray(force, dtype=np.float64)
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = np.zeros(n)
    on = ci == rows
    diag[rows[on]] = v[on]
    bad = np.nonzero((diag == 0.0) & cons)[0]
    if len(bad):  # the crate stops at the first such dof in index order
        idx = int(bad[0])
        number = int(node_numbers[idx // 6]) if node_numbers is not None else 0
        raise OracleError("There are no stiffness to withstand displacement %s applied to node %d!"
                          % (DOF_NAMES[idx % 6], number))
    is_a = (diag != 0.0) & ~cons
    is_b = (diag != 0.0) & cons
    aa, bb = np.nonzero(is_a)[0], np.nonzero(is_b)[0]
    if len(bb) == 0:
        raise OracleError("No restraints")
    pos = np.full(n, -1, dtype=np.int64)
    pos[aa] = np.arange(len(aa))
    pos[bb] = np.arange(len(bb))
    keep = v != 0.0
    mats = {}
    for name, rmask, cmask, nrows in (("aa", is_a, is_a, len(aa)), ("ab", is_a, is_b, len(aa)),
                                      ("ba", is_b, is_a, len(bb)), ("bb", is_b, is_b, len(bb))):
        sel = keep & rmask[rows] & cmask[ci]
        r_loc = pos[rows[sel]]
        out_rp = np.zeros(nrows + 1, dtype=np.int64)
        out_rp[1:] = np.cumsum(np.bincount(r_loc, minlength=nrows))
        mats[name] = (out_rp, pos[ci[sel]].astype(np.int32), v[sel].copy())
    if len(mats["aa"][2]) == 0:
        raise OracleError("Sparse separation: K_aa is empty (structure has no free stiffness?)")
    return mats, aa.astype(np.int64), bb.astype(np.int64), force[aa].copy(), disp[bb].copy()


# ---------------------------------------------------------------- global analysis
# methods_for_global_analysis.rs:28-133, 135-275.  pcg_jacobi_csr and
# pcg_block_jacobi_csr live in the absent crate iterative_solvers_smpl 0.1.5;
# they are restated as the textbook preconditioned conjugate gradient (Saad,
# Iterative Methods for Sparse Linear Systems, 2nd ed., Algorithm 9.1) with
# x0 = 0, stopped when ||r||_2 <= max(rel_tol * ||b||_2, abs_tol) and
# counting the x updates.  The stopping rule and the block inverse are
# assumptions (parity unpinned there); the tests pin the restatement with the
# crate's own 1-iteration truss tests, direct solves and CG invariants.

def csr_matvec(mat, x, n_rows):
    rp, ci, v = mat
    rows = np.repeat(np.arange(n_rows), np.diff(rp))
    y = np.zeros(n_rows)
    np.add.at(y, rows, v * x[ci])
    return y


def find_b_sparse(r_a, k_ab, u_b):
    """methods_for_global_analysis.rs:28-47: b = r_a - K_ab u_b."""
    return np.asarray(r_a, dtype=np.float64) - csr_matvec(k_ab, np.asarray(u_b, dtype=np.float64), len(r_a))


def block_starts(k_aa_indexes):
    """methods_for_global_analysis.rs:135-160: K_aa rows grouped by node."""
    aa = np.asarray(k_aa_indexes)
    if len(aa) == 0:
        return np.zeros(0, dtype=np.int64)
    node = aa // 6
    return np.concatenate([[0], np.nonzero(np.diff(node))[0] + 1]).astype(np.int64)


def pcg(k_aa, b, max_iter, rel_tol, abs_tol, starts=None):
    """Preconditioned CG on K_aa x = b: Jacobi (starts is None) or block
    Jacobi over [starts[i], starts[i + 1]).  Returns (x, iterations,
    converged)."""
    n = len(b)
    rp, ci, v = k_aa
    rows = np.repeat(np.arange(n), np.diff(rp))
    if starts is None:
        d = np.zeros(n)
        on = rows == ci
        d[rows[on]] = v[on]
        prec = lambda r: r / d
    else:
        bounds = list(starts) + [n]
        inv = []
        for s0, s1 in zip(bounds[:-1], bounds[1:]):
            blk = np.zeros((s1 - s0, s1 - s0))
            for r in range(s0, s1):
                for k in range(rp[r], rp[r + 1]):
                    if s0 <= ci[k] < s1:
                        blk[r - s0, ci[k] - s0] = v[k]
            inv.append((s0, s1, np.linalg.inv(blk)))

        def prec(r):
            z = np.empty_like(r)
            for s0, s1, bi in inv:
                z[s0:s1] = bi @ r[s0:s1]
            return z
    x = np.zeros(n)
    r = np.asarray(b, dtype=np.float64).copy()
    stop = max(rel_tol * np.linalg.norm(b), abs_tol)
    if np.linalg.norm(r) <= stop:
        return x, 0, True
    z = prec(r)
    p = z.copy()
    rz = r @ z
    for it in range(1, max_iter + 1):
        ap = np.zeros(n)
        np.add.at(ap, rows, v * p[ci])
        alpha = rz / (p @ ap)
        x += alpha * p
        r -= alpha * ap
        if np.linalg.norm(r) <= stop:
            return x, it, True
        z = prec(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iter, False


def find_r_r_sparse(k_ba, u_a, k_bb, u_b, r_b):
    """methods_for_global_analysis.rs:92-133: r_r = K_ba u_a + K_bb u_b - r_b."""
    n_bb = len(u_b)
    return csr_matvec(k_ba, np.asarray(u_a), n_bb) + csr_matvec(k_bb, np.asarray(u_b), n_bb) - np.asarray(r_b)


class OracleModel:
    """The crate's FEM<V> (src/fem/fem.rs:15-65) restated on the CPU."""

    TRUSS, BEAM, PLATE = 0, 1, 2

    def __init__(self, rel_tol, abs_tol, nodes_number, f32=False):
        self._h = lib().om_create(rel_tol, abs_tol, nodes_number, int(f32))
        self.nodes_number = nodes_number

    def __del__(self):
        if getattr(self, "_h", None):
            lib().om_destroy(self._h)
            self._h = None

    def _err(self):
        return lib().om_last_error(self._h).decode()

    def _check(self, n_ok, n):
        if n_ok != n:
            raise OracleError(self._err())
        return n_ok

    # --- batched SoA adds: prefix-commit, raise on the first failing element
    def add_nodes(self, numbers, xyz):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        return self._check(lib().om_add_nodes(self._h, len(numbers), _ptr(numbers), _ptr(xyz)), len(numbers))

    def add_trusses(self, numbers, conn, E, A, A2=None):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        n = len(numbers)
        conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(n, 2)
        E = np.ascontiguousarray(np.broadcast_to(E, (n,)), dtype=np.float64)
        A = np.ascontiguousarray(np.broadcast_to(A, (n,)), dtype=np.float64)
        A2 = np.full(n, np.nan) if A2 is None else np.ascontiguousarray(np.broadcast_to(A2, (n,)), dtype=np.float64)
        return self._check(lib().om_add_trusses(self._h, n, _ptr(numbers), _ptr(conn), _ptr(E), _ptr(A), _ptr(A2)), n)

    def add_beams(self, numbers, conn, props):
        """props: (n, 11) = E, nu, A, I11, I22, I12, It, shear_factor, v1x, v1y, v1z."""
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        n = len(numbers)
        conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(n, 2)
        props = np.ascontiguousarray(props, dtype=np.float64).reshape(n, 11)
        return self._check(lib().om_add_beams(self._h, n, _ptr(numbers), _ptr(conn), _ptr(props)), n)

    def add_plates(self, numbers, conn, props):
        """props: (n, 4) = E, nu, thickness, shear_factor."""
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        n = len(numbers)
        conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(n, 4)
        props = np.ascontiguousarray(props, dtype=np.float64).reshape(n, 4)
        return self._check(lib().om_add_plates(self._h, n, _ptr(numbers), _ptr(conn), _ptr(props)), n)

    # --- the crate's one-at-a-time API
    def add_node(self, number, x, y, z):
        self.add_nodes([number], [[x, y, z]])

    def add_truss(self, number, n1, n2, E, A, A2=None):
        self.add_trusses([number], [[n1, n2]], [E], [A], None if A2 is None else [A2])

    def add_beam(self, number, n1, n2, E, nu, A, i11, i22, i12, it, k, v1):
        self.add_beams([number], [[n1, n2]], [[E, nu, A, i11, i22, i12, it, k, *v1]])

    def add_plate(self, number, n1, n2, n3, n4, E, nu, t, k):
        self.add_plates([number], [[n1, n2, n3, n4]], [[E, nu, t, k]])

    def rotation(self, etype, number):
        out = np.zeros(9)
        if lib().om_get_rotation(self._h, etype, number, _ptr(out)) != 0:
            raise OracleError(self._err())
        return out

    def element_matrix(self, etype, number, global_frame=True):
        out = np.zeros(24 * 24)
        n = lib().om_element_matrix(self._h, etype, number, int(global_frame), _ptr(out))
        if n < 0:
            raise OracleError("element %d of type %d does not exist" % (number, etype))
        return out[: n * n].reshape(n, n)

    def csr(self, structural=True):
        """Global 6N x 6N stiffness as CSR (row_ptr int64, col_idx int32, val f64).

        structural=False gives the crate's value-based matrix (exact zeros dropped)."""
        nnz = lib().om_export(self._h, 1 if structural else 0)
        ndof = lib().om_ndof(self._h)
        row_ptr = np.zeros(ndof + 1, dtype=np.int64)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz, dtype=np.float64)
        lib().om_export_copy(self._h, _ptr(row_ptr), _ptr(col), _ptr(val))
        return row_ptr, col, val

    def csr_magnitude(self):
        """Term magnitude of every entry of csr(structural=True), in its order:
        the sum over the entry's element contributions of the stiffness terms
        taken with the absolute value of each factor (|T|^T k^ |T|).  Two
        evaluations of an entry that sum the same terms in different orders
        differ by at most about 2 n u times this (u = 2^-53, n roundings per
        term), however much the signed sum cancels."""
        nnz = lib().om_export(self._h, 1)
        out = np.zeros(nnz, dtype=np.float64)
        if lib().om_export_magnitude(self._h, _ptr(out), nnz) != nnz:
            raise OracleError("magnitude export size mismatch")
        return out

    def add_displacement(self, number, dof, value):
        """FEM::add_displacement (methods_for_bc_data_handle.rs:176-203); dof 0..5 = X..ThZ."""
        if lib().om_add_displacement(self._h, number, dof, float(value)):
            raise OracleError(self._err())

    def add_concentrated_load(self, number, dof, value):
        """FEM::add_concentrated_load (methods_for_bc_data_handle.rs:34-56)."""
        if lib().om_add_concentrated_load(self._h, number, dof, float(value)):
            raise OracleError(self._err())

    def add_uniformly_distributed_line_load(self, number, dof, value):
        """FEM::add_uniformly_distributed_line_load (methods_for_bc_data_handle.rs:58-100)."""
        if lib().om_add_line_load(self._h, number, dof, float(value)):
            raise OracleError(self._err())

    def add_uniformly_distributed_surface_load(self, number, dof, value):
        """FEM::add_uniformly_distributed_surface_load (methods_for_bc_data_handle.rs:102-174)."""
        if lib().om_add_surface_load(self._h, number, dof, float(value)):
            raise OracleError(self._err())

    def forces(self, n_nodes):
        out = np.zeros(6 * n_nodes)
        lib().om_forces(self._h, _ptr(out))
        return out

    def separate_sparse(self, structural=True):
        """separate_stiffness_matrix_sparse_iterative on the exported CSR:
        {"aa","ab","ba","bb": (row_ptr, col_idx, val)}, k_aa_indexes, k_bb_indexes,
        r_a, u_b."""
        self.csr(structural=structural)
        out = np.zeros(6, dtype=np.int64)
        if lib().om_separate(self._h, _ptr(out)):
            raise OracleError(self._err())
        n_aa, n_bb = int(out[0]), int(out[1])
        rows = [n_aa, n_aa, n_bb, n_bb]
        mats = {}
        for w, name in enumerate(("aa", "ab", "ba", "bb")):
            rp = np.zeros(rows[w] + 1, dtype=np.int64)
            ci = np.zeros(int(out[2 + w]), dtype=np.int32)
            v = np.zeros(int(out[2 + w]), dtype=np.float64)
            lib().om_separated_copy(self._h, w, _ptr(rp), _ptr(ci), _ptr(v))
            mats[name] = (rp, ci, v)
        aa, bb = np.zeros(n_aa, dtype=np.int64), np.zeros(n_bb, dtype=np.int64)
        ra, ub = np.zeros(n_aa), np.zeros(n_bb)
        lib().om_separated_vectors(self._h, _ptr(aa), _ptr(bb), _ptr(ra), _ptr(ub))
        return mats, aa, bb, ra, ub

    def element_forces(self, etype, u):
        """extract_element_analysis_result (truss.rs:281-333, beam.rs:803-993,
        plate.rs:1196-1409) for the elements of `etype` in insertion order,
        from global displacements u (6 per node): (n, 1 | 10 | 8)."""
        per = {0: 1, 1: 10, 2: 8}[etype]
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(per * max(1, lib().om_num_elements(self._h, int(etype))))
        w = lib().om_element_forces(self._h, int(etype), _ptr(u), _ptr(out))
        return out[:w].reshape(-1, per)

    def time_values(self, threads=1, reps=3, warmup=0):
        """Median seconds (steady clock) of `reps` host value recomputes
        (element matrices on `threads` OpenMP threads, insertion-order sums)
        after `warmup` untimed ones, and the number of values that differ
        from the sequential CSR export (0 expected)."""
        bad = np.zeros(1, dtype=np.int64)
        sec = lib().om_time_values(self._h, int(threads), int(warmup), int(reps), _ptr(bad))
        return sec, int(bad[0])

    def structural_violations(self):
        return lib().om_structural_violations(self._h)

    def nonstructural_max_rel(self):
        return lib().om_nonstructural_max_rel(self._h)

    def num_contributions(self):
        return lib().om_num_contributions(self._h)
