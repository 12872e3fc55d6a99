"""MI355X (gfx950) element-stiffness and global-assembly path of the Rust crate
`finite_element_method`, behind the crate's model API.

    from finite_element_method_amd import FEM
    m = FEM.create(1e-4, 1e-12, 2)          # FEM::<f64>::create  (src/fem/fem.rs:34)
    m.add_node(1, 0.0, 0.0, 0.0)            # add_node            (methods_for_node_data_handle.rs:66)
    m.add_node(2, 30.0, 0.0, 0.0)
    m.add_truss(1, 1, 2, 1e6, 2.0, None)    # add_truss           (methods_for_truss_data_handle.rs:49)
    row_ptr, col_idx, val = m.stiffness_csr()   # the global K, assembled on the GPU

Errors the crate returns as `Err(String)` raise ``FemError`` with the crate's
message.  Everything here is a thin ctypes layer over ``libfem_amd.so`` (C ABI
in include/fem_amd.h); there is no CPU fallback: without the library or a GPU
the assembly raises.
"""

import ctypes
import enum
import os

import numpy as np

__all__ = ["FEM", "FEMBatch", "FemError", "DeviceError", "DOFParameter", "TRUSS", "BEAM", "PLATE", "lib",
           "is_points_of_quadrilateral_on_the_same_line", "is_points_of_quadrilateral_on_the_same_plane",
           "convex_hull_on_four_points_on_plane"]

TRUSS, BEAM, PLATE = 0, 1, 2
SEPARATED = ("aa", "ab", "ba", "bb")


class DOFParameter(enum.IntEnum):
    """The crate's DOFParameter (methods_for_bc_data_handle.rs:6-14)."""
    X = 0
    Y = 1
    Z = 2
    ThX = 3
    ThY = 4
    ThZ = 5


def _dof(d):
    return int(DOFParameter[d]) if isinstance(d, str) else int(d)
PARAMS_PER = {TRUSS: 3, BEAM: 11, PLATE: 4}
FEM_OK, FEM_ERR_VALIDATION, FEM_ERR_STATE, FEM_ERR_DEVICE, FEM_ERR_ARGUMENT = 0, 1, 2, 3, 4

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfem_amd.so")
_lib = None


class FemError(Exception):
    """A validation error: the crate would return Err(String) with this message."""


class DeviceError(RuntimeError):
    """The device path failed (no GPU, HIP error, or library missing)."""


def lib():
    """Loads libfem_amd.so (built by finite_element_method_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DeviceError("libfem_amd.so is not built: run `python -m finite_element_method_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i64, u32, dbl, ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int64, ctypes.c_uint32, ctypes.c_double, ctypes.c_int
    psz = ctypes.POINTER(ctypes.c_size_t)
    sigs = {
        "fem_create": (ci, [dbl, dbl, u32, ctypes.POINTER(vp)]),
        "fem_reset": (ci, [vp, u32]),
        "fem_destroy": (None, [vp]),
        "fem_last_error": (ctypes.c_char_p, [vp]),
        "fem_add_nodes": (ci, [vp, sz, vp, vp, psz]),
        "fem_add_trusses": (ci, [vp, sz, vp, vp, vp, vp, vp, psz]),
        "fem_add_beams": (ci, [vp, sz, vp, vp, vp, psz]),
        "fem_add_plates": (ci, [vp, sz, vp, vp, vp, psz]),
        "fem_get_rotation": (ci, [vp, ci, u32, vp]),
        "fem_assemble": (ci, [vp, vp]),
        "fem_assemble_into": (ci, [vp, vp, vp]),
        "fem_build_pattern": (ci, [vp]),
        "fem_csr_info": (ci, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "fem_csr_device": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "fem_csr_to_host": (ci, [vp, vp, vp, vp]),
        "fem_last_assembly_ms": (dbl, [vp]),
        "fem_assembly_path": (ctypes.c_char_p, [vp]),
        "fem_set_option": (ci, [vp, ctypes.c_char_p, i64]),
        "fem_model_size": (ci, [vp, psz, psz, psz, psz]),
        "fem_backward": (ci, [vp, vp, vp, vp, vp, vp, vp]),
        "fem_compliance_sensitivity": (ci, [vp, vp, vp, vp, vp, vp, vp]),
        "fem_forward_backward": (ci, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "fem_last_backward_ms": (dbl, [vp]),
        "fem_update_params": (ci, [vp, ci, vp, vp]),
        "fem_update_nodes": (ci, [vp, vp, vp]),
        "fem_add_displacements": (ci, [vp, sz, vp, vp, vp, psz]),
        "fem_add_concentrated_loads": (ci, [vp, sz, vp, vp, vp, psz]),
        "fem_add_uniformly_distributed_line_loads": (ci, [vp, sz, vp, vp, vp, psz]),
        "fem_add_uniformly_distributed_surface_loads": (ci, [vp, sz, vp, vp, vp, psz]),
        "fem_separate": (ci, [vp, vp, ctypes.POINTER(i64), ctypes.POINTER(i64), vp]),
        "fem_separated_device": (ci, [vp, ci, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "fem_separated_to_host": (ci, [vp, ci, vp, vp, vp]),
        "fem_separated_vectors_device": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                              ctypes.POINTER(vp)]),
        "fem_separated_vectors_to_host": (ci, [vp, vp, vp, vp, vp]),
        "fem_last_separate_ms": (dbl, [vp]),
        "fem_solve_pcg": (ci, [vp, ci, ctypes.c_uint64, vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ci)]),
        "fem_solution_device": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "fem_solution_to_host": (ci, [vp, vp, vp]),
        "fem_last_solve_ms": (dbl, [vp]),
        "fem_compose_global_result": (ci, [vp]),
        "fem_global_result": (ci, [vp, vp, vp, vp, vp]),
        "fem_solution_displacements": (ci, [vp, vp, ctypes.POINTER(vp)]),
        "fem_element_forces": (ci, [vp, vp, vp, vp, vp, vp]),
        "fem_batch_create": (ci, [vp, sz, ctypes.POINTER(vp)]),
        "fem_batch_destroy": (None, [vp]),
        "fem_batch_last_error": (ctypes.c_char_p, [vp]),
        "fem_batch_size": (ci, [vp, psz, psz]),
        "fem_batch_autotune": (ci, [vp]),
        "fem_batch_assemble": (ci, [vp, vp]),
        "fem_batch_last_ms": (dbl, [vp]),
        "fem_batch_config": (ctypes.c_char_p, [vp]),
        "fem_batch_mesh_csr": (ci, [vp, sz, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(vp),
                                    ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i64),
                                    ctypes.POINTER(i64)]),
        "fem_batch_mesh_csr_to_host": (ci, [vp, sz, vp, vp, vp]),
        "fem_quad_points_on_same_line": (ci, [vp, ctypes.POINTER(ci)]),
        "fem_quad_points_on_same_plane": (ci, [vp, dbl, ctypes.POINTER(ci)]),
        "fem_convex_hull_on_four_points_on_plane": (ci, [vp, vp, dbl, dbl, vp, psz, ctypes.c_char_p, sz]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _stream_handle(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)  # torch.cuda.Stream


class FEM:
    """The crate's ``FEM<f64>`` (src/fem/fem.rs:15-65) with GPU assembly."""

    def __init__(self, rel_tol, abs_tol, nodes_number):
        h = ctypes.c_void_p()
        rc = lib().fem_create(float(rel_tol), float(abs_tol), int(nodes_number), ctypes.byref(h))
        if rc != FEM_OK:
            raise DeviceError("fem_create failed")
        self._h = h
        self._sep_sizes = (0, 0)
        self.rel_tol, self.abs_tol, self.nodes_number = float(rel_tol), float(abs_tol), int(nodes_number)

    @classmethod
    def create(cls, rel_tol, abs_tol, nodes_number):
        return cls(rel_tol, abs_tol, nodes_number)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.fem_destroy(h)
            self._h = None

    def _raise(self, rc):
        msg = lib().fem_last_error(self._h).decode()
        if rc == FEM_ERR_VALIDATION:
            raise FemError(msg)
        raise DeviceError(msg)

    def _call(self, rc):
        if rc != FEM_OK:
            self._raise(rc)

    def reset(self, nodes_number):
        """FEM::reset (src/fem/fem.rs:155-169)."""
        self._call(lib().fem_reset(self._h, int(nodes_number)))
        self.nodes_number = int(nodes_number)

    # ----------------------------------------------------------- batched (SoA) adds
    def _batched(self, fn, n, *arrays):
        n_ok = ctypes.c_size_t(0)
        rc = fn(self._h, n, *[_ptr(a) for a in arrays], ctypes.byref(n_ok))
        if rc != FEM_OK:
            self._raise(rc)
        return n_ok.value

    def add_nodes(self, numbers, xyz):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        if len(numbers) != len(xyz):
            raise ValueError("numbers and xyz differ in length")
        return self._batched(lib().fem_add_nodes, len(numbers), numbers, xyz)

    def add_trusses(self, numbers, conn, E, A, A2=None):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        n = len(numbers)
        conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(n, 2)
        E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, dtype=np.float64), (n,)))
        A = np.ascontiguousarray(np.broadcast_to(np.asarray(A, dtype=np.float64), (n,)))
        A2 = None if A2 is None else np.ascontiguousarray(np.broadcast_to(np.asarray(A2, dtype=np.float64), (n,)))
        return self._batched(lib().fem_add_trusses, n, numbers, conn, E, A, A2)

    def add_beams(self, numbers, conn, props):
        """props: (n, 11) = E, nu, A, I11, I22, I12, It, shear_factor, v1x, v1y, v1z."""
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        n = len(numbers)
        conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(n, 2)
        props = np.ascontiguousarray(props, dtype=np.float64).reshape(n, 11)
        return self._batched(lib().fem_add_beams, n, numbers, conn, props)

    def add_plates(self, numbers, conn, props):
        """props: (n, 4) = E, nu, thickness, shear_factor."""
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        n = len(numbers)
        conn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(n, 4)
        props = np.ascontiguousarray(props, dtype=np.float64).reshape(n, 4)
        return self._batched(lib().fem_add_plates, n, numbers, conn, props)

    # ----------------------------------------------------------- the crate's API
    def add_node(self, number, x, y, z):
        """FEM::add_node (methods_for_node_data_handle.rs:66-79)."""
        self.add_nodes([number], [[x, y, z]])

    def add_truss(self, number, node_1_number, node_2_number, young_modulus, area, optional_area_2=None):
        """FEM::add_truss (methods_for_truss_data_handle.rs:49-128)."""
        self.add_trusses([number], [[node_1_number, node_2_number]], [young_modulus], [area],
                         None if optional_area_2 is None else [optional_area_2])

    def add_beam(self, number, node_1_number, node_2_number, young_modulus, poisson_ratio, area, i11, i22, i12, it,
                 shear_factor, local_axis_1_direction):
        """FEM::add_beam (methods_for_beam_data_handle.rs:49-142)."""
        self.add_beams([number], [[node_1_number, node_2_number]],
                       [[young_modulus, poisson_ratio, area, i11, i22, i12, it, shear_factor,
                         *local_axis_1_direction]])

    def add_plate(self, number, node_1_number, node_2_number, node_3_number, node_4_number, young_modulus,
                  poisson_ratio, thickness, shear_factor):
        """FEM::add_plate (methods_for_plate_data_handle.rs:62-217)."""
        self.add_plates([number], [[node_1_number, node_2_number, node_3_number, node_4_number]],
                        [[young_modulus, poisson_ratio, thickness, shear_factor]])

    def _rotation(self, etype, number):
        out = np.zeros(9)
        self._call(lib().fem_get_rotation(self._h, etype, int(number), _ptr(out)))
        return out

    def get_truss_rotation_matrix_elements(self, number):
        """src/fem/fem.rs:171-180."""
        return self._rotation(TRUSS, number)

    def get_beam_rotation_matrix_elements(self, number):
        """src/fem/fem.rs:182-191."""
        return self._rotation(BEAM, number)

    def get_plate_rotation_matrix_elements(self, number):
        """src/fem/fem.rs:193-202."""
        return self._rotation(PLATE, number)

    # ----------------------------------------------------------- assembly
    def build_pattern(self):
        self._call(lib().fem_build_pattern(self._h))

    def assemble(self, stream=None):
        """Assembles the global stiffness on the GPU (asynchronous on `stream`)."""
        self._call(lib().fem_assemble(self._h, _stream_handle(stream)))

    def assemble_into(self, val, stream=None):
        """Assembles into a caller-owned float64 CUDA tensor of nnz values."""
        self.build_pattern()
        nnz = self.csr_info()[1]
        if val.numel() != nnz or not val.is_cuda or val.dtype.itemsize != 8 or not val.is_contiguous():
            raise ValueError("expected a contiguous float64 CUDA tensor of %d values" % nnz)
        self._call(lib().fem_assemble_into(self._h, ctypes.c_void_p(val.data_ptr()), _stream_handle(stream)))

    def csr_info(self):
        n_rows, nnz = ctypes.c_int64(), ctypes.c_int64()
        self._call(lib().fem_csr_info(self._h, ctypes.byref(n_rows), ctypes.byref(nnz)))
        return n_rows.value, nnz.value

    def stiffness_csr(self, assemble=True):
        """Global stiffness as host CSR arrays (row_ptr int64, col_idx int32, val float64)."""
        if assemble:
            self.assemble()
        n_rows, nnz = self.csr_info()
        rp = np.zeros(n_rows + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz, dtype=np.float64)
        self._call(lib().fem_csr_to_host(self._h, _ptr(rp), _ptr(ci), _ptr(v)))
        return rp, ci, v

    def device_csr(self):
        """Borrowed device pointers (row_ptr, col_idx, val) of the library's CSR."""
        rp, ci, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._call(lib().fem_csr_device(self._h, ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(v)))
        return rp.value, ci.value, v.value

    def last_assembly_ms(self):
        return lib().fem_last_assembly_ms(self._h)

    def set_option(self, key, value):
        """Tuning knobs of the device paths (see fem_set_option in include/fem_amd.h)."""
        self._call(lib().fem_set_option(self._h, key.encode(), int(value)))

    def assembly_path(self):
        return lib().fem_assembly_path(self._h).decode()

    # ----------------------------------------------------------- boundary conditions
    def add_displacements(self, numbers, dofs, values):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        dofs = np.ascontiguousarray([_dof(d) for d in np.atleast_1d(dofs)], dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        if not len(numbers) == len(dofs) == len(values):
            raise ValueError("numbers, dofs and values differ in length")
        return self._batched(lib().fem_add_displacements, len(numbers), numbers, dofs, values)

    def add_concentrated_loads(self, numbers, dofs, values):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        dofs = np.ascontiguousarray([_dof(d) for d in np.atleast_1d(dofs)], dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        if not len(numbers) == len(dofs) == len(values):
            raise ValueError("numbers, dofs and values differ in length")
        return self._batched(lib().fem_add_concentrated_loads, len(numbers), numbers, dofs, values)

    def add_displacement(self, node_number, dof_parameter, value):
        """FEM::add_displacement (methods_for_bc_data_handle.rs:176-203)."""
        self.add_displacements([node_number], [dof_parameter], [value])

    def add_concentrated_load(self, node_number, dof_parameter, value):
        """FEM::add_concentrated_load (methods_for_bc_data_handle.rs:34-56)."""
        self.add_concentrated_loads([node_number], [dof_parameter], [value])

    def _distributed(self, fn, numbers, dofs, values):
        numbers = np.ascontiguousarray(numbers, dtype=np.uint32)
        dofs = np.ascontiguousarray([_dof(d) for d in np.atleast_1d(dofs)], dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        if not len(numbers) == len(dofs) == len(values):
            raise ValueError("numbers, dofs and values differ in length")
        return self._batched(fn, len(numbers), numbers, dofs, values)

    def add_uniformly_distributed_line_loads(self, beam_numbers, dofs, values):
        return self._distributed(lib().fem_add_uniformly_distributed_line_loads, beam_numbers, dofs, values)

    def add_uniformly_distributed_surface_loads(self, plate_numbers, dofs, values):
        return self._distributed(lib().fem_add_uniformly_distributed_surface_loads, plate_numbers, dofs, values)

    def add_uniformly_distributed_line_load(self, beam_element_number, dof_parameter, value):
        """FEM::add_uniformly_distributed_line_load (methods_for_bc_data_handle.rs:58-100)."""
        self.add_uniformly_distributed_line_loads([beam_element_number], [dof_parameter], [value])

    def add_uniformly_distributed_surface_load(self, plate_element_number, dof_parameter, value):
        """FEM::add_uniformly_distributed_surface_load (methods_for_bc_data_handle.rs:102-174)."""
        self.add_uniformly_distributed_surface_loads([plate_element_number], [dof_parameter], [value])

    def loads(self):
        """The load vector (6 per node, node insertion order) as
        extract_global_analysis_result reports it."""
        return np.array([r[3] for r in self.extract_global_analysis_result()])

    def separate(self, assemble=True, stream=None):
        """Partitions the assembled stiffness on the GPU
        (separate_stiffness_matrix_sparse_iterative,
        methods_for_separate_stiffness_matrix.rs:217-320).  Returns
        (n_aa, n_bb, {"aa" | "ab" | "ba" | "bb": nnz}); synchronous."""
        if assemble:
            self.assemble(stream)
        n_aa, n_bb = ctypes.c_int64(), ctypes.c_int64()
        nnz = np.zeros(4, dtype=np.int64)
        self._call(lib().fem_separate(self._h, _stream_handle(stream), ctypes.byref(n_aa), ctypes.byref(n_bb),
                                      _ptr(nnz)))
        self._sep_sizes = (n_aa.value, n_bb.value)
        return n_aa.value, n_bb.value, dict(zip(SEPARATED, (int(x) for x in nnz)))

    def separate_sparse(self, assemble=True, stream=None):
        """Separation copied to the host: ({"aa","ab","ba","bb": (row_ptr,
        col_idx, val)}, k_aa_indexes, k_bb_indexes, r_a, u_b) -- the crate's
        SeparatedStiffnessMatrixSparse with compose_r_a_vector /
        compose_u_b_vector (:323-345)."""
        n_aa, n_bb, nnz = self.separate(assemble, stream)
        mats = {}
        for w, name in enumerate(SEPARATED):
            rows = n_aa if w < 2 else n_bb
            rp = np.zeros(rows + 1, dtype=np.int64)
            ci = np.zeros(nnz[name], dtype=np.int32)
            v = np.zeros(nnz[name], dtype=np.float64)
            self._call(lib().fem_separated_to_host(self._h, w, _ptr(rp), _ptr(ci), _ptr(v)))
            mats[name] = (rp, ci, v)
        aa, bb = np.zeros(n_aa, dtype=np.int64), np.zeros(n_bb, dtype=np.int64)
        ra, ub = np.zeros(n_aa), np.zeros(n_bb)
        self._call(lib().fem_separated_vectors_to_host(self._h, _ptr(aa), _ptr(bb), _ptr(ra), _ptr(ub)))
        return mats, aa, bb, ra, ub

    def separated_device(self, which):
        """Borrowed device pointers (row_ptr, col_idx, val) of K_aa | K_ab | K_ba | K_bb."""
        w = SEPARATED.index(which) if isinstance(which, str) else int(which)
        rp, ci, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._call(lib().fem_separated_device(self._h, w, ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(v)))
        return rp.value, ci.value, v.value

    def separated_vectors_device(self):
        """Borrowed device pointers (k_aa_indexes, k_bb_indexes, r_a, u_b)."""
        p = [ctypes.c_void_p() for _ in range(4)]
        self._call(lib().fem_separated_vectors_device(self._h, *[ctypes.byref(x) for x in p]))
        return tuple(x.value for x in p)

    def last_separate_ms(self):
        return lib().fem_last_separate_ms(self._h)

    # ----------------------------------------------------------- global analysis
    def solve_pcg(self, preconditioner="jacobi", max_iter=1000, stream=None):
        """PCG on the last separation (find_ua_vector_iterative_pcg_{jacobi,
        block_jacobi}_sparse + find_r_r_vector_sparse,
        methods_for_global_analysis.rs:186-333).  Returns (u_a, r_r,
        iterations, converged) with host copies of u_a and r_r."""
        pre = {"jacobi": 0, "block_jacobi": 1}.get(preconditioner, preconditioner)
        it, conv = ctypes.c_uint64(), ctypes.c_int()
        self._call(lib().fem_solve_pcg(self._h, int(pre), int(max_iter), _stream_handle(stream), ctypes.byref(it),
                                       ctypes.byref(conv)))
        n_aa, n_bb = self._sep_sizes
        ua, rr = np.zeros(n_aa), np.zeros(n_bb)
        self._call(lib().fem_solution_to_host(self._h, _ptr(ua), _ptr(rr)))
        return ua, rr, int(it.value), bool(conv.value)

    def last_solve_ms(self):
        return lib().fem_last_solve_ms(self._h)

    def compose_global_analysis_result(self):
        """methods_for_global_analysis.rs:335-358 for the last solve."""
        self._call(lib().fem_compose_global_result(self._h))

    def extract_global_analysis_result(self):
        """methods_for_global_analysis.rs:360-384: [(node, DOFParameter, displacement, load)]
        in node insertion order."""
        n = 6 * self.model_size()[0]
        nums, dofs = np.zeros(n, dtype=np.uint32), np.zeros(n, dtype=np.int32)
        disp, load = np.zeros(n), np.zeros(n)
        self._call(lib().fem_global_result(self._h, _ptr(nums), _ptr(dofs), _ptr(disp), _ptr(load)))
        return [(int(a), DOFParameter(int(b)), float(c), float(d)) for a, b, c, d in zip(nums, dofs, disp, load)]

    FORCES_PER = {TRUSS: 1, BEAM: 10, PLATE: 8}
    FORCE_NAMES = {
        TRUSS: ("ForceR",),
        BEAM: ("ForceR", "ForceS", "ForceT", "MomentR", "MomentS", "MomentS", "MomentS", "MomentT", "MomentT",
               "MomentT"),
        PLATE: ("MembraneForceR", "MembraneForceS", "MembraneForceRS", "BendingMomentR", "BendingMomentS",
                "BendingMomentRS", "ShearForceRT", "ShearForceST"),
    }

    def element_forces(self, u=None, stream=None):
        """Element forces (extract_elements_analysis_result) from global
        displacements u (a float64 CUDA tensor of 6 per node; default: the
        last solve's).  Returns {TRUSS: (n, 1), BEAM: (n, 10), PLATE: (n, 8)}
        float64 CUDA tensors, elements in insertion order."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        if u is None:
            p = ctypes.c_void_p()
            self._call(lib().fem_solution_displacements(self._h, _stream_handle(stream), ctypes.byref(p)))
            uptr = p
        else:
            if u.numel() != 6 * self.nodes_number or u.dtype != torch.float64 or not u.is_cuda or \
                    not u.is_contiguous():
                raise ValueError("expected a contiguous float64 CUDA tensor of %d values" % (6 * self.nodes_number))
            uptr = ctypes.c_void_p(u.data_ptr())
        counts = self.model_size()[1:]
        outs = {t: torch.empty((n, self.FORCES_PER[t]), dtype=torch.float64, device="cuda")
                for t, n in zip((TRUSS, BEAM, PLATE), counts)}
        ptrs = [ctypes.c_void_p(outs[t].data_ptr()) if outs[t].numel() else None for t in (TRUSS, BEAM, PLATE)]
        self._call(lib().fem_element_forces(self._h, uptr, *ptrs, _stream_handle(stream)))
        return outs

    # ----------------------------------------------------------- differentiation
    def model_size(self):
        """(n_nodes, n_truss, n_beam, n_plate) committed so far."""
        v = [ctypes.c_size_t() for _ in range(4)]
        self._call(lib().fem_model_size(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def _adjoint(self, fn, g, params, xyz, stream, expect):
        import torch

        n_nodes, *counts = self.model_size()
        if g.numel() != expect:
            raise ValueError("expected %d values, got %d" % (expect, g.numel()))
        dev = g.device
        grads = {}
        ptrs = []
        for t, n in zip((TRUSS, BEAM, PLATE), counts):
            if params and n:
                grads[t] = torch.empty((n, PARAMS_PER[t]), dtype=torch.float64, device=dev)
                ptrs.append(ctypes.c_void_p(grads[t].data_ptr()))
            else:
                ptrs.append(None)
        dxyz = torch.empty((n_nodes, 3), dtype=torch.float64, device=dev) if xyz else None
        if g.dtype != torch.float64 or not g.is_cuda or not g.is_contiguous():
            raise ValueError("expected a contiguous float64 CUDA tensor")
        if stream is None:
            stream = torch.cuda.current_stream(dev)
        self._call(fn(self._h, ctypes.c_void_p(g.data_ptr()), *ptrs,
                      None if dxyz is None else ctypes.c_void_p(dxyz.data_ptr()), _stream_handle(stream)))
        return grads, dxyz

    def backward(self, dl_dk, params=True, xyz=True, stream=None):
        """Gradients of a scalar L(K) from dL/dK (float64 CUDA tensor of nnz
        values in CSR order): ({TRUSS: (n, 3), BEAM: (n, 11), PLATE: (n, 4)}
        parameter gradients in the add_* layouts, (N, 3) node-coordinate
        gradient or None).  See fem_backward in include/fem_amd.h."""
        self.build_pattern()
        return self._adjoint(lib().fem_backward, dl_dk, params, xyz, stream, self.csr_info()[1])

    def forward_backward(self, val, dl_dk, params=True, xyz=True, stream=None):
        """Assembles into `val` (as assemble_into) and returns the gradients of
        L(K) from dl_dk (as backward) in one call; the two passes run at the
        same time on two streams.  See fem_forward_backward."""
        self.build_pattern()
        nnz = self.csr_info()[1]
        if val.numel() != nnz or not val.is_cuda or val.dtype.itemsize != 8 or not val.is_contiguous():
            raise ValueError("expected a contiguous float64 CUDA tensor of %d values" % nnz)
        fn = lambda h, g, *rest: lib().fem_forward_backward(h, ctypes.c_void_p(val.data_ptr()), g, *rest)
        return self._adjoint(fn, dl_dk, params, xyz, stream, nnz)

    def compliance_sensitivity(self, u, params=True, xyz=True, stream=None):
        """-u_e^T (dK_e/dp) u_e summed per parameter, for nodal displacements
        u (float64 CUDA tensor of 6N values, dof order of K)."""
        return self._adjoint(lib().fem_compliance_sensitivity, u, params, xyz, stream, 6 * self.model_size()[0])

    def last_backward_ms(self):
        return lib().fem_last_backward_ms(self._h)

    def update_params(self, etype, params, stream=None):
        """Replaces one element type's parameters (float64 CUDA tensor, add_* layout)."""
        params = params.detach().contiguous()
        n = self.model_size()[1 + int(etype)]
        if params.numel() != n * PARAMS_PER[etype] or not params.is_cuda or params.dtype.itemsize != 8:
            raise ValueError("expected a float64 CUDA tensor of %d x %d" % (n, PARAMS_PER[etype]))
        self._call(lib().fem_update_params(self._h, int(etype), ctypes.c_void_p(params.data_ptr()),
                                           _stream_handle(stream)))
        self._version = getattr(self, "_version", 0) + 1

    def update_nodes(self, xyz, stream=None):
        """Replaces all node coordinates (float64 CUDA tensor, N x 3)."""
        xyz = xyz.detach().contiguous()
        n = self.model_size()[0]
        if xyz.numel() != 3 * n or not xyz.is_cuda or xyz.dtype.itemsize != 8:
            raise ValueError("expected a float64 CUDA tensor of %d x 3" % n)
        self._call(lib().fem_update_nodes(self._h, ctypes.c_void_p(xyz.data_ptr()), _stream_handle(stream)))
        self._version = getattr(self, "_version", 0) + 1

    # ----------------------------------------------------------- convenience
    def load_mesh(self, mesh):
        """Adds a finite_element_method_amd.meshgen.Mesh (nodes then elements)."""
        self.add_nodes(mesh.node_numbers, mesh.xyz)
        if mesh.kind == "truss":
            self.add_trusses(mesh.numbers, mesh.conn, mesh.props[:, 0], mesh.props[:, 1], mesh.props[:, 2])
        elif mesh.kind == "beam":
            self.add_beams(mesh.numbers, mesh.conn, mesh.props)
        else:
            self.add_plates(mesh.numbers, mesh.conn, mesh.props)
        return self

    @classmethod
    def from_mesh(cls, mesh):
        return cls(mesh.rel_tol, mesh.abs_tol, mesh.n_nodes).load_mesh(mesh)


class FEMBatch:
    """Independent meshes assembled by one hipGraph launch (fem_batch_* in
    include/fem_amd.h).  The models' nodes and elements are copied at
    construction; each mesh keeps its own CSR."""

    def __init__(self, models):
        arr = (ctypes.c_void_p * len(models))(*[m._h.value for m in models])
        h = ctypes.c_void_p()
        rc = lib().fem_batch_create(arr, len(models), ctypes.byref(h))
        if rc != FEM_OK:
            raise DeviceError("fem_batch_create failed (%d)" % rc)
        self._h = h
        self.n_meshes = len(models)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.fem_batch_destroy(h)
            self._h = None

    def _call(self, rc):
        if rc != FEM_OK:
            raise DeviceError(lib().fem_batch_last_error(self._h).decode())

    def autotune(self):
        """Times candidate launch shapes and the groups' schedule (concurrent
        or serial), keeps the fastest, captures the graph.  Tuned results are
        cached per (gfx arch, batch signature) in the process and in the tune
        cache file (FEM_AMD_TUNE_CACHE, default ~/.cache/fem_amd/tune.txt)."""
        self._call(lib().fem_batch_autotune(self._h))
        return self.config()

    def assemble(self, stream=None):
        """One graph launch assembling every mesh (asynchronous on `stream`)."""
        self._call(lib().fem_batch_assemble(self._h, _stream_handle(stream)))

    def last_ms(self):
        return lib().fem_batch_last_ms(self._h)

    def config(self):
        return lib().fem_batch_config(self._h).decode()

    def groups(self):
        n, g = ctypes.c_size_t(), ctypes.c_size_t()
        self._call(lib().fem_batch_size(self._h, ctypes.byref(n), ctypes.byref(g)))
        return g.value

    def mesh_info(self, i):
        """(n_rows, nnz, val_device_ptr) of mesh i."""
        nr, nz, vb, cb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        rp, ci, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._call(lib().fem_batch_mesh_csr(self._h, int(i), ctypes.byref(nr), ctypes.byref(nz), ctypes.byref(rp),
                                            ctypes.byref(ci), ctypes.byref(v), ctypes.byref(vb), ctypes.byref(cb)))
        return nr.value, nz.value, v.value

    def mesh_csr(self, i):
        """Mesh i's CSR on the host (row_ptr int64, col_idx int32, val float64)."""
        n_rows, nnz, _ = self.mesh_info(i)
        rp = np.zeros(n_rows + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz, dtype=np.float64)
        self._call(lib().fem_batch_mesh_csr_to_host(self._h, int(i), _ptr(rp), _ptr(ci), _ptr(v)))
        return rp, ci, v


# ------------------------------------------------------------------ quadrilateral predicates
def _quad_points(points):
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(4, 3)
    return pts


def is_points_of_quadrilateral_on_the_same_line(point_1, point_2, point_3, point_4):
    """quadrilateral_4n_element_functions.rs:14-86 (re-exported, src/lib.rs:19-23)."""
    pts = _quad_points([point_1, point_2, point_3, point_4])
    out = ctypes.c_int()
    if lib().fem_quad_points_on_same_line(_ptr(pts), ctypes.byref(out)) != FEM_OK:
        raise ValueError("fem_quad_points_on_same_line failed")
    return bool(out.value)


def is_points_of_quadrilateral_on_the_same_plane(point_1, point_2, point_3, point_4, abs_tol):
    """quadrilateral_4n_element_functions.rs:88-129 (re-exported, src/lib.rs:19-23)."""
    pts = _quad_points([point_1, point_2, point_3, point_4])
    out = ctypes.c_int()
    if lib().fem_quad_points_on_same_plane(_ptr(pts), float(abs_tol), ctypes.byref(out)) != FEM_OK:
        raise ValueError("fem_quad_points_on_same_plane failed")
    return bool(out.value)


def convex_hull_on_four_points_on_plane(point_numbers, points, rel_tol, abs_tol):
    """quadrilateral_4n_element_functions.rs:173-250 (re-exported, src/lib.rs:19-23):
    the point numbers of the convex hull of four points, in the crate's order.
    Raises FemError with the crate's message when points 2-4 have no frame."""
    nums = np.ascontiguousarray(point_numbers, dtype=np.uint32).reshape(4)
    pts = _quad_points(points)
    hull = np.zeros(4, dtype=np.uint32)
    n = ctypes.c_size_t()
    err = ctypes.create_string_buffer(256)
    rc = lib().fem_convex_hull_on_four_points_on_plane(_ptr(nums), _ptr(pts), float(rel_tol), float(abs_tol),
                                                        _ptr(hull), ctypes.byref(n), err, len(err))
    if rc == FEM_ERR_VALIDATION:
        raise FemError(err.value.decode())
    if rc != FEM_OK:
        raise ValueError("fem_convex_hull_on_four_points_on_plane failed (%d)" % rc)
    return [int(x) for x in hull[: n.value]]
