/* fem_amd.h -- C ABI of the MI355X (gfx950) element-stiffness and global-assembly
 * library, a drop-in for the assembly path of the Rust crate
 * `finite_element_method` (0.9.12).
 *
 * Conventions
 *   - Opaque handle, `int` status: 0 = FEM_OK.  On failure the handle owns a
 *     message readable with fem_last_error(); validation messages are the
 *     crate's own strings.
 *   - Batched adds take SoA arrays and commit a prefix: elements before the
 *     first failing one are added, `*n_ok` receives how many.  This is the
 *     crate's per-call Result semantics applied element by element.
 *   - Assembly is deferred: the crate assembles eagerly inside add_*, but its
 *     global matrix is crate-private (src/fem/fem.rs:70-72) and is only seen
 *     through separate_stiffness_matrix_*, so deferring is unobservable.
 *   - The global matrix is 6N x 6N (N = nodes_number, dof = 6 * node_index + k,
 *     k in X, Y, Z, ThX, ThY, ThZ), stored as CSR: int64 row_ptr, int32
 *     col_idx, double values, rows and columns ascending.
 *   - Streams are passed as `void*` (a hipStream_t; NULL = default stream).
 *   - One handle per thread; a handle is not thread-safe.
 */
#ifndef FEM_AMD_H
#define FEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_model fem_model;

enum fem_status {
  FEM_OK = 0,
  FEM_ERR_VALIDATION = 1, /* the crate would return Err(String) */
  FEM_ERR_STATE = 2,      /* call not valid in the handle's current state */
  FEM_ERR_DEVICE = 3,     /* a HIP call failed or no device is available */
  FEM_ERR_ARGUMENT = 4    /* null pointer, size mismatch, limits */
};

enum fem_element_type { FEM_TRUSS = 0, FEM_BEAM = 1, FEM_PLATE = 2 };

/* ---- model life cycle ------------------------------------------------- */

/* FEM::<V>::create(rel_tol, abs_tol, nodes_number)      src/fem/fem.rs:34-65 */
int fem_create(double rel_tol, double abs_tol, uint32_t nodes_number, fem_model** out);
/* FEM::reset(nodes_number)                               src/fem/fem.rs:155-169 */
int fem_reset(fem_model* m, uint32_t nodes_number);
void fem_destroy(fem_model* m);
/* message of the last failing call (owned by the handle, valid until the next call) */
const char* fem_last_error(const fem_model* m);

/* ---- nodes and elements ----------------------------------------------- */

/* FEM::add_node(number, x, y, z), n nodes            methods_for_node_data_handle.rs:66-79
 * xyz: n x 3 row-major */
int fem_add_nodes(fem_model* m, size_t n, const uint32_t* numbers, const double* xyz, size_t* n_ok);

/* FEM::add_truss(number, n1, n2, E, A, Option<A2>)     methods_for_truss_data_handle.rs:49-128
 * conn: n x 2 node numbers; A2[i] = NaN means None; A2 may be NULL (all None) */
int fem_add_trusses(fem_model* m, size_t n, const uint32_t* numbers, const uint32_t* conn, const double* E,
                    const double* A, const double* A2, size_t* n_ok);

/* FEM::add_beam(number, n1, n2, E, nu, A, I11, I22, I12, It, shear_factor,
 *               local_axis_1_direction)                 methods_for_beam_data_handle.rs:49-142
 * props: n x 11 row-major = E, nu, A, I11, I22, I12, It, shear_factor, v1x, v1y, v1z */
int fem_add_beams(fem_model* m, size_t n, const uint32_t* numbers, const uint32_t* conn, const double* props,
                  size_t* n_ok);

/* FEM::add_plate(number, n1, n2, n3, n4, E, nu, thickness, shear_factor)
 *                                                        methods_for_plate_data_handle.rs:62-217
 * conn: n x 4 node numbers; props: n x 4 = E, nu, thickness, shear_factor */
int fem_add_plates(fem_model* m, size_t n, const uint32_t* numbers, const uint32_t* conn, const double* props,
                   size_t* n_ok);

/* Counts of committed nodes and elements (the crate keeps them in its HashMaps,
 * fem.rs:15-32; any pointer may be NULL). */
int fem_model_size(const fem_model* m, size_t* n_nodes, size_t* n_truss, size_t* n_beam, size_t* n_plate);

/* FEM::get_{truss,beam,plate}_rotation_matrix_elements(number) -> [V; 9]
 *                                                        src/fem/fem.rs:171-202
 * Computed on the device; row-major. */
int fem_get_rotation(fem_model* m, int type, uint32_t number, double out[9]);

/* ---- quadrilateral predicates (re-exported by the crate, src/lib.rs:19-23) --
 *
 * points: 4 x 3 row-major (points 1..4 of the quadrilateral).  No handle. */

/* is_points_of_quadrilateral_on_the_same_line(p1, p2, p3, p4)
 *                                   quadrilateral_4n_element_functions.rs:14-86
 * *out = 1 when the cross product of some corner's two edges has zero norm. */
int fem_quad_points_on_same_line(const double* points, int* out);
/* is_points_of_quadrilateral_on_the_same_plane(p1, p2, p3, p4, abs_tol)
 *                                   quadrilateral_4n_element_functions.rs:88-129
 * *out = 1 when point 1 lies in the plane of points 2, 3, 4 (abs_tol snapping). */
int fem_quad_points_on_same_plane(const double* points, double abs_tol, int* out);
/* convex_hull_on_four_points_on_plane(point_numbers, points, rel_tol, abs_tol)
 *                                   quadrilateral_4n_element_functions.rs:173-250,
 *                                   convex_hull_on_plane.rs:14-214
 * Ok: hull[0 .. *n_hull) = the hull's point numbers in the crate's order
 * (lowest point first, then by polar angle; *n_hull <= 4).  Err (no frame for
 * points 2-4): FEM_ERR_VALIDATION, the message copied into err (err_cap bytes,
 * NUL-terminated; err may be NULL). */
int fem_convex_hull_on_four_points_on_plane(const uint32_t* numbers, const double* points, double rel_tol,
                                            double abs_tol, uint32_t* hull, size_t* n_hull, char* err,
                                            size_t err_cap);

/* ---- assembly ---------------------------------------------------------- */

/* Computes the global stiffness on the device (replaces the eager scatter of
 * add_truss/add_beam/add_plate, methods_for_*_data_handle.rs).  The first call
 * after a model change builds the sparsity pattern and the element-to-slot
 * maps; later calls only recompute values.  Asynchronous on `stream`. */
int fem_assemble(fem_model* m, void* stream);

/* Same, writing the values into caller-owned device memory of nnz doubles
 * (the pattern must exist: call fem_assemble or fem_build_pattern first). */
int fem_assemble_into(fem_model* m, double* val_device, void* stream);

/* Builds the pattern only (synchronous). */
int fem_build_pattern(fem_model* m);

/* Pattern size: n_rows = 6 * nodes_number. */
int fem_csr_info(fem_model* m, int64_t* n_rows, int64_t* nnz);

/* Borrowed device pointers to the library's CSR (valid until the next model
 * change, reset or destroy). */
int fem_csr_device(fem_model* m, const int64_t** row_ptr, const int32_t** col_idx, const double** val);

/* Synchronous copy of the CSR to host buffers (any pointer may be NULL). */
int fem_csr_to_host(fem_model* m, int64_t* row_ptr, int32_t* col_idx, double* val);

/* Timing of the last fem_assemble / fem_assemble_into on its stream, from
 * HIP events around the value kernels only (ms); -1 when unavailable. */
double fem_last_assembly_ms(fem_model* m);

/* ---- differentiation (no counterpart in the crate) ---------------------
 *
 * Both entries recompute each element's stiffness on the device (nothing is
 * stored by the forward) and are deterministic (no atomics).  All pointers are
 * device memory; every output pointer may be NULL.  Asynchronous on `stream`.
 *   dl_dparam_truss: n_truss x 3 = dL/d(E, A, A2)   (A2 slot 0 when A2 is None)
 *   dl_dparam_beam:  n_beam x 11 = dL/d(E, nu, A, I11, I22, I12, It, k, v1x, v1y, v1z);
 *                    I11/I22 are given when I12 == 0 and I11 != I22, NaN otherwise;
 *                    I12 and v1 (they rotate the section) are NaN
 *   dl_dparam_plate: n_plate x 4 = dL/d(E, nu, thickness, shear_factor)
 *   dl_dxyz:         N x 3 node coordinates, node insertion order; FEM_ERR_ARGUMENT
 *                    for models with beams (the crate's beam roll angle, beam.rs:185-202,
 *                    is acos of a sign-blind cosine: not differentiable at roll 0 or pi)
 * Element stiffness definitions: truss.rs:95-189, beam.rs:260-605, plate.rs:160-724,
 * frames truss.rs:66-93, beam.rs:162-258, quadrilateral_4n_element_functions.rs:131-171. */

/* Backward of the assembly: dl_dk is dL/dK on the CSR pattern (nnz doubles). */
int fem_backward(fem_model* m, const double* dl_dk, double* dl_dparam_truss, double* dl_dparam_beam,
                 double* dl_dparam_plate, double* dl_dxyz, void* stream);

/* Compliance sensitivities: with nodal displacements u (6N doubles, dof order of
 * the global matrix) every output receives sum_e -u_e^T (dK_e/dp) u_e. */
int fem_compliance_sensitivity(fem_model* m, const double* u, double* s_truss, double* s_beam, double* s_plate,
                               double* s_xyz, void* stream);
/* Forward and backward in one call (no counterpart in the crate): the
 * values into val_device as fem_assemble_into, and the gradients of
 * L(K) from dl_dk as fem_backward.  The backward does not depend on the
 * values, so it runs on a second stream of the library, forked from and
 * joined back to `stream`; the results are the same as the two calls'. */
int fem_forward_backward(fem_model* m, double* val_device, const double* dl_dk, double* dl_dparam_truss,
                         double* dl_dparam_beam, double* dl_dparam_plate, double* dl_dxyz, void* stream);

/* Timing of the last backward / sensitivity call (HIP events, ms); -1 when unavailable. */
double fem_last_backward_ms(fem_model* m);

/* Replace one element type's parameters (device memory, same layout as the
 * fem_add_* props) or all node coordinates (device memory, N x 3).  Values are
 * not re-validated.  Frames are recomputed; if a frame's zero pattern changes
 * the sparsity pattern is rebuilt by the next assembly. */
int fem_update_params(fem_model* m, int type, const double* params_device, void* stream);
int fem_update_nodes(fem_model* m, const double* xyz_device, void* stream);

/* ---- boundary conditions and sparse separation ------------------------
 *
 * dof 0..5 = X, Y, Z, ThX, ThY, ThZ (DOFParameter, methods_for_bc_data_handle.rs:6-14);
 * global dof index = 6 * node_index + dof. */

/* FEM::add_displacement(node, dof, value), n entries  methods_for_bc_data_handle.rs:176-203
 * A second displacement on the same dof fails with the crate's message. */
int fem_add_displacements(fem_model* m, size_t n, const uint32_t* numbers, const int32_t* dofs, const double* values,
                          size_t* n_ok);
/* FEM::add_concentrated_load(node, dof, value), n entries  methods_for_bc_data_handle.rs:34-56
 * Loads on the same dof add up. */
int fem_add_concentrated_loads(fem_model* m, size_t n, const uint32_t* numbers, const int32_t* dofs,
                               const double* values, size_t* n_ok);

/* FEM::add_uniformly_distributed_line_load(beam, dof, value), n entries
 *                                  methods_for_bc_data_handle.rs:58-100, beam.rs:775-797
 * FEM::add_uniformly_distributed_surface_load(plate, dof, value), n entries
 *                                  methods_for_bc_data_handle.rs:102-174, plate.rs:1145-1185
 * The element's nodal loads (one point for beams, the four plate points) are
 * computed on the device and added to the load vector in call order; the
 * loads act along the global dof given. */
int fem_add_uniformly_distributed_line_loads(fem_model* m, size_t n, const uint32_t* beam_numbers,
                                             const int32_t* dofs, const double* values, size_t* n_ok);
int fem_add_uniformly_distributed_surface_loads(fem_model* m, size_t n, const uint32_t* plate_numbers,
                                                const int32_t* dofs, const double* values, size_t* n_ok);

/* FEM::separate_stiffness_matrix_sparse_iterative()
 *                              methods_for_separate_stiffness_matrix.rs:217-320
 * plus compose_r_a_vector / compose_u_b_vector (:323-345), on the device, from
 * the values of the last fem_assemble.  Dofs with a zero (or absent) diagonal
 * are excluded; constrained dofs form b, the others a; entries equal to zero
 * are dropped.  Outputs: n_aa, n_bb and the non-zero counts of K_aa, K_ab,
 * K_ba, K_bb (any pointer may be NULL).  Fails with the crate's messages
 * (constrained dof without stiffness, "No restraints", empty K_aa).
 * Synchronizes `stream`. */
int fem_separate(fem_model* m, void* stream, int64_t* n_aa, int64_t* n_bb, int64_t nnz[4]);
/* The four matrices as CSR with local indices, rows and columns ascending
 * (which: 0 = K_aa, 1 = K_ab, 2 = K_ba, 3 = K_bb).  _device lends the
 * library's buffers until the next fem_separate; _to_host copies. */
int fem_separated_device(fem_model* m, int which, const int64_t** row_ptr, const int32_t** col_idx,
                         const double** values);
int fem_separated_to_host(fem_model* m, int which, int64_t* row_ptr, int32_t* col_idx, double* values);
/* k_aa_indexes, k_bb_indexes (global dof indexes, ascending), r_a (forces at
 * the a dofs) and u_b (imposed displacements at the b dofs). */
int fem_separated_vectors_device(fem_model* m, const int64_t** k_aa_indexes, const int64_t** k_bb_indexes,
                                 const double** r_a, const double** u_b);
int fem_separated_vectors_to_host(fem_model* m, int64_t* k_aa_indexes, int64_t* k_bb_indexes, double* r_a,
                                  double* u_b);
/* HIP-event time of the last fem_separate's kernels (ms); -1 when unavailable. */
double fem_last_separate_ms(fem_model* m);

/* ---- sparse global analysis ------------------------------------------- */

/* FEM::find_ua_vector_iterative_pcg_jacobi_sparse (preconditioner 0) and
 * find_ua_vector_iterative_pcg_block_jacobi_sparse (preconditioner 1), then
 * find_r_r_vector_sparse        methods_for_global_analysis.rs:186-333
 * on the last fem_separate: b = r_a - K_ab u_b, K_aa u_a = b by
 * preconditioned CG from u_a = 0 (Jacobi, or the inverse of each node's
 * block of K_aa, blocks as build_block_starts_from_k_aa_indexes :135-160),
 * stopped when ||r|| <= max(rel_tol ||b||, abs_tol) or after max_iter
 * updates; r_r = (K_ba u_a + K_bb u_b) - r_b.  Outputs: the number of
 * updates and whether the stopping test was met.  Synchronizes `stream`. */
int fem_solve_pcg(fem_model* m, int preconditioner, uint64_t max_iter, void* stream, uint64_t* iterations,
                  int* converged);
/* No counterpart in the crate: fem_solve_pcg for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same solve of the last fem_separate, bit for bit,
 * with every work buffer (the PCG vectors, the preconditioner, the reduction
 * partials and scalars) carved from one static device array of 64 MiB per GPU,
 * held until the process ends; u_a (n_aa doubles) and r_r (n_bb doubles) are
 * caller device memory.  Calls no hipMalloc/hipFree.  FEM_ERR_ARGUMENT when the
 * work buffers exceed the array: about 56 bytes per a-dof with Jacobi, 108
 * with node-block Jacobi (so up to ~1.1M / ~0.6M a-dofs).  Calls that overlap
 * on one GPU race on the array.  The handle's stored solution (fem_solution_*)
 * is not changed.  Synchronizes `stream`. */
int fem_solve_pcg_noalloc(fem_model* m, int preconditioner, uint64_t max_iter, double* u_a, double* r_r,
                          void* stream, uint64_t* iterations, int* converged);
/* u_a (n_aa) and r_r (n_bb) of the last solve: borrowed device pointers
 * (valid until the next solve) or a host copy. */
int fem_solution_device(fem_model* m, const double** u_a, const double** r_r);
int fem_solution_to_host(fem_model* m, double* u_a, double* r_r);
/* HIP-event time of the last fem_solve_pcg (ms); -1 when unavailable. */
double fem_last_solve_ms(fem_model* m);
/* FEM::compose_global_analysis_result         methods_for_global_analysis.rs:335-358
 * displacements[k_aa_indexes] = u_a, forces[k_bb_indexes] = r_r. */
int fem_compose_global_result(fem_model* m);
/* Global displacements of the last solve (6 per node: imposed values, u_a at
 * the k_aa indexes); a borrowed device pointer valid until the next call. */
int fem_solution_displacements(fem_model* m, void* stream, const double** u);
/* FEM::extract_elements_analysis_result      methods_for_element_analysis.rs:27-58
 * per element of each type, in insertion order, from device displacements
 * u_device (6 per node):
 *   truss (1):  ForceR                                     truss.rs:281-333
 *   beam (10):  ForceR, ForceS, ForceT, MomentR, MomentS at node 1 / average /
 *               node 2, MomentT at node 1 / average / node 2  beam.rs:803-993
 *   plate (8):  MembraneForceR, S, RS, BendingMomentR, S, RS, ShearForceRT, ST
 *                                                          plate.rs:1196-1409
 * Outputs are device arrays (or NULL to skip a type).  Asynchronous on `stream`. */
int fem_element_forces(fem_model* m, const double* u_device, double* truss_out, double* beam_out, double* plate_out,
                       void* stream);
/* FEM::extract_global_analysis_result         methods_for_global_analysis.rs:360-384
 * 6 entries per node in insertion order: (node number, dof, displacement,
 * load); arrays of 6 x n_nodes, any may be NULL. */
int fem_global_result(fem_model* m, uint32_t* node_numbers, int32_t* dofs, double* displacements, double* loads);

/* ---- batched assembly (no counterpart in the crate) ---------------------
 *
 * Many independent meshes assembled by one hipGraph launch.  The meshes'
 * nodes and elements are copied (at create) into per-group disjoint-union
 * models: flat plate-only meshes (fused plate kernel), truss and beam meshes
 * (node-owner kernels) and all others (generic path); the input handles are
 * not modified and may be destroyed afterwards.  Each mesh keeps its own CSR:
 * its rows and values are contiguous slices of its group's CSR, and every
 * value is summed in the mesh's own element order by the same value path as
 * when the mesh is assembled alone, so the results are identical.  All
 * meshes must share rel_tol and abs_tol. */
typedef struct fem_batch fem_batch;
int fem_batch_create(fem_model* const* models, size_t n, fem_batch** out);
void fem_batch_destroy(fem_batch* b);
const char* fem_batch_last_error(const fem_batch* b);
int fem_batch_size(const fem_batch* b, size_t* n_meshes, size_t* n_groups);
/* Times candidate launch shapes per group (row-staging plate kernel: owned
 * nodes per patch and resident workgroups per CU; block-lane plate kernels:
 * workgroups per CU; generic path: block size) and the schedule of the groups
 * (concurrent graph branches, or one after another on one stream) on whole
 * graph replays, keeps the fastest and captures the graph.  Results are cached
 * per (gfx architecture, batch signature) for the life of the process and in a
 * cache file: FEM_AMD_TUNE_CACHE, else $XDG_CACHE_HOME/fem_amd/tune.txt, else
 * $HOME/.cache/fem_amd/tune.txt (one line per entry; skipped when not
 * writable).  A cached entry is applied without timing.  Synchronous. */
int fem_batch_autotune(fem_batch* b);
/* One hipGraphLaunch of every group's value kernels on `stream`. */
int fem_batch_assemble(fem_batch* b, void* stream);
/* HIP-event time of the last fem_batch_assemble (ms); -1 when unavailable. */
double fem_batch_last_ms(fem_batch* b);
/* Description of the chosen launch shapes. */
const char* fem_batch_config(const fem_batch* b);
/* Mesh `mesh`'s CSR on the device, borrowed: row_ptr (n_rows + 1 entries) and
 * col_idx (nnz) hold group positions and group columns; subtract *val_base from
 * row_ptr and *col_base from col_idx for mesh-local ones.  val points at the
 * mesh's nnz values. */
int fem_batch_mesh_csr(fem_batch* b, size_t mesh, int64_t* n_rows, int64_t* nnz, const int64_t** row_ptr,
                       const int32_t** col_idx, const double** val, int64_t* val_base, int64_t* col_base);
/* Synchronous copy of mesh `mesh`'s CSR with mesh-local positions and columns. */
int fem_batch_mesh_csr_to_host(fem_batch* b, size_t mesh, int64_t* row_ptr, int32_t* col_idx, double* val);

/* Tuning options (no counterpart in the crate; they change speed, never values):
 *   "plate_kernel"       fused kernel of flat plate meshes: 3 = balanced lanes
 *                        (plate_lanes, default; meshes whose nodes have at
 *                        most four elements, else plate_blocks), 1 = one lane
 *                        per owned node-pair block (plate_blocks), 0 = row
 *                        staging (plate_fused), 2 = one lane per element
 *                        corner (plate_corners); bitwise equal values
 *   "plate_patch_order"  patches of the balanced-lane kernel: 0 = automatic
 *                        (node-index order unless it scatters the patches'
 *                        elements, then Morton order), 1 = node-index order,
 *                        2 = Morton order; bitwise equal values
 *   "plate_patch_nodes"  owned nodes per workgroup of the row-staging kernel: 4..64
 *   "plate_wg_per_cu"    resident workgroups per CU of the fused plate kernels,
 *                        looping over the patches: 0..64 (0 = one workgroup
 *                        per patch; default 0 for plate_blocks, 2 for plate_fused)
 *   "force_generic"      1 = always use the generic value path
 *   "separate_lanes"     0 (default) = fem_separate's four-entry kernels where
 *                        the values span under 4 GiB (count 4 and fill 8 lanes
 *                        per CSR row), else one entry per lane with 4 or 8
 *                        lanes per row from the mean row length; 4, 8, 16, 32 =
 *                        the one-entry kernels with that many lanes.  The
 *                        separated matrices are bitwise equal either way
 *   "wide_loads"         1 (default) = where the values span under 4 GiB,
 *                        the backward reads each row run of a dL/dK block, and
 *                        the PCG's SpMV four values per lane, with 16-byte
 *                        range-checked loads; 0 = one 8-byte load per value
 *                        (the paths taken above 4 GiB).  Gradients are bitwise
 *                        equal either way; PCG iterates differ by rounding (the
 *                        SpMV sums a row in another order)
 *   "time_assembly"      1 (default) = events around every assembly launch
 *                        for fem_last_assembly_ms; 0 = none (two fewer
 *                        packets between back-to-back launches)
 * Setting a plate or path option invalidates the pattern (rebuilt by the next
 * assembly). */
int fem_set_option(fem_model* m, const char* key, int64_t value);

/* Name of the value path of the last assembly: "plate_lanes", "plate_blocks",
 * "plate_fused", "plate_corners", "bar_owner" or "generic". */
const char* fem_assembly_path(fem_model* m);

#ifdef __cplusplus
}
#endif

#endif /* FEM_AMD_H */
