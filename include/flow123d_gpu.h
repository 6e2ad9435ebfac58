/*
 * C ABI of the MI355X steady mixed-hybrid Darcy solve and explicit upwind
 * transport step.  Plain pointers and sizes only; every call returns 0 on
 * success or a negative status, with the message in fl_last_error().  A
 * positive status is a warning: the solve calls return FL_NOT_CONVERGED when
 * CG stopped at its iteration limit; the solution (the last iterate) is still
 * available, as in the reference, which prints KSP's convergence reason and
 * goes on (src/la/solve.cc:454-458).
 * Handles are not thread-safe; each handle owns one HIP stream.
 *
 * Statuses with a fixed meaning:
 *   FL_NOT_CONVERGED        1  iteration limit reached (warning, see above)
 *   FL_E_SINGULAR_ELEMENT -10  a local RT0 matrix is singular: |det| below
 *                              NUM_ZERO (the reference stops there,
 *                              src/flow/mh_fe_values.cc:54-59 with
 *                              src/system/math_fce.cc:147,168,211) or not
 *                              finite (a zero-volume element); the message
 *                              names the element id
 *   FL_E_BAD_FIELD        -11  a conductivity or cross-section value is not
 *                              finite and positive (the message names the
 *                              element, and the realization for a batch)
 *   FL_E_DIVERGED         -12  the CG residual became NaN or infinite
 *   FL_E_CALLBACK         -13  a data callback of the caller returned non-zero;
 *                              the step it was to supply was not taken
 *   FL_E_SORPTION         -14  the reaction module's sorption met a negative
 *                              total mass (the reference stops there,
 *                              src/reaction/isotherm.cc:62) or an isotherm
 *                              equation with no root in its bracket (where
 *                              toms748_solve throws, isotherm.cc:127)
 *
 * Each entry point names the Flow123d 1.7 interface it replaces
 * (paths relative to the Flow123d source tree).
 *
 * Arrays follow the reference's numbering exactly:
 *   E bulk elements in mesh-file order; S sides numbered element by element
 *   (MH_DofHandler::reinit, src/flow/mh_dofhandler.cc:12-21); F edges in the
 *   order Mesh::make_neighbours_and_edges creates them (src/mesh/mesh.cc:343-500);
 *   NBC boundary elements = file boundary elements then implicit ones;
 *   the MH solution is [side fluxes (S) | element pressures (E) | edge traces (F)]
 *   (src/flow/darcy_flow_mh.cc:330-366).
 */
#ifndef FLOW123D_GPU_H
#define FLOW123D_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FL_NOT_CONVERGED 1
#define FL_E_SINGULAR_ELEMENT (-10)
#define FL_E_BAD_FIELD (-11)
#define FL_E_DIVERGED (-12)
#define FL_E_CALLBACK (-13)
#define FL_E_SORPTION (-14)

typedef struct fl_mesh fl_mesh;
typedef struct fl_darcy fl_darcy;
typedef struct fl_transport fl_transport;
typedef struct fl_element_data fl_element_data;
typedef struct fl_old_bcd fl_old_bcd;

/* Message of the last failed call on this thread. */
const char *fl_last_error(void);

/* ---- mesh ---------------------------------------------------------------- */

/* Mesh::Mesh + GmshMeshReader::read_mesh + Mesh::setup_topology
 * (src/mesh/msh_gmshreader.cc:76-85, src/mesh/mesh.cc:227-245). */
int fl_mesh_read_gmsh(const char *path, fl_mesh **out);

/* Same, from arrays already in memory: the element list in file order with
 * gmsh types (1 line, 2 triangle, 4 tetrahedron, 15 point), region ids and
 * node indices (0-based, 4 per element, -1 padded); physical names optional
 * (labels starting with '.' are boundary regions, msh_gmshreader.cc:236-239). */
int fl_mesh_from_arrays(int n_nodes, const double *xyz, int n_elem, const int *elem_id, const int *gmsh_type,
                        const int *region_id, const int *nodes4, int n_phys, const int *phys_id,
                        const char *const *phys_name, fl_mesh **out);

void fl_mesh_free(fl_mesh *m);

/* counts[9] = n_nodes, E, NBC, S, F, n_boundaries, n_vb, n_regions, n_file_bc
 * (Mesh::n_elements/n_sides/n_edges/n_vb_neighbours, mesh.h:108-332). */
int fl_mesh_counts(const fl_mesh *m, int *counts);

/* Topology arrays (the bit-exact numbering contract):
 * elem_dim[E], elem_id[E], elem_region[E] (region id), elem_nodes[4E],
 * elem_edge[4E] (Element::edge_idx_), elem_bnd[4E] (Element::boundary_idx_),
 * edge_ptr[F+1] + edge_side[S] (Edge::side_, as global side dofs),
 * vb_elem[V] + vb_edge[V] (Neighbour::element_, Neighbour::edge_idx_),
 * bnd_edge[B] + bnd_bcele[B] (Boundary::edge_idx_, Boundary::bc_ele_idx_),
 * bc_region[NBC] (region id of each boundary element). */
int fl_mesh_topology(const fl_mesh *m, int *elem_dim, int *elem_id, int *elem_region, int *elem_nodes,
                     int *elem_edge, int *elem_bnd, int *edge_ptr, int *edge_side, int *vb_elem, int *vb_edge,
                     int *bnd_edge, int *bnd_bcele, int *bc_region);

/* Region table (RegionDB, src/mesh/region.cc:105-171): ids, boundary flag,
 * index among boundary or bulk regions; labels '\n'-separated into buf. */
int fl_mesh_regions(const fl_mesh *m, int *ids, int *is_boundary, int *sub_idx, char *buf, int buflen);

/* Mesh::elements_id_maps (src/mesh/mesh.cc:781-812): gmsh ids of the E bulk
 * elements (boundary = 0; an error unless strictly increasing) or of the NBC
 * boundary elements, -1 for the implicit ones (boundary = 1). */
int fl_mesh_element_ids(const fl_mesh *m, int boundary, int *ids);

/* ---- element-wise data in gmsh files ------------------------------------- */

/* FieldElementwise::init_from_input + set_mesh
 * (src/fields/field_elementwise_impl.hh:67-76, 121-137) over a
 * GmshMeshReader opened on gmsh_file (src/mesh/msh_gmshreader.cc:51-57): the
 * field `field_name` with n_comp values per element on the bulk (boundary = 0)
 * or boundary (boundary = 1) elements of m.  Values start at zero.  The mesh
 * is only read here. */
int fl_element_data_open(const fl_mesh *m, const char *gmsh_file, const char *field_name, int n_comp, int boundary,
                         fl_element_data **out);

/* FieldElementwise::set_time (field_elementwise_impl.hh:92-117) ->
 * GmshMeshReader::read_element_data (msh_gmshreader.cc:306-397): applies, in
 * file order, every not yet applied $ElementData section of the field whose
 * time is <= time * (1 + 2 eps); rows are matched to the element ids in
 * increasing order.  *updated = 1 when a section was applied.  time = +inf
 * changes nothing.  A field with no section at all is an error on the first call. */
int fl_element_data_set_time(fl_element_data *f, double time, int *updated);

/* FieldElementwise::value over all elements (field_elementwise_impl.hh:144-156):
 * values[n_entities * n_comp], element by element. */
int fl_element_data_values(const fl_element_data *f, double *values);

/* info[3] = n_entities, n_comp, rows stored by the last section applied;
 * the warnings of the last set_time, '\n'-separated, into buf. */
int fl_element_data_info(const fl_element_data *f, int *info, char *buf, int buflen);

void fl_element_data_free(fl_element_data *f);

/* OutputMSH::write_head (src/io/output_msh.cc:53-98, 424-435): creates path
 * with $MeshFormat, $Nodes (index + 1, %g) and $Elements (bulk elements,
 * index + 1, gmsh type, 3 tags: region, region, partition). */
int fl_msh_write_head(const fl_mesh *m, const char *path);

/* OutputMSH::write_msh_elem_data with write_msh_ascii_cont_data
 * (output_msh.cc:100-188, 383-408): appends to path one $ElementData section
 * "name_[units]" at `time` with time index `step`, n_values rows
 * "id v1 .. vn" (ids 1..n_values; rows of more than one value end with a
 * blank).  precision = significant digits of the time and the values; the
 * reference's stream default is 6. */
int fl_msh_write_element_data(const char *path, const char *name, const char *units, double time, int step,
                              int n_comp, int n_values, const double *values, int precision);

/* OldBcdInput::read_flow (src/flow/old_bcd.cc:44-205): the legacy flow
 * boundary file (.fbc) onto the NBC boundary elements of m: bc_type (0 none,
 * 1 Dirichlet, 2 Neumann, 3 Newton/Robin), bc_pressure, bc_flux, bc_sigma, ready
 * for fl_darcy_set_data; elements without a row keep zeros.  The handle keeps
 * the file's row ids for fl_old_bcd_read_transport. */
int fl_old_bcd_read_flow(const fl_mesh *m, const char *path, int *bc_type, double *bc_pressure, double *bc_flux,
                         double *bc_sigma, fl_old_bcd **out);

/* OldBcdInput::read_transport (old_bcd.cc:209-250): the legacy transport
 * boundary file (.tbc), bc_conc[n_subst][NBC] as fl_transport_create takes it. */
int fl_old_bcd_read_transport(const fl_old_bcd *b, const char *path, int n_subst, double *bc_conc);

void fl_old_bcd_free(fl_old_bcd *b);

/* OutputMSH::write_msh_node_data (output_msh.cc:327-381): appends one
 * $NodeData section, laid out as fl_msh_write_element_data's rows. */
int fl_msh_write_node_data(const char *path, const char *name, const char *units, double time, int step,
                           int n_comp, int n_values, const double *values, int precision);

/* ---- balance files --------------------------------------------------------- */

/* DarcyFlowMHOutput::water_balance (src/flow/darcy_flow_mh_output.cc:671-773):
 * appends one water-balance block to path from per-region arrays in the
 * region order of m (boundary index / bulk index): bal, plus, minus
 * [n_boundary_regions], src [n_bulk_regions]. */
int fl_write_water_balance(const fl_mesh *m, const char *path, double time, const double *bal, const double *plus,
                           const double *minus, const double *src);

/* MassBalance::output (src/transport/mass_balance.cc:242-354): appends one
 * mass-balance block; arrays as fl_transport_mass_balance returns them
 * (bal/plus/minus[n_subst][n_boundary_regions], mass/src[n_subst][n_bulk_regions],
 * totals[n_subst][4] = initial mass, integrated flux, integrated sources,
 * current mass), subst_names[n_subst]. */
int fl_write_mass_balance(const fl_mesh *m, const char *path, double time, double initial_time, int n_subst,
                          const char *const *subst_names, const double *bal, const double *plus, const double *minus,
                          const double *mass, const double *src, const double *totals);

/* ---- steady MH Darcy flow -------------------------------------------------- */

/* DarcyFlowMH_Steady::DarcyFlowMH_Steady (src/flow/darcy_flow_mh.cc:249-367):
 * builds the device mesh, the condensed edge-system pattern and the
 * precomputed scatter lists. */
int fl_darcy_create(fl_mesh *m, fl_darcy **out);
void fl_darcy_free(fl_darcy *d);

/* Field values the reference reads in assembly_steady_mh_matrix
 * (darcy_flow_mh.cc:525-694, EqData fields :152-183):
 * per bulk element: conductivity K, cross_section cs, sigma, water_source_density f,
 * anisotropy (9 per element, row-major; NULL = identity);
 * per boundary element: bc_type (0 none, 1 dirichlet, 2 neumann, 3 robin),
 * bc_pressure, bc_flux, bc_robin_sigma.  FL_E_BAD_FIELD unless every K and
 * cs is finite and positive (the reference does not check them). */
int fl_darcy_set_data(fl_darcy *d, const double *K, const double *cs, const double *aniso, const double *sigma,
                      const double *f, const int *bc_type, const double *bc_pressure, const double *bc_flux,
                      const double *bc_robin_sigma);

/* Replace the conductivity only (new Monte Carlo realization). */
int fl_darcy_set_conductivity(fl_darcy *d, const double *K);

/* Solver options (Solver record, src/la/solve.cc:69-84): r_tol, a_tol, max_it;
 * precision: 0 = FP64 CG (the drop-in path), 1 = FP32 CG with FP64 reliable
 * updates (the stop test is always on the true FP64 residual);
 * norm: residual norm of the stop test ||r|| <= max(r_tol ||b||, a_tol):
 * 0 = the 2-norm, 1 = the natural norm of the Jacobi preconditioner,
 * sqrt(r^T D^-1 r) with D the diagonal of the edge system (PETSc's
 * KSP_NORM_NATURAL; the reference's default KSP CG tests a preconditioned
 * norm, solve.cc:370).  On strongly heterogeneous K the natural norm tracks the
 * flux error far better than the 2-norm.  Residuals reported back are in the
 * same norm. */
typedef struct {
    double r_tol;
    double a_tol;
    int max_it;
    int precision;
    int norm;
} fl_solver_opts;

/* Result of a solve: CG iterations, final residual 2-norm, rhs 2-norm,
 * converged flag, device times of the phases in ms, and the average device
 * durations of the CG iteration's two halves (direction update + SpMV, and
 * the vector update), sampled with HIP events on the solver stream every 8th
 * iteration. */
typedef struct {
    int iterations;
    int converged;
    double residual;
    double rhs_norm;
    double ms_assembly;
    double ms_solve;
    double ms_resolve;
    double ms_total;
    double ms_spmv_avg;
    double ms_update_avg;
    int kernel_samples;
    int reliable_updates;
} fl_solve_info;

/* DarcyFlowMH_Steady::update_solution (darcy_flow_mh.cc:375-433).  Returns
 * FL_NOT_CONVERGED with info->converged = 0 at the iteration limit,
 * FL_E_SINGULAR_ELEMENT or FL_E_DIVERGED as described at the top. */
int fl_darcy_update_solution(fl_darcy *d, const fl_solver_opts *opts, fl_solve_info *info);

/* No counterpart in the reference.  For callers that cannot allocate during
 * the call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fl_darcy_update_solution in FP64 (opts->precision must
 * be 0) that allocates nothing and creates no event.  Its CG partial sums and
 * state live in a file-scope device array of the library, 5 x 8192 + 8
 * doubles (328 KiB), held by each GPU the library runs on until the process
 * ends; a handle whose CG grid exceeds 8192 blocks is rejected.  Every call on
 * a GPU shares that array, so calls that overlap on one GPU race.  The kernel
 * duration fields of info stay 0.  Same statuses and results as
 * fl_darcy_update_solution. */
int fl_darcy_update_solution_noalloc(fl_darcy *d, const fl_solver_opts *opts, fl_solve_info *info);

/* DarcyFlowMH_Unsteady (lumped = 0) or DarcyFlowLMH_Unsteady (lumped = 1)
 * setup_time_term (darcy_flow_mh.cc:1601-1658, 1680-1749) on the data of
 * fl_darcy_set_data: storativity[E], init_pressure[E].  The solution becomes
 * the initial state: element pressures = init_pressure (MH), or edge traces =
 * the sum of init_pressure / n_sides over each edge's sides (LMH), the rest 0. */
int fl_darcy_set_unsteady(fl_darcy *d, int lumped, const double *storativity, const double *init_pressure);

/* One implicit Euler step of length dt from the current solution
 * (update_solution with modify_system, darcy_flow_mh.cc:1660-1677 / 1751-1770,
 * and the LMH postprocess :1772-1827): MH adds S|e|/dt to each element's
 * pressure row and S|e|/dt p_prev to its right-hand side; LMH adds
 * |e| S cs / dt / n_sides per side to its edge's row and the same times the
 * previous trace to its right-hand side, then corrects every side flux by the
 * lumped storage and source terms.  A plain fl_darcy_update_solution afterwards
 * solves the steady problem again. */
int fl_darcy_unsteady_step(fl_darcy *d, double dt, const fl_solver_opts *opts, fl_solve_info *info);

/* ---- time coupling of the equations (Equation, equation.hh:113-146) -------
 * Each Darcy handle carries the reference's TimeGovernor (time_governor.cc).
 * By default the steady one (darcy_flow_mh.cc:263): t = 0 and dt = inf until
 * fl_darcy_update_solution, then t = inf, dt = 0 and is_end (a steady handle
 * still solves again on every fl_darcy_update_solution, e.g. after new data).
 * fl_darcy_set_time(end_time, init_dt, save_step) after fl_darcy_set_unsteady
 * gives the unsteady governor of DarcyFlowMH_Unsteady with the input keys
 * init_dt (fixed step) and the output save_step (its marks; <= 0: none),
 * its step fixed until the next mark at once (darcy_flow_mh.cc:1606-1608):
 * every fl_darcy_update_solution is then one implicit step of the
 * governor's next dt, and nothing once it is at the end.
 * planned_time = t + estimate_dt, solved_time = t, choose_next_time =
 * fix_dt_until_mark, set_time_upper_constraint = set_upper_constraint. */
int fl_darcy_set_time(fl_darcy *d, double end_time, double init_dt, double save_step);
int fl_darcy_time(fl_darcy *d, double *time, double *dt, int *is_end);
int fl_darcy_set_time_upper_constraint(fl_darcy *d, double dt);
int fl_darcy_planned_time(fl_darcy *d, double *time);
int fl_darcy_solved_time(fl_darcy *d, double *time);
int fl_darcy_choose_next_time(fl_darcy *d);

/* DarcyFlowMH_Steady::get_solution_vector (darcy_flow_mh.cc:485-500):
 * copies the [q | p | lambda] vector (S+E+F doubles) to out. */
int fl_darcy_get_solution(fl_darcy *d, double *out);

/* Device pointer of the same vector (valid until the next update_solution);
 * the analogue of get_parallel_solution_vector (darcy_flow_mh.cc:502-505). */
const double *fl_darcy_device_solution(fl_darcy *d);

/* DarcyFlowMH_Steady::solution_precision (darcy_flow_mh.cc:463-482). */
double fl_darcy_solution_precision(fl_darcy *d, double r_tol, double a_tol);

/* Per-element mass balance of the element row of the MH system
 * (darcy_flow_mh.cc:603-640): sum_i q_i + sum_vb v (p_e - lambda) - |e| cs f. */
int fl_darcy_element_balance(fl_darcy *d, double *out);

/* DarcyFlowMHOutput::water_balance (darcy_flow_mh_output.cc:671-773):
 * per boundary region (by boundary sub-index) total, outflow, inflow. */
int fl_darcy_water_balance(fl_darcy *d, double *balance, double *outflow, double *inflow);

/* The condensed SPD edge system (the reference's schur2 system after
 * SchurComplement::scale(-1), schur.cc:702-707): CSR with F+1 row pointers,
 * nnz columns / values, and the rhs.  nnz via fl_darcy_edge_nnz. */
int fl_darcy_edge_nnz(fl_darcy *d);
int fl_darcy_edge_system(fl_darcy *d, int *ptr, int *col, double *val, double *rhs);

/* DarcyFlowMHOutput::make_element_scalar / make_element_vector
 * (src/flow/darcy_flow_mh_output.cc:261-325) for the current solution:
 * pressure[E], piezo_head[E] (pressure + z of the element centre) and
 * velocity[3E], the Darcy velocity at the element centre from the RT0 side
 * fluxes divided by the cross-section (MHFEValues::RT0_value,
 * mh_fe_values.cc:76-118). */
int fl_darcy_element_output(fl_darcy *d, double *pressure, double *piezo_head, double *velocity);

/* DarcyFlowMHOutput::make_node_scalar_param (darcy_flow_mh_output.cc:385-480):
 * pressure_nodes[n_nodes], the element pressures and side traces around each
 * node averaged with the reference's distance weights. */
int fl_darcy_node_pressure(fl_darcy *d, double *pressure_nodes);

/* DarcyFlowMHOutput::water_balance for the current solution: boundary side
 * fluxes and the |e| cs f sources per region, one block appended to path. */
int fl_darcy_write_water_balance(fl_darcy *d, const char *path, double time);

/* ---- batched Monte-Carlo solves ------------------------------------------- */

/* n_real conductivity realizations of one Darcy problem solved together (no
 * counterpart in the reference: a caller would otherwise loop
 * DarcyFlowMH_Steady::update_solution, darcy_flow_mh.cc:375-433, over K
 * fields).  The batch shares the mesh, cross-sections, sources and boundary
 * conditions of `base` (fl_darcy_set_data must have been called); only K
 * differs.  The base handle must outlive the batch. */
typedef struct fl_darcy_batch fl_darcy_batch;
int fl_darcy_batch_create(fl_darcy *base, int n_real, fl_darcy_batch **out);
/* The same with n_slots concurrent realization lanes (rounded up to 64;
 * 0 = min(n_real, 256)): a slot whose realization has converged is refilled
 * with the next one, hardest realizations (largest ln K variance) first. */
int fl_darcy_batch_create_slots(fl_darcy *base, int n_real, int n_slots, fl_darcy_batch **out);
void fl_darcy_batch_free(fl_darcy_batch *b);
/* K[n_real][E]; device_ptr != 0: K is a device pointer. */
int fl_darcy_batch_set_conductivity(fl_darcy_batch *b, const double *K, int device_ptr);

typedef struct {
    int n_real, chunks, reliable_updates, fp64_finish, max_iterations, n_converged;
    long long total_iterations; /* summed over realizations */
    double ms_setup, ms_solve, ms_resolve, ms_total;
    double ms_op_avg, ms_update_avg; /* live event samples of the operator / update kernels */
    int grid, xcd_map, chunk;        /* launch shape used */
    int compactions;                 /* times live realizations were packed into fewer 64-wide chunks */
    int handovers;                   /* realizations finished by the single-solve FP64 path */
    long long handover_iterations;   /* CG iterations they took there */
    double ms_handover;              /* wall time of that phase */
    double ms_coarse_avg;            /* two-level: coarse solve + direction kernel per iteration (sampled) */
    double ms_coarse_setup;          /* two-level: coarse operators and inverses, all refills */
    /* wall-time phases of the solve (ms from its start): the queue ran empty, the
     * two-level kernels narrowed to 16 / 4 lanes per row (0: never), the batch
     * loop ended (hand-over issued), the hand-over path finished */
    double ms_queue_empty, ms_narrow16, ms_narrow4, ms_loop_end, ms_handover_end;
    int tile_rows; /* two-level FP32 operator: rows per tile of the tiling the tuner kept */
} fl_batch_info;

/* Solve every realization to ||b - S x|| <= max(r_tol ||b||, a_tol) (opts as
 * for fl_darcy_update_solution; precision 1 = FP32 CG with FP64 reliable
 * updates).  Per-realization iterations / converged flags / final FP64
 * relative residuals go to the optional arrays [n_real]. */
int fl_darcy_batch_update_solution(fl_darcy_batch *b, const fl_solver_opts *opts, fl_batch_info *info,
                                   int *iterations, int *converged, double *residual);
/* Per-realization outcome of the last batch solve: status (1 converged,
 * 4 iteration limit, 5 breakdown) and the number of FP64 reliable updates
 * (precision 1); arrays [n_real], either may be NULL. */
int fl_darcy_batch_diagnostics(fl_darcy_batch *b, int *status, int *reliable_updates);
/* [q | p | lambda] of realization r, as fl_darcy_get_solution. */
int fl_darcy_batch_get_solution(fl_darcy_batch *b, int r, double *out);
/* Device solutions, realization-interleaved: value i of realization r at
 * [i * n_padded + r], n_padded = n_real rounded up to 64. */
const double *fl_darcy_batch_device_solution(fl_darcy_batch *b, int *n_padded);
/* Launch shape (grid blocks per 64 realizations, multiple of 8 up to 2048;
 * row mapping 0 strided, 1 XCD-contiguous, 2 XCD- and block-contiguous;
 * iterations per captured graph chunk, 0 automatic: 64 with the two-level
 * preconditioner, 16 without) and the autotuner that picks grid and mapping
 * by timing the operator kernel. */
int fl_darcy_batch_set_launch(fl_darcy_batch *b, int grid, int xcd_map, int chunk);
int fl_darcy_batch_autotune(fl_darcy_batch *b);
/* Once the queue is empty and at most n realizations are still iterating,
 * finish them with the single-solve FP64 CG (warm start from the batch
 * iterate), up to 4 side by side on solvers and streams of their own (copies
 * of the base solver on the same mesh, made at the first hand-over and kept
 * by the batch handle); 0 disables; negative restores the default: 16, or 8
 * when the two-level tail runs narrowed (fl_darcy_batch_set_narrow). */
int fl_darcy_batch_set_handover(fl_darcy_batch *b, int n);
/* Two-level preconditioner only: once the queue is empty and at most 16 (then
 * 4) realizations are still iterating, they continue in the batch with 16 (4)
 * lanes per edge row, keeping the coarse space; the hand-over threshold above
 * still applies first.  on = 1 (default) enables it. */
int fl_darcy_batch_set_narrow(fl_darcy_batch *b, int on);
/* Two-level preconditioner of the batched CG: ncol 0 = Jacobi only, 1 = plus
 * one constant coarse column per aggregate, 2 = plus a second column phi per
 * aggregate (fl_darcy_batch_set_coarse_basis; e.g. the learned basis of
 * fl_gnn_*).  M^-1 = D^-1 + Z (Z^T S Z)^-1 Z^T, one fixed SPD operator per
 * realization (Fletcher-Reeves CG).  The coarse inverses are formed on the
 * GPU (blocked FP64 Cholesky, L^-1, L^-T L^-1) and applied from their packed
 * lower 64 x 64 tiles in FP32. */
int fl_darcy_batch_set_coarse(fl_darcy_batch *b, int ncol);
/* phi[n_real][F] (FP32, device pointer, edge order); normalised per aggregate. */
int fl_darcy_batch_set_coarse_basis(fl_darcy_batch *b, const float *phi);
/* Aggregate of every edge (edge order) into edge_agg[F] (may be NULL);
 * returns the number of aggregates. */
int fl_darcy_batch_aggregates(fl_darcy_batch *b, int *edge_agg);
/* Edge traces lambda of every realization of the last solve into
 * out[n_real][F] (device pointer, edge order). */
int fl_darcy_batch_lambda(fl_darcy_batch *b, double *out);
/* Inverse of np SPD matrices a[np][n][n] (host, row-major, lower triangle
 * read, n <= 2048) by the coarse space's batched Cholesky + inverse kernels,
 * rounded to FP32 into inv[np][n][n] (host); info[np]: 0, or 1 for a matrix
 * that is not positive definite (its inverse is returned as zeros).  No
 * counterpart in the reference: it checks the coarse-space factorization
 * against a host inverse. */
int fl_coarse_invert(const double *a, int n, int np, float *inv, int *info);

/* ---- learned coarse basis (graph network) --------------------------------- */

/* A 3-layer message-passing network (hidden width 64) on the element-edge
 * graph of the batch's mesh that maps a conductivity field to the second
 * coarse column phi of fl_darcy_batch_set_coarse(b, 2).  No counterpart in
 * the reference.  BF16 MFMA forward and backward kernels with FP32
 * accumulation, FP32 master weights and Adam, all on the GPU.  Trained on
 * (K, lambda) pairs from the solver: the loss is the share of lambda that the
 * columns {1, phi} of each aggregate cannot represent. */
typedef struct fl_gnn fl_gnn;
int fl_gnn_create(fl_darcy_batch *b, unsigned seed, fl_gnn **out);
void fl_gnn_free(fl_gnn *g);
int fl_gnn_n_params(fl_gnn *g);
int fl_gnn_get_params(fl_gnn *g, float *out); /* host arrays, n_params */
int fl_gnn_set_params(fl_gnn *g, const float *in);
/* one realization: K[E] (device FP64) -> phi[F] (device FP32, edge order; may be NULL) */
int fl_gnn_forward(fl_gnn *g, const double *K, float *phi);
/* loss of the last forward against lambda[F] (device FP64, edge order);
 * the gradient is added to the accumulated gradient */
int fl_gnn_loss_grad(fl_gnn *g, const double *lambda, double *loss);
int fl_gnn_zero_grad(fl_gnn *g);
int fl_gnn_get_grad(fl_gnn *g, float *out); /* host, n_params */
int fl_gnn_adam_step(fl_gnn *g, double lr);
/* Training on n pairs K[n][E], lambda[n][F] (device): `epochs` passes, an
 * Adam step every `batch` samples; mean loss per epoch into epoch_loss. */
int fl_gnn_train(fl_gnn *g, const double *K, const double *lambda, int n, int epochs, int batch, double lr,
                 double *epoch_loss);
/* phi[R][F] for K[R][E] (device arrays) */
int fl_gnn_infer(fl_gnn *g, const double *K, int R, float *phi);
/* inputs of the last forward (host FP32 [E][16], [F][16]; either may be NULL) */
int fl_gnn_features(fl_gnn *g, float *xe, float *xf);

/* ---- explicit upwind transport -------------------------------------------- */

/* ConvectionTransport::ConvectionTransport (src/transport/transport.cc:105-188)
 * with TransportOperatorSplitting's time governor: n_subst substances (1..8),
 * mobile porosity por[E], cross_section cs[E], bc_conc[n_subst][NBC],
 * init_conc[n_subst][E], end_time, save_step (output marks). */
int fl_transport_create(fl_mesh *m, int n_subst, const double *por, const double *cs, const double *bc_conc,
                        const double *init_conc, double end_time, double save_step, fl_transport **out);
void fl_transport_free(fl_transport *t);

/* TransportBase::set_velocity_field (transport_operator_splitting.hh:90-92):
 * side fluxes [S] from host memory, or straight from a Darcy handle's device
 * solution. */
int fl_transport_set_velocity(fl_transport *t, const double *side_flux);
int fl_transport_set_velocity_from_darcy(fl_transport *t, fl_darcy *d);

/* ConvectionTransport::create_transport_matrix_mpi (transport.cc:788-943) for
 * the current velocity field, done ahead of the first step; cfl receives
 * cfl_max_step = 1 / max_i |a_ii| (transport.cc:911-921). */
int fl_transport_prepare(fl_transport *t, double *cfl);

/* TransportOperatorSplitting::update_solution (transport_operator_splitting.cc:244-287):
 * ConvectionTransport::set_target_time + compute_one_step until target_time.
 * steps (optional) receives the number of steps taken. */
int fl_transport_update_solution(fl_transport *t, double target_time, int *steps);

/* ConvectionTransport::get_concentration_matrix (transport.h:107-173): out[n_subst][E]. */
int fl_transport_get_conc(fl_transport *t, double *out);

/* time(), dt() of the convection's governor (its CFL steps), cfl_max_step */
int fl_transport_time(fl_transport *t, double *time, double *dt, double *cfl);

/* The operator-splitting equation's own TimeGovernor (transport_operator_splitting.cc:213:
 * marks at 0, end_time and every save_step of fl_transport_create), as the
 * coupling loop of HC_ExplicitSequential::run_simulation
 * (hc_explicit_sequential.cc:186-231) drives it: set_time_upper_constraint
 * (the water's dt), planned_time (t + estimate_dt), solved_time (t),
 * choose_next_time (fix_dt_until_mark), and fl_transport_update_solution_step
 * = TransportOperatorSplitting::update_solution (:244-287): the governor's
 * next_time, then the convection's CFL steps up to it.  This governor follows
 * only these calls (fl_transport_update_solution with an explicit target
 * leaves it). */
int fl_transport_time_governor(fl_transport *t, double *time, double *dt, int *is_end);
int fl_transport_set_time_upper_constraint(fl_transport *t, double dt);
int fl_transport_planned_time(fl_transport *t, double *time);
int fl_transport_solved_time(fl_transport *t, double *time);
int fl_transport_choose_next_time(fl_transport *t);
int fl_transport_update_solution_step(fl_transport *t, int *steps);

/* MassBalance::calculate (mass_balance.cc:79-239) state after the last step:
 * bal/plus/minus[n_subst][n_boundary_regions], mass[n_subst][n_bulk_regions],
 * src[n_subst][n_bulk_regions] (source term of the last step),
 * totals[n_subst][4] = initial mass, integrated boundary flux, integrated
 * sources, current mass. */
int fl_transport_mass_balance(fl_transport *t, double *bal, double *plus, double *minus, double *mass, double *src,
                              double *totals);

/* MassBalance::output at the current time of the transport: one block
 * appended to path.  Before the first step the block shows the initial state
 * (initial mass = current mass, interval [t, t]), as the reference's output at
 * the start time does. */
int fl_transport_write_mass_balance(fl_transport *t, const char *path, const char *const *subst_names);

/* Time-dependent boundary concentrations (a bc_conc field given as a list of
 * data times, e.g. old_boundary_file entries with `time`): bc_conc[n_subst][NBC]
 * takes effect at the start of the next step, as data_.set_time does at the start
 * of ConvectionTransport::compute_one_step (transport.cc:598-627: the boundary
 * vector is rebuilt and scaled by the step estimate).  Call it after the
 * fl_transport_update_solution that reached the data time; balances read at that
 * output time still use the previous values, as the reference's do. */
int fl_transport_set_bc_conc(fl_transport *t, const double *bc_conc);

/* Linear_reaction (src/reaction/linear_reaction.cc:144-300) as
 * TransportOperatorSplitting drives it (transport_operator_splitting.cc:244-287):
 * n_decays first-order decays, decay i with parent substance parent[i],
 * products products[product_ptr[i] .. product_ptr[i+1]) with branch_ratio[..],
 * and half_life[i] (a kinetic constant k is half_life = ln 2 / k,
 * linear_reaction.cc:214-224).  At every fl_transport_update_solution the
 * matrix R is rebuilt from the step estimate dt: identity, R[p][p] =
 * 0.5^(dt/half_life), R[p][product] = (1 - 0.5^(dt/half_life)) * ratio; after
 * each convection step every element's concentrations become c R.
 * n_decays = 0 switches reactions off. */
int fl_transport_set_linear_reactions(fl_transport *t, int n_decays, const int *parent, const int *product_ptr,
                                      const int *products, const double *half_life, const double *branch_ratio);

/* Pade_approximant (src/reaction/pade_approximant.cc:97-290): the decays of
 * fl_transport_set_linear_reactions with the step matrix the (nom_pol_deg,
 * den_pol_deg) Pade approximant of exp(A dt), A[p][p] = -ln2 dt/half_life,
 * A[p][product] = ln2 dt/half_life * ratio: N(A) and D(A) by Horner, P =
 * (D^-1 N)^T, each element's concentrations become P c after every convection
 * step.  Each polynomial is evaluated with its own degree. */
int fl_transport_set_pade_reactions(fl_transport *t, int n_decays, const int *parent, const int *product_ptr,
                                    const int *products, const double *half_life, const double *branch_ratio,
                                    int nom_pol_deg, int den_pol_deg);

/* Concentration sources (ConvectionTransport::compute_concentration_sources,
 * transport.cc:501-544): sources_density, sources_sigma, sources_conc, each
 * [n_subst][E]; the step adds (density + max(conc_src - c, 0) sigma) dt. */
int fl_transport_set_sources(fl_transport *t, const double *density, const double *sigma, const double *conc);

/* Sorption of the reaction module (Sorption, reaction/sorption.cc:71-285 with
 * Isotherm, isotherm.cc:33-258; the "adsorptions" record of
 * TransportOperatorSplitting, transport_operator_splitting.cc:178-192), applied
 * after every step and after the decays: species k on substance subst[k];
 * sorption_type (0 none, 1 linear, 2 freundlich, 3 langmuir), mult_coef,
 * second_param [n_bulk_regions][n]; rock_density [n_bulk_regions]; molar_mass and
 * solubility [n] (solubility null or <= 0: no limit); solvent_dens; substeps
 * (the reference's default 1000) table intervals.  scale_sorbed is
 * phi (1 - por_m - por_imm) rock_density molar_mass with por_imm and phi of
 * fl_transport_set_dual_porosity_sorption (defaults 0 and 1, sorption.cc:172-213).
 * A bulk region on which the porosities por_m, por_imm and phi are constant is
 * tabulated (Sorption::make_tables, sorption.cc:148-170); on any other region
 * every element's isotherm is solved directly (Sorption::compute_reaction
 * without a table, sorption.cc:236-260, isotherm.cc:184-200).  The direct solve
 * brackets the root on [0, critical mass / por_m + 1e-5] as the reference does
 * (isotherm.cc:93-99): without a positive solubility that interval is empty and
 * the first step stops with FL_E_SORPTION, where the reference's root finder
 * throws.  Dual porosity off.  Set before the first step; the isotherms are made
 * at the first step. */
int fl_transport_set_reaction_sorption(fl_transport *t, int n, const int *subst, const int *sorption_type,
                                       const double *mult_coef, const double *second_param,
                                       const double *rock_density, const double *molar_mass,
                                       const double *solubility, double solvent_dens, int substeps);

/* The same sorption with element-wise parameter fields (the reference's
 * bulk_data given as element data, so no region has constant parameters and
 * every element is solved directly, sorption.cc:236-260): sorption_type,
 * mult_coef, second_param [n][E]; rock_density [E]; molar_mass, solubility [n]. */
int fl_transport_set_reaction_sorption_elementwise(fl_transport *t, int n, const int *subst, const int *sorption_type,
                                                   const double *mult_coef, const double *second_param,
                                                   const double *rock_density, const double *molar_mass,
                                                   const double *solubility, double solvent_dens);

/* Sorbed concentrations of the reaction module's species, out[n][E]. */
int fl_transport_get_reaction_sorbed(fl_transport *t, double *out);

/* Boundary concentrations given as a function of time (a time-dependent
 * FieldFormula for bc_conc, field_formula_impl.hh:70-122, read by
 * ConvectionTransport::compute_one_step at every step's start, transport.cc:598-627):
 * fn(t, bc_conc[n_subst][NBC], user) is called on the host at the start time t of
 * every step and its data drive that step; a non-zero return stops the update
 * with FL_E_CALLBACK before that step.  fn null: the last data stay fixed. */
typedef int (*fl_bc_conc_fn)(double time, double *bc_conc, void *user);
int fl_transport_set_bc_conc_function(fl_transport *t, fl_bc_conc_fn fn, void *user);

/* Dual porosity and sorption of ConvectionTransport (dual_porosity and
 * sorption_enable; transport_dual_porosity / transport_sorption / compute_sorption,
 * transport.cc:1026-1160), applied after every convection step and before the
 * decays (which then act on the immobile phase too, linear_reaction.cc:295-301).
 * por_imm[E] and phi[E] (null: 1.0); alpha, sorp_type (0 none, 1 linear,
 * 2 freundlich, 3 langmuir; reaction/isotherm.hh:71-76), sorp_coef0, sorp_coef1
 * each [n_subst][E], null where the model reading them is off.  Set before the
 * first step; the immobile and sorbed phases start at zero.  The region mass of
 * the balance then sums every phase in use times por_m (transport.cc:1354-1362). */
int fl_transport_set_dual_porosity_sorption(fl_transport *t, int dual_porosity, int sorption, const double *por_imm,
                                            const double *alpha, const double *phi, const int *sorp_type,
                                            const double *sorp_coef0, const double *sorp_coef1);

/* get_concentration_matrix()[phase] (transport.h:58-63, 107-173): phase 0 mobile,
 * 1 immobile, 2 mobile sorbed, 3 immobile sorbed; out[n_subst][E] (zero for a
 * phase not in use). */
int fl_transport_get_conc_phase(fl_transport *t, int phase, double *out);

/* Scaled transport matrix tm = I + dt*A (create_transport_matrix_mpi,
 * transport.cc:788-943, after the scaling at :628-648). */
int fl_transport_nnz(fl_transport *t);
int fl_transport_matrix(fl_transport *t, int *ptr, int *col, double *val);

/* Device time (ms) of the transport steps so far and their count. */
int fl_transport_stats(fl_transport *t, double *ms_steps, int *steps);

/* ---- measurement (no counterpart in the reference) ------------------------ */

/* Achievable HBM bandwidth of the current device, the denominator beside the
 * vendor peak in roofline fractions: gbs[0] copy (read + write counted),
 * gbs[1] read only, gbs[2] write only, in GB/s, over buffers of `bytes` each
 * (>= 1 MiB, a multiple of 16; allocated and freed by the call), the best of
 * `reps` timed launches after a warm-up. */
int fl_device_stream_bandwidth(size_t bytes, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif
