/*
 * fesom_tracer_amd.h -- C ABI of the MI355X tracer-advection library
 * (libfesom_tracer_amd.so).
 *
 * Part 1 keeps the 23 entry points of the reference's tracer_c_interface
 * (lib/tracer_c_interface.F90) with the same names, argument meaning and
 * return codes (0 = ok, -1 = not initialised / invalid id / missing
 * directory), writing errors to stderr as "### error: ...". Like the
 * reference, the library holds one module-level instance and is not
 * thread-safe. Arrays are Fortran column-major float64, (nl-1, nod2D) for
 * tracers.
 *
 * Part 2 adds what the reference's array-initialisation path lacks to run a
 * step (mesh topology, schemes, element velocity, multi-tracer stepping,
 * diagnostics). These have no counterpart in the reference; each says so.
 */
#ifndef FESOM_TRACER_AMD_H
#define FESOM_TRACER_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------
 * Part 1: the reference C ABI
 * ------------------------------------------------------------------- */

/* lib/tracer_c_interface.F90:46-84. One process drives one GPU; no MPI is
 * linked. Records the communicator and returns 0. */
int tracer_init_mpi(int mpi_comm);

/* lib/tracer_c_interface.F90:89-106. Returns 0. */
int tracer_finalize_mpi(void);

/* lib/tracer_c_interface.F90:113-157. Reads the derived-type binary
 * restarts <restart_dir>/np1/t_mesh.0, t_partit.0, t_tracer.0, t_dynamics.0
 * (lib/io_restart_derivedtype.F90:154-199; the layout of FESOM's default
 * double-precision build, one rank): mesh, tracers (values, valuesAB,
 * schemes, tra_adv_ph/pv), velocity. -1 with an error if the directory is
 * missing, a file is unreadable, or its items do not fill its record
 * exactly (another build or layout). */
int tracer_init(const char *restart_dir);

/* lib/tracer_c_interface.F90:163-220. One step of tracer 1 on the GPU:
 * zero the advective tendencies, do_oce_adv_tra, del_ttf += dttf_h+dttf_v,
 * values += del_ttf/hnode_new. del_ttf is zeroed every step unless the
 * option "accumulate_del_ttf" is set, which reproduces the reference (it
 * never zeroes it). */
int tracer_advect_step(double dt);

/* lib/tracer_c_interface.F90:228-248: min, max and sum of all
 * (nl-1) x nod2D values of tracer 1 (inactive levels included). As the
 * reference's sum(values), the sum is NaN when a value is NaN; min and max
 * skip NaN values (tracer_get_diagnostics counts them). */
int tracer_get_stats(double *tmin, double *tmax, double *tsum);

/* lib/tracer_c_interface.F90:255-272: nz = nl-1, nn = myDim_nod2D */
int tracer_get_size(int *nz, int *nn);

/* lib/tracer_c_interface.F90:280-302: tracer 1 values, (nz, nn) column-major */
int tracer_get_values(double *values, int nz, int nn);

/* lib/tracer_c_interface.F90:308-348. Writes the restarts that
 * tracer_init reads to <output_dir>/np1/ (created if needed) and
 * <output_dir>/meta.time, with the current tracers, and marks the library
 * uninitialised. Fields the tracer path does not carry are written
 * unallocated, scalars at FESOM's defaults. */
int tracer_finalize(const char *output_dir);

/* lib/tracer_c_interface.F90:356-386: init, nsteps of tracer_advect_step
 * (printing min/max/sum after each), finalize into the same directory */
int tracer_run_workflow(const char *restart_dir, int nsteps, double dt);

/* lib/tracer_c_interface.F90:391-402: resets the flags (and frees the GPU state) */
int tracer_cleanup(void);

/* lib/tracer_c_interface.F90:415-433 */
int tracer_set_mesh_dims(int nl, int nod2D, int edge2D, int elem2D);

/* lib/tracer_c_interface.F90:441-469: nlevels_nod2D / ulevels_nod2D */
int tracer_set_levels(const int *nlevels, const int *ulevels, int nod2D);

/* lib/tracer_c_interface.F90:477-503: sets hnode_new only, (nl-1, nod2D) */
int tracer_set_thickness(const double *hnode, int nl, int nod2D);

/* lib/tracer_c_interface.F90:510-530 */
int tracer_set_partition(int myDim_nod2D, int eDim_nod2D);

/* lib/tracer_c_interface.F90:539-604: schemes default to UPW1/UPW1/NONE,
 * tra_adv_ph = tra_adv_pv = 1 */
int tracer_allocate_tracers(int num_tracers, int nl, int nod2D, int AB_order);

/* lib/tracer_c_interface.F90:613-644: tracer_id is 1-based; copies into
 * values and valuesAB */
int tracer_set_values(int tracer_id, const double *values, int nl, int nod2D);

/* lib/tracer_c_interface.F90:652-684 */
int tracer_set_velocity_zero(int nl, int nod2D, int edge2D);

/* lib/tracer_c_interface.F90:694-744. The buffer holds (nl-1)*edge2D*2
 * doubles; as in the reference (which passes dynamics%uv to
 * do_oce_adv_tra's explicit-shape vel(2, nl-1, elem2D) dummy), the first
 * 2*(nl-1)*elem2D of them are read as vel(2, nl-1, elem2D). w is
 * (nl, nod2D); w_e and w_i are set to zero as the reference does. */
int tracer_set_velocity(const double *uv, const double *w, int nl, int nod2D, int edge2D);

/* lib/tracer_c_interface.F90:749-759 */
int tracer_init_complete(void);

/* lib/tracer_c_interface.F90:769-821: mesh, partition and velocity from the
 * restarts under <restart_dir>/np1 (no tracers; allocate them next) */
int tracer_load_mesh_partition(const char *restart_dir);

/* lib/tracer_c_interface.F90:830-868 (mesh_setup, lib/oce_mesh.F90:49-94 on
 * one rank): FESOM mesh files in mesh_path -- nod2d.out (lon, lat in degrees),
 * elem2d.out, aux3d.out (nl, zbar), elvls.out, nlvls.out, and edges.out /
 * edge_tri.out / edgenum.out (built from the elements when absent). Node
 * coordinates are rotated by FESOM's default Euler angles (50, 15, -90
 * degrees; tracer_set_option("force_rotation", 0) keeps them as read), the
 * longitude is cyclic (360 degrees), and the geometry is spherical
 * (mesh_areas / mesh_auxiliary_arrays). nlvls.out must equal the deepest
 * level of each node's elements. -1 when the directory or a required file
 * is missing or inconsistent, or a node has more than 16 edges or elements. Follow with tracer_allocate_tracers,
 * tracer_set_values and a velocity, as after tracer_generate_mesh. */
int tracer_load_mesh_from_files(const char *mesh_path);

/* lib/tracer_c_interface.F90:875-921: hnode = hnode_new = helem = zbar(k)-zbar(k+1) */
int tracer_init_ale_arrays(void);

/* lib/tracer_c_interface.F90:931-997: allocates dynamics/work arrays; here
 * also uploads mesh, velocity and tracers to the GPU */
int tracer_init_for_advection(void);

/* ---------------------------------------------------------------------
 * Part 2: extensions (no counterpart in the reference)
 * ------------------------------------------------------------------- */

/* No reference counterpart. Builds a doubly periodic (periodic != 0) or
 * closed nx x ny node rectangle of Lx x Ly metres split into right
 * triangles with FESOM's orientation (edge_tri(1) left of the edge,
 * clockwise elements), nl interfaces of thickness dz (stretched by
 * 'stretch'), per-node bottom levels drawn from [lev_lo, lev_hi] with
 * 'seed' (elements = min of nodes, nodes = max of elements), nodes and
 * elements ordered along a Hilbert curve when sfc != 0. Replaces the
 * reference's generate_analytic_mesh (lib/analytic_mesh.F90:114). */
int tracer_generate_mesh(int nx, int ny, int periodic, double Lx, double Ly, int nl, double dz,
                         double stretch, int lev_lo, int lev_hi, unsigned long long seed, int sfc);

/* No reference counterpart. Scheme strings as tra_adv_hor/ver/lim
 * (lib/MOD_TRACER.F90:18-20): hor 'UPW1'|'MUSCL'|'MFCT', ver
 * 'UPW1'|'QR4C'|'PPM'|'CDIFF', lim 'NONE'|'FCT'. -1 on an unknown scheme
 * (the reference aborts: lib/oce_adv_tra_driver.F90:253-254). */
int tracer_set_schemes(int tracer_id, const char *hor, const char *ver, const char *lim,
                       double tra_adv_ph, double tra_adv_pv);

/* No reference counterpart: velocity in do_oce_adv_tra's layout,
 * uv(2, nl-1, elem2D), w and w_e (nl, nod2D) */
int tracer_set_velocity_elem(const double *uv, const double *w, const double *w_e, int nl,
                             int nod2D, int elem2D);

/* No reference counterpart: analytic divergence-free velocity (background
 * current + rotation cells + vertical shear) on the current mesh, w from
 * discrete continuity, w_e = w. Writes the dt giving the requested CFL. */
int tracer_make_velocity(double u_bg, double v_bg, double cell_amp, double shear, double cfl,
                         double *dt_out);

/* No reference counterpart: caller-supplied valuesAB (1-based id) */
int tracer_set_values_ab(int tracer_id, const double *values, int nl, int nod2D);

/* No reference counterpart: values of any tracer (1-based id) */
int tracer_get_values_id(int tracer_id, double *values, int nz, int nn);

/* No reference counterpart. Options:
 *   "precision"          64 (default), 32, 16 (FP32 with FP16 gradients)
 *   "refresh_valuesAB"   1 (default): valuesAB := values each step; 0: as set
 *   "accumulate_del_ttf" 0 (default): del_ttf is zeroed every step, as the
 *                        analytic driver does (src/fesom_analytic.F90:407-413);
 *                        1: it is carried over from step to step, as the
 *                        reference's tracer_advect_step does (it never zeroes
 *                        it, lib/tracer_c_interface.F90:194-212), one del_ttf
 *                        per tracer, starting from zero when switched on (and
 *                        when the device state is rebuilt, e.g. by a precision
 *                        switch); not with the LEARNED limiter
 *   "use_graph"          1 (default): replay multi-step runs from a hipGraph
 *   "gradient_fan"       1 (default): MUSCL's nodal mean gradients are formed
 *                        over each node's fan of elements (when every node
 *                        has one); 0: element by element (the same values to
 *                        rounding)
 *   "block_waves"        4 (default): waves per workgroup of the chunk kernels (1, 2, 4, 8)
 *   "batch_tracers"      4 (default): most tracers per launch (1, 2, 4)
 *   "force_rotation"     1 (default): tracer_load_mesh_from_files rotates the
 *                        node coordinates to the model grid; 0: as read
 *   "device_geometry"    0 (default); 1: tracer_generate_mesh and
 *                        tracer_load_mesh_from_files derive the geometry
 *                        (mesh_areas, mesh_auxiliary_arrays) with GPU
 *                        kernels; planar meshes come out bit for bit as
 *                        with 0 */
int tracer_set_option(const char *name, int value);

/* lib/oce_adv_tra_driver.F90:23-294 (do_oce_adv_tra), as its Fortran callers
 * use it: one advection of tracer tracer_id (1-based) on the GPU from its
 * current values and valuesAB, the velocity and the schemes set, whose
 * horizontal and vertical tendencies are ADDED into dttf_h and dttf_v
 * ((nl-1, nod2D) column-major, the wet layers of each column; the caller's
 * del_ttf_advhoriz / del_ttf_advvert), with the tracer left unchanged. The
 * work arrays of tracer_get_work_array then hold this call's fct_LO and
 * fluxes. Not with the LEARNED limiter. */
int tracer_do_oce_adv_tra(int tracer_id, double dt, double *dttf_h, double *dttf_v, int nl, int nod2D);

/* No reference counterpart: the allocation-free form of tracer_do_oce_adv_tra,
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). dttf_h and dttf_v are
 * DEVICE pointers to (nl-1, nod2D) column-major doubles, into whose wet
 * layers the tendencies are added on the GPU. The call allocates nothing: its
 * scratch (the two packed tendency arrays) lives in a file-scope device
 * array of 64 MiB that each GPU holds until the process ends, so it takes
 * meshes of at most 4,194,304 packed node layers (each column's wet layers
 * rounded up to even, columns packed whole into 128-layer chunks) and refuses
 * larger ones with -1 and an error; and the device state must exist already
 * (-1 otherwise). Every call on one GPU shares that scratch: calls that
 * overlap on one GPU race. The call runs on the library's own stream: work
 * other streams still have pending on dttf_h / dttf_v must be finished
 * first. Returns after the GPU work is done. */
int tracer_do_oce_adv_tra_noalloc(int tracer_id, double dt, double *dttf_h, double *dttf_v, int nl, int nod2D);

/* No reference counterpart: nsteps steps of ALL tracers */
int tracer_advect_steps(int nsteps, double dt);

/* No reference counterpart: out[7] = {min, max, sum over all values,
 * min and max over active cells, mass = sum T*hnode_new*areasvol, NaN count} */
int tracer_get_diagnostics(int tracer_id, double *out);

/* No reference counterpart: size (elements) and kind (0 int32, 1 float64)
 * of a named mesh array in the reference layout; -1 if unknown */
int tracer_mesh_array_info(const char *name, long long *count, int *kind);
/* No reference counterpart: copy a named mesh array out (0-based indices,
 * -1 for no element) */
int tracer_get_mesh_array(const char *name, void *out, long long count);

/* No reference counterpart (the reference's host fills its t_mesh in
 * Fortran, lib/MOD_MESH.F90): a mesh whose arrays the caller supplies, e.g.
 * a FESOM host handing over its own geometry. tracer_mesh_begin fixes the
 * dimensions; tracer_set_mesh_array copies one array in, by the names,
 * layouts and kinds of tracer_get_mesh_array (int32 or float64, 0-based
 * indices, -1 for "no element"; count must be the exact size); and
 * tracer_mesh_end validates and adopts the mesh, after which tracers are
 * allocated and velocities set as after tracer_generate_mesh.
 *   required: edges, edge_tri, elem2D_nodes, nlevels, nlevels_nod2D,
 *     nod_in_elem2D_num, nod_in_elem2D, zbar, elem_cos, elem_area,
 *     gradient_sca, edge_dxdy, edge_cross_dxdy, area, hnode, helem
 *   optional, with FESOM's derivation or default when absent:
 *     nlevels_nod2D_min, ulevels_nod2D_max (find_levels_min_e2n),
 *     nboundary_lay, ulevels / ulevels_nod2D (1), edge_up_dn_tri (-1: no
 *     up/down-wind element), areasvol (= area), hnode_new (= hnode),
 *     Z, Z_3d_n, zbar_3d_n (from zbar), depth, coord_nod2D (0)
 * tracer_set_mesh_array fails on an unknown name or a wrong count, and
 * tracer_mesh_end on a missing array, an index out of range, or an element
 * whose layers are not inside each of its nodes' columns, or a node with
 * more than 16 edges or elements; either failure leaves no mesh (begin
 * again). FCT and LEARNED need nod_in_elem2D to list, for every node, the
 * elements of one fan around it; tracer_set_schemes refuses them otherwise,
 * and the other schemes run. */
int tracer_mesh_begin(int nl, int nod2D, int elem2D, int edge2D, int maxAdj);
int tracer_set_mesh_array(const char *name, const void *data, long long count);
int tracer_mesh_end(void);

/* No reference counterpart: parse one derived-type restart file of the
 * given kind ("t_mesh", "t_partit", "t_tracer", "t_dynamics") and write it
 * back to out_path in canonical form (a format check; byte-identical for a
 * file FESOM's default build wrote). -1 with an error otherwise. */
int tracer_restart_file_copy(const char *kind, const char *in_path, const char *out_path);

/* No reference counterpart: a named GPU work array of the last step of one
 * tracer (1-based id) in the reference layout: fct_LO, adv_flux_hor (HO-LO
 * edge flux before limiting; the HO flux without a limiter), adv_flux_ver
 * (HO-LO vertical flux before limiting), fct_plus / fct_minus (R+/R-), gnod
 * (nodal mean gradient, (2,nl-1,nod2D)) */
int tracer_get_work_array(int tracer_id, const char *name, double *out, long long count);

/* No reference counterpart: error of the current precision against another
 * one (64, 32 or 16). A second device state at ref_precision starts from the
 * current tracers; both advance nsteps steps of all tracers in lockstep and
 * after every step out[4*(step*ntracers + t)..+3] receives, over the wet
 * layers of tracer t: max|a-b|, ||a-b||_2, max|a-b|/max|b|, ||a-b||_2/||b||_2
 * (a = this precision, b = the reference). The current state advances too. */
int tracer_precision_error(int ref_precision, int nsteps, double dt, double *out);

/* ---- Learned limiter (no reference counterpart). tracer_set_schemes(id,
 * hor, ver, "LEARNED", ...) selects it: a network 10-64-64-1 (ReLU, ReLU,
 * sigmoid; bf16 weights on the matrix cores, FP32 accumulation) predicts the
 * limiter coefficient of every horizontal edge level from its LO/HO fluxes,
 * node values, Zalesak headroom and cell geometry; the coefficient is applied
 * in flux form (mass exact); every edge of a node level whose bounds the
 * prediction would break takes the smaller of the network's and Zalesak's
 * coefficients, which keeps both of its nodes inside their bounds.
 * Parameters are FP32, flattened as W1[10][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3 (tracer_mlp_param_count() values). */
int tracer_mlp_param_count(void);
int tracer_mlp_set_params(const float *params, int n);
int tracer_mlp_get_params(float *params, int n);
/* nsteps FCT steps of all tracers, appending one training row (features and
 * Zalesak's coefficient) per edge level of every FCT/LEARNED tracer to the
 * on-device training set of capacity max_rows; *rows_out = rows held */
int tracer_learn_collect(int nsteps, double dt, long long max_rows, long long *rows_out);
int tracer_learn_clear(void);
/* Adam on the training set (minibatches of batch rows, seeded order);
 * losses[epochs] = mean squared error per epoch */
int tracer_learn_train(int epochs, long long batch, double lr, unsigned seed, double *losses);
/* the network on rows x 10 features (row-major floats) -> out[rows] */
int tracer_mlp_forward(const float *X, long long rows, float *out);
/* gradient of scale * sum((sigmoid(z) - y)^2) over the rows with respect to
 * the parameters (no update); rows with y < 0 are ignored */
int tracer_mlp_gradient(const float *X, const float *y, long long rows, float scale, float *grad, double *loss);
/* fraction of wet layers of a tracer where the last LEARNED step fell back
 * to Zalesak */
int tracer_learn_fallback(int tracer_id, double *fraction);

/* ---- Batched form (no reference counterpart): several independent periodic
 * meshes, each with its own tracers, advanced together by one hipGraph
 * launch per step (the meshes run as concurrent branches of the graph).
 * tracer_batch_add_periodic_mesh builds the mesh (bottom levels U{lev_lo..
 * lev_hi}, seeded), the analytic velocity at the given CFL (its dt in
 * *dt_out) and ntracers tracers: T, S, then passive Gaussian blobs, 0/1 steps
 * and white noise in turn, seeded; it returns the mesh index (>= 0) or -1. */
int tracer_batch_create(void);
int tracer_batch_destroy(void);
int tracer_batch_add_periodic_mesh(int nx, int ny, int nl, double Lx, double Ly, double dz, int lev_lo, int lev_hi,
                                   unsigned long long seed, int ntracers, int precision, const char *hor,
                                   const char *ver, const char *lim, double cfl, double *dt_out);
/* nsteps steps of every mesh with its own dt (use_graph: replay the captured step) */
int tracer_batch_run(int nsteps, int use_graph);
/* hipEvent time of nsteps graph launches, milliseconds */
int tracer_batch_time(int nsteps, double *ms_out);
/* every mesh's engine runs the autotuner (tracer_autotune) on reps steps of
 * its own and keeps its fastest schedule; table[2*i..] = {waves, tracers per
 * launch} of mesh i (up to max_rows meshes); returns the number of meshes */
int tracer_batch_autotune(int reps, double *table, int max_rows);
/* one launch schedule for every mesh (waves per workgroup 2, 4, 8; tracers
 * per launch 1, 2, 4) */
int tracer_batch_set_schedule(int waves, int max_nt);
int tracer_batch_mesh_info(int mesh, int *nl, int *nod2D, int *ntracers, long long *active_cells, double *dt);
int tracer_batch_get_values(int mesh, int tracer_id, double *values);
int tracer_batch_diagnostics(int mesh, int tracer_id, double *out);
/* a named array of batch mesh `mesh`, by the names, layouts and kinds of
 * tracer_get_mesh_array (mesh arrays, "uv", "w", "w_e"): *size (entries)
 * and *kind (0 int32, 1 float64) when not NULL, then, when out is not NULL,
 * the data (count >= *size) */
int tracer_batch_get_mesh_array(int mesh, const char *name, void *out, long long count, long long *size, int *kind);

/* No reference counterpart: the autotuner. Times every launch schedule of
 * the chunk kernels (waves per workgroup 2, 4, 8 x tracers per launch 4, 2,
 * 1) on reps graph-replayed steps of the current state, keeps the fastest
 * and restores the tracers. table[3*i..] = {waves, tracers per launch,
 * ms per step}; returns the number of rows (or -1). The schedule can also be
 * set with tracer_set_option("block_waves" | "batch_tracers", value). */
int tracer_autotune(double dt, int reps, double *table, int max_rows);

/* No reference counterpart: wall time of nsteps steps of all tracers on the
 * library's stream (hipEvent), milliseconds */
int tracer_time_steps(int nsteps, double dt, double *ms_out);

/* No reference counterpart: per-kernel mean time (ms) of one step, measured
 * with hipEvents on the library's stream; names written ';'-separated */
int tracer_time_kernels(double dt, int reps, char *names, int names_len, double *ms, int max_kernels);

/* ---- Output for the reference's plotting scripts (lib/mesh_output.F90:19-161,
 * read by visualization/plot_scalars.py). tracer_write_mesh writes
 * mesh_info.txt, coord_nod2D.bin (float64 (2,nod2D), metres), elem2D_nodes.bin
 * (int32 (3,elem2D), 0-based), zbar.bin and depth.bin (float64) into
 * output_dir (write_mesh_for_python :19-88). tracer_output_open writes
 * scalar_info.txt (nod2D, nz, nsteps, num_tracers, wp=8, tracer names:
 * temperature, salinity, tracer_3, ...) and opens one <name>.bin per tracer
 * (open_scalar_output :92-128); every tracer_output_step appends one float64
 * (nz,nod2D) frame per tracer (write_scalar_step :131-146); the reader expects
 * nsteps+1 frames (the initial state and one per step). */
int tracer_write_mesh(const char *output_dir);
int tracer_output_open(const char *output_dir, int nsteps);
int tracer_output_step(void);
int tracer_output_close(void);

/* No reference counterpart: the practical HBM ceiling. Times reps copies of
 * a `bytes` device buffer with a hand-written 16-byte streaming kernel over a
 * few launch shapes; *gbs = best (read + write bytes) / time in GB/s, and
 * variant[3] (may be NULL) = {workgroups, 16-byte words in flight per lane,
 * nontemporal hint} of the best. Allocates and frees two buffers. */
int tracer_hbm_copy_rate(long long bytes, int reps, double *gbs, int *variant);

/* No reference counterpart: the last error message */
const char *tracer_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
