/*
 * emfem_amd: MI355X-native implementation of EMFEM's forward-modelling hot path.
 *
 * C ABI.  Plain pointers and sizes only; every call returns an int status
 * (0 on success, negative on failure) like the reference's PetscErrorCode
 * functions, and the message of the last failure is available from
 * emfem_last_error().  One context owns one HIP stream and all device buffers;
 * calls on a context are synchronous at return and a context must be used by
 * one host thread at a time (the reference's PETSc objects carry the same rule).
 *
 * Indexing is zero-based int32 (PetscInt default, TetGen `z` switch,
 * src/em_mesh.cc:79-81).  Complex vectors cross the boundary as split real and
 * imaginary arrays (PETScBlockVector {re, im}, src/em_defs.h:33-35) on the host,
 * or as interleaved (re, im) pairs per edge and right-hand side on the device
 * (row-major [n_edges][2 * n_rhs]).
 *
 * Each declaration names the reference interface it replaces (file:line,
 * relative to the reference checkout).
 */
#ifndef EMFEM_AMD_H
#define EMFEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  EMFEM_ERR_USER matches EM_ERR_USER (src/em_defs.h:76). */
#define EMFEM_OK 0
#define EMFEM_ERR_USER (-1)
#define EMFEM_ERR_HIP (-2)
#define EMFEM_ERR_NOT_CONVERGED (-3)
#define EMFEM_ERR_STATE (-4)
#define EMFEM_ERR_ARG (-5)

/* Polarisation codes (PolarType, src/em_defs.h:39) and the MT transmitter index
 * (src/em_fwd.cc:766). */
#define EMFEM_XY_POLAR (-1)
#define EMFEM_YX_POLAR (-2)
#define EMFEM_MT_TIDX (-3)

typedef struct emfem_context emfem_context;

/* ---- context and options --------------------------------------------------- */

/* create_context (src/em_ctx.cc:7-84): allocates the context and its HIP stream on
 * the current device. */
int emfem_create_context(emfem_context **ctx);

/* destroy_context (src/em_ctx.cc:86-98). */
int emfem_destroy_context(emfem_context *ctx);

/* Message of the last failed call on ctx (SETERRQ text in the reference). */
const char *emfem_last_error(const emfem_context *ctx);

/* process_options (src/em_ctx.cc:100-175): parses PETSc options-file text
 * ("-key value" tokens, '#' comments) for the EMFEM keys and the A_/B_ solver
 * keys this implementation honours; unknown keys are kept and reported by
 * emfem_unused_options.  May be called more than once; later values win. */
int emfem_process_options(emfem_context *ctx, const char *options_text);

/* Current value of an option as text (after defaults); returns EMFEM_ERR_ARG for
 * an unknown key. */
int emfem_get_option(const emfem_context *ctx, const char *key, char *buf, int buflen);

/* Space-separated list of option keys that were given but are not used, or whose
 * given value is not honoured: -inner_pc_type direct / mixed with -pc_threshold and
 * -direct_solver_type select the reference's direct inner solver
 * (src/em_la.cc:147-155), which does not exist here; the inner solve then stays
 * FCG + AMS (the default, -inner_pc_type ams). */
int emfem_unused_options(const emfem_context *ctx, char *buf, int buflen);

/* ---- mesh, model and survey ------------------------------------------------- */

/* read_mesh + Mesh::create (src/em_io.cc:155-200, src/em_mesh.cc:56-85): points
 * [n_points*3], tets [n_tets*4] (det([1; x_v]) > 0), attributes [n_tets] (row of
 * the resistivity table).  Derives edges (ascending orientation,
 * src/em_mesh.cc:448-455), face neighbours (src/em_mesh.cc:463-481) and boundary
 * markers (src/em_mesh.cc:483-502), and uploads the mesh to the device.  Edges are
 * numbered in lexicographic order of (lower vertex, higher vertex) unless
 * emfem_set_mesh_edges supplies them. */
int emfem_set_mesh(emfem_context *ctx, int n_points, const double *points, int n_tets, const int *tets,
                   const int *attributes);

/* Supplies TetGen's edge list and tet->edge map (tetgenio edgelist /
 * tet2edgelist, src/em_mesh.cc:448-461) so that edge numbering matches a TetGen
 * run exactly.  Call after emfem_set_mesh. */
int emfem_set_mesh_edges(emfem_context *ctx, int n_edges, const int *edgelist, const int *tet2edge);

/* read_rho (src/em_io.cc:9-86): aniso_form 0..3 with {1,2,3,6} columns per row. */
int emfem_set_rho(emfem_context *ctx, int n_rho, int aniso_form, const double *rho);

/* read_emd parts (src/em_io.cc:202-235): frequencies (Hz), transmitters (7 numbers
 * each: x y z azimuth dip current length, src/em_defs.h:25) and receivers. */
int emfem_set_frequencies(emfem_context *ctx, int n, const double *freqs);
int emfem_set_transmitters(emfem_context *ctx, int n, const double *tx);
int emfem_set_receivers(emfem_context *ctx, int n, const double *rx);

/* info[0..4] = n_vertices, n_tets, n_edges, nnz (0 before create_linear_system),
 * n_colours of the atomic-free assembly schedule. */
int emfem_get_mesh_info(const emfem_context *ctx, int *info);

/* Topology as the context holds it: edgelist [n_edges*2], tet2edge [n_tets*6],
 * neighbor [n_tets*4], edge_bdr [n_edges].  Any pointer may be NULL. */
int emfem_get_topology(const emfem_context *ctx, int *edgelist, int *tet2edge, int *neighbor,
                       unsigned char *edge_bdr);

/* Mesh::find_cell_around_point (src/em_mesh.cc:527-566): tet containing p, or -1
 * (host search). */
int emfem_find_cell(const emfem_context *ctx, const double *p);

/* Mesh::find_closest_vertex + find_cell_around_point (src/em_mesh.cc:518-566) for n
 * points [n*3] at once, on the device: tets[i] = containing tet or -1.  The
 * breadth-first search per point is bounded at 4096 queued tets; a point that
 * reaches the bound fails the call with EMFEM_ERR_USER.  The receivers and dipoles of
 * the right-hand sides and responses are located this way. */
int emfem_locate_points(emfem_context *ctx, int n, const double *points, int *tets);

/* update_background_model (src/em_fwd.cc:135-159): 1-D layered columns under the
 * four top corners, used by the MT boundary data. */
int emfem_update_background_model(emfem_context *ctx);

/* ---- linear system ---------------------------------------------------------- */

/* create_linear_system (src/em_fwd.cc:248-295): builds the CSR pattern
 * (src/em_fwd.cc:191-246) on the host, uploads it, and allocates n_slots
 * right-hand-side and solution columns on the device. */
int emfem_create_linear_system(emfem_context *ctx, int n_slots);

/* destroy_linear_system (src/em_fwd.cc:297-316). */
int emfem_destroy_linear_system(emfem_context *ctx);

/* assemble_matrix (src/em_fwd.cc:318-400): C and M on the device for frequency fidx.
 * The preconditioner of the previous matrix is no longer applied (a solve needs
 * emfem_create_pc again); with -pc_reuse true its device state is kept for that call
 * to refresh. */
int emfem_assemble_matrix(emfem_context *ctx, int fidx);

/* Copies the assembled CSR to the host: rowptr [n_edges+1], col [nnz], C [nnz],
 * M [nnz].  Any pointer may be NULL. */
int emfem_get_matrix(const emfem_context *ctx, int *rowptr, int *col, double *C, double *M);

/* assemble_rhs_mt (src/em_fwd.cc:488-580) for polarisation mode (-1 XY, -2 YX). */
int emfem_assemble_rhs_mt(emfem_context *ctx, int fidx, int mode, int slot);

/* assemble_rhs_csem (src/em_fwd.cc:616-671). */
int emfem_assemble_rhs_csem(emfem_context *ctx, int fidx, int tidx, int slot);

/* assemble_rhs_dual (src/em_fwd.cc:673-717). */
int emfem_assemble_rhs_dual(emfem_context *ctx, int slot);

/* Host copies of a right-hand side / solution slot (s.re, s.im / e.re, e.im). */
int emfem_set_rhs(emfem_context *ctx, int slot, const double *re, const double *im);
int emfem_get_rhs(const emfem_context *ctx, int slot, double *re, double *im);
int emfem_get_solution(const emfem_context *ctx, int slot, double *re, double *im);

/* create_pc (src/em_la.cc:113-191): B = C + M, the auxiliary-space (AMS)
 * preconditioner for B and the outer FGMRES set-up.  Precision of the inner solve
 * follows the option -inner_precision (fp32 | fp64).  On a linear system whose
 * previous preconditioner was built with the same set-up options, its device state
 * is refreshed with the new values rather than built again (-pc_reuse, stat
 * "pc_reused"). */
int emfem_create_pc(emfem_context *ctx);

/* Starts the set-up for frequency fidx ahead of time, for drivers that loop over
 * frequencies: the matrix of fidx is assembled on a second stream of the context,
 * the auxiliary-space Galerkin products are formed there and the host AMG set-up is
 * built from them on a worker thread, while the caller goes on (typically with the
 * solve of the current frequency).  A later emfem_assemble_matrix(fidx) takes the
 * prepared matrix and the emfem_create_pc after it the prepared host set-up (stats
 * "assemble_prepared", "pc_prepared"); the results are those of the same calls
 * without it.  A no-op before the first emfem_create_pc of the linear system and
 * with -ams_setup host; discarded by the assembly of another frequency and by any
 * change of the mesh, conductivities, frequencies or linear system.  No counterpart
 * in the reference, which sets each frequency up in turn (src/em_fwd.cc:1341-1360). */
int emfem_prepare_pc(emfem_context *ctx, int fidx);

/* destroy_pc (src/em_la.cc:193-206). */
int emfem_destroy_pc(emfem_context *ctx);

/* solve_linear_system (src/em_la.cc:35-63), batched: solves slots
 * [first, first+n) together, each to its own relative tolerance rtols[i] (or
 * rtol for all when rtols is NULL), zero initial guess, at most max_it outer
 * iterations.  The stopping test is the reference's (FGMRES's implicit residual
 * <= rtol ||s||); when it passes but the true residual does not, the solve restarts
 * from the true residual until that stops halving, and then accepts the right-hand
 * side only if its true residual is within -true_residual_slack (10) x rtol.
 * iters[i] and relres[i] (the true FP64 residual ||s - A e|| / ||s||) are written
 * when the pointers are non-NULL.  Returns EMFEM_ERR_NOT_CONVERGED (after storing
 * the partial solution, with the worst true residual in the error message) if any
 * right-hand side missed these tests within max_it, which the reference never
 * reports.  Device workspace is kept by the context across calls (released by
 * emfem_destroy_linear_system), so repeated solves do not allocate. */
int emfem_solve_linear_system(emfem_context *ctx, int first, int n, int max_it, double rtol,
                              const double *rtols, int *iters, double *relres);

/* Multi-model batch: solve_linear_system (src/em_la.cc:35-63) for slots
 * [first, first+n) of n_ctx contexts at once.  The reference runs its models one at a
 * time or in separate MPI task groups (src/em_ctx.cc:17-43); here the contexts'
 * assembled systems (all on this device) are stacked block-diagonally, the
 * auxiliary-space set-up is built for the stack, and one FGMRES runs every launch over
 * all models while each (model, slot) keeps its own Arnoldi process, inner FCG and
 * stopping test at rtol.  Solutions are written to each context's slots.  Outputs
 * (any may be NULL) are indexed k * n + j for model k, slot first + j: outer
 * iterations, true FP64 relative residual, and the wall time (ms) from the start of
 * the call (stack set-up included) until that right-hand side was accepted.  With
 * -batch_retire true (the default) the solve runs in stages: each stops once half of
 * the right-hand sides still open have passed, models whose right-hand sides have all
 * passed retire, and the rest are restacked and continue from their iterates; at 8
 * models or fewer one stage runs to completion.  -batch_retire false runs one
 * lock-step solve.
 * Options (precision, restart, cycle, batch_retire) are taken from ctxs[0].  Returns
 * EMFEM_ERR_NOT_CONVERGED like emfem_solve_linear_system. */
int emfem_solve_batch(emfem_context **ctxs, int n_ctx, int first, int n, int max_it, double rtol, int *iters,
                      double *relres, double *solve_ms);

/* matshell_mult_a (src/em_la.cc:65-87) on device arrays: y = A x for n_rhs
 * interleaved right-hand sides ([n_edges][2*n_rhs]).  Also MATOP_MULT_TRANSPOSE
 * (A is symmetric, src/em_la.cc:169). */
int emfem_matshell_mult_a(emfem_context *ctx, int n_rhs, const double *x_dev, double *y_dev);

/* pc_apply_b (src/em_la.cc:208-228) on device arrays: x = diag(B,B)^-1 b by the
 * inner Krylov solve (FCG + AMS) for n_rhs interleaved right-hand sides. */
int emfem_pc_apply_b(emfem_context *ctx, int n_rhs, const double *b_dev, double *x_dev);

/* Vector layouts of emfem_create_vecs. */
#define EMFEM_LAYOUT_INTERLEAVED 0
#define EMFEM_LAYOUT_SPLIT 1

/* matshell_createvecs_a (MATOP_CREATE_VECS, src/em_la.cc:89-111, registered at
 * src/em_la.cc:170): zeroed device vectors for n_rhs right-hand sides of the
 * operator.  EMFEM_LAYOUT_SPLIT gives the VecNest {re, im} halves the reference
 * creates: *x = re and *x_im = im, n_rhs * n_edges doubles each (vector j at offset
 * j * n_edges).  EMFEM_LAYOUT_INTERLEAVED gives one [n_edges][2 * n_rhs] array in
 * *x (x_im may be NULL; it is set to NULL).  Free with emfem_device_free. */
int emfem_create_vecs(emfem_context *ctx, int n_rhs, int layout, double **x, double **x_im);

/* matshell_mult_a (src/em_la.cc:65-87) on the VecNest halves: y = A x with
 * x = {x_re, x_im} and y = {y_re, y_im} split device arrays (EMFEM_LAYOUT_SPLIT),
 * n_rhs vectors each.  In place (x == y) is allowed. */
int emfem_matshell_mult_a_split(emfem_context *ctx, int n_rhs, const double *x_re, const double *x_im,
                                double *y_re, double *y_im);

/* No counterpart in the reference: emfem_matshell_mult_a_split for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result; the interleaved
 * staging lives in a fixed device array of the library (64 MiB per GPU, held until
 * the process ends) instead of context buffers, so the call allocates and frees
 * nothing.  n_edges * n_rhs must not exceed 2097152 (EMFEM_ERR_ARG otherwise).  All
 * calls on one GPU share that array: calls that overlap in time on one GPU (from
 * several host threads or streams) race. */
int emfem_matshell_mult_a_split_noalloc(emfem_context *ctx, int n_rhs, const double *x_re, const double *x_im,
                                        double *y_re, double *y_im);

/* pc_apply_b (src/em_la.cc:208-228) on the VecNest halves: {x_re, x_im} =
 * B^-1 {b_re, b_im} (split device arrays, n_rhs vectors each). */
int emfem_pc_apply_b_split(emfem_context *ctx, int n_rhs, const double *b_re, const double *b_im, double *x_re,
                           double *x_im);

/* ---- estimator, responses, marking ----------------------------------------- */

/* estimate_error (src/em_fwd.cc:1102-1170): squared indicators per tet for each of
 * the n solution slots; errors is [n][n_tets] on the host. */
int emfem_estimate_error(emfem_context *ctx, int fidx, int n, const int *slots, double *errors);

/* interpolate_fields (src/em_fwd.cc:719-744) at n points: E and H as
 * [n][3][2] (re, im) arrays.  The containing tet is found by
 * emfem_find_cell's rule. */
int emfem_interpolate_fields(emfem_context *ctx, int fidx, int slot, int n, const double *points,
                             double *E, double *H);

/* calculate_response_mt (src/em_fwd.cc:746-802): obs is [n_obs][4] =
 * (otype, fidx, tidx, ridx); rsp [n_obs][2] is written for rows with this fidx and
 * tidx -3.  Apparent resistivity and phase (codes 212, 222) are answered as
 * log(rho_a) + i*phi, which the reference leaves unimplemented. */
int emfem_calculate_response_mt(emfem_context *ctx, int fidx, int slot_xy, int slot_yx, int n_obs,
                                const int *obs, double *rsp);

/* calculate_response_csem (src/em_fwd.cc:804-871). */
int emfem_calculate_response_csem(emfem_context *ctx, int fidx, int tidx, int slot, int n_obs,
                                  const int *obs, double *rsp);

/* refine_fixed_fraction / refine_fixed_number (src/em_fwd.cc:1182-1240); flags are
 * OR-ed into flag[n]. */
int emfem_refine_fixed_fraction(int n, const double *error, double fraction, unsigned char *flag);
int emfem_refine_fixed_number(int n, const double *error, double fraction, unsigned char *flag);

/* ---- inputs, outputs, refinement and the forward driver -------------------- */

/* read_mesh (src/em_io.cc:155-200) with read_rho (src/em_io.cc:9-86): the .mesh
 * text format (points, tets with attribute, resistivity block); '#' starts a
 * comment (parse_string, src/em_utils.cc:34-58).  Also keeps the mesh as the
 * coarse mesh that emfem_forward restarts every task from (src/em_fwd.cc:1330). */
int emfem_read_mesh(emfem_context *ctx, const char *path);

/* read_emd (src/em_io.cc:202-287): frequencies, transmitters, receivers and the
 * observation list of an .emd file. */
int emfem_read_emd(emfem_context *ctx, const char *path);

/* The observation part of read_emd: obs [n][4] = (otype, fidx, tidx, ridx) and
 * values [n][4] = (re, im, err_re, err_im) as in the file (amplitude/phase codes
 * are stored as log amplitude + i phase, src/em_io.cc:249-261); values may be
 * NULL. */
int emfem_set_observations(emfem_context *ctx, int n, const int *obs, const double *values);

/* save_rsp (src/em_io.cc:316-399): writes the .rsp response file. */
int emfem_save_rsp(emfem_context *ctx, const char *path);

/* Responses of the last emfem_forward as [n_obs][2] (re, im). */
int emfem_get_responses(const emfem_context *ctx, double *rsp);

/* save_mesh (src/em_io.cc:289-314) with Mesh::save_vtk (src/em_mesh.cc:184-248):
 * legacy ASCII VTK of the current mesh with the subdomain, attribute, resistivity
 * columns and error indicators.  tidx < 0 (MT): error0 = error_xy and error1 =
 * error_yx; tidx >= 0 (CSEM): error0 = error, error1 unused.  Indicator arrays hold
 * one value per tet.  emfem_forward writes these per cycle when -save_mesh is true
 * and -oprefix is set. */
int emfem_save_mesh(emfem_context *ctx, const char *path, int tidx, const double *error0,
                    const double *error1);

/* em_forward (src/em_fwd.cc:1441-1471) after the inputs are loaded: background
 * model, task distribution (src/em_fwd.cc:1273-1318), forward_mt / forward_csem
 * (src/em_fwd.cc:1320-1439) with receiver/transmitter-area refinement, uniform
 * refinement and goal-oriented adaptive cycles.  CSEM transmitters of one
 * frequency that share a mesh are solved as one block of right-hand sides. */
int emfem_forward(emfem_context *ctx);

/* refine_mesh's mesh update (src/em_fwd.cc:1268 -> Mesh::refine_tetgen,
 * src/em_mesh.cc:101-139): every flagged tet is halved by conforming longest-edge
 * bisection (TetGen is not used); topology is rebuilt. */
int emfem_refine_mesh(emfem_context *ctx, const unsigned char *flags);

/* Mesh::refine_uniform (src/em_mesh.cc:141-182): 1 -> 8 split of every tet. */
int emfem_refine_uniform(emfem_context *ctx);

/* The current mesh: points [n_vertices*3], tets [n_tets*4], attributes [n_tets]. */
int emfem_get_mesh(const emfem_context *ctx, double *points, int *tets, int *attributes);

/* ---- learned indicator surrogate (no counterpart in the reference) ----------
 * A per-element message-passing network over the face-neighbour graph (input layer
 * of 16 padded features, 3 message-passing layers of width 64, bf16 matrix-core
 * GEMMs with fp32 accumulation) that predicts, from per-tet features of the mesh,
 * the conductivity, the frequency, the receivers and the primal XY / YX solutions, the
 * dual solution's squared indicator eta2_dual of estimate_error
 * (src/em_fwd.cc:1102-1180).  emfem_forward uses it, when -surrogate_cycles is true, in
 * the intermediate adaptive MT cycles in place of the dual solve: the indicators
 * sqrt(eta2_xy * eta2_dual) (src/em_fwd.cc:1172-1177) take the exact primal factors
 * and the predicted dual.  The last cycle always solves the dual, compares the
 * indicator with the predicted and with the solved dual by Spearman rank correlation
 * (stat "surrogate_rank_corr") and rejects the surrogate for the rest of the
 * context's life below -surrogate_min_corr (0.9; stat "surrogate_rejected"). */

/* New network with Glorot-uniform weights from seed. */
int emfem_surrogate_init(emfem_context *ctx, unsigned seed);

/* *n = parameter count (network weights, then 16 feature means, 16 feature standard
 * deviations, target mean and standard deviation of log10 indicator); copies them out
 * to get and/or in from set (either may be NULL). */
int emfem_surrogate_params(emfem_context *ctx, float *get, const float *set, int *n);

/* Raw features [n_tets][16] of the current mesh for the XY / YX solution slots. */
int emfem_surrogate_features(emfem_context *ctx, int fidx, int slot_xy, int slot_yx, float *feats);

/* Full-batch Adam training on n elements: raw features [n][16], face neighbours
 * [n][4] (-1 for none; symmetric), target eta2_dual [n] (> 0; fitted in log10 after
 * standardization).  loss[e] (may be NULL) receives the mean squared error of epoch e. */
int emfem_surrogate_train(emfem_context *ctx, int n, const float *feats, const int *neighbors,
                          const double *indicator, int epochs, double lr, double *loss);

/* Predicted eta2_dual [n_tets] for the current mesh and solution slots. */
int emfem_surrogate_predict(emfem_context *ctx, int fidx, int slot_xy, int slot_yx, double *indicator);

/* The raw network on n elements of standardized features [n][16] and neighbours
 * [n][4]: outputs y [n], and with target [n] the mean squared error and its gradient
 * with respect to the network weights (any output may be NULL). */
int emfem_surrogate_eval(emfem_context *ctx, int n, const float *feats, const int *neighbors, const float *target,
                         float *y, double *loss, float *grad);

/* ---- device buffers -------------------------------------------------------- */

/* Device memory on the context's GPU for callers that have no GPU framework of
 * their own (the device-pointer entry points above take such buffers).  No
 * reference counterpart: PETSc would hand its own device vectors. */
int emfem_device_alloc(emfem_context *ctx, unsigned long long bytes, void **ptr);
int emfem_device_free(emfem_context *ctx, void *ptr);
int emfem_copy_to_device(emfem_context *ctx, void *dst, const void *src, unsigned long long bytes);
int emfem_copy_to_host(emfem_context *ctx, void *dst, const void *src, unsigned long long bytes);

/* ---- measurement ---------------------------------------------------------- */

/* Named counters and timers of the last calls (e.g. "assemble_ms", "solve_ms",
 * "outer_its", "inner_its", "vcycles"; after emfem_forward "forward_ms" and the
 * driver's phases "fwd_system_ms", "fwd_assemble_ms", "fwd_pc_ms", "fwd_solve_ms",
 * "fwd_estimate_ms", "fwd_refine_ms", "fwd_responses_ms", like the reference's
 * log_view events, src/em_ctx.cc:70-81; after emfem_create_pc "spmm_col16", 1 when
 * B's row-block SpMMs read 16-bit columns); returns EMFEM_ERR_ARG for an unknown
 * name. */
int emfem_get_stat(const emfem_context *ctx, const char *name, double *value);

/* Device time (ms per launch) and algorithmic bytes of the hot kernels:
 * which = "assemble" (all colours of the last assembly), "estimate", and, after a
 * solve run with -use_graph false, "b_spmm" (the smoother B SpMM in the cycle
 * precision, averaged over HIP events around each of its first 512 launches in that
 * solve's cycles), "op_spmm" (complex operator) and "fcg_spmm" (fp64 B SpMM of the
 * FCG), the last two from HIP events around 20 back-to-back launches on that solve's
 * operands; all on the context's stream.  The bytes of b_spmm and fcg_spmm count the
 * column index at 2 + 4 / 64 bytes per nonzero where the 16-bit form is read. */
int emfem_kernel_time(const emfem_context *ctx, const char *which, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif

#endif
