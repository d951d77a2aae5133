/*
 * tea_leaf_engine.h -- C-ABI of the TeaLeaf solver engine for MI355X (gfx950).
 *
 * The first block declares the 29 entry points that the TeaLeaf Fortran driver
 * calls through implicit interfaces (lower-case name + trailing underscore,
 * every argument by reference: INTEGER = int32, REAL(KIND=8) = double).  Each
 * one replaces the CUDA definition cited beside it, with the same argument
 * meaning.  The library owns every field array (single chunk, one GPU); the
 * caller owns only scalars, coefficient arrays, state tables and host halo
 * buffers.  Calls that return a scalar are synchronous; the others may return
 * before the GPU work finishes.  Fatal errors print a message and abort, as
 * the reference's DIE/MPI_Abort does, unless tea_engine_set_error_mode(1) asked
 * for a sticky error instead (see the extension block).
 *
 * The kernel side reads its own settings from the input deck (file "tea.in"
 * in the working directory unless tea_engine_set_deck_path() names another):
 * tl_use_ppcg / tl_use_chebyshev / tl_use_cg / tl_use_jacobi (priority in
 * that order), tl_preconditioner_type=none|jac_diag|jac_block, halo_depth,
 * profiler_on, cuda_device / hip_device, and the mixed-precision key
 * tl_inner_precision=fp32 (init_cuda.cu:63-163).
 *
 * The second block holds extensions that have no counterpart in the
 * reference: field access for tests and the warm start, the profiler, error
 * mode and the mixed-precision inner solver.
 */
#ifndef TEA_LEAF_ENGINE_H
#define TEA_LEAF_ENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants shared with the Fortran driver (data.f90:37-75) ---------- */
enum {
    TEA_CHUNK_LEFT = 1, TEA_CHUNK_RIGHT = 2, TEA_CHUNK_BOTTOM = 3, TEA_CHUNK_TOP = 4,
    TEA_EXTERNAL_FACE = -1
};
enum {
    TEA_FIELD_DENSITY = 1, TEA_FIELD_ENERGY0 = 2, TEA_FIELD_ENERGY1 = 3,
    TEA_FIELD_U = 4, TEA_FIELD_P = 5, TEA_FIELD_SD = 6, TEA_FIELD_R = 7,
    TEA_NUM_FIELDS = 7
};
enum { TEA_G_RECT = 1, TEA_G_CIRC = 2, TEA_G_POINT = 3 };
enum { TEA_CONDUCTIVITY = 1, TEA_RECIP_CONDUCTIVITY = 2 };
enum { TEA_PREC_NONE = 1, TEA_PREC_JAC_DIAG = 2, TEA_PREC_JAC_BLOCK = 3 };
enum { TEA_SOLVER_JACOBI = 0, TEA_SOLVER_CG = 1, TEA_SOLVER_CHEBYSHEV = 2, TEA_SOLVER_PPCG = 3 };

/* ---- setup (replaces init_cuda.cu and the *_kernel_cuda.cu set-up files) - */

/* init_cuda.cu:35-43.  Creates the chunk for cells [x_min..x_max] x [y_min..y_max]
 * (1-based, inclusive), reads the deck, selects the device, allocates and
 * zero-fills every field (halo cells included). */
void initialise_cuda_(int *x_min, int *x_max, int *y_min, int *y_max);

/* initialise_chunk_kernel_cuda.cu:30-34.  Vertex/cell coordinates, volume, areas. */
void initialise_chunk_kernel_cuda_(double *xmin, double *ymin, double *dx, double *dy);

/* generate_chunk_kernel_cuda.cu:30-47.  Paints state 1 everywhere, then states
 * 2..n in order (later states overwrite), then u = u0 = energy0*density. */
void generate_chunk_kernel_cuda_(const int *number_of_states,
                                 const double *state_density, const double *state_energy,
                                 const double *state_xmin, const double *state_xmax,
                                 const double *state_ymin, const double *state_ymax,
                                 const double *state_radius, const int *state_geometry,
                                 const int *g_rect, const int *g_circ, const int *g_point);

/* set_field_kernel_cuda.cu:4-8.  energy1 = energy0. */
void set_field_kernel_cuda_(void);

/* ---- summary and halos --------------------------------------------------- */

/* field_summary_kernel_cuda.cu:30-34.  Interior sums of volume, volume*density,
 * volume*density*energy1 and volume*density*u. */
void field_summary_kernel_cuda_(double *vol, double *mass, double *ie, double *temp);

/* update_halo_kernel_cuda.cu:31-37.  Reflective update of the selected fields
 * (fields[7], 1 = selected) on every face whose neighbour is EXTERNAL_FACE, to
 * the given depth: the halo cell at distance k from a face takes the value of
 * the interior cell at distance k-1.  Bottom and top first, then left and
 * right, so corners are reflected twice. */
void update_halo_kernel_cuda_(const int *chunk_neighbours, const int *fields, const int *depth);

/* pack_kernel_cuda.cu:7-19.  Gather (pack) or scatter (unpack) the selected
 * fields' face strips of the given depth into/out of host_buffer at element
 * offsets[f]; face = TEA_CHUNK_*.  Buffer layout is pack_kernel.cuknl's: left and
 * bottom strips are stored mirrored, right and top strips in natural order. */
void cuda_pack_buffers_(int *fields, int *offsets, int *depth, int *face, double *host_buffer);
void cuda_unpack_buffers_(int *fields, int *offsets, int *depth, int *face, double *host_buffer);

/* ---- solver: common ------------------------------------------------------ */

/* tea_leaf_kernel_cuda.cu:213-219,230-252.  rx = dt/dx^2, ry = dt/dy^2 (outputs);
 * Kx, Ky face coefficients from density (coefficient 1: w = density,
 * 2: w = 1/density); zero Kx/Ky on external faces flagged in zero_boundary[4]
 * (left,right,bottom,top) unless reflective; then u = u0 = energy1*density. */
void tea_leaf_common_init_kernel_cuda_(const int *coefficient, double *dt, double *rx, double *ry,
                                       int *zero_boundary, int *reflective_boundary);

/* tea_leaf_kernel_cuda.cu:227-229,286-297.  r = u0 - A u. */
void tea_leaf_calc_residual_cuda_(void);

/* tea_leaf_kernel_cuda.cu:221-225,279-284.  energy1 = u/density. */
void tea_leaf_common_finalise_kernel_cuda_(void);

/* tea_leaf_kernel_cuda.cu:34-38,56-87.  norm_array 0: u0.u0, 1: r.r,
 * 2: r.z (r.r without a preconditioner). */
void tea_leaf_calc_2norm_kernel_cuda_(int *norm_array, double *norm);

/* ---- solver: CG ---------------------------------------------------------- */

/* tea_leaf_kernel_cuda.cu:139-143,164-186.  Preconditioner set-up, z = M^-1 r,
 * p = z (or r); returns rro = r.p. */
void tea_leaf_cg_init_kernel_cuda_(double *rro);
/* tea_leaf_kernel_cuda.cu:145-149,188-195.  w = A p; returns pw = p.w. */
void tea_leaf_cg_calc_w_kernel_cuda_(double *pw);
/* tea_leaf_kernel_cuda.cu:150-154,197-205.  u += alpha p; r -= alpha w;
 * z = M^-1 r; returns rrn = r.z (r.r without a preconditioner). */
void tea_leaf_cg_calc_ur_kernel_cuda_(double *alpha, double *rrn);
/* tea_leaf_kernel_cuda.cu:155-159,207-211.  p = z + beta p (r + beta p). */
void tea_leaf_cg_calc_p_kernel_cuda_(double *beta);

/* ---- solver: Chebyshev --------------------------------------------------- */

/* tea_leaf_kernel_cuda.cu:40-47,105-123.  Uploads n_coefs (alpha, beta) pairs;
 * w = A u, r = u0 - w, p = M^-1 r / theta, u += p. */
void tea_leaf_cheby_init_kernel_cuda_(const double *ch_alphas, const double *ch_betas,
                                      int *n_coefs, const double *theta);
/* tea_leaf_kernel_cuda.cu:49-53,125-135.  With k = cheby_calc_step-1:
 * w = A u, r = u0 - w, p = alpha_k p + beta_k M^-1 r, u += p. */
void tea_leaf_cheby_iterate_kernel_cuda_(const int *cheby_calc_step);

/* ---- solver: Jacobi ------------------------------------------------------ */

/* tea_leaf_kernel_cuda.cu:215-228.  One Jacobi sweep; returns sum |u - u_old|. */
void tea_leaf_jacobi_solve_kernel_cuda_(double *error);

/* ---- solver: PPCG -------------------------------------------------------- */

/* tea_leaf_kernel_cuda.cu:303-308,357-362.  Uploads the inner-step coefficients. */
void tea_leaf_ppcg_init_constants_cuda_(const double *ch_alphas, const double *ch_betas,
                                        int *n_inner_steps);
/* tea_leaf_kernel_cuda.cu:310-315,366-375.  step 1: p = z = 0, z = M^-1 r,
 * p = z (or r), rro = r.p.  step 2: z = M^-1 r, rro = 0.  step 3: p = z, rro = r.p. */
void tea_leaf_ppcg_init_kernel_cuda_(const int *step, double *rro);
/* tea_leaf_kernel_cuda.cu:316-320,377-383.  sd = M^-1 r / theta (the reference
 * driver never calls it; kept for link compatibility). */
void tea_leaf_ppcg_init_sd_kernel_cuda_(const double *theta);
/* tea_leaf_kernel_cuda.cu:322-327,386-391.  sd = (z or r)/theta, rtemp = r, utemp = sd. */
void tea_leaf_ppcg_init_sd_new_kernel_cuda_(const double *theta);
/* tea_leaf_kernel_cuda.cu:348-354,418-486.  One inner step k = ppcg_cur_step-1:
 * rtemp -= A sd; sd = alpha_k sd + beta_k M^-1 rtemp; utemp += sd, over the
 * matrix-powers extent given by bounds_extra and chunk_neighbours[4]. */
void tea_leaf_ppcg_inner_kernel_cuda_(int *ppcg_cur_step, int *bounds_extra, int *chunk_neighbours);
/* tea_leaf_kernel_cuda.cu:331-334,395-399.  r_store = r. */
void tea_leaf_ppcg_store_r_kernel_cuda_(void);
/* tea_leaf_kernel_cuda.cu:337-340,402-406.  z = utemp. */
void tea_leaf_ppcg_update_z_kernel_cuda_(void);
/* tea_leaf_kernel_cuda.cu:343-346,409-416.  returns (r - r_store).z. */
void tea_leaf_ppcg_calc_rrn_kernel_cuda_(double *norm);
/* tea_leaf_kernel_cuda.cu:488-500.  returns r.z (always r.z). */
void tea_leaf_ppcg_calc_2norm_kernel_cuda_(double *norm);
/* tea_leaf_kernel_cuda.cu:502-513.  p = z + beta p. */
void tea_leaf_ppcg_calc_p_kernel_cuda_(double *beta);

/* ======================================================================== *
 * Extensions (no counterpart in the reference).                            *
 * ======================================================================== */

/* Extra field ids for tea_engine_{get,set}_field beyond TEA_FIELD_*. */
enum {
    TEA_XFIELD_U0 = 8, TEA_XFIELD_W = 9, TEA_XFIELD_Z = 10, TEA_XFIELD_MI = 11,
    TEA_XFIELD_KX = 12, TEA_XFIELD_KY = 13, TEA_XFIELD_RTEMP = 14, TEA_XFIELD_UTEMP = 15,
    TEA_XFIELD_RSTORE = 16, TEA_XFIELD_CP = 17, TEA_XFIELD_BFP = 18, TEA_XFIELD_VOLUME = 19,
    TEA_XFIELD_COUNT = 20
};

/* Deck file read by initialise_cuda_ (default "tea.in" in the working directory). */
void tea_engine_set_deck_path(const char *path);

/* 0 (default): fatal errors print and abort, as the reference does.
 * 1: fatal errors print, set a sticky message and make the failing call return. */
void tea_engine_set_error_mode(int mode);
/* Copies the sticky error message (empty if none) and returns 1 if one is set. */
int tea_engine_last_error(char *buf, int buflen);
void tea_engine_clear_error(void);

/* Chunk geometry after initialise_cuda_: nx, ny, halo depth, solver, preconditioner. */
void tea_engine_info(int *nx, int *ny, int *halo_depth, int *solver, int *preconditioner);

/* Copy a whole field, halo included, to/from a host array of
 * (nx+2h)*(ny+2h) doubles laid out row-major with x fastest (the reference's
 * THARR2D layout, cuda_kernel_header.hpp:27-29).  Returns 0 on success. */
int tea_engine_get_field(int field_id, double *host);
int tea_engine_set_field(int field_id, const double *host);

/* Kernel profiler (the reference's profiler_on, cuda_common.hpp:97-122):
 * enable/disable per-launch HIP event timing, reset counters, and read
 * back count and total milliseconds for one kernel name.  Returns 0 if the
 * name was seen.  tea_engine_profile_names writes newline-separated names.
 * tea_engine_profile_only(prefix) times only launches whose name starts with
 * prefix (NULL or "": all); a timed launch costs the run a few microseconds
 * of idle GPU around it, so tea_engine_profile_sample(every) times one in
 * `every` of each name's launches (counts are of the timed ones). */
void tea_engine_profile_enable(int on);
void tea_engine_profile_only(const char *prefix);
void tea_engine_profile_sample(int every);
void tea_engine_profile_reset(void);
int tea_engine_profile_get(const char *kernel, long *count, double *total_ms);
int tea_engine_profile_names(char *buf, int buflen);

/* The stream all engine work is queued on (a hipStream_t), for callers that
 * record their own events around engine calls. */
void *tea_engine_stream(void);

/* Blocks until all queued engine work is done. */
void tea_engine_synchronize(void);

/* Releases every device allocation; the next call must be initialise_cuda_. */
void tea_engine_finalize(void);

/* ---- fused PPCG outer iteration -------------------------------------------- *
 * No counterpart in the reference: each call is the composition of reference
 * entry points given beside it, with identical results, in fewer passes over
 * HBM and fewer host round trips.  Used by the host driver for PPCG decks
 * without reflective boundaries (the reference driver's update_halo of p
 * between ppcg_calc_p and cg_calc_w is a no-op there).  Follow the
 * precision mode of tea_engine_set_precision.
 *   tea_engine_ppcg_begin: [ppcg_calc_p(beta) if with_p], cg_calc_w (pw_rro[0]),
 *                          ppcg_store_r, ppcg_calc_2norm (pw_rro[1])
 *   tea_engine_ppcg_ur:    cg_calc_ur(alpha) (rrn), ppcg_init_sd_new(theta)
 *   tea_engine_ppcg_end:   ppcg_update_z, ppcg_calc_rrn (rrn)
 *   tea_engine_ppcg_begin_ur: tea_engine_ppcg_begin, then tea_engine_ppcg_ur
 *                          with alpha = pw_rro[1] / pw_rro[0] formed on the
 *                          device (one host wait instead of two) */
void tea_engine_ppcg_begin(int with_p, double beta, double *pw_rro);
void tea_engine_ppcg_ur(double alpha, double theta, double *rrn);
void tea_engine_ppcg_end(double *rrn);
void tea_engine_ppcg_begin_ur(int with_p, double beta, double theta, double *pw_rro, double *rrn);
/* tea_engine_ppcg_end(rrn_end), then tea_engine_ppcg_begin_ur(1, *rrn_end /
 * rro_prev, theta, pw_rro, rrn) with the division done on the device (one host
 * wait instead of two); rro_prev is the r.z of the previous begin. */
void tea_engine_ppcg_end_begin_ur(double rro_prev, double theta, double *rrn_end, double *pw_rro, double *rrn);
/* tea_engine_ppcg_end_begin_ur, then inner steps 1..inner_steps
 * (tea_engine_ppcg_inner_n passes as tea_engine_ppcg_inner_plan splits them),
 * returning once the end_begin_ur values are known while the passes still
 * run.  A caller that then finds it needs a residual replacement first lets
 * tea_engine_mp_replace overwrite what the passes wrote (sd, rtemp, utemp)
 * and runs them again. */
void tea_engine_ppcg_end_begin_ur_inner(double rro_prev, double theta, int inner_steps, double *rrn_end,
                                        double *pw_rro, double *rrn);
/* One whole outer iteration, as the reference's own PPCG branch runs it
 * (tea_solve.f90:324-340 without halo exchanges): tea_engine_ppcg_begin_ur
 * (pw_rro, rrn_ur), inner steps 1..inner_steps as tea_engine_ppcg_inner_plan
 * splits them, then tea_engine_ppcg_end (rrn_end), queued back to back with
 * one host wait, at the end, instead of two. */
void tea_engine_ppcg_iteration(int with_p, double beta, double theta, int inner_steps, double *pw_rro,
                               double *rrn_ur, double *rrn_end);
/* ---- fused CG iteration ------------------------------------------------------ *
 * No counterpart in the reference: [tea_leaf_cg_calc_p_kernel_cuda_(beta) if
 * with_p], tea_leaf_cg_calc_w_kernel_cuda_ (pw), then
 * tea_leaf_cg_calc_ur_kernel_cuda_(rro / pw) (rrn), with identical results, in
 * two passes over HBM and one host wait (the division by pw is done on the
 * device as the host does it).  For CG-family decks without reflective
 * boundaries (the reference driver's update_halo of p between calc_p and
 * calc_w is a no-op there). */
void tea_engine_cg_begin_ur(int with_p, double beta, double rro, double *pw, double *rrn);
/* ---- Chebyshev iterations in passes ------------------------------------------- *
 * No counterpart in the reference: tea_leaf_cheby_iterate_kernel_cuda_ for
 * cheby_calc_step = step .. step + nsteps - 1 (tea_leaf_kernel_cuda.cu:125-135),
 * with identical results, in streaming passes of several iterations over HBM
 * (temporal blocking).  For decks without reflective boundaries (the reference
 * driver's update_halo of u between iterations is a no-op there). */
void tea_engine_cheby_iterate_n(int step, int nsteps);

/* Inner steps `step` .. `step + nsteps - 1` (1-based; nsteps 1..5) in one
 * pass over HBM (temporal blocking): the same results as
 * tea_leaf_ppcg_inner_kernel_cuda_ called for each of the steps with
 * bounds_extra 0 and all four faces external (one chunk).
 * tea_engine_ppcg_inner2(step) is tea_engine_ppcg_inner_n(step, 2).
 * tea_engine_ppcg_inner_plan(remaining) returns how many of `remaining`
 * consecutive steps the next call should take: the first pass of the
 * cheapest grouping for the current precision mode (e.g. 10 steps: 4, 3, 3). */
void tea_engine_ppcg_inner_n(int step, int nsteps);
void tea_engine_ppcg_inner2(int step);
int tea_engine_ppcg_inner_plan(int remaining);

/* ---- mixed precision: FP32 inner solver under FP64 refinement ------------ *
 * tea_engine_set_precision(32) switches the CG and PPCG entry points
 * (cg_init, cg_calc_w/ur/p, ppcg_init, init_sd(_new), inner, store_r,
 * update_z, calc_rrn, ppcg_calc_2norm, ppcg_calc_p) to FP32 shadow fields
 * (allocated on first use); their scalars stay FP64 and their dot products
 * accumulate in compensated FP64.  calc_residual, calc_2norm, field_summary,
 * the set-up and finalise calls always use the FP64 fields.  Chebyshev and
 * Jacobi calls are rejected in FP32 mode.  (64) switches back. */
int tea_engine_set_precision(int bits);
/* tea_engine_set_precision(64) at the end of an FP32 solve (no reference
 * counterpart): the FP32 inner-solver fields (w, z, sd, rtemp, utemp) are
 * left with unspecified values instead of being brought up to date, for a
 * caller whose next FP32 solve writes them before reading them. */
int tea_engine_mp_leave(void);
/* FP32 copies of the FP64 face coefficients (after tea_leaf_common_init). */
void tea_engine_mp_setup(void);
/* r32 = (float) r (the FP64 residual field), u32 = 0 (the FP32 correction);
 * returns rr = r32.r32. */
void tea_engine_mp_begin(double *rr);
/* u += (double) u32. */
void tea_engine_mp_end(void);
/* A residual replacement of the FP32 solver in one call (no reference
 * counterpart; PPCG, FP32 mode): tea_engine_mp_end,
 * tea_engine_set_precision(64), tea_leaf_calc_residual_cuda_,
 * tea_leaf_calc_2norm_kernel_cuda_(1) -> *rr64, tea_engine_set_precision(32),
 * tea_engine_mp_begin -> *rr32, tea_leaf_ppcg_init_kernel_cuda_ (step 2) and
 * tea_leaf_ppcg_init_sd_new_kernel_cuda_(theta), with the same results.
 * Leaves the FP32 mode on. */
void tea_engine_mp_replace(double theta, double *rr64, double *rr32);
/* Copy an FP32 shadow field to/from a host array of (nx+2h)*(ny+2h) floats. */
int tea_engine_get_field32(int field_id, float *host);
int tea_engine_set_field32(int field_id, const float *host);

/* ---- the 2-norm without allocation ------------------------------------------ *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the sum of a*b over the interior, as
 * tea_leaf_calc_2norm_kernel_cuda_ forms it (the same compensated reduction,
 * bit-identical results), on caller-owned FP64 device arrays of ny+2*halo rows
 * of `pitch` elements (element [r*pitch + c] is cell (c, r), halo included;
 * the interior is c in [halo, halo+nx), r in [halo, halo+ny)).  The result is
 * written to dev_norm (device memory) by work queued on `stream` (a
 * hipStream_t; NULL: the null stream); the call does not wait.  Allocates
 * nothing: the reduction's partials live in a 64 KiB device array of the
 * library that each GPU holds until the process ends, and every call on one
 * GPU shares it, so calls that overlap on one GPU race.  Needs no chunk
 * (initialise_cuda_ not required).  Limits: 1 <= nx, ny <= 65536, 0 <= halo <= 64,
 * pitch >= nx + 2*halo.  Returns 0, or -1 with a message on stderr and in
 * tea_engine_last_error for arguments beyond the limits (never aborts). */
int tea_engine_calc_2norm_noalloc(const double *a, const double *b, int nx, int ny, int halo, long pitch,
                                  double *dev_norm, void *stream);

#ifdef __cplusplus
}
#endif
#endif
