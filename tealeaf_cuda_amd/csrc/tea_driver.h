/*
 * tea_driver.h -- host driver of the TeaLeaf solver engine.
 *
 * The driver is the program above the C-ABI: it reads a tea.in deck, runs the
 * time loop and the solver control (CG, Chebyshev, PPCG, Jacobi), and calls the
 * 29 kernel entry points of a backend library (libtea_engine.so on the GPU, or
 * any library exporting the same symbols) in the same order as the TeaLeaf
 * Fortran driver (tea_leaf.f90, diffuse.f90, start.f90, tea_solve.f90).
 */
#ifndef TEA_DRIVER_H
#define TEA_DRIVER_H

#ifdef __cplusplus
extern "C" {
#endif

#define TEA_MAX_STEPS 4096
#define TEA_MAX_SUMMARIES 1024

typedef struct {
    int n_steps;                                  /* timesteps completed */
    int iterations[TEA_MAX_STEPS];                /* "Iteration count" (n-1) per step */
    int ppcg_inner[TEA_MAX_STEPS];                /* PPCG inner iterations per step */
    double conduction_error[TEA_MAX_STEPS];       /* error / initial residual */
    double exact_error[TEA_MAX_STEPS];            /* ||u0 - A u|| / ||r0|| (tl_check_result) */
    double initial_residual[TEA_MAX_STEPS];       /* ||r0|| */
    double step_seconds[TEA_MAX_STEPS];           /* wall time of the step */
    int refinements[TEA_MAX_STEPS];               /* refinement: residual replacements / inner solves */
    int warm_start[TEA_MAX_STEPS];                /* 1 if the learned initial guess was used */
    int n_summaries;
    int summary_step[TEA_MAX_SUMMARIES];
    double summary[TEA_MAX_SUMMARIES][6];         /* time, volume, mass, density, energy, u */
    double eigmin, eigmax;                        /* eigenvalue estimate (after scaling) */
    double total_seconds;
    int error_code;                               /* 0 = ok */
    char error_msg[512];
} tea_run_result;

/* Called after every timestep (after the step's work has completed on the
 * device) with the 1-based step number; NULL to remove. */
typedef void (*tea_step_callback)(int step, void *user);
void tea_driver_set_step_callback(tea_step_callback cb, void *user);

/* Runs the deck at deck_path with the kernel backend library backend_lib
 * (a path to a shared library exporting the tea_leaf_engine.h symbols).
 * If out_path is non-NULL a tea.out-style log is written there.
 * Returns 0 on success; on failure result->error_msg says why. */
int tea_driver_run(const char *deck_path, const char *backend_lib, const char *out_path,
                   tea_run_result *result);

#ifdef __cplusplus
}
#endif
#endif
