/*
 * tea_oracle.c -- CPU restatement of the TeaLeaf kernel layer (TEST INFRASTRUCTURE).
 *
 * This is the oracle.  It exports the same 29 Fortran-ABI entry points as the
 * HIP engine (include/tea_leaf_engine.h) and restates, in plain C, the
 * arithmetic of the reference CUDA kernels (the kernel_files .cuknl sources) in the same
 * expression order, and the launch sequence of tea_leaf_kernel_cuda.cu,
 * init_cuda.cu and the other kernel host files.  Only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it; the product path
 * never does.
 *
 * Build with -ffp-contract=off (no fused multiply-add unless written as fma()).
 *
 * Reductions (set with the environment variable TEA_ORACLE_DOT):
 *   "dot2"   (default) compensated dot products (Ogita-Rump-Oishi Dot2: exact
 *            products via fma, error-free TwoSum accumulation), combined
 *            row by row; the rounded result is the correctly rounded sum for
 *            all practical purposes, so it does not depend on summation order
 *            and matches the HIP engine's reductions bit for bit;
 *   "serial" plain left-to-right sum in row-major order (the order of the
 *            upstream Fortran kernels whose outputs the Benchmarks tea_bm_N.out
 *            files hold);
 *   "plain"  plain OpenMP reduction (the CPU baseline).
 *
 * Deliberate departures from the reference, all outside the observable
 * interior results of a valid deck:
 *   - init_jac_diag is evaluated on the interior only (the reference's +-h
 *     extent reads past the array end, tea_leaf_common.cuknl:131-147); halo Mi
 *     values are never read by an interior update.
 *   - A block-Jacobi strip cut short by the top of the extent (rows mod 4 != 0)
 *     is solved as a shorter tridiagonal system; the reference reads halo
 *     rows there and produces NaN.
 *   - Chebyshev coefficient storage grows when more coefficients arrive
 *     (the reference keeps the first allocation, tea_leaf_kernel_cuda.cu:93-103).
 *   - The reflective halo update follows the mirror rule of the pack kernels
 *     (pack_kernel.cuknl:26) because the reference's own update_halo path looks
 *     up a kernel name that does not exist (update_halo_kernel_cuda.cu:51).
 *
 * Parity pins for each function are listed in the test suite
 * (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

enum { CHUNK_LEFT = 1, CHUNK_RIGHT = 2, CHUNK_BOTTOM = 3, CHUNK_TOP = 4, EXTERNAL_FACE = -1 };
enum { F_DENSITY = 1, F_ENERGY0, F_ENERGY1, F_U, F_P, F_SD, F_R,
       X_U0, X_W, X_Z, X_MI, X_KX, X_KY, X_RTEMP, X_UTEMP, X_RSTORE, X_CP, X_BFP, X_VOLUME, X_COUNT };
enum { PREC_NONE = 1, PREC_DIAG = 2, PREC_BLOCK = 3 };
enum { SOLVER_JACOBI = 0, SOLVER_CG = 1, SOLVER_CHEBY = 2, SOLVER_PPCG = 3 };
enum { COEF_CONDUCTIVITY = 1, COEF_RECIP_CONDUCTIVITY = 2 };
enum { DOT_DOT2 = 0, DOT_SERIAL = 1, DOT_PLAIN = 2 };
#define JAC_BLOCK 4

static struct {
    int nx, ny, h, pitch, rows;
    int solver, prec;
    int dot_mode;
    double *f[X_COUNT];        /* 2D fields, index = c + r*pitch */
    double *zl;                /* block-Jacobi result kept out of z (the reference's shared-memory z_l) */
    float *f32[X_COUNT];       /* FP32 shadow fields of the mixed-precision inner solver */
    float *zl32;
    int fp32;                  /* CG / PPCG entry points work on the FP32 fields */
    double *xarea, *yarea;
    double *cellx, *celly, *celldx, *celldy, *vertexx, *vertexy, *vertexdx, *vertexdy;
    double *ch_a, *ch_b;
    int n_coefs;
    int initialised;
} g;

static char g_deck_path[4096] = "tea.in";
static int g_error_mode = 0;
static char g_error[1024] = "";

#define IDX(c, r) ((size_t)(c) + (size_t)(r) * (size_t)g.pitch)

static void die(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    fprintf(stderr, "tea_oracle: %s\n", g_error);
    if (g_error_mode == 0) abort();
}
#define CHECK_INIT(ret) do { if (!g.initialised) { die("%s called before initialise_cuda_", __func__); return ret; } } while (0)

/* Kernels are registered per solver (init_cuda.cu:251-282); launching one that
 * is not registered throws in the reference (std::map::at in CUDALAUNCH). */
#define S_CG_FAMILY ((1 << SOLVER_CG) | (1 << SOLVER_CHEBY) | (1 << SOLVER_PPCG))
#define REQUIRE_SOLVER(mask) do { if (!((mask) & (1 << g.solver))) { \
        die("%s: kernel not registered for the solver selected in the deck", __func__); return; } } while (0)

/* ---------------- deck reading: init_cuda.cu:63-163, cuda_strings.cu:7-83 ---- */

/* First line without '!' containing name: the text after '=' or the name itself. */
static int match_param(const char *name, char *out, size_t outlen)
{
    FILE *fp = fopen(g_deck_path, "r");
    char line[4096];
    out[0] = 0;
    if (!fp) return -1;
    while (fgets(line, sizeof line, fp)) {
        size_t n = strlen(line);
        while (n && (line[n - 1] == '\n' || line[n - 1] == '\r')) line[--n] = 0;
        if (strchr(line, '!')) continue;
        if (strstr(line, name)) {
            char *eq = strchr(line, '=');
            snprintf(out, outlen, "%s", eq ? eq + 1 : name);
            break;
        }
    }
    fclose(fp);
    return 0;
}
static int param_enabled(const char *name) { char v[4096]; match_param(name, v, sizeof v); return v[0] != 0; }
static int read_int(const char *name)
{
    char v[4096]; int x;
    match_param(name, v, sizeof v);
    if (!v[0] || sscanf(v, " %d", &x) != 1) return -1;
    return x;
}
static void read_string(const char *name, char *v, size_t n)
{
    match_param(name, v, n);
    for (char *p = v; *p; p++) *p = (char)tolower((unsigned char)*p);
}

/* ---------------- compensated arithmetic ------------------------------------ */

typedef struct { double s, c; } dd_t;
/* acc += a*b, with the product and the sum carried error-free (Dot2). */
static inline void dd_add_prod(dd_t *acc, double a, double b)
{
    const double p = a * b;
    const double pe = fma(a, b, -p);
    const double s = acc->s + p;
    const double bb = s - acc->s;
    const double se = (acc->s - (s - bb)) + (p - bb);
    acc->s = s;
    acc->c += se + pe;
}
static inline void dd_merge(dd_t *acc, dd_t x)
{
    const double s = acc->s + x.s;
    const double bb = s - acc->s;
    const double se = (acc->s - (s - bb)) + (x.s - bb);
    acc->s = s;
    acc->c += (se + x.c);
}

/* Reduction over rows [Y0,Y1] x cols [X0,X1].  BODY sees c, r, i = IDX(c,r),
 * may update fields, and sets the product factors ta, tb of this cell's term. */
#define REDUCE_EXTENT(RESULT, X0, X1, Y0, Y1, ...)                                      \
    do {                                                                                 \
        const int x0_ = (X0), x1_ = (X1), y0_ = (Y0), y1_ = (Y1);                        \
        if (g.dot_mode == DOT_SERIAL) {                                                  \
            double s_ = 0.0;                                                             \
            for (int r = y0_; r <= y1_; r++)                                             \
                for (int c = x0_; c <= x1_; c++) {                                       \
                    const size_t i = IDX(c, r); (void)i;                                 \
                    double ta = 0.0, tb = 0.0;                                           \
                    __VA_ARGS__;                                                         \
                    s_ += ta * tb;                                                       \
                }                                                                        \
            RESULT = s_;                                                                 \
        } else if (g.dot_mode == DOT_PLAIN) {                                            \
            double s_ = 0.0;                                                             \
            _Pragma("omp parallel for reduction(+:s_) schedule(static)")                 \
            for (int r = y0_; r <= y1_; r++)                                             \
                for (int c = x0_; c <= x1_; c++) {                                       \
                    const size_t i = IDX(c, r); (void)i;                                 \
                    double ta = 0.0, tb = 0.0;                                           \
                    __VA_ARGS__;                                                         \
                    s_ += ta * tb;                                                       \
                }                                                                        \
            RESULT = s_;                                                                 \
        } else {                                                                         \
            const int nr_ = y1_ >= y0_ ? y1_ - y0_ + 1 : 0;                              \
            dd_t *rows_ = (dd_t *)calloc((size_t)(nr_ > 0 ? nr_ : 1), sizeof(dd_t));     \
            _Pragma("omp parallel for schedule(static)")                                 \
            for (int r = y0_; r <= y1_; r++) {                                           \
                dd_t a_ = {0.0, 0.0};                                                    \
                for (int c = x0_; c <= x1_; c++) {                                       \
                    const size_t i = IDX(c, r); (void)i;                                 \
                    double ta = 0.0, tb = 0.0;                                           \
                    __VA_ARGS__;                                                         \
                    dd_add_prod(&a_, ta, tb);                                            \
                }                                                                        \
                rows_[r - y0_] = a_;                                                     \
            }                                                                            \
            dd_t t_ = {0.0, 0.0};                                                        \
            for (int k_ = 0; k_ < nr_; k_++) dd_merge(&t_, rows_[k_]);                   \
            free(rows_);                                                                 \
            RESULT = t_.s + t_.c;                                                        \
        }                                                                                \
    } while (0)

#define FOR_EXTENT(X0, X1, Y0, Y1, ...)                                                 \
    do {                                                                                 \
        const int x0_ = (X0), x1_ = (X1), y0_ = (Y0), y1_ = (Y1);                        \
        _Pragma("omp parallel for schedule(static)")                                     \
        for (int r = y0_; r <= y1_; r++)                                                 \
            for (int c = x0_; c <= x1_; c++) {                                           \
                const size_t i = IDX(c, r); (void)i;                                     \
                __VA_ARGS__;                                                             \
            }                                                                            \
    } while (0)

/* interior extent (kernel_info_t with 0,0,0,0: init_cuda.cu:222-292) */
#define IX0 (g.h)
#define IX1 (g.h + g.nx - 1)
#define IY0 (g.h)
#define IY1 (g.h + g.ny - 1)

#define DENS  g.f[F_DENSITY]
#define E0    g.f[F_ENERGY0]
#define E1    g.f[F_ENERGY1]
#define U     g.f[F_U]
#define P     g.f[F_P]
#define SD    g.f[F_SD]
#define R     g.f[F_R]
#define U0    g.f[X_U0]
#define W     g.f[X_W]
#define Z     g.f[X_Z]
#define MI    g.f[X_MI]
#define KX    g.f[X_KX]
#define KY    g.f[X_KY]
#define RTEMP g.f[X_RTEMP]
#define UTEMP g.f[X_UTEMP]
#define RSTORE g.f[X_RSTORE]
#define CP    g.f[X_CP]
#define BFP   g.f[X_BFP]
#define VOL   g.f[X_VOLUME]

static double *zl_f64(void) { return g.zl; }
static float *zl_f32(void) { return g.zl32; }

#define REAL double
#define S(n) n##_f64
#define FLD(id) g.f[id]
#include "tea_oracle_solver.inc"
#undef REAL
#undef S
#undef FLD
#define REAL float
#define S(n) n##_f32
#define FLD(id) g.f32[id]
#include "tea_oracle_solver.inc"
#undef REAL
#undef S
#undef FLD

/* FP64 names used by the Chebyshev, Jacobi and residual kernels */
#define apply_A apply_A_f64
#define coef_b coef_b_f64
#define block_solve_extent block_solve_extent_f64

/* ---------------- set-up ------------------------------------------------------ */

static void free_all(void)
{
    for (int k = 0; k < X_COUNT; k++) { free(g.f[k]); g.f[k] = NULL; }
    free(g.zl); g.zl = NULL;
    for (int k = 0; k < X_COUNT; k++) { free(g.f32[k]); g.f32[k] = NULL; }
    free(g.zl32); g.zl32 = NULL;
    g.fp32 = 0;
    free(g.xarea); free(g.yarea);
    free(g.cellx); free(g.celly); free(g.celldx); free(g.celldy);
    free(g.vertexx); free(g.vertexy); free(g.vertexdx); free(g.vertexdy);
    free(g.ch_a); free(g.ch_b);
    g.xarea = g.yarea = g.cellx = g.celly = g.celldx = g.celldy = NULL;
    g.vertexx = g.vertexy = g.vertexdx = g.vertexdy = g.ch_a = g.ch_b = NULL;
    g.n_coefs = 0;
    g.initialised = 0;
}

/* init_cuda.cu:35-167 and initBuffers :295-371 (zero-filled arrays). */
EXPORT void initialise_cuda_(int *x_min, int *x_max, int *y_min, int *y_max)
{
    free_all();
    const char *mode = getenv("TEA_ORACLE_DOT");
    g.dot_mode = DOT_DOT2;
    if (mode && !strcmp(mode, "serial")) g.dot_mode = DOT_SERIAL;
    else if (mode && !strcmp(mode, "plain")) g.dot_mode = DOT_PLAIN;

    FILE *fp = fopen(g_deck_path, "r");
    if (!fp) { die("Input file not found: %s", g_deck_path); return; }
    fclose(fp);

    g.nx = *x_max - *x_min + 1;
    g.ny = *y_max - *y_min + 1;
    g.h = read_int("halo_depth");
    if (g.h < 1) { die("Halo exchange depth unspecified or was too small"); return; }
    if (param_enabled("tl_use_ppcg")) g.solver = SOLVER_PPCG;
    else if (param_enabled("tl_use_chebyshev")) g.solver = SOLVER_CHEBY;
    else if (param_enabled("tl_use_cg")) g.solver = SOLVER_CG;
    else g.solver = SOLVER_JACOBI;
    char prec[4096];
    read_string("tl_preconditioner_type", prec, sizeof prec);
    if (strstr(prec, "jac_diag")) g.prec = PREC_DIAG;
    else if (strstr(prec, "jac_block")) g.prec = PREC_BLOCK;
    else g.prec = PREC_NONE;

    g.pitch = g.nx + 2 * g.h;
    g.rows = g.ny + 2 * g.h;
    const size_t n2 = (size_t)g.pitch * (size_t)g.rows;
    for (int k = 1; k < X_COUNT; k++) g.f[k] = (double *)calloc(n2, sizeof(double));
    g.zl = (double *)calloc(n2, sizeof(double));
    g.xarea = (double *)calloc((size_t)(g.pitch + 1) * (size_t)g.rows, sizeof(double));
    g.yarea = (double *)calloc((size_t)g.pitch * (size_t)(g.rows + 1), sizeof(double));
    /* one spare element: the vertex kernel writes index nx+2h (initialise_chunk_kernel.cuknl:21,43) */
    g.cellx = (double *)calloc((size_t)g.pitch + 2, sizeof(double));
    g.celldx = (double *)calloc((size_t)g.pitch + 2, sizeof(double));
    g.vertexx = (double *)calloc((size_t)g.pitch + 2, sizeof(double));
    g.vertexdx = (double *)calloc((size_t)g.pitch + 2, sizeof(double));
    g.celly = (double *)calloc((size_t)g.rows + 2, sizeof(double));
    g.celldy = (double *)calloc((size_t)g.rows + 2, sizeof(double));
    g.vertexy = (double *)calloc((size_t)g.rows + 2, sizeof(double));
    g.vertexdy = (double *)calloc((size_t)g.rows + 2, sizeof(double));
    g.initialised = 1;
}

/* initialise_chunk_kernel.cuknl:3-70 over the +-h extent. */
EXPORT void initialise_chunk_kernel_cuda_(double *xmin, double *ymin, double *dx, double *dy)
{
    CHECK_INIT();
    const double d_xmin = *xmin, d_ymin = *ymin, d_dx = *dx, d_dy = *dy;
    const int h = g.h;
    for (int c = 0; c <= g.nx + 2 * h; c++) {
        g.vertexx[c] = d_xmin + d_dx * (double)((c - 1) - h + 1);
        g.vertexdx[c] = d_dx;
        const double vplus = d_xmin + d_dx * (double)(c - h + 1);
        g.cellx[c] = 0.5 * (g.vertexx[c] + vplus);
        g.celldx[c] = d_dx;
    }
    for (int r = 0; r <= g.ny + 2 * h; r++) {
        g.vertexy[r] = d_ymin + d_dy * (double)((r - 1) - h + 1);
        g.vertexdy[r] = d_dy;
        const double vplus = d_ymin + d_dy * (double)(r - h + 1);
        g.celly[r] = 0.5 * (g.vertexy[r] + vplus);
        g.celldy[r] = d_dy;
    }
    for (int r = 0; r < g.rows; r++)
        for (int c = 0; c < g.pitch; c++) {
            VOL[IDX(c, r)] = d_dx * d_dy;
            g.xarea[(size_t)c + (size_t)r * (size_t)(g.pitch + 1)] = d_dy;
            g.yarea[IDX(c, r)] = d_dx;
        }
}

/* generate_chunk_kernel_cuda.cu:47-115 with generate_chunk_kernel.cuknl:5-99. */
EXPORT void generate_chunk_kernel_cuda_(const int *number_of_states,
                                        const double *state_density, const double *state_energy,
                                        const double *state_xmin, const double *state_xmax,
                                        const double *state_ymin, const double *state_ymax,
                                        const double *state_radius, const int *state_geometry,
                                        const int *g_rect, const int *g_circ, const int *g_point)
{
    CHECK_INIT();
    const int ns = *number_of_states;
    if (ns < 1) { die("generate_chunk: no states"); return; }
    FOR_EXTENT(IX0, IX1, IY0, IY1, { E0[i] = state_energy[0]; DENS[i] = state_density[0]; });
    for (int s = 1; s < ns; s++) {
        const double x_cent = state_xmin[s], y_cent = state_ymin[s];
        const int geom = state_geometry[s];
        FOR_EXTENT(IX0, IX1, IY0, IY1, {
            int hit = 0;
            if (*g_rect == geom) {
                hit = g.vertexx[1 + c] >= state_xmin[s] && g.vertexx[c] < state_xmax[s] &&
                      g.vertexy[1 + r] >= state_ymin[s] && g.vertexy[r] < state_ymax[s];
            } else if (geom == *g_circ) {
                const double x_pos = g.cellx[c] - x_cent;
                const double y_pos = g.celly[r] - y_cent;
                const double radius = sqrt(x_pos * x_pos + y_pos * y_pos);
                hit = radius <= state_radius[s];
            } else if (geom == *g_point) {
                hit = g.vertexx[c] == x_cent && g.vertexy[r] == y_cent;
            }
            if (hit) { E0[i] = state_energy[s]; DENS[i] = state_density[s]; }
        });
    }
    FOR_EXTENT(IX0, IX1, IY0, IY1, { U[i] = E0[i] * DENS[i]; U0[i] = E0[i] * DENS[i]; });
}

/* set_field_kernel.cuknl:5-16 */
EXPORT void set_field_kernel_cuda_(void)
{
    CHECK_INIT();
    FOR_EXTENT(IX0, IX1, IY0, IY1, { E1[i] = E0[i]; });
}

/* field_summary_kernel.cuknl:5-58: four sums over the interior. */
EXPORT void field_summary_kernel_cuda_(double *vol, double *mass, double *ie, double *temp)
{
    CHECK_INIT();
    double v = 0, m = 0, e = 0, t = 0;
    REDUCE_EXTENT(v, IX0, IX1, IY0, IY1, { ta = VOL[i]; tb = 1.0; });
    REDUCE_EXTENT(m, IX0, IX1, IY0, IY1, { ta = VOL[i]; tb = DENS[i]; });
    REDUCE_EXTENT(e, IX0, IX1, IY0, IY1, { ta = VOL[i] * DENS[i]; tb = E1[i]; });
    REDUCE_EXTENT(t, IX0, IX1, IY0, IY1, { ta = VOL[i] * DENS[i]; tb = U[i]; });
    *vol = v; *mass = m; *ie = e; *temp = t;
}

/* ---------------- halos -------------------------------------------------------- */

static double *field_by_id(int id)
{
    if (id < 1 || id >= X_COUNT) return NULL;
    return g.f[id];
}

/* Reflective update: halo cell at distance k from an external face takes the
 * interior cell at distance k-1 (pack_kernel.cuknl:26 mirror), bottom and top
 * over columns [h-d, h+nx-1+d], then left and right over rows [h-d, h+ny-1+d]. */
static void reflect_field(double *a, const int *nb, int d)
{
    const int h = g.h, nx = g.nx, ny = g.ny;
    if (nb[CHUNK_BOTTOM - 1] == EXTERNAL_FACE)
        for (int k = 1; k <= d; k++)
            for (int c = h - d; c <= h + nx - 1 + d; c++) a[IDX(c, h - k)] = a[IDX(c, h + k - 1)];
    if (nb[CHUNK_TOP - 1] == EXTERNAL_FACE)
        for (int k = 1; k <= d; k++)
            for (int c = h - d; c <= h + nx - 1 + d; c++) a[IDX(c, h + ny - 1 + k)] = a[IDX(c, h + ny - k)];
    if (nb[CHUNK_LEFT - 1] == EXTERNAL_FACE)
        for (int r = h - d; r <= h + ny - 1 + d; r++)
            for (int k = 1; k <= d; k++) a[IDX(h - k, r)] = a[IDX(h + k - 1, r)];
    if (nb[CHUNK_RIGHT - 1] == EXTERNAL_FACE)
        for (int r = h - d; r <= h + ny - 1 + d; r++)
            for (int k = 1; k <= d; k++) a[IDX(h + nx - 1 + k, r)] = a[IDX(h + nx - k, r)];
}

EXPORT void update_halo_kernel_cuda_(const int *chunk_neighbours, const int *fields, const int *depth)
{
    CHECK_INIT();
    const int d = *depth;
    if (d < 1 || d > g.h) { die("update_halo: invalid depth %d (halo depth %d)", d, g.h); return; }
    for (int f = 1; f <= 7; f++)
        if (fields[f - 1] == 1) reflect_field(field_by_id(f), chunk_neighbours, d);
}

/* pack_kernel.cuknl:5-147 layouts.  Left/right strips: (ny+2d) rows x d columns,
 * element (j, k) at offset + j + k*d; bottom/top strips: d rows x (nx+2d)
 * columns, element (j, k) at offset + j + k*(nx+2d).  Left and bottom strips
 * hold the face-adjacent cells mirrored (j or k = d-1 is the first interior
 * cell); right and top strips are in natural order. */
static void pack_unpack(int *fields, int *offsets, int d, int face, int pack, double *buf)
{
    const int h = g.h, nx = g.nx, ny = g.ny;
    if (d < 1 || d > g.h) { die("pack: invalid depth %d", d); return; }
    if (face < CHUNK_LEFT || face > CHUNK_TOP) { die("Invalid face identifier %d passed to mpi buffer packing", face); return; }
    const int side = (face <= CHUNK_RIGHT) ? d * (ny + 2 * d) : d * (nx + 2 * d);
    for (int f = 1; f <= 7; f++) {
        if (!fields[f - 1]) continue;
        if (offsets[f - 1] < 0 || offsets[f - 1] > 7 * side) {
            die("Tried to pack/unpack field %d but invalid offset %d given", f - 1, offsets[f - 1]);
            return;
        }
        double *a = field_by_id(f);
        double *b = buf + offsets[f - 1];
        if (face == CHUNK_LEFT || face == CHUNK_RIGHT) {
            for (int k = 0; k < ny + 2 * d; k++)
                for (int j = 0; j < d; j++) {
                    const int r = h - d + k;
                    int c;
                    if (face == CHUNK_LEFT) c = pack ? (h + d - 1 - j) : (h - d + j);
                    else c = pack ? (h + nx - d + j) : (h + nx + d - 1 - j);
                    if (pack) b[j + k * d] = a[IDX(c, r)];
                    else a[IDX(c, r)] = b[j + k * d];
                }
        } else {
            for (int k = 0; k < d; k++)
                for (int j = 0; j < nx + 2 * d; j++) {
                    const int c = h - d + j;
                    int r;
                    if (face == CHUNK_BOTTOM) r = pack ? (h + d - 1 - k) : (h - d + k);
                    else r = pack ? (h + ny - d + k) : (h + ny + d - 1 - k);
                    if (pack) b[j + k * (nx + 2 * d)] = a[IDX(c, r)];
                    else a[IDX(c, r)] = b[j + k * (nx + 2 * d)];
                }
        }
    }
}

EXPORT void cuda_pack_buffers_(int *fields, int *offsets, int *depth, int *face, double *host_buffer)
{
    CHECK_INIT();
    pack_unpack(fields, offsets, *depth, *face, 1, host_buffer);
}

EXPORT void cuda_unpack_buffers_(int *fields, int *offsets, int *depth, int *face, double *host_buffer)
{
    CHECK_INIT();
    pack_unpack(fields, offsets, *depth, *face, 0, host_buffer);
}

/* ---------------- solver: common ------------------------------------------------ */

/* tea_leaf_kernel_cuda.cu:16-32 (calcrxry), :254-283, tea_leaf_common.cuknl:60-129,
 * generate_chunk_kernel.cuknl:69-83 (init_u from energy1). */
EXPORT void tea_leaf_common_init_kernel_cuda_(const int *coefficient, double *dt, double *rx, double *ry,
                                              int *zero_boundary, int *reflective_boundary)
{
    CHECK_INIT();
    const int coef = *coefficient;
    if (coef != COEF_CONDUCTIVITY && coef != COEF_RECIP_CONDUCTIVITY) {
        die("Unknown coefficient %d passed to tea leaf", coef);
        return;
    }
    const double dx = g.celldx[g.h], dy = g.celldy[g.h];
    *rx = *dt / (dx * dx);
    *ry = *dt / (dy * dy);
    const double rxv = *rx, ryv = *ry;
    const int h = g.h;
    const size_t pp = (size_t)g.pitch;
    /* init_common over x,y in [0,+h] */
    FOR_EXTENT(h, h + g.nx - 1 + h, h, h + g.ny - 1 + h, {
        double dc = 0, dl = 0, dd = 0;
        if (coef == COEF_CONDUCTIVITY) {
            dc = DENS[i];
            if (i - 1 > 0) dl = DENS[i - 1];
            if (i - pp > 0) dd = DENS[i - pp];
        } else {
            dc = 1.0 / DENS[i];
            dl = 1.0 / DENS[i - 1];
            dd = 1.0 / DENS[i - pp];
        }
        double kx = (dl + dc) / (2.0 * dl * dc);
        kx *= rxv;
        double ky = (dd + dc) / (2.0 * dd * dc);
        ky *= ryv;
        KX[i] = kx;
        KY[i] = ky;
    });
    if (!*reflective_boundary) {
        const int zl = zero_boundary[0], zr = zero_boundary[1], zb = zero_boundary[2], zt = zero_boundary[3];
        FOR_EXTENT(0, g.pitch - 1, 0, g.rows - 1, {
            if (zl && c <= h) KX[i] = 0;
            if (zr && c > (g.nx - 1) + h) KX[i] = 0;
            if (zb && r <= h) KY[i] = 0;
            if (zt && r > (g.ny - 1) + h) KY[i] = 0;
        });
    }
    FOR_EXTENT(IX0, IX1, IY0, IY1, { U[i] = E1[i] * DENS[i]; U0[i] = E1[i] * DENS[i]; });
}

/* tea_leaf_common.cuknl:19-39 */
EXPORT void tea_leaf_calc_residual_cuda_(void)
{
    CHECK_INIT();
    FOR_EXTENT(IX0, IX1, IY0, IY1, {
        const double smvp = apply_A(U, i);
        R[i] = U0[i] - smvp;
    });
}

/* tea_leaf_common.cuknl:5-17 */
EXPORT void tea_leaf_common_finalise_kernel_cuda_(void)
{
    CHECK_INIT();
    FOR_EXTENT(IX0, IX1, IY0, IY1, { E1[i] = U[i] / DENS[i]; });
}

/* tea_leaf_kernel_cuda.cu:56-87 */
EXPORT void tea_leaf_calc_2norm_kernel_cuda_(int *norm_array, double *norm)
{
    CHECK_INIT();
    if (*norm_array < 0 || *norm_array > 2) {
        die("Invalid value '%d' for norm_array passed, should be [0, 1, 2]", *norm_array);
        return;
    }
    double s = 0;
    if (g.fp32) {   /* the FP32 solver's fields (its right-hand side, residual, z) */
        const float *u0 = g.f32[X_U0], *r = g.f32[F_R], *z = (g.prec != PREC_NONE) ? g.f32[X_Z] : g.f32[F_R];
        const float *a = *norm_array == 0 ? u0 : r, *b = *norm_array == 0 ? u0 : (*norm_array == 1 ? r : z);
        REDUCE_EXTENT(s, IX0, IX1, IY0, IY1, { ta = a[i]; tb = b[i]; });
    } else {
        const double *a, *b;
        if (*norm_array == 0) { a = U0; b = U0; }
        else if (*norm_array == 1) { a = R; b = R; }
        else { a = R; b = (g.prec != PREC_NONE) ? Z : R; }
        REDUCE_EXTENT(s, IX0, IX1, IY0, IY1, { ta = a[i]; tb = b[i]; });
    }
    *norm = s;
}

/* ---------------- solver: CG ----------------------------------------------------- */

/* tea_leaf_kernel_cuda.cu:164-186; tea_leaf_cg.cuknl:5-39; tea_leaf_common.cuknl:131-147 */
EXPORT void tea_leaf_cg_init_kernel_cuda_(double *rro)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    if (g.fp32) cg_init_f32(rro); else cg_init_f64(rro);
}

/* tea_leaf_cg.cuknl:43-68 */
EXPORT void tea_leaf_cg_calc_w_kernel_cuda_(double *pw)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    if (g.fp32) cg_calc_w_f32(pw); else cg_calc_w_f64(pw);
}

/* tea_leaf_cg.cuknl:72-137 */
EXPORT void tea_leaf_cg_calc_ur_kernel_cuda_(double *alpha, double *rrn)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    if (g.fp32) cg_calc_ur_f32(*alpha, rrn); else cg_calc_ur_f64(*alpha, rrn);
}

/* tea_leaf_cg.cuknl:141-161 */
EXPORT void tea_leaf_cg_calc_p_kernel_cuda_(double *beta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    if (g.fp32) calc_p_f32(*beta, g.prec != PREC_NONE ? g.f32[X_Z] : g.f32[F_R]);
    else calc_p_f64(*beta, g.prec != PREC_NONE ? g.f[X_Z] : g.f[F_R]);
}

/* ---------------- solver: Chebyshev ------------------------------------------------ */

static void upload_coefs(const double *a, const double *b, int n)
{
    if (n < 0) n = 0;
    if (n > g.n_coefs) {
        free(g.ch_a); free(g.ch_b);
        g.ch_a = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        g.ch_b = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        g.n_coefs = n;
    }
    if (n) { memcpy(g.ch_a, a, (size_t)n * sizeof(double)); memcpy(g.ch_b, b, (size_t)n * sizeof(double)); }
}

/* tea_leaf_kernel_cuda.cu:105-123; tea_leaf_cheby.cuknl:5-82 */
EXPORT void tea_leaf_cheby_init_kernel_cuda_(const double *ch_alphas, const double *ch_betas,
                                             int *n_coefs, const double *theta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_CHEBY);
    upload_coefs(ch_alphas, ch_betas, *n_coefs);
    if (g.fp32) cheby_step_f32(1, *theta, 0.0, 0.0); else cheby_step_f64(1, *theta, 0.0, 0.0);
}

/* tea_leaf_kernel_cuda.cu:125-135; tea_leaf_cheby.cuknl:84-152 */
EXPORT void tea_leaf_cheby_iterate_kernel_cuda_(const int *cheby_calc_step)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_CHEBY);
    const int k = *cheby_calc_step - 1;
    if (k < 0 || k >= g.n_coefs) { die("cheby_iterate: step %d outside the %d uploaded coefficients", k + 1, g.n_coefs); return; }
    if (g.fp32) cheby_step_f32(0, 0.0, g.ch_a[k], g.ch_b[k]); else cheby_step_f64(0, 0.0, g.ch_a[k], g.ch_b[k]);
}

/* The engine's run of Chebyshev iterations, as the composition it names. */
EXPORT void tea_engine_cheby_iterate_n(int step, int nsteps)
{
    if (nsteps < 1) { die("cheby_iterate_n: %d steps", nsteps); return; }
    for (int s = step; s < step + nsteps; s++) tea_leaf_cheby_iterate_kernel_cuda_(&s);
}

/* ---------------- solver: Jacobi ----------------------------------------------------- */

/* tea_leaf_kernel_cuda.cu:221-228; tea_leaf_jacobi.cuknl:5-46 (un lives in the Mi buffer) */
EXPORT void tea_leaf_jacobi_solve_kernel_cuda_(double *error)
{
    CHECK_INIT();
    if (g.fp32) { die("%s: not available in the FP32 solver mode", __func__); return; }
    REQUIRE_SOLVER(1 << SOLVER_JACOBI);
    double *un = MI;
    const size_t pp = (size_t)g.pitch;
    FOR_EXTENT(IX0, IX1, IY0, IY1, { un[i] = U[i]; });
    double s = 0;
    REDUCE_EXTENT(s, IX0, IX1, IY0, IY1, {
        U[i] = (U0[i]
                + KX[i + 1] * un[i + 1]
                + KX[i] * un[i - 1]
                + KY[i + pp] * un[i + pp]
                + KY[i] * un[i - pp])
               / (1.0 + (KX[i] + KX[i + 1])
                      + (KY[i] + KY[i + pp]));
        ta = fabs(U[i] - un[i]); tb = 1.0;
    });
    *error = s;
}

/* ---------------- solver: PPCG ------------------------------------------------------- */

/* tea_leaf_kernel_cuda.cu:357-362 */
EXPORT void tea_leaf_ppcg_init_constants_cuda_(const double *ch_alphas, const double *ch_betas, int *n_inner_steps)
{
    CHECK_INIT();
    upload_coefs(ch_alphas, ch_betas, *n_inner_steps);
}

/* tea_leaf_kernel_cuda.cu:366-375; tea_leaf_ppcg.cuknl:319-420 */
EXPORT void tea_leaf_ppcg_init_kernel_cuda_(const int *step, double *rro)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    const int st = *step;
    if (st < 1 || st > 3) { die("ppcg_init: invalid step %d", st); return; }
    if (g.fp32) ppcg_init_f32(st, rro); else ppcg_init_f64(st, rro);
}

/* tea_leaf_ppcg.cuknl:5-55 (never called by the reference driver) */
EXPORT void tea_leaf_ppcg_init_sd_kernel_cuda_(const double *theta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (g.fp32) ppcg_init_sd_f32(*theta); else ppcg_init_sd_f64(*theta);
}

/* tea_leaf_ppcg.cuknl:58-89 */
EXPORT void tea_leaf_ppcg_init_sd_new_kernel_cuda_(const double *theta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (g.fp32) ppcg_init_sd_new_f32(*theta); else ppcg_init_sd_new_f64(*theta);
}

/* Matrix-powers extent of one inner step (tea_leaf_kernel_cuda.cu:418-452). */
static void ppcg_extent(int bounds_extra, const int *nb, int *x0, int *x1, int *y0, int *y1)
{
    const int h = g.h;
    *x0 = (nb[CHUNK_LEFT - 1] == EXTERNAL_FACE) ? h : h - bounds_extra;
    *x1 = h + g.nx - 1 + ((nb[CHUNK_RIGHT - 1] == EXTERNAL_FACE) ? 0 : bounds_extra);
    *y0 = (nb[CHUNK_BOTTOM - 1] == EXTERNAL_FACE) ? h : h - bounds_extra;
    *y1 = h + g.ny - 1 + ((nb[CHUNK_TOP - 1] == EXTERNAL_FACE) ? 0 : bounds_extra);
}

/* tea_leaf_kernel_cuda.cu:418-486; tea_leaf_ppcg.cuknl:154-244 */
EXPORT void tea_leaf_ppcg_inner_kernel_cuda_(int *ppcg_cur_step, int *bounds_extra, int *chunk_neighbours)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    const int k = *ppcg_cur_step - 1;
    if (k < 0 || k >= g.n_coefs) { die("ppcg_inner: step %d outside the %d uploaded coefficients", k + 1, g.n_coefs); return; }
    if (*bounds_extra < 0 || *bounds_extra >= g.h) { die("ppcg_inner: invalid bounds_extra %d", *bounds_extra); return; }
    int x0, x1, y0, y1;
    ppcg_extent(*bounds_extra, chunk_neighbours, &x0, &x1, &y0, &y1);
    if (g.fp32) ppcg_inner_f32(g.ch_a[k], g.ch_b[k], x0, x1, y0, y1);
    else ppcg_inner_f64(g.ch_a[k], g.ch_b[k], x0, x1, y0, y1);
}

/* tea_leaf_ppcg.cuknl:249-260 */
EXPORT void tea_leaf_ppcg_store_r_kernel_cuda_(void)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (g.fp32) copy_f32(g.f32[X_RSTORE], g.f32[F_R]); else copy_f64(g.f[X_RSTORE], g.f[F_R]);
}

/* tea_leaf_ppcg.cuknl:263-274 */
EXPORT void tea_leaf_ppcg_update_z_kernel_cuda_(void)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (g.fp32) copy_f32(g.f32[X_Z], g.f32[X_UTEMP]); else copy_f64(g.f[X_Z], g.f[X_UTEMP]);
}

/* tea_leaf_ppcg.cuknl:279-297 */
EXPORT void tea_leaf_ppcg_calc_rrn_kernel_cuda_(double *norm)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (g.fp32) ppcg_calc_rrn_f32(norm); else ppcg_calc_rrn_f64(norm);
}

/* tea_leaf_kernel_cuda.cu:494-500 */
EXPORT void tea_leaf_ppcg_calc_2norm_kernel_cuda_(double *norm)
{
    CHECK_INIT();
    if (g.fp32) ppcg_calc_2norm_f32(norm); else ppcg_calc_2norm_f64(norm);
}

/* tea_leaf_ppcg.cuknl:301-313 */
EXPORT void tea_leaf_ppcg_calc_p_kernel_cuda_(double *beta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (g.fp32) calc_p_f32(*beta, g.f32[X_Z]); else calc_p_f64(*beta, g.f[X_Z]);
}

/* ---------------- extensions ---------------------------------------------------------- */

EXPORT void tea_engine_set_deck_path(const char *path) { snprintf(g_deck_path, sizeof g_deck_path, "%s", path ? path : "tea.in"); }
EXPORT void tea_engine_set_error_mode(int mode) { g_error_mode = mode; }
EXPORT int tea_engine_last_error(char *buf, int buflen)
{
    if (buf && buflen > 0) snprintf(buf, (size_t)buflen, "%s", g_error);
    return g_error[0] != 0;
}
EXPORT void tea_engine_clear_error(void) { g_error[0] = 0; }
EXPORT void tea_engine_info(int *nx, int *ny, int *halo_depth, int *solver, int *preconditioner)
{
    *nx = g.nx; *ny = g.ny; *halo_depth = g.h; *solver = g.solver; *preconditioner = g.prec;
}
EXPORT int tea_engine_get_field(int id, double *host)
{
    if (!g.initialised || id < 1 || id >= X_COUNT) return -1;
    memcpy(host, g.f[id], (size_t)g.pitch * (size_t)g.rows * sizeof(double));
    return 0;
}
EXPORT int tea_engine_set_field(int id, const double *host)
{
    if (!g.initialised || id < 1 || id >= X_COUNT) return -1;
    memcpy(g.f[id], host, (size_t)g.pitch * (size_t)g.rows * sizeof(double));
    return 0;
}
EXPORT void tea_engine_profile_enable(int on) { (void)on; }
EXPORT void tea_engine_profile_only(const char *prefix) { (void)prefix; }
EXPORT void tea_engine_profile_sample(int every) { (void)every; }
EXPORT void tea_engine_profile_reset(void) {}
EXPORT int tea_engine_profile_get(const char *k, long *count, double *ms) { (void)k; *count = 0; *ms = 0; return -1; }
EXPORT int tea_engine_profile_names(char *buf, int n) { if (buf && n > 0) buf[0] = 0; return 0; }
EXPORT void *tea_engine_stream(void) { return NULL; }
EXPORT void tea_engine_synchronize(void) {}
EXPORT void tea_engine_finalize(void) { free_all(); }

/* ---------------- mixed precision (no reference counterpart) ---------------------------- */

/* Fused PPCG outer iteration of the engine: the compositions it stands for */
EXPORT void tea_engine_ppcg_begin(int with_p, double beta, double *pw_rro)
{
    if (with_p) tea_leaf_ppcg_calc_p_kernel_cuda_(&beta);
    tea_leaf_cg_calc_w_kernel_cuda_(&pw_rro[0]);
    tea_leaf_ppcg_store_r_kernel_cuda_();
    tea_leaf_ppcg_calc_2norm_kernel_cuda_(&pw_rro[1]);
}

EXPORT void tea_engine_ppcg_ur(double alpha, double theta, double *rrn)
{
    tea_leaf_cg_calc_ur_kernel_cuda_(&alpha, rrn);
    tea_leaf_ppcg_init_sd_new_kernel_cuda_(&theta);
}

EXPORT void tea_engine_ppcg_begin_ur(int with_p, double beta, double theta, double *pw_rro, double *rrn)
{
    tea_engine_ppcg_begin(with_p, beta, pw_rro);
    tea_engine_ppcg_ur(pw_rro[1] / pw_rro[0], theta, rrn);
}

/* The fused CG iteration of the engine, as the composition it names:
 * [cg_calc_p(beta)], cg_calc_w (pw), cg_calc_ur(rro / pw) (rrn). */
EXPORT void tea_engine_cg_begin_ur(int with_p, double beta, double rro, double *pw, double *rrn)
{
    if (with_p) tea_leaf_cg_calc_p_kernel_cuda_(&beta);
    tea_leaf_cg_calc_w_kernel_cuda_(pw);
    double alpha = rro / *pw;
    tea_leaf_cg_calc_ur_kernel_cuda_(&alpha, rrn);
}

EXPORT void tea_engine_ppcg_end(double *rrn);
EXPORT void tea_engine_ppcg_end_begin_ur(double rro_prev, double theta, double *rrn_end, double *pw_rro,
                                         double *rrn)
{
    tea_engine_ppcg_end(rrn_end);
    tea_engine_ppcg_begin_ur(1, *rrn_end / rro_prev, theta, pw_rro, rrn);
}

EXPORT void tea_engine_ppcg_inner_n(int step, int nsteps);
EXPORT void tea_engine_ppcg_end_begin_ur_inner(double rro_prev, double theta, int inner_steps, double *rrn_end,
                                               double *pw_rro, double *rrn)
{
    tea_engine_ppcg_end_begin_ur(rro_prev, theta, rrn_end, pw_rro, rrn);
    if (inner_steps > 0) tea_engine_ppcg_inner_n(1, inner_steps);
}

EXPORT void tea_engine_ppcg_end(double *rrn)
{
    tea_leaf_ppcg_update_z_kernel_cuda_();
    tea_leaf_ppcg_calc_rrn_kernel_cuda_(rrn);
}

/* The engine's whole outer iteration, as the composition it names. */
EXPORT void tea_engine_ppcg_iteration(int with_p, double beta, double theta, int inner_steps, double *pw_rro,
                                      double *rrn_ur, double *rrn_end)
{
    tea_engine_ppcg_begin_ur(with_p, beta, theta, pw_rro, rrn_ur);
    if (inner_steps > 0) tea_engine_ppcg_inner_n(1, inner_steps);
    tea_engine_ppcg_end(rrn_end);
}

EXPORT void tea_engine_ppcg_inner_n(int step, int nsteps)
{
    int neighbours[4] = {EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE};
    for (int k = 0; k < nsteps; k++) {
        int be = 0, s = step + k;
        tea_leaf_ppcg_inner_kernel_cuda_(&s, &be, neighbours);
    }
}

EXPORT void tea_engine_ppcg_inner2(int step) { tea_engine_ppcg_inner_n(step, 2); }

/* Any grouping of the steps gives the same results; one step per call here. */
EXPORT int tea_engine_ppcg_inner_plan(int remaining) { return remaining > 0 ? 1 : 0; }

EXPORT int tea_engine_set_precision(int bits)
{
    if (!g.initialised) { die("%s called before initialise_cuda_", __func__); return -1; }
    if (bits != 32 && bits != 64) { die("tea_engine_set_precision: %d bits is not 32 or 64", bits); return -1; }
    if (bits == 32 && !g.f32[1]) {
        const size_t n2 = (size_t)g.pitch * (size_t)g.rows;
        for (int k = 1; k < X_COUNT; k++) g.f32[k] = (float *)calloc(n2, sizeof(float));
        g.zl32 = (float *)calloc(n2, sizeof(float));
    }
    g.fp32 = bits == 32;
    return 0;
}

/* Nothing is owed here: every field is stored as it is formed. */
EXPORT int tea_engine_mp_leave(void) { return tea_engine_set_precision(64); }

/* FP32 copies of the face coefficients, whole arrays */
EXPORT void tea_engine_mp_setup(void)
{
    CHECK_INIT();
    if (!g.f32[1]) { die("tea_engine_mp_setup: call tea_engine_set_precision(32) first"); return; }
    const size_t n2 = (size_t)g.pitch * (size_t)g.rows;
    for (size_t i = 0; i < n2; i++) { g.f32[X_KX][i] = (float)KX[i]; g.f32[X_KY][i] = (float)KY[i]; }
}

/* r32 = (float) r, the FP32 right-hand side u0_32 = r32 too (Chebyshev
 * reads it), u32 = 0 (the correction), rr = r32.r32 */
EXPORT void tea_engine_mp_begin(double *rr)
{
    CHECK_INIT();
    if (!g.f32[1]) { die("tea_engine_mp_begin: call tea_engine_set_precision(32) first"); return; }
    float *R32 = g.f32[F_R], *U32 = g.f32[F_U], *U032 = g.f32[X_U0];
    double s = 0;
    REDUCE_EXTENT(s, IX0, IX1, IY0, IY1, {
        R32[i] = (float)R[i];
        U032[i] = R32[i];
        U32[i] = 0.0f;
        ta = R32[i]; tb = R32[i];
    });
    *rr = s;
}

/* u += (double) u32 */
EXPORT void tea_engine_mp_end(void)
{
    CHECK_INIT();
    if (!g.f32[1]) { die("tea_engine_mp_end: call tea_engine_set_precision(32) first"); return; }
    const float *U32 = g.f32[F_U];
    FOR_EXTENT(IX0, IX1, IY0, IY1, { U[i] += (double)U32[i]; });
}

/* tea_engine_mp_replace: the calls it stands for, in order */
EXPORT void tea_engine_mp_replace(double theta, double *rr64, double *rr32)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (!g.fp32 || !g.f32[1]) { die("tea_engine_mp_replace: call from the FP32 solver mode"); return; }
    int one = 1, two = 2;
    double unused = 0.0;
    tea_engine_mp_end();
    tea_engine_set_precision(64);
    tea_leaf_calc_residual_cuda_();
    tea_leaf_calc_2norm_kernel_cuda_(&one, rr64);
    tea_engine_set_precision(32);
    tea_engine_mp_begin(rr32);
    tea_leaf_ppcg_init_kernel_cuda_(&two, &unused);
    tea_leaf_ppcg_init_sd_new_kernel_cuda_(&theta);
}

EXPORT int tea_engine_get_field32(int id, float *host)
{
    if (!g.initialised || !g.f32[1] || id < 1 || id >= X_COUNT) return -1;
    memcpy(host, g.f32[id], (size_t)g.pitch * (size_t)g.rows * sizeof(float));
    return 0;
}

EXPORT int tea_engine_set_field32(int id, const float *host)
{
    if (!g.initialised || !g.f32[1] || id < 1 || id >= X_COUNT) return -1;
    memcpy(g.f32[id], host, (size_t)g.pitch * (size_t)g.rows * sizeof(float));
    return 0;
}
