// engine_solver.hpp -- the CG / PPCG kernel family, generic in the field type.
//
// T = double is the reference's FP64 path; T = float is the inner solver of the
// mixed-precision refinement.  Arithmetic is done in T in the expression order
// of the reference kernels; dot products always accumulate in compensated FP64
// (a product of two floats is exact in FP64, so FP32 dot products come out
// correctly rounded too).  Scalars (alpha, beta, theta) arrive in FP64 and are
// rounded to T once per launch.
#pragma once
#include "engine_device.hpp"

namespace tea {

// Loop over this workgroup's tiles; the body sees the lane's strip s and i0,
// the flat index of the strip's first cell.
#define STRIP_BEGIN                                                    \
    for (TileLoop tl_(e); tl_.more(); tl_.advance()) {                 \
        const Strip s = tl_.strip();                                   \
        const long i0 = (long)s.c + (long)s.r0 * pitch;                \
        (void)i0;
#define STRIP_END }

// Generic a = b copy (set_field_kernel.cuknl:5-16, ppcg store_r / update_z,
// jacobi copy_u).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_copy(Ext e, long pitch, T *__restrict__ dst, const T *__restrict__ src)
{
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) dst[i0 + k * pitch] = src[i0 + k * pitch];
    STRIP_END
}

// tea_leaf_common.cuknl:41-58: sum a*b
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_dot(Ext e, long pitch, const T *__restrict__ a, const T *__restrict__ b,
                                               Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) dd_add_prod(acc[0], (double)a[i0 + k * pitch], (double)b[i0 + k * pitch]);
    STRIP_END
    block_reduce<1>(acc, red);
}

// tea_leaf_common.cuknl:131-147 (interior): Mi = 1/diag
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_jac_diag(Ext e, long pitch, T *__restrict__ Mi, const T *__restrict__ Kx,
                                                    const T *__restrict__ Ky)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const T diag = diag_A(Kx, Ky, i, pitch);
        Mi[i] = T(1.0) / diag;
    }
    STRIP_END
}

// tea_block_jacobi.cuknl:83-111: Thomas factors cp, bfp per strip (kept for the
// fields the reference exposes; the solves recompute them in registers)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_block_init(Ext e, long pitch, T *__restrict__ cp, T *__restrict__ bfp,
                                                      const T *__restrict__ Kx, const T *__restrict__ Ky)
{
    STRIP_BEGIN
    if (s.len == 0) continue;
    cp[i0] = (-Ky[i0 + pitch]) / diag_A(Kx, Ky, i0, pitch);
    for (int k = 1; k < s.len; k++) {
        const long i = i0 + k * pitch;
        bfp[i] = T(1.0) / (diag_A(Kx, Ky, i, pitch) - (-Ky[i]) * cp[i - pitch]);
        cp[i] = (-Ky[i + pitch]) * bfp[i];
    }
    STRIP_END
}

// z = M^-1 r on a strip held in registers (block: Thomas; diag: r*Mi).
template <typename T, int PREC>
__device__ __forceinline__ void precond_strip(const T (&rv)[STRIP], T (&zv)[STRIP], int len, long i0, long pitch,
                                              const T *__restrict__ Mi, const T *__restrict__ Kx,
                                              const T *__restrict__ Ky)
{
    if (PREC == PREC_BLOCK) {
        if (len > 0) block_solve_strip(rv, zv, len, Kx, Ky, i0, pitch);
    } else if (PREC == PREC_DIAG) {
#pragma unroll
        for (int k = 0; k < STRIP; k++)
            if (k < len) zv[k] = rv[k] * Mi[i0 + k * pitch];
    } else {
#pragma unroll
        for (int k = 0; k < STRIP; k++) zv[k] = rv[k];
    }
}

// ---- CG ---------------------------------------------------------------------------

// tea_leaf_cg.cuknl:5-39 (+ block_solve for jac_block): z = M^-1 r, p = z or r, sum r*p
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_cg_init_p(Ext e, long pitch, T *__restrict__ p, const T *__restrict__ r,
                                                     T *__restrict__ z, const T *__restrict__ Mi,
                                                     const T *__restrict__ Kx, const T *__restrict__ Ky, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) rv[k] = r[i0 + k * pitch];
    if (PREC != PREC_NONE) precond_strip<T, PREC>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            T pv;
            if (PREC != PREC_NONE) {
                z[i] = zv[k];
                pv = zv[k];
            } else {
                pv = rv[k];
            }
            p[i] = pv;
            dd_add_prod(acc[0], (double)rv[k], (double)pv);
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// tea_leaf_cg.cuknl:43-68: w = A p, sum p*w
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_cg_calc_w(Ext e, long pitch, const T *__restrict__ p, T *__restrict__ w,
                                                     const T *__restrict__ Kx, const T *__restrict__ Ky, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T wv = apply_A(p, Kx, Ky, i, pitch);
            w[i] = wv;
            dd_add_prod(acc[0], (double)p[i], (double)wv);
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// tea_leaf_cg.cuknl:72-137: u += a p; r -= a w; z = M^-1 r; sum r*z (r*r)
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_cg_calc_ur(Ext e, long pitch, double alpha_in, T *__restrict__ u,
                                                      const T *__restrict__ p, T *__restrict__ r,
                                                      const T *__restrict__ w, T *__restrict__ z,
                                                      const T *__restrict__ Mi, const T *__restrict__ Kx,
                                                      const T *__restrict__ Ky, Red red)
{
    const T alpha = (T)alpha_in;
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            u[i] += alpha * p[i];
            rv[k] = r[i] - alpha * w[i];
            r[i] = rv[k];
        }
    }
    if (PREC != PREC_NONE) precond_strip<T, PREC>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            if (PREC != PREC_NONE) {
                z[i0 + k * pitch] = zv[k];
                dd_add_prod(acc[0], (double)rv[k], (double)zv[k]);
            } else {
                dd_add_prod(acc[0], (double)rv[k], (double)rv[k]);
            }
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// tea_leaf_cg.cuknl:141-161 and tea_leaf_ppcg.cuknl:301-313: p = src + beta p
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_calc_p(Ext e, long pitch, double beta_in, T *__restrict__ p,
                                                  const T *__restrict__ src)
{
    const T beta = (T)beta_in;
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) {
            const long i = i0 + k * pitch;
            p[i] = src[i] + beta * p[i];
        }
    STRIP_END
}

// ---- PPCG -------------------------------------------------------------------------

// tea_leaf_ppcg.cuknl:319-420
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_init(Ext e, long pitch, T *__restrict__ p, const T *__restrict__ r,
                                                     T *__restrict__ z, const T *__restrict__ Mi,
                                                     const T *__restrict__ Kx, const T *__restrict__ Ky, int step,
                                                     Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) rv[k] = r[i0 + k * pitch];
    if (PREC != PREC_NONE && (step == 1 || step == 2)) precond_strip<T, PREC>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            if (step == 1) {
                // p = z = 0, then z = M^-1 r and p = z (or p = r)
                if (PREC != PREC_NONE) {
                    z[i] = zv[k];
                    p[i] = zv[k];
                } else {
                    z[i] = T(0.0);
                    p[i] = rv[k];
                }
                dd_add_prod(acc[0], (double)rv[k], (double)p[i]);
            } else if (step == 2) {
                if (PREC != PREC_NONE) z[i] = zv[k];
            } else {
                const T pv = z[i];
                p[i] = pv;
                dd_add_prod(acc[0], (double)rv[k], (double)pv);
            }
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// tea_leaf_ppcg.cuknl:5-55 (init_sd, not called by the reference driver)
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_init_sd(Ext e, long pitch, const T *__restrict__ r,
                                                        T *__restrict__ sd, const T *__restrict__ Mi,
                                                        const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                        double theta_in)
{
    const T theta = (T)theta_in;
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) rv[k] = r[i0 + k * pitch];
    if (PREC == PREC_BLOCK) precond_strip<T, PREC_BLOCK>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            if (PREC == PREC_BLOCK) sd[i] = zv[k] / theta;
            else if (PREC == PREC_DIAG) sd[i] = rv[k] * Mi[i] / theta;
            else sd[i] = rv[k] / theta;
        }
    }
    STRIP_END
}

// tea_leaf_ppcg.cuknl:58-89: sd = (z or r) * (1/theta); rtemp = r; utemp = sd
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_init_sd_new(Ext e, long pitch, const T *__restrict__ r,
                                                            T *__restrict__ sd, const T *__restrict__ src,
                                                            T *__restrict__ rtemp, T *__restrict__ utemp,
                                                            double theta_in)
{
    // theta_r = 1.0/theta is formed in the field type, as the reference's
    // shared theta_r is (tea_leaf_ppcg.cuknl:68-69)
    const T theta_r = T(1.0) / (T)theta_in;
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T sdv = src[i] * theta_r;
            sd[i] = sdv;
            rtemp[i] = r[i];
            utemp[i] = sdv;
        }
    }
    STRIP_END
}

// tea_leaf_ppcg.cuknl:154-173: rtemp -= A sd
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_update_r(Ext e, long pitch, T *__restrict__ rtemp,
                                                         const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                         const T *__restrict__ sd)
{
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T result = apply_A(sd, Kx, Ky, i, pitch);
            rtemp[i] -= result;
        }
    }
    STRIP_END
}

// tea_leaf_ppcg.cuknl:178-244: sd = alpha sd + beta M^-1 rtemp; utemp += sd
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_calc_sd(Ext e, long pitch, T *__restrict__ sd,
                                                        const T *__restrict__ rtemp, T *__restrict__ utemp,
                                                        const T *__restrict__ Mi, const T *__restrict__ Kx,
                                                        const T *__restrict__ Ky, double alpha_in, double beta_in)
{
    const T alpha = (T)alpha_in, beta = (T)beta_in;
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) rv[k] = rtemp[i0 + k * pitch];
    if (PREC == PREC_BLOCK) precond_strip<T, PREC_BLOCK>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            T sdv;
            if (PREC == PREC_BLOCK) sdv = alpha * sd[i] + beta * zv[k];
            else if (PREC == PREC_DIAG) sdv = alpha * sd[i] + beta * rv[k] * Mi[i];
            else sdv = alpha * sd[i] + beta * rv[k];
            sd[i] = sdv;
            utemp[i] += sdv;
        }
    }
    STRIP_END
}

// One PPCG inner step in one pass (tea_leaf_ppcg.cuknl:154-173 then :178-244):
// rtemp -= A sd; sd' = alpha sd + beta M^-1 rtemp; utemp += sd'.  sd is read
// from sd_in and the new sd written to sd_out, so every cell sees the previous
// step's sd exactly as the two separate launches of the reference do.
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_inner_fused(Ext e, long pitch, const T *__restrict__ sd_in,
                                                            T *__restrict__ sd_out, T *__restrict__ rtemp,
                                                            T *__restrict__ utemp, const T *__restrict__ Mi,
                                                            const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                            double alpha_in, double beta_in)
{
    const T alpha = (T)alpha_in, beta = (T)beta_in;
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T result = apply_A(sd_in, Kx, Ky, i, pitch);
            rv[k] = rtemp[i] - result;
            rtemp[i] = rv[k];
        }
    }
    if (PREC == PREC_BLOCK) precond_strip<T, PREC_BLOCK>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            T sdv;
            if (PREC == PREC_BLOCK) sdv = alpha * sd_in[i] + beta * zv[k];
            else if (PREC == PREC_DIAG) sdv = alpha * sd_in[i] + beta * rv[k] * Mi[i];
            else sdv = alpha * sd_in[i] + beta * rv[k];
            sd_out[i] = sdv;
            utemp[i] += sdv;
        }
    }
    STRIP_END
}

// Two PPCG inner steps in one pass (no reference counterpart: the composition
// of two k_ppcg_inner_fused launches with coefficients (a1, b1) then (a2, b2),
// bit-identical to them).  A workgroup of NW waves owns an output tile of
// WAVE-2 columns x (NW-2) strips; wave j takes strip j-1 of the tile (waves 0
// and NW-1 the strips below and above it), lane l column l-1 (lanes 0 and
// WAVE-1 the columns left and right of it).  Step 1 runs on the whole
// (WAVE) x (NW strips) region, so every sd value step 2 needs at the tile's
// cells is at hand: the strip's own rows in registers, the rows above and below
// through LDS, the columns left and right through lane shuffles.  Step 2 runs
// on the tile only.  rtemp is read from rt_in and written to rt_out because
// step 1 also reads it on the neighbouring tiles' cells; utemp is read and
// written on the tile's own cells only.  sd outside the extent keeps its value.
template <typename T, int PREC, int NW>
__global__ __launch_bounds__(NW * WAVE) void k_ppcg_inner2(Ext e, long pitch, const T *__restrict__ sd_in,
                                                          T *__restrict__ sd_out, const T *__restrict__ rt_in,
                                                          T *__restrict__ rt_out, T *__restrict__ utemp,
                                                          const T *__restrict__ Mi, const T *__restrict__ Kx,
                                                          const T *__restrict__ Ky, double a1_in, double b1_in,
                                                          double a2_in, double b2_in)
{
    constexpr int OUT_X = WAVE - 2, OUT_Y = (NW - 2) * STRIP;
    __shared__ T lds_lo[NW][WAVE], lds_hi[NW][WAVE];  // each strip's first and last row of sd after step 1
    const T a1 = (T)a1_in, b1 = (T)b1_in, a2 = (T)a2_in, b2 = (T)b2_in;
    const int lane = threadIdx.x & (WAVE - 1), j = threadIdx.x / WAVE;
    const int ntx = (e.x1 - e.x0 + OUT_X) / OUT_X, nty = (e.y1 - e.y0 + OUT_Y) / OUT_Y;
    const int ntiles = (e.x1 < e.x0 || e.y1 < e.y0) ? 0 : ntx * nty;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int ty = t / ntx, tx = t - ty * ntx;
        const int c = e.x0 + tx * OUT_X - 1 + lane;
        const int r0 = e.y0 + ty * OUT_Y + (j - 1) * STRIP;
        const long i0 = (long)c + (long)r0 * pitch;
        const bool col_in = c >= e.x0 && c <= e.x1;
        const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;   // the strip's rows inside the extent
        len = len < 0 ? 0 : (len > STRIP ? STRIP : len);
        const int act = col_in ? len : 0;
        const bool out = j >= 1 && j <= NW - 2 && lane >= 1 && lane <= WAVE - 2;

        // ---- step 1 on the whole region
        T sd1[STRIP], rt1[STRIP];
        {
            T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const long i = i0 + k * pitch;
                const int rr = r0 + k;
                sd1[k] = T(0.0);
                if (col_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1) sd1[k] = sd_in[i];
                if (k < act) {
                    const T result = apply_A(sd_in, Kx, Ky, i, pitch);
                    rv[k] = rt_in[i] - result;
                }
            }
            if (PREC == PREC_BLOCK && act > 0) precond_strip<T, PREC_BLOCK>(rv, zv, act, i0, pitch, Mi, Kx, Ky);
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                rt1[k] = rv[k];
                if (k < act) {
                    const long i = i0 + k * pitch;
                    if (PREC == PREC_BLOCK) sd1[k] = a1 * sd1[k] + b1 * zv[k];
                    else if (PREC == PREC_DIAG) sd1[k] = a1 * sd1[k] + b1 * rv[k] * Mi[i];
                    else sd1[k] = a1 * sd1[k] + b1 * rv[k];
                }
            }
        }
        lds_lo[j][lane] = sd1[0];
        lds_hi[j][lane] = sd1[STRIP - 1];
        __syncthreads();

        // ---- step 2 on the tile
        {
            T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
            const int act2 = out ? act : 0;
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const T left = __shfl_up(sd1[k], 1, WAVE);
                const T right = __shfl_down(sd1[k], 1, WAVE);
                if (k < act2) {
                    const long i = i0 + k * pitch;
                    const T below = k == 0 ? lds_hi[j - 1][lane] : sd1[k - 1];
                    const T above = k == STRIP - 1 ? lds_lo[j + 1][lane] : sd1[k + 1];
                    const T result = (T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i])) * sd1[k]
                                     - (Ky[i + pitch] * above + Ky[i] * below)
                                     - (Kx[i + 1] * right + Kx[i] * left);
                    rv[k] = rt1[k] - result;
                }
            }
            if (PREC == PREC_BLOCK && act2 > 0) precond_strip<T, PREC_BLOCK>(rv, zv, act2, i0, pitch, Mi, Kx, Ky);
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                if (k < act2) {
                    const long i = i0 + k * pitch;
                    T sd2;
                    if (PREC == PREC_BLOCK) sd2 = a2 * sd1[k] + b2 * zv[k];
                    else if (PREC == PREC_DIAG) sd2 = a2 * sd1[k] + b2 * rv[k] * Mi[i];
                    else sd2 = a2 * sd1[k] + b2 * rv[k];
                    const T ut1 = utemp[i] + sd1[k];
                    sd_out[i] = sd2;
                    rt_out[i] = rv[k];
                    utemp[i] = ut1 + sd2;
                }
            }
        }
        __syncthreads();
    }
}

// Block-Jacobi strip solve of one column from its own diagonal and Ky values
// (diag[k] = diag_A at row k, ky[k] = Ky at row k, k = 0..len): the
// expressions of block_solve_strip.
__device__ __forceinline__ void strip_solve_col(const float (&rhs)[STRIP], float (&z)[STRIP], int len,
                                                const float (&diag)[STRIP], const float (&ky)[STRIP + 1])
{
    float cp[STRIP], bfp[STRIP], dp[STRIP];
    cp[0] = (-ky[1]) / diag[0];
    dp[0] = rhs[0] / diag[0];
#pragma unroll
    for (int k = 1; k < STRIP; k++) {
        if (k < len) {
            bfp[k] = 1.0f / (diag[k] - (-ky[k]) * cp[k - 1]);
            cp[k] = (-ky[k + 1]) * bfp[k];
            dp[k] = (rhs[k] - (-ky[k]) * dp[k - 1]) * bfp[k];
        }
    }
#pragma unroll
    for (int k = STRIP - 1; k >= 0; k--) {
        if (k == len - 1) z[k] = dp[k];
        else if (k < len - 1) z[k] = dp[k] - cp[k] * z[k + 1];
    }
}

// k_ppcg_inner2 for FP32 fields with two columns per lane (8-byte accesses:
// an FP32 row segment of a wave is then as long as an FP64 one).  Lane l
// holds columns X0-2+2l and X0-1+2l of an output tile of 2*WAVE-4 columns
// starting at X0 (lane 0's and lane WAVE-1's pairs are the halo); otherwise
// the same scheme and arithmetic as k_ppcg_inner2, bit-identical to two
// k_ppcg_inner_fused<float> launches.
template <int PREC, int NW>
__global__ __launch_bounds__(NW * WAVE) void k_ppcg_inner2_x2(Ext e, long pitch, const float *__restrict__ sd_in,
                                                             float *__restrict__ sd_out,
                                                             const float *__restrict__ rt_in,
                                                             float *__restrict__ rt_out, float *__restrict__ utemp,
                                                             const float *__restrict__ Mi,
                                                             const float *__restrict__ Kx,
                                                             const float *__restrict__ Ky, double a1_in,
                                                             double b1_in, double a2_in, double b2_in)
{
    constexpr int OUT_X = 2 * WAVE - 4, OUT_Y = (NW - 2) * STRIP;
    __shared__ float2 lds_lo[NW][WAVE], lds_hi[NW][WAVE];
    const float a1 = (float)a1_in, b1 = (float)b1_in, a2 = (float)a2_in, b2 = (float)b2_in;
    const int lane = threadIdx.x & (WAVE - 1), j = threadIdx.x / WAVE;
    const int ntx = (e.x1 - e.x0 + OUT_X) / OUT_X, nty = (e.y1 - e.y0 + OUT_Y) / OUT_Y;
    const int ntiles = (e.x1 < e.x0 || e.y1 < e.y0) ? 0 : ntx * nty;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int ty = t / ntx, tx = t - ty * ntx;
        const int c = e.x0 + tx * OUT_X - 2 + 2 * lane;   // the lane's first column
        const int r0 = e.y0 + ty * OUT_Y + (j - 1) * STRIP;
        const long i0 = (long)c + (long)r0 * pitch;
        const bool pair_mem = c >= e.x0 - 2 && c <= e.x1 + 1;
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        len = len < 0 ? 0 : (len > STRIP ? STRIP : len);
        int act[2];
        act[0] = (c >= e.x0 && c <= e.x1) ? len : 0;
        act[1] = (c + 1 >= e.x0 && c + 1 <= e.x1) ? len : 0;
        const bool out = j >= 1 && j <= NW - 2 && lane >= 1 && lane <= WAVE - 2;
        const bool right_lane = lane < WAVE - 1;

        // ---- loads: sd rows r0-1 .. r0+STRIP, rtemp, Kx, Ky (rows r0 .. r0+STRIP), utemp
        float sv[2][STRIP + 2], kxv[2][STRIP], kyv[2][STRIP + 1], rtv[2][STRIP], utv[2][STRIP], miv[2][STRIP];
#pragma unroll
        for (int q = 0; q < STRIP + 2; q++) {
            const int rr = r0 - 1 + q;
            float2 v = make_float2(0.0f, 0.0f);
            if (pair_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1) v = *(const float2 *)(sd_in + i0 + (long)(q - 1) * pitch);
            sv[0][q] = v.x;
            sv[1][q] = v.y;
        }
        const int lanelen = act[0] > act[1] ? act[0] : act[1];
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) {
            float2 v = make_float2(0.0f, 0.0f);
            if (k <= lanelen && lanelen > 0) v = *(const float2 *)(Ky + i0 + (long)k * pitch);
            kyv[0][k] = v.x;
            kyv[1][k] = v.y;
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            float2 kx = make_float2(0.0f, 0.0f), rt = make_float2(0.0f, 0.0f), ut = make_float2(0.0f, 0.0f),
                   mi = make_float2(0.0f, 0.0f);
            if (k < lanelen) {
                const long i = i0 + (long)k * pitch;
                kx = *(const float2 *)(Kx + i);
                rt = *(const float2 *)(rt_in + i);
                if (out) ut = *(const float2 *)(utemp + i);
                if (PREC == PREC_DIAG) mi = *(const float2 *)(Mi + i);
            }
            kxv[0][k] = kx.x; kxv[1][k] = kx.y;
            rtv[0][k] = rt.x; rtv[1][k] = rt.y;
            utv[0][k] = ut.x; utv[1][k] = ut.y;
            miv[0][k] = mi.x; miv[1][k] = mi.y;
        }
        // Kx on the right face of the second column (the next lane's first Kx)
        float kxr[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const float sh = __shfl_down(kxv[0][k], 1, WAVE);
            kxr[k] = 0.0f;
            if (k < act[1]) kxr[k] = (right_lane && c + 2 <= e.x1) ? sh : Kx[i0 + (long)k * pitch + 2];
        }
        float diag[2][STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            diag[0][k] = 1.0f + (kyv[0][k + 1] + kyv[0][k]) + (kxv[1][k] + kxv[0][k]);
            diag[1][k] = 1.0f + (kyv[1][k + 1] + kyv[1][k]) + (kxr[k] + kxv[1][k]);
        }

        // ---- step 1 on the whole region
        float sd1[2][STRIP], rt1[2][STRIP];
        {
            float rv[2][STRIP], zv[2][STRIP];
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const float from_left = __shfl_up(sv[1][k + 1], 1, WAVE);
                const float from_right = __shfl_down(sv[0][k + 1], 1, WAVE);
                const long i = i0 + (long)k * pitch;
                rv[0][k] = rv[1][k] = zv[0][k] = zv[1][k] = 0.0f;
                sd1[0][k] = sv[0][k + 1];
                sd1[1][k] = sv[1][k + 1];
                if (k < act[0]) {
                    const float left = lane > 0 ? from_left : sd_in[i - 1];
                    const float res = diag[0][k] * sv[0][k + 1] - (kyv[0][k + 1] * sv[0][k + 2] + kyv[0][k] * sv[0][k])
                                      - (kxv[1][k] * sv[1][k + 1] + kxv[0][k] * left);
                    rv[0][k] = rtv[0][k] - res;
                }
                if (k < act[1]) {
                    const float right = right_lane ? from_right : sd_in[i + 2];
                    const float res = diag[1][k] * sv[1][k + 1] - (kyv[1][k + 1] * sv[1][k + 2] + kyv[1][k] * sv[1][k])
                                      - (kxr[k] * right + kxv[1][k] * sv[0][k + 1]);
                    rv[1][k] = rtv[1][k] - res;
                }
            }
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (PREC == PREC_BLOCK && act[q] > 0) {
                    float dq[STRIP], kq[STRIP + 1];
#pragma unroll
                    for (int k = 0; k < STRIP; k++) dq[k] = diag[q][k];
#pragma unroll
                    for (int k = 0; k < STRIP + 1; k++) kq[k] = kyv[q][k];
                    strip_solve_col(rv[q], zv[q], act[q], dq, kq);
                }
#pragma unroll
                for (int k = 0; k < STRIP; k++) {
                    rt1[q][k] = rv[q][k];
                    if (k < act[q]) {
                        if (PREC == PREC_BLOCK) sd1[q][k] = a1 * sd1[q][k] + b1 * zv[q][k];
                        else if (PREC == PREC_DIAG) sd1[q][k] = a1 * sd1[q][k] + b1 * rv[q][k] * miv[q][k];
                        else sd1[q][k] = a1 * sd1[q][k] + b1 * rv[q][k];
                    }
                }
            }
        }
        lds_lo[j][lane] = make_float2(sd1[0][0], sd1[1][0]);
        lds_hi[j][lane] = make_float2(sd1[0][STRIP - 1], sd1[1][STRIP - 1]);
        __syncthreads();

        // ---- step 2 on the tile
        {
            const int act2[2] = {out ? act[0] : 0, out ? act[1] : 0};
            float rv[2][STRIP], zv[2][STRIP];
            const float2 lo_below = lds_hi[j > 0 ? j - 1 : 0][lane];
            const float2 hi_above = lds_lo[j < NW - 1 ? j + 1 : NW - 1][lane];
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const float from_left = __shfl_up(sd1[1][k], 1, WAVE);
                const float from_right = __shfl_down(sd1[0][k], 1, WAVE);
                rv[0][k] = rv[1][k] = zv[0][k] = zv[1][k] = 0.0f;
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    if (k < act2[q]) {
                        const float below = k == 0 ? (q ? lo_below.y : lo_below.x) : sd1[q][k - 1];
                        const float above = k == STRIP - 1 ? (q ? hi_above.y : hi_above.x) : sd1[q][k + 1];
                        const float left = q ? sd1[0][k] : from_left;
                        const float right = q ? from_right : sd1[1][k];
                        const float kr = q ? kxr[k] : kxv[1][k];
                        const float res = diag[q][k] * sd1[q][k] - (kyv[q][k + 1] * above + kyv[q][k] * below)
                                          - (kr * right + kxv[q][k] * left);
                        rv[q][k] = rt1[q][k] - res;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (PREC == PREC_BLOCK && act2[q] > 0) {
                    float dq[STRIP], kq[STRIP + 1];
#pragma unroll
                    for (int k = 0; k < STRIP; k++) dq[k] = diag[q][k];
#pragma unroll
                    for (int k = 0; k < STRIP + 1; k++) kq[k] = kyv[q][k];
                    strip_solve_col(rv[q], zv[q], act2[q], dq, kq);
                }
            }
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                float sdo[2], uto[2];
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    if (PREC == PREC_BLOCK) sdo[q] = a2 * sd1[q][k] + b2 * zv[q][k];
                    else if (PREC == PREC_DIAG) sdo[q] = a2 * sd1[q][k] + b2 * rv[q][k] * miv[q][k];
                    else sdo[q] = a2 * sd1[q][k] + b2 * rv[q][k];
                    const float ut1 = utv[q][k] + sd1[q][k];
                    uto[q] = ut1 + sdo[q];
                }
                const long i = i0 + (long)k * pitch;
                if (k < act2[0] && k < act2[1]) {
                    *(float2 *)(sd_out + i) = make_float2(sdo[0], sdo[1]);
                    *(float2 *)(rt_out + i) = make_float2(rv[0][k], rv[1][k]);
                    *(float2 *)(utemp + i) = make_float2(uto[0], uto[1]);
                } else {
#pragma unroll
                    for (int q = 0; q < 2; q++)
                        if (k < act2[q]) {
                            sd_out[i + q] = sdo[q];
                            rt_out[i + q] = rv[q][k];
                            utemp[i + q] = uto[q];
                        }
                }
            }
        }
        __syncthreads();
    }
}

// tea_leaf_ppcg.cuknl:279-297: sum (r - r_store)*z
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_calc_rrn(Ext e, long pitch, const T *__restrict__ rstore,
                                                    const T *__restrict__ r, const T *__restrict__ z, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) {
            const long i = i0 + k * pitch;
            dd_add_prod(acc[0], (double)(r[i] - rstore[i]), (double)z[i]);
        }
    STRIP_END
    block_reduce<1>(acc, red);
}

// ---- fused PPCG outer iteration (no reference counterpart: each kernel below is
// the composition of reference kernels named in its comment, with their
// expressions unchanged, so results are bit-identical to the separate launches)

// [ppcg_calc_p: p' = z + beta p] (with_p), cg_calc_w: w = A p', p'.w,
// ppcg_store_r: rstore = r, ppcg_calc_2norm: r.z.  p' goes to the spare
// buffer pn (the caller swaps it in): p is read at neighbours that other
// lanes rewrite.  Cells outside the interior keep their p, as calc_p only
// updates the interior.  Each lane forms p' down its own column (the strip
// and the rows above and below it) and takes its x neighbours from the
// adjacent lanes; the wave's edge lanes form those themselves.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_begin(Ext e, long pitch, int with_p, double beta_in,
                                                      const T *__restrict__ p, T *__restrict__ pn,
                                                      const T *__restrict__ z, T *__restrict__ w,
                                                      const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                      const T *__restrict__ r, T *__restrict__ rstore, Red red)
{
    const T beta = (T)beta_in;
    const int lane = threadIdx.x & (WAVE - 1);
    DD acc[2] = {{0, 0}, {0, 0}};
    auto pv = [&](long j, int c, int rr) -> T {
        const bool in = c >= e.x0 && c <= e.x1 && rr >= e.y0 && rr <= e.y1;
        return (with_p && in) ? z[j] + beta * p[j] : p[j];
    };
    STRIP_BEGIN
    T col[STRIP + 2];
#pragma unroll
    for (int q = 0; q < STRIP + 2; q++) {
        col[q] = T(0.0);
        if (s.len > 0 && q <= s.len + 1) col[q] = pv(i0 + (long)(q - 1) * pitch, s.c, s.r0 - 1 + q);
    }
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        const T from_left = __shfl_up(col[k + 1], 1, WAVE);
        const T from_right = __shfl_down(col[k + 1], 1, WAVE);
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const int c = s.c, rr = s.r0 + k;
            const T left = lane > 0 ? from_left : pv(i - 1, c - 1, rr);
            const T right = (lane < WAVE - 1 && c + 1 <= e.x1) ? from_right : pv(i + 1, c + 1, rr);
            const T pc = col[k + 1];
            const T wv = (T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i])) * pc
                         - (Ky[i + pitch] * col[k + 2] + Ky[i] * col[k])
                         - (Kx[i + 1] * right + Kx[i] * left);
            pn[i] = pc;
            w[i] = wv;
            const T rv = r[i];
            rstore[i] = rv;
            dd_add_prod(acc[0], (double)pc, (double)wv);
            dd_add_prod(acc[1], (double)rv, (double)z[i]);
        }
    }
    STRIP_END
    block_reduce<2>(acc, red);
}

// k_ppcg_begin for FP32 fields with two columns per lane (8-byte accesses);
// tiles of 2*TILE_X x TILE_Y cells.  The same arithmetic cell by cell.
__global__ __launch_bounds__(BLOCK) void k_ppcg_begin_x2(Ext e, long pitch, int with_p, double beta_in,
                                                         const float *__restrict__ p, float *__restrict__ pn,
                                                         const float *__restrict__ z, float *__restrict__ w,
                                                         const float *__restrict__ Kx, const float *__restrict__ Ky,
                                                         const float *__restrict__ r, float *__restrict__ rstore,
                                                         Red red)
{
    const float beta = (float)beta_in;
    const int lane = threadIdx.x & (WAVE - 1);
    DD acc[2] = {{0, 0}, {0, 0}};
    auto inside = [&](int c, int rr) { return c >= e.x0 && c <= e.x1 && rr >= e.y0 && rr <= e.y1; };
    auto pv = [&](long j, int c, int rr) -> float { return (with_p && inside(c, rr)) ? z[j] + beta * p[j] : p[j]; };
    for (TileLoopT<2 * TILE_X> tl_(e); tl_.more(); tl_.advance()) {
        const Strip s = tl_.strip();
        const int c = s.c;
        const long i0 = (long)c + (long)s.r0 * pitch;
        const int act0 = s.len, act1 = (c + 1 <= e.x1) ? s.len : 0;
        // p' down both columns, rows r0-1 .. r0+len; z on the strip's rows
        float col[2][STRIP + 2], zs[2][STRIP];
#pragma unroll
        for (int q = 0; q < STRIP + 2; q++) {
            col[0][q] = col[1][q] = 0.0f;
            if (act0 > 0 && q <= act0 + 1) {
                const long j = i0 + (long)(q - 1) * pitch;
                const int rr = s.r0 - 1 + q;
                const float2 pp = *(const float2 *)(p + j);
                const float2 zz = *(const float2 *)(z + j);
                col[0][q] = (with_p && inside(c, rr)) ? zz.x + beta * pp.x : pp.x;
                col[1][q] = (with_p && inside(c + 1, rr)) ? zz.y + beta * pp.y : pp.y;
                if (q >= 1 && q <= STRIP) {
                    zs[0][q - 1] = zz.x;
                    zs[1][q - 1] = zz.y;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const float from_left = __shfl_up(col[1][k + 1], 1, WAVE);
            const float from_right = __shfl_down(col[0][k + 1], 1, WAVE);
            if (k < act0) {
                const long i = i0 + (long)k * pitch;
                const int rr = s.r0 + k;
                const float2 kx = *(const float2 *)(Kx + i);
                const float2 ky0 = *(const float2 *)(Ky + i);
                const float2 ky1 = *(const float2 *)(Ky + i + pitch);
                const float2 rv = *(const float2 *)(r + i);
                const bool rlane = lane < WAVE - 1 && c + 2 <= e.x1;
                const float left = lane > 0 ? from_left : pv(i - 1, c - 1, rr);
                const float w0 = (1.0f + (ky1.x + ky0.x) + (kx.y + kx.x)) * col[0][k + 1]
                                 - (ky1.x * col[0][k + 2] + ky0.x * col[0][k])
                                 - (kx.y * col[1][k + 1] + kx.x * left);
                dd_add_prod(acc[0], (double)col[0][k + 1], (double)w0);
                dd_add_prod(acc[1], (double)rv.x, (double)zs[0][k]);
                if (k < act1) {
                    const float right = rlane ? from_right : pv(i + 2, c + 2, rr);
                    const float kxr = Kx[i + 2];
                    const float w1 = (1.0f + (ky1.y + ky0.y) + (kxr + kx.y)) * col[1][k + 1]
                                     - (ky1.y * col[1][k + 2] + ky0.y * col[1][k])
                                     - (kxr * right + kx.y * col[0][k + 1]);
                    *(float2 *)(pn + i) = make_float2(col[0][k + 1], col[1][k + 1]);
                    *(float2 *)(w + i) = make_float2(w0, w1);
                    *(float2 *)(rstore + i) = rv;
                    dd_add_prod(acc[0], (double)col[1][k + 1], (double)w1);
                    dd_add_prod(acc[1], (double)rv.y, (double)zs[1][k]);
                } else {
                    pn[i] = col[0][k + 1];
                    w[i] = w0;
                    rstore[i] = rv.x;
                }
            }
        }
    }
    block_reduce<2>(acc, red);
}

// k_ppcg_ur_sd for FP32 fields with two columns per lane.
template <int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_ur_sd_x2(Ext e, long pitch, double alpha_in, double theta_in,
                                                         float *__restrict__ u, const float *__restrict__ p,
                                                         float *__restrict__ r, const float *__restrict__ w,
                                                         float *__restrict__ z, const float *__restrict__ Mi,
                                                         const float *__restrict__ Kx, const float *__restrict__ Ky,
                                                         float *__restrict__ sd, float *__restrict__ rtemp,
                                                         float *__restrict__ utemp, Red red)
{
    const float alpha = (float)alpha_in;
    const float theta_r = 1.0f / (float)theta_in;
    DD acc[1] = {{0, 0}};
    for (TileLoopT<2 * TILE_X> tl_(e); tl_.more(); tl_.advance()) {
        const Strip s = tl_.strip();
        const int c = s.c;
        const long i0 = (long)c + (long)s.r0 * pitch;
        int act[2] = {s.len, (c + 1 <= e.x1) ? s.len : 0};
        float rv[2][STRIP], zv[2][STRIP], mi[2][STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            rv[0][k] = rv[1][k] = zv[0][k] = zv[1][k] = mi[0][k] = mi[1][k] = 0.0f;
            if (k < act[0]) {
                const long i = i0 + (long)k * pitch;
                const float2 pp = *(const float2 *)(p + i), ww = *(const float2 *)(w + i);
                float2 uu = *(const float2 *)(u + i), rr = *(const float2 *)(r + i);
                uu.x += alpha * pp.x;
                uu.y += alpha * pp.y;
                rv[0][k] = rr.x - alpha * ww.x;
                rv[1][k] = rr.y - alpha * ww.y;
                if (PREC == PREC_DIAG) {
                    const float2 m = *(const float2 *)(Mi + i);
                    mi[0][k] = m.x;
                    mi[1][k] = m.y;
                }
                if (k < act[1]) {
                    *(float2 *)(u + i) = uu;
                    *(float2 *)(r + i) = make_float2(rv[0][k], rv[1][k]);
                } else {
                    u[i] = uu.x;
                    r[i] = rv[0][k];
                }
            }
        }
        if (PREC == PREC_BLOCK && act[0] > 0) {
            float diag[2][STRIP], ky[2][STRIP + 1];
#pragma unroll
            for (int k = 0; k < STRIP + 1; k++) {
                ky[0][k] = ky[1][k] = 0.0f;
                if (k <= act[0]) {
                    const float2 v = *(const float2 *)(Ky + i0 + (long)k * pitch);
                    ky[0][k] = v.x;
                    ky[1][k] = v.y;
                }
            }
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                diag[0][k] = diag[1][k] = 1.0f;
                if (k < act[0]) {
                    const long i = i0 + (long)k * pitch;
                    const float2 kx = *(const float2 *)(Kx + i);
                    const float kxr = Kx[i + 2];
                    diag[0][k] = 1.0f + (ky[0][k + 1] + ky[0][k]) + (kx.y + kx.x);
                    diag[1][k] = 1.0f + (ky[1][k + 1] + ky[1][k]) + (kxr + kx.y);
                }
            }
#pragma unroll
            for (int q = 0; q < 2; q++)
                if (act[q] > 0) strip_solve_col(rv[q], zv[q], act[q], diag[q], ky[q]);
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            if (k < act[0]) {
                const long i = i0 + (long)k * pitch;
                float src[2], sdv[2];
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    if (PREC == PREC_BLOCK) src[q] = zv[q][k];
                    else if (PREC == PREC_DIAG) src[q] = rv[q][k] * mi[q][k];
                    else src[q] = rv[q][k];
                    sdv[q] = src[q] * theta_r;
                }
                const int nq = k < act[1] ? 2 : 1;
                for (int q = 0; q < nq; q++)
                    dd_add_prod(acc[0], (double)rv[q][k], (double)(PREC != PREC_NONE ? src[q] : rv[q][k]));
                if (nq == 2) {
                    if (PREC != PREC_NONE) *(float2 *)(z + i) = make_float2(src[0], src[1]);
                    *(float2 *)(sd + i) = make_float2(sdv[0], sdv[1]);
                    *(float2 *)(rtemp + i) = make_float2(rv[0][k], rv[1][k]);
                    *(float2 *)(utemp + i) = make_float2(sdv[0], sdv[1]);
                } else {
                    if (PREC != PREC_NONE) z[i] = src[0];
                    sd[i] = sdv[0];
                    rtemp[i] = rv[0][k];
                    utemp[i] = sdv[0];
                }
            }
        }
    }
    block_reduce<1>(acc, red);
}

// cg_calc_ur (u += a p; r -= a w; z = M^-1 r; r.z or r.r) then
// ppcg_init_sd_new (sd = (z or r) / theta; rtemp = r; utemp = sd).
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_ur_sd(Ext e, long pitch, double alpha_in, double theta_in,
                                                      T *__restrict__ u, const T *__restrict__ p,
                                                      T *__restrict__ r, const T *__restrict__ w,
                                                      T *__restrict__ z, const T *__restrict__ Mi,
                                                      const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                      T *__restrict__ sd, T *__restrict__ rtemp,
                                                      T *__restrict__ utemp, Red red)
{
    const T alpha = (T)alpha_in;
    const T theta_r = T(1.0) / (T)theta_in;
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            u[i] += alpha * p[i];
            rv[k] = r[i] - alpha * w[i];
            r[i] = rv[k];
        }
    }
    if (PREC != PREC_NONE) precond_strip<T, PREC>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            T src;
            if (PREC != PREC_NONE) {
                z[i] = zv[k];
                dd_add_prod(acc[0], (double)rv[k], (double)zv[k]);
                src = zv[k];
            } else {
                dd_add_prod(acc[0], (double)rv[k], (double)rv[k]);
                src = rv[k];
            }
            const T sdv = src * theta_r;
            sd[i] = sdv;
            rtemp[i] = rv[k];
            utemp[i] = sdv;
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// ppcg_update_z (z = utemp) then ppcg_calc_rrn ((r - rstore).z).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_end(Ext e, long pitch, const T *__restrict__ utemp,
                                                    T *__restrict__ z, const T *__restrict__ r,
                                                    const T *__restrict__ rstore, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T zv = utemp[i];
            z[i] = zv;
            dd_add_prod(acc[0], (double)(r[i] - rstore[i]), (double)zv);
        }
    STRIP_END
    block_reduce<1>(acc, red);
}

// ---- mixed precision glue (no reference counterpart) -------------------------------

// FP32 copies of the FP64 face coefficients over the whole array (halo included).
__global__ __launch_bounds__(BLOCK) void k_mp_coef(Ext e, long pitch, long pitch32, const double *__restrict__ Kx,
                                                   const double *__restrict__ Ky, float *__restrict__ Kx32,
                                                   float *__restrict__ Ky32)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const long j = (long)s.c + (long)(s.r0 + k) * pitch32;
        Kx32[j] = (float)Kx[i];
        Ky32[j] = (float)Ky[i];
    }
    STRIP_END
}

// r32 = (float) r64; u32 = 0; sum r32*r32 (compensated FP64).
__global__ __launch_bounds__(BLOCK) void k_mp_begin(Ext e, long pitch, long pitch32, const double *__restrict__ r,
                                                    float *__restrict__ r32, float *__restrict__ u32, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const long j = (long)s.c + (long)(s.r0 + k) * pitch32;
        const float rv = (float)r[i];
        r32[j] = rv;
        u32[j] = 0.0f;
        dd_add_prod(acc[0], (double)rv, (double)rv);
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// u64 += (double) u32
__global__ __launch_bounds__(BLOCK) void k_mp_end(Ext e, long pitch, long pitch32, double *__restrict__ u,
                                                  const float *__restrict__ u32)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const long j = (long)s.c + (long)(s.r0 + k) * pitch32;
        u[i] += (double)u32[j];
    }
    STRIP_END
}

}  // namespace tea
