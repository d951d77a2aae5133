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

#include <type_traits>

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

// z = M^-1 r (block Jacobi) and sd = z / theta, each where its pointer is not
// null: the values a fused ur (k_ppcg_ur_sd_wp, k_ppcg_ur_sd_x2_wp) forms and
// leaves unstored when the first inner pass forms them itself (no reference
// counterpart as a kernel; the same expressions, so the same values).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_minv_r(Ext e, long pitch, const T *__restrict__ r, T *__restrict__ z,
                                                  T *__restrict__ sd, double theta_in, const T *__restrict__ Kx,
                                                  const T *__restrict__ Ky)
{
    const T theta_r = T(1.0) / (T)theta_in;
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) rv[k] = r[i0 + k * pitch];
    precond_strip<T, PREC_BLOCK>(rv, zv, s.len, i0, pitch, (const T *)nullptr, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            if (z) z[i] = zv[k];
            if (sd) sd[i] = zv[k] * theta_r;
        }
    }
    STRIP_END
}

// z = M^-1 r with jac_diag, Mi = 1/diag formed from Kx and Ky with the
// expression Mi was set with (tea_leaf_common.cuknl:141-145): the z a fused
// CG iteration (k_cg_ur_wp) leaves unstored, written out for a reader (no
// reference counterpart as a kernel; the same values).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_minv_r_diag(Ext e, long pitch, const T *__restrict__ r, T *__restrict__ z,
                                                       const T *__restrict__ Kx, const T *__restrict__ Ky)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        z[i] = r[i] * (T(1.0) / diag_A(Kx, Ky, i, pitch));
    }
    STRIP_END
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

// Block-Jacobi strip solve of one column from its own diagonal and Ky values
// (diag[k] = diag_A at row k, ky[k] = Ky at row k, k = 0..len): the
// expressions of block_solve_strip.
template <typename T>
__device__ __forceinline__ void strip_solve_col(const T (&rhs)[STRIP], T (&z)[STRIP], int len,
                                                const T (&diag)[STRIP], const T (&ky)[STRIP + 1])
{
    T cp[STRIP], bfp[STRIP], dp[STRIP];
    cp[0] = (-ky[1]) / diag[0];
    dp[0] = rhs[0] / diag[0];
#pragma unroll
    for (int k = 1; k < STRIP; k++) {
        if (k < len) {
            bfp[k] = T(1.0) / (diag[k] - (-ky[k]) * cp[k - 1]);
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

// The two halves of strip_solve_col: the factors cp, bfp depend on the
// strip's diagonal and Ky only (the reference's block_init,
// tea_block_jacobi.cuknl:80-111), so a pass of several inner steps forms them
// once; strip_apply is the per-application part (block_solve_func,
// tea_block_jacobi.cuknl:9-42).  Same expressions, so the same values.
template <typename T>
__device__ __forceinline__ void strip_factor(int len, const T (&diag)[STRIP], const T (&ky)[STRIP + 1],
                                             T (&cp)[STRIP], T (&bfp)[STRIP])
{
    cp[0] = (-ky[1]) / diag[0];
    bfp[0] = T(0.0);
#pragma unroll
    for (int k = 1; k < STRIP; k++) {
        bfp[k] = cp[k] = T(0.0);
        if (k < len) {
            bfp[k] = T(1.0) / (diag[k] - (-ky[k]) * cp[k - 1]);
            cp[k] = (-ky[k + 1]) * bfp[k];
        }
    }
}

template <typename T>
__device__ __forceinline__ void strip_apply(const T (&rhs)[STRIP], T (&z)[STRIP], int len, T diag0,
                                            const T (&ky)[STRIP + 1], const T (&cp)[STRIP], const T (&bfp)[STRIP])
{
    T dp[STRIP];
    dp[0] = rhs[0] / diag0;
#pragma unroll
    for (int k = 1; k < STRIP; k++)
        if (k < len) dp[k] = (rhs[k] - (-ky[k]) * dp[k - 1]) * bfp[k];
#pragma unroll
    for (int k = STRIP - 1; k >= 0; k--) {
        if (k == len - 1) z[k] = dp[k];
        else if (k < len - 1) z[k] = dp[k] - cp[k] * z[k + 1];
    }
}

// strip_apply for FP32 with dp0 = rhs0 / diag0 formed as (float)(rhs0 * rd),
// rd = 1 / (double)diag0 rounded to double once per strip: the same float as
// the IEEE division.  The exact quotient of two floats lies at least about
// 2^-49 (relative) from every midpoint between floats (it is never one unless
// it is a float itself), and rhs0 * rd carries a relative error below 2^-52,
// so both round to the same float.
__device__ __forceinline__ void strip_apply_rd(const float (&rhs)[STRIP], float (&z)[STRIP], int len, double rd,
                                               const float (&ky)[STRIP + 1], const float (&cp)[STRIP],
                                               const float (&bfp)[STRIP])
{
    float dp[STRIP];
    dp[0] = (float)((double)rhs[0] * rd);
#pragma unroll
    for (int k = 1; k < STRIP; k++)
        if (k < len) dp[k] = (rhs[k] - (-ky[k]) * dp[k - 1]) * bfp[k];
#pragma unroll
    for (int k = STRIP - 1; k >= 0; k--) {
        if (k == len - 1) z[k] = dp[k];
        else if (k < len - 1) z[k] = dp[k] - cp[k] * z[k + 1];
    }
}

// Coefficients (alpha_k, beta_k) of consecutive inner steps, passed by value.
constexpr int MAX_PASS_STEPS = 5;
struct StepCoefs {
    double a[MAX_PASS_STEPS], b[MAX_PASS_STEPS];
    double theta;   // passes that form sd = M^-1 r / theta from r (FROM_R)
};

// Tile of a pass of NS inner steps with NW waves per workgroup: HL halo
// lanes at each side (X2: two columns per lane), NS-1 halo strips at the top
// and bottom; OUT_X x OUT_Y cells are written.
template <bool X2, int NW, int NS>
struct InnerTile {
    static constexpr int HL = X2 ? NS / 2 : NS - 1;
    static constexpr int HW = NS - 1;
    static constexpr int OUT_X = X2 ? 2 * (WAVE - 2 * HL) : WAVE - 2 * HL;
    static constexpr int OUT_Y = (NW - 2 * HW) * STRIP;
};

// NS = 1..5 PPCG inner steps in one pass (temporal blocking; no reference
// counterpart), FP32 fields, two columns per lane (8-byte accesses): the
// composition of NS k_ppcg_inner_fused<float> launches with coefficients
// (cf.a[k], cf.b[k]), bit-identical to them.  A workgroup of NW waves loads
// a region of 2*WAVE columns x NW strips (wave j the strip j, lane l the
// columns 2l and 2l+1; sd one row beyond the strip at each end, so the
// block-Jacobi strips are whole); step k (k = 1..NS) runs on that region minus k-1
// strips at the top and bottom and k-1 columns at each side, so the tile
// written at the end is NW - 2(NS-1) strips high and 2*WAVE - 4*(NS/2)
// columns wide.  Between steps each wave passes its strip's first and last
// rows of sd to the waves below and above through LDS (two buffers by step
// parity, one barrier per step); columns come from lane shuffles (memory at
// the wave's edge lanes in step 1, stale values there afterwards, which only
// reach cells outside the tile).  sd and rtemp are read from the _in and
// written to the _out buffers (the neighbouring tiles read them), utemp only
// on the tile's own cells.  Cells outside the extent keep their sd, as in the
// reference, where inner steps update the interior only.
template <int PREC, int NW, int NS>
__global__ __launch_bounds__(NW * WAVE) void k_ppcg_inner_n_x2(Ext e, long pitch, const float *__restrict__ sd_in,
                                                              float *__restrict__ sd_out,
                                                              const float *__restrict__ rt_in,
                                                              float *__restrict__ rt_out, float *__restrict__ utemp,
                                                              const float *__restrict__ Mi,
                                                              const float *__restrict__ Kx,
                                                              const float *__restrict__ Ky, StepCoefs cf)
{
    static_assert(NS >= 1 && NS <= MAX_PASS_STEPS && NW - 2 * (NS - 1) >= 1, "tile too small for NS steps");
    using Tile = InnerTile<true, NW, NS>;
    constexpr int HL = Tile::HL, HW = Tile::HW, OUT_X = Tile::OUT_X, OUT_Y = Tile::OUT_Y;
    __shared__ float2 lds[2][2][NW][WAVE];   // [step parity][first row, last row][wave][lane]
    const int lane = threadIdx.x & (WAVE - 1), j = threadIdx.x / WAVE;
    const int ntx = (e.x1 - e.x0 + OUT_X) / OUT_X, nty = (e.y1 - e.y0 + OUT_Y) / OUT_Y;
    const int ntiles = (e.x1 < e.x0 || e.y1 < e.y0) ? 0 : ntx * nty;
    const int jb = j > 0 ? j - 1 : 0, ja = j < NW - 1 ? j + 1 : NW - 1;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int ty = t / ntx, tx = t - ty * ntx;
        const int c = e.x0 + tx * OUT_X - 2 * HL + 2 * lane;   // the lane's first column
        const int r0 = e.y0 + ty * OUT_Y + (j - HW) * STRIP;
        const long i0 = (long)c + (long)r0 * pitch;
        const bool pair_mem = c >= e.x0 - 2 && c <= e.x1 + 1;
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        len = len < 0 ? 0 : (len > STRIP ? STRIP : len);
        int act[2];
        act[0] = (c >= e.x0 && c <= e.x1) ? len : 0;
        act[1] = (c + 1 >= e.x0 && c + 1 <= e.x1) ? len : 0;
        const bool out = j >= HW && j <= NW - 1 - HW && lane >= HL && lane <= WAVE - 1 - HL;
        const bool right_lane = lane < WAVE - 1;

        // ---- loads: sd rows r0-1 .. r0+STRIP, rtemp, Kx, Ky (rows r0 .. r0+STRIP), utemp
        float sv[2][STRIP + 2], kxv[2][STRIP], kyv[2][STRIP + 1], rtc[2][STRIP], utv[2][STRIP], miv[2][STRIP];
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
            rtc[0][k] = rt.x; rtc[1][k] = rt.y;
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
        float sdc[2][STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            sdc[0][k] = sv[0][k + 1];
            sdc[1][k] = sv[1][k + 1];
        }
        float cpf[2][STRIP], bff[2][STRIP];   // block-Jacobi factors, formed once for all steps
        if (PREC == PREC_BLOCK) {
#pragma unroll
            for (int q = 0; q < 2; q++)
                if (act[q] > 0) strip_factor(act[q], diag[q], kyv[q], cpf[q], bff[q]);
        }

#pragma unroll
        for (int st = 0; st < NS; st++) {
            const float a = (float)cf.a[st], b = (float)cf.b[st];
            const bool last = st == NS - 1;
            const int ak[2] = {(!last || out) ? act[0] : 0, (!last || out) ? act[1] : 0};
            float lo[2], hi[2];   // sd of the row below and above the strip
            if (st == 0) {
                lo[0] = sv[0][0]; lo[1] = sv[1][0];
                hi[0] = sv[0][STRIP + 1]; hi[1] = sv[1][STRIP + 1];
            } else {
                const float2 vb = lds[(st - 1) & 1][1][jb][lane], va = lds[(st - 1) & 1][0][ja][lane];
                lo[0] = vb.x; lo[1] = vb.y;
                hi[0] = va.x; hi[1] = va.y;
            }
            float rv[2][STRIP], zv[2][STRIP];
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const float from_left = __shfl_up(sdc[1][k], 1, WAVE);
                const float from_right = __shfl_down(sdc[0][k], 1, WAVE);
                const long i = i0 + (long)k * pitch;
                rv[0][k] = rv[1][k] = zv[0][k] = zv[1][k] = 0.0f;
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    if (k < ak[q]) {
                        const float below = k == 0 ? lo[q] : sdc[q][k - 1];
                        const float above = k == STRIP - 1 ? hi[q] : sdc[q][k + 1];
                        float left, right;
                        if (q == 0) {
                            left = (st > 0 || lane > 0) ? from_left : sd_in[i - 1];
                            right = sdc[1][k];
                        } else {
                            left = sdc[0][k];
                            right = (st > 0 || right_lane) ? from_right : sd_in[i + 2];
                        }
                        const float kr = q ? kxr[k] : kxv[1][k];
                        const float res = diag[q][k] * sdc[q][k] - (kyv[q][k + 1] * above + kyv[q][k] * below)
                                          - (kr * right + kxv[q][k] * left);
                        rv[q][k] = rtc[q][k] - res;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (PREC == PREC_BLOCK && ak[q] > 0) strip_apply(rv[q], zv[q], ak[q], diag[q][0], kyv[q], cpf[q], bff[q]);
#pragma unroll
                for (int k = 0; k < STRIP; k++) {
                    if (k < ak[q]) {
                        float sdn;
                        if (PREC == PREC_BLOCK) sdn = a * sdc[q][k] + b * zv[q][k];
                        else if (PREC == PREC_DIAG) sdn = a * sdc[q][k] + b * rv[q][k] * miv[q][k];
                        else sdn = a * sdc[q][k] + b * rv[q][k];
                        sdc[q][k] = sdn;
                        rtc[q][k] = rv[q][k];
                        utv[q][k] = utv[q][k] + sdn;
                    }
                }
            }
            if (!last) {
                lds[st & 1][0][j][lane] = make_float2(sdc[0][0], sdc[1][0]);
                lds[st & 1][1][j][lane] = make_float2(sdc[0][STRIP - 1], sdc[1][STRIP - 1]);
                __syncthreads();
            }
        }

        // ---- the tile's cells
        if (out) {
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const long i = i0 + (long)k * pitch;
                if (k < act[0] && k < act[1]) {
                    *(float2 *)(sd_out + i) = make_float2(sdc[0][k], sdc[1][k]);
                    *(float2 *)(rt_out + i) = make_float2(rtc[0][k], rtc[1][k]);
                    *(float2 *)(utemp + i) = make_float2(utv[0][k], utv[1][k]);
                } else {
#pragma unroll
                    for (int q = 0; q < 2; q++)
                        if (k < act[q]) {
                            sd_out[i + q] = sdc[q][k];
                            rt_out[i + q] = rtc[q][k];
                            utemp[i + q] = utv[q][k];
                        }
                }
            }
        }
        __syncthreads();
    }
}

// NS = 1..5 PPCG inner steps in one pass, one column per lane (any field
// type; used for FP64): the composition of NS k_ppcg_inner_fused<T>
// launches, bit-identical to them.  The scheme of k_ppcg_inner_n_x2: the
// tile written at the end is NW - 2(NS-1) strips high and WAVE - 2(NS-1)
// columns wide.
template <typename T, int PREC, int NW, int NS>
__global__ __launch_bounds__(NW * WAVE) void k_ppcg_inner_n(Ext e, long pitch, const T *__restrict__ sd_in,
                                                           T *__restrict__ sd_out, const T *__restrict__ rt_in,
                                                           T *__restrict__ rt_out, T *__restrict__ utemp,
                                                           const T *__restrict__ Mi, const T *__restrict__ Kx,
                                                           const T *__restrict__ Ky, StepCoefs cf)
{
    static_assert(NS >= 1 && NS <= MAX_PASS_STEPS && NW - 2 * (NS - 1) >= 1, "tile too small for NS steps");
    using Tile = InnerTile<false, NW, NS>;
    constexpr int HL = Tile::HL, HW = Tile::HW, OUT_X = Tile::OUT_X, OUT_Y = Tile::OUT_Y;
    __shared__ T lds[2][2][NW][WAVE];   // [step parity][first row, last row][wave][lane]
    const int lane = threadIdx.x & (WAVE - 1), j = threadIdx.x / WAVE;
    const int ntx = (e.x1 - e.x0 + OUT_X) / OUT_X, nty = (e.y1 - e.y0 + OUT_Y) / OUT_Y;
    const int ntiles = (e.x1 < e.x0 || e.y1 < e.y0) ? 0 : ntx * nty;
    const int jb = j > 0 ? j - 1 : 0, ja = j < NW - 1 ? j + 1 : NW - 1;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int ty = t / ntx, tx = t - ty * ntx;
        const int c = e.x0 + tx * OUT_X - HL + lane;
        const int r0 = e.y0 + ty * OUT_Y + (j - HW) * STRIP;
        const long i0 = (long)c + (long)r0 * pitch;
        const bool col_in = c >= e.x0 && c <= e.x1;
        const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        len = len < 0 ? 0 : (len > STRIP ? STRIP : len);
        const int act = col_in ? len : 0;
        const bool out = j >= HW && j <= NW - 1 - HW && lane >= HL && lane <= WAVE - 1 - HL;
        const bool right_lane = lane < WAVE - 1;

        // ---- loads: sd rows r0-1 .. r0+STRIP, rtemp, Kx, Ky (rows r0 .. r0+STRIP), utemp
        T sv[STRIP + 2], kxv[STRIP], kyv[STRIP + 1], rtc[STRIP], utv[STRIP], miv[STRIP];
#pragma unroll
        for (int q = 0; q < STRIP + 2; q++) {
            const int rr = r0 - 1 + q;
            sv[q] = T(0.0);
            if (col_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1) sv[q] = sd_in[i0 + (long)(q - 1) * pitch];
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) {
            kyv[k] = T(0.0);
            if (k <= act && act > 0) kyv[k] = Ky[i0 + (long)k * pitch];
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            kxv[k] = rtc[k] = utv[k] = miv[k] = T(0.0);
            if (k < act) {
                const long i = i0 + (long)k * pitch;
                kxv[k] = Kx[i];
                rtc[k] = rt_in[i];
                if (out) utv[k] = utemp[i];
                if (PREC == PREC_DIAG) miv[k] = Mi[i];
            }
        }
        // Kx on the cell's right face (the next lane's Kx)
        T kxr[STRIP], diag[STRIP], sdc[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const T sh = __shfl_down(kxv[k], 1, WAVE);
            kxr[k] = T(0.0);
            if (k < act) kxr[k] = (right_lane && c + 1 <= e.x1) ? sh : Kx[i0 + (long)k * pitch + 1];
            diag[k] = T(1.0) + (kyv[k + 1] + kyv[k]) + (kxr[k] + kxv[k]);
            sdc[k] = sv[k + 1];
        }
        // block-Jacobi factors formed once for all steps, where the registers
        // allow (more than three FP64 steps would spill)
        constexpr bool HOIST = sizeof(T) == 4 || NS <= 3;
        T cpf[STRIP], bff[STRIP];
        if (HOIST && PREC == PREC_BLOCK && act > 0) strip_factor<T>(act, diag, kyv, cpf, bff);

#pragma unroll
        for (int st = 0; st < NS; st++) {
            const T a = (T)cf.a[st], b = (T)cf.b[st];
            const bool last = st == NS - 1;
            const int ak = (!last || out) ? act : 0;
            T lo, hi;   // sd of the row below and above the strip
            if (st == 0) {
                lo = sv[0];
                hi = sv[STRIP + 1];
            } else {
                lo = lds[(st - 1) & 1][1][jb][lane];
                hi = lds[(st - 1) & 1][0][ja][lane];
            }
            T rv[STRIP], zv[STRIP];
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const T from_left = __shfl_up(sdc[k], 1, WAVE);
                const T from_right = __shfl_down(sdc[k], 1, WAVE);
                const long i = i0 + (long)k * pitch;
                rv[k] = zv[k] = T(0.0);
                if (k < ak) {
                    const T below = k == 0 ? lo : sdc[k - 1];
                    const T above = k == STRIP - 1 ? hi : sdc[k + 1];
                    const T left = (st > 0 || lane > 0) ? from_left : sd_in[i - 1];
                    const T right = (st > 0 || right_lane) ? from_right : sd_in[i + 1];
                    const T res = diag[k] * sdc[k] - (kyv[k + 1] * above + kyv[k] * below)
                                  - (kxr[k] * right + kxv[k] * left);
                    rv[k] = rtc[k] - res;
                }
            }
            if (PREC == PREC_BLOCK && ak > 0) {
                if (HOIST) strip_apply<T>(rv, zv, ak, diag[0], kyv, cpf, bff);
                else strip_solve_col<T>(rv, zv, ak, diag, kyv);
            }
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                if (k < ak) {
                    T sdn;
                    if (PREC == PREC_BLOCK) sdn = a * sdc[k] + b * zv[k];
                    else if (PREC == PREC_DIAG) sdn = a * sdc[k] + b * rv[k] * miv[k];
                    else sdn = a * sdc[k] + b * rv[k];
                    sdc[k] = sdn;
                    rtc[k] = rv[k];
                    utv[k] = utv[k] + sdn;
                }
            }
            if (!last) {
                lds[st & 1][0][j][lane] = sdc[0];
                lds[st & 1][1][j][lane] = sdc[STRIP - 1];
                __syncthreads();
            }
        }

        if (out) {
#pragma unroll
            for (int k = 0; k < STRIP; k++)
                if (k < act) {
                    const long i = i0 + (long)k * pitch;
                    sd_out[i] = sdc[k];
                    rt_out[i] = rtc[k];
                    utemp[i] = utv[k];
                }
        }
        __syncthreads();
    }
}

// ---- streaming pass of NS inner steps ------------------------------------------------

// A field seen through a raw buffer resource of `bytes` bytes (gfx9
// descriptor word 3: 0x00020000): an offset at or beyond `bytes` reads 0 and
// writes nothing, with no memory access.
constexpr unsigned OOB = 0xFFFFFFF0u;
template <typename T> struct FieldBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ FieldBuf(const T *base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000))
    {
    }
    __device__ __forceinline__ T load(unsigned off) const
    {
        if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
        else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    }
    __device__ __forceinline__ void store(unsigned off, T v) const
    {
        if constexpr (sizeof(T) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b64(
                __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, (int)off, 0, 0);
    }
};

// One block-Jacobi strip (4 rows of one column) as read from memory: the
// fields, the coefficients, and for lane 0 the left neighbour column's sd.
template <typename T> struct MarchRaw {
    T sd[STRIP], rt[STRIP], ut[STRIP], kx[STRIP], ky[STRIP + 1], mi[STRIP], sde[STRIP];
};
// A strip inside the pipeline: what it carries from step to step and the
// coefficients every step uses (the right-face Kx, the next lane's Kx, and
// the diagonal are re-formed per step); len: its rows inside the extent.
template <typename T> struct MarchStrip {
    T sd[STRIP], rt[STRIP], ut[STRIP], kx[STRIP], ky[STRIP + 1], cp[STRIP], bfp[STRIP], mi[STRIP];
    int r0, len;
};

constexpr int MARCH_WAVES = 4;   // independent waves per workgroup
// Columns a streaming pass of ns steps writes per band of WAVE lanes: ns - 1
// halo lanes at the left (lane 0 reads its left neighbour's step-0 sd from
// memory) and ns at the right (the last lane lacks its right neighbour's Kx,
// so its values are stale from the first step on).  Where the pass forms sd
// from r itself (from_r), one more on each side: lane 0 has no neighbour's sd
// to read, and the last lane's own sd is formed with the missing Kx.
constexpr int march_out_x(int ns, bool from_r = false) { return WAVE - 2 * ns + 1 - (from_r ? 2 : 0); }

// Whole-wave lane shifts by one (DPP wave_shr:1 / wave_shl:1, VALU moves with
// no LDS round trip): lane l receives v from lane l-1 (from_below) or l+1
// (from_above); the end lane (lane 0, resp. the last lane), which has no
// source, receives 0 (bound_ctrl), so no copy of v has to be made first.
// The streaming pass never reads those end values.
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ float lane_from_below(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
__device__ __forceinline__ float lane_from_above(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
__device__ __forceinline__ double lane_from_below(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shr1((int)(b & 0xffffffffLL)), hi = dpp_shr1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lane_from_above(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shl1((int)(b & 0xffffffffLL)), hi = dpp_shl1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// NS = 1..5 PPCG inner steps in one streaming pass (temporal blocking; no
// reference counterpart): the composition of NS k_ppcg_inner_fused<T>
// launches with coefficients (cf.a[k], cf.b[k]), bit-identical to them.
// Each wave owns a band of WAVE columns (one per lane; NS-1 halo lanes at the
// left and NS at the right, so WAVE - 2NS + 1 columns are written) and a
// chunk of `chunk` strips,
// and walks down it one strip per iteration: at iteration i strip i enters
// (read one iteration ahead) and strip i-1-m moves from step m to step m+1
// for m = 0..NS-1, so strip i-NS leaves the pipeline finished.  Step m+1 of a
// strip needs step m of the strip itself and of the rows just below and above
// it (the previous strip's last row, kept from the iteration before, and the
// next strip's first row, formed earlier in this iteration); columns come from
// lane shifts (memory at lane 0 in the first step, stale values at the edge
// lanes afterwards, which only reach halo lanes; the last lane has no right
// neighbour's Kx, so its values are stale from the first step, hence one more
// halo lane at the right).  The strips in the pipeline move
// down one place per iteration (register moves: a ring of slots unrolled NS+1
// times needs more registers than the 256 that keep two waves per SIMD).
// Reads and writes go through bounded buffer resources, so the guards of the
// reference semantics (rows and columns outside the extent, halo lanes) are
// offsets, not branches.  The NS strips before the
// chunk and the NS after it are read and advanced only as far as the chunk's
// strips need.  No barriers: the waves are independent.  sd, rtemp and utemp
// are read from the _in and written to the _out buffers (utemp's may be one
// buffer: a cell's utemp is written only by the wave that owns the cell, after
// that wave has read it); cells
// outside the extent keep their sd.
template <typename T, int PREC, int NS>
__global__ __launch_bounds__(MARCH_WAVES *WAVE) void k_ppcg_inner_march(
    Ext e, long pitch, int chunk, const T *__restrict__ sd_in, T *__restrict__ sd_out, const T *__restrict__ rt_in,
    T *__restrict__ rt_out, const T *ut_in, T *ut_out, const T *__restrict__ Mi, const T *__restrict__ Kx,
    const T *__restrict__ Ky, StepCoefs cf)
{
    static_assert(NS >= 1 && NS <= MAX_PASS_STEPS, "1..5 steps per pass");
    constexpr int HL = NS - 1, HR = NS, OUT_X = march_out_x(NS);
    static_assert(OUT_X == WAVE - HL - HR, "band layout");
    const int lane = threadIdx.x & (WAVE - 1);
    if (e.x1 < e.x0 || e.y1 < e.y0) return;
    const int nbands = (e.x1 - e.x0 + OUT_X) / OUT_X;
    const int nstrips = (e.y1 - e.y0 + STRIP) / STRIP;
    const int nchunks = (nstrips + chunk - 1) / chunk;
    // the wave's index, known to the compiler as uniform across the wave, so
    // the band, chunk and strip bookkeeping and the loop stay scalar
    const int w = blockIdx.x * MARCH_WAVES + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    if (w >= nbands * nchunks) return;
    const int band = w % nbands, ck = w / nbands;
    const int c = e.x0 + band * OUT_X - HL + lane;
    const bool col_in = c >= e.x0 && c <= e.x1;
    const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
    const bool out_lane = lane >= HL && lane <= WAVE - 1 - HR;
    // Fields are read and written through buffer resources over their rows
    // 0 .. y1+1 (the host launches this kernel only for fields under 4 GiB):
    // an access the reference semantics exclude gets the offset OOB, which
    // reads 0 and writes nothing, so no access needs a branch.
    const unsigned row_b = (unsigned)pitch * (unsigned)sizeof(T);
    const unsigned bytes = (unsigned)(e.y1 + 2) * row_b;
    const FieldBuf<T> b_sd(sd_in, bytes), b_rt(rt_in, bytes), b_ut(ut_in, bytes), b_mi(Mi, bytes), b_kx(Kx, bytes),
        b_ky(Ky, bytes), b_sdo(sd_out, bytes), b_rto(rt_out, bytes), b_uto(ut_out, bytes);
    auto off = [&](int r0) { return ((unsigned)c + (unsigned)r0 * (unsigned)pitch) * (unsigned)sizeof(T); };
    const int s_first = ck * chunk;
    const int s_last = (s_first + chunk < nstrips ? s_first + chunk : nstrips) - 1;

    auto strip_len = [&](int r0) {   // rows of the strip inside the extent (uniform across the wave)
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        return len < 0 ? 0 : (len > STRIP ? STRIP : len);
    };
    // Issues the reads of strip s.
    auto fetch = [&](int s, MarchRaw<T> &m) {
        const int r0 = e.y0 + s * STRIP, len = strip_len(r0), act = col_in ? len : 0;
        const unsigned b0 = off(r0);
        constexpr unsigned W = sizeof(T);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const int rr = r0 + k;
            m.sd[k] = b_sd.load((col_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1) ? b0 + k * row_b : OOB);
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = b_ky.load((k <= act && act > 0) ? b0 + k * row_b : OOB);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const bool a = k < act;
            const unsigned bk = b0 + k * row_b;
            m.kx[k] = b_kx.load(col_mem && k < len ? bk : OOB);   // column x1+1 too: the right face of x1
            m.rt[k] = b_rt.load(a ? bk : OOB);
            m.ut[k] = b_ut.load(a && out_lane ? bk : OOB);
            m.mi[k] = PREC == PREC_DIAG ? b_mi.load(a ? bk : OOB) : T(0.0);
            m.sde[k] = b_sd.load(a && lane == 0 ? bk - W : OOB);
        }
    };
    // Forms the pipeline strip of strip s from its reads.
    auto prepare = [&](int s, const MarchRaw<T> &raw, MarchStrip<T> &m) {
        m.r0 = e.y0 + s * STRIP;
        m.len = strip_len(m.r0);
        const int act = col_in ? m.len : 0;
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            m.sd[k] = raw.sd[k];
            m.rt[k] = raw.rt[k];
            m.ut[k] = raw.ut[k];
            m.kx[k] = raw.kx[k];
            m.mi[k] = raw.mi[k];
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = raw.ky[k];
        T diag[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const T kxr = lane_from_above(m.kx[k]);   // Kx on the right face: the next lane's
            diag[k] = T(1.0) + (m.ky[k + 1] + m.ky[k]) + (kxr + m.kx[k]);
        }
        if (PREC == PREC_BLOCK && act > 0) strip_factor<T>(act, diag, m.ky, m.cp, m.bfp);
    };

    // Moves strip m from step st to st+1 (lo, hi: sd of the rows below and
    // above it at step st; sde: lane 0's left neighbour column at step 0).
    auto advance = [&](MarchStrip<T> &m, T lo, T hi, const T (&sde)[STRIP], int st) {
        const T a = (T)cf.a[st], b = (T)cf.b[st];
        const int ak = (st < NS - 1 || out_lane) && col_in ? m.len : 0;
        T rv[STRIP], zv[STRIP], diag[STRIP], kxr[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const T from_left = lane_from_below(m.sd[k]);
            const T from_right = lane_from_above(m.sd[k]);
            kxr[k] = lane_from_above(m.kx[k]);
            diag[k] = T(1.0) + (m.ky[k + 1] + m.ky[k]) + (kxr[k] + m.kx[k]);
            rv[k] = zv[k] = T(0.0);
            if (k < ak) {
                const T below = k == 0 ? lo : m.sd[k - 1];
                const T above = k == STRIP - 1 ? hi : m.sd[k + 1];
                const T left = (st > 0 || lane > 0) ? from_left : sde[k];
                const T res = diag[k] * m.sd[k] - (m.ky[k + 1] * above + m.ky[k] * below)
                              - (kxr[k] * from_right + m.kx[k] * left);
                rv[k] = m.rt[k] - res;
            }
        }
        if (PREC == PREC_BLOCK && ak > 0) strip_apply<T>(rv, zv, ak, diag[0], m.ky, m.cp, m.bfp);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            if (k < ak) {
                T sdn;
                if (PREC == PREC_BLOCK) sdn = a * m.sd[k] + b * zv[k];
                else if (PREC == PREC_DIAG) sdn = a * m.sd[k] + b * rv[k] * m.mi[k];
                else sdn = a * m.sd[k] + b * rv[k];
                m.sd[k] = sdn;
                m.rt[k] = rv[k];
                m.ut[k] = m.ut[k] + sdn;
            }
        }
    };

    MarchStrip<T> H[NS];   // H[m]: strip i-1-m at step m
    T P[NS];               // P[m]: last row of strip i-2-m at step m
    T sde0[STRIP];         // edge-lane neighbours of the strip in H[0]
#pragma unroll
    for (int m = 0; m < NS; m++) {
        H[m] = MarchStrip<T>{};
        H[m].len = 0;
        H[m].r0 = e.y0;
        P[m] = T(0.0);
    }
#pragma unroll
    for (int k = 0; k < STRIP; k++) sde0[k] = T(0.0);
    MarchRaw<T> nxt;
    fetch(s_first - NS, nxt);
    for (int i = s_first - NS; i <= s_last + NS; i++) {
        MarchStrip<T> cur;
        T sde_cur[STRIP];
        prepare(i, nxt, cur);
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde_cur[k] = nxt.sde[k];
        if (i < s_last + NS) fetch(i + 1, nxt);   // read one strip ahead
        T hi = cur.sd[0];
#pragma unroll
        for (int m = 0; m < NS; m++) {
            const T last = H[m].sd[STRIP - 1];
            advance(H[m], P[m], hi, sde0, m);   // strip i-1-m: step m -> m+1
            P[m] = last;
            hi = H[m].sd[0];
        }
        const MarchStrip<T> &f = H[NS - 1];   // strip i-NS, finished
        const int sf = i - NS;
        const bool out = out_lane && sf >= s_first && sf <= s_last;
        const unsigned b0 = off(f.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const unsigned bk = out && col_in && k < f.len ? b0 + k * row_b : OOB;
            b_sdo.store(bk, f.sd[k]);
            b_rto.store(bk, f.rt[k]);
            b_uto.store(bk, f.ut[k]);
        }
#pragma unroll
        for (int m = NS - 1; m > 0; m--) H[m] = H[m - 1];
        H[0] = cur;
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde0[k] = sde_cur[k];
    }
}

// The value of lane 0 as a wave-uniform value (held in scalar registers).
__device__ __forceinline__ float lane0_value(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ double lane0_value(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL)), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The streaming pass of k_ppcg_inner_march with its pipeline held as a ring
// (no reference counterpart; the same results bit for bit).  Strip s lives
// in ring slot (s - first) mod (NS+1) from the iteration it enters to the one
// it leaves, and is advanced in place, so no strip moves between registers:
// the loop body is instantiated NS+1 times, one per ring phase.  Row guards
// are uniform per strip (its rows inside the extent) and branch, the column
// guard is one select of sd per cell (a cell outside the extent keeps its sd;
// its rtemp and utemp are never stored), and an advance a chunk's strips do
// not need (the pipeline fill above the chunk: strip s needs steps up to
// NS - its distance from the chunk) is skipped.
template <typename T, int PREC, int NS, bool END = false>
__global__ __launch_bounds__(MARCH_WAVES *WAVE) void k_ppcg_inner_ring(
    Ext e, long pitch, int chunk, const T *__restrict__ sd_in, T *__restrict__ sd_out, const T *__restrict__ rt_in,
    T *__restrict__ rt_out, const T *ut_in, T *ut_out, const T *__restrict__ Mi, const T *__restrict__ Kx,
    const T *__restrict__ Ky, StepCoefs cf, const T *__restrict__ r_end,
    const T *__restrict__ rs_end, Red red)
{
    static_assert(NS >= 1 && NS <= MAX_PASS_STEPS, "1..5 steps per pass");
    constexpr int HL = NS - 1, HR = NS, OUT_X = march_out_x(NS), RN = NS + 1;
    static_assert(OUT_X == WAVE - HL - HR, "band layout");
    const int lane = threadIdx.x & (WAVE - 1);
    if (e.x1 < e.x0 || e.y1 < e.y0) return;
    const int nbands = (e.x1 - e.x0 + OUT_X) / OUT_X;
    const int nstrips = (e.y1 - e.y0 + STRIP) / STRIP;
    const int nchunks = (nstrips + chunk - 1) / chunk;
    const int w = blockIdx.x * MARCH_WAVES + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    // END: the last workgroup's spare waves still join the reduction
    const bool active = w < nbands * nchunks;
    if (!END && !active) return;
    const int band = w % nbands, ck = w / nbands;
    const int c = e.x0 + band * OUT_X - HL + lane;
    const bool col_in = c >= e.x0 && c <= e.x1;
    const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
    const bool out_lane = lane >= HL && lane <= WAVE - 1 - HR;
    const unsigned row_b = (unsigned)pitch * (unsigned)sizeof(T);
    const unsigned bytes = (unsigned)(e.y1 + 2) * row_b;
    const FieldBuf<T> b_sd(sd_in, bytes), b_rt(rt_in, bytes), b_ut(ut_in, bytes), b_mi(Mi, bytes), b_kx(Kx, bytes),
        b_ky(Ky, bytes), b_sdo(sd_out, bytes), b_rto(rt_out, bytes), b_uto(ut_out, bytes), b_re(r_end, bytes),
        b_rse(rs_end, bytes);
    auto off = [&](int r0) { return ((unsigned)c + (unsigned)r0 * (unsigned)pitch) * (unsigned)sizeof(T); };
    const int s_first = ck * chunk;
    const int s_last = (s_first + chunk < nstrips ? s_first + chunk : nstrips) - 1;

    auto strip_len = [&](int r0) {   // rows of the strip inside the extent (uniform across the wave)
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        return len < 0 ? 0 : (len > STRIP ? STRIP : len);
    };
    // Issues the reads of strip s (the offsets of k_ppcg_inner_march).
    auto fetch = [&](int s, MarchRaw<T> &m) {
        const int r0 = e.y0 + s * STRIP, len = strip_len(r0), act = col_in ? len : 0;
        const unsigned b0 = off(r0);
        constexpr unsigned W = sizeof(T);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const int rr = r0 + k;
            m.sd[k] = b_sd.load((col_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1) ? b0 + k * row_b : OOB);
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = b_ky.load((k <= act && act > 0) ? b0 + k * row_b : OOB);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const bool a = k < act;
            const unsigned bk = b0 + k * row_b;
            m.kx[k] = b_kx.load(col_mem && k < len ? bk : OOB);
            m.rt[k] = b_rt.load(a ? bk : OOB);
            m.ut[k] = b_ut.load(a && out_lane ? bk : OOB);
            m.mi[k] = PREC == PREC_DIAG ? b_mi.load(a ? bk : OOB) : T(0.0);
            m.sde[k] = b_sd.load(a && lane == 0 ? bk - W : OOB);
        }
    };
    // Forms ring slot m from strip s's reads; the block-Jacobi factors over the
    // strip's rows inside the extent (a lane outside the extent reads zero
    // coefficients, so its factors are finite and never stored).
    auto prepare = [&](int s, const MarchRaw<T> &raw, MarchStrip<T> &m) {
        m.r0 = e.y0 + s * STRIP;
        m.len = strip_len(m.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            m.sd[k] = raw.sd[k];
            m.rt[k] = raw.rt[k];
            m.ut[k] = raw.ut[k];
            m.kx[k] = raw.kx[k];
            m.mi[k] = raw.mi[k];
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = raw.ky[k];
        if (PREC == PREC_BLOCK && m.len > 0) {
            T diag[STRIP];
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const T kxr = lane_from_above(m.kx[k]);
                diag[k] = T(1.0) + (m.ky[k + 1] + m.ky[k]) + (kxr + m.kx[k]);
            }
            if (m.len == STRIP) strip_factor<T>(STRIP, diag, m.ky, m.cp, m.bfp);
            else strip_factor<T>(m.len, diag, m.ky, m.cp, m.bfp);
        }
    };
    // Moves strip m from step st to st+1 in place (lo, hi: sd of the rows
    // below and above it at step st; sde: lane 0's left neighbour column at
    // step 0).  The expressions of k_ppcg_inner_march's advance; LEN: the
    // strip's rows inside the extent (STRIP, or 0 for a shorter strip, whose
    // length m.len is then read at run time).
    auto advance_len = [&](auto len_c, MarchStrip<T> &m, T lo, T hi, const T (&sde)[STRIP], int st) {
        constexpr int LEN = decltype(len_c)::value;
        const int len = LEN == STRIP ? STRIP : m.len;
        const T a = (T)cf.a[st], b = (T)cf.b[st];
        T rv[STRIP], zv[STRIP], diag[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const T from_left = lane_from_below(m.sd[k]);
            const T from_right = lane_from_above(m.sd[k]);
            const T kxr = lane_from_above(m.kx[k]);
            diag[k] = T(1.0) + (m.ky[k + 1] + m.ky[k]) + (kxr + m.kx[k]);
            const T below = k == 0 ? lo : m.sd[k - 1];
            const T above = k == STRIP - 1 ? hi : m.sd[k + 1];
            const T left = (st > 0 || lane > 0) ? from_left : sde[k];
            const T res = diag[k] * m.sd[k] - (m.ky[k + 1] * above + m.ky[k] * below)
                          - (kxr * from_right + m.kx[k] * left);
            rv[k] = m.rt[k] - res;
        }
        if (PREC == PREC_BLOCK) strip_apply<T>(rv, zv, len, diag[0], m.ky, m.cp, m.bfp);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            if (LEN == STRIP || k < len) {
                T sdn;
                if (PREC == PREC_BLOCK) sdn = a * m.sd[k] + b * zv[k];
                else if (PREC == PREC_DIAG) sdn = a * m.sd[k] + b * rv[k] * m.mi[k];
                else sdn = a * m.sd[k] + b * rv[k];
                m.sd[k] = col_in ? sdn : m.sd[k];
                m.rt[k] = rv[k];
                m.ut[k] = m.ut[k] + sdn;
            }
        }
    };
    auto advance = [&](MarchStrip<T> &m, T lo, T hi, const T (&sde)[STRIP], int st) {
        if (m.len == STRIP) advance_len(std::integral_constant<int, STRIP>{}, m, lo, hi, sde, st);
        else if (m.len > 0) advance_len(std::integral_constant<int, 0>{}, m, lo, hi, sde, st);
    };

    MarchStrip<T> R[RN];   // R[j]: the strip entered at the iterations of phase j
    T P[NS];               // P[m]: last row at step m of the strip behind the one at age m
    T sde0[STRIP];         // lane 0's left neighbours of the strip at age 0 (uniform: scalar registers)
#pragma unroll
    for (int j = 0; j < RN; j++) {
        R[j] = MarchStrip<T>{};
        R[j].len = 0;
        R[j].r0 = e.y0;
    }
#pragma unroll
    for (int m = 0; m < NS; m++) P[m] = T(0.0);
#pragma unroll
    for (int k = 0; k < STRIP; k++) sde0[k] = T(0.0);
    const int i0 = s_first - NS, i1 = s_last + NS;
    MarchRaw<T> nxt;
    DD eacc[1] = {{0.0, 0.0}};
    // r and rstore at the cells of the strip that leaves finished at
    // iteration i (strip i-NS, whose slot still holds its rows), 0 elsewhere
    auto end_loads = [&](const auto &f, int i, T (&er)[STRIP], T (&es)[STRIP]) {
        const int sf = i - NS;
        const bool out = out_lane && sf >= s_first && sf <= s_last;
        const unsigned b0 = off(f.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const unsigned bk = out && col_in && k < f.len ? b0 + k * row_b : OOB;
            er[k] = b_re.load(bk);
            es[k] = b_rse.load(bk);
        }
    };
    if (active) fetch(i0, nxt);
    // One iteration at ring phase PH = (i - i0) mod (NS+1): strip i enters
    // slot PH, the strip of age m (i-1-m, slot PH-1-m) moves from step m to
    // m+1, and strip i-NS (slot PH+1) leaves finished.
    auto body = [&](auto ph, int i) {
        constexpr int PH = decltype(ph)::value;
        T er[STRIP], es[STRIP];   // END: r and rstore of strip i-NS, read before anything else this iteration
        if constexpr (END) end_loads(R[(PH + 1) % RN], i, er, es);
        MarchStrip<T> &cur = R[PH];
        prepare(i, nxt, cur);
        T sde_cur[STRIP];   // lane 0's values, read out before the next fetch overwrites them
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde_cur[k] = lane0_value(nxt.sde[k]);
        if (i < i1) fetch(i + 1, nxt);   // read one strip ahead
        T hi = cur.sd[0];
#pragma unroll
        for (int m = 0; m < NS; m++) {
            MarchStrip<T> &h = R[(PH - 1 - m + 2 * RN) % RN];
            const T last = h.sd[STRIP - 1];
            const int s = i - 1 - m;
            const int dist = s < s_first ? s_first - s : (s > s_last ? s - s_last : 0);
            if (dist <= NS - 1 - m) advance(h, P[m], hi, sde0, m);
            P[m] = last;
            hi = h.sd[0];
        }
        const MarchStrip<T> &f = R[(PH + 1) % RN];   // strip i-NS, finished
        const int sf = i - NS;
        const bool out = out_lane && sf >= s_first && sf <= s_last;
        const unsigned b0 = off(f.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const unsigned bk = out && col_in && k < f.len ? b0 + k * row_b : OOB;
            b_sdo.store(bk, f.sd[k]);
            b_rto.store(bk, f.rt[k]);
            b_uto.store(bk, f.ut[k]);
        }
        if constexpr (END) {   // k_ppcg_end's (r - rstore).z with z = the utemp just written
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const bool o = out && col_in && k < f.len;
                dd_add_prod(eacc[0], (double)(er[k] - es[k]), o ? (double)f.ut[k] : 0.0);
            }
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde0[k] = sde_cur[k];
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    for (int i = i0; active;) {
        body(I0{}, i);
        if (++i > i1) break;
        body(I1{}, i);
        if (++i > i1) break;
        if constexpr (RN > 2) {
            body(I2{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 3) {
            body(I3{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 4) {
            body(I4{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 5) {
            body(I5{}, i);
            if (++i > i1) break;
        }
    }
    if constexpr (END) block_reduce<1>(eacc, red);
}

// NS Chebyshev iterations (tea_leaf_cheby.cuknl:84-152 then :71-82, the
// expressions of k_cheby_march: w = A u, r = u0 - w, p = alpha_k p + beta_k
// M^-1 r, u = u + p) in one streaming pass, with the pipeline of
// k_ppcg_inner_ring: bands of WAVE lanes (NS - 1 halo lanes at the left, NS
// at the right) walk down chunks of strips, each strip advanced NS times in
// its ring slot (u the stencil field in place of sd, p in place of rtemp, u0
// read-only in place of utemp; no reference counterpart as a kernel, the same
// values bit for bit).  u and p go to their spare buffers (neighbouring waves
// read both); the last iteration's w and r, which the reference leaves in
// their fields, are written by the strip's last advance.
template <typename T, int PREC, int NS>
__global__ __launch_bounds__(MARCH_WAVES *WAVE) void k_cheby_ring(
    Ext e, long pitch, int chunk, const T *__restrict__ u_in, T *__restrict__ u_out, const T *__restrict__ p_in,
    T *__restrict__ p_out, const T *__restrict__ u0, T *__restrict__ r_out, T *__restrict__ w_out,
    const T *__restrict__ Mi, const T *__restrict__ Kx, const T *__restrict__ Ky, StepCoefs cf)
{
    static_assert(NS >= 1 && NS <= MAX_PASS_STEPS, "1..5 iterations per pass");
    constexpr int HL = NS - 1, HR = NS, OUT_X = march_out_x(NS), RN = NS + 1;
    static_assert(OUT_X == WAVE - HL - HR, "band layout");
    const int lane = threadIdx.x & (WAVE - 1);
    if (e.x1 < e.x0 || e.y1 < e.y0) return;
    const int nbands = (e.x1 - e.x0 + OUT_X) / OUT_X;
    const int nstrips = (e.y1 - e.y0 + STRIP) / STRIP;
    const int nchunks = (nstrips + chunk - 1) / chunk;
    const int w = blockIdx.x * MARCH_WAVES + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    if (w >= nbands * nchunks) return;
    const int band = w % nbands, ck = w / nbands;
    const int c = e.x0 + band * OUT_X - HL + lane;
    const bool col_in = c >= e.x0 && c <= e.x1;
    const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
    const bool out_lane = lane >= HL && lane <= WAVE - 1 - HR;
    const unsigned row_b = (unsigned)pitch * (unsigned)sizeof(T);
    const unsigned bytes = (unsigned)(e.y1 + 2) * row_b;
    const FieldBuf<T> b_u(u_in, bytes), b_p(p_in, bytes), b_u0(u0, bytes), b_mi(Mi, bytes), b_kx(Kx, bytes),
        b_ky(Ky, bytes), b_uo(u_out, bytes), b_po(p_out, bytes), b_ro(r_out, bytes), b_wo(w_out, bytes);
    auto off = [&](int r0) { return ((unsigned)c + (unsigned)r0 * (unsigned)pitch) * (unsigned)sizeof(T); };
    const int s_first = ck * chunk;
    const int s_last = (s_first + chunk < nstrips ? s_first + chunk : nstrips) - 1;

    auto strip_len = [&](int r0) {   // rows of the strip inside the extent (uniform across the wave)
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        return len < 0 ? 0 : (len > STRIP ? STRIP : len);
    };
    // Issues the reads of strip s: u (sd's offsets: the halo rows and columns
    // read as stored), p and u0 at the cells inside the extent, lane 0's left
    // column of u.
    auto fetch = [&](int s, MarchRaw<T> &m) {
        const int r0 = e.y0 + s * STRIP, len = strip_len(r0), act = col_in ? len : 0;
        const unsigned b0 = off(r0);
        constexpr unsigned W = sizeof(T);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const int rr = r0 + k;
            m.sd[k] = b_u.load((col_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1) ? b0 + k * row_b : OOB);
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = b_ky.load((k <= act && act > 0) ? b0 + k * row_b : OOB);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const bool a = k < act;
            const unsigned bk = b0 + k * row_b;
            m.kx[k] = b_kx.load(col_mem && k < len ? bk : OOB);
            m.rt[k] = b_p.load(a ? bk : OOB);
            m.ut[k] = b_u0.load(a ? bk : OOB);
            m.mi[k] = PREC == PREC_DIAG ? b_mi.load(a ? bk : OOB) : T(0.0);
            m.sde[k] = b_u.load(a && lane == 0 ? bk - W : OOB);
        }
    };
    auto prepare = [&](int s, const MarchRaw<T> &raw, MarchStrip<T> &m) {
        m.r0 = e.y0 + s * STRIP;
        m.len = strip_len(m.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            m.sd[k] = raw.sd[k];
            m.rt[k] = raw.rt[k];
            m.ut[k] = raw.ut[k];
            m.kx[k] = raw.kx[k];
            m.mi[k] = raw.mi[k];
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = raw.ky[k];
        if (PREC == PREC_BLOCK && m.len > 0) {
            T diag[STRIP];
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const T kxr = lane_from_above(m.kx[k]);
                diag[k] = T(1.0) + (m.ky[k + 1] + m.ky[k]) + (kxr + m.kx[k]);
            }
            if (m.len == STRIP) strip_factor<T>(STRIP, diag, m.ky, m.cp, m.bfp);
            else strip_factor<T>(m.len, diag, m.ky, m.cp, m.bfp);
        }
    };
    // One Chebyshev iteration of strip m in place (lo, hi: u of the rows
    // below and above it; sde: lane 0's left column at iteration 0); the last
    // one (st == NS-1) of a strip this wave outputs also writes w and r.
    auto advance_len = [&](auto len_c, MarchStrip<T> &m, T lo, T hi, const T (&sde)[STRIP], int st, bool out_s) {
        constexpr int LEN = decltype(len_c)::value;
        const int len = LEN == STRIP ? STRIP : m.len;
        const T alpha = (T)cf.a[st], beta = (T)cf.b[st];
        T wv[STRIP], rv[STRIP], zv[STRIP], diag[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const T from_left = lane_from_below(m.sd[k]);
            const T from_right = lane_from_above(m.sd[k]);
            const T kxr = lane_from_above(m.kx[k]);
            diag[k] = T(1.0) + (m.ky[k + 1] + m.ky[k]) + (kxr + m.kx[k]);
            const T below = k == 0 ? lo : m.sd[k - 1];
            const T above = k == STRIP - 1 ? hi : m.sd[k + 1];
            const T left = (st > 0 || lane > 0) ? from_left : sde[k];
            wv[k] = diag[k] * m.sd[k] - (m.ky[k + 1] * above + m.ky[k] * below) - (kxr * from_right + m.kx[k] * left);
            rv[k] = m.ut[k] - wv[k];
        }
        if (PREC == PREC_BLOCK) strip_apply<T>(rv, zv, len, diag[0], m.ky, m.cp, m.bfp);
        const bool wr = st == NS - 1 && out_s;
        const unsigned b0 = off(m.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            if (LEN == STRIP || k < len) {
                T pv;
                if (PREC == PREC_BLOCK) pv = alpha * m.rt[k] + beta * zv[k];
                else if (PREC == PREC_DIAG) pv = alpha * m.rt[k] + beta * m.mi[k] * rv[k];
                else pv = alpha * m.rt[k] + beta * rv[k];
                if (wr) {
                    const unsigned bk = col_in ? b0 + k * row_b : OOB;
                    b_wo.store(bk, wv[k]);
                    b_ro.store(bk, rv[k]);
                }
                m.rt[k] = col_in ? pv : m.rt[k];
                m.sd[k] = col_in ? m.sd[k] + pv : m.sd[k];
            }
        }
    };
    auto advance = [&](MarchStrip<T> &m, T lo, T hi, const T (&sde)[STRIP], int st, bool out_s) {
        if (m.len == STRIP) advance_len(std::integral_constant<int, STRIP>{}, m, lo, hi, sde, st, out_s);
        else if (m.len > 0) advance_len(std::integral_constant<int, 0>{}, m, lo, hi, sde, st, out_s);
    };

    MarchStrip<T> R[RN];
    T P[NS];
    T sde0[STRIP];
#pragma unroll
    for (int j = 0; j < RN; j++) {
        R[j] = MarchStrip<T>{};
        R[j].len = 0;
        R[j].r0 = e.y0;
    }
#pragma unroll
    for (int m = 0; m < NS; m++) P[m] = T(0.0);
#pragma unroll
    for (int k = 0; k < STRIP; k++) sde0[k] = T(0.0);
    const int i0 = s_first - NS, i1 = s_last + NS;
    MarchRaw<T> nxt;
    fetch(i0, nxt);
    auto body = [&](auto ph, int i) {
        constexpr int PH = decltype(ph)::value;
        MarchStrip<T> &cur = R[PH];
        prepare(i, nxt, cur);
        T sde_cur[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde_cur[k] = lane0_value(nxt.sde[k]);
        if (i < i1) fetch(i + 1, nxt);
        T hi = cur.sd[0];
#pragma unroll
        for (int m = 0; m < NS; m++) {
            MarchStrip<T> &h = R[(PH - 1 - m + 2 * RN) % RN];
            const T last = h.sd[STRIP - 1];
            const int s = i - 1 - m;
            const int dist = s < s_first ? s_first - s : (s > s_last ? s - s_last : 0);
            if (dist <= NS - 1 - m) advance(h, P[m], hi, sde0, m, out_lane && dist == 0);
            P[m] = last;
            hi = h.sd[0];
        }
        const MarchStrip<T> &f = R[(PH + 1) % RN];   // strip i-NS, finished
        const int sf = i - NS;
        const bool out = out_lane && sf >= s_first && sf <= s_last;
        const unsigned b0 = off(f.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const unsigned bk = out && col_in && k < f.len ? b0 + k * row_b : OOB;
            b_uo.store(bk, f.sd[k]);
            b_po.store(bk, f.rt[k]);
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde0[k] = sde_cur[k];
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    for (int i = i0;;) {
        body(I0{}, i);
        if (++i > i1) break;
        body(I1{}, i);
        if (++i > i1) break;
        if constexpr (RN > 2) {
            body(I2{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 3) {
            body(I3{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 4) {
            body(I4{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 5) {
            body(I5{}, i);
            if (++i > i1) break;
        }
    }
}

// k_ppcg_inner_ring with the face coefficients of the strips in flight in
// LDS instead of registers (no reference counterpart; the same results bit
// for bit).  A strip's Kx (4 rows), Ky (5 rows) and, in FP32, diagonal (4
// rows, formed once as it enters) are written to the wave's LDS ring slot
// when it enters and read back at each of its steps (the right face's Kx from
// the next lane's column, lane 63 reading a zero column as the lane shift
// gives it), which frees 9 registers per strip, so longer passes keep two
// waves per SIMD, and saves the diagonal's four additions per cell and step
// (FP32 five-step pass 164.5 -> 159.1 us at 4096^2).  Each wave owns its part of LDS; its writes and reads
// are in program order, so no barrier is needed.
// FROM_R (block Jacobi): the pass starts from sd = M^-1 r / theta, rtemp = r
// and utemp = sd, as a fused ur (k_ppcg_ur_sd_x2_wp, k_ppcg_ur_sd_wp) forms
// them, formed here from r (read as rtemp) when each strip enters, by the
// expressions of the inner steps' z = M^-1 rtemp: that ur then stores
// neither sd nor z.  Only sd's halo ring is read.
template <typename T, int PREC, int NS, bool END = false, bool FROM_R = false>
__global__ __launch_bounds__(MARCH_WAVES *WAVE) void k_ppcg_inner_ringl(
    Ext e, long pitch, int chunk, const T *__restrict__ sd_in, T *__restrict__ sd_out, const T *__restrict__ rt_in,
    T *__restrict__ rt_out, const T *ut_in, T *ut_out, const T *__restrict__ Mi, const T *__restrict__ Kx,
    const T *__restrict__ Ky, StepCoefs cf, const T *__restrict__ r_end,
    const T *__restrict__ rs_end, Red red)
{
    static_assert(NS >= 1 && NS <= MAX_PASS_STEPS, "1..5 steps per pass");
    static_assert(!FROM_R || PREC == PREC_BLOCK, "sd formed from r with block Jacobi");
    constexpr int HL = FROM_R ? NS : NS - 1, HR = FROM_R ? NS + 1 : NS, OUT_X = march_out_x(NS, FROM_R), RN = NS + 1;
    // LDS rows per slot: Kx 0..3, Ky 0..4 and, in FP32, the diagonal 0..3 (in FP64 the
    // diagonal is formed again at each step: measured faster there)
    constexpr bool CACHE_DIAG = sizeof(T) == 4;
    constexpr int DG = 2 * STRIP + 1, NK = CACHE_DIAG ? 3 * STRIP + 1 : DG, LW = WAVE + 1;
    static_assert(OUT_X == WAVE - HL - HR, "band layout");
    __shared__ T lds[MARCH_WAVES][RN][NK][LW];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    if (e.x1 < e.x0 || e.y1 < e.y0) return;
    const int nbands = (e.x1 - e.x0 + OUT_X) / OUT_X;
    const int nstrips = (e.y1 - e.y0 + STRIP) / STRIP;
    const int nchunks = (nstrips + chunk - 1) / chunk;
    const int w = blockIdx.x * MARCH_WAVES + wv;
    // END: the last workgroup's spare waves still join the reduction
    const bool active = w < nbands * nchunks;
    if (!END && !active) return;
    T(*L)[NK][LW] = lds[wv];
    if (lane == WAVE - 1) {   // the column past the last lane reads as zero
#pragma unroll
        for (int j = 0; j < RN; j++)
#pragma unroll
            for (int q = 0; q < NK; q++) L[j][q][WAVE] = T(0.0);
    }
    const int band = w % nbands, ck = w / nbands;
    const int c = e.x0 + band * OUT_X - HL + lane;
    const bool col_in = c >= e.x0 && c <= e.x1;
    const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
    const bool out_lane = lane >= HL && lane <= WAVE - 1 - HR;
    const unsigned row_b = (unsigned)pitch * (unsigned)sizeof(T);
    const unsigned bytes = (unsigned)(e.y1 + 2) * row_b;
    const FieldBuf<T> b_sd(sd_in, bytes), b_rt(rt_in, bytes), b_ut(ut_in, bytes), b_mi(Mi, bytes), b_kx(Kx, bytes),
        b_ky(Ky, bytes), b_sdo(sd_out, bytes), b_rto(rt_out, bytes), b_uto(ut_out, bytes), b_re(r_end, bytes),
        b_rse(rs_end, bytes);
    auto off = [&](int r0) { return ((unsigned)c + (unsigned)r0 * (unsigned)pitch) * (unsigned)sizeof(T); };
    const int s_first = ck * chunk;
    const int s_last = (s_first + chunk < nstrips ? s_first + chunk : nstrips) - 1;

    struct Slot {
        T sd[STRIP], rt[STRIP], ut[STRIP], cp[STRIP], bfp[STRIP], mi[STRIP];
        double rd;   // FP32 block Jacobi: 1 / diag0 in double (strip_apply_rd)
        int r0, len;
    };
    auto strip_len = [&](int r0) {   // rows of the strip inside the extent (uniform across the wave)
        int len = r0 < e.y0 ? 0 : e.y1 - r0 + 1;
        return len < 0 ? 0 : (len > STRIP ? STRIP : len);
    };
    // Issues the reads of strip s (the offsets of k_ppcg_inner_march).
    auto fetch = [&](int s, MarchRaw<T> &m) {
        const int r0 = e.y0 + s * STRIP, len = strip_len(r0), act = col_in ? len : 0;
        const unsigned b0 = off(r0);
        constexpr unsigned W = sizeof(T);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const int rr = r0 + k;
            // FROM_R: the cells inside the extent (k < act) take sd formed from r
            const bool rd_sd = col_mem && rr >= e.y0 - 1 && rr <= e.y1 + 1 && !(FROM_R && k < act);
            m.sd[k] = b_sd.load(rd_sd ? b0 + k * row_b : OOB);
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) m.ky[k] = b_ky.load((k <= act && act > 0) ? b0 + k * row_b : OOB);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const bool a = k < act;
            const unsigned bk = b0 + k * row_b;
            m.kx[k] = b_kx.load(col_mem && k < len ? bk : OOB);
            m.rt[k] = b_rt.load(a ? bk : OOB);
            m.ut[k] = b_ut.load(!FROM_R && a && out_lane ? bk : OOB);
            m.mi[k] = PREC == PREC_DIAG ? b_mi.load(a ? bk : OOB) : T(0.0);
            m.sde[k] = b_sd.load(!FROM_R && a && lane == 0 ? bk - W : OOB);
        }
    };
    // Forms ring slot j from strip s's reads: the fields into registers, the
    // coefficients into the slot's LDS rows, the block-Jacobi factors (FROM_R:
    // and sd = utemp = M^-1 r / theta inside the extent).
    auto prepare = [&](int s, const MarchRaw<T> &raw, Slot &m, T (*Lj)[LW]) {
        m.r0 = e.y0 + s * STRIP;
        m.len = strip_len(m.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            m.sd[k] = raw.sd[k];
            m.rt[k] = raw.rt[k];
            m.ut[k] = raw.ut[k];
            m.mi[k] = raw.mi[k];
            Lj[k][lane] = raw.kx[k];
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) Lj[STRIP + k][lane] = raw.ky[k];
        T diag[STRIP];   // formed once per strip (read back at each step with CACHE_DIAG)
        if (CACHE_DIAG || PREC == PREC_BLOCK) {
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const T kxr = lane_from_above(raw.kx[k]);
                diag[k] = T(1.0) + (raw.ky[k + 1] + raw.ky[k]) + (kxr + raw.kx[k]);
                if constexpr (CACHE_DIAG) Lj[DG + k][lane] = diag[k];
            }
        }
        if (PREC == PREC_BLOCK && m.len > 0) {
            if (m.len == STRIP) strip_factor<T>(STRIP, diag, raw.ky, m.cp, m.bfp);
            else strip_factor<T>(m.len, diag, raw.ky, m.cp, m.bfp);
            if constexpr (sizeof(T) == 4) m.rd = 1.0 / (double)diag[0];
            if constexpr (FROM_R) {
                const T theta_r = T(1.0) / (T)cf.theta;
                T zv[STRIP];
                if constexpr (sizeof(T) == 4) strip_apply_rd(m.rt, zv, m.len, m.rd, raw.ky, m.cp, m.bfp);
                else strip_apply<T>(m.rt, zv, m.len, diag[0], raw.ky, m.cp, m.bfp);
#pragma unroll
                for (int k = 0; k < STRIP; k++) {
                    if (col_in && k < m.len) {
                        m.sd[k] = zv[k] * theta_r;
                        m.ut[k] = m.sd[k];
                    }
                }
            }
        }
    };
    // Moves the strip in slot m (coefficients in Lj) from step st to st+1 in
    // place: the expressions of k_ppcg_inner_march's advance.
    auto advance_len = [&](auto len_c, Slot &m, T (*Lj)[LW], T lo, T hi, const T (&sde)[STRIP], int st) {
        constexpr int LEN = decltype(len_c)::value;
        const int len = LEN == STRIP ? STRIP : m.len;
        const T a = (T)cf.a[st], b = (T)cf.b[st];
        T kx[STRIP], kxr[STRIP], ky[STRIP + 1];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            kx[k] = Lj[k][lane];
            kxr[k] = Lj[k][lane + 1];
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) ky[k] = Lj[STRIP + k][lane];
        T rv[STRIP], zv[STRIP], diag[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const T from_left = lane_from_below(m.sd[k]);
            const T from_right = lane_from_above(m.sd[k]);
            if constexpr (CACHE_DIAG) diag[k] = Lj[DG + k][lane];
            else diag[k] = T(1.0) + (ky[k + 1] + ky[k]) + (kxr[k] + kx[k]);
            const T below = k == 0 ? lo : m.sd[k - 1];
            const T above = k == STRIP - 1 ? hi : m.sd[k + 1];
            const T left = (st > 0 || lane > 0) ? from_left : sde[k];
            const T res = diag[k] * m.sd[k] - (ky[k + 1] * above + ky[k] * below)
                          - (kxr[k] * from_right + kx[k] * left);
            rv[k] = m.rt[k] - res;
        }
        if constexpr (PREC == PREC_BLOCK && sizeof(T) == 4) strip_apply_rd(rv, zv, len, m.rd, ky, m.cp, m.bfp);
        else if (PREC == PREC_BLOCK) strip_apply<T>(rv, zv, len, diag[0], ky, m.cp, m.bfp);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            if (LEN == STRIP || k < len) {
                T sdn;
                if (PREC == PREC_BLOCK) sdn = a * m.sd[k] + b * zv[k];
                else if (PREC == PREC_DIAG) sdn = a * m.sd[k] + b * rv[k] * m.mi[k];
                else sdn = a * m.sd[k] + b * rv[k];
                m.sd[k] = col_in ? sdn : m.sd[k];
                m.rt[k] = rv[k];
                m.ut[k] = m.ut[k] + sdn;
            }
        }
    };
    auto advance = [&](Slot &m, T (*Lj)[LW], T lo, T hi, const T (&sde)[STRIP], int st) {
        if (m.len == STRIP) advance_len(std::integral_constant<int, STRIP>{}, m, Lj, lo, hi, sde, st);
        else if (m.len > 0) advance_len(std::integral_constant<int, 0>{}, m, Lj, lo, hi, sde, st);
    };

    Slot R[RN];
    T P[NS];
    T sde0[STRIP];
#pragma unroll
    for (int j = 0; j < RN; j++) {
        R[j] = Slot{};
        R[j].len = 0;
        R[j].r0 = e.y0;
    }
#pragma unroll
    for (int m = 0; m < NS; m++) P[m] = T(0.0);
#pragma unroll
    for (int k = 0; k < STRIP; k++) sde0[k] = T(0.0);
    const int i0 = s_first - NS, i1 = s_last + NS;
    MarchRaw<T> nxt;
    DD eacc[1] = {{0.0, 0.0}};
    // r and rstore at the cells of the strip that leaves finished at
    // iteration i (strip i-NS, whose slot still holds its rows), 0 elsewhere
    auto end_loads = [&](const auto &f, int i, T (&er)[STRIP], T (&es)[STRIP]) {
        const int sf = i - NS;
        const bool out = out_lane && sf >= s_first && sf <= s_last;
        const unsigned b0 = off(f.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const unsigned bk = out && col_in && k < f.len ? b0 + k * row_b : OOB;
            er[k] = b_re.load(bk);
            es[k] = b_rse.load(bk);
        }
    };
    if (active) fetch(i0, nxt);
    auto body = [&](auto ph, int i) {
        constexpr int PH = decltype(ph)::value;
        T er[STRIP], es[STRIP];   // END: r and rstore of strip i-NS, read before anything else this iteration
        if constexpr (END) end_loads(R[(PH + 1) % RN], i, er, es);
        prepare(i, nxt, R[PH], L[PH]);
        T sde_cur[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde_cur[k] = lane0_value(nxt.sde[k]);
        T hi = R[PH].sd[0];
        if (i < i1) fetch(i + 1, nxt);   // read one strip ahead
#pragma unroll
        for (int m = 0; m < NS; m++) {
            const int j = (PH - 1 - m + 2 * RN) % RN;
            Slot &h = R[j];
            const T last = h.sd[STRIP - 1];
            const int s = i - 1 - m;
            const int dist = s < s_first ? s_first - s : (s > s_last ? s - s_last : 0);
            if (dist <= NS - 1 - m) advance(h, L[j], P[m], hi, sde0, m);
            P[m] = last;
            hi = h.sd[0];
        }
        const Slot &f = R[(PH + 1) % RN];   // strip i-NS, finished
        const int sf = i - NS;
        const bool out = out_lane && sf >= s_first && sf <= s_last;
        const unsigned b0 = off(f.r0);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const unsigned bk = out && col_in && k < f.len ? b0 + k * row_b : OOB;
            b_sdo.store(bk, f.sd[k]);
            b_rto.store(bk, f.rt[k]);
            b_uto.store(bk, f.ut[k]);
        }
        if constexpr (END) {   // k_ppcg_end's (r - rstore).z with z = the utemp just written
#pragma unroll
            for (int k = 0; k < STRIP; k++) {
                const bool o = out && col_in && k < f.len;
                dd_add_prod(eacc[0], (double)(er[k] - es[k]), o ? (double)f.ut[k] : 0.0);
            }
        }
#pragma unroll
        for (int k = 0; k < STRIP; k++) sde0[k] = sde_cur[k];
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    for (int i = i0; active;) {
        body(I0{}, i);
        if (++i > i1) break;
        body(I1{}, i);
        if (++i > i1) break;
        if constexpr (RN > 2) {
            body(I2{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 3) {
            body(I3{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 4) {
            body(I4{}, i);
            if (++i > i1) break;
        }
        if constexpr (RN > 5) {
            body(I5{}, i);
            if (++i > i1) break;
        }
    }
    if constexpr (END) block_reduce<1>(eacc, red);
}

// The fused begin (k_ppcg_begin) as a streaming pass (no reference
// counterpart; the same values): [p' = z + beta p] on the fly, w = A p'
// (stored if w != null), sum p'.w, and with STORE_R rstore = r and sum r.z.
// Each wave owns a band of WAVE columns (lanes 1..WAVE-2 written; lanes 0 and
// WAVE-1 form p' for the band's neighbour columns) and a run of `rows` rows,
// and walks down it keeping p' of the rows above and below in registers, so
// every field row is read once per band (k_ppcg_begin forms p' of a strip's
// rows above and below from memory again).  Columns come from lane shifts.
// Z_FROM_R (jac_diag): `z` holds r, and z = r * (1/diag) is formed on the fly
// with the expressions k_cg_ur_wp forms it with.
constexpr int BEGIN_OUT_X = WAVE - 2;
template <typename T, bool STORE_R, bool Z_FROM_R = false>
__global__ __launch_bounds__(MARCH_WAVES *WAVE) void k_begin_march(Ext e, long pitch, int rows, int with_p,
                                                                   double beta_in, const double *__restrict__ beta_num,
                                                                   const T *__restrict__ p, T *__restrict__ pn,
                                                                   const T *__restrict__ z, T *__restrict__ w,
                                                                   const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                                   const T *__restrict__ r, T *__restrict__ rstore,
                                                                   Red red)
{
    static_assert(!(Z_FROM_R && STORE_R), "r.z with z formed from r is not needed");
    const T beta = (T)(beta_num ? beta_num[0] / beta_in : beta_in);
    const int lane = threadIdx.x & (WAVE - 1);
    const int nbands = (e.x1 - e.x0 + BEGIN_OUT_X) / BEGIN_OUT_X;
    const int ny = e.y1 - e.y0 + 1;
    const int nruns = (ny + rows - 1) / rows;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int wid = blockIdx.x * MARCH_WAVES + wv;
    DD acc[2] = {{0, 0}, {0, 0}};
    if (wid < nbands * nruns) {
        const int band = wid % nbands, run = wid / nbands;
        const int c = e.x0 + band * BEGIN_OUT_X - 1 + lane;
        const bool col_in = c >= e.x0 && c <= e.x1;
        const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
        const bool out = col_in && lane >= 1 && lane <= WAVE - 2;
        const int ya = e.y0 + run * rows;
        const int yb = (ya + rows - 1 < e.y1 ? ya + rows - 1 : e.y1);
        // p' at (c, row): z + beta p inside the extent (with_p), p elsewhere
        auto pval = [&](int row) -> T {
            if (!col_mem || row < e.y0 - 1 || row > e.y1 + 1) return T(0.0);
            const long i = (long)c + (long)row * pitch;
            const bool in = col_in && row >= e.y0 && row <= e.y1;
            if (!(with_p && in)) return p[i];
            const T zv = Z_FROM_R ? z[i] * (T(1.0) / diag_A(Kx, Ky, i, pitch)) : z[i];
            return zv + beta * p[i];
        };
        auto kyv = [&](int row) -> T {   // Ky at (c, row) (rows y0 .. y1+1 are faces of the extent)
            return (col_in && row >= e.y0 && row <= e.y1 + 1) ? Ky[(long)c + (long)row * pitch] : T(0.0);
        };
        T pm = pval(ya - 1), pc = pval(ya);
        T ky0 = kyv(ya);
        if constexpr (Z_FROM_R) {
            // The diagonal of the row ahead from registers: its Kx and the Ky
            // of the row after it are loaded one row early and then serve
            // that row's own stencil (no re-read).  The last lane reads the
            // Kx of the column to its right itself.
            T ky1 = kyv(ya + 1);
            T kx0 = col_mem ? Kx[(long)c + (long)ya * pitch] : T(0.0);
            for (int row = ya; row <= yb; row++) {
                const long i = (long)c + (long)row * pitch, i1 = i + pitch;
                const T ky2 = kyv(row + 2);
                const T kxn = col_mem ? Kx[i1] : T(0.0);   // row + 1 <= y1 + 1: a halo row at most
                const T kxn_r = lane_from_above(kxn);
                const T kx1 = lane_from_above(kx0);
                T pp = T(0.0);
                if (col_mem && row + 1 <= e.y1 + 1) {
                    if (with_p && col_in && row + 1 <= e.y1) {
                        const T kr = lane == WAVE - 1 ? Kx[i1 + 1] : kxn_r;
                        const T diag = T(1.0) + (ky2 + ky1) + (kr + kxn);
                        pp = z[i1] * (T(1.0) / diag) + beta * p[i1];
                    } else {
                        pp = p[i1];
                    }
                }
                const T left = lane_from_below(pc), right = lane_from_above(pc);
                const T wval =
                    (T(1.0) + (ky1 + ky0) + (kx1 + kx0)) * pc - (ky1 * pp + ky0 * pm) - (kx1 * right + kx0 * left);
                if (out) {
                    pn[i] = pc;
                    if (w) w[i] = wval;
                    dd_add_prod(acc[0], (double)pc, (double)wval);
                }
                pm = pc;
                pc = pp;
                ky0 = ky1;
                ky1 = ky2;
                kx0 = kxn;
            }
        } else {
            // STORE_R: z of the row ahead is loaded once, for p' there and for
            // that row's r.z one row later (zc), not read again
            T zc = (STORE_R && col_in) ? z[(long)c + (long)ya * pitch] : T(0.0);
            for (int row = ya; row <= yb; row++) {
                const long i = (long)c + (long)row * pitch;
                T pp, zn = T(0.0);
                if constexpr (STORE_R) {
                    const long i1 = i + pitch;
                    const bool in1 = col_in && row + 1 <= e.y1;
                    if (in1) zn = z[i1];
                    pp = T(0.0);
                    if (col_mem && row + 1 <= e.y1 + 1) pp = (with_p && in1) ? zn + beta * p[i1] : p[i1];
                } else {
                    pp = pval(row + 1);
                }
                const T ky1 = kyv(row + 1);
                const T kx0 = col_mem ? Kx[i] : T(0.0);
                const T kx1 = lane_from_above(kx0);   // Kx on the right face: the next column's
                const T left = lane_from_below(pc), right = lane_from_above(pc);
                const T wval =
                    (T(1.0) + (ky1 + ky0) + (kx1 + kx0)) * pc - (ky1 * pp + ky0 * pm) - (kx1 * right + kx0 * left);
                if (out) {
                    pn[i] = pc;
                    if (w) w[i] = wval;
                    dd_add_prod(acc[0], (double)pc, (double)wval);
                    if (STORE_R) {
                        const T rv = r[i];
                        rstore[i] = rv;
                        dd_add_prod(acc[1], (double)rv, (double)zc);
                    }
                }
                pm = pc;
                pc = pp;
                ky0 = ky1;
                zc = zn;
            }
        }
    }
    block_reduce<2>(acc, red);
}

// One Chebyshev iteration (tea_leaf_cheby.cuknl:84-152: w = A u, r = u0 - w,
// p = alpha p + beta M^-1 r; then calc_u: u += p) as one streaming pass (no
// reference counterpart as a kernel; the same values): u + p goes to u_out
// (neighbouring bands read u), w and r are not stored (the caller owes them:
// both follow from the u read here, which the caller keeps).  Bands of 62
// columns walk down runs of rows with the rows of u above and below in
// registers.  No preconditioner or the diagonal one (block Jacobi needs whole
// strips and keeps the separate kernels).
template <typename T, int PREC>
__global__ __launch_bounds__(MARCH_WAVES *WAVE) void k_cheby_march(Ext e, long pitch, int rows,
                                                                   const T *__restrict__ u, T *__restrict__ u_out,
                                                                   const T *__restrict__ u0, T *__restrict__ p,
                                                                   const T *__restrict__ Mi, const T *__restrict__ Kx,
                                                                   const T *__restrict__ Ky, double alpha_in,
                                                                   double beta_in)
{
    static_assert(PREC != PREC_BLOCK, "block Jacobi keeps the separate kernels");
    const T alpha = (T)alpha_in, beta = (T)beta_in;
    const int lane = threadIdx.x & (WAVE - 1);
    const int nbands = (e.x1 - e.x0 + BEGIN_OUT_X) / BEGIN_OUT_X;
    const int ny = e.y1 - e.y0 + 1;
    const int nruns = (ny + rows - 1) / rows;
    const int wid = blockIdx.x * MARCH_WAVES + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    if (wid >= nbands * nruns) return;
    const int band = wid % nbands, run = wid / nbands;
    const int c = e.x0 + band * BEGIN_OUT_X - 1 + lane;
    const bool col_in = c >= e.x0 && c <= e.x1;
    const bool col_mem = c >= e.x0 - 1 && c <= e.x1 + 1;
    const bool out = col_in && lane >= 1 && lane <= WAVE - 2;
    const int ya = e.y0 + run * rows;
    const int yb = (ya + rows - 1 < e.y1 ? ya + rows - 1 : e.y1);
    auto uval = [&](int row) -> T {
        return (col_mem && row >= e.y0 - 1 && row <= e.y1 + 1) ? u[(long)c + (long)row * pitch] : T(0.0);
    };
    auto kyv = [&](int row) -> T {
        return (col_in && row >= e.y0 && row <= e.y1 + 1) ? Ky[(long)c + (long)row * pitch] : T(0.0);
    };
    T um = uval(ya - 1), uc = uval(ya);
    T ky0 = kyv(ya);
    for (int row = ya; row <= yb; row++) {
        const long i = (long)c + (long)row * pitch;
        const T up = uval(row + 1);
        const T ky1 = kyv(row + 1);
        const T kx0 = col_mem ? Kx[i] : T(0.0);
        const T kx1 = lane_from_above(kx0);
        const T left = lane_from_below(uc), right = lane_from_above(uc);
        const T diag = T(1.0) + (ky1 + ky0) + (kx1 + kx0);
        const T wv = diag * uc - (ky1 * up + ky0 * um) - (kx1 * right + kx0 * left);
        if (out) {
            const T rv = u0[i] - wv;
            T pv;
            // Mi = 1/diag, the expression it was set with (tea_leaf_common.cuknl:141-145), not read
            if (PREC == PREC_DIAG) pv = alpha * p[i] + beta * (T(1.0) / diag) * rv;
            else pv = alpha * p[i] + beta * rv;
            p[i] = pv;
            u_out[i] = uc + pv;
        }
        um = uc;
        uc = up;
        ky0 = ky1;
    }
}

// w = A u and r = u0 - w on the interior, for u the iterate a fused
// Chebyshev iteration read (the w and r it left owed): the expressions of
// k_cheby_p.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_cheby_wr(Ext e, long pitch, const T *__restrict__ u,
                                                    const T *__restrict__ u0, T *__restrict__ r, T *__restrict__ w,
                                                    const T *__restrict__ Kx, const T *__restrict__ Ky)
{
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T wv = apply_A(u, Kx, Ky, i, pitch);
            w[i] = wv;
            r[i] = u0[i] - wv;
        }
    }
    STRIP_END
}

// sum r*r with r = u0 - A u formed on the fly (the r a fused Chebyshev
// iteration left owed): calc_2norm(1) of the r k_cheby_wr would write.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_cheby_rr(Ext e, long pitch, const T *__restrict__ u,
                                                    const T *__restrict__ u0, const T *__restrict__ Kx,
                                                    const T *__restrict__ Ky, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T rv = u0[i] - apply_A(u, Kx, Ky, i, pitch);
            dd_add_prod(acc[0], (double)rv, (double)rv);
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
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
                                                      const double *__restrict__ beta_num,
                                                      const T *__restrict__ p, T *__restrict__ pn,
                                                      const T *__restrict__ z, T *__restrict__ w,
                                                      const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                      const T *__restrict__ r, T *__restrict__ rstore, Red red)
{
    const T beta = (T)(beta_num ? beta_num[0] / beta_in : beta_in);
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
            if (w) w[i] = wv;
            dd_add_prod(acc[0], (double)pc, (double)wv);
            if (rstore) {   // null: CG's begin (no store_r, no r.z)
                const T rv = r[i];
                rstore[i] = rv;
                dd_add_prod(acc[1], (double)rv, (double)z[i]);
            }
        }
    }
    STRIP_END
    block_reduce<2>(acc, red);
}

// k_ppcg_begin for FP32 fields with two columns per lane (8-byte accesses);
// tiles of 2*TILE_X x TILE_Y cells.  The same arithmetic cell by cell.  w
// null: w is not stored (k_ppcg_ur_sd_x2_wp forms it again from p').
__global__ __launch_bounds__(BLOCK) void k_ppcg_begin_x2(Ext e, long pitch, int with_p, double beta_in,
                                                         const double *__restrict__ beta_num,
                                                         const float *__restrict__ p, float *__restrict__ pn,
                                                         const float *__restrict__ z, float *__restrict__ w,
                                                         const float *__restrict__ Kx, const float *__restrict__ Ky,
                                                         const float *__restrict__ r, float *__restrict__ rstore,
                                                         Red red)
{
    const float beta = (float)(beta_num ? beta_num[0] / beta_in : beta_in);
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
                    if (w) *(float2 *)(w + i) = make_float2(w0, w1);
                    *(float2 *)(rstore + i) = rv;
                    dd_add_prod(acc[0], (double)col[1][k + 1], (double)w1);
                    dd_add_prod(acc[1], (double)rv.y, (double)zs[1][k]);
                } else {
                    pn[i] = col[0][k + 1];
                    if (w) w[i] = w0;
                    rstore[i] = rv.x;
                }
            }
        }
    }
    block_reduce<2>(acc, red);
}

// k_ppcg_ur_sd for FP32 fields with two columns per lane.
template <int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_ur_sd_x2(Ext e, long pitch, double alpha_in,
                                                         const double *__restrict__ pw_rro, double theta_in,
                                                         float *__restrict__ u, const float *__restrict__ p,
                                                         float *__restrict__ r, const float *__restrict__ w,
                                                         float *__restrict__ z, const float *__restrict__ Mi,
                                                         const float *__restrict__ Kx, const float *__restrict__ Ky,
                                                         float *__restrict__ sd, float *__restrict__ rtemp,
                                                         float *__restrict__ utemp, Red red)
{
    const float alpha = (float)(pw_rro ? pw_rro[1] / pw_rro[0] : alpha_in);
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
                    if (rtemp) {   // null: rtemp = r and utemp = sd left to the first inner pass
                        *(float2 *)(rtemp + i) = make_float2(rv[0][k], rv[1][k]);
                        *(float2 *)(utemp + i) = make_float2(sdv[0], sdv[1]);
                    }
                } else {
                    if (PREC != PREC_NONE) z[i] = src[0];
                    sd[i] = sdv[0];
                    if (rtemp) {
                        rtemp[i] = rv[0][k];
                        utemp[i] = sdv[0];
                    }
                }
            }
        }
    }
    block_reduce<1>(acc, red);
}

// k_ppcg_ur_sd_x2<PREC_BLOCK> with w = A p formed here, by the expression of
// k_ppcg_begin_x2, instead of read (begin then leaves w unwritten): p down
// both columns of the strip and the rows above and below it, x neighbours from
// the adjacent lanes, the edge lanes' from memory.  p is read, not written, so
// neighbours need no spare buffer; the cells' values are those of
// k_ppcg_begin_x2 followed by k_ppcg_ur_sd_x2<PREC_BLOCK>.
__global__ __launch_bounds__(BLOCK) void k_ppcg_ur_sd_x2_wp(Ext e, long pitch, double alpha_in,
                                                            const double *__restrict__ pw_rro, double theta_in,
                                                            float *__restrict__ u, const float *__restrict__ p,
                                                            float *__restrict__ r, float *__restrict__ z,
                                                            const float *__restrict__ Kx, const float *__restrict__ Ky,
                                                            float *__restrict__ sd, float *__restrict__ rtemp,
                                                            float *__restrict__ utemp, Red red)
{
    const float alpha = (float)(pw_rro ? pw_rro[1] / pw_rro[0] : alpha_in);
    const float theta_r = 1.0f / (float)theta_in;
    const int lane = threadIdx.x & (WAVE - 1);
    DD acc[1] = {{0, 0}};
    for (TileLoopT<2 * TILE_X> tl_(e); tl_.more(); tl_.advance()) {
        const Strip s = tl_.strip();
        const int c = s.c;
        const long i0 = (long)c + (long)s.r0 * pitch;
        int act[2] = {s.len, (c + 1 <= e.x1) ? s.len : 0};
        const bool rlane = lane < WAVE - 1 && c + 2 <= e.x1;
        float col[2][STRIP + 2], ky[2][STRIP + 1];
#pragma unroll
        for (int q = 0; q < STRIP + 2; q++) {
            col[0][q] = col[1][q] = 0.0f;
            if (act[0] > 0 && q <= act[0] + 1) {
                const float2 pp = *(const float2 *)(p + i0 + (long)(q - 1) * pitch);
                col[0][q] = pp.x;
                col[1][q] = pp.y;
            }
        }
#pragma unroll
        for (int k = 0; k < STRIP + 1; k++) {
            ky[0][k] = ky[1][k] = 0.0f;
            if (act[0] > 0 && k <= act[0]) {
                const float2 v = *(const float2 *)(Ky + i0 + (long)k * pitch);
                ky[0][k] = v.x;
                ky[1][k] = v.y;
            }
        }
        float rv[2][STRIP], zv[2][STRIP], diag[2][STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            const float from_left = __shfl_up(col[1][k + 1], 1, WAVE);
            const float from_right = __shfl_down(col[0][k + 1], 1, WAVE);
            rv[0][k] = rv[1][k] = zv[0][k] = zv[1][k] = 0.0f;
            diag[0][k] = diag[1][k] = 1.0f;
            if (k < act[0]) {
                const long i = i0 + (long)k * pitch;
                const float2 kx = *(const float2 *)(Kx + i);
                const float left = lane > 0 ? from_left : p[i - 1];
                const float w0 = (1.0f + (ky[0][k + 1] + ky[0][k]) + (kx.y + kx.x)) * col[0][k + 1]
                                 - (ky[0][k + 1] * col[0][k + 2] + ky[0][k] * col[0][k])
                                 - (kx.y * col[1][k + 1] + kx.x * left);
                float2 uu = *(const float2 *)(u + i), rr = *(const float2 *)(r + i);
                uu.x += alpha * col[0][k + 1];
                rv[0][k] = rr.x - alpha * w0;
                diag[0][k] = 1.0f + (ky[0][k + 1] + ky[0][k]) + (kx.y + kx.x);
                if (k < act[1]) {
                    const float kxr = Kx[i + 2];
                    const float right = rlane ? from_right : p[i + 2];
                    const float w1 = (1.0f + (ky[1][k + 1] + ky[1][k]) + (kxr + kx.y)) * col[1][k + 1]
                                     - (ky[1][k + 1] * col[1][k + 2] + ky[1][k] * col[1][k])
                                     - (kxr * right + kx.y * col[0][k + 1]);
                    uu.y += alpha * col[1][k + 1];
                    rv[1][k] = rr.y - alpha * w1;
                    diag[1][k] = 1.0f + (ky[1][k + 1] + ky[1][k]) + (kxr + kx.y);
                    *(float2 *)(u + i) = uu;
                    *(float2 *)(r + i) = make_float2(rv[0][k], rv[1][k]);
                } else {
                    u[i] = uu.x;
                    r[i] = rv[0][k];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (act[q] > 0) strip_solve_col(rv[q], zv[q], act[q], diag[q], ky[q]);
#pragma unroll
        for (int k = 0; k < STRIP; k++) {
            if (k < act[0]) {
                const long i = i0 + (long)k * pitch;
                float sdv[2];
#pragma unroll
                for (int q = 0; q < 2; q++) sdv[q] = zv[q][k] * theta_r;
                const int nq = k < act[1] ? 2 : 1;
                for (int q = 0; q < nq; q++) dd_add_prod(acc[0], (double)rv[q][k], (double)zv[q][k]);
                // z, sd null: both left to the first inner pass, which forms them from r
                if (nq == 2) {
                    if (z) *(float2 *)(z + i) = make_float2(zv[0][k], zv[1][k]);
                    if (sd) *(float2 *)(sd + i) = make_float2(sdv[0], sdv[1]);
                    if (rtemp) {   // null: rtemp = r and utemp = sd left to the first inner pass
                        *(float2 *)(rtemp + i) = make_float2(rv[0][k], rv[1][k]);
                        *(float2 *)(utemp + i) = make_float2(sdv[0], sdv[1]);
                    }
                } else {
                    if (z) z[i] = zv[0][k];
                    if (sd) sd[i] = sdv[0];
                    if (rtemp) {
                        rtemp[i] = rv[0][k];
                        utemp[i] = sdv[0];
                    }
                }
            }
        }
    }
    block_reduce<1>(acc, red);
}

// k_ppcg_ur_sd<T, PREC_BLOCK> with w = A p formed here, by the expression of
// k_ppcg_begin (which then leaves w unwritten), instead of read: p down the
// lane's column (the strip and the rows above and below it), x neighbours
// from the adjacent lanes, the edge lanes' from memory.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_ur_sd_wp(Ext e, long pitch, double alpha_in,
                                                         const double *__restrict__ pw_rro, double theta_in,
                                                         T *__restrict__ u, const T *__restrict__ p,
                                                         T *__restrict__ r, T *__restrict__ z,
                                                         const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                         T *__restrict__ sd, T *__restrict__ rtemp,
                                                         T *__restrict__ utemp, Red red)
{
    const T alpha = (T)(pw_rro ? pw_rro[1] / pw_rro[0] : alpha_in);
    const T theta_r = T(1.0) / (T)theta_in;
    const int lane = threadIdx.x & (WAVE - 1);
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    const int c = s.c;
    T col[STRIP + 2];
#pragma unroll
    for (int q = 0; q < STRIP + 2; q++) {
        col[q] = T(0.0);
        if (s.len > 0 && q <= s.len + 1) col[q] = p[i0 + (long)(q - 1) * pitch];
    }
    const bool rlane = lane < WAVE - 1 && c + 1 <= e.x1;
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        const T from_left = __shfl_up(col[k + 1], 1, WAVE);
        const T from_right = __shfl_down(col[k + 1], 1, WAVE);
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T left = lane > 0 ? from_left : p[i - 1];
            const T right = rlane ? from_right : p[i + 1];
            const T pc = col[k + 1];
            const T wv = (T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i])) * pc
                         - (Ky[i + pitch] * col[k + 2] + Ky[i] * col[k])
                         - (Kx[i + 1] * right + Kx[i] * left);
            u[i] += alpha * pc;
            rv[k] = r[i] - alpha * wv;
            r[i] = rv[k];
        }
    }
    precond_strip<T, PREC_BLOCK>(rv, zv, s.len, i0, pitch, (const T *)nullptr, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            if (z) z[i] = zv[k];   // z, sd null: left to the first inner pass, which forms them from r
            dd_add_prod(acc[0], (double)rv[k], (double)zv[k]);
            const T sdv = zv[k] * theta_r;
            if (sd) sd[i] = sdv;
            if (rtemp) {   // null: rtemp = r and utemp = sd left to the first inner pass
                rtemp[i] = rv[k];
                utemp[i] = sdv;
            }
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// CG's calc_ur (tea_leaf_cg.cuknl:72-137) with w = A p formed again from p
// (no reference counterpart as a kernel: the fused CG iteration's begin
// leaves w unwritten), alpha = rro / pw formed on the device from begin's
// p.w: u += alpha p; r -= alpha w; z = M^-1 r; sum r*z (r*r).  The same
// expressions as cg_calc_w and cg_calc_ur, so the same values.  jac_diag:
// Mi = 1/diag is formed from the stencil's own diagonal (the expression of
// tea_leaf_common.cuknl:141-145, which is how Mi was set) instead of read,
// and z is not stored when z is null (the next begin forms it from r).
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_cg_ur_wp(Ext e, long pitch, double rro, const double *__restrict__ pw,
                                                    T *__restrict__ u, const T *__restrict__ p, T *__restrict__ r,
                                                    T *__restrict__ z, const T *__restrict__ Mi,
                                                    const T *__restrict__ Kx, const T *__restrict__ Ky, Red red)
{
    const T alpha = (T)(rro / pw[0]);
    const int lane = threadIdx.x & (WAVE - 1);
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    const int c = s.c;
    T col[STRIP + 2];
#pragma unroll
    for (int q = 0; q < STRIP + 2; q++) {
        col[q] = T(0.0);
        if (s.len > 0 && q <= s.len + 1) col[q] = p[i0 + (long)(q - 1) * pitch];
    }
    const bool rlane = lane < WAVE - 1 && c + 1 <= e.x1;
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0}, mi[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        const T from_left = __shfl_up(col[k + 1], 1, WAVE);
        const T from_right = __shfl_down(col[k + 1], 1, WAVE);
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T left = lane > 0 ? from_left : p[i - 1];
            const T right = rlane ? from_right : p[i + 1];
            const T pc = col[k + 1];
            const T diag = T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i]);
            const T wv = diag * pc
                         - (Ky[i + pitch] * col[k + 2] + Ky[i] * col[k])
                         - (Kx[i + 1] * right + Kx[i] * left);
            if (PREC == PREC_DIAG) mi[k] = T(1.0) / diag;
            u[i] += alpha * pc;
            rv[k] = r[i] - alpha * wv;
            r[i] = rv[k];
        }
    }
    if (PREC == PREC_DIAG) {
#pragma unroll
        for (int k = 0; k < STRIP; k++) zv[k] = rv[k] * mi[k];
    } else if (PREC != PREC_NONE) {
        precond_strip<T, PREC>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
    }
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            if (PREC != PREC_NONE) {
                if (z) z[i0 + k * pitch] = zv[k];
                dd_add_prod(acc[0], (double)rv[k], (double)zv[k]);
            } else {
                dd_add_prod(acc[0], (double)rv[k], (double)rv[k]);
            }
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// cg_calc_ur (u += a p; r -= a w; z = M^-1 r; r.z or r.r) then
// ppcg_init_sd_new (sd = (z or r) / theta; rtemp = r; utemp = sd).  With
// pw_rro (device memory: p.w and r.z of k_ppcg_begin) a = rro / pw is formed
// here, as the host would form it, instead of taken from alpha_in.
template <typename T, int PREC>
__global__ __launch_bounds__(BLOCK) void k_ppcg_ur_sd(Ext e, long pitch, double alpha_in,
                                                      const double *__restrict__ pw_rro, double theta_in,
                                                      T *__restrict__ u, const T *__restrict__ p,
                                                      T *__restrict__ r, const T *__restrict__ w,
                                                      T *__restrict__ z, const T *__restrict__ Mi,
                                                      const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                      T *__restrict__ sd, T *__restrict__ rtemp,
                                                      T *__restrict__ utemp, Red red)
{
    const T alpha = (T)(pw_rro ? pw_rro[1] / pw_rro[0] : alpha_in);
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
            if (rtemp) {   // null: rtemp = r and utemp = sd left to the first inner pass
                rtemp[i] = rv[k];
                utemp[i] = sdv;
            }
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// ppcg_update_z (z = utemp) then ppcg_calc_rrn ((r - rstore).z).  z null:
// the store is left out (the caller overwrites z before anything reads it).
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
            if (z) z[i] = zv;
            dd_add_prod(acc[0], (double)(r[i] - rstore[i]), (double)zv);
        }
    STRIP_END
    block_reduce<1>(acc, red);
}

// k_ppcg_end followed by the next fused begin's store_r and r.z (no reference
// counterpart as a kernel; the same values): rstore = r and sum r.utemp,
// which the begin would form from the same r (nothing changes r in between)
// and utemp (the z it reads); the begin then reads neither r nor rstore.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_end_rr(Ext e, long pitch, const T *__restrict__ utemp,
                                                       T *__restrict__ z, const T *__restrict__ r,
                                                       T *__restrict__ rstore, Red red)
{
    DD acc[2] = {{0, 0}, {0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T zv = utemp[i], rv = r[i];
            if (z) z[i] = zv;
            dd_add_prod(acc[0], (double)(rv - rstore[i]), (double)zv);
            rstore[i] = rv;
            dd_add_prod(acc[1], (double)rv, (double)zv);
        }
    STRIP_END
    block_reduce<2>(acc, red);
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

// r32 = (float) r64; the FP32 right-hand side u0_32 = r32 too (Chebyshev reads
// it); u32 = 0; sum r32*r32 (compensated FP64).
__global__ __launch_bounds__(BLOCK) void k_mp_begin(Ext e, long pitch, long pitch32, const double *__restrict__ r,
                                                    float *__restrict__ r32, float *__restrict__ u0_32,
                                                    float *__restrict__ u32, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const long j = (long)s.c + (long)(s.r0 + k) * pitch32;
        const float rv = (float)r[i];
        r32[j] = rv;
        u0_32[j] = rv;
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

// a = v on the extent
template <typename T> __global__ __launch_bounds__(BLOCK) void k_fill(Ext e, long pitch, T *__restrict__ a, T v)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) a[i0 + k * pitch] = v;
    STRIP_END
}

// A residual replacement of the FP32 inner solver in one pass: the composition
// of k_mp_end (u' = u + (double) u32 on the interior), k_calc_residual
// (r = u0 - A u'), the FP64 r.r, k_mp_begin (r32 = (float) r, u0_32 = r32,
// sum r32*r32), k_ppcg_init step 2 (z32 = M^-1 r32 with the FP32
// coefficients) and k_ppcg_init_sd_new (sd32 = (z32 or r32) * (1/theta),
// rtemp32 = r32, utemp32 = sd32), with their expressions, so the same values.
// u' goes to u_out (u is read at the neighbours other lanes own); u32 = 0 is
// left to a following launch (u32 is read at those neighbours too).  Each lane
// forms u' down its column (the strip and the rows above and below it) and
// takes its x neighbours from the adjacent lanes; edge lanes form those
// themselves.  OWE (block Jacobi): z32, sd32, rtemp32 and utemp32 are not
// formed; the caller leaves them owed from r32, which the first inner pass
// reads to form them with the same expressions.
template <int PREC, bool OWE = false>
__global__ __launch_bounds__(BLOCK) void k_mp_replace(
    Ext e, long pitch, long pitch32, const double *__restrict__ u, double *__restrict__ u_out,
    const double *__restrict__ u0, double *__restrict__ r, const double *__restrict__ Kx,
    const double *__restrict__ Ky, const float *__restrict__ u32, float *__restrict__ r32, float *__restrict__ u0_32,
    float *__restrict__ z32, float *__restrict__ sd32, float *__restrict__ rtemp32, float *__restrict__ utemp32,
    const float *__restrict__ Mi32, const float *__restrict__ Kx32, const float *__restrict__ Ky32, double theta_in,
    Red red)
{
    const float theta_r = 1.0f / (float)theta_in;
    const int lane = threadIdx.x & (WAVE - 1);
    DD acc[2] = {{0, 0}, {0, 0}};
    auto inside = [&](int c, int rr) { return c >= e.x0 && c <= e.x1 && rr >= e.y0 && rr <= e.y1; };
    auto unew = [&](int c, int rr) -> double {
        const long i = (long)c + (long)rr * pitch;
        return inside(c, rr) ? u[i] + (double)u32[(long)c + (long)rr * pitch32] : u[i];
    };
    STRIP_BEGIN
    const int c = s.c;
    const long j0 = (long)c + (long)s.r0 * pitch32;
    double col[STRIP + 2];
#pragma unroll
    for (int q = 0; q < STRIP + 2; q++) {
        col[q] = 0.0;
        if (s.len > 0 && q <= s.len + 1) col[q] = unew(c, s.r0 - 1 + q);
    }
    const bool rlane = lane < WAVE - 1 && c + 1 <= e.x1;
    float rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        const double from_left = __shfl_up(col[k + 1], 1, WAVE);
        const double from_right = __shfl_down(col[k + 1], 1, WAVE);
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const int rr = s.r0 + k;
            const double left = lane > 0 ? from_left : unew(c - 1, rr);
            const double right = rlane ? from_right : unew(c + 1, rr);
            const double smvp = (1.0 + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i])) * col[k + 1]
                                - (Ky[i + pitch] * col[k + 2] + Ky[i] * col[k])
                                - (Kx[i + 1] * right + Kx[i] * left);
            const double rd = u0[i] - smvp;
            u_out[i] = col[k + 1];
            r[i] = rd;
            dd_add_prod(acc[0], rd, rd);
            rv[k] = (float)rd;
            dd_add_prod(acc[1], (double)rv[k], (double)rv[k]);
        }
    }
    if (PREC != PREC_NONE && !OWE) precond_strip<float, PREC>(rv, zv, s.len, j0, pitch32, Mi32, Kx32, Ky32);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long j = j0 + k * pitch32;
            r32[j] = rv[k];
            u0_32[j] = rv[k];
            if (OWE) continue;
            if (PREC != PREC_NONE) z32[j] = zv[k];
            const float sdv = (PREC != PREC_NONE ? zv[k] : rv[k]) * theta_r;
            sd32[j] = sdv;
            rtemp32[j] = rv[k];
            utemp32[j] = sdv;
        }
    }
    STRIP_END
    block_reduce<2>(acc, red);
}

}  // namespace tea
