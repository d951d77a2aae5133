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
// updates the interior.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ppcg_begin(Ext e, long pitch, int with_p, double beta_in,
                                                      const T *__restrict__ p, T *__restrict__ pn,
                                                      const T *__restrict__ z, T *__restrict__ w,
                                                      const T *__restrict__ Kx, const T *__restrict__ Ky,
                                                      const T *__restrict__ r, T *__restrict__ rstore, Red red)
{
    const T beta = (T)beta_in;
    DD acc[2] = {{0, 0}, {0, 0}};
    auto pv = [&](long j, int c, int rr) -> T {
        const bool in = c >= e.x0 && c <= e.x1 && rr >= e.y0 && rr <= e.y1;
        return (with_p && in) ? z[j] + beta * p[j] : p[j];
    };
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const int c = s.c, rr = s.r0 + k;
            const T pc = pv(i, c, rr);
            const T wv = (T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i])) * pc
                         - (Ky[i + pitch] * pv(i + pitch, c, rr + 1) + Ky[i] * pv(i - pitch, c, rr - 1))
                         - (Kx[i + 1] * pv(i + 1, c + 1, rr) + Kx[i] * pv(i - 1, c - 1, rr));
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
