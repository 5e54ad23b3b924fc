// gfx950 kernels of the two-phase AMR-LBM step: FP64 (the reference's
// arithmetic) and FP32 (the production form) instantiations.
//
// Layout (one slot per reference block id, slot = id - 1):
//   pop[b][slot][set][q][pcell(x,y)]  x,y in 0..17 (1-cell ghost ring), set 0 = h, 1 = g
//   C[slot][y+1][x+1]            x,y in -1..18 (2 ghost rings)
//   PU[slot][k][j-1][i-1]        k = P, Ux, Uy
// One workgroup of 256 threads owns one 16 x 16 block: thread t <-> cell
// (t % 16 + 1, t / 16 + 1), so each population row is one 128-byte run.
//
// Arithmetic follows the reference expression by expression (left-to-right
// sums, integer weights converted before the multiply, divisions kept as
// divisions) and the library is built with -ffp-contract=off, so results
// match the reference build to the last bit wherever the GPU's IEEE division
// and square root round as the host's do.
//
// FP32 stores h - w_q and C - 1 (deviations from the liquid at rest; g and P
// are deviations from zero already). Every weight set of the fills, the
// stream, prolong and restrict sums to 1 and bounce-back swaps populations of
// equal weight, so those kernels are the same code on either storage; only
// the collision and U, P have FP32 forms of their own (deviation form).
#include <type_traits>

#include "amrlbm_internal.h"

namespace amrlbm {

// D2Q9 directions (D2Q9_Mod.f90:6-7). Compile-time tables: in the unrolled
// population loops every use folds to a constant (LDS offsets, products by
// +-1 exact in IEEE arithmetic), where a __constant__ array was loaded and
// converted at run time.
constexpr int c_ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int c_ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
// opposite population (BC_Cell.f90:15, J = (i-2) + 4 MOD(i/2,2)), 0-based
constexpr int c_opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
__constant__ int c_dsx[8] = {-1, 1, 0, 0, -1, -1, 1, 1};
__constant__ int c_dsy[8] = {0, 0, -1, 1, -1, 1, -1, 1};

// 1 / dx^p for the power-of-two cell size dx: 2^(-p log2 dx), one exponent
// adjustment with exactly the value of the IEEE division (cells are never
// finer than 1 nor coarser than 2^14, so no exponent range is left)
__device__ __forceinline__ double inv_pow2(int dx, int p) { return ldexp(1.0, -p * (__ffs(dx) - 1)); }
__device__ __forceinline__ float inv_pow2f(int dx, int p) { return ldexpf(1.0f, -p * (__ffs(dx) - 1)); }
template <class T>
__device__ __forceinline__ T inv_dx_of(int dx)
{
    if constexpr (std::is_same<T, float>::value)
        return inv_pow2f(dx, 1);
    else
        return inv_pow2(dx, 1);
}

// lattice weights Wa = [16,4,4,4,4,1,1,1,1]/36 (D2Q9_Mod.f90:8)
__device__ __forceinline__ double wa(int q) { return q == 0 ? 16.0 / 36.0 : (q < 5 ? 4.0 / 36.0 : 1.0 / 36.0); }

__device__ __forceinline__ int fidx(int set, int q, int x, int y) { return (set * 9 + q) * GS + pcell(x, y); }
__device__ __forceinline__ int cidx(int x, int y) { return (y + 1) * CW + (x + 1); }
__device__ __forceinline__ int pidx(int i, int j) { return (i - 1) + NX * (j - 1); }

// ---------------------------------------------------------------------------
// Cell-level numerics (LBM_Subroutines.f90)
// ---------------------------------------------------------------------------

// x / d, correctly rounded, from y = RN(1 / d) by two Markstein corrections
// (P. W. Markstein, IBM J. Res. Dev. 34(1), 1990, and Muller et al., Handbook
// of Floating-Point Arithmetic, Markstein's theorem): RN(x y) is within 1.5
// ulp of x / d; one correction q + (x - d q) y makes it faithful (within one
// ulp); with a faithful q the residual x - d q is exact in an FMA and
// RN(q + (x - d q) y) is RN(x / d). The residual is taken negated, d q - x
// (RN is odd, so the same value up to sign), which keeps the sign of a zero x.
// One multiply and four FMAs instead of the generic eleven-instruction
// division, with the same result for x = +-0 and every 2^-969 <= |x| < 2^1020
// (below, the residual can fall under the subnormal grid); NaN stays NaN,
// +-Inf gives NaN.
__device__ __forceinline__ double div_rn(double x, double d, double y)
{
    double q = x * y;
    double r = fma(q, d, -x);
    q = fma(-r, y, q);
    r = fma(q, d, -x);
    return fma(-r, y, q);
}
template <int D> __device__ __forceinline__ double div_c(double x) { return div_rn(x, (double)D, 1.0 / (double)D); }

// Laplacian_AMR (LBM_Subroutines.f90:316-333) + Chemical_Potential_AMR (:94-112)
__device__ __forceinline__ double chem_pot(const double *sC, int i, int j, int dx, const Phys &ph)
{
    double d2c = div_c<6>(sC[cidx(i - 1, j - 1)] + sC[cidx(i + 1, j - 1)] + sC[cidx(i - 1, j + 1)] +
                          sC[cidx(i + 1, j + 1)] +
                          4.0 * (sC[cidx(i, j - 1)] + sC[cidx(i - 1, j)] + sC[cidx(i + 1, j)] + sC[cidx(i, j + 1)]) -
                          20.0 * sC[cidx(i, j)]);
    double c = sC[cidx(i, j)];
    // x / dx^2 as x * (1 / dx^2): dx is a power of two, so both are the correctly
    // rounded x / dx^2 (the same bits, subnormals included)
    return 4.0 * ph.beta * c * (c - 1.0) * (c - 0.5) - ph.kappa * d2c * inv_pow2(dx, 2);
}

// Gradient_ED_AMR (LBM_Subroutines.f90:281-289)
__device__ __forceinline__ void gradient_ed(const double e[9], double &gx, double &gy)
{
    gx = div_c<3>(e[1] - e[3]) + div_c<12>(e[5] - e[6] - e[7] + e[8]);
    gy = div_c<3>(e[2] - e[4]) + div_c<12>(e[5] + e[6] - e[7] - e[8]);
}

// ED_C_AMR (LBM_Subroutines.f90:244-258); / dx as * (1 / dx), exact for the
// power-of-two dx
__device__ __forceinline__ void ed_c(const double *sC, int i, int j, int dx, double e[9])
{
    const double inv_dx = inv_pow2(dx, 1);
#pragma unroll
    for (int k = 0; k < 9; k++)
        e[k] = (sC[cidx(i + c_ex[k], j + c_ey[k])] - sC[cidx(i - c_ex[k], j - c_ey[k])]) * 0.5 * inv_dx;
}

// ED_M_AMR (LBM_Subroutines.f90:262-277)
__device__ __forceinline__ void ed_m(const double *sC, int i, int j, int dx, double e[9])
{
    const double inv_dx = inv_pow2(dx, 1);
#pragma unroll
    for (int k = 0; k < 9; k++)
        e[k] = (-sC[cidx(i - c_ex[k], j - c_ey[k])] - 3.0 * sC[cidx(i, j)] + 5.0 * sC[cidx(i + c_ex[k], j + c_ey[k])] -
                sC[cidx(i + 2 * c_ex[k], j + 2 * c_ey[k])]) *
               0.25 * inv_dx;
}

// Equilibrium_h_g_AMR (LBM_Subroutines.f90:2-21)
__device__ __forceinline__ void equilibrium(double C, double U, double V, double ni, double nj, const Phys &ph,
                                            double Gamma[9], double heq[9])
{
    double tmp = div_rn(12.0 * (1.0 - C), ph.W, ph.inv_W) * ph.M0;
    double U2 = U * U + V * V;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        double eU = (double)c_ex[i] * U + (double)c_ey[i] * V;
        double eN = (double)c_ex[i] * ni + (double)c_ey[i] * nj;
        Gamma[i] = wa(i) * ((1.0 + eU * (3.0 + 4.5 * eU)) - 1.5 * U2);
        heq[i] = C * (Gamma[i] + tmp * wa(i) * eN);
    }
}

// hlp of LBM_Subroutines.f90:154-155 and :164-165
__device__ __forceinline__ void forcing(const double Gamma[9], double dRho3, double mu, const double e[9], double Ux,
                                        double Uy, double gx, double gy, double Fb, double hlp[9])
{
#pragma unroll
    for (int i = 0; i < 9; i++)
        hlp[i] = ((Gamma[i] - wa(i)) * dRho3 + Gamma[i] * mu) * (e[i] - (Ux * gx + Uy * gy)) +
                 ((double)c_ey[i] - Uy) * Gamma[i] * Fb;
}

// MRT_Collision_AMR (LBM_Subroutines.f90:348-361) with Convert_D2Q9_standard
// (:365-380) and Reconvt_D2Q9_standard (:384-401)
__device__ __forceinline__ void mrt(double s9, const double geq[9], double g[9])
{
    double f[9], o[9], t[9];
#pragma unroll
    for (int i = 0; i < 9; i++) f[i] = g[i] - geq[i];
    o[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    o[1] = -4.0 * f[0] - f[1] - f[2] - f[3] - f[4] + 2.0 * (f[5] + f[6] + f[7] + f[8]);
    o[2] = 4.0 * f[0] - 2.0 * (f[1] + f[2] + f[3] + f[4]) + f[5] + f[6] + f[7] + f[8];
    o[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    o[4] = -2.0 * (f[1] - f[3]) + f[5] - f[6] - f[7] + f[8];
    o[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    o[6] = -2.0 * (f[2] - f[4]) + f[5] + f[6] - f[7] - f[8];
    o[7] = f[1] - f[2] + f[3] - f[4];
    o[8] = f[5] - f[6] + f[7] - f[8];
#pragma unroll
    for (int i = 0; i < 7; i++) t[i] = o[i] * 1.0;
    t[7] = o[7] * s9;
    t[8] = o[8] * s9;
    o[0] = 4.0 * (t[0] - t[1] + t[2]);
    o[1] = 4.0 * t[0] - t[1] - 2.0 * t[2] + 6.0 * (t[3] - t[4]) + 9.0 * t[7];
    o[2] = 4.0 * t[0] - t[1] - 2.0 * t[2] + 6.0 * (t[5] - t[6]) - 9.0 * t[7];
    o[3] = 4.0 * t[0] - t[1] - 2.0 * t[2] - 6.0 * (t[3] - t[4]) + 9.0 * t[7];
    o[4] = 4.0 * t[0] - t[1] - 2.0 * t[2] - 6.0 * (t[5] - t[6]) - 9.0 * t[7];
    o[5] = 4.0 * t[0] + 2.0 * t[1] + t[2] + 6.0 * t[3] + 3.0 * t[4] + 6.0 * t[5] + 3.0 * t[6] + 9.0 * t[8];
    o[6] = 4.0 * t[0] + 2.0 * t[1] + t[2] - 6.0 * t[3] - 3.0 * t[4] + 6.0 * t[5] + 3.0 * t[6] - 9.0 * t[8];
    o[7] = 4.0 * t[0] + 2.0 * t[1] + t[2] - 6.0 * t[3] - 3.0 * t[4] - 6.0 * t[5] - 3.0 * t[6] + 9.0 * t[8];
    o[8] = 4.0 * t[0] + 2.0 * t[1] + t[2] + 6.0 * t[3] + 3.0 * t[4] - 6.0 * t[5] - 3.0 * t[6] - 9.0 * t[8];
#pragma unroll
    for (int i = 0; i < 9; i++) g[i] = g[i] - div_c<36>(o[i]);
}

// Collision_h_g_AMR (LBM_Subroutines.f90:116-189) for cell (i,j); h, g in registers
__device__ __forceinline__ void collide_cell(const double *sC, int i, int j, int dx, double P, double Ux, double Uy,
                                             const Phys &ph, double h[9], double g[9])
{
    const double dRho3 = ph.drho3;
    double e[9], hlp[9], Gamma[9], heq[9], geq[9], gx, gy, ni, nj;
    double mu = chem_pot(sC, i, j, dx, ph);
    double c = sC[cidx(i, j)];
    ed_c(sC, i, j, dx, e);
    gradient_ed(e, gx, gy);
    double tmp = sqrt(gx * gx + gy * gy + 1.0e-32);   // normal_AMR (:25-37)
    ni = gx / tmp;
    nj = gy / tmp;
    equilibrium(c, Ux, Uy, ni, nj, ph, Gamma, heq);
#pragma unroll
    for (int k = 0; k < 9; k++) h[k] = h[k] * (1.0 - ph.wc) + heq[k] * ph.wc;
    double Rho = ph.rhol + c * (ph.rhoh - ph.rhol);
    double Fb = (ph.rhoh - Rho) * ph.gy;
    forcing(Gamma, dRho3, mu, e, Ux, Uy, gx, gy, Fb, hlp);
#pragma unroll
    for (int k = 0; k < 9; k++) geq[k] = P * wa(k) + div_c<3>((Gamma[k] - wa(k)) * Rho) - 0.5 * hlp[k];
    ed_m(sC, i, j, dx, e);
    gradient_ed(e, gx, gy);
    forcing(Gamma, dRho3, mu, e, Ux, Uy, gx, gy, Fb, hlp);
    double tau;
    if (c < 0.0)
        tau = ph.taul + 0.5;
    else if (c > 1.0)
        tau = ph.tauh + 0.5;
    else {
        tau = c * (ph.inv_tauh - ph.inv_taul) + ph.inv_taul;
        tau = 1.0 / tau + 0.5;
    }
    mrt(1.0 / tau, geq, g);
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = g[k] + hlp[k];
}

// Macroscopic_Values_g_AMR (LBM_Subroutines.f90:205-240) with Gradient_AMR (:293-312)
__device__ __forceinline__ void macro_g_cell(const double *sC, int i, int j, int dx, const double g[9], const Phys &ph,
                                             double &P, double &Ux, double &Uy)
{
    const double dRho6 = ph.drho6;
    const double inv_dx = inv_pow2(dx, 1);   // exact: dx is a power of two
    double gx = (div_c<3>(sC[cidx(i + 1, j)] - sC[cidx(i - 1, j)]) +
                 div_c<12>(sC[cidx(i + 1, j + 1)] + sC[cidx(i + 1, j - 1)] - sC[cidx(i - 1, j + 1)] -
                           sC[cidx(i - 1, j - 1)])) *
                inv_dx;
    double gy = (div_c<3>(sC[cidx(i, j + 1)] - sC[cidx(i, j - 1)]) +
                 div_c<12>(sC[cidx(i + 1, j + 1)] + sC[cidx(i - 1, j + 1)] - sC[cidx(i + 1, j - 1)] -
                           sC[cidx(i - 1, j - 1)])) *
                inv_dx;
    double mu = chem_pot(sC, i, j, dx, ph);
    double Rho = ph.rhol + sC[cidx(i, j)] * (ph.rhoh - ph.rhol);
    double Fb = (ph.rhoh - Rho) * ph.gy;
    double Fx = mu * gx;
    double Fy = mu * gy + Fb;
    double sx = 0.0, sy = 0.0, s = 0.0;
#pragma unroll
    for (int k = 1; k < 9; k++) sx += g[k] * (double)c_ex[k];
#pragma unroll
    for (int k = 1; k < 9; k++) sy += g[k] * (double)c_ey[k];
#pragma unroll
    for (int k = 0; k < 9; k++) s += g[k];
    Ux = (3.0 * sx + 0.5 * Fx) / Rho;
    Uy = (3.0 * sy + 0.5 * Fy) / Rho;
    P = s + dRho6 * (Ux * gx + Uy * gy);
}

// ---------------------------------------------------------------------------
// FP32 cell numerics in deviation form: sC holds C~ = C - 1, h holds h - w_q,
// g, P, U are stored as they are. Same formulas as above with C = C~ + 1
// substituted where C enters absolutely (chemical potential, density,
// equilibrium, relaxation time), constant quotients taken as products, and
// the multiply-adds of each expression fused (one rounding each). The fusions
// are fixed per source expression (contract(on): the front end marks each
// a * b + c), not left to the back end across statements, so every kernel
// that inlines these functions computes the same bits.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float waf(int q) { return q == 0 ? 16.0f / 36.0f : (q < 5 ? 4.0f / 36.0f : 1.0f / 36.0f); }

// 4 beta C (C - 1)(C - 1/2) - kappa lap(C) / dx^2 (the Laplacian's weights sum to 0)
__device__ __forceinline__ float chem_pot(const float *sC, int i, int j, float inv_dx2, const PhysF &ph)
{
#pragma clang fp contract(on)
    float d2c = (sC[cidx(i - 1, j - 1)] + sC[cidx(i + 1, j - 1)] + sC[cidx(i - 1, j + 1)] + sC[cidx(i + 1, j + 1)] +
                 4.0f * (sC[cidx(i, j - 1)] + sC[cidx(i - 1, j)] + sC[cidx(i + 1, j)] + sC[cidx(i, j + 1)]) -
                 20.0f * sC[cidx(i, j)]) *
                (1.0f / 6.0f);
    float ct = sC[cidx(i, j)];
    return 4.0f * ph.beta * (ct + 1.0f) * ct * (ct + 0.5f) - ph.kappa * d2c * inv_dx2;
}

__device__ __forceinline__ void gradient_ed(const float e[9], float &gx, float &gy)
{
#pragma clang fp contract(on)
    gx = (e[1] - e[3]) * (1.0f / 3.0f) + (e[5] - e[6] - e[7] + e[8]) * (1.0f / 12.0f);
    gy = (e[2] - e[4]) * (1.0f / 3.0f) + (e[5] + e[6] - e[7] - e[8]) * (1.0f / 12.0f);
}

__device__ __forceinline__ void ed_c(const float *sC, int i, int j, float inv_dx, float e[9])
{
#pragma clang fp contract(on)
#pragma unroll
    for (int k = 0; k < 9; k++)
        e[k] = (sC[cidx(i + c_ex[k], j + c_ey[k])] - sC[cidx(i - c_ex[k], j - c_ey[k])]) * (0.5f * inv_dx);
}

__device__ __forceinline__ void ed_m(const float *sC, int i, int j, float inv_dx, float e[9])
{
#pragma clang fp contract(on)
#pragma unroll
    for (int k = 0; k < 9; k++)
        e[k] = (-sC[cidx(i - c_ex[k], j - c_ey[k])] - 3.0f * sC[cidx(i, j)] + 5.0f * sC[cidx(i + c_ex[k], j + c_ey[k])] -
                sC[cidx(i + 2 * c_ex[k], j + 2 * c_ey[k])]) *
               (0.25f * inv_dx);
}

// Gamma - w and heq - w (Equilibrium_h_g_AMR, LBM_Subroutines.f90:2-21):
// heq - w = C~ Gamma + (Gamma - w) + C (12 (1 - C) / W) M0 w e.n
// The nine sum to C~ exactly in real arithmetic. The binary lattice weights
// sum to 1 + 7.5e-9, so summed term by term they would move C~ by that much
// times C~ (-1 in the gas) every collision: a steady drift of the gas's phase
// mass. heq_0 - w_0 is therefore C~ minus the other eight, which leaves only
// rounding.
__device__ __forceinline__ void equilibrium(float ct, float U, float V, float ni, float nj, const PhysF &ph,
                                            float gmw[9], float heqd[9])
{
#pragma clang fp contract(on)
    const float C = ct + 1.0f;
    const float tmp = -ct * ph.k12W;
    const float U2 = U * U + V * V;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const float eU = (float)c_ex[i] * U + (float)c_ey[i] * V;
        const float eN = (float)c_ex[i] * ni + (float)c_ey[i] * nj;
        gmw[i] = waf(i) * (eU * (3.0f + 4.5f * eU) - 1.5f * U2);
        heqd[i] = ct * (waf(i) + gmw[i]) + gmw[i] + C * tmp * waf(i) * eN;
    }
    float rest = 0.0f;
#pragma unroll
    for (int i = 1; i < 9; i++) rest += heqd[i];
    heqd[0] = ct - rest;
}

__device__ __forceinline__ void forcing(const float gmw[9], float drho3, float mu, const float e[9], float Ux,
                                        float Uy, float gx, float gy, float Fb, float hlp[9])
{
#pragma clang fp contract(on)
    const float ug = Ux * gx + Uy * gy;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const float G = waf(i) + gmw[i];
        hlp[i] = (gmw[i] * drho3 + G * mu) * (e[i] - ug) + ((float)c_ey[i] - Uy) * G * Fb;
    }
}

__device__ __forceinline__ void mrt(float s9, const float geq[9], float g[9])
{
#pragma clang fp contract(on)
    float f[9], o[9], t[9];
#pragma unroll
    for (int i = 0; i < 9; i++) f[i] = g[i] - geq[i];
    o[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    o[1] = -4.0f * f[0] - f[1] - f[2] - f[3] - f[4] + 2.0f * (f[5] + f[6] + f[7] + f[8]);
    o[2] = 4.0f * f[0] - 2.0f * (f[1] + f[2] + f[3] + f[4]) + f[5] + f[6] + f[7] + f[8];
    o[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    o[4] = -2.0f * (f[1] - f[3]) + f[5] - f[6] - f[7] + f[8];
    o[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    o[6] = -2.0f * (f[2] - f[4]) + f[5] + f[6] - f[7] - f[8];
    o[7] = f[1] - f[2] + f[3] - f[4];
    o[8] = f[5] - f[6] + f[7] - f[8];
#pragma unroll
    for (int i = 0; i < 7; i++) t[i] = o[i];
    t[7] = o[7] * s9;
    t[8] = o[8] * s9;
    o[0] = 4.0f * (t[0] - t[1] + t[2]);
    o[1] = 4.0f * t[0] - t[1] - 2.0f * t[2] + 6.0f * (t[3] - t[4]) + 9.0f * t[7];
    o[2] = 4.0f * t[0] - t[1] - 2.0f * t[2] + 6.0f * (t[5] - t[6]) - 9.0f * t[7];
    o[3] = 4.0f * t[0] - t[1] - 2.0f * t[2] - 6.0f * (t[3] - t[4]) + 9.0f * t[7];
    o[4] = 4.0f * t[0] - t[1] - 2.0f * t[2] - 6.0f * (t[5] - t[6]) - 9.0f * t[7];
    o[5] = 4.0f * t[0] + 2.0f * t[1] + t[2] + 6.0f * t[3] + 3.0f * t[4] + 6.0f * t[5] + 3.0f * t[6] + 9.0f * t[8];
    o[6] = 4.0f * t[0] + 2.0f * t[1] + t[2] - 6.0f * t[3] - 3.0f * t[4] + 6.0f * t[5] + 3.0f * t[6] - 9.0f * t[8];
    o[7] = 4.0f * t[0] + 2.0f * t[1] + t[2] - 6.0f * t[3] - 3.0f * t[4] - 6.0f * t[5] - 3.0f * t[6] + 9.0f * t[8];
    o[8] = 4.0f * t[0] + 2.0f * t[1] + t[2] + 6.0f * t[3] + 3.0f * t[4] - 6.0f * t[5] - 3.0f * t[6] - 9.0f * t[8];
#pragma unroll
    for (int i = 0; i < 9; i++) g[i] = g[i] - o[i] * (1.0f / 36.0f);
}

// Collision_h_g_AMR (LBM_Subroutines.f90:116-189), deviation form
__device__ __forceinline__ void collide_cell(const float *sC, int i, int j, int dx, float P, float Ux, float Uy,
                                             const PhysF &ph, float h[9], float g[9])
{
#pragma clang fp contract(on)
    const float inv_dx = inv_pow2f(dx, 1);
    float e[9], hlp[9], gmw[9], heqd[9], geq[9], gx, gy;
    const float mu = chem_pot(sC, i, j, inv_dx * inv_dx, ph);
    const float ct = sC[cidx(i, j)];
    ed_c(sC, i, j, inv_dx, e);
    gradient_ed(e, gx, gy);
    const float rn = rsqrtf(gx * gx + gy * gy + 1.0e-32f);   // normal_AMR (:25-37)
    equilibrium(ct, Ux, Uy, gx * rn, gy * rn, ph, gmw, heqd);
#pragma unroll
    for (int k = 0; k < 9; k++) h[k] = h[k] * (1.0f - ph.wc) + heqd[k] * ph.wc;
    const float Rho = ph.rhoh + ct * ph.drho;
    const float Fb = -ct * ph.drho * ph.gy;
    forcing(gmw, ph.drho3, mu, e, Ux, Uy, gx, gy, Fb, hlp);
    const float rho3 = Rho * (1.0f / 3.0f);
#pragma unroll
    for (int k = 0; k < 9; k++) geq[k] = P * waf(k) + gmw[k] * rho3 - 0.5f * hlp[k];
    ed_m(sC, i, j, inv_dx, e);
    gradient_ed(e, gx, gy);
    forcing(gmw, ph.drho3, mu, e, Ux, Uy, gx, gy, Fb, hlp);
    float tau;
    if (ct < -1.0f)
        tau = ph.taul + 0.5f;
    else if (ct > 0.0f)
        tau = ph.tauh + 0.5f;
    else
        tau = 1.0f / (ct * (ph.inv_tauh - ph.inv_taul) + ph.inv_tauh) + 0.5f;
    mrt(1.0f / tau, geq, g);
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = g[k] + hlp[k];
}

// Macroscopic_Values_g_AMR (LBM_Subroutines.f90:205-240), deviation form
__device__ __forceinline__ void macro_g_cell(const float *sC, int i, int j, int dx, const float g[9], const PhysF &ph,
                                             float &P, float &Ux, float &Uy)
{
#pragma clang fp contract(on)
    const float inv_dx = inv_pow2f(dx, 1);
    const float gx = ((sC[cidx(i + 1, j)] - sC[cidx(i - 1, j)]) * (1.0f / 3.0f) +
                      (sC[cidx(i + 1, j + 1)] + sC[cidx(i + 1, j - 1)] - sC[cidx(i - 1, j + 1)] - sC[cidx(i - 1, j - 1)]) *
                          (1.0f / 12.0f)) *
                     inv_dx;
    const float gy = ((sC[cidx(i, j + 1)] - sC[cidx(i, j - 1)]) * (1.0f / 3.0f) +
                      (sC[cidx(i + 1, j + 1)] + sC[cidx(i - 1, j + 1)] - sC[cidx(i + 1, j - 1)] - sC[cidx(i - 1, j - 1)]) *
                          (1.0f / 12.0f)) *
                     inv_dx;
    const float mu = chem_pot(sC, i, j, inv_dx * inv_dx, ph);
    const float ct = sC[cidx(i, j)];
    const float Rho = ph.rhoh + ct * ph.drho;
    const float Fb = -ct * ph.drho * ph.gy;
    const float Fx = mu * gx;
    const float Fy = mu * gy + Fb;
    float sx = 0.0f, sy = 0.0f, s = 0.0f;
#pragma unroll
    for (int k = 1; k < 9; k++) sx += g[k] * (float)c_ex[k];
#pragma unroll
    for (int k = 1; k < 9; k++) sy += g[k] * (float)c_ey[k];
#pragma unroll
    for (int k = 0; k < 9; k++) s += g[k];
    const float ir = 1.0f / Rho;
    Ux = (3.0f * sx + 0.5f * Fx) * ir;
    Uy = (3.0f * sy + 0.5f * Fy) * ir;
    P = s + ph.drho6 * (Ux * gx + Uy * gy);
}

// the physics arguments of either precision, so one kernel template serves both
template <class T> struct PhysOf;
template <> struct PhysOf<double> {
    static __device__ __forceinline__ const Phys &get(const Phys &p, const PhysF &) { return p; }
    static __device__ __forceinline__ const Phys *table(const DevMeshT<double> &m) { return m.phs; }
};
template <> struct PhysOf<float> {
    static __device__ __forceinline__ const PhysF &get(const Phys &, const PhysF &p) { return p; }
    static __device__ __forceinline__ const PhysF *table(const DevMeshT<float> &m) { return m.phsf; }
};

// storage shift of the FP32 path: C is stored as C - 1, h_q as h_q - w_q
template <class T> __device__ __forceinline__ double c_shift() { return std::is_same<T, float>::value ? 1.0 : 0.0; }

// ---------------------------------------------------------------------------
// Ghost fills (Fill_Ghost_Cells_{C,C2,hg}.f90, Interpolations.f90:2-138,
// BC_Cell.f90) by ghost program: every value a block's fill needs is built
// from interior data of the previous kernel, the coarse neighbours' ghost
// cells included, so one launch covers every level.
// ---------------------------------------------------------------------------
// byte offset of the recipes in the hg kernel's LDS, after 9 population
// arrays of lcf cells of esz bytes (16-byte aligned)
__host__ __device__ constexpr size_t hg_nodes_offset(int lcf, size_t esz) { return (9 * (size_t)lcf * esz + 15) & ~size_t(15); }

// Ghost cell t of the ring-1 population block: sides L,R,B,T then the corners.
__device__ __forceinline__ int ring1_cell(int t)
{
    int x, y;
    if (t < 64) {
        int d = t / 16, pos = t % 16 + 1;
        x = d == 0 ? 0 : (d == 1 ? 17 : pos);
        y = d == 2 ? 0 : (d == 3 ? 17 : pos);
    } else {
        int c = t - 64;
        x = c < 2 ? 0 : 17;
        y = (c == 0 || c == 2) ? 0 : 17;
    }
    return y * FW + x;
}

// axis weights of Biquadratic_from_Fine / _from_Coarse (Interpolations.f90:13-31, :85-101)
template <class T>
__device__ __forceinline__ void wpat(int p, T w[3])
{
    w[0] = p == 0 ? T(3) : (p == 1 ? T(-1) : (p == 2 ? T(5) : T(-3)));
    w[1] = p < 2 ? T(6) : T(30);
    w[2] = p == 0 ? T(-1) : (p == 1 ? T(3) : (p == 2 ? T(-3) : T(5)));
}

// end of stage s of a ghost program. A select over the constant indices
// keeps the header in registers; indexing end[] with a variable would place
// the header in per-lane scratch memory, written and read back through L2.
__device__ __forceinline__ int stage_end(const PHdr &h, int s)
{
    int e = 0;
#pragma unroll
    for (int t = 0; t < MAX_STAGES; t++)
        if (s == t) e = h.end[t];
    return e;
}

// node load as three 16-byte loads
__device__ __forceinline__ PNode load_node(const PNode *p)
{
    const int4 *q = reinterpret_cast<const int4 *>(p);
    int4 a = q[0], b = q[1], c = q[2];
    PNode n;
    n.dest = a.x; n.info = a.y; n.ref[0] = a.z; n.ref[1] = a.w;
    n.ref[2] = b.x; n.ref[3] = b.y; n.ref[4] = b.z; n.ref[5] = b.w;
    n.ref[6] = c.x; n.ref[7] = c.y; n.ref[8] = c.z; n.own = c.w;
    return n;
}

// A load of a value another workgroup wrote in the same launch (COH: the
// persistent segment kernel), as an agent-scope atomic load: it reads the
// device-coherent value past a stale copy in this XCD's L2 without an L2
// invalidation. Otherwise (values of an earlier kernel) a plain load.
template <bool COH, class T>
__device__ __forceinline__ T ldg(const T *p)
{
    if (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
// a store other workgroups read in the same launch: agent-scope atomic,
// written through this XCD's L2, so that no L2 write-back is needed
template <class T>
__device__ __forceinline__ void stg_coh(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the recipe's operands that live in global memory, fetched before any stage
// barrier (they are interior values left by the previous kernel)
template <bool COH = false, class T>
__device__ __forceinline__ void prefetch_node(const PNode &nd, const T *__restrict__ G, int gs, T gv[9])
{
    const int type = nd.info & 15;
    const int nref = type == NT_COPY ? 1 : (type >= NT_BB_BOT ? 0 : 9);
#pragma unroll
    for (int r = 0; r < 9; r++) gv[r] = (r < nref && nd.ref[r] >= 0) ? ldg<COH>(G + nd.ref[r] + gs) : T(0);
}

// one recipe from prefetched global operands gv and LDS L (+ shift ls);
// GLOBAL_ONLY: a first-stage recipe, whose operands are all in gv
template <bool GLOBAL_ONLY = false, class T>
__device__ __forceinline__ T eval_pref(const PNode &nd, const T gv[9], const T *L, int ls)
{
    const int type = nd.info & 15;
#define PV(r) (GLOBAL_ONLY || nd.ref[r] >= 0 ? gv[r] : L[-1 - nd.ref[r] + ls])
    if (type == NT_COPY) return PV(0);
    T wx[3], wy[3];
    wpat((nd.info >> 4) & 3, wx);
    wpat((nd.info >> 6) & 3, wy);
    T t[3];
#pragma unroll
    for (int a = 0; a < 3; a++) t[a] = wy[0] * PV(3 * a) + wy[1] * PV(3 * a + 1) + wy[2] * PV(3 * a + 2);
    T o = wx[0] * t[0] + wx[1] * t[1] + wx[2] * t[2];
#undef PV
    return type == NT_FINE ? o / T(64) : o / T(1024);
}

// one recipe: copy, or the tensor product (pass along y, then x) divided by
// 64 (from finer) or 1024 (from coarser)
template <bool COH = false, class T>
__device__ __forceinline__ T eval_node(const PNode &nd, const T *__restrict__ G, int gs, const T *L, int ls)
{
    const int type = nd.info & 15;
#define PV(r) ((r) >= 0 ? ldg<COH>(G + (r) + gs) : L[-1 - (r) + ls])
    if (type == NT_COPY) return PV(nd.ref[0]);
    T wx[3], wy[3];
    wpat((nd.info >> 4) & 3, wx);
    wpat((nd.info >> 6) & 3, wy);
    T t[3];
#pragma unroll
    for (int a = 0; a < 3; a++) t[a] = wy[0] * PV(nd.ref[3 * a]) + wy[1] * PV(nd.ref[3 * a + 1]) + wy[2] * PV(nd.ref[3 * a + 2]);
    T o = wx[0] * t[0] + wx[1] * t[1] + wx[2] * t[2];
#undef PV
    return type == NT_FINE ? o / T(64) : o / T(1024);
}

// a first-stage recipe: every operand is an interior value in memory
template <bool COH = false, class T>
__device__ __forceinline__ T eval_global(const PNode &nd, const T *__restrict__ G, int gs)
{
    const int type = nd.info & 15;
    if (type == NT_COPY) return ldg<COH>(G + nd.ref[0] + gs);
    T wx[3], wy[3];
    wpat((nd.info >> 4) & 3, wx);
    wpat((nd.info >> 6) & 3, wy);
    T t[3];
#pragma unroll
    for (int a = 0; a < 3; a++)
        t[a] = wy[0] * ldg<COH>(G + nd.ref[3 * a] + gs) + wy[1] * ldg<COH>(G + nd.ref[3 * a + 1] + gs) +
               wy[2] * ldg<COH>(G + nd.ref[3 * a + 2] + gs);
    T o = wx[0] * t[0] + wx[1] * t[1] + wx[2] * t[2];
    return type == NT_FINE ? o / T(64) : o / T(1024);
}

// population q of set `set` for one recipe; sF holds that set's 9 arrays,
// lcf cells apart.
// Bounce-back rows take the opposite population of the first row
// (Bounce_Back_3/4, BC_Cell.f90:2-42).
template <bool GLOBAL_ONLY, bool COH = false, class T>
__device__ __forceinline__ T eval_hg(const PNode &nd, int set, int q, const T *__restrict__ Fc, const T *sF, int lcf)
{
    const int type = nd.info & 15;
    const int c = set * 9 + q;
    if (type < NT_BB_BOT)
        return GLOBAL_ONLY ? eval_global<COH>(nd, Fc, c * GS) : eval_node<COH>(nd, Fc, c * GS, sF, q * lcf);
    const int e = c_ex[q] * c_ey[q];
    const int r = nd.ref[1 + (type == NT_BB_TOP ? -e : e)];
    const int qo = c_opp[q];
    return (GLOBAL_ONLY || r >= 0) ? ldg<COH>(Fc + r + (set * 9 + qo) * GS) : sF[-1 - r + qo * lcf];
}

// t / ns for 0 <= t < 18 ns and 0 < ns <= NODES_HG, as a multiply and a
// shift: m = 2^22 / ns + 1, taken once per stage. Exact over that range: m /
// 2^22 exceeds 1 / ns by less than 2^-22, so t m / 2^22 exceeds t / ns by less
// than 18 NODES_HG / 2^22 < 1 / NODES_HG <= 1 / ns, which keeps it below the
// next multiple of 1 / ns; the products stay below 2^31.
static_assert(18 * NODES_HG * NODES_HG < (1 << 22), "the quotient shift needs a wider margin");
__device__ __forceinline__ unsigned quot_mul(int ns) { return (1u << 22) / (unsigned)ns + 1u; }
__device__ __forceinline__ int quot(int t, unsigned m) { return (int)(((unsigned)t * m) >> 22); }

// One stage of an hg ghost program for one population set: recipes [b, e)
// of the block (staged in sN) applied to the set's 9 populations. Items run
// recipe-fastest, so the lanes of a wavefront read neighbouring cells of one
// population array; U items per thread have their operands in flight together.
template <bool GLOBAL_ONLY, int U, class T>
__device__ __forceinline__ void hg_stage(const PNode *sN, int b, int e, int set, const T *__restrict__ Fc, T *sF,
                                         int lcf, T *F, bool store, int tid)
{
    const int ns = e - b;
    if (ns <= 0) return;
    const unsigned mq = quot_mul(ns);
    for (int base = 0; base < 9 * ns; base += 256 * U) {
        T v[U];
#pragma unroll
        for (int r = 0; r < U; r++) {
            const int t = base + tid + 256 * r;
            if (t < 9 * ns) {
                const int q = quot(t, mq);
                v[r] = eval_hg<GLOBAL_ONLY>(sN[b + t - q * ns], set, q, Fc, sF, lcf);
            }
        }
#pragma unroll
        for (int r = 0; r < U; r++) {
            const int t = base + tid + 256 * r;
            if (t < 9 * ns) {
                const int q = quot(t, mq);
                const PNode &nd = sN[b + t - q * ns];
                sF[q * lcf + nd.dest] = v[r];
                if (store && nd.own >= 0) F[(set * 9 + q) * GS + nd.own] = v[r];
            }
        }
    }
}

// Fill_Ghost_Cells_C (kind C1) / Fill_Ghost_Cells_C2 (kind C2) of every leaf
// in list by its ghost program, then the op on the block: the collision
// (after C2), Macroscopic_Values_g_AMR (after C), or both (OP_MACRO_G_COLLIDE:
// U, P of the previous step then this step's collision; between two steps
// with no regrid, the C fill that ends a step and the C2 fill that starts the
// next compute the same first ghost ring from the same C, so U, P come out
// identical). The collision and U,P read only this block's C. With
// prog.hdr == null the ghosts are read from memory (no fill). ENS: an
// ensemble mesh, whose fluid constants are the block's instance's (m.phs,
// m.phsf); else the kernel arguments'.
template <class T, int OP, bool ENS>
__global__ __launch_bounds__(256) void k_prog_c(DevMeshT<T> m, int cur, const int *__restrict__ list, Prog prog,
                                                Phys ph64, PhysF ph32, double *err_slot, StepCtl ctl)
{
    const auto &ph_arg = PhysOf<T>::get(ph64, ph32);
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    // a judged state needing refinement before this launch's: the segment stops there
    const int skip = *m.skip || (ctl.judge && *ctl.judge < ctl.ord);
    const int slot = list[k];
    const bool fill = prog.hdr != nullptr;
    PHdr hd;
    hd.off = 0;
    hd.nst = 0;
    if (fill) hd = prog.hdr[k];
    if (skip) return;
    const BlockDev me = m.blk[slot];
    const auto &ph = ENS ? PhysOf<T>::table(m)[me.inst] : ph_arg;
    T *C = m.C + (size_t)slot * CS;
    __shared__ T sC[LC_C];
    const int i = tid % NX + 1, j = tid / NX + 1;
    T *F = m.pop[cur] + (size_t)slot * POPB;
    T *Fo = ctl.out >= 0 ? m.pop[ctl.out] + (size_t)slot * POPB : F;   // the collision's output
    T *PU = m.PU + (size_t)slot * PUB;
    T h[9], g[9], P = T(0), Ux = T(0), Uy = T(0);
    if (OP == OP_COLLIDE || OP == OP_MACRO_G_COLLIDE) {
#pragma unroll
        for (int q = 0; q < 9; q++) h[q] = F[fidx(0, q, i, j)];
    }
    if (OP == OP_COLLIDE) {
        P = PU[pidx(i, j)];
        Ux = PU[PS + pidx(i, j)];
        Uy = PU[2 * PS + pidx(i, j)];
    }
    if (OP != OP_NONE) {
#pragma unroll
        for (int q = 0; q < 9; q++) g[q] = F[fidx(1, q, i, j)];
    }
    // this thread's first recipe; first-stage recipes read only interior
    // values in memory, so they are done before the first barrier
    const int total = hd.nst > 0 ? stage_end(hd, hd.nst - 1) : 0;
    const int n1 = hd.nst > 0 ? hd.end[0] : 0;
    PNode nd0;
    T gv[9];
    if (tid < total) {
        nd0 = load_node(prog.nodes + hd.off + tid);
        prefetch_node(nd0, (const T *)m.C, 0, gv);
    }
    if (fill) {
        const T cin = C[cidx(i, j)];
        if (tid < n1) {
            const T v = eval_pref<true>(nd0, gv, (const T *)sC, 0);
            sC[nd0.dest] = v;
            if (nd0.dest < CS) C[nd0.dest] = v;
        }
        sC[cidx(i, j)] = cin;
    } else {
        for (int t = tid; t < CS; t += 256) sC[t] = C[t];
    }
    __syncthreads();
    if (fill) {
        int begin = 0;
        for (int s = 0; s < hd.nst; s++) {
            const int end = stage_end(hd, s);
            if (s > 0 && tid >= begin && tid < end) {
                const T v = eval_pref(nd0, gv, (const T *)sC, 0);
                sC[nd0.dest] = v;
                if (nd0.dest < CS) C[nd0.dest] = v;
            }
            for (int t = tid + 256; t < total; t += 256) {
                if (t < begin || t >= end) continue;
                const PNode nd = load_node(prog.nodes + hd.off + t);
                const T v = eval_node(nd, (const T *)m.C, 0, (const T *)sC, 0);
                sC[nd.dest] = v;
                if (nd.dest < CS) C[nd.dest] = v;
            }
            begin = end;
            __syncthreads();
        }
    }
    if (err_slot || ctl.judge_on) {
        // Err_Interface (Refine_Derefine.f90:327-352) from the filled ring
        const T a = sC[cidx(i + 1, j)] - sC[cidx(i - 1, j)];
        const T b = sC[cidx(i, j + 1)] - sC[cidx(i, j - 1)];
        T dc = a * a + b * b;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const T x = __shfl_xor(dc, o);
            if (x > dc) dc = x;
        }
        if (!err_slot) {
            // the judge alone: sqrt (correctly rounded), the division by 2 dx (a
            // power of two) and the widening are monotone, so the block's
            // Err_Interface passes the tolerance iff one wavefront's maximum
            // does, and each wavefront tests its own with no barrier
            if ((tid & 63) == 0) {
                const double err = (double)(sqrt(dc) / (T(2) * (T)me.dx));
                const int lo = ENS ? m.lv[2 * me.inst] : ctl.min_level, hi = ENS ? m.lv[2 * me.inst + 1] : ctl.max_level;
                if (me.level < lo || (me.level < hi && err > ctl.refin_tol)) atomicMin(ctl.judge, ctl.ord);
            }
        } else {
            __shared__ T wmax[4];
            if ((tid & 63) == 0) wmax[tid >> 6] = dc;
            __syncthreads();
            if (tid == 0) {
                T e = wmax[0];
                for (int w = 1; w < 4; w++)
                    if (wmax[w] > e) e = wmax[w];
                const double err = (double)(sqrt(e) / (T(2) * (T)me.dx));
                err_slot[slot] = err;
                if (ctl.judge_on) {
                    // the judge: Flag_for_refinement's criterion (Refine_Derefine.f90:18-54) on this leaf
                    const int lo = ENS ? m.lv[2 * me.inst] : ctl.min_level,
                              hi = ENS ? m.lv[2 * me.inst + 1] : ctl.max_level;
                    if (me.level < lo || (me.level < hi && err > ctl.refin_tol)) atomicMin(ctl.judge, ctl.ord);
                }
            }
        }
    }
    if (OP == OP_MACRO_G_COLLIDE || OP == OP_MACRO_G) {
        macro_g_cell(sC, i, j, me.dx, g, ph, P, Ux, Uy);
        PU[pidx(i, j)] = P;
        PU[PS + pidx(i, j)] = Ux;
        PU[2 * PS + pidx(i, j)] = Uy;
    }
    if (OP == OP_COLLIDE || OP == OP_MACRO_G_COLLIDE) {
        collide_cell(sC, i, j, me.dx, P, Ux, Uy, ph, h, g);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            Fo[fidx(0, q, i, j)] = h[q];
            Fo[fidx(1, q, i, j)] = g[q];
        }
    }
}

// LDS layout of a population block: component c, cell (x,y) at c*lcf + y*FW + x
// (lcf = FS + the aux cells of the launch's largest ghost program)
__device__ __forceinline__ int lidx(int c, int x, int y, int lcf) { return c * lcf + y * FW + x; }

// Perfect_Shift_AMR / Streaming_LaxWendroff_AMR (LBM_Subroutines.f90:41-90)
// of one cell from the LDS copy of the block into G, and C = sum(h) (:193-201)
template <class T>
__device__ __forceinline__ T stream_cell(const T *sF, int lcf, T *__restrict__ G, int set, int i, int j, int dx)
{
    const T cfl = inv_dx_of<T>(dx);   // 1 / dx, exact
    const T a = cfl / T(2) * (T(1) + cfl), b = T(1) - cfl * cfl, c = cfl / T(2) * (T(1) - cfl);
    T s = T(0);
#pragma unroll
    for (int q = 0; q < 9; q++) {
        T v;
        if (q == 0)
            v = sF[lidx(q, i, j, lcf)];
        else if (cfl > T(0.9))
            v = sF[lidx(q, i - c_ex[q], j - c_ey[q], lcf)];
        else
            v = a * sF[lidx(q, i - c_ex[q], j - c_ey[q], lcf)] + b * sF[lidx(q, i, j, lcf)] -
                c * sF[lidx(q, i + c_ex[q], j + c_ey[q], lcf)];
        G[fidx(set, q, i, j)] = v;
        s += v;
    }
    return s;
}

// Fill_Ghost_Cells_hg of every leaf in list by its ghost program (bounce-back
// rows included), then optionally the stream into the other population buffer
// and C = sum(h). The two population sets never read each other here, so a
// block is two workgroups, one per set (0: h, 1: g). Workgroup w takes list
// entry 8 (w / 16) + w % 8 and set (w / 8) % 2, so both of entry k's
// workgroups run on XCD k % 8, where every other launch runs entry k. Each
// recipe covers a ghost cell of the block or a coarse neighbour's ghost cell
// it needs.
// flags: 1 fill (else ghosts from memory), 2 stream, 4 C = sum h,
// 8 carry the ghost ring into the new buffer (as the in-place reference stream),
// 16 store the filled ghost ring in memory (the phase-by-phase fill; a fill
// followed by the stream in the same kernel keeps it in LDS only: no kernel
// reads a population ghost cell from memory, and the next fill rewrites it).
// U0 / U1: recipes per thread whose operands are in flight together in the
// first stage / later stages (a launch variant; the results do not depend on it).
template <class T, int U0, int U1>
__global__ __launch_bounds__(256) void k_prog_hg(DevMeshT<T> m, int cur, const int *__restrict__ list, int n,
                                                 Prog prog, int flags, const int *judge, int ord)
{
    const int w = blockIdx.x;
    const int k = 8 * (w >> 4) + (w & 7), set = (w >> 3) & 1;
    if (k >= n) return;
    const int tid = threadIdx.x;
    // the stream of step ord - 1 runs unless the judge stopped the segment at or before its state
    const int skip = *m.skip || (judge && *judge < ord);
    const int slot = list[k];
    PHdr hd;
    hd.off = 0;
    hd.nst = 0;
    if (flags & 1) hd = prog.hdr[k];
    if (skip) return;
    const BlockDev me = m.blk[slot];
    const T *__restrict__ Fc = m.pop[cur];
    T *F = m.pop[cur] + (size_t)slot * POPB;
    // LDS sized at launch (hg_lds_bytes): the 9 populations of the set, FS
    // cells plus the program's aux cells each, then the block's recipes
    extern __shared__ __align__(16) unsigned char hg_lds[];
    const int lcf = FS + prog.naux;
    T *sF = reinterpret_cast<T *>(hg_lds);
    PNode *sN = reinterpret_cast<PNode *>(hg_lds + hg_nodes_offset(lcf, sizeof(T)));
    const int i = tid % NX + 1, j = tid / NX + 1;
    const int nn = hd.nst > 0 ? stage_end(hd, hd.nst - 1) : 0;
    {   // the block's recipes, staged in LDS
        const int4 *src = reinterpret_cast<const int4 *>(prog.nodes + hd.off);
        int4 *dst = reinterpret_cast<int4 *>(sN);
        for (int t = tid; t < 3 * nn; t += 256) dst[t] = src[t];
    }
#pragma unroll
    for (int q = 0; q < 9; q++) sF[lidx(q, i, j, lcf)] = F[fidx(set, q, i, j)];
    if (!(flags & 1)) {
        for (int t = tid; t < 68 * 9; t += 256) {   // the ring is stored in ring1_cell order
            int q = t / 68, r = t - q * 68;
            sF[q * lcf + ring1_cell(r)] = F[(set * 9 + q) * GS + NX * NY + r];
        }
    }
    __syncthreads();
    // the first stage reads interior values in memory only; later stages
    // also read cells of earlier stages
    int begin = 0;
    for (int s = 0; s < hd.nst; s++) {
        const int end = stage_end(hd, s);
        if (s == 0)
            hg_stage<true, U0>(sN, begin, end, set, Fc, sF, lcf, F, (flags & 16) != 0, tid);
        else
            hg_stage<false, U1>(sN, begin, end, set, Fc, sF, lcf, F, (flags & 16) != 0, tid);
        begin = end;
        __syncthreads();
    }
    if (!(flags & 2)) return;
    T *G = m.pop[cur ^ 1] + (size_t)slot * POPB;
    const T csum = stream_cell((const T *)sF, lcf, G, set, i, j, me.dx);
    if ((flags & 4) && set == 0) m.C[(size_t)slot * CS + cidx(i, j)] = csum;
    if (flags & 8) {
        for (int t = tid; t < 68 * 9; t += 256) {
            int q = t / 68, r = t - q * 68;
            G[(set * 9 + q) * GS + NX * NY + r] = sF[q * lcf + ring1_cell(r)];
        }
    }
}

// Macroscopic_Values_h_AMR (LBM_Subroutines.f90:193-201)
template <class T>
__global__ __launch_bounds__(256) void k_macro_h(DevMeshT<T> m, int cur, const int *__restrict__ list)
{
    const int slot = list[blockIdx.x];
    const int i = threadIdx.x % NX + 1, j = threadIdx.x / NX + 1;
    const T *F = m.pop[cur] + (size_t)slot * POPB;
    T s = T(0);
#pragma unroll
    for (int q = 0; q < 9; q++) s += F[fidx(0, q, i, j)];
    m.C[(size_t)slot * CS + cidx(i, j)] = s;
}

// Err_Interface (Refine_Derefine.f90:327-352): sqrt(max |C(i+1)-C(i-1)|^2 ...)/(2dx)
template <class T>
__global__ __launch_bounds__(256) void k_err(DevMeshT<T> m, const int *__restrict__ list, double *err, double *err_slot)
{
    const int slot = list[blockIdx.x];
    const int i = threadIdx.x % NX + 1, j = threadIdx.x / NX + 1;
    const T *C = m.C + (size_t)slot * CS;
    T a = C[cidx(i + 1, j)] - C[cidx(i - 1, j)];
    T b = C[cidx(i, j + 1)] - C[cidx(i, j - 1)];
    T dc = a * a + b * b;
    __shared__ T red[256];
    red[threadIdx.x] = dc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            T o = red[threadIdx.x + s];
            if (o > red[threadIdx.x]) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double e = (double)(sqrt(red[0]) / (T(2) * (T)m.blk[slot].dx));
        err[blockIdx.x] = e;
        if (err_slot) err_slot[slot] = e;
    }
}

// Whether Refine_Derefine_Mesh (Refine_Derefine.f90:2-14) would change the
// mesh, and the flags it would set, decided on the device from Err_Interface
// of every leaf:
//  - Flag_for_refinement (:18-54) and Enforce_1_Level_Jump (:58-93): the
//    criterion's flags, then the least fixed point under "a leaf one level
//    coarser next to a flagged leaf is flagged", reached breadth first from
//    the criterion's flags (any order reaches the same set);
//  - Flag_for_derefinement (:97-145): a parent of leaves none of which is
//    flagged for refinement and all of whose children pass the tolerance;
//  - its level-by-level check (:149-194): a candidate at level L is cancelled
//    iff one of the two children next to it of a non-leaf neighbour (pairs
//    :155-163) is flagged for refinement, or is a non-leaf without a
//    surviving derefinement flag; those flags are final once the pass over
//    level L+1 is done, so each candidate gathers its watch blocks once and
//    the passes only test LDS bits.
// One workgroup. changed[0] = 1 iff the mesh changes, changed[1] = 1 iff any
// refinement flag is set. With flags (kRegridFlagWords ints): flags[0] = 1
// when the block slots of the flags[1] refinement flags follow from flags[3]
// and those of the flags[2] derefinement flags from flags[kRegridDerefAt];
// 0 when there were more than the lists hold (RC_REF flagged leaves, RC_CAND
// candidates; the host then flags itself).
// With judge (the stop ordinal of a judged segment): changed[2] = *judge, and
// a stop before jstop (a state inside the segment needed refinement, and the
// segment's last state was never reached) answers "changed, refinement" with
// no flags (they belong to that earlier state).
// Ordered stream compaction in a 1024-thread workgroup: the threads whose
// pred holds append value to list after *count, in thread order (wavefront
// ballots, a prefix sum over the 16 waves' counts, each lane's rank among the
// flagged lanes of its wave); entries past cap are counted, not stored.
// Every thread of the workgroup calls it (it holds barriers).
__device__ __forceinline__ void append_ordered(bool pred, int value, int *list, int cap, int *count, int *wave_cnt)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(pred);
    if (lane == 0) wave_cnt[wave] = __popcll(bal);
    __syncthreads();
    int ix = *count;
    for (int w = 0; w < wave; w++) ix += wave_cnt[w];
    ix += __popcll(bal & ((1ull << lane) - 1ull));
    if (pred && ix < cap) list[ix] = value;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = *count;
        for (int w = 0; w < 16; w++) t += wave_cnt[w];
        *count = t;
    }
    __syncthreads();
}

// With INFO, the block table and each leaf's Err_Interface are read from
// memory once, into a packed word per block slot in LDS (bits 0-4: level + 1,
// 0 for a free slot; 5: leaf; 6: parent of leaves; 7: level at or above the
// block's minimum; 8: a leaf's Err_Interface above deref_tol; 9-31: first
// child + 1), so that the later passes, which follow children and candidates,
// wait on LDS rather than on memory; without it (more slots than LDS holds)
// they read the table.
constexpr int kInfoLevel = 0x1f, kInfoLeaf = 1 << 5, kInfoHasLeaf = 1 << 6, kInfoLoOk = 1 << 7, kInfoErrD = 1 << 8,
              kInfoChildShift = 9;

template <bool INFO>
__global__ __launch_bounds__(1024) void k_regrid_check(DevMeshT<double> m, int nblks, const double *__restrict__ err_slot,
                                                      int min_levels, int max_levels, int levelmin, int levelmax,
                                                      double refin_tol, double deref_tol, int *changed,
                                                      const int *__restrict__ judge, int jstop, int *flags)
{
    extern __shared__ unsigned masks[];
    __shared__ int cand[RC_CAND], clev[RC_CAND], watch[RC_CAND * 16], rlist[RC_REF];
    __shared__ int ncand, nref, nder, wave_cnt[16];
    const int tid = threadIdx.x, nw = (nblks + 31) >> 5;
    unsigned *dmask = masks, *rmask = masks + nw, *info = masks + 2 * nw;
    auto rbit = [&](int c) { return (rmask[c >> 5] >> (c & 31)) & 1u; };
    // what the passes after the first read of a block slot
    auto level_of = [&](int n) { return INFO ? (int)(info[n] & kInfoLevel) - 1 : m.blk[n].level; };
    auto leaf_of = [&](int n) { return INFO ? (info[n] & kInfoLeaf) != 0 : m.blk[n].isleaf != 0; };
    if (judge) {
        const int J = *judge;
        if (tid == 0) changed[2] = J;
        if (J < jstop) {
            if (tid == 0) {
                changed[0] = changed[1] = 1;
                if (flags) flags[0] = 0;
            }
            return;
        }
    }
    for (int w = tid; w < 2 * nw; w += 1024) masks[w] = 0u;
    if (tid == 0) ncand = nref = nder = 0;
    __syncthreads();
    // the criterion's refinement flags, listed in ascending slot order
    for (int base = 0; base < nblks; base += 1024) {
        const int n = base + tid;
        bool f = false;
        if (n < nblks) {
            const BlockDev b = m.blk[n];
            const int lo = m.lv ? m.lv[2 * b.inst] : min_levels, hi = m.lv ? m.lv[2 * b.inst + 1] : max_levels;
            const bool leaf = b.level >= 0 && b.isleaf;
            const double e = leaf ? err_slot[n] : 0.0;
            if (leaf) f = b.level < lo || (b.level < hi && e > refin_tol);
            if (INFO)
                info[n] = (unsigned)(b.level + 1) | (leaf ? kInfoLeaf : 0) | (b.hasleaf ? kInfoHasLeaf : 0) |
                          (b.level >= lo ? kInfoLoOk : 0) | (leaf && e > deref_tol ? kInfoErrD : 0) |
                          ((unsigned)(b.child0 + 1) << kInfoChildShift);
        }
        if (f) atomicOr(&rmask[n >> 5], 1u << (n & 31));
        append_ordered(f, n, rlist, RC_REF, &nref, wave_cnt);
    }
    // the 2:1 closure, one breadth-first round per pass
    int head = 0;
    bool lists_ok = true;
    while (true) {
        const int tail = nref;
        if (tail > RC_REF) {
            lists_ok = false;
            break;
        }
        for (int t = head + tid; t < tail; t += 1024) {
            const int n = rlist[t];
            const int L = level_of(n);
            for (int k = 0; k < 8; k++) {
                const NbDesc e = m.nbd[(size_t)n * 8 + k];
                if (e.id <= 0 || !e.isleaf || e.level != L - 1) continue;
                const int q = e.id - 1;
                const unsigned bit = 1u << (q & 31);
                if (!(atomicOr(&rmask[q >> 5], bit) & bit)) {
                    const int ix = atomicAdd(&nref, 1);
                    if (ix < RC_REF) rlist[ix] = q;
                }
            }
        }
        __syncthreads();
        const int now = nref;
        __syncthreads();   // every thread has read it before the next round adds to it
        if (now == tail) break;
        head = tail;
    }
    const int nr = nref;
    // derefinement candidates, listed in ascending slot order
    for (int base = 0; base < nblks; base += 1024) {
        const int n = base + tid;
        bool d = false;
        if (n < nblks) {
            if (INFO) {
                const unsigned w = info[n];
                const int c0 = (int)(w >> kInfoChildShift) - 1;
                if ((w & kInfoLoOk) && (w & kInfoHasLeaf) && c0 >= 0) {
                    d = true;
                    for (int c = c0; c < c0 + 4; c++)
                        if (rbit(c) || (info[c] & kInfoErrD)) d = false;
                }
            } else {
                const BlockDev b = m.blk[n];
                const int lo = m.lv ? m.lv[2 * b.inst] : min_levels;
                if (b.level >= lo && b.hasleaf && b.child0 >= 0) {
                    d = true;
                    for (int c = b.child0; c < b.child0 + 4; c++) {
                        const double e = m.blk[c].isleaf ? err_slot[c] : 0.0;
                        if (rbit(c) || e > deref_tol) d = false;
                    }
                }
            }
        }
        if (d) atomicOr(&dmask[n >> 5], 1u << (n & 31));
        append_ordered(d, n, cand, RC_CAND, &ncand, wave_cnt);
    }
    const int nc = ncand;
    for (int ci = tid; ci < nc && ci < RC_CAND; ci += 1024) clev[ci] = level_of(cand[ci]);
    __syncthreads();
    if (nc > RC_CAND) {   // the host decides and flags
        if (tid == 0) {
            changed[0] = 1;
            changed[1] = nr > 0;
            if (flags) flags[0] = 0;
        }
        return;
    }
    // children adjacent to direction k (1-based, Refine_Derefine.f90:155-163):
    // a flagged one cancels the candidate at once, a non-leaf one is watched
    const int pair_c1[8] = {2, 1, 3, 1, 4, 2, 3, 1};
    const int pair_c2[8] = {4, 3, 4, 2, 4, 2, 3, 1};
    for (int t = tid; t < 8 * nc; t += 1024) {
        const int ci = t >> 3, k = t & 7, n = cand[ci];
        int w0 = -1, w1 = -1;
        const NbDesc e = m.nbd[(size_t)n * 8 + k];
        if (e.id > 0 && !e.isleaf && e.child0 >= 0) {
            const int c1 = e.child0 + pair_c1[k] - 1, c2 = e.child0 + pair_c2[k] - 1;
            if (rbit(c1) || rbit(c2)) atomicAnd(&dmask[n >> 5], ~(1u << (n & 31)));
            if (!leaf_of(c1)) w0 = c1;
            if (!leaf_of(c2)) w1 = c2;
        }
        watch[ci * 16 + 2 * k] = w0;
        watch[ci * 16 + 2 * k + 1] = w1;
    }
    __syncthreads();
    for (int L = levelmax - 1; L >= levelmin - 1; L--) {
        for (int ci = tid; ci < nc; ci += 1024) {
            if (clev[ci] != L) continue;
            bool keep = true;
            for (int r = 0; r < 16; r++) {
                const int c = watch[ci * 16 + r];
                if (c >= 0 && !((dmask[c >> 5] >> (c & 31)) & 1u)) keep = false;
            }
            const int n = cand[ci];
            if (!keep) atomicAnd(&dmask[n >> 5], ~(1u << (n & 31)));
        }
        __syncthreads();
    }
    // the surviving derefinement flags (candidates only carry dmask bits), in
    // ascending slot order
    for (int base = 0; base < nc; base += 1024) {
        const int ci = base + tid;
        const int n = ci < nc ? cand[ci] : 0;
        const bool keep = ci < nc && ((dmask[n >> 5] >> (n & 31)) & 1u);
        append_ordered(keep, n, (flags && lists_ok) ? flags + kRegridDerefAt : wave_cnt, (flags && lists_ok) ? RC_CAND : 0,
                       &nder, wave_cnt);
    }
    if (flags && lists_ok)
        for (int t = tid; t < nr; t += 1024) flags[3 + t] = rlist[t];
    __syncthreads();
    if (tid == 0) {
        changed[0] = (nr > 0 || nder > 0) ? 1 : 0;
        changed[1] = nr > 0 ? 1 : 0;
        if (flags) {
            flags[0] = lists_ok ? 1 : 0;
            flags[1] = nr;
            flags[2] = nder;
        }
    }
}

// Restrict_Bilinear (Restrict.f90:14-65) with Bilinear_from_Fine(_f)
// (Interpolations.f90:142-194): parent interior = 2x2 mean of the children,
// parent ghosts zeroed. pairs[2k] = parent slot, pairs[2k+1] = first child slot.
template <class T>
__global__ __launch_bounds__(256) void k_restrict(DevMeshT<T> m, int cur, const int *__restrict__ pairs)
{
    const int ps = pairs[2 * blockIdx.x], c0 = pairs[2 * blockIdx.x + 1];
    const int tid = threadIdx.x;
    T *F = m.pop[cur] + (size_t)ps * POPB;
    T *C = m.C + (size_t)ps * CS;
    T *PU = m.PU + (size_t)ps * PUB;
    for (int t = tid; t < POPB; t += 256) F[t] = T(0);
    for (int t = tid; t < CS; t += 256) C[t] = T(0);
    for (int t = tid; t < PUB; t += 256) PU[t] = T(0);
    __syncthreads();
    const int X = tid % NX + 1, Y = tid / NX + 1;
    const int cs = c0 + (X > NX / 2) + 2 * (Y > NY / 2);
    const int lx = 2 * X - 1 - NX * (X > NX / 2), ly = 2 * Y - 1 - NY * (Y > NY / 2);   // (2X-1, 2Y-1) in the child
    auto mean4 = [](T a, T b, T c, T d) {   // in(i,2Y-1), in(i,2Y), in(i+1,2Y-1), in(i+1,2Y)
        return T(0.5) * (T(0.5) * (a + b) + T(0.5) * (c + d));
    };
    const T *Fch = m.pop[cur] + (size_t)cs * POPB;
    for (int comp = 0; comp < 18; comp++) {
        int set = comp / 9, q = comp % 9;
        F[fidx(set, q, X, Y)] = mean4(Fch[fidx(set, q, lx, ly)], Fch[fidx(set, q, lx, ly + 1)],
                                      Fch[fidx(set, q, lx + 1, ly)], Fch[fidx(set, q, lx + 1, ly + 1)]);
    }
    const T *Cch = m.C + (size_t)cs * CS;
    C[cidx(X, Y)] = mean4(Cch[cidx(lx, ly)], Cch[cidx(lx, ly + 1)], Cch[cidx(lx + 1, ly)], Cch[cidx(lx + 1, ly + 1)]);
    const T *Pch = m.PU + (size_t)cs * PUB;
    for (int k = 0; k < 3; k++)
        PU[k * PS + pidx(X, Y)] = mean4(Pch[k * PS + pidx(lx, ly)], Pch[k * PS + pidx(lx, ly + 1)],
                                        Pch[k * PS + pidx(lx + 1, ly)], Pch[k * PS + pidx(lx + 1, ly + 1)]);
}

// Prolong_Bilinear (Prolong.f90:12-39): children of a refined block from
// Bilinear_Interpolate_Extrapolate of h, g (:43-70), Bilinear_Interpolation of
// C with one ghost ring (:74-101), then U, P by Macroscopic_Values_g_AMR
// (Fill_Children_Data, :105-130). New children start zeroed
// (AMR_Rising_2D_OMP.f90:572-577). pairs as in k_restrict.
template <class T>
__global__ __launch_bounds__(256) void k_prolong(DevMeshT<T> m, int cur, const int *__restrict__ pairs, Phys ph64,
                                                 PhysF ph32)
{
    const int ps = pairs[2 * blockIdx.x], c0 = pairs[2 * blockIdx.x + 1];
    const auto *tab = PhysOf<T>::table(m);   // an ensemble mesh: the parent's instance's constants
    const auto &ph = tab ? tab[m.blk[ps].inst] : PhysOf<T>::get(ph64, ph32);
    const int tid = threadIdx.x;
    const T *Fp = m.pop[cur] + (size_t)ps * POPB;
    const T *Cp = m.C + (size_t)ps * CS;
    __shared__ T sC[CS];
    for (int k = 0; k < 4; k++) {
        T *F = m.pop[cur] + (size_t)(c0 + k) * POPB;
        for (int t = tid; t < POPB; t += 256) F[t] = T(0);
        T *C = m.C + (size_t)(c0 + k) * CS;
        for (int t = tid; t < CS; t += 256) C[t] = T(0);
        T *PU = m.PU + (size_t)(c0 + k) * PUB;
        for (int t = tid; t < PUB; t += 256) PU[t] = T(0);
    }
    __syncthreads();
    const int I = tid % NX + 1, J = tid / NX + 1;
    for (int k = 0; k < 4; k++) {
        const int i = I + NX * (k & 1), j = J + NY * (k >> 1);   // fine index 1..2n
        int Xi = i / 2; Xi = Xi > NX - 1 ? NX - 1 : (Xi < 1 ? 1 : Xi);
        int Yj = j / 2; Yj = Yj > NY - 1 ? NY - 1 : (Yj < 1 ? 1 : Yj);
        const T ai = (T)(2 * i - 4 * Xi + 1), bi = T(4) - ai;
        const T aj = (T)(2 * j - 4 * Yj + 1), bj = T(4) - aj;
        T *F = m.pop[cur] + (size_t)(c0 + k) * POPB;
        for (int comp = 0; comp < 18; comp++) {
            int set = comp / 9, q = comp % 9;
            T t0 = bi * Fp[fidx(set, q, Xi, Yj)] + ai * Fp[fidx(set, q, Xi + 1, Yj)];
            T t1 = bi * Fp[fidx(set, q, Xi, Yj + 1)] + ai * Fp[fidx(set, q, Xi + 1, Yj + 1)];
            F[fidx(set, q, I, J)] = (bj * t0 + aj * t1) / T(16);
        }
    }
    // C on (0:nx+1)^2 of each child from the parent's (0:nx+1)^2
    for (int t = tid; t < 4 * FS; t += 256) {
        int k = t / FS, r = t - k * FS, Ic = r % FW, Jc = r / FW;
        int i = Ic + NX * (k & 1), j = Jc + NY * (k >> 1);      // 0..2n+1
        int Xi = i / 2, Yj = j / 2;
        T ai = (T)(2 * i - 4 * Xi + 1), bi = T(4) - ai;
        T aj = (T)(2 * j - 4 * Yj + 1), bj = T(4) - aj;
        T t0 = bi * Cp[cidx(Xi, Yj)] + ai * Cp[cidx(Xi + 1, Yj)];
        T t1 = bi * Cp[cidx(Xi, Yj + 1)] + ai * Cp[cidx(Xi + 1, Yj + 1)];
        m.C[(size_t)(c0 + k) * CS + cidx(Ic, Jc)] = (bj * t0 + aj * t1) / T(16);
    }
    __syncthreads();
    for (int k = 0; k < 4; k++) {
        const T *Cc = m.C + (size_t)(c0 + k) * CS;
        for (int t = tid; t < CS; t += 256) sC[t] = Cc[t];
        __syncthreads();
        const T *F = m.pop[cur] + (size_t)(c0 + k) * POPB;
        T g[9], P, Ux, Uy;
#pragma unroll
        for (int q = 0; q < 9; q++) g[q] = F[fidx(1, q, I, J)];
        macro_g_cell((const T *)sC, I, J, m.blk[ps].dx / 2, g, ph, P, Ux, Uy);
        T *PU = m.PU + (size_t)(c0 + k) * PUB;
        PU[pidx(I, J)] = P;
        PU[PS + pidx(I, J)] = Ux;
        PU[2 * PS + pidx(I, J)] = Uy;
        __syncthreads();
    }
}

// Initialize_AMR_Rising_Bubble (AMR_Rising_2D_OMP.f90:186-225): with C set
// (both rings), U = 0, P = -sigma/R C, h = heq, g = geq with Fb = 0.
// Computed in FP64 for either storage. With nb > 0 bubbles (x0, y0, R per
// bubble, finest cells) C itself is set here first, on both ghost rings:
// 0.5 + 0.5 tanh(2 d / W) with d the distance to the nearest bubble surface,
// min_k(|x - x_k| - R_k), and R in P is the nearest bubble's; else C comes
// from memory and R is the argument, or with inst_R (an ensemble mesh)
// inst_R[instance]. An ensemble mesh takes its instances' fluid constants.
template <class T>
__global__ __launch_bounds__(256) void k_init_fields(DevMeshT<T> m, int cur, const int *__restrict__ list,
                                                     Phys ph_arg, double R, const double *__restrict__ bub, int nb,
                                                     const double *__restrict__ inst_R)
{
    const int slot = list[blockIdx.x];
    const int tid = threadIdx.x;
    __shared__ double sC[CS];
    __shared__ double sR[CS];
    T *C = m.C + (size_t)slot * CS;
    const BlockDev b = m.blk[slot];
    const Phys &ph = m.phs ? m.phs[b.inst] : ph_arg;
    if (inst_R) R = inst_R[b.inst];
    if (nb > 0) {
        for (int t = tid; t < CS; t += 256) {
            const int x = t % CW - 1, y = t / CW - 1;
            // cell centre as the host set-up computes it (bubble_c)
            const double Xc = (double)(b.I * b.dx * NX) + (double)b.dx * ((double)x - 0.5);
            const double Yc = (double)(b.J * b.dx * NY) + (double)b.dx * ((double)y - 0.5);
            double dmin = 0.0, rk = 0.0;
            for (int q = 0; q < nb; q++) {
                const double dxq = Xc - bub[3 * q], dyq = Yc - bub[3 * q + 1];
                const double d = sqrt(dxq * dxq + dyq * dyq) - bub[3 * q + 2];
                if (q == 0 || d < dmin) {
                    dmin = d;
                    rk = bub[3 * q + 2];
                }
            }
            const double c = 0.5 + 0.5 * tanh(2.0 * dmin / ph.W);
            sC[t] = c;
            sR[t] = rk;
            C[t] = (T)(c - c_shift<T>());
        }
    } else {
        for (int t = tid; t < CS; t += 256) sC[t] = (double)C[t] + c_shift<T>();
    }
    __syncthreads();
    const int i = tid % NX + 1, j = tid / NX + 1, dx = b.dx;
    const double dRho3 = (ph.rhoh - ph.rhol) / 3.0;
    double c = sC[cidx(i, j)];
    const double Rc = nb > 0 ? sR[cidx(i, j)] : R;
    double P = -(ph.sigma / Rc * c), Ux = 0.0, Uy = 0.0;
    T *PU = m.PU + (size_t)slot * PUB;
    PU[pidx(i, j)] = (T)P;
    PU[PS + pidx(i, j)] = (T)Ux;
    PU[2 * PS + pidx(i, j)] = (T)Uy;
    double mu = chem_pot(sC, i, j, dx, ph);
    double e[9], hlp[9], Gamma[9], heq[9], gx, gy;
    ed_c(sC, i, j, dx, e);
    gradient_ed(e, gx, gy);
    double tmp = sqrt(gx * gx + gy * gy + 1.0e-32);
    double ni = gx / tmp, nj = gy / tmp;
    double Rho = ph.rhol + c * (ph.rhoh - ph.rhol);
    equilibrium(c, Ux, Uy, ni, nj, ph, Gamma, heq);
    forcing(Gamma, dRho3, mu, e, Ux, Uy, gx, gy, 0.0, hlp);
    T *F = m.pop[cur] + (size_t)slot * POPB;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if (std::is_same<T, float>::value)
            F[fidx(0, q, i, j)] = (T)(heq[q] - wa(q));
        else
            F[fidx(0, q, i, j)] = (T)heq[q];
        F[fidx(1, q, i, j)] = (T)(P * wa(q) + (Gamma[q] - wa(q)) * Rho / 3.0 - 0.5 * hlp[q]);
    }
}

// FP64 diagnostics, stage 1: per-leaf partials in a fixed reduction order
// (min rho, max rho, max|Uy|, mass, area, x-moment, y-moment, Uy-moment,
// non-finite). In an ensemble mesh the block's instance's constants and the
// instance's own frame (y from its first root).
template <class T>
__global__ __launch_bounds__(256) void k_diag_partial(DevMeshT<T> m, const int *__restrict__ list, Phys ph_arg,
                                                      double *part)
{
    const int slot = list[blockIdx.x];
    const int tid = threadIdx.x;
    const int i = tid % NX + 1, j = tid / NX + 1;
    const BlockDev b = m.blk[slot];
    const Phys &ph = m.phs ? m.phs[b.inst] : ph_arg;
    const int J = m.phs ? b.J - b.inst * (m.tile_roots << b.level) : b.J;
    double c = (double)m.C[(size_t)slot * CS + cidx(i, j)] + c_shift<T>();
    double uy = (double)m.PU[(size_t)slot * PUB + 2 * PS + pidx(i, j)];
    double rho = ph.rhol + c * (ph.rhoh - ph.rhol);
    double dx2 = (double)(b.dx * b.dx);
    double w = (1.0 - c) * dx2;
    double xc = (double)(b.I * b.dx * NX) + (double)b.dx * ((double)i - 0.5);
    double yc = (double)(J * b.dx * NY) + (double)b.dx * ((double)j - 0.5);
    __shared__ double red[9][256];
    red[0][tid] = rho;
    red[1][tid] = rho;
    red[2][tid] = fabs(uy);
    red[3][tid] = rho * dx2;
    red[4][tid] = w;
    red[5][tid] = w * xc;
    red[6][tid] = w * yc;
    red[7][tid] = w * uy;
    red[8][tid] = (fabs(c) <= 1.79e308) ? 0.0 : 1.0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] = fmin(red[0][tid], red[0][tid + s]);
            red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
            red[2][tid] = fmax(red[2][tid], red[2][tid + s]);
#pragma unroll
            for (int k = 3; k < 9; k++) red[k][tid] = red[k][tid] + red[k][tid + s];
        }
        __syncthreads();
    }
    if (tid < 9) part[(size_t)blockIdx.x * 9 + tid] = red[tid][0];
}

// stage 2: one workgroup folds the per-leaf partials in leaf-list order
__global__ __launch_bounds__(256) void k_diag_final(const double *part, int n, double *out)
{
    const int tid = threadIdx.x;
    __shared__ double red[9][256];
    double a[9] = {10.0, -10.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = tid; k < n; k += 256) {
        const double *p = part + (size_t)k * 9;
        a[0] = fmin(a[0], p[0]);
        a[1] = fmax(a[1], p[1]);
        a[2] = fmax(a[2], p[2]);
        for (int f = 3; f < 9; f++) a[f] += p[f];
    }
    for (int f = 0; f < 9; f++) red[f][tid] = a[f];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] = fmin(red[0][tid], red[0][tid + s]);
            red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
            red[2][tid] = fmax(red[2][tid], red[2][tid + s]);
            for (int f = 3; f < 9; f++) red[f][tid] = red[f][tid] + red[f][tid + s];
        }
        __syncthreads();
    }
    if (tid < 9) out[tid] = red[tid][0];
}

// stage 2 per instance of an ensemble mesh: workgroup q folds the partials of
// instance q's leaves in leaf-list order (the same tree as k_diag_final)
__global__ __launch_bounds__(256) void k_diag_final_inst(DevMeshT<double> m, const int *__restrict__ list,
                                                         const double *part, int n, double *out)
{
    const int tid = threadIdx.x, q = blockIdx.x;
    __shared__ double red[9][256];
    double a[9] = {10.0, -10.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = tid; k < n; k += 256) {
        if (m.blk[list[k]].inst != q) continue;
        const double *p = part + (size_t)k * 9;
        a[0] = fmin(a[0], p[0]);
        a[1] = fmax(a[1], p[1]);
        a[2] = fmax(a[2], p[2]);
        for (int f = 3; f < 9; f++) a[f] += p[f];
    }
    for (int f = 0; f < 9; f++) red[f][tid] = a[f];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] = fmin(red[0][tid], red[0][tid + s]);
            red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
            red[2][tid] = fmax(red[2][tid], red[2][tid + s]);
            for (int f = 3; f < 9; f++) red[f][tid] = red[f][tid] + red[f][tid + s];
        }
        __syncthreads();
    }
    if (tid < 9) out[(size_t)q * 9 + tid] = red[tid][0];
}

// ---- phase-field mass correction (FP64 on the device) ----------------------
// The launches of a stepping segment return at once when the skip word is set
// or the judge stopped the segment before state `ord` (as the step kernels
// do, so a segment launched ahead of a decision leaves the state whole);
// those of a mesh change (gate 0) always run.
__device__ __forceinline__ bool mass_skipped(const int *skip, int gate, const int *judge, int ord)
{
    return gate && (*skip || (judge && *judge < ord));
}

// stage 1: per leaf, in FP64, part[2k] = sum of C dx^2 over the block's
// cells (the phase-field mass), part[2k+1] = sum of max(0, 4 C (1 - C)) dx^2
// (the interface weight a correction is spread with); a fixed reduction tree
template <class T>
__global__ __launch_bounds__(256) void k_mass_partial(DevMeshT<T> m, const int *__restrict__ list, double *part,
                                                      int gate, const int *judge, int ord)
{
    if (mass_skipped(m.skip, gate, judge, ord)) return;
    const int slot = list[blockIdx.x];
    const int tid = threadIdx.x;
    const int i = tid % NX + 1, j = tid / NX + 1;
    const double dx = (double)m.blk[slot].dx, dx2 = dx * dx;
    const double c = (double)m.C[(size_t)slot * CS + cidx(i, j)] + c_shift<T>();
    double a = c * dx2, w = fmax(0.0, 4.0 * c * (1.0 - c)) * dx2;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {   // a butterfly: every lane ends with the same sums
        a += __shfl_xor(a, o);
        w += __shfl_xor(w, o);
    }
    __shared__ double red[2][4];
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = a;
        red[1][tid >> 6] = w;
    }
    __syncthreads();
    if (tid < 2) part[2 * (size_t)blockIdx.x + tid] = ((red[tid][0] + red[tid][1]) + red[tid][2]) + red[tid][3];
}

// stage 2, workgroup q = instance q: S, W = the sums over the instance's
// leaves in leaf-list order, then by mode (mass[4q .. 4q+3] = target, S at
// the last fold, W, lambda):
//   MASS_SAVE    mass[4q+1] = S (before a mesh change)
//   MASS_TARGET  target = S (the budget starts here)
//   MASS_BUDGET  target += S - mass[4q+1] (the change a mesh change made)
//   MASS_CORRECT lambda = (target - S) / W (0 without interface cells)
__global__ __launch_bounds__(256) void k_mass_final(DevMeshT<double> m, const int *__restrict__ list,
                                                    const double *__restrict__ part, int n, int ninst, double *mass,
                                                    int mode, int gate, const int *judge, int ord)
{
    if (mass_skipped(m.skip, gate, judge, ord)) return;
    const int q = blockIdx.x, tid = threadIdx.x;
    double a = 0.0, w = 0.0;
    for (int k = tid; k < n; k += 256) {
        if (ninst > 1 && m.blk[list[k]].inst != q) continue;
        a += part[2 * (size_t)k];
        w += part[2 * (size_t)k + 1];
    }
    __shared__ double red[2][256];
    red[0][tid] = a;
    red[1][tid] = w;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] += red[0][tid + s];
            red[1][tid] += red[1][tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        double *b = mass + 4 * q;
        const double S = red[0][0], W = red[1][0];
        if (mode == MASS_SAVE) {
            b[1] = S;
        } else if (mode == MASS_TARGET) {
            b[0] = S;
            b[1] = S;
            b[2] = W;
            b[3] = 0.0;
        } else if (mode == MASS_BUDGET) {
            b[0] += S - b[1];
        } else {
            b[1] = S;
            b[2] = W;
            b[3] = W > 0.0 ? (b[0] - S) / W : 0.0;
        }
    }
}

// stage 3: with dC = lambda max(0, 4 C (1 - C)), evaluated in FP64 on every
// leaf cell of the state in buffer cur: C += dC and h_q += Gamma_q(u) dC
// (h's equilibrium share C Gamma_q, Equilibrium_h_g_AMR, LBM_Subroutines.f90:
// 2-21: sum_q h_q moves by dC, the phase field's flux by u dC); the density
// moves by drho = (rhoh - rhol) dC, and g_q += (Gamma_q(u) - w_q) drho / 3
// keeps the cell's velocity and pressure: g's first moment is rho u / 3, its
// zeroth P (Macroscopic_Values_g_AMR, :205-240). u: the cell's last U.
template <class T>
__global__ __launch_bounds__(256) void k_mass_apply(DevMeshT<T> m, int cur, const int *__restrict__ list,
                                                    const double *__restrict__ mass, Phys ph_arg, const int *judge,
                                                    int ord)
{
    if (mass_skipped(m.skip, 1, judge, ord)) return;
    const int slot = list[blockIdx.x];
    const int inst = m.blk[slot].inst;
    const double lam = mass[4 * inst + 3];
    if (lam == 0.0) return;
    const Phys &ph = m.phs ? m.phs[inst] : ph_arg;
    const int tid = threadIdx.x;
    const int i = tid % NX + 1, j = tid / NX + 1;
    T *C = m.C + (size_t)slot * CS + cidx(i, j);
    const double c = (double)*C + c_shift<T>();
    const double dc = lam * fmax(0.0, 4.0 * c * (1.0 - c));
    const double drho3 = (ph.rhoh - ph.rhol) * dc / 3.0;
    const T *PU = m.PU + (size_t)slot * PUB;
    const double ux = (double)PU[PS + pidx(i, j)], uy = (double)PU[2 * PS + pidx(i, j)];
    const double uu = ux * ux + uy * uy;
    T *F = m.pop[cur] + (size_t)slot * POPB;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const double eu = (double)c_ex[q] * ux + (double)c_ey[q] * uy;
        const double gw = wa(q) * (3.0 * eu + 4.5 * eu * eu - 1.5 * uu);   // Gamma_q - w_q
        F[fidx(0, q, i, j)] = (T)((double)F[fidx(0, q, i, j)] + (wa(q) + gw) * dc);
        F[fidx(1, q, i, j)] = (T)((double)F[fidx(1, q, i, j)] + gw * drho3);
    }
    *C = (T)((double)*C + dc);
}

// Ghost-program upload: descriptor k copies desc[3k+2] 16-byte words from
// src + desc[3k] to dst + desc[3k+1] (one workgroup each).
// One upload batch: segment k (one workgroup) copies desc[k].nwords 4-byte
// words from src + desc[k].src_word to the device address desc[k].dst.
__global__ __launch_bounds__(256) void k_scatter_segments(const int *__restrict__ src,
                                                          const ScatterSeg *__restrict__ desc)
{
    const ScatterSeg d = desc[blockIdx.x];
    int *dst = reinterpret_cast<int *>(d.dst);
    const int *from = src + d.src_word;
    for (int t = threadIdx.x; t < d.nwords; t += 256) dst[t] = from[t];
}

// ---------------------------------------------------------------------------
// The persistent segment kernel: every step of a regrid segment in one
// launch. Workgroup k owns leaf list[k] for the whole segment and keeps its
// populations in registers from one step to the next; only what other blocks
// read goes through memory: the post-collision populations (for their hg
// fills) and C (for their C2 fills). Per step:
//   C2 fill (ghost program) -> [U, P] + collision -> post-collision h, g to the
//   exchange buffer -> grid barrier -> hg fill (ghost program, both sets) ->
//   stream in LDS, C = sum h -> C to memory -> grid barrier
// and at the end of the segment the C fill, U, P and Err_Interface, as the
// two-kernel step graph's last launch, with the streamed populations stored
// where that graph leaves them. The arithmetic is the same cell functions on
// the same operands in the same order, so every value is bit for bit the
// graph's. Needs every workgroup resident at once (a cooperative launch,
// which the runtime refuses rather than under-schedules).
// ---------------------------------------------------------------------------

// Grid barrier with a bounded wait, two levels deep so that no memory word
// takes more than one atomic per workgroup group: workgroup k arrives at the
// counter of group k % 8 (its XCD under round-robin dispatch); the last of a
// group arrives at the top counter; the last of those publishes the barrier's
// ordinal to the 8 group flags, which the workgroups poll. sync[0..7] group
// counters, sync[8] top counter, sync[16 + 16 g] group g's flag (each on its
// own 64-byte line), sync[15] the fault word. A wait longer than
// kBarrierSpins polls, or a fault word set by another workgroup, sets the
// fault word and returns false: the caller exits, and so do the other
// workgroups at their next barrier.
// Visibility: every value one workgroup writes for another in the segment
// kernel is written with an agent-scope atomic store (through the XCD's L2 to
// the device-coherent level) and read with an agent-scope atomic load, and
// each wave waits for its stores to be acknowledged (s_waitcnt) before the
// workgroup arrives. So no L2 write-back or invalidation is needed: an
// agent-scope release / acquire per workgroup, which writes back and
// invalidates the XCD's whole L2, measured 70 us per barrier at 300
// workgroups.
constexpr unsigned kBarrierSpins = 1u << 22;

__device__ __forceinline__ bool grid_barrier(unsigned *sync, unsigned epoch)
{
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);   // this wave's stores acknowledged
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nwg = gridDim.x, grp = blockIdx.x & 7u;
        const unsigned gsize = (nwg - grp + 7u) >> 3, ngrp = nwg < 8u ? nwg : 8u;
        int *fault = reinterpret_cast<int *>(sync + 15);
        unsigned *flag = sync + 16 + 16 * grp;
        const unsigned g_old = __hip_atomic_fetch_add(sync + grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g_old + 1u == gsize * epoch) {
            const unsigned t_old = __hip_atomic_fetch_add(sync + 8, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t_old + 1u == ngrp * epoch)
                for (unsigned f = 0; f < 8u; f++)
                    __hip_atomic_store(sync + 16 + 16 * f, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        int ok = 1;
        for (unsigned spins = 0; __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch; spins++) {
            if ((spins & 255u) == 255u && __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (spins >= kBarrierSpins) {
                __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// One stage of the hg ghost program for both population sets (recipes [b, e)
// in sN, items (set, population, recipe) with the recipe fastest); sF holds
// the 18 population arrays, lcf cells apart; X is the exchange buffer.
template <bool GLOBAL_ONLY, int U, class T>
__device__ __forceinline__ void hg_stage_both(const PNode *sN, int b, int e, const T *__restrict__ X, T *sF, int lcf,
                                              int tid)
{
    const int ns = e - b, nitem = 18 * ns;
    if (ns <= 0) return;
    const unsigned mq = quot_mul(ns);
    for (int base = 0; base < nitem; base += 256 * U) {
        T v[U];
#pragma unroll
        for (int r = 0; r < U; r++) {
            const int t = base + tid + 256 * r;
            if (t < nitem) {
                const int c = quot(t, mq), set = c >= 9 ? 1 : 0;
                v[r] = eval_hg<GLOBAL_ONLY, true>(sN[b + t - c * ns], set, c - 9 * set, X, sF + 9 * set * lcf, lcf);
            }
        }
#pragma unroll
        for (int r = 0; r < U; r++) {
            const int t = base + tid + 256 * r;
            if (t < nitem) {
                const int c = quot(t, mq);
                sF[c * lcf + sN[b + t - c * ns].dest] = v[r];
            }
        }
    }
}

// Perfect_Shift_AMR / Streaming_LaxWendroff_AMR (LBM_Subroutines.f90:41-90) of
// one cell of one set from its LDS copy into registers (stream_cell's
// arithmetic)
template <class T>
__device__ __forceinline__ void stream_cell_regs(const T *sF, int lcf, int i, int j, int dx, T f[9])
{
    const T cfl = inv_dx_of<T>(dx);   // 1 / dx, exact
    const T a = cfl / T(2) * (T(1) + cfl), b = T(1) - cfl * cfl, c = cfl / T(2) * (T(1) - cfl);
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if (q == 0)
            f[q] = sF[lidx(q, i, j, lcf)];
        else if (cfl > T(0.9))
            f[q] = sF[lidx(q, i - c_ex[q], j - c_ey[q], lcf)];
        else
            f[q] = a * sF[lidx(q, i - c_ex[q], j - c_ey[q], lcf)] + b * sF[lidx(q, i, j, lcf)] -
                   c * sF[lidx(q, i + c_ex[q], j + c_ey[q], lcf)];
    }
}

// A C fill (Fill_Ghost_Cells_C / _C2 by ghost program, as k_prog_c) of the
// block into sC, its own ghost values also into memory. nd0: this thread's
// first recipe (kept from one step to the next), cin: the block's C at this
// thread's cell.
template <class T>
__device__ __forceinline__ void fill_c_block(const PHdr &hd, const PNode &nd0, const PNode *__restrict__ nodes,
                                             const T *__restrict__ Cg, T *C, T *sC, T cin, int i, int j, int tid)
{
    const int total = hd.nst > 0 ? stage_end(hd, hd.nst - 1) : 0;
    const int n1 = hd.nst > 0 ? hd.end[0] : 0;
    T gv[9];
    if (tid < total) prefetch_node<true>(nd0, Cg, 0, gv);
    if (tid < n1) {
        const T v = eval_pref<true>(nd0, gv, (const T *)sC, 0);
        sC[nd0.dest] = v;
        if (nd0.dest < CS) C[nd0.dest] = v;
    }
    sC[cidx(i, j)] = cin;
    __syncthreads();
    int begin = 0;
    for (int s = 0; s < hd.nst; s++) {
        const int end = stage_end(hd, s);
        if (s > 0 && tid >= begin && tid < end) {
            const T v = eval_pref(nd0, gv, (const T *)sC, 0);
            sC[nd0.dest] = v;
            if (nd0.dest < CS) C[nd0.dest] = v;
        }
        for (int t = tid + 256; t < total; t += 256) {
            if (t < begin || t >= end) continue;
            const PNode nd = load_node(nodes + hd.off + t);
            const T v = eval_node<true>(nd, Cg, 0, (const T *)sC, 0);
            sC[nd.dest] = v;
            if (nd.dest < CS) C[nd.dest] = v;
        }
        begin = end;
        __syncthreads();
    }
}

// the grid barrier's counters and flags to zero before a segment kernel (its
// fault word, sync[15], stays until the host has read it)
__global__ __launch_bounds__(64) void k_sync_reset(unsigned *sync)
{
    for (int t = threadIdx.x; t < kSegSyncWords; t += 64)
        if (t != 15) sync[t] = 0u;
}

// byte offset of the staged hg recipes in the segment kernel's LDS, after 18
// population arrays of lcf cells
__host__ __device__ constexpr size_t seg_nodes_offset(int lcf, size_t esz) { return (18 * (size_t)lcf * esz + 15) & ~size_t(15); }

template <class T, int U0, int U1, bool ENS>
__global__ __launch_bounds__(256, 2) void k_segment(DevMeshT<T> m, int cur, const int *__restrict__ list, Prog pc2,
                                                 Prog phg, Prog pc1, Phys ph64, PhysF ph32, int len,
                                                 double *err_slot, unsigned *sync)
{
    const auto &ph_arg = PhysOf<T>::get(ph64, ph32);
    const int k = blockIdx.x, tid0 = threadIdx.x, tid = tid0;
    const int skip = *m.skip;
    const int slot = list[k];
    const PHdr h2 = pc2.hdr[k], hh = phg.hdr[k], h1 = pc1.hdr[k];
    if (skip) return;   // uniform: every workgroup reads the same word before any barrier
    const BlockDev me = m.blk[slot];
    const auto &ph = ENS ? PhysOf<T>::table(m)[me.inst] : ph_arg;
    const int i0 = tid % NX + 1, j0 = tid / NX + 1, i = i0, j = j0;
    const T *__restrict__ Cg = m.C;
    T *C = m.C + (size_t)slot * CS;
    T *PU = m.PU + (size_t)slot * PUB;
    const T *F0 = m.pop[cur] + (size_t)slot * POPB;
    const T *__restrict__ X = m.pop[cur ^ 1];   // exchange buffer: post-collision populations
    T *Xb = m.pop[cur ^ 1] + (size_t)slot * POPB;

    extern __shared__ __align__(16) unsigned char seg_lds[];
    const int lcf = FS + phg.naux;
    T *sF = reinterpret_cast<T *>(seg_lds);
    PNode *sN = reinterpret_cast<PNode *>(seg_lds + seg_nodes_offset(lcf, sizeof(T)));
    __shared__ T sC[LC_C];

    // the block's hg recipes, staged once for the whole segment
    const int nn = hh.nst > 0 ? stage_end(hh, hh.nst - 1) : 0;
    {
        const int4 *src = reinterpret_cast<const int4 *>(phg.nodes + hh.off);
        int4 *dst = reinterpret_cast<int4 *>(sN);
        for (int t = tid; t < 3 * nn; t += 256) dst[t] = src[t];
    }
    // the block's populations between steps: the own cells of sF (the
    // stream's result), so that only the collision holds them in registers
    const int n2 = h2.nst > 0 ? stage_end(h2, h2.nst - 1) : 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        sF[lidx(q, i, j, lcf)] = F0[fidx(0, q, i, j)];
        sF[lidx(9 + q, i, j, lcf)] = F0[fidx(1, q, i, j)];
    }
    T csum = C[cidx(i, j)];
    unsigned epoch = 0;
    for (int s = 0; s < len; s++) {
        // the cell, its size and the thread taken afresh each step (an empty
        // asm the compiler cannot see through), so that the stencils'
        // addresses and the constants derived from them are not hoisted out
        // of the loop and held in registers for the whole segment
        int i = i0, j = j0, dx = me.dx, tid = tid0;
        asm volatile("" : "+v"(i), "+v"(j), "+v"(dx), "+v"(tid));
        // Fill_Ghost_Cells_C2, then U, P of the previous step (none before the
        // first: U, P are current in memory) and Collision_h_g_AMR
        PNode nd2;
        if (tid < n2) nd2 = load_node(pc2.nodes + h2.off + tid);
        fill_c_block(h2, nd2, pc2.nodes, Cg, C, sC, csum, i, j, tid);
        {
            T h[9], g[9], P, Ux, Uy;
#pragma unroll
            for (int q = 0; q < 9; q++) {
                h[q] = sF[lidx(q, i, j, lcf)];
                g[q] = sF[lidx(9 + q, i, j, lcf)];
            }
            if (s > 0) {
                macro_g_cell((const T *)sC, i, j, dx, g, ph, P, Ux, Uy);
            } else {
                P = PU[pidx(i, j)];
                Ux = PU[PS + pidx(i, j)];
                Uy = PU[2 * PS + pidx(i, j)];
            }
            collide_cell((const T *)sC, i, j, dx, P, Ux, Uy, ph, h, g);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                stg_coh(Xb + fidx(0, q, i, j), h[q]);
                stg_coh(Xb + fidx(1, q, i, j), g[q]);
                sF[lidx(q, i, j, lcf)] = h[q];
                sF[lidx(9 + q, i, j, lcf)] = g[q];
            }
        }
        if (!grid_barrier(sync, ++epoch)) return;
        // Fill_Ghost_Cells_hg of both sets from the neighbours' post-collision
        // values, then the stream and C = sum h (Macroscopic_Values_h_AMR)
        int begin = 0;
        for (int st = 0; st < hh.nst; st++) {
            const int end = stage_end(hh, st);
            if (st == 0)
                hg_stage_both<true, U0>(sN, begin, end, X, sF, lcf, tid);
            else
                hg_stage_both<false, U1>(sN, begin, end, X, sF, lcf, tid);
            begin = end;
            __syncthreads();
        }
        {
            T f[9];
            stream_cell_regs((const T *)sF, lcf, i, j, dx, f);
            csum = T(0);
#pragma unroll
            for (int q = 0; q < 9; q++) csum += f[q];
            T f2[9];
            stream_cell_regs((const T *)(sF + 9 * lcf), lcf, i, j, dx, f2);
            __syncthreads();   // every cell's stream has read its neighbours
#pragma unroll
            for (int q = 0; q < 9; q++) {
                sF[lidx(q, i, j, lcf)] = f[q];
                sF[lidx(9 + q, i, j, lcf)] = f2[q];
            }
        }
        stg_coh(C + cidx(i, j), csum);
        if (!grid_barrier(sync, ++epoch)) return;
    }
    // Fill_Ghost_Cells_C, Macroscopic_Values_g_AMR and Err_Interface
    // (Refine_Derefine.f90:327-352) for a regrid decision after the segment
    PNode nd1;
    if (tid < (h1.nst > 0 ? stage_end(h1, h1.nst - 1) : 0)) nd1 = load_node(pc1.nodes + h1.off + tid);
    fill_c_block(h1, nd1, pc1.nodes, Cg, C, sC, csum, i, j, tid);
    {
        const T a = sC[cidx(i + 1, j)] - sC[cidx(i - 1, j)];
        const T b = sC[cidx(i, j + 1)] - sC[cidx(i, j - 1)];
        T dc = a * a + b * b;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const T x = __shfl_xor(dc, o);
            if (x > dc) dc = x;
        }
        __shared__ T wmax[4];
        if ((tid & 63) == 0) wmax[tid >> 6] = dc;
        __syncthreads();
        if (tid == 0) {
            T e = wmax[0];
            for (int w = 1; w < 4; w++)
                if (wmax[w] > e) e = wmax[w];
            err_slot[slot] = (double)(sqrt(e) / (T(2) * (T)me.dx));
        }
    }
    T g[9], P, Ux, Uy;
#pragma unroll
    for (int q = 0; q < 9; q++) g[q] = sF[lidx(9 + q, i, j, lcf)];
    macro_g_cell((const T *)sC, i, j, me.dx, g, ph, P, Ux, Uy);
    PU[pidx(i, j)] = P;
    PU[PS + pidx(i, j)] = Ux;
    PU[2 * PS + pidx(i, j)] = Uy;
    T *Fend = m.pop[cur] + (size_t)slot * POPB;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        Fend[fidx(0, q, i, j)] = sF[lidx(q, i, j, lcf)];
        Fend[fidx(1, q, i, j)] = g[q];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class T, bool ENS>
static void prog_c(const DevMesh &dm, int cur, const int *list, int n, int op, const Prog &prog, const Phys &ph,
                   hipStream_t s, double *err_slot, const StepCtl &ctl)
{
    const DevMeshT<T> m = typed<T>(dm);
    dim3 g(n), b(256);
    if (op == OP_MACRO_G_COLLIDE)
        hipLaunchKernelGGL((k_prog_c<T, OP_MACRO_G_COLLIDE, ENS>), g, b, 0, s, m, cur, list, prog, ph, dm.phf,
                           err_slot, ctl);
    else if (op == OP_COLLIDE)
        hipLaunchKernelGGL((k_prog_c<T, OP_COLLIDE, ENS>), g, b, 0, s, m, cur, list, prog, ph, dm.phf, err_slot, ctl);
    else if (op == OP_MACRO_G)
        hipLaunchKernelGGL((k_prog_c<T, OP_MACRO_G, ENS>), g, b, 0, s, m, cur, list, prog, ph, dm.phf, err_slot, ctl);
    else
        hipLaunchKernelGGL((k_prog_c<T, OP_NONE, ENS>), g, b, 0, s, m, cur, list, prog, ph, dm.phf, err_slot, ctl);
}

void launch_prog_c(const DevMesh &m, int cur, const int *list, int n, int kind, int op, const Prog &prog,
                   const Phys &ph, hipStream_t s, double *err_slot, const StepCtl &ctl)
{
    if (n <= 0) return;
    (void)kind;   // the program decides the ring; the op needs C with the rings it reads
    const bool ens = m.phs != nullptr;
    if (m.f32)
        ens ? prog_c<float, true>(m, cur, list, n, op, prog, ph, s, err_slot, ctl)
            : prog_c<float, false>(m, cur, list, n, op, prog, ph, s, err_slot, ctl);
    else
        ens ? prog_c<double, true>(m, cur, list, n, op, prog, ph, s, err_slot, ctl)
            : prog_c<double, false>(m, cur, list, n, op, prog, ph, s, err_slot, ctl);
}

__global__ void k_fill_int(int *p, int v) { *p = v; }

void launch_fill_int(int *p, int v, hipStream_t s) { hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, s, p, v); }

// the ghost-program headers of the leaves in launch order, from the headers
// kept by block slot: out[kind][k] = slot[kind][list[k]]
__global__ __launch_bounds__(256) void k_gather_headers(const PHdr *s0, const PHdr *s1, const PHdr *s2, PHdr *o0,
                                                        PHdr *o1, PHdr *o2, const int *__restrict__ list, int n)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) {
        const int slot = list[k];
        o0[k] = s0[slot];
        o1[k] = s1[slot];
        o2[k] = s2[slot];
    }
}

void launch_gather_headers(PHdr *const slot[3], PHdr *const out[3], const int *list, int n, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_gather_headers, dim3((n + 255) / 256), dim3(256), 0, s, slot[0], slot[1], slot[2], out[0],
                           out[1], out[2], list, n);
}

// the hg kernel of operand depth d: 2/1, 4/2 or 8/4 recipes in flight per thread
template <class T>
static void hg_launch(int d, dim3 grid, size_t lds, hipStream_t s, const DevMeshT<T> &m, int cur, const int *list, int n,
                      const Prog &prog, int flags, const int *judge, int ord)
{
    if (d == 0)
        hipLaunchKernelGGL((k_prog_hg<T, 2, 1>), grid, dim3(256), lds, s, m, cur, list, n, prog, flags, judge, ord);
    else if (d == 2)
        hipLaunchKernelGGL((k_prog_hg<T, 8, 4>), grid, dim3(256), lds, s, m, cur, list, n, prog, flags, judge, ord);
    else
        hipLaunchKernelGGL((k_prog_hg<T, 4, 2>), grid, dim3(256), lds, s, m, cur, list, n, prog, flags, judge, ord);
}

void launch_prog_hg(const DevMesh &m, int cur, const int *list, int n, const Prog &prog, bool stream_after,
                    bool macro_h, bool copy_ghosts, hipStream_t s, const int *judge, int ord)
{
    if (n <= 0) return;
    int flags = (prog.hdr ? 1 : 0) | (stream_after ? 2 : 0) | (macro_h ? 4 : 0) | (copy_ghosts ? 8 : 0) |
                (prog.hdr && !stream_after ? 16 : 0);
    const dim3 grid(16 * ((n + 7) / 8));   // two sets per list entry, entries padded to a multiple of 8
    // LDS for the launch's largest program only: fewer aux cells, more workgroups per CU
    const int lcf = FS + prog.naux;
    const size_t esz = m.f32 ? sizeof(float) : sizeof(double);
    const size_t lds = hg_nodes_offset(lcf, esz) + (size_t)prog.nmax * sizeof(PNode);
    if (m.f32)
        hg_launch<float>(m.hg_depth, grid, lds, s, typed<float>(m), cur, list, n, prog, flags, judge, ord);
    else
        hg_launch<double>(m.hg_depth, grid, lds, s, typed<double>(m), cur, list, n, prog, flags, judge, ord);
}

void launch_macro_h(const DevMesh &m, int cur, const int *list, int n, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_macro_h<float>, dim3(n), dim3(256), 0, s, typed<float>(m), cur, list);
    else
        hipLaunchKernelGGL(k_macro_h<double>, dim3(n), dim3(256), 0, s, typed<double>(m), cur, list);
}

void launch_err(const DevMesh &m, const int *list, int n, double *err, double *err_slot, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_err<float>, dim3(n), dim3(256), 0, s, typed<float>(m), list, err, err_slot);
    else
        hipLaunchKernelGGL(k_err<double>, dim3(n), dim3(256), 0, s, typed<double>(m), list, err, err_slot);
}

int regrid_check_max_blocks()
{
    // dynamic LDS: two bit masks over the block slots, beside the kernel's
    // static lists, within the 160 KB a gfx950 workgroup may hold
    return (int)(((160 * 1024 - kRegridStaticLds - 2 * sizeof(unsigned)) / (2 * sizeof(unsigned))) * 32);
}

hipError_t launch_regrid_check(const DevMesh &m, int nblks, const double *err_slot, int min_levels, int max_levels,
                               int levelmin, int levelmax, double refin_tol, double deref_tol, int *changed,
                               hipStream_t s, const int *judge, int jstop, int *flags)
{
    if (nblks > regrid_check_max_blocks()) return hipErrorInvalidValue;
    const size_t masks = sizeof(unsigned) * 2 * (size_t)((nblks + 31) / 32 + 1);
    // the packed block words beside the masks where the LDS holds them
    const size_t with_info = masks + sizeof(unsigned) * (size_t)nblks;
    const bool info = with_info <= 160 * 1024 - kRegridStaticLds - 256;
    const size_t lds = info ? with_info : masks;
    const void *fn = info ? reinterpret_cast<const void *>(&k_regrid_check<true>)
                          : reinterpret_cast<const void *>(&k_regrid_check<false>);
    // static and dynamic LDS beyond the default 64 KB per workgroup is asked for explicitly
    static size_t granted[2] = {64 * 1024 - kRegridStaticLds - 256, 64 * 1024 - kRegridStaticLds - 256};   // 256: the compiler's own alignment
    if (lds > granted[info]) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        granted[info] = lds;
    }
    if (info)
        hipLaunchKernelGGL(k_regrid_check<true>, dim3(1), dim3(1024), lds, s, typed<double>(m), nblks, err_slot,
                           min_levels, max_levels, levelmin, levelmax, refin_tol, deref_tol, changed, judge, jstop, flags);
    else
        hipLaunchKernelGGL(k_regrid_check<false>, dim3(1), dim3(1024), lds, s, typed<double>(m), nblks, err_slot,
                           min_levels, max_levels, levelmin, levelmax, refin_tol, deref_tol, changed, judge, jstop, flags);
    return hipGetLastError();
}

// the decision after a judged segment that ends between full regrids:
// changed[2] = *judge, and "changed, refinement" iff the judge stopped it
__global__ void k_judge_check(const int *judge, int jstop, int *changed)
{
    const int J = *judge;
    changed[2] = J;
    if (J < jstop) changed[0] = changed[1] = 1;
}

void launch_judge_check(const int *judge, int jstop, int *changed, hipStream_t s)
{
    hipLaunchKernelGGL(k_judge_check, dim3(1), dim3(1), 0, s, judge, jstop, changed);
}

// the learned predictor's refinement flag per block slot: logit 0 of leaf k
// (launch order) above the threshold, on a block whose Err_Interface already
// passes gate (the interface's profile reaches into it)
__global__ __launch_bounds__(256) void k_pred_flags(const float *__restrict__ logits, const int *__restrict__ list,
                                                    int n, float threshold, const double *__restrict__ err_slot,
                                                    double gate, int *pred)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) {
        const int slot = list[k];
        pred[slot] = (logits[2 * k] > threshold && err_slot[slot] > gate) ? 1 : 0;
    }
}

void launch_pred_flags(const float *logits, const int *list, int n, float threshold, const double *err_slot,
                       double gate, int *pred, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_pred_flags, dim3((n + 255) / 256), dim3(256), 0, s, logits, list, n, threshold, err_slot,
                           gate, pred);
}

void launch_restrict(const DevMesh &m, int cur, const int *pairs, int n, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_restrict<float>, dim3(n), dim3(256), 0, s, typed<float>(m), cur, pairs);
    else
        hipLaunchKernelGGL(k_restrict<double>, dim3(n), dim3(256), 0, s, typed<double>(m), cur, pairs);
}

void launch_prolong(const DevMesh &m, int cur, const int *pairs, int n, const Phys &ph, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_prolong<float>, dim3(n), dim3(256), 0, s, typed<float>(m), cur, pairs, ph, m.phf);
    else
        hipLaunchKernelGGL(k_prolong<double>, dim3(n), dim3(256), 0, s, typed<double>(m), cur, pairs, ph, m.phf);
}

void launch_init_fields(const DevMesh &m, int cur, const int *list, int n, const Phys &ph, double R, hipStream_t s,
                        const double *bub, int nb, const double *inst_R)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_init_fields<float>, dim3(n), dim3(256), 0, s, typed<float>(m), cur, list, ph, R, bub, nb,
                           inst_R);
    else
        hipLaunchKernelGGL(k_init_fields<double>, dim3(n), dim3(256), 0, s, typed<double>(m), cur, list, ph, R, bub,
                           nb, inst_R);
}

void launch_diag(const DevMesh &m, const int *list, int n, const Phys &ph, double *partial, double *out, hipStream_t s,
                 int ninst, double *out_inst)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_diag_partial<float>, dim3(n), dim3(256), 0, s, typed<float>(m), list, ph, partial);
    else
        hipLaunchKernelGGL(k_diag_partial<double>, dim3(n), dim3(256), 0, s, typed<double>(m), list, ph, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, s, partial, n, out);
    if (ninst > 0 && out_inst)
        hipLaunchKernelGGL(k_diag_final_inst, dim3(ninst), dim3(256), 0, s, typed<double>(m), list, partial, n,
                           out_inst);
}

void launch_mass_sum(const DevMesh &m, const int *list, int n, int ninst, double *part, double *mass, int mode,
                     int gate, const int *judge, int ord, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_mass_partial<float>, dim3(n), dim3(256), 0, s, typed<float>(m), list, part, gate, judge,
                           ord);
    else
        hipLaunchKernelGGL(k_mass_partial<double>, dim3(n), dim3(256), 0, s, typed<double>(m), list, part, gate, judge,
                           ord);
    hipLaunchKernelGGL(k_mass_final, dim3(ninst), dim3(256), 0, s, typed<double>(m), list, part, n, ninst, mass, mode,
                       gate, judge, ord);
}

void launch_mass_apply(const DevMesh &m, int cur, const int *list, int n, const double *mass, const Phys &ph,
                       const int *judge, int ord, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_mass_apply<float>, dim3(n), dim3(256), 0, s, typed<float>(m), cur, list, mass, ph, judge,
                           ord);
    else
        hipLaunchKernelGGL(k_mass_apply<double>, dim3(n), dim3(256), 0, s, typed<double>(m), cur, list, mass, ph,
                           judge, ord);
}

void launch_scatter_segments(const void *src, const ScatterSeg *desc, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter_segments, dim3(n), dim3(256), 0, s, static_cast<const int *>(src), desc);
}

// ---- persistent segment kernel -------------------------------------------
size_t segment_lds_bytes(const DevMesh &m, const Prog &hg)
{
    const size_t esz = m.f32 ? sizeof(float) : sizeof(double);
    return seg_nodes_offset(FS + hg.naux, esz) + (size_t)hg.nmax * sizeof(PNode);
}

template <class T, bool ENS>
static const void *segment_kernel(int depth)
{
    if (depth == 0) return reinterpret_cast<const void *>(&k_segment<T, 2, 1, ENS>);
    if (depth == 2) return reinterpret_cast<const void *>(&k_segment<T, 8, 4, ENS>);
    return reinterpret_cast<const void *>(&k_segment<T, 4, 2, ENS>);
}

static const void *segment_kernel_of(const DevMesh &m)
{
    const bool ens = m.phs != nullptr;
    if (m.f32) return ens ? segment_kernel<float, true>(m.hg_depth) : segment_kernel<float, false>(m.hg_depth);
    return ens ? segment_kernel<double, true>(m.hg_depth) : segment_kernel<double, false>(m.hg_depth);
}

int segment_capacity(const DevMesh &m, const Prog &hg)
{
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const size_t lds = segment_lds_bytes(m, hg);
    // dynamic LDS beyond the default 64 KB per workgroup is asked for explicitly
    if (hipFuncSetAttribute(segment_kernel_of(m), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, segment_kernel_of(m), 256, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return per_cu * ncu;
}

hipError_t launch_segment(const DevMesh &dm, int cur, const int *list, int n, const Prog &c2, const Prog &hg,
                          const Prog &c1, const Phys &ph, int len, double *err_slot, unsigned *sync, hipStream_t s)
{
    const size_t lds = segment_lds_bytes(dm, hg);
    const PhysF phf = dm.phf;
    void *args[11];
    DevMeshT<double> m64 = typed<double>(dm);
    DevMeshT<float> m32 = typed<float>(dm);
    args[0] = dm.f32 ? static_cast<void *>(&m32) : static_cast<void *>(&m64);
    args[1] = &cur;
    args[2] = &list;
    args[3] = const_cast<Prog *>(&c2);
    args[4] = const_cast<Prog *>(&hg);
    args[5] = const_cast<Prog *>(&c1);
    args[6] = const_cast<Phys *>(&ph);
    args[7] = const_cast<PhysF *>(&phf);
    args[8] = &len;
    args[9] = &err_slot;
    args[10] = &sync;
    // a plain launch: the host has checked that n workgroups fit on the GPU
    // at once, and each barrier's wait is bounded (a cooperative launch, which
    // guarantees residency, cost 100-140 us per segment here)
    hipLaunchKernelGGL(k_sync_reset, dim3(1), dim3(64), 0, s, sync);
    return hipLaunchKernel(segment_kernel_of(dm), dim3(n), dim3(256), args, lds, s);
}

}  // namespace amrlbm
