/*
 * amroracle.c -- CPU restatement (plain C, FP64, scalar, single thread) of the
 * reference's two-phase AMR-LBM step (AbbasHub/AMR-LBM-OpenMP-2D).
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests/ compare the HIP
 * library against it, bench.py times it as the CPU baseline. Nothing in the
 * product links it.
 *
 * Every routine follows the reference routine cited beside it, in the same
 * floating-point operation order (Fortran evaluates a+b+c as (a+b)+c; integer
 * factors are converted to double before the multiply), so that results agree
 * with the compiled reference to the last bit wherever the reference's compiler
 * does not reorder. Build with -O2 -ffp-contract=off and no -march flag (no FMA),
 * which is how the reference is built here.
 *
 * The eight directional ghost fills of each of Fill_Ghost_Cells_{C,C2,hg}.f90
 * are restated as one rule per axis (fine_axis / coarse_axis below): each
 * directional routine of the reference picks 3 source cells per axis, in an
 * order that runs from the interface outward on the normal axis and upward on
 * the tangential axis, interpolates along y first and x second, and divides at
 * the end. The snapshots written by the reference harness pin this restatement
 * direction by direction (tests/test_oracle.py).
 */
#include "amroracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* D2Q9 lattice (D2Q9_Mod.f90:6-8) */
static const int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static double Wa[9];
static int wa_ready = 0;
static void lattice_init(void)
{
    if (wa_ready) return;
    static const int num[9] = {16, 4, 4, 4, 4, 1, 1, 1, 1};
    for (int i = 0; i < 9; i++) Wa[i] = (double)num[i] / 36.0;
    wa_ready = 1;
}

/* ---- layout helpers (reference shapes, column-major) ---- */
#define NX (p->nx)
#define NY (p->ny)
/* h,g(9, 0:nx+1, 0:ny+1) */
static inline int fidx(const orc_params *p, int q, int x, int y) { return q + 9 * (x + (NX + 2) * y); }
/* C(-1:nx+2, -1:ny+2) */
static inline int cidx(const orc_params *p, int x, int y) { return (x + 1) + (NX + 4) * (y + 1); }
/* P(nx,ny), 1-based */
static inline int pidx(const orc_params *p, int i, int j) { return (i - 1) + NX * (j - 1); }
static inline size_t fsize(const orc_params *p) { return (size_t)9 * (NX + 2) * (NY + 2); }
static inline size_t csize(const orc_params *p) { return (size_t)(NX + 4) * (NY + 4); }
static inline size_t psize(const orc_params *p) { return (size_t)NX * NY; }

void orc_reference_physics(orc_params *p, double D, int Bo, double Ar, double gy, double W,
                           double rhol, double rhoh, double visc_ratio)
{
    /* Set_Rising_Bubble_Parameters (AMR_Rising_2D_OMP.f90:241-259) */
    p->rhol = rhol;
    p->rhoh = rhoh;
    p->gy = gy;
    p->sigma = gy * (rhoh - rhol) * (D * D) / (double)Bo;
    double muh = sqrt(gy * rhoh * (rhoh - rhol) * (D * D * D)) / Ar;
    double mul = muh / visc_ratio;
    p->tauh = 3.0 * (muh / rhoh);
    p->taul = 3.0 * (mul / rhol);
    p->W = W;
    p->beta = 12.0 * p->sigma / W;
    p->kappa = 1.5 * p->sigma * W;
}

/* ======================================================================
 * LBM_Subroutines.f90
 * ====================================================================== */

/* Equilibrium_h_g_AMR (LBM_Subroutines.f90:2-21) */
static void equilibrium(const orc_params *p, double C, double U, double V, double ni, double nj,
                        double Gamma[9], double heq[9])
{
    double tmp = ((12.0 * (1.0 - C)) / p->W) * p->M0;
    double U2 = U * U + V * V;
    for (int i = 0; i < 9; i++) {
        double eU = (double)ex[i] * U + (double)ey[i] * V;
        double eN = (double)ex[i] * ni + (double)ey[i] * nj;
        Gamma[i] = Wa[i] * ((1.0 + eU * (3.0 + 4.5 * eU)) - 1.5 * U2);
        heq[i] = C * (Gamma[i] + tmp * Wa[i] * eN);
    }
}

/* normal_AMR (LBM_Subroutines.f90:25-37) */
static void normal(double dcdx, double dcdy, double *ni, double *nj)
{
    double tmp = sqrt(dcdx * dcdx + dcdy * dcdy + 1.0e-32);
    *ni = dcdx / tmp;
    *nj = dcdy / tmp;
}

/* Perfect_Shift_AMR (LBM_Subroutines.f90:41-62) */
void orc_perfect_shift_block(const orc_params *p, double *f)
{
    double *fnew = malloc(sizeof(double) * 8 * NX * NY);
    for (int Y = 1; Y <= NY; Y++)
        for (int X = 1; X <= NX; X++)
            for (int i = 1; i < 9; i++)
                fnew[(i - 1) + 8 * ((X - 1) + NX * (Y - 1))] = f[fidx(p, i, X - ex[i], Y - ey[i])];
    for (int Y = 1; Y <= NY; Y++)
        for (int X = 1; X <= NX; X++)
            for (int i = 1; i < 9; i++)
                f[fidx(p, i, X, Y)] = fnew[(i - 1) + 8 * ((X - 1) + NX * (Y - 1))];
    free(fnew);
}

/* Streaming_LaxWendroff_AMR (LBM_Subroutines.f90:66-90) */
void orc_lax_wendroff_block(const orc_params *p, double cfl, double *f)
{
    double *fnew = malloc(sizeof(double) * 8 * NX * NY);
    double a = cfl / 2.0 * (1.0 + cfl);
    double b = 1.0 - cfl * cfl;
    double c = cfl / 2.0 * (1.0 - cfl);
    for (int Y = 1; Y <= NY; Y++)
        for (int X = 1; X <= NX; X++)
            for (int i = 1; i < 9; i++)
                fnew[(i - 1) + 8 * ((X - 1) + NX * (Y - 1))] =
                    a * f[fidx(p, i, X - ex[i], Y - ey[i])] + b * f[fidx(p, i, X, Y)] -
                    c * f[fidx(p, i, X + ex[i], Y + ey[i])];
    for (int Y = 1; Y <= NY; Y++)
        for (int X = 1; X <= NX; X++)
            for (int i = 1; i < 9; i++)
                f[fidx(p, i, X, Y)] = fnew[(i - 1) + 8 * ((X - 1) + NX * (Y - 1))];
    free(fnew);
}

/* Laplacian_AMR (LBM_Subroutines.f90:316-333) */
static double laplacian(const orc_params *p, const double *C, int i, int j)
{
    return (C[cidx(p, i - 1, j - 1)] + C[cidx(p, i + 1, j - 1)] + C[cidx(p, i - 1, j + 1)] +
            C[cidx(p, i + 1, j + 1)] +
            4.0 * (C[cidx(p, i, j - 1)] + C[cidx(p, i - 1, j)] + C[cidx(p, i + 1, j)] +
                   C[cidx(p, i, j + 1)]) -
            20.0 * C[cidx(p, i, j)]) / 6.0;
}

/* Chemical_Potential_AMR (LBM_Subroutines.f90:94-112), one cell */
static double chem_pot(const orc_params *p, int dx, const double *C, int i, int j)
{
    double c = C[cidx(p, i, j)];
    double d2c = laplacian(p, C, i, j);
    return 4.0 * p->beta * c * (c - 1.0) * (c - 0.5) - p->kappa * d2c / (double)(dx * dx);
}

/* ED_C_AMR (LBM_Subroutines.f90:244-258) */
static void ed_c(const orc_params *p, int dx, const double *C, int i, int j, double e[9])
{
    for (int k = 0; k < 9; k++)
        e[k] = C[cidx(p, i + ex[k], j + ey[k])] - C[cidx(p, i - ex[k], j - ey[k])];
    for (int k = 0; k < 9; k++) e[k] = e[k] * 0.5 / (double)dx;
}

/* ED_M_AMR (LBM_Subroutines.f90:262-277) */
static void ed_m(const orc_params *p, int dx, const double *C, int i, int j, double e[9])
{
    for (int k = 0; k < 9; k++)
        e[k] = -C[cidx(p, i - ex[k], j - ey[k])] - 3.0 * C[cidx(p, i, j)] +
               5.0 * C[cidx(p, i + ex[k], j + ey[k])] - C[cidx(p, i + 2 * ex[k], j + 2 * ey[k])];
    for (int k = 0; k < 9; k++) e[k] = e[k] * 0.25 / (double)dx;
}

/* Gradient_ED_AMR (LBM_Subroutines.f90:281-289) */
static void gradient_ed(const double e[9], double *dx, double *dy)
{
    *dx = (e[1] - e[3]) / 3.0 + (e[5] - e[6] - e[7] + e[8]) / 12.0;
    *dy = (e[2] - e[4]) / 3.0 + (e[5] + e[6] - e[7] - e[8]) / 12.0;
}

/* Convert_D2Q9_standard (LBM_Subroutines.f90:365-380) */
static void convert_d2q9(const double *f, double *o)
{
    o[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    o[1] = -4.0 * f[0] - f[1] - f[2] - f[3] - f[4] + 2.0 * (f[5] + f[6] + f[7] + f[8]);
    o[2] = 4.0 * f[0] - 2.0 * (f[1] + f[2] + f[3] + f[4]) + f[5] + f[6] + f[7] + f[8];
    o[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    o[4] = -2.0 * (f[1] - f[3]) + f[5] - f[6] - f[7] + f[8];
    o[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    o[6] = -2.0 * (f[2] - f[4]) + f[5] + f[6] - f[7] - f[8];
    o[7] = f[1] - f[2] + f[3] - f[4];
    o[8] = f[5] - f[6] + f[7] - f[8];
}

/* Reconvt_D2Q9_standard (LBM_Subroutines.f90:384-401) */
static void reconvt_d2q9(const double *f, double *o)
{
    o[0] = 4.0 * (f[0] - f[1] + f[2]);
    o[1] = 4.0 * f[0] - f[1] - 2.0 * f[2] + 6.0 * (f[3] - f[4]) + 9.0 * f[7];
    o[2] = 4.0 * f[0] - f[1] - 2.0 * f[2] + 6.0 * (f[5] - f[6]) - 9.0 * f[7];
    o[3] = 4.0 * f[0] - f[1] - 2.0 * f[2] - 6.0 * (f[3] - f[4]) + 9.0 * f[7];
    o[4] = 4.0 * f[0] - f[1] - 2.0 * f[2] - 6.0 * (f[5] - f[6]) - 9.0 * f[7];
    o[5] = 4.0 * f[0] + 2.0 * f[1] + f[2] + 6.0 * f[3] + 3.0 * f[4] + 6.0 * f[5] + 3.0 * f[6] + 9.0 * f[8];
    o[6] = 4.0 * f[0] + 2.0 * f[1] + f[2] - 6.0 * f[3] - 3.0 * f[4] + 6.0 * f[5] + 3.0 * f[6] - 9.0 * f[8];
    o[7] = 4.0 * f[0] + 2.0 * f[1] + f[2] - 6.0 * f[3] - 3.0 * f[4] - 6.0 * f[5] - 3.0 * f[6] + 9.0 * f[8];
    o[8] = 4.0 * f[0] + 2.0 * f[1] + f[2] + 6.0 * f[3] + 3.0 * f[4] - 6.0 * f[5] - 3.0 * f[6] - 9.0 * f[8];
    for (int i = 0; i < 9; i++) o[i] = o[i] / 36.0;
}

/* MRT_Collision_AMR (LBM_Subroutines.f90:348-361) */
void orc_mrt_collision(double s9, const double geq[9], double g[9])
{
    double d[9], m[9], t[9];
    for (int i = 0; i < 9; i++) d[i] = g[i] - geq[i];
    convert_d2q9(d, m);
    for (int i = 0; i < 7; i++) t[i] = m[i] * 1.0;
    t[7] = m[7] * s9;
    t[8] = m[8] * s9;
    reconvt_d2q9(t, m);
    for (int i = 0; i < 9; i++) g[i] = g[i] - m[i];
}

/* the forcing term hlp of LBM_Subroutines.f90:154-155 / 164-165 */
static void forcing(const double Gamma[9], double dRho3, double mu, const double e[9], double Ux,
                    double Uy, double dcdx, double dcdy, double Fb, double hlp[9])
{
    for (int i = 0; i < 9; i++)
        hlp[i] = ((Gamma[i] - Wa[i]) * dRho3 + Gamma[i] * mu) * (e[i] - (Ux * dcdx + Uy * dcdy)) +
                 ((double)ey[i] - Uy) * Gamma[i] * Fb;
}

/* Collision_h_g_AMR (LBM_Subroutines.f90:116-189) */
void orc_collision_block(const orc_params *p, int dx, const double *C, const double *P,
                         const double *Ux, const double *Uy, double *h, double *g)
{
    lattice_init();
    double wc = 1.0 / (0.5 + p->M0 * 3.0);   /* D2Q9_Mod.f90:11 */
    double dRho3 = (p->rhoh - p->rhol) / 3.0;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            double e[9], hlp[9], Gamma[9], heq[9], geq[9], dcdx, dcdy, ni, nj;
            double mu = chem_pot(p, dx, C, i, j);
            double c = C[cidx(p, i, j)];
            double ux = Ux[pidx(p, i, j)], uy = Uy[pidx(p, i, j)], pp = P[pidx(p, i, j)];
            ed_c(p, dx, C, i, j, e);
            gradient_ed(e, &dcdx, &dcdy);
            normal(dcdx, dcdy, &ni, &nj);
            equilibrium(p, c, ux, uy, ni, nj, Gamma, heq);
            double *hc = &h[fidx(p, 0, i, j)];
            for (int k = 0; k < 9; k++) hc[k] = hc[k] * (1.0 - wc) + heq[k] * wc;
            double Rho = p->rhol + c * (p->rhoh - p->rhol);
            double Fb = (p->rhoh - Rho) * p->gy;
            forcing(Gamma, dRho3, mu, e, ux, uy, dcdx, dcdy, Fb, hlp);
            for (int k = 0; k < 9; k++)
                geq[k] = pp * Wa[k] + (Gamma[k] - Wa[k]) * Rho / 3.0 - 0.5 * hlp[k];
            ed_m(p, dx, C, i, j, e);
            gradient_ed(e, &dcdx, &dcdy);
            forcing(Gamma, dRho3, mu, e, ux, uy, dcdx, dcdy, Fb, hlp);
            double tau;
            if (c < 0.0) tau = p->taul + 0.5;
            else if (c > 1.0) tau = p->tauh + 0.5;
            else {
                tau = c * (1.0 / p->tauh - 1.0 / p->taul) + 1.0 / p->taul;
                tau = 1.0 / tau + 0.5;
            }
            double s9 = 1.0 / tau;
            double *gc = &g[fidx(p, 0, i, j)];
            orc_mrt_collision(s9, geq, gc);
            for (int k = 0; k < 9; k++) gc[k] = gc[k] + hlp[k];
        }
}

/* Macroscopic_Values_h_AMR (LBM_Subroutines.f90:193-201): C = SUM(h, dim 1) */
void orc_macro_h_block(const orc_params *p, const double *h, double *C)
{
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            const double *hc = &h[fidx(p, 0, i, j)];
            double s = 0.0;
            for (int k = 0; k < 9; k++) s += hc[k];
            C[cidx(p, i, j)] = s;
        }
}

/* Macroscopic_Values_g_AMR (LBM_Subroutines.f90:205-240) with Gradient_AMR (:293-312) */
void orc_macro_g_block(const orc_params *p, int dx, const double *g, const double *C, double *P,
                       double *Ux, double *Uy)
{
    lattice_init();
    double dRho6 = (p->rhoh - p->rhol) / 6.0;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
#define CC(a, b) C[cidx(p, a, b)]
            double dcdx = ((CC(i + 1, j) - CC(i - 1, j)) / 3.0 +
                           (CC(i + 1, j + 1) + CC(i + 1, j - 1) - CC(i - 1, j + 1) - CC(i - 1, j - 1)) / 12.0) /
                          (double)dx;
            double dcdy = ((CC(i, j + 1) - CC(i, j - 1)) / 3.0 +
                           (CC(i + 1, j + 1) + CC(i - 1, j + 1) - CC(i + 1, j - 1) - CC(i - 1, j - 1)) / 12.0) /
                          (double)dx;
            double mu = chem_pot(p, dx, C, i, j);
            double Rho = p->rhol + CC(i, j) * (p->rhoh - p->rhol);
#undef CC
            double Fb = (p->rhoh - Rho) * p->gy;
            double Fx = mu * dcdx;
            double Fy = mu * dcdy + Fb;
            const double *gc = &g[fidx(p, 0, i, j)];
            double sx = 0.0, sy = 0.0, s = 0.0;
            for (int k = 1; k < 9; k++) sx += gc[k] * (double)ex[k];
            for (int k = 1; k < 9; k++) sy += gc[k] * (double)ey[k];
            for (int k = 0; k < 9; k++) s += gc[k];
            double ux = (3.0 * sx + 0.5 * Fx) / Rho;
            double uy = (3.0 * sy + 0.5 * Fy) / Rho;
            Ux[pidx(p, i, j)] = ux;
            Uy[pidx(p, i, j)] = uy;
            P[pidx(p, i, j)] = s + dRho6 * (ux * dcdx + uy * dcdy);
        }
}

/* ======================================================================
 * Interpolations.f90 (general forms, used by closed-form tests and prolong/restrict)
 * ====================================================================== */

/* Biquadratic_from_Fine (Interpolations.f90:2-34); in(i2,j2) -> out(X2,Y2) */
void orc_biquadratic_from_fine(int i2, int j2, int X2, int Y2, const double *in, double *out)
{
    (void)j2;
    double *tmp = malloc(sizeof(double) * i2 * Y2);
    const int b = 6;
    for (int X = 1; X <= Y2; X++) {
        int J = (X + 1) % 2, i = 2 * X - 1 - J, a = 3 - 4 * J, c = 2 - a;
        for (int r = 0; r < i2; r++)
            tmp[r + i2 * (X - 1)] = (double)a * in[r + i2 * (i - 1)] + (double)b * in[r + i2 * i] +
                                    (double)c * in[r + i2 * (i + 1)];
    }
    for (int X = 1; X <= X2; X++) {
        int J = (X + 1) % 2, i = 2 * X - 1 - J, a = 3 - 4 * J, c = 2 - a;
        for (int s = 0; s < Y2; s++)
            out[(X - 1) + X2 * s] = (double)a * tmp[(i - 1) + i2 * s] + (double)b * tmp[i + i2 * s] +
                                    (double)c * tmp[(i + 1) + i2 * s];
    }
    for (int k = 0; k < X2 * Y2; k++) out[k] = out[k] / 64.0;
    free(tmp);
}

/* Biquadratic_from_Coarse (Interpolations.f90:74-104); in(I2,J2) -> out(x2,y2) */
void orc_biquadratic_from_coarse(int I2, int J2, int x2, int y2, const double *in, double *out)
{
    (void)J2;
    double *tmp = malloc(sizeof(double) * I2 * y2);
    const int b = 30;
    for (int x = 1; x <= y2; x++) {
        int I = (x + 1) / 2, a = 5 - 8 * ((x + 1) % 2), c = 2 - a;
        for (int r = 0; r < I2; r++)
            tmp[r + I2 * (x - 1)] = (double)a * in[r + I2 * (I - 1)] + (double)b * in[r + I2 * I] +
                                    (double)c * in[r + I2 * (I + 1)];
    }
    for (int x = 1; x <= x2; x++) {
        int I = (x + 1) / 2, a = 5 - 8 * ((x + 1) % 2), c = 2 - a;
        for (int s = 0; s < y2; s++)
            out[(x - 1) + x2 * s] = (double)a * tmp[(I - 1) + I2 * s] + (double)b * tmp[I + I2 * s] +
                                    (double)c * tmp[(I + 1) + I2 * s];
    }
    for (int k = 0; k < x2 * y2; k++) out[k] = out[k] / 1024.0;
    free(tmp);
}

/* Bilinear_from_Fine (Interpolations.f90:142-166); in(i2,j2) -> out(X2,Y2) */
void orc_bilinear_from_fine(int i2, int j2, int X2, int Y2, const double *in, double *out)
{
    (void)j2;
    double *tmp = malloc(sizeof(double) * i2 * Y2);
    for (int X = 1; X <= Y2; X++) {
        int i = 2 * X - 1;
        for (int r = 0; r < i2; r++)
            tmp[r + i2 * (X - 1)] = 0.5 * (in[r + i2 * (i - 1)] + in[r + i2 * i]);
    }
    for (int X = 1; X <= X2; X++) {
        int i = 2 * X - 1;
        for (int s = 0; s < Y2; s++)
            out[(X - 1) + X2 * s] = 0.5 * (tmp[(i - 1) + i2 * s] + tmp[i + i2 * s]);
    }
    free(tmp);
}

/* Bilinear_Interpolate_Extrapolate (Prolong.f90:43-70), scalar form:
 * in(nx,ny) -> out(2nx,2ny) */
void orc_bilinear_interp_extrap(int nx, int ny, const double *in, double *out)
{
    double *tmp = malloc(sizeof(double) * 2 * nx * ny);
    for (int i = 1; i <= 2 * nx; i++) {
        int X = i / 2;
        if (X > nx - 1) X = nx - 1;
        if (X < 1) X = 1;
        int a = 2 * i - 4 * X + 1, b = 4 - a;
        for (int s = 0; s < ny; s++)
            tmp[(i - 1) + 2 * nx * s] = (double)b * in[(X - 1) + nx * s] + (double)a * in[X + nx * s];
    }
    for (int i = 1; i <= 2 * ny; i++) {
        int X = i / 2;
        if (X > ny - 1) X = ny - 1;
        if (X < 1) X = 1;
        int a = 2 * i - 4 * X + 1, b = 4 - a;
        for (int r = 0; r < 2 * nx; r++)
            out[r + 2 * nx * (i - 1)] = (double)b * tmp[r + 2 * nx * (X - 1)] + (double)a * tmp[r + 2 * nx * X];
    }
    for (int k = 0; k < 4 * nx * ny; k++) out[k] = out[k] / 16.0;
    free(tmp);
}

/* ======================================================================
 * Mesh store (Parameters.f90) and neighbour pointers (Neighbors.f90)
 * ====================================================================== */

static int ptr_value(const orc_mesh *m, orc_ptr pt)
{
    if (pt.kind == 0) return pt.a;
    return m->blk[pt.a].child[pt.b - 1];
}

int orc_nghbr_value(const orc_mesh *m, int n, int k) { return ptr_value(m, m->blk[n].nghbr[k - 1]); }

static orc_ptr lit(int v)
{
    orc_ptr r = {0, v, 0};
    return r;
}
static orc_ptr slot(int b, int c)
{
    orc_ptr r = {1, b, c};
    return r;
}

orc_mesh *orc_create(const orc_params *p)
{
    lattice_init();
    orc_mesh *m = calloc(1, sizeof(orc_mesh));
    m->p = *p;
    m->blk = calloc((size_t)p->max_blocks + 1, sizeof(orc_block));
    m->leaf = calloc((size_t)p->max_blocks + 1, sizeof(int));
    m->freeq = calloc((size_t)p->max_blocks + 1, sizeof(int));
    return m;
}

static void free_arrays(orc_block *b)
{
    free(b->h); free(b->g); free(b->C); free(b->P); free(b->Ux); free(b->Uy);
    b->h = b->g = b->C = b->P = b->Ux = b->Uy = NULL;
}

static void alloc_arrays(const orc_params *p, orc_block *b)
{
    free_arrays(b);
    b->h = calloc(fsize(p), sizeof(double));
    b->g = calloc(fsize(p), sizeof(double));
    b->C = calloc(csize(p), sizeof(double));
    b->P = calloc(psize(p), sizeof(double));
    b->Ux = calloc(psize(p), sizeof(double));
    b->Uy = calloc(psize(p), sizeof(double));
}

void orc_destroy(orc_mesh *m)
{
    if (!m) return;
    for (int n = 1; n <= m->p.max_blocks; n++) free_arrays(&m->blk[n]);
    free(m->blk);
    free(m->leaf);
    free(m->freeq);
    free(m);
}

/* Neighbor (Neighbors.f90:19-37) */
static void neighbor(orc_mesh *m, int p, int n, int n2, int c)
{
    orc_ptr pp = m->blk[p].nghbr[n2 - 1];
    if (ptr_value(m, pp) < 0) m->blk[n].nghbr[n2 - 1] = pp;
    else m->blk[n].nghbr[n2 - 1] = slot(ptr_value(m, pp), c);
    int n1 = n2 + 1 - 2 * ((n2 - 1) % 2);
    m->blk[n].nghbr[n1 - 1] = slot(p, c);
}

/* Diagonal_Neighbor (Neighbors.f90:41-122); SW/NW/SE/NE are the wall codes -5..-8 */
static void diagonal_neighbor(orc_mesh *m, int p, int n, int c)
{
    orc_block *P = &m->blk[p], *N = &m->blk[n];
#define PV(k) ptr_value(m, P->nghbr[(k)-1])
    if (c == 1) {
        N->nghbr[4] = PV(5) < 0 ? P->nghbr[4] : slot(PV(5), 4);
        N->nghbr[5] = PV(1) < 0 ? lit(-6) : slot(PV(1), 4);
        N->nghbr[6] = PV(3) < 0 ? lit(-7) : slot(PV(3), 4);
        N->nghbr[7] = slot(p, 4);
    } else if (c == 2) {
        N->nghbr[4] = PV(3) < 0 ? lit(-5) : slot(PV(3), 3);
        N->nghbr[5] = slot(p, 3);
        N->nghbr[6] = PV(7) < 0 ? P->nghbr[6] : slot(PV(7), 3);
        N->nghbr[7] = PV(2) < 0 ? lit(-8) : slot(PV(2), 3);
    } else if (c == 3) {
        N->nghbr[4] = PV(1) < 0 ? lit(-5) : slot(PV(1), 2);
        N->nghbr[5] = PV(6) < 0 ? P->nghbr[5] : slot(PV(6), 2);
        N->nghbr[6] = slot(p, 2);
        N->nghbr[7] = PV(4) < 0 ? lit(-8) : slot(PV(4), 2);
    } else {
        N->nghbr[4] = slot(p, 1);
        N->nghbr[5] = PV(4) < 0 ? lit(-6) : slot(PV(4), 1);
        N->nghbr[6] = PV(2) < 0 ? lit(-7) : slot(PV(2), 1);
        N->nghbr[7] = PV(8) < 0 ? P->nghbr[7] : slot(PV(8), 1);
    }
#undef PV
}

/* Find_Neighbors (Neighbors.f90:2-15) */
static void find_neighbors(orc_mesh *m, int par, int n, int I, int J)
{
    int child = 1 + I + 2 * J;
    neighbor(m, par, n, 1 + I, child + 1 - 2 * I);
    neighbor(m, par, n, 3 + J, child + 2 - 4 * J);
    diagonal_neighbor(m, par, n, child);
}

/* Add_Child_Blocks (AMR_Rising_2D_OMP.f90:517-584) */
static int add_child_blocks(orc_mesh *m, int par, int freeblk)
{
    const orc_params *p = &m->p;
    int N = freeblk - 1;
    if (N + 4 > p->max_blocks) {
        m->overflow = 1;
        return -1;
    }
    for (int i = 1; i <= 4; i++) m->blk[par].child[i - 1] = N + i;
    m->blk[par].isleaf = 0;
    m->blk[par].hasleaf = 1;
    int grandpar = m->blk[par].parent;
    if (grandpar > 0) m->blk[grandpar].hasleaf = 0;
    for (int J = 0; J <= 1; J++)
        for (int I = 0; I <= 1; I++) {
            N = N + 1;
            orc_block *b = &m->blk[N];
            b->level = m->blk[par].level + 1;
            b->dx = m->blk[par].dx / 2;
            b->I = 2 * m->blk[par].I + I;
            b->J = 2 * m->blk[par].J + J;
            b->parent = par;
            b->has_child = 1;
            for (int k = 0; k < 4; k++) b->child[k] = N;
            find_neighbors(m, par, N, I, J);
            b->isleaf = 1;
            b->hasleaf = 0;
            b->refin = 0;
            b->deref = 0;
            alloc_arrays(p, b);
        }
    if (N > m->nblks) m->nblks = N;
    return 0;
}

/* Array_of_Leaf_Blocks (AMR_Rising_2D_OMP.f90:588-610) */
static void array_of_leaf_blocks(orc_mesh *m)
{
    int i = 0;
    for (int n = 1; n <= m->nblks; n++)
        if (m->blk[n].isleaf) m->leaf[++i] = n;
    m->nleaf = i;
}

/* Determine_Min_Max_Levels (AMR_Rising_2D_OMP.f90:434-452) */
void orc_determine_min_max_levels(orc_mesh *m)
{
    int lmin = 100, lmax = -100;
    for (int k = 1; k <= m->nleaf; k++) {
        int L = m->blk[m->leaf[k]].level;
        if (L < lmin) lmin = L;
        if (L > lmax) lmax = L;
    }
    m->levelmin = lmin;
    m->levelmax = lmax;
}

/* ======================================================================
 * Ghost fills: Fill_Ghost_Cells_hg.f90, Fill_Ghost_Cells_C.f90,
 * Fill_Ghost_Cells_C2.f90, with Interpolations.f90:2-138 and BC_Cell.f90
 * ====================================================================== */

enum { FILL_C1 = 0, FILL_C2 = 1, FILL_HG = 2 };

/* direction d (0-based) of the 8 neighbours: 1=L 2=R 3=B 4=T 5=LB 6=LT 7=RB 8=RT
 * (AMR_Rising_2D_OMP.f90:97-101) */
static const int dsx[8] = {-1, 1, 0, 0, -1, -1, 1, 1};
static const int dsy[8] = {0, 0, -1, 1, -1, 1, -1, 1};

/* one axis of Biquadratic_from_Fine as used by the directional fills: source
 * coordinates in the 2n-wide fine frame of the finer neighbour, listed in the
 * order the reference gathers them (Fill_Left_hg, Fill_Ghost_Cells_hg.f90:74-81:
 * normal axis from the interface outward, tangential axis upward). */
static void fine_axis(int s, int t, int n, int f[3], int w[3])
{
    int k = (s == 0) ? t : (s < 0 ? 1 - t : t - n);
    int m0;
    if (k & 1) { m0 = 2 * k - 1; w[0] = 3; w[1] = 6; w[2] = -1; }
    else       { m0 = 2 * k - 2; w[0] = -1; w[1] = 6; w[2] = 3; }
    for (int a = 0; a < 3; a++) f[a] = (s < 0) ? 2 * n + 1 - (m0 + a) : (m0 + a);
}

/* one axis of Biquadratic_from_Coarse: source coordinates in the coarser
 * neighbour's frame, in the reference's gather order (Fill_Left_hg:91-96,
 * Fill_Left_Bottom_hg:311-315); pos = the block's position in its parent */
static void coarse_axis(int s, int t, int pos, int n, int c[3], int w[3], int *side)
{
    int fp = pos * n + t;
    int cp = (fp + 1) >> 1;
    int sn = (cp < 1) ? -1 : (cp > n ? 1 : 0);
    int cc = cp - sn * n;
    int upper = ((fp & 1) == 0);
    *side = sn;
    if (s < 0) {  /* descending: first entry lies toward the block being filled */
        c[0] = cc + 1; c[1] = cc; c[2] = cc - 1;
        if (upper) { w[0] = 5; w[1] = 30; w[2] = -3; } else { w[0] = -3; w[1] = 30; w[2] = 5; }
    } else {
        c[0] = cc - 1; c[1] = cc; c[2] = cc + 1;
        if (upper) { w[0] = -3; w[1] = 30; w[2] = 5; } else { w[0] = 5; w[1] = 30; w[2] = -3; }
    }
}

/* value accessor: component q of field in block b at (x,y) */
static inline double fval(const orc_mesh *m, int b, int kind, int q, int x, int y)
{
    const orc_params *p = &m->p;
    const orc_block *B = &m->blk[b];
    if (kind == FILL_HG) return (q < 9 ? B->h : B->g)[fidx(p, q % 9, x, y)];
    return B->C[cidx(p, x, y)];
}
static inline void fset(orc_mesh *m, int b, int kind, int q, int x, int y, double v)
{
    const orc_params *p = &m->p;
    orc_block *B = &m->blk[b];
    if (kind == FILL_HG) (q < 9 ? B->h : B->g)[fidx(p, q % 9, x, y)] = v;
    else B->C[cidx(p, x, y)] = v;
}

static void fill_direction(orc_mesh *m, int n, int L, int d, int N1, int kind)
{
    const orc_params *p = &m->p;
    int R = (kind == FILL_C2) ? 2 : 1;
    int ncomp = (kind == FILL_HG) ? 18 : 1;
    int sx = dsx[d], sy = dsy[d];
    int x0 = sx < 0 ? 1 - R : (sx > 0 ? NX + 1 : 1), x1 = sx < 0 ? 0 : (sx > 0 ? NX + R : NX);
    int y0 = sy < 0 ? 1 - R : (sy > 0 ? NY + 1 : 1), y1 = sy < 0 ? 0 : (sy > 0 ? NY + R : NY);
    int Lnb = m->blk[N1].level;
    if (L == Lnb) {
        if (m->blk[N1].isleaf) {
            /* case (i): copy of the neighbour's edge cells (Fill_Left_hg:66-70) */
            for (int y = y0; y <= y1; y++)
                for (int x = x0; x <= x1; x++)
                    for (int q = 0; q < ncomp; q++)
                        fset(m, n, kind, q, x, y, fval(m, N1, kind, q, x - sx * NX, y - sy * NY));
        } else {
            /* case (ii): biquadratic from the neighbour's children (Fill_Left_hg:72-85) */
            int c1 = m->blk[N1].child[0];
            for (int y = y0; y <= y1; y++)
                for (int x = x0; x <= x1; x++) {
                    int fx[3], wx[3], fy[3], wy[3];
                    fine_axis(sx, x, NX, fx, wx);
                    fine_axis(sy, y, NY, fy, wy);
                    for (int q = 0; q < ncomp; q++) {
                        double t[3];
                        for (int a = 0; a < 3; a++) {
                            int ix = fx[a] > NX, lx = fx[a] - ix * NX;
                            double v[3];
                            for (int b = 0; b < 3; b++) {
                                int iy = fy[b] > NY, ly = fy[b] - iy * NY;
                                v[b] = fval(m, c1 + ix + 2 * iy, kind, q, lx, ly);
                            }
                            t[a] = (double)wy[0] * v[0] + (double)wy[1] * v[1] + (double)wy[2] * v[2];
                        }
                        double o = (double)wx[0] * t[0] + (double)wx[1] * t[1] + (double)wx[2] * t[2];
                        fset(m, n, kind, q, x, y, o / 64.0);
                    }
                }
        }
    } else if (L == Lnb + 1) {
        /* case (iii): biquadratic from the coarser neighbour, its ghosts included
         * (Fill_Left_hg:89-99) */
        int px = m->blk[n].I % 2, py = m->blk[n].J % 2;
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) {
                int cx[3], wx[3], cy[3], wy[3], sidex, sidey;
                coarse_axis(sx, x, px, NX, cx, wx, &sidex);
                coarse_axis(sy, y, py, NY, cy, wy, &sidey);
                for (int q = 0; q < ncomp; q++) {
                    double t[3];
                    for (int a = 0; a < 3; a++)
                        t[a] = (double)wy[0] * fval(m, N1, kind, q, cx[a], cy[0]) +
                               (double)wy[1] * fval(m, N1, kind, q, cx[a], cy[1]) +
                               (double)wy[2] * fval(m, N1, kind, q, cx[a], cy[2]);
                    double o = (double)wx[0] * t[0] + (double)wx[1] * t[1] + (double)wx[2] * t[2];
                    fset(m, n, kind, q, x, y, o / 1024.0);
                }
            }
    } else {
        /* the reference prints an error and continues (Fill_Left_hg:101-103) */
        m->level_errors++;
    }
}

/* Bounce_Back_3 / Bounce_Back_4 (BC_Cell.f90:2-42) */
static void bounce_back(const orc_params *p, double *f, int top)
{
    int Y = top ? NY + 1 : 0, Yi = top ? NY : 1;
    for (int X = 0; X <= NX + 1; X++) {
        f[fidx(p, 0, X, Y)] = f[fidx(p, 0, X, Yi)];
        for (int i = 2; i <= 9; i++) {
            int e = ex[i - 1] * ey[i - 1];
            int Xn = top ? X - e : X + e;
            if (Xn < 0) Xn = 0;
            if (Xn > NX + 1) Xn = NX + 1;
            int J = (i - 2) + 4 * ((i / 2) % 2);
            f[fidx(p, i - 1, X, Y)] = f[fidx(p, J - 1, Xn, Yi)];
        }
    }
}

/* Symmetric_C_3/4 and Symmetric_C2_3/4 (BC_Cell.f90:46-104) */
static void symmetric_c(const orc_params *p, double *C, int top, int two)
{
    if (!two) {
        for (int X = 0; X <= NX + 1; X++) {
            if (!top) C[cidx(p, X, 0)] = C[cidx(p, X, 1)];
            else C[cidx(p, X, NY + 1)] = C[cidx(p, X, NY)];
        }
    } else {
        for (int X = -1; X <= NX + 2; X++) {
            if (!top) {
                C[cidx(p, X, -1)] = C[cidx(p, X, 2)];
                C[cidx(p, X, 0)] = C[cidx(p, X, 1)];
            } else {
                C[cidx(p, X, NY + 1)] = C[cidx(p, X, NY)];
                C[cidx(p, X, NY + 2)] = C[cidx(p, X, NY - 1)];
            }
        }
    }
}

/* Fill_Ghost_Cells_{C,C2,hg} dispatcher (Fill_Ghost_Cells_hg.f90:2-49) */
static void fill_ghost_cells(orc_mesh *m, int kind)
{
    const orc_params *p = &m->p;
    for (int L = m->levelmin; L <= m->levelmax; L++)
        for (int k = 1; k <= m->nleaf; k++) {
            int n = m->leaf[k];
            if (m->blk[n].level != L) continue;
            int nb[8];
            for (int d = 0; d < 8; d++) nb[d] = orc_nghbr_value(m, n, d + 1);
            for (int d = 0; d < 8; d++)
                if (nb[d] > 0) fill_direction(m, n, L, d, nb[d], kind);
            for (int top = 0; top <= 1; top++) {
                if (nb[2 + top] >= 0) continue;
                if (kind == FILL_HG) {
                    bounce_back(p, m->blk[n].h, top);
                    bounce_back(p, m->blk[n].g, top);
                } else {
                    symmetric_c(p, m->blk[n].C, top, kind == FILL_C2);
                }
            }
        }
}

void orc_fill_ghost_cells_c2(orc_mesh *m) { fill_ghost_cells(m, FILL_C2); }
void orc_fill_ghost_cells_c(orc_mesh *m) { fill_ghost_cells(m, FILL_C1); }
void orc_fill_ghost_cells_hg(orc_mesh *m) { fill_ghost_cells(m, FILL_HG); }

/* ======================================================================
 * Per-step loops (AMR_Rising_2D_OMP.f90:328-354, 391-425)
 * ====================================================================== */

void orc_collide_all(orc_mesh *m)
{
    for (int k = 1; k <= m->nleaf; k++) {
        orc_block *b = &m->blk[m->leaf[k]];
        orc_collision_block(&m->p, b->dx, b->C, b->P, b->Ux, b->Uy, b->h, b->g);
    }
}

void orc_stream_all(orc_mesh *m)
{
    for (int k = 1; k <= m->nleaf; k++) {
        orc_block *b = &m->blk[m->leaf[k]];
        double cfl = 1.0 / b->dx;
        if (cfl > 0.9) {
            orc_perfect_shift_block(&m->p, b->h);
            orc_perfect_shift_block(&m->p, b->g);
        } else {
            orc_lax_wendroff_block(&m->p, cfl, b->h);
            orc_lax_wendroff_block(&m->p, cfl, b->g);
        }
    }
}

void orc_macroscopic_h_all(orc_mesh *m)
{
    for (int k = 1; k <= m->nleaf; k++) {
        orc_block *b = &m->blk[m->leaf[k]];
        orc_macro_h_block(&m->p, b->h, b->C);
    }
}

void orc_macroscopic_g_all(orc_mesh *m)
{
    for (int k = 1; k <= m->nleaf; k++) {
        orc_block *b = &m->blk[m->leaf[k]];
        orc_macro_g_block(&m->p, b->dx, b->g, b->C, b->P, b->Ux, b->Uy);
    }
}

/* ======================================================================
 * Regrid: Refine_Derefine.f90, Prolong.f90, Restrict.f90
 * ====================================================================== */

/* Err_Interface (Refine_Derefine.f90:327-352) */
double orc_err_interface(const orc_params *p, int dx, const double *C)
{
    double err = 0.0;
    for (int J = 1; J <= NY; J++)
        for (int I = 1; I <= NX; I++) {
            double a = C[cidx(p, I + 1, J)] - C[cidx(p, I - 1, J)];
            double b = C[cidx(p, I, J + 1)] - C[cidx(p, I, J - 1)];
            double DC = a * a + b * b;
            if (DC > err) err = DC;   /* MAX(err,DC) */
        }
    return sqrt(err) / (2.0 * dx);
}

static void enforce_1_level_jump(orc_mesh *m, int n);

/* Check_Level_Jump (Refine_Derefine.f90:74-93) */
static void check_level_jump(orc_mesh *m, int n, int I)
{
    int b = orc_nghbr_value(m, n, I);
    if (b > 0 && !m->blk[b].refin && m->blk[b].level == m->blk[n].level - 1 && m->blk[b].isleaf) {
        m->blk[b].refin = 1;
        enforce_1_level_jump(m, b);
    }
}

/* Enforce_1_Level_Jump (Refine_Derefine.f90:58-70) */
static void enforce_1_level_jump(orc_mesh *m, int n)
{
    for (int I = 1; I <= 8; I++) check_level_jump(m, n, I);
}

/* Flag_for_refinement (Refine_Derefine.f90:18-54) */
static void flag_for_refinement(orc_mesh *m, double tol)
{
    const orc_params *p = &m->p;
    for (int n = 1; n <= m->nblks; n++) {
        orc_block *b = &m->blk[n];
        b->refin = 0;
        if (b->isleaf) {
            if (b->level < p->min_levels) b->refin = 1;
            else if (b->level < p->max_levels) {
                if (orc_err_interface(p, b->dx, b->C) > tol) b->refin = 1;
            }
        }
    }
    for (int n = 1; n <= m->nblks; n++)
        if (m->blk[n].refin) enforce_1_level_jump(m, n);
}

/* Check_Level_Jump_Deref_2D (Refine_Derefine.f90:169-194) */
static void check_level_jump_deref(orc_mesh *m, int n, int nbdir, int c1, int c2)
{
    if (!m->blk[n].deref) return;
    int b = orc_nghbr_value(m, n, nbdir);
    if (b > 0 && !m->blk[b].isleaf) {
        int c = m->blk[b].child[c1 - 1];
        for (int I = 1; I <= 2; I++) {
            if (m->blk[c].refin || (!m->blk[c].isleaf && !m->blk[c].deref)) {
                m->blk[n].deref = 0;
                return;
            }
            c = m->blk[b].child[c2 - 1];
        }
    }
}

/* enforce_1_Level_jump_deref_2D (Refine_Derefine.f90:149-165) */
static void enforce_deref(orc_mesh *m, int n)
{
    check_level_jump_deref(m, n, 1, 2, 4);
    check_level_jump_deref(m, n, 2, 1, 3);
    check_level_jump_deref(m, n, 3, 3, 4);
    check_level_jump_deref(m, n, 4, 1, 2);
    check_level_jump_deref(m, n, 5, 4, 4);
    check_level_jump_deref(m, n, 6, 2, 2);
    check_level_jump_deref(m, n, 7, 3, 3);
    check_level_jump_deref(m, n, 8, 1, 1);
}

/* Flag_for_derefinement (Refine_Derefine.f90:97-145) */
static void flag_for_derefinement(orc_mesh *m, double tol)
{
    const orc_params *p = &m->p;
    for (int n = 1; n <= m->nblks; n++) {
        orc_block *b = &m->blk[n];
        b->deref = 0;
        if (!b->hasleaf) continue;
        if (b->level < p->min_levels) continue;
        int c1 = b->child[0];
        int any = 0;
        for (int c = c1; c <= c1 + 3; c++) any |= m->blk[c].refin;
        if (any) continue;
        b->deref = 1;
        for (int c = c1; c <= c1 + 3; c++)
            if (orc_err_interface(p, m->blk[c].dx, m->blk[c].C) > tol) {
                b->deref = 0;
                break;
            }
    }
    for (int L = m->levelmax - 1; L >= m->levelmin - 1; L--)
        for (int n = 1; n <= m->nblks; n++)
            if (m->blk[n].level == L && m->blk[n].deref) enforce_deref(m, n);
}

/* Restrict_Bilinear (Restrict.f90:14-65) */
static void restrict_block(orc_mesh *m, int n)
{
    const orc_params *p = &m->p;
    orc_block *B = &m->blk[n];
    alloc_arrays(p, B);   /* zeroed, ghosts included (Restrict.f90:25-34) */
    int W2 = 2 * NX, H2 = 2 * NY;
    double *fin = malloc(sizeof(double) * W2 * H2);
    double *fout = malloc(sizeof(double) * NX * NY);
    int child = B->child[0];
    /* h and g: Bilinear_from_Fine_f, component by component gives the same sums */
    for (int field = 0; field < 6; field++) {
        int ncomp = field < 2 ? 9 : 1;
        for (int q = 0; q < ncomp; q++) {
            for (int cj = 0; cj <= 1; cj++)
                for (int ci = 0; ci <= 1; ci++) {
                    orc_block *c = &m->blk[child + ci + 2 * cj];
                    for (int j = 1; j <= NY; j++)
                        for (int i = 1; i <= NX; i++) {
                            double v;
                            switch (field) {
                            case 0: v = c->h[fidx(p, q, i, j)]; break;
                            case 1: v = c->g[fidx(p, q, i, j)]; break;
                            case 2: v = c->C[cidx(p, i, j)]; break;
                            case 3: v = c->P[pidx(p, i, j)]; break;
                            case 4: v = c->Ux[pidx(p, i, j)]; break;
                            default: v = c->Uy[pidx(p, i, j)]; break;
                            }
                            fin[(i - 1 + ci * NX) + W2 * (j - 1 + cj * NY)] = v;
                        }
                }
            orc_bilinear_from_fine(W2, H2, NX, NY, fin, fout);
            for (int j = 1; j <= NY; j++)
                for (int i = 1; i <= NX; i++) {
                    double v = fout[(i - 1) + NX * (j - 1)];
                    switch (field) {
                    case 0: B->h[fidx(p, q, i, j)] = v; break;
                    case 1: B->g[fidx(p, q, i, j)] = v; break;
                    case 2: B->C[cidx(p, i, j)] = v; break;
                    case 3: B->P[pidx(p, i, j)] = v; break;
                    case 4: B->Ux[pidx(p, i, j)] = v; break;
                    default: B->Uy[pidx(p, i, j)] = v; break;
                    }
                }
        }
    }
    free(fin);
    free(fout);
}

/* Remove_Children (Refine_Derefine.f90:215-258) */
static void remove_children(orc_mesh *m, int par)
{
    int c1 = m->blk[par].child[0], c4 = c1 + 3;
    for (int c = c1; c <= c4; c++) {
        m->blk[c].level = -1;
        m->blk[c].isleaf = 0;
    }
    for (int k = 0; k < 4; k++) m->blk[par].child[k] = par;
    m->blk[par].isleaf = 1;
    m->blk[par].hasleaf = 0;
    int gp = m->blk[par].parent;
    if (gp > 0) {
        int s1 = m->blk[gp].child[0], s4 = m->blk[gp].child[3], all = 1;
        for (int s = s1; s <= s4; s++) all &= m->blk[s].isleaf;
        if (all) m->blk[gp].hasleaf = 1;
    }
    for (int c = c1; c <= c4; c++) {
        free_arrays(&m->blk[c]);
        m->blk[c].has_child = 0;
    }
    m->freeq[m->ftail++] = c1;
    m->kk += 4;
}

/* Derefine (Refine_Derefine.f90:198-211) */
static void derefine(orc_mesh *m)
{
    for (int n = 1; n <= m->nblks; n++)
        if (m->blk[n].deref) {
            restrict_block(m, n);
            remove_children(m, n);
        }
}

/* Prolong_Bilinear (Prolong.f90:12-39), Bilinear_Interpolation (:74-101),
 * Fill_Children_Data (:105-130) */
static void prolong(orc_mesh *m, int n, int c)
{
    const orc_params *p = &m->p;
    orc_block *B = &m->blk[n];
    int W2 = 2 * NX, H2 = 2 * NY;
    double *in = malloc(sizeof(double) * NX * NY);
    double *out = malloc(sizeof(double) * W2 * H2);
    for (int q = 0; q < 18; q++) {
        double *f = q < 9 ? B->h : B->g;
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) in[(i - 1) + NX * (j - 1)] = f[fidx(p, q % 9, i, j)];
        orc_bilinear_interp_extrap(NX, NY, in, out);
        for (int cj = 0; cj <= 1; cj++)
            for (int ci = 0; ci <= 1; ci++) {
                orc_block *C = &m->blk[c + ci + 2 * cj];
                double *fc = q < 9 ? C->h : C->g;
                for (int J = 1; J <= NY; J++)
                    for (int I = 1; I <= NX; I++)
                        fc[fidx(p, q % 9, I, J)] = out[(I - 1 + ci * NX) + W2 * (J - 1 + cj * NY)];
            }
    }
    /* C with one ghost ring: Bilinear_Interpolation (0:nx+1)^2 -> (0:2nx+1)^2 */
    int cw = NX + 2, fw = 2 * NX + 2;
    double *cc = malloc(sizeof(double) * cw * (NY + 2));
    double *tmp = malloc(sizeof(double) * fw * (NY + 2));
    double *cf = malloc(sizeof(double) * fw * (2 * NY + 2));
    for (int j = 0; j <= NY + 1; j++)
        for (int i = 0; i <= NX + 1; i++) cc[i + cw * j] = B->C[cidx(p, i, j)];
    for (int i = 0; i <= 2 * NX + 1; i++) {
        int X = i / 2, a = 2 * i - 4 * X + 1, b = 4 - a;
        for (int s = 0; s <= NY + 1; s++)
            tmp[i + fw * s] = (double)b * cc[X + cw * s] + (double)a * cc[(X + 1) + cw * s];
    }
    for (int i = 0; i <= 2 * NY + 1; i++) {
        int X = i / 2, a = 2 * i - 4 * X + 1, b = 4 - a;
        for (int r = 0; r <= 2 * NX + 1; r++)
            cf[r + fw * i] = (double)b * tmp[r + fw * X] + (double)a * tmp[r + fw * (X + 1)];
    }
    for (int k = 0; k < fw * (2 * NY + 2); k++) cf[k] = cf[k] / 16.0;
    for (int cj = 0; cj <= 1; cj++)
        for (int ci = 0; ci <= 1; ci++) {
            orc_block *C = &m->blk[c + ci + 2 * cj];
            for (int J = 0; J <= NY + 1; J++)
                for (int I = 0; I <= NX + 1; I++) C->C[cidx(p, I, J)] = cf[(I + ci * NX) + fw * (J + cj * NY)];
        }
    free_arrays(B);   /* Prolong.f90:26 */
    for (int k = 0; k < 4; k++) {
        orc_block *C = &m->blk[c + k];
        orc_macro_g_block(p, C->dx, C->g, C->C, C->P, C->Ux, C->Uy);   /* Prolong.f90:127 */
    }
    free(in); free(out); free(cc); free(tmp); free(cf);
}

/* Refine (Refine_Derefine.f90:262-294) */
static void refine(orc_mesh *m)
{
    for (int L = m->levelmin; L <= m->levelmax; L++) {
        int nend = m->nblks;   /* DO n = 1, Nblks: bound fixed at loop entry */
        for (int n = 1; n <= nend; n++) {
            if (m->blk[n].level != L || !m->blk[n].refin) continue;
            int c1;
            if (m->fhead < m->ftail) {
                c1 = m->freeq[m->fhead];
                if (add_child_blocks(m, n, c1)) return;
                m->kk -= 4;
                m->fhead++;
            } else {
                c1 = m->nblks + 1;
                if (add_child_blocks(m, n, c1)) return;
            }
            prolong(m, n, c1);
        }
    }
}

/* Refine_Derefine_Mesh (Refine_Derefine.f90:2-14) */
void orc_refine_derefine_mesh(orc_mesh *m, double refin_tol, double deref_tol)
{
    flag_for_refinement(m, refin_tol);
    flag_for_derefinement(m, deref_tol);
    derefine(m);
    refine(m);
    array_of_leaf_blocks(m);
}

/* ======================================================================
 * Set-up (AMR_Rising_2D_OMP.f90:43-231, 493-513)
 * ====================================================================== */

/* Initialize_Base_Block (AMR_Rising_2D_OMP.f90:79-115) */
static void init_base_block(orc_mesh *m, int N)
{
    orc_block *b = &m->blk[N];
    m->nblks = N;
    b->level = 0;
    b->dx = m->p.L0 / m->p.nx;
    b->I = 0;
    b->J = N - 1;
    b->has_child = 1;
    for (int k = 0; k < 4; k++) b->child[k] = -1;
    b->parent = -1;
    for (int I = 1; I <= 8; I++) b->nghbr[I - 1] = lit(-I);
    b->nghbr[0] = lit(N);
    b->nghbr[1] = lit(N);
    b->isleaf = 1;
    b->hasleaf = 0;
    b->refin = 0;
    b->deref = 0;
}

/* the analytic phase field of Initialize_AMR_Rising_Bubble
 * (AMR_Rising_2D_OMP.f90:174-185), both ghost rings included */
static void bubble_c(orc_mesh *m, orc_block *B, double R, double X0, double Y0)
{
    const orc_params *p = &m->p;
    int dx = B->dx;
    for (int j = -1; j <= NY + 2; j++)
        for (int i = -1; i <= NX + 2; i++) {
            double Xc = (double)(B->I * dx * NX) + (double)dx * ((double)i - 0.5);
            double Yc = (double)(B->J * dx * NY) + (double)dx * ((double)j - 0.5);
            double Ri = sqrt((Xc - X0) * (Xc - X0) + (Yc - Y0) * (Yc - Y0));
            B->C[cidx(p, i, j)] = 0.5 + 0.5 * tanh(2.0 * (Ri - R) / p->W);
        }
}

void orc_set_bubble_c(orc_mesh *m, double R, double X0, double Y0)
{
    for (int k = 1; k <= m->nleaf; k++) bubble_c(m, &m->blk[m->leaf[k]], R, X0, Y0);
}

/* Initialize_AMR_Rising_Bubble (AMR_Rising_2D_OMP.f90:142-231) */
static void init_bubble(orc_mesh *m, double R, double X0, double Y0)
{
    const orc_params *p = &m->p;
    double dRho3 = (p->rhoh - p->rhol) / 3.0;
    for (int k = 1; k <= m->nleaf; k++) {
        orc_block *B = &m->blk[m->leaf[k]];
        int dx = B->dx;
        bubble_c(m, B, R, X0, Y0);
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                B->Ux[pidx(p, i, j)] = 0.0;
                B->Uy[pidx(p, i, j)] = 0.0;
                B->P[pidx(p, i, j)] = -(p->sigma / R * B->C[cidx(p, i, j)]);
            }
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                double e[9], hlp[9], Gamma[9], dcdx, dcdy, ni, nj;
                double mu = chem_pot(p, dx, B->C, i, j);
                ed_c(p, dx, B->C, i, j, e);
                gradient_ed(e, &dcdx, &dcdy);
                normal(dcdx, dcdy, &ni, &nj);
                double c = B->C[cidx(p, i, j)];
                double Rho = p->rhol + c * (p->rhoh - p->rhol);
                double Fb = 0.0;
                double ux = B->Ux[pidx(p, i, j)], uy = B->Uy[pidx(p, i, j)], pp = B->P[pidx(p, i, j)];
                equilibrium(p, c, ux, uy, ni, nj, Gamma, &B->h[fidx(p, 0, i, j)]);
                forcing(Gamma, dRho3, mu, e, ux, uy, dcdx, dcdy, Fb, hlp);
                for (int q = 0; q < 9; q++)
                    B->g[fidx(p, q, i, j)] = pp * Wa[q] + (Gamma[q] - Wa[q]) * Rho / 3.0 - 0.5 * hlp[q];
            }
    }
}

int orc_setup_rising_bubble(orc_mesh *m, double R, double X0, double Y0)
{
    const orc_params *p = &m->p;
    m->levelmin = p->min_levels;
    m->levelmax = p->min_levels;
    for (int N = 1; N <= p->nroots; N++) init_base_block(m, N);
    /* roots stacked in y (AMR_Rising_2D_OMP.f90:119-138 for three roots) */
    for (int N = 1; N <= p->nroots; N++) {
        if (N > 1) {
            m->blk[N].nghbr[2] = lit(N - 1);
            m->blk[N].nghbr[4] = lit(N - 1);
            m->blk[N].nghbr[6] = lit(N - 1);
        }
        if (N < p->nroots) {
            m->blk[N].nghbr[3] = lit(N + 1);
            m->blk[N].nghbr[5] = lit(N + 1);
            m->blk[N].nghbr[7] = lit(N + 1);
        }
    }
    /* Initial_Mesh (AMR_Rising_2D_OMP.f90:493-513) */
    for (int L = 1; L <= m->levelmin; L++) {
        int nend = m->nblks;
        for (int n = 1; n <= nend; n++)
            if (m->blk[n].level == L - 1)
                if (add_child_blocks(m, n, m->nblks + 1)) return -1;
    }
    array_of_leaf_blocks(m);
    init_bubble(m, R, X0, Y0);
    while (m->levelmax < p->max_levels) {
        orc_refine_derefine_mesh(m, 0.002, 0.001);
        if (m->overflow) return -1;
        init_bubble(m, R, X0, Y0);
        m->levelmax++;
    }
    m->T = 0;
    orc_determine_min_max_levels(m);
    return 0;
}

void orc_step(orc_mesh *m, int regrid_every)
{
    orc_determine_min_max_levels(m);
    orc_fill_ghost_cells_c2(m);
    orc_collide_all(m);
    orc_fill_ghost_cells_hg(m);
    orc_stream_all(m);
    orc_macroscopic_h_all(m);
    orc_fill_ghost_cells_c(m);
    orc_macroscopic_g_all(m);
    if ((m->T + 1) % regrid_every == 0) orc_refine_derefine_mesh(m, 0.002, 0.001);
    m->T++;
    orc_determine_min_max_levels(m);
}

void orc_set_step(orc_mesh *m, int T) { m->T = T; }

void orc_run(orc_mesh *m, int nsteps, int regrid_every)
{
    for (int s = 0; s < nsteps; s++) orc_step(m, regrid_every);
}

void orc_diagnostics(const orc_mesh *m, double out[9])
{
    const orc_params *p = &m->p;
    double rmin = 10, rmax = -10, umax = 0, mass = 0, vol = 0, xw = 0, yw = 0, vw = 0, bad = 0;
    for (int k = 1; k <= m->nleaf; k++) {
        const orc_block *B = &m->blk[m->leaf[k]];
        int dx = B->dx;
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                double c = B->C[cidx(p, i, j)];
                double Rho = p->rhol + c * (p->rhoh - p->rhol);
                if (Rho < rmin) rmin = Rho;
                if (Rho > rmax) rmax = Rho;
                mass = mass + Rho * (double)(dx * dx);
                double w = (1.0 - c) * (double)(dx * dx);
                double xc = (double)(B->I * dx * NX) + (double)dx * ((double)i - 0.5);
                double yc = (double)(B->J * dx * NY) + (double)dx * ((double)j - 0.5);
                vol += w;
                xw += w * xc;
                yw += w * yc;
                vw += w * B->Uy[pidx(p, i, j)];
                if (!(fabs(c) <= 1.79e308)) bad += 1;
                double u = fabs(B->Uy[pidx(p, i, j)]);
                if (u > umax) umax = u;
            }
    }
    out[0] = rmin; out[1] = rmax; out[2] = umax; out[3] = mass; out[4] = vol;
    out[5] = xw / vol; out[6] = yw / vol; out[7] = vw / vol; out[8] = bad;
}

/* ======================================================================
 * Snapshots in the reference harness's layout (oracle/ref_harness.f90)
 * ====================================================================== */

int orc_save_snapshot(const orc_mesh *m, const char *path)
{
    const orc_params *p = &m->p;
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    int hdr[16] = {0};
    int nfree = m->ftail - m->fhead;
    hdr[0] = 1095586380; hdr[1] = 1; hdr[2] = NX; hdr[3] = NY; hdr[4] = p->L0;
    hdr[5] = p->min_levels; hdr[6] = p->max_levels; hdr[7] = m->nblks; hdr[8] = m->nleaf;
    hdr[9] = m->T; hdr[10] = m->levelmin; hdr[11] = m->levelmax; hdr[12] = m->kk; hdr[13] = nfree;
    hdr[14] = p->max_blocks;
    fwrite(hdr, sizeof(int), 16, fp);
    if (nfree > 0) fwrite(&m->freeq[m->fhead], sizeof(int), nfree, fp);
    for (int n = 1; n <= m->nblks; n++) {
        const orc_block *b = &m->blk[n];
        int iv[22] = {0};
        iv[0] = b->level; iv[1] = b->dx; iv[2] = b->I; iv[3] = b->J; iv[4] = b->parent;
        iv[5] = b->isleaf; iv[6] = b->hasleaf; iv[7] = b->refin; iv[8] = b->deref;
        if (b->level >= 0) {
            if (b->has_child) for (int k = 0; k < 4; k++) iv[9 + k] = b->child[k];
            for (int k = 1; k <= 8; k++) iv[12 + k] = orc_nghbr_value(m, n, k);
        }
        iv[21] = b->h != NULL;
        fwrite(iv, sizeof(int), 22, fp);
        if (b->h) {
            fwrite(b->h, sizeof(double), fsize(p), fp);
            fwrite(b->g, sizeof(double), fsize(p), fp);
            fwrite(b->C, sizeof(double), csize(p), fp);
            fwrite(b->P, sizeof(double), psize(p), fp);
            fwrite(b->Ux, sizeof(double), psize(p), fp);
            fwrite(b->Uy, sizeof(double), psize(p), fp);
        }
    }
    fclose(fp);
    return 0;
}

int orc_load_snapshot(orc_mesh *m, const char *path)
{
    const orc_params *p = &m->p;
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    int hdr[16];
    if (fread(hdr, sizeof(int), 16, fp) != 16 || hdr[0] != 1095586380 || hdr[2] != NX || hdr[3] != NY) {
        fclose(fp);
        return -2;
    }
    for (int n = 1; n <= p->max_blocks; n++) {
        free_arrays(&m->blk[n]);
        memset(&m->blk[n], 0, sizeof(orc_block));
    }
    m->nblks = hdr[7]; m->nleaf = hdr[8]; m->T = hdr[9]; m->levelmin = hdr[10];
    m->levelmax = hdr[11]; m->kk = hdr[12];
    int nfree = hdr[13];
    m->fhead = 0; m->ftail = nfree;
    if (nfree > 0 && fread(m->freeq, sizeof(int), nfree, fp) != (size_t)nfree) { fclose(fp); return -3; }
    int (*nbv)[8] = calloc((size_t)m->nblks + 1, sizeof(*nbv));
    for (int n = 1; n <= m->nblks; n++) {
        int iv[22];
        if (fread(iv, sizeof(int), 22, fp) != 22) { free(nbv); fclose(fp); return -3; }
        orc_block *b = &m->blk[n];
        b->level = iv[0]; b->dx = iv[1]; b->I = iv[2]; b->J = iv[3]; b->parent = iv[4];
        b->isleaf = iv[5]; b->hasleaf = iv[6]; b->refin = iv[7]; b->deref = iv[8];
        b->has_child = b->level >= 0;
        for (int k = 0; k < 4; k++) b->child[k] = iv[9 + k];
        for (int k = 0; k < 8; k++) {
            nbv[n][k] = iv[13 + k];
            b->nghbr[k] = lit(iv[13 + k]);
        }
        if (iv[21]) {
            alloc_arrays(p, b);
            size_t ok = fread(b->h, sizeof(double), fsize(p), fp);
            ok += fread(b->g, sizeof(double), fsize(p), fp);
            ok += fread(b->C, sizeof(double), csize(p), fp);
            ok += fread(b->P, sizeof(double), psize(p), fp);
            ok += fread(b->Ux, sizeof(double), psize(p), fp);
            ok += fread(b->Uy, sizeof(double), psize(p), fp);
            if (ok != 2 * fsize(p) + csize(p) + 3 * psize(p)) { free(nbv); fclose(fp); return -3; }
        }
    }
    fclose(fp);
    array_of_leaf_blocks(m);
    if (m->nleaf != hdr[8]) return -5;
    /* Rebuild the alias structure: a non-root block's pointers are what
     * Find_Neighbors set when it was created, and the parent's pointers then
     * resolved to the same blocks as now (a refined block has no coarser
     * neighbour under 2:1 balance). Replay Find_Neighbors parent-first. */
    for (int L = 1; L <= 30; L++)
        for (int n = 1; n <= m->nblks; n++) {
            orc_block *b = &m->blk[n];
            if (b->level != L) continue;
            find_neighbors(m, b->parent, n, b->I % 2, b->J % 2);
        }
    int bad = 0;
    for (int n = 1; n <= m->nblks; n++) {
        if (m->blk[n].level < 0) continue;
        for (int k = 1; k <= 8; k++)
            if (orc_nghbr_value(m, n, k) != nbv[n][k - 1]) bad++;
    }
    free(nbv);
    return bad ? -4 : 0;
}
