/*
 * ela3d_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU path
 * (seismicCPML_3d_isotropic_simpelmodel.cpp, tagged CPU below), used as the
 * parity checker of the HIP library and as the CPU baseline in bench.py.
 * It is never linked into, called by or shipped with the product library.
 *
 * Each function cites the reference lines it follows.  The five stencil loops
 * keep the reference's loop order, 0-based-translated ranges and float
 * expression grouping; the typed body lives in oracle_impl.inc and is
 * instantiated for float and double.  C-PML tables are evaluated in FP64.
 *
 * What pins this restatement (tests/test_oracle_*.py):
 *   tables   : closed forms (d0, alpha_max, profile powers, mirror symmetry,
 *              zero-frequency C-PML response 1/s(0) = alpha/(d+K alpha),
 *              discrete reflection coefficient), compat quirks by construction.
 *   wavelets : closed forms and derivative relations between them.
 *   stencils : discrete plane-wave eigenmodes (P and S, along x, y and z)
 *              matched to the exact staggered-grid dispersion relation; 1-D
 *              heterogeneous reductions for all nine (axis, polarisation)
 *              pairs, which pin the neighbour pair of every material average;
 *              a one-step closed form of the ghost-plane sigma_xz/sigma_yz
 *              update (CPU:445); zero-in/zero-out; power-of-two linearity;
 *              translation invariance; energy conservation with the C-PML
 *              off; C-PML absorption against an enlarged domain.
 *   source / receivers : injection formulas checked on one step from rest.
 *   free surface (no reference counterpart; the reference states the intent at
 *              HOM:853) : vertical shear standing waves and horizontal SH waves
 *              exact with the stress image, the plane-stress law on plane 0,
 *              the Rayleigh-wave speed (tests/test_oracle_free_surface.py).
 */
#include "ela3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643

/* ---------------------------------------------------------------------- */
/* C-PML profile tables, CPU:142-361 (x 158-224, y 226-291, z 293-361).    */
/* ---------------------------------------------------------------------- */
int orc_tables(const orc_cfg* c, int axis, double* out)
{
    int n = axis == 0 ? c->nx : (axis == 1 ? c->ny : c->nz);
    double h = axis == 0 ? c->dx : (axis == 1 ? c->dy : c->dz);
    double *d = out, *K = out + n, *al = out + 2 * n, *a = out + 3 * n, *b = out + 4 * n;
    double *dh = out + 5 * n, *Kh = out + 6 * n, *alh = out + 7 * n, *ah = out + 8 * n,
           *bh = out + 9 * n;
    int i;
    if (n <= 0 || h <= 0.0) return -1;
    /* defaults, CPU:169-180 */
    for (i = 0; i < n; ++i) {
        d[i] = 0.0; K[i] = 1.0; al[i] = 0.0; a[i] = 0.0; b[i] = 1.0;
        dh[i] = 0.0; Kh[i] = 1.0; alh[i] = 0.0; ah[i] = 0.0; bh[i] = 1.0;
    }
    if (c->npml <= 0) return 0; /* no absorbing layer at all */

    double L = c->npml * h;                                   /* CPU:147-149 */
    double alpha_max = c->alpha_max > 0.0 ? c->alpha_max : ORC_PI * c->f0; /* CPU:78 */
    double d0 = -(c->npower + 1.0) * c->vpml * log(c->rcoef) / (2.0 * L); /* CPU:154-156 */
    double lo = L;                                            /* CPU:182 */
    double hi = (n - 1) * h - L;                              /* CPU:183 */
    /* normalising thickness per (side, grid) and the min-side half abscissa */
    double Lmin_full = L, Lmin_half = L, Lmax_half = L;
    int zq = (axis == 2) && (c->compat & ORC_COMPAT_ZHALF_MIN);
    double Ly = c->npml * c->dy;
    if (zq) { /* CPU:329, 332, 357: z divides by thickness_PML_y */
        Lmin_full = Ly; Lmin_half = Ly; Lmax_half = Ly;
    }
    /* a free surface at k = 0 has no z-min layer ("disable pml zmin for free
       surface condition", HOM:853) */
    int min_side = !(axis == 2 && c->free_surface);
    for (i = 0; i < n; ++i) {
        double x = h * i;
        double s, abscissa;
        abscissa = lo - x;                                    /* min side, full */
        if (min_side && abscissa >= 0.0) {
            s = abscissa / Lmin_full;
            d[i] = d0 * pow(s, c->npower);
            K[i] = 1.0 + (c->k_max - 1.0) * pow(s, c->npower);
            al[i] = alpha_max * (1.0 - s) + 0.1 * alpha_max;
        }
        if (zq) /* CPU:331: stale yval = DELTAY*(NIMY-1) from the y loop */
            abscissa = Ly - (c->dy * (c->ny - 1) + c->dy / 2.0);
        else
            abscissa = lo - (x + h / 2.0);                    /* min side, half */
        if (min_side && abscissa >= 0.0) {
            s = abscissa / Lmin_half;
            dh[i] = d0 * pow(s, c->npower);
            Kh[i] = 1.0 + (c->k_max - 1.0) * pow(s, c->npower);
            alh[i] = alpha_max * (1.0 - s) + 0.1 * alpha_max;
        }
        abscissa = x - hi;                                    /* max side, full */
        if (abscissa >= 0.0) {
            s = abscissa / L;
            d[i] = d0 * pow(s, c->npower);
            K[i] = 1.0 + (c->k_max - 1.0) * pow(s, c->npower);
            al[i] = alpha_max * (1.0 - s) + 0.1 * alpha_max;
        }
        abscissa = x + h / 2.0 - hi;                          /* max side, half */
        if (abscissa >= 0.0) {
            s = abscissa / Lmax_half;
            dh[i] = d0 * pow(s, c->npower);
            Kh[i] = 1.0 + (c->k_max - 1.0) * pow(s, c->npower);
            alh[i] = alpha_max * (1.0 - s) + 0.1 * alpha_max;
        }
        if (al[i] < 0.0) al[i] = 0.0;                         /* CPU:217-218 */
        if (alh[i] < 0.0) alh[i] = 0.0;
        b[i] = exp(-(d[i] / K[i] + al[i]) * c->dt);           /* CPU:219-220 */
        bh[i] = exp(-(dh[i] / Kh[i] + alh[i]) * c->dt);
        if (fabs(d[i]) > 1e-6)                                /* CPU:222 */
            a[i] = d[i] * (b[i] - 1.0) / (K[i] * (d[i] + K[i] * al[i]));
        if (fabs(dh[i]) > 1e-6)                               /* CPU:223 (clean intent) */
            ah[i] = dh[i] * (bh[i] - 1.0) / (Kh[i] * (dh[i] + Kh[i] * alh[i]));
        if (axis == 0 && (c->compat & ORC_COMPAT_XHALF_A))
            ah[i] = 0.0; /* CPU:223 assigns calpha_x_half, leaving ca_x_half at 0 */
    }
    return 0;
}

/* Source time functions at t=(it-1)dt, a=pi^2 f0^2 (CPU:608-610, HOM:464, UNI:441). */
int orc_wavelet(const orc_cfg* c, int n, double* out)
{
    double a = ORC_PI * ORC_PI * c->f0 * c->f0;
    for (int it = 1; it <= n; ++it) {
        double t = (it - 1) * c->dt;
        double u = t - c->t0;
        double g = exp(-a * u * u);
        double s;
        switch (c->wavelet) {
        case ORC_WAVELET_GAUSSIAN: s = c->factor * g; break;
        case ORC_WAVELET_DGAUSSIAN: s = -c->factor * 2.0 * a * u * g; break;
        case ORC_WAVELET_RICKER: s = c->factor * (1.0 - 2.0 * a * u * u) * g; break;
        default: return -1;
        }
        out[it - 1] = s;
    }
    return 0;
}

static int orc_threads(const orc_cfg* c)
{
#ifdef _OPENMP
    return c->threads > 0 ? c->threads : omp_get_max_threads();
#else
    (void)c;
    return 1;
#endif
}

/* typed bodies */
#define REAL float
#define SFX(name) name##_f32
#include "oracle_impl.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX(name) name##_f64
#include "oracle_impl.inc"
#undef REAL
#undef SFX
