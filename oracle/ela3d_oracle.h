/*
 * ela3d_oracle.h -- CPU restatement of the reference SEISMIC_CPML 3-D isotropic
 * elastic time loop (seismicCPML_3d_isotropic_simpelmodel.cpp), used ONLY as
 * test infrastructure: the parity checker for the HIP path, and the CPU
 * baseline timed beside it.  Nothing in the product links or calls this.
 *
 * Layout: 0-based, x-fastest, offset = i + j*nx + k*nx*ny.  The reference
 * indexes 1-based (i + j*NIMX + k*NIMX*NIMY, CPU:381) over (NX+1)^3 buffers,
 * which is the same array shifted by a constant (CPU:41-43, 80-106).
 *
 * Field ids (shared with the product's C ABI):
 *   0 vx 1 vy 2 vz 3 sxx 4 syy 5 szz 6 sxy 7 sxz 8 syz
 *   memory variables, stress pass (CPU:395-405, 432-437, 459-464, 486-491)
 *   9 dvx_dx 10 dvy_dy 11 dvz_dz 12 dvy_dx 13 dvx_dy 14 dvz_dx 15 dvx_dz
 *   16 dvz_dy 17 dvy_dz
 *   memory variables, velocity pass (CPU:518-524, 547-553, 577-583)
 *   18 dsxx_dx 19 dsxy_dy 20 dsxz_dz 21 dsxy_dx 22 dsyy_dy 23 dsyz_dz
 *   24 dsxz_dx 25 dsyz_dy 26 dszz_dz
 */
#ifndef ELA3D_ORACLE_H
#define ELA3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* quirk switches reproducing the reference CPU program bit-for-bit in behaviour */
#define ORC_COMPAT_XHALF_A   1u  /* CPU:223: ca_x_half never assigned (stays 0)          */
#define ORC_COMPAT_ZHALF_MIN 2u  /* CPU:325-336: z tables normalised by the y thickness, */
                                 /* z-half min side tested with the stale y abscissa   */
#define ORC_COMPAT_GHOST     4u  /* CPU:445: sxz/syz loop reaches k0=nz-1, zero ghost   */

/* source kinds */
#define ORC_SRC_REFERENCE_SIGMA 0 /* CPU:620-622 sxx,syy += fx*dt*(l+2m); szz += fy*dt*(l+2m) */
#define ORC_SRC_FORCE           1 /* CPU:614-617 vx += fx*dt/rho_x; vy += fy*dt/rho_y          */
#define ORC_SRC_EXPLOSIVE       2 /* sxx=syy=szz += s*dt*(l+2m)                              */
#define ORC_SRC_FORCE3          3 /* seismicCPML_3d_isotropic_cudazerocopy.cu:451-453         */

/* wavelets */
#define ORC_WAVELET_GAUSSIAN  0  /* HOM:464            */
#define ORC_WAVELET_DGAUSSIAN 1  /* CPU:610            */
#define ORC_WAVELET_RICKER    2  /* UNI:441            */

typedef struct orc_cfg {
    int nx, ny, nz;
    double dx, dy, dz, dt;
    int nstep;
    int npml;                 /* 0 disables the C-PML entirely */
    double npower, k_max, alpha_max, rcoef, vpml;
    double f0, t0, factor, angle_deg;
    int src_kind, wavelet;
    int isrc, jsrc, ksrc;     /* 0-based */
    int nrec;
    const int* rec;           /* [nrec][3] 0-based i,j,k */
    int sample_every;         /* seismogram cadence (reference sampgat, CPU:32) */
    int energy_every;         /* 0: no energy */
    unsigned compat;
    int dirichlet;            /* zero vx,vy,vz on the six faces after the v update */
    int free_surface;         /* traction-free surface at k = 0, no z-min C-PML (the
                                 reference's stated intent, HOM:853; not implemented there) */
    int threads;              /* OpenMP threads, <=0: runtime default */
} orc_cfg;

/* C-PML profile tables of one axis (0=x,1=y,2=z), FP64, CPU:142-361.
 * out: 10*n doubles: d, K, alpha, a, b (full grid) then the same on the half grid. */
int orc_tables(const orc_cfg* c, int axis, double* out);

/* source time function s(t) at t=(it-1)dt for it=1..n, FP64 (CPU:608-610, UNI:441) */
int orc_wavelet(const orc_cfg* c, int n, double* out);

/* One stress pass (CPU:377-499) / one velocity pass (CPU:501-590) on full-grid
 * state.  f[9] fields, m[18] memory variables (all nx*ny*nz). No source. */
void orc_stress_f32(const orc_cfg* c, const float* cp, const float* cs, const float* rho,
                    float* const* f, float* const* m);
void orc_velocity_f32(const orc_cfg* c, const float* rho, float* const* f, float* const* m);
void orc_stress_f64(const orc_cfg* c, const double* cp, const double* cs, const double* rho,
                    double* const* f, double* const* m);
void orc_velocity_f64(const orc_cfg* c, const double* rho, double* const* f, double* const* m);

/* Total energy of a state, FP64 sums over the non-PML box [npml, n-1-npml)^3,
 * out[0]=kinetic, out[1]=potential (J, times the cell volume). */
void orc_energy_f32(const orc_cfg* c, const float* cp, const float* cs, const float* rho,
                    float* const* f, double* out);
void orc_energy_f64(const orc_cfg* c, const double* cp, const double* cs, const double* rho,
                    double* const* f, double* out);

/* Run steps it0+1 .. it0+nsteps of the time loop (CPU:376-622) on the given state.
 * seismo: [nstep/sample_every][nrec][3] doubles (indexed by absolute step),
 * energy: [nstep/energy_every][2].  Either may be NULL.  Returns 0 or -1. */
int orc_run_f32(const orc_cfg* c, const float* cp, const float* cs, const float* rho,
                int it0, int nsteps, float* const* f, float* const* m,
                double* seismo, double* energy);
int orc_run_f64(const orc_cfg* c, const double* cp, const double* cs, const double* rho,
                int it0, int nsteps, double* const* f, double* const* m,
                double* seismo, double* energy);

#ifdef __cplusplus
}
#endif
#endif
