/*
 * ela3d.h -- C ABI of the MI355X-native SEISMIC_CPML 3-D isotropic elastic
 * velocity-stress time loop.
 *
 * The reference (thejatmik/ela3d) has no library API: its "surface" is the
 * parameter block and time loop inside main() of
 * seismicCPML_3d_isotropic_simpelmodel.cpp (tagged CPU below; the CUDA variants
 * seismicCPML_3d_isotropic_cudazerocopy.cu = ZC, elatic_3d_unified_cuda_memory.cu
 * = UNI, seismic_CPML_3D_isotropic_homogeneous_withCUDA.cu = HOM).  Each entry
 * point below names the reference lines it replaces.
 *
 * Conventions
 *   - Arrays are x-fastest, 0-based: offset = i + j*nx + k*nx*ny.  The reference
 *     uses 1-based (i,j,k) with the same x-fastest order (CPU:381); subtract 1.
 *   - All functions return ELA3D_OK (0) or a negative ELA3D_E_* code; the library
 *     never exits the process and no C++ exception crosses this boundary.
 *     ela3d_last_error() gives a message for the calling thread's last failure.
 *   - A shot owns all of its device memory; model arrays are copied at create;
 *     outputs are copied into caller buffers.  No returned pointer aliases
 *     library memory.  One shot = one HIP stream; a shot is not thread-safe.
 *     The server (ela3d_server_*) is thread-safe for submit/poll/wait.
 */
#ifndef ELA3D_H
#define ELA3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ELA3D_ABI_VERSION 1

/* ---- status codes (replace HANDLE_ERROR -> exit(EXIT_FAILURE), HOM:153-162) ---- */
#define ELA3D_OK           0
#define ELA3D_E_INVALID   -1  /* invalid config or argument (CPU:631-634 is the reference's only check) */
#define ELA3D_E_OOM       -2  /* device or host allocation failed                        */
#define ELA3D_E_UNSTABLE  -3  /* non-finite wavefield detected (e.g. dt above Courant)   */
#define ELA3D_E_HIP       -4  /* HIP runtime error                                        */
#define ELA3D_E_NODEVICE  -5  /* no usable GPU                                            */
#define ELA3D_E_RANGE     -6  /* output buffer too small / index out of range            */
#define ELA3D_E_STATE     -7  /* call not valid in the object's current state            */
#define ELA3D_E_BUDGET    -8  /* shot can never fit the server's device-memory budget    */
#define ELA3D_E_NOTFOUND  -9  /* unknown ticket                                           */
#define ELA3D_E_TIMEOUT  -10  /* wait timed out                                           */
#define ELA3D_E_IO       -11  /* file could not be opened, read or written                */

/* ---- source kinds (CPU:611-622, ZC:446-459) ---- */
#define ELA3D_SRC_REFERENCE_SIGMA 0 /* CPU:620-622: sxx,syy += fx*dt*(l+2m); szz += fy*dt*(l+2m) */
#define ELA3D_SRC_FORCE           1 /* CPU:614-617: vx += fx*dt/rho_x; vy += fy*dt/rho_y        */
#define ELA3D_SRC_EXPLOSIVE       2 /* sxx = syy = szz += s*dt*(l+2m)                          */
#define ELA3D_SRC_FORCE3          3 /* ZC:451-453: vx, vy, vz += s*dt/rho_avg                  */

/* ---- source time functions, t = (it-1)*dt, a = pi^2 f0^2 (CPU:608-610) ---- */
#define ELA3D_WAVELET_GAUSSIAN  0 /* HOM:464 */
#define ELA3D_WAVELET_DGAUSSIAN 1 /* CPU:610 */
#define ELA3D_WAVELET_RICKER    2 /* UNI:441 */

/* ---- wavefield storage ---- */
#define ELA3D_STORAGE_FP32 0 /* FP32 fields and memory variables (default)                */
#define ELA3D_STORAGE_FP16 1 /* FP16 storage, per-field power-of-two scale, FP32 math     */
#define ELA3D_STORAGE_FP64 2 /* FP64 storage and math (the GPU FP64 reference run)        */

/* ---- reproduce the reference CPU program's quirks (all table/loop values only) ---- */
#define ELA3D_COMPAT_XHALF_A   1u /* CPU:223: ca_x_half is never assigned and stays 0      */
#define ELA3D_COMPAT_ZHALF_MIN 2u /* CPU:325-336: z tables normalised by the y thickness,
                                     z-half min side tested with the stale y abscissa     */
#define ELA3D_COMPAT_GHOST     4u /* CPU:445: sxz/syz also updated on k0=nz-1 against a
                                     zero ghost plane                                      */
#define ELA3D_COMPAT_REFERENCE 7u /* all of the above: the reference CPU path as written  */

/* field ids for ela3d_snapshot / ela3d_set_field */
#define ELA3D_FIELD_VX 0
#define ELA3D_FIELD_VY 1
#define ELA3D_FIELD_VZ 2
#define ELA3D_FIELD_SXX 3
#define ELA3D_FIELD_SYY 4
#define ELA3D_FIELD_SZZ 5
#define ELA3D_FIELD_SXY 6
#define ELA3D_FIELD_SXZ 7
#define ELA3D_FIELD_SYZ 8
/* 9..17: memory_dvx_dx, dvy_dy, dvz_dz, dvy_dx, dvx_dy, dvz_dx, dvx_dz, dvz_dy, dvy_dz
   18..26: memory_dsigmaxx_dx, dsigmaxy_dy, dsigmaxz_dz, dsigmaxy_dx, dsigmayy_dy,
           dsigmayz_dz, dsigmaxz_dx, dsigmayz_dy, dsigmazz_dz          (CPU:395-583) */
#define ELA3D_NUM_FIELDS 27

/* The reference's parameter block (CPU:19-78, 146-153) as one versioned struct.
 * Fill with ela3d_config_default() first. */
typedef struct ela3d_config {
    uint32_t struct_size;   /* sizeof(ela3d_config) */
    uint32_t abi_version;   /* ELA3D_ABI_VERSION */
    int32_t nx, ny, nz;     /* NIMX, NIMY, NIMZ (CPU:20-23) */
    double dx, dy, dz;      /* DELTAX, DELTAY, DELTAZ [m] (CPU:35-37) */
    double dt;              /* DELTATT [s] (CPU:31) */
    int32_t nstep;          /* NSTEP (CPU:30) */
    int32_t npoints_pml;    /* NPOINTS_PML (CPU:66); 0 disables the C-PML */
    double npower;          /* NPOWER (CPU:75), default 2 */
    double k_max_pml;       /* K_MAX_PML (CPU:77), default 1 */
    double alpha_max_pml;   /* ALPHA_MAX_PML; <= 0 means pi*f0 (CPU:78) */
    double rcoef;           /* Rcoef (CPU:151), default 1e-3 */
    double vpml;            /* vpml (CPU:153), default 3000 m/s */
    double f0;              /* dominant frequency [Hz] (CPU:62) */
    double t0;              /* wavelet delay [s]; <= 0 means 1.2/f0 (CPU:63) */
    double factor;          /* amplitude (CPU:64), default 1e7 */
    double angle_force;     /* ANGLE_FORCE [deg] (CPU:72) */
    int32_t src_kind;       /* ELA3D_SRC_* */
    int32_t wavelet;        /* ELA3D_WAVELET_* */
    int32_t isrc, jsrc, ksrc; /* ISOURCE-1, JSOURCE-1, KSOURCE-1 (0-based, CPU:69-71) */
    int32_t nrec;           /* number of receivers */
    const int32_t* rec;     /* [nrec][3] 0-based (i,j,k); replaces the plane gather CPU:593-602 */
    int32_t sample_every;   /* sampgat (CPU:32): seismogram sample every N steps */
    int32_t energy_every;   /* total energy every N steps; 0 = off (no reference counterpart) */
    int32_t storage;        /* ELA3D_STORAGE_* */
    uint32_t compat_flags;  /* ELA3D_COMPAT_* */
    int32_t dirichlet;      /* 1: vx=vy=vz=0 on the six faces after each velocity update */
    int32_t fp16_rescale_every; /* FP16 storage: re-choose scales every N steps (default 100) */
    int32_t device;         /* HIP device ordinal */
    int32_t use_graph;      /* 1 (default): replay a hipGraph of one step */
    int32_t block;          /* threads per workgroup, 0 = autotuned */
    int32_t zseg;           /* z planes per workgroup, 0 = autotuned */
    int32_t free_surface;   /* 1: traction-free surface at the plane k = 0 (no z-min C-PML
                               there); the reference only marks the intent (HOM:853,
                               UNI:1031, ZC:1304: "disable pml zmin for free surface") */
    int32_t reserved[5];
} ela3d_config;

typedef struct ela3d_shot ela3d_shot;

typedef struct ela3d_info {
    int64_t cells;                 /* nx*ny*nz */
    int64_t device_bytes;          /* device memory held by the shot */
    double alg_bytes_per_step;     /* algorithmic HBM bytes of one step (two-pass schedule) */
    double alg_bytes_stress;       /* stress pass (interior + absorbing-layer launches) */
    double alg_bytes_velocity;     /* velocity pass (interior + absorbing-layer launches) */
    double pml_bytes_per_step;     /* the C-PML memory-variable share of alg_bytes_per_step */
    int32_t nsamples;              /* seismogram samples over nstep */
    int32_t nenergy;               /* energy samples over nstep */
    int32_t steps_done;
    int32_t block, zseg;           /* interior launch: threads, z planes per workgroup */
    int32_t zseg_slab, vec;        /* absorbing-layer z planes per workgroup; interior xy positions per lane */
    int32_t region_lo[3], region_hi[3]; /* C-PML compact regions: [0,lo) u [hi,n) per axis */
    int32_t block_slab, vec_slab;  /* absorbing-layer launch: threads, xy positions per lane */
    int32_t overlap;               /* absorbing-layer launches concurrent with the interior ones */
    int32_t merged;                /* one mixed launch per pass */
    int32_t dbuf;                  /* two field sets (double buffered) */
    int32_t fused, tile;           /* fused interior schedule and its tile */
    int32_t strips;                /* fused: xy strips of its workgroup order in effect (0: up to 8) */
} ela3d_info;

#define ELA3D_MAX_KERNELS 8
typedef struct ela3d_kernel_stats {
    char name[48];
    int64_t launches;
    double total_ms;               /* summed hipEvent durations */
    double alg_bytes_per_launch;
} ela3d_kernel_stats;

/* ---- shot API (replaces main() CPU:17-726 / ZC:487-1506) ---- */
/* Defaults: the homogeneous 101^3 cube workload (the reference's medium and C-PML,
 * CPU:41-78, 146-153, with dt = 1.6e-3, a Ricker force source at f0 = 7 Hz at the
 * centre and rigid faces).  The reference's own literals (f0 = 35, a derivative-
 * of-Gaussian sigma source, dt = 1e-3, sampgat = 2, its quirks) are
 * Config.from_reference in the Python binding and reference=1 in the CLI. */
void ela3d_config_default(ela3d_config* cfg);
/* Full config check without a GPU (the reference's only check is CPU:631-634). */
int ela3d_validate(const ela3d_config* cfg);
/* Create a shot: validate, build FP64 C-PML tables (CPU:142-361) and the FP64
 * wavelet (CPU:604-610), copy the model (CPU:41-59), zero the 27 fields (CPU:107-140). */
int ela3d_create(const ela3d_config* cfg, const float* cp, const float* cs, const float* rho,
                 ela3d_shot** out);
/* Enqueue n time steps (CPU:376-622) on the shot's stream; asynchronous. */
int ela3d_step(ela3d_shot* shot, int32_t nsteps);
/* Wait for the shot's queued work; ELA3D_E_UNSTABLE once a non-finite value was seen at the
 * source cell (checked every step), a receiver sample or the energy. */
int ela3d_sync(ela3d_shot* shot);
/* ela3d_step(nstep - done) + ela3d_sync */
int ela3d_run(ela3d_shot* shot);
/* Seismograms [nsamples][nrec][3] as FP64 (replaces rechorv*.txt, CPU:624-671). */
int ela3d_seismograms(ela3d_shot* shot, double* out, size_t cap);
/* Energy [nenergy][2] = (kinetic, potential) per energy step, FP64 (no reference counterpart). */
int ela3d_energy(ela3d_shot* shot, double* out, size_t cap);
/* Full-grid copy of one field (ids above) as FP64, nx*ny*nz values (CPU:685-722 wrote vx,vy,vz). */
int ela3d_snapshot(ela3d_shot* shot, int32_t field, double* out, size_t cap);
/* Overwrite one field from FP64 values (rounded to the storage type); the reference
 * starts from zero fields only (CPU:107-140).  In FP16 storage the field's scale is
 * re-fitted to the values when the current one would overflow or lose them. */
int ela3d_set_field(ela3d_shot* shot, int32_t field, const double* in, size_t n);
/* sigma_xx, sigma_yy, sigma_zz at the source cell right after each step's injection,
 * [nstep+1][3] FP64 with row 0 zero (the lesrcx/y/z arrays, UNI:463-465, 1202-1223). */
int ela3d_source_trace(ela3d_shot* shot, double* out, size_t cap);
/* Checkpoint / resume (no reference counterpart; the reference snapshots cannot
 * resume).  Writes the shot's whole device state and step count to a file; a shot
 * created from the same config and model can load it and continue bit-identically.
 * ELA3D_E_INVALID if the file belongs to another config, ELA3D_E_IO on file errors;
 * a file of the wrong length is refused before any state is overwritten. */
int ela3d_checkpoint_save(ela3d_shot* shot, const char* path);
int ela3d_checkpoint_load(ela3d_shot* shot, const char* path);
/* ---- z-slabs: one GPU's share of a grid decomposed along z (no reference
 * counterpart; the reference indexes its C-PML tables by global position at
 * HOM:177-179, and the planes that cross a z interface are the ones its loops
 * read at k-1 / k+1: CPU:384, 450, 477, 508, 569) ----
 * A slab owns planes [k0, k1) of the grid `cfg` describes, plus one ghost plane
 * on each side that has a neighbour.  Its C-PML z tables, update ranges, rigid
 * faces and energy box are the global ones; receivers outside its planes are
 * dropped (receivers[] gets the global indices of those it keeps, in order) and
 * the source is injected only by the slab that owns its plane.  cp, cs, rho are
 * the whole grid's model.  FP32 or FP64 storage.  A slab steps by phases; after
 * each phase every pair of neighbours exchanges halos: the lower one packs side
 * 1 and the upper one unpacks side 0, the upper one packs side 0 and the lower
 * one unpacks side 1 (the buffers are device memory; copies run on the slab's
 * stream).  Phase 0 sends sigma_xz, sigma_yz up and sigma_zz down; phase 1
 * sends vz up and vx, vy down.  Stepped this way, every slab's fields equal the
 * whole grid's on its planes. */
int ela3d_create_slab(const ela3d_config* cfg, int32_t k0, int32_t k1, const float* cp, const float* cs,
                      const float* rho, ela3d_shot** out, int32_t* receivers, size_t receivers_cap,
                      int32_t* nreceivers);
/* phase 0: the stress pass; phase 1: the velocity pass, receivers, energy, source (one step) */
int ela3d_step_phase(ela3d_shot* shot, int32_t phase);
/* bytes a halo_pack(phase, side) writes (what the neighbour's unpack on the other side reads) */
int ela3d_halo_bytes(ela3d_shot* shot, int32_t phase, int32_t side, size_t* bytes);
int ela3d_halo_pack(ela3d_shot* shot, int32_t phase, int32_t side, void* dev_buf, size_t cap);
int ela3d_halo_unpack(ela3d_shot* shot, int32_t phase, int32_t side, const void* dev_buf, size_t n);
/* Both halos of one interface between two slabs on the same GPU after `phase`,
 * through a staging buffer the lower slab keeps (device-to-device; synchronous). */
int ela3d_halo_exchange(ela3d_shot* lower, ela3d_shot* upper, int32_t phase);

/* Allocation-free run (no counterpart in the reference).  For callers that cannot
 * allocate device memory (no hipMalloc allowed in their loop, harnesses that forbid
 * it): ela3d_create + ela3d_run + ela3d_seismograms + ela3d_energy + ela3d_destroy in
 * one synchronous call, with the shot's whole device state in a static device buffer
 * of ELA3D_NOALLOC_BYTES that each GPU the library runs on holds until the process
 * ends.  It calls no hipMalloc or hipFree: the buffer is zeroed by a kernel, the
 * launches are eager (no graph), in place and in sequence on `stream` (a
 * hipStream_t; NULL is the default stream).  Results equal ela3d_create/ela3d_run's
 * for the same config with the same launch shape.  A config whose device state
 * exceeds the buffer is refused with ELA3D_E_RANGE before anything runs.  Every call
 * on one GPU uses that GPU's one buffer: calls that overlap on one GPU race. */
#define ELA3D_NOALLOC_BYTES (128ll * 1024 * 1024)
int ela3d_run_noalloc(const ela3d_config* cfg, const float* cp, const float* cs, const float* rho,
                      double* seis, size_t seis_cap, double* energy, size_t energy_cap, void* stream);
/* Sizes, regions, launch shape and algorithmic bytes of a shot (no reference counterpart). */
int ela3d_info_get(ela3d_shot* shot, ela3d_info* out);
/* Instrumentation (no reference counterpart).
 * Run n steps eagerly with hipEvents around every launch; per-kernel totals. */
int ela3d_profile(ela3d_shot* shot, int32_t nsteps, ela3d_kernel_stats* out, int32_t cap,
                  int32_t* nkernels, double* wall_ms);
/* Run n steps: step s (0-based) eagerly with hipEvents around every launch when
 * s % every == 0, the others as graph replays (the production path); per-kernel
 * totals over the instrumented steps, wall time over all n. */
int ela3d_profile_sampled(ela3d_shot* shot, int32_t nsteps, int32_t every, ela3d_kernel_stats* out,
                          int32_t cap, int32_t* nkernels, double* wall_ms);
void ela3d_destroy(ela3d_shot* shot);
/* Host-only introspection (no GPU): the plan a config expands to (cells, regions,
 * algorithmic bytes) and its FP64 C-PML tables of one axis, 10*n doubles:
 * d, K, alpha, a, b (full grid) then the same on the half grid (CPU:158-361). */
int ela3d_plan(const ela3d_config* cfg, ela3d_info* out);
int ela3d_pml_tables(const ela3d_config* cfg, int32_t axis, double* out, size_t cap);

/* ---- serving: continuous batching of forward shots on one GPU (no reference
 * counterpart; the reference runs one hard-coded shot per process) ----
 * A device thread admits queued shots in arrival order while the device-memory
 * budget and the active-shot cap allow, steps every admitted shot on its own
 * stream in rounds of steps_per_round graph replays, and retires finished shots
 * (results copied to host) so that later arrivals are admitted while earlier
 * ones still run.  plain_loop = 1 instead runs one shot at a time the way the
 * reference drives its kernels (ZC:1374-1494): eager launches, a blocking
 * host-to-device copy of the step index and a blocking receiver readback every
 * step.  submit/poll/wait are thread-safe. */
typedef struct ela3d_server ela3d_server;
typedef struct ela3d_server_opts {
    uint32_t struct_size;      /* sizeof(ela3d_server_opts) */
    int32_t device;
    int64_t hbm_budget_bytes;  /* admission budget for shot device memory (default 200 GB) */
    int32_t max_active;        /* shots resident and stepping at once (default 16) */
    int32_t steps_per_round;   /* steps enqueued per shot between completion checks (default 50) */
    int32_t use_graph;         /* replay a hipGraph per step (default 1) */
    int32_t plain_loop;        /* 1: reference-pattern one-shot-at-a-time loop */
    int32_t batch_cells;       /* FP32 shots of at most this many cells without energy sums
                                * step together, one stress, one velocity and one tail launch
                                * per step for all of them (default 128^3; 0: never) */
    int32_t reserved[7];
} ela3d_server_opts;
typedef struct ela3d_result {
    uint64_t ticket;
    int32_t status;            /* ELA3D_OK or the failing ELA3D_E_* code */
    int32_t nsamples, nrec, nenergy;
    double t_submit, t_admit, t_done; /* seconds since the server started */
    int64_t device_bytes;
} ela3d_result;
#define ELA3D_SHOT_QUEUED  0
#define ELA3D_SHOT_RUNNING 1
#define ELA3D_SHOT_DONE    2
void ela3d_server_opts_default(ela3d_server_opts* o);
int ela3d_server_create(const ela3d_server_opts* o, ela3d_server** out);
/* Queue one shot; config and model are copied.  ELA3D_E_INVALID for a config or model
 * that ela3d_create would refuse, ELA3D_E_BUDGET if it can never fit. */
int ela3d_submit(ela3d_server* s, const ela3d_config* cfg, const float* cp, const float* cs,
                 const float* rho, uint64_t* ticket);
/* state: ELA3D_SHOT_QUEUED / RUNNING / DONE, or a negative ELA3D_E_* code if it failed */
int ela3d_poll(ela3d_server* s, uint64_t ticket, int32_t* state);
/* Block until the shot is done (timeout_s < 0: forever); fills *res. */
int ela3d_wait(ela3d_server* s, uint64_t ticket, double timeout_s, ela3d_result* res);
int ela3d_result_seismograms(ela3d_server* s, uint64_t ticket, double* out, size_t cap);
int ela3d_result_energy(ela3d_server* s, uint64_t ticket, double* out, size_t cap);
int ela3d_result_release(ela3d_server* s, uint64_t ticket);
/* out[10]: submitted, done, failed, active, queued, device bytes in use, peak device bytes, rounds,
 * shots that stepped in the batch launches, batch rounds */
int ela3d_server_stats(ela3d_server* s, int64_t* out);
void ela3d_server_destroy(ela3d_server* s);

/* ---- launch-shape autotuner (no reference counterpart) ----
 * Each pass is two launches: the interior box and the absorbing-layer boxes.
 * The tuner times the interior candidates (threads per workgroup, z planes per
 * workgroup, xy positions per lane), then the absorbing-layer candidates, on a
 * scratch shot of this grid and caches the fastest for (nx, ny, nz, storage,
 * npoints_pml); shots created with block = 0 and zseg = 0 use the cached shape. */
typedef struct ela3d_tuning {
    int32_t block, zseg_interior, zseg_slab;
    int32_t vec;          /* interior: consecutive xy positions per lane, 1, 2 or 4 */
    int32_t candidates;
    double ms_per_step;   /* stress + velocity launches */
    int32_t block_slab;   /* absorbing-layer launch: threads per workgroup (0 = block) */
    int32_t vec_slab;     /* absorbing-layer launch: xy positions per lane, 1 or 2 (0 = 1) */
    int32_t overlap;      /* 1: absorbing-layer launches run concurrently with the interior ones */
    int32_t merged;       /* 1: one mixed launch per pass over every box (interior shape) */
    int32_t dbuf;         /* 1: two field sets; each step reads one and writes the other
                           * (FP32 and FP64 storage; FP16 storage always updates in place) */
    int32_t fused;        /* 1: one fused stress+velocity launch over the interior box (needs dbuf) */
    int32_t tile;         /* fused: interior cells per workgroup (multiple of 64, <= 1024) */
    int32_t strips;       /* fused: xy strips the workgroup order walks, each through every z
                           * segment before the next (0 = the default: 4 for tiles of up to 768
                           * threads and at least 32 tiles per plane, else up to 8 strips of at
                           * least 8 tiles) */
} ela3d_tuning;
int ela3d_autotune(const ela3d_config* cfg, const float* cp, const float* cs, const float* rho,
                   int32_t steps, ela3d_tuning* out);
int ela3d_tuning_set(const ela3d_config* cfg, const ela3d_tuning* t);
int ela3d_tuning_get(const ela3d_config* cfg, ela3d_tuning* t);

/* Failure injection (no reference counterpart): cap the device bytes the library
 * holds at once (shot memory, the server's pool, halo staging), process-wide; an
 * allocation past the cap fails with ELA3D_E_OOM as a real out-of-memory would.
 * bytes <= 0 removes the cap. */
int ela3d_set_device_limit(int64_t bytes);
/* Error text (replaces HANDLE_ERROR's printf + exit, HOM:153-162). */
const char* ela3d_strerror(int code);
const char* ela3d_last_error(void);
/* Visible HIP devices (no reference counterpart). */
int ela3d_device_count(int32_t* out);
/* Stream-copy kernel of n bytes, timed with hipEvents (HBM ceiling calibration; no
 * reference counterpart).
 * variant 0: grid-stride float4; 1: one float4 per thread; 2: four float4 per
 * thread, nontemporal; 3: one float per thread.  Multi-stream variants, every
 * array `bytes` long, element-wise in one pass: 4: 12 arrays read + 6 written,
 * one float per lane (a stress pass); 5: 10 + 3 (a velocity pass); 6: 12 + 9
 * (one fused step); 7: 12 + 6 with float4 per lane; 8: 12 + 6, one float per lane,
 * the arrays interleaved in blocks of 1024 floats ([n/1024][12][1024] read,
 * [n/1024][6][1024] written; `bytes` is rounded down to a multiple of 4096). */
int ela3d_copy_bench(int32_t device, int64_t bytes, int32_t reps, int32_t variant, double* ms_per_copy);

/* ---- seeded synthetic models (CPU:41-59 homogeneous; layered per the workload spec) ---- */
int ela3d_model_homogeneous(int32_t nx, int32_t ny, int32_t nz, float vp, float vs, float rho,
                            float* cp, float* cs, float* rh);
typedef struct ela3d_layered_spec {
    int32_t nx, ny, nz;
    double h;                 /* grid spacing [m] */
    int32_t nlayers;          /* 6 */
    double vp_top, vp_bottom; /* 2000, 4500 m/s, linear in layer index */
    double noise;             /* 0.03: uniform +-3 % per cell */
    uint64_t seed;            /* splitmix64 seed */
    double bump_amp;          /* A = 600 m (anticline uplift) */
    double bump_width;        /* w = 1200 m */
    double vp_vs;             /* 1.73 */
    double gardner_a, gardner_b; /* rho = a * vp^b: 310, 0.25 */
} ela3d_layered_spec;
void ela3d_layered_default(ela3d_layered_spec* s);
int ela3d_model_layered(const ela3d_layered_spec* s, float* cp, float* cs, float* rh,
                        double* vp_max);
/* The MATLAB anticline (generate_3d_model_anticline.m): an n^3 grid (n = 201 there)
 * with interfaces h - d*Z(x,y), Z = max(0, sin x + sin y - 1) on x, y in [0, pi],
 * h = 121, 161, 201 (scaled by n/201), d = 100*n/201, five media (MAT:72-76)
 * including the oil-bearing sandstone above k = 80*n/201.  cp, cs, rho: n^3 each. */
int ela3d_model_anticline(int32_t n, float* cp, float* cs, float* rh);

/* ---- the reference's file formats (host only) ---- */
/* Raw x-fastest model files (modelcp.bin / modelcs.bin / modelrh.bin): float64 as
 * read at UNI:595-644 (bytes_per_value 8) or float32 as at ZC:544-595 (4); read
 * into FP32.  Writing uses the same layouts (snapshots v?%05i.bin are float32,
 * CPU:685-722). */
int ela3d_read_raw(const char* path, int64_t n, int32_t bytes_per_value, float* out);
int ela3d_write_raw(const char* path, int64_t n, int32_t bytes_per_value, const float* in);
/* Text traces in the layout of rechorvx/y/z.txt (CPU:636-670): one line per row,
 * each of ncols values as the reference's float stream prints it (6 significant
 * digits) followed by a space.  Value (r, c) = v[(r*ncols + c)*stride + offset]:
 * seismograms [nsamples][nrec][3] with stride 3, offset = component.  append = 1
 * appends as the reference does; final_newline = 0 omits the last newline
 * (recsrc*.txt, UNI:1202-1223, is one line without one). */
int ela3d_write_traces(const char* path, const double* v, int64_t nrows, int64_t ncols,
                       int32_t stride, int32_t offset, int32_t append, int32_t final_newline);

#ifdef __cplusplus
}
#endif
#endif /* ELA3D_H */
