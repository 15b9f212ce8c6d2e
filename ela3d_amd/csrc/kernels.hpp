// kernels.hpp -- argument blocks of the gfx950 stencil kernels (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

namespace ela3d {

typedef _Float16 half_t;

// box modes
constexpr int kInterior = 0;  // y and z outside the C-PML: only x is checked per lane
constexpr int kSlab = 1;      // y or z inside the C-PML: every axis checked
constexpr int kMixed = 2;     // launch mode only: boxes of both kinds, each block runs its box's body
constexpr int kFused = 3;     // launch mode only: one fused stress+velocity step over the interior box

// FP16 storage keeps one power-of-two scale per field (stored = value * scale)
// in a device table of kScaleTable floats, so replayed graphs see live values:
//   [0,27) current scale of the data in memory, [27,54) its inverse,
//   [54,81) scale of the data being written this step, [81,108) its inverse,
//   [108,135) each field's largest |value| of the previous step (0 at first).
// The tail kernel makes next current at the end of a step and picks new next
// scales from the maxima the stencil kernels fold into kMaxSlots spread slots.
constexpr int kScaleTable = 5 * 27;
constexpr int kMaxSlots = 64;

// A launch covers up to kMaxBox boxes.  A box is a flat xy range [p0,p1)
// (p = i + j*nx) times a z range [k0,k1); its blocks march z in segments of
// zseg planes, one thread per xy position.
constexpr int kMaxBox = 5;
struct Box {
    int p0, p1, k0, k1;
    int nchunk;  // ceil((p1-p0)/block)
    int zseg;    // planes per block
    int blk0;    // first logical block of this box
    int mode;    // kInterior or kSlab
    int nstrip;  // xy strips of the tile order (0: up to 8 of at least 8 chunks, see tile_of)
};

// Fields are read through fi and written through fo, memory variables read
// through mi and written through mo.  In place, fi == fo and mi == mo.  Double
// buffered, the stress pass reads v and sigma of the current set and writes
// sigma to the next set, and the velocity pass reads the new sigma and the
// current v and writes v to the next set.
template <typename T>
struct StepArgs {
    const T* fi[9];    // vx vy vz sxx syy szz sxy sxz syz, padded planes (see shot.cpp)
    T* fo[9];
    const T* mi[18];   // memory variables, compact per axis family (see shot.cpp)
    T* mo[18];
    const float* cp;
    const float* cs;
    const float* rho;
    const double* tab[3];  // per axis [6][n]: a, b, 1/K (full grid), a, b, 1/K (half grid)
    const float* tabf;     // FP32 x then y tables [6][nx], [6][ny]: a, b, 1/K, a, b, 1/K (full, half grid)
    int nx, ny, nz, nxy;
    int nxyp;                // padded plane stride of the 9 fields and the model (elements)
    int g0;                  // guard elements before plane 0; f[] and cp/cs/rho point at the guard
    int lo[3], hi[3], w[3];  // compact C-PML regions [0,lo) u [hi,n), width w
    int kmax_shear;          // last plane of the sxz/syz update
    int dirichlet;
    int fs;                  // traction-free surface on plane 0 (absorbing-layer kernels)
    int emargin;             // energy box [emargin, n-1-emargin) along x and y
    int ek0, ek1;            // energy box planes [ek0, ek1) (a slab's local indices of the global box)
    double dt, idx, idy, idz;
    const float* scl;        // FP16 storage scales, kScaleTable floats (layout below)
    // FP16 storage with a stress source: the source cell's sigma_xx, sigma_yy,
    // sigma_zz live unscaled in side[0..2] (FP32) instead of the FP16 arrays,
    // which hold 0 there, so the static stress a non-zero-mean wavelet builds up
    // at that one cell neither loses precision nor sets the fields' scales.
    // sidx: the cell's element offset (as cell()).  side == nullptr: off.
    float* side;
    unsigned sidx;
    unsigned* maxslot;       // [kMaxSlots][27] |value| maxima as float bits (FP16 storage)
    double* epart;           // energy partials [nblocks][2]
    Box box[kMaxBox];
    int nbox, nblocks, block;
    int vec;                 // consecutive xy positions per lane: 1, 2 or 4
    int mode;                // kInterior or kSlab (every box of the launch), kMixed or kFused
    // kFused launches address both field sets from one pointer per array: fi[q]
    // and mi[q] hold the first copy of each array (the second copy follows it),
    // a step reads at element offset fdin (fields) / mdin (memory variables) from
    // them and writes at fdout / mdout; fo and mo are unused.
    unsigned fdin, fdout, mdin, mdout;
    int tile;                // kFused: threads per workgroup = region cells (own cells + 2 halo rows)
    int ldsreg;              // kFused: cells of one LDS stress plane (= tile)
};

template <typename T>
struct TailArgs {
    T* f[9];
    const float* cp;
    const float* rho;
    const int* rec;
    int nrec;
    double* seis;
    int nsamples, sample_every;
    const double* epart;
    int nepart;
    double* energy;
    int nenergy, energy_every;
    double evol;             // cell volume
    const double* wavelet;
    int nwavelet;
    int src_kind, src, nx, nxy;  // nxy: plane stride of the device layout
    double sin_t, cos_t, dt;
    int* step;               // steps completed (device counter)
    int* flag;               // set to 1 when a sample or the energy is not finite
    float* scl;              // FP16 storage scale table (see kScaleTable)
    float* side;             // FP16 with a stress source: the source cell's normal stresses (see StepArgs)
    unsigned* maxslot;       // [kMaxSlots][27] maxima folded by this step's stencil kernels
    int rescale_every;       // FP16 storage: re-choose every scale each N steps
    double* srctrace;        // [nstep+1][3] sigma_xx, sigma_yy, sigma_zz at the source after injection
    int nsrctrace;           // rows of srctrace
};

template <typename T>
void launch_stress(const StepArgs<T>& a, hipStream_t s);
template <typename T>
void launch_velocity(const StepArgs<T>& a, bool energy, hipStream_t s);
// fused step over the interior box; needs double-buffered fields and the
// absorbing-layer stress launch of the same step to have run.  A workgroup has
// `tile` threads (at most kFusedMaxThreads, kFusedMaxThreadsF64 in FP64), one per
// cell of its region: tile - 2*nx own cells and one halo row below and above.
constexpr int kFusedMaxThreads = 1024;
constexpr int kFusedMaxThreadsF64 = 512;
template <typename T>
void launch_fused(const StepArgs<T>& a, bool energy, hipStream_t s);
// LDS bytes of a fused launch (0: not launchable with these sizes: tile not a
// multiple of 64 in [64, max], or fewer than 64 own cells)
template <typename T>
size_t fused_lds_bytes(int nx, int tile);
template <typename T>
void launch_tail(const TailArgs<T>& a, hipStream_t s);
// A batch of shots stepped together (each a merged launch over all its boxes,
// in-place, no energy): args[s] in device memory; shot s owns the logical
// blocks [blk0[s], blk0[s+1]) of one launch of blk0[nshot] blocks of `block`
// threads, VEC 1.  The tails run one workgroup per shot.
template <typename T>
void launch_stress_multi(const StepArgs<T>* args, const int* blk0, int nshot, int nblocks, int block, hipStream_t s);
template <typename T>
void launch_velocity_multi(const StepArgs<T>* args, const int* blk0, int nshot, int nblocks, int block,
                           hipStream_t s);
template <typename T>
void launch_tail_multi(const TailArgs<T>* args, int nshot, hipStream_t s);
constexpr int kMaxBatch = 64;  // shots in one batch launch
void launch_copy(const float4* src, float4* dst, long n4, int variant, hipStream_t s);
// zero `bytes` bytes at p (p 16-byte aligned) with a kernel on stream s
void launch_zero(void* p, size_t bytes, hipStream_t s);
// Halo of a z interface: nf planes, each the first n elements at fields[q],
// copied to (pack) or from (unpack) the contiguous buffer buf[nf][n].
template <typename T>
void launch_halo(T* const* fields, int nf, long n, T* buf, bool pack, hipStream_t s);
// First level of the energy reduction on energy steps: the n per-block partial
// pairs in `in` are summed in a fixed order by `nout` workgroups into
// out[nout][2], which the tail kernel then sums (kEpartBlocks at most).
constexpr int kEpartBlocks = 256;
constexpr int kEpartDirect = 2048;  // up to this many partials the tail sums them itself
void launch_epart(const double* in, int n, double* out, int nout, hipStream_t s);
// multi-stream calibration (variants 4-7): arrays of n floats packed back to back at base
constexpr long kStreamBlock = 1024;  // floats per block of the interleaved variant 8
int stream_counts(int variant, int& nreads, int& nwrites);  // returns floats per lane
void launch_streams(float* base, long n, int variant, hipStream_t s);

}  // namespace ela3d
