// shot.hpp -- one forward shot resident on one GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "core.hpp"
#include "kernels.hpp"

struct ela3d_shot;  // opaque C handle, defined in shot.cpp

namespace ela3d {

#define HIP_CHECK(expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::ela3d::Error{e_ == hipErrorOutOfMemory ? ELA3D_E_OOM : ELA3D_E_HIP,   \
                                 std::string(#expr) + ": " + hipGetErrorString(e_)};      \
    } while (0)

// Shape of one launch (the stress and velocity passes share it).
struct LaunchGeom {
    Box box[kMaxBox];
    int nbox = 0, nblocks = 0, block = 256, vec = 1;
    int mode = kInterior;  // kInterior, kSlab, kMixed or kFused
    int tile = 0;          // kFused: cells per workgroup
};

// Launch tuning chosen for a grid (see tune.cpp): the interior launch and the
// absorbing-layer launch of each pass are shaped independently.
struct Tuning {
    int block = 256;        // interior launch
    int zseg_interior = 16;
    int vec = 2;            // consecutive xy positions per lane (1, 2, 4)
    int block_slab = 256;   // absorbing-layer launch
    int zseg_slab = 8;
    int vec_slab = 1;       // 1 or 2
    int overlap = 0;        // 1: absorbing-layer launches concurrent with the interior ones
    int merged = 0;         // 1: one mixed launch per pass over all boxes (interior shape;
                            //    z layers with zseg_slab)
    int dbuf = 0;           // 1: two field sets, each step reads one and writes the other
    int fused = 0;          // 1: one fused stress+velocity launch over the interior box
                            //    (needs dbuf; zseg_interior planes per workgroup)
    int tile = 1024;        // fused: interior cells of the flattened plane per workgroup
    int strips = 0;         // fused: xy strips of the tile order (0: the default of tile_of)
};

// Device layout of the 9 fields and the 3 model arrays: planes of plane_stride
// elements (nx*ny rounded up to 64) between zeroed guards of guard_elems.
int64_t plane_stride(int nx, int ny);
int64_t guard_elems(int nx);

struct ArenaLayout {
    size_t f[9], m[18], f2[9], m2[18], model[3], tab[3];
    size_t wav, rec, seis, energy, epart, epart2, ctl, scl, slot, srctrace, tabf, side, total;
    int epart_cap;
};
ArenaLayout arena_layout(const ShotPlan& p, int elem_bytes, int nepart, bool dbuf);

// A shot's share of a batch launch (FP32 storage, in place, no energy sums):
// its merged-launch arguments for both passes, its tail, and its block count.
struct BatchMember {
    StepArgs<float> stress, velocity;
    TailArgs<float> tail;
    int nblocks = 0;
};

class Shot {
public:
    virtual ~Shot() = default;
    virtual void enqueue_steps(int n) = 0;          // asynchronous
    // n steps driven the way a plain host loop drives them: per step a blocking
    // host-to-device copy of the step index, eager launches, and a blocking
    // readback of that step's receiver samples (synchronous).
    virtual void step_plain(int n) = 0;
    virtual int sync() = 0;                         // ELA3D_OK or ELA3D_E_UNSTABLE
    // queue a copy of the instability flag into pinned host memory; after the
    // stream has passed that point, unstable_seen() reports it without a sync
    virtual void enqueue_flag_readback(hipStream_t on = nullptr) = 0;  // on: another stream
    // Arguments of this shot as a member of a batch launch of `block` threads
    // (one merged launch per pass over all its boxes); false if the shot cannot
    // be batched (not FP32, double-buffered or with energy sums).
    virtual bool batch_member(int block, BatchMember& out) = 0;
    virtual bool unstable_seen() const = 0;
    virtual void seismograms(double* out) = 0;      // nsamples*nrec*3
    virtual void energy(double* out) = 0;           // nenergy*2
    virtual void source_trace(double* out) = 0;     // (nstep+1)*3
    // raw access to the whole device state (every buffer of the single allocation,
    // device_bytes() long) for checkpoints; set_steps_done after a full restore
    virtual void read_state(size_t offset, size_t bytes, void* host) = 0;
    virtual void write_state(size_t offset, size_t bytes, const void* host) = 0;
    virtual void set_steps_done(int n) = 0;
    virtual void snapshot(int field, double* out) = 0;
    virtual void set_field(int field, const double* in) = 0;
    // n steps; every `every`-th one eager with hipEvents around each launch (the
    // per-launch totals), the others graph replays; wall_ms spans all n
    virtual void profile(int n, std::vector<ela3d_kernel_stats>& st, double& wall_ms, int every = 1) = 0;
    virtual void info(ela3d_info& out) = 0;
    virtual void retune(const Tuning& t) = 0;       // new launch shape (drops captured graphs)
    virtual hipStream_t stream() const = 0;
    // run the absorbing-layer launches concurrently with the interior ones (default)
    // or in sequence on the main stream (drops captured graphs)
    virtual void set_overlap(bool on) = 0;
    // One half of a step of the in-place two-pass schedule, eagerly: phase 0 the
    // stress pass, phase 1 the velocity pass and the tail (receivers, energy,
    // source).  A z-slab runs its steps this way, with halo exchanges between.
    virtual void enqueue_phase(int phase) = 0;
    // Halo of a z-slab after `phase`: the fields it sends to its neighbour on
    // `side` (0 below, 1 above) from its boundary plane on that side, or (unpack)
    // receives from that neighbour into its ghost plane there.  `buf` is device
    // memory of halo_bytes(phase); the copy is enqueued on the shot's stream.
    virtual size_t halo_bytes(int phase, int side) const = 0;
    virtual void halo_pack(int phase, int side, void* buf) = 0;
    virtual void halo_unpack(int phase, int side, const void* buf) = 0;
    virtual int steps_done() const = 0;
    virtual int64_t device_bytes() const = 0;
    const ShotPlan& plan() const { return plan_; }

protected:
    ShotPlan plan_;
};

// Device memory a shot is placed in instead of an allocation of its own: an
// arena the caller owns (the allocation-free entry point's static buffer).  The
// shot zeroes the part it uses with a kernel and never frees it.
struct ShotMem {
    void* arena = nullptr;  // 256-byte aligned; nullptr: the shot allocates its own
    size_t bytes = 0;       // usable bytes at arena
    bool inplace = false;   // one field set: no fused, double-buffered or overlapped schedule
    int* host_flag = nullptr;  // pinned host int for the instability-flag readback
                               // (nullptr: the shot allocates one when first asked)
};

// double_buffer: allocate a second field set even if the chosen tuning does not
// need one (the tuner's scratch shot, which tries every schedule)
std::unique_ptr<Shot> make_shot(ShotPlan plan, const float* cp, const float* cs, const float* rho,
                                hipStream_t external_stream = nullptr, bool double_buffer = false,
                                const ShotMem& mem = ShotMem());
// Device memory the library holds (shot arenas, the server's pool, halo staging)
// is taken and returned through these.  A process-wide limit on the bytes held at
// once (ela3d_set_device_limit) lets tests force ELA3D_E_OOM; device_alloc throws
// Error{ELA3D_E_OOM} past it or when hipMalloc fails.
void* device_alloc(size_t bytes);
void device_free(void* p, size_t bytes);
void set_device_limit(int64_t bytes);  // <= 0: no limit

// device address of the allocation-free entry point's static buffer on `device`
// (ELA3D_NOALLOC_BYTES long; looked up once per device)
void* noalloc_scratch(int device);
Tuning choose_tuning(const ShotPlan& p);
bool cached_tuning(const ela3d_config& c, Tuning& out);
void store_tuning(const ela3d_config& c, const Tuning& t, double ms);
// launch shapes of the absorbing-layer boxes (slab) and the interior box
void geometry(const ShotPlan& p, const Tuning& t, LaunchGeom& slab, LaunchGeom& interior);
// whether the fused schedule can run this plan with this tile (empty: yes)
std::string fused_unsupported(const ShotPlan& p, int tile);

}  // namespace ela3d
