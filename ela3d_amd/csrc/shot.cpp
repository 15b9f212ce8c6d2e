// shot.cpp -- device-resident shot: buffers, launch geometry, step replay, readback.
//
// HBM layout of one shot (x-fastest, 0-based, offset = i + j*nx + k*nx*ny):
//   9 wavefields  [nz][ny][nx]  storage type T
//   3 model arrays cp, cs, rho  [nz][ny][nx]  FP32
//   18 C-PML memory variables in compact slabs holding only the absorbing layers:
//     x family (d/dx: ids 9,12,14,18,21,24)  [nz][ny][wx]
//     y family (d/dy: ids 10,13,16,19,22,25) [nz][wy][nx]
//     z family (d/dz: ids 11,15,17,20,23,26) [wz][ny][nx]
//   where an axis of n points keeps indices [0,lo) u [hi,n) (width w = lo+n-hi).
//   FP64: C-PML a/b/K tables [3 axes][6][n], wavelet [nstep], seismograms
//   [nsamples][nrec][3], energy [nenergy][2], energy partials.
#include "shot.hpp"
#include "geom.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace ela3d {

// axis family of each memory variable (index 0..17 -> field id 9..26)
static const int kMemAxis[18] = {0, 1, 2, 0, 1, 0, 2, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2};

namespace {
std::atomic<int64_t> g_device_held{0};
std::atomic<int64_t> g_device_limit{0};  // 0: no limit
}  // namespace

void set_device_limit(int64_t bytes) { g_device_limit = bytes > 0 ? bytes : 0; }

void* device_alloc(size_t bytes)
{
    const int64_t lim = g_device_limit.load();
    const int64_t held = g_device_held.fetch_add(int64_t(bytes)) + int64_t(bytes);
    if (lim > 0 && held > lim) {
        g_device_held -= int64_t(bytes);
        throw Error{ELA3D_E_OOM, "device memory limit: " + std::to_string(bytes) + " more bytes would exceed " +
                                     std::to_string(lim)};
    }
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_device_held -= int64_t(bytes);
        throw Error{e == hipErrorOutOfMemory ? ELA3D_E_OOM : ELA3D_E_HIP,
                    std::string("hipMalloc: ") + hipGetErrorString(e)};
    }
    return p;
}

void device_free(void* p, size_t bytes)
{
    if (!p) return;
    (void)hipFree(p);
    g_device_held -= int64_t(bytes);
}

Tuning choose_tuning(const ShotPlan& p)
{
    Tuning t;
    const ela3d_config& c = p.cfg;
    if (c.block == 0 && c.zseg == 0 && cached_tuning(c, t)) return t;
    t.block = c.block > 0 ? c.block : 256;
    t.block_slab = t.block;
    t.vec = 1;
    t.vec_slab = 1;
    if (c.zseg > 0) {
        t.zseg_interior = c.zseg;
        t.zseg_slab = c.zseg;
        return t;
    }
    // enough workgroups to fill 256 CUs several times over; short z-segments
    // measured best on MI355X (more waves in flight beats longer register queues)
    const int64_t nchunk = (int64_t(c.nx) * (p.ax[1].hi - p.ax[1].lo) + t.block * t.vec - 1) / (t.block * t.vec);
    const int nzi = std::max(1, p.ax[2].hi - p.ax[2].lo);
    int zs = 8;
    while (zs > 2 && nchunk * ((nzi + zs - 1) / zs) < 8192) zs /= 2;
    t.zseg_interior = zs;
    t.zseg_slab = 4;
    return t;
}

int64_t plane_stride(int nx, int ny) { return (int64_t(nx) * ny + 63) / 64 * 64; }
int64_t guard_elems(int nx) { return (int64_t(nx) + 64 + 63) / 64 * 64; }

static void add_box(LaunchGeom& g, const BoxRange& r, int zseg)
{
    Box& b = g.box[g.nbox++];
    const int per = g.block * g.vec;
    b.p0 = r.p0; b.p1 = r.p1; b.k0 = r.k0; b.k1 = r.k1;
    b.nchunk = (r.p1 - r.p0 + per - 1) / per;
    b.zseg = std::max(1, zseg);
    b.blk0 = g.nblocks;
    b.mode = r.mode == 0 ? kInterior : kSlab;
    b.nstrip = 0;
    g.nblocks += b.nchunk * ((r.k1 - r.k0 + b.zseg - 1) / b.zseg);
}

// Two launches per pass: the absorbing-layer launch covers the z-min, z-max,
// y-min and y-max boxes, the interior launch the interior box.  Merged: one
// mixed launch with the absorbing-layer boxes first in block order (their
// slower blocks start early), then the interior box.  Box starts are
// multiples of 64 elements, so with the padded plane stride every workgroup
// tile begins on a 256-byte line boundary and VEC groups are aligned.
void geometry(const ShotPlan& p, const Tuning& t, LaunchGeom& slab, LaunchGeom& interior)
{
    slab = LaunchGeom();
    interior = LaunchGeom();
    slab.block = t.block_slab;
    slab.vec = t.vec_slab;
    slab.mode = kSlab;
    interior.block = t.block;
    interior.vec = t.vec;
    interior.mode = t.fused ? kFused : t.merged ? kMixed : kInterior;
    if (t.fused) {  // one workgroup per (tile, z segment): `tile` threads, one per region cell
        interior.block = t.tile;
        interior.vec = 1;
        interior.tile = t.tile;
    }
    BoxRange r[5];
    const int n = split_boxes(p, r);
    for (int q = 0; q < n; ++q) {  // absorbing-layer boxes come first in split_boxes order
        const bool zlayer = r[q].k1 <= p.ax[2].lo || r[q].k0 >= p.ax[2].hi;
        if (r[q].mode == 0) {
            if (t.fused) {
                Box& b = interior.box[interior.nbox++];
                b.p0 = r[q].p0; b.p1 = r[q].p1; b.k0 = r[q].k0; b.k1 = r[q].k1;
                const int own = fused_own(t.tile, p.cfg.nx);  // own cells per workgroup
                b.nchunk = (r[q].p1 - r[q].p0 + own - 1) / own;
                b.zseg = std::max(1, t.zseg_interior);
                b.blk0 = 0;
                b.mode = kInterior;
                // tile order: with two workgroups per CU (tiles of up to 768 threads),
                // 4 strips, each marched by two XCDs, keep more of an XCD's workgroups on
                // the same z segment than one strip per XCD, so the halo rows they share
                // are re-read from its L2 (classic, 768-thread tiles: 867 -> 858 us);
                // with 1,024-thread tiles the default order is faster (885 against 895 us)
                const int ns = t.strips > 0 ? t.strips : (t.tile <= 768 && b.nchunk >= 32 ? 4 : 0);
                b.nstrip = std::max(0, std::min(ns, b.nchunk));
                interior.nblocks = b.nchunk * ((b.k1 - b.k0 + b.zseg - 1) / b.zseg);
            } else {
                add_box(interior, r[q], t.zseg_interior);
            }
        } else {  // z layers march in short segments; the y layers as the interior (with the
                  // fused interior step, whose segments are long, as the z layers)
            add_box(t.merged ? interior : slab, r[q], (zlayer || t.fused) ? t.zseg_slab : t.zseg_interior);
        }
    }
}

std::string fused_unsupported(const ShotPlan& p, int tile)
{
    const ela3d_config& c = p.cfg;
    if (c.storage == ELA3D_STORAGE_FP16) return "the fused schedule runs FP32 and FP64 storage only";
    // (a free surface at k = 0 is a one-plane box of the absorbing-layer launches)
    if (p.ax[1].lo < 1 || (p.ax[2].lo < 1 && !c.free_surface) || p.ax[1].hi > c.ny - 1 || p.ax[2].hi > c.nz - 1)
        return "the fused schedule needs absorbing layers on every y and z face";
    {  // both copies of an array must lie within 4 GiB of its first byte (32-bit offsets)
        const size_t elem = storage_bytes(c.storage);
        const size_t arr = size_t(plane_stride(c.nx, c.ny)) * c.nz + 2 * size_t(guard_elems(c.nx));
        if (2 * (arr * elem + 256) >= (size_t(1) << 32)) return "the fused schedule needs field arrays below 2 GiB";
    }
    const size_t lds = c.storage == ELA3D_STORAGE_FP64 ? fused_lds_bytes<double>(c.nx, tile)
                                                       : fused_lds_bytes<float>(c.nx, tile);
    if (lds == 0)
        return "fused tile (threads per workgroup) must be a multiple of 64, at most 1024 (512 in FP64), "
               "and leave at least 64 own cells after two halo rows of nx";
    BoxRange r[5];
    const int n = split_boxes(p, r);
    for (int q = 0; q < n; ++q)
        if (r[q].mode == 0) return "";
    return "the grid has no interior box";
}

static size_t mem_elems(const ShotPlan& p, int axis)
{
    const ela3d_config& c = p.cfg;
    const PmlAxis& A = p.ax[axis];
    if (axis == 0) return size_t(c.nz) * c.ny * A.width();
    if (axis == 1) return size_t(c.nz) * A.width() * c.nx;
    return size_t(A.width()) * c.ny * c.nx;
}

// Byte offsets of every buffer in a shot's single device allocation.  Each
// buffer starts on a 256-byte boundary; the energy-partial buffer has room for
// any launch shape the tuner may pick (block >= 64, vec >= 1, zseg >= 1).
ArenaLayout arena_layout(const ShotPlan& p, int elem, int nepart, bool dbuf)
{
    const ela3d_config& c = p.cfg;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    ArenaLayout L;
    const int64_t nxyp = plane_stride(c.nx, c.ny);
    const size_t padded = size_t(nxyp) * c.nz + 2 * size_t(guard_elems(c.nx));  // guard | planes | guard
    // The second set of a double-buffered shot sits right after the first copy of
    // each array, one array size above it: the fused kernel addresses both sets
    // from one base pointer per array with a 32-bit offset.
    size_t off = 0;
    for (int q = 0; q < 9; ++q) {
        L.f[q] = off;
        off += al(padded * elem);
        L.f2[q] = L.f[q];
        if (dbuf) { L.f2[q] = off; off += al(padded * elem); }
    }
    for (int q = 0; q < 18; ++q) {
        L.m[q] = off;
        off += al(mem_elems(p, kMemAxis[q]) * elem);
        L.m2[q] = L.m[q];
        if (dbuf) { L.m2[q] = off; off += al(mem_elems(p, kMemAxis[q]) * elem); }
    }
    for (int q = 0; q < 3; ++q) { L.model[q] = off; off += al(padded * sizeof(float)); }
    const int ns[3] = {c.nx, c.ny, c.nz};
    for (int q = 0; q < 3; ++q) { L.tab[q] = off; off += al(6 * size_t(ns[q]) * sizeof(double)); }
    L.wav = off; off += al(std::max<size_t>(1, p.wavelet.size()) * sizeof(double));
    L.rec = off; off += al(std::max<size_t>(1, p.rec.size()) * sizeof(int));
    const size_t nseis = size_t(p.nsamples) * c.nrec * 3;
    L.seis = off; off += al(std::max<size_t>(1, nseis) * sizeof(double));
    L.energy = off; off += al(std::max<size_t>(1, size_t(p.nenergy) * 2) * sizeof(double));
    L.epart_cap = int((nxyp / 64 + kMaxBox + 1) * int64_t(c.nz) + 8);
    L.epart = off; off += al(size_t(std::max(L.epart_cap, nepart)) * 2 * sizeof(double));
    L.epart2 = off; off += al(size_t(kEpartBlocks) * 2 * sizeof(double));
    L.ctl = off; off += al(2 * sizeof(int));
    L.scl = off; off += al(kScaleTable * sizeof(float));
    L.slot = off; off += al(size_t(kMaxSlots) * 27 * sizeof(unsigned));
    L.srctrace = off; off += al((size_t(c.nstep) + 1) * 3 * sizeof(double));
    L.tabf = off; off += al(6 * size_t(c.nx + c.ny) * sizeof(float));
    L.side = off; off += al(3 * sizeof(float));
    L.total = off;
    return L;
}

// FP16 storage always updates in place: its kernels store through the input
// pointers, and it has no fused schedule, the one that needs two field sets.
static bool double_buffers(const ShotPlan& p, bool asked)
{
    return asked && p.cfg.storage != ELA3D_STORAGE_FP16;
}

int64_t shot_device_bytes(const ShotPlan& p)
{
    const Tuning t = choose_tuning(p);
    return int64_t(arena_layout(p, storage_bytes(p.cfg.storage), 0, double_buffers(p, t.dbuf != 0)).total);
}

template <typename T>
class ShotImpl final : public Shot {
public:
    ShotImpl(ShotPlan plan, const float* cp, const float* cs, const float* rho, hipStream_t ext, bool dbuf,
             const ShotMem& mem)
    {
        plan_ = std::move(plan);
        mem_ = mem;
        const ela3d_config& c = plan_.cfg;
        HIP_CHECK(hipSetDevice(c.device));
        device_ = c.device;
        try {
            init(cp, cs, rho, ext, dbuf);
        } catch (...) {  // a throwing constructor runs no destructor: free what was made
            release();
            throw;
        }
    }

    ~ShotImpl() override { release(); }

private:
    void init(const float* cp, const float* cs, const float* rho, hipStream_t ext, bool dbuf)
    {
        if (ext || mem_.arena) {  // a shot in a caller's arena runs on the caller's stream (0: default)
            stream_ = ext;
            own_stream_ = false;
        } else {
            HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
            own_stream_ = true;
        }
        tune_ = choose_tuning(plan_);
        if (tune_.fused && !fused_unsupported(plan_, tune_.tile).empty()) tune_.fused = 0;
        if (mem_.inplace) {
            tune_.fused = 0;
            tune_.dbuf = 0;
            tune_.overlap = 0;
            dbuf = false;
        }
        dbuf_ = double_buffers(plan_, dbuf || tune_.dbuf != 0 || tune_.fused != 0);
        tune_.dbuf = dbuf_ ? 1 : 0;
        overlap_ = tune_.overlap != 0;
        if (overlap_) make_aux();
        geometry(plan_, tune_, geo_[1], geo_[0]);
        allocate();
        upload(cp, cs, rho);
        initial_scales(cp, rho);
        set_args();
    }

    void release()
    {
        (void)hipSetDevice(device_);
        if (stream_ || mem_.arena) (void)hipStreamSynchronize(stream_);
        if (aux_) (void)hipStreamSynchronize(aux_);
        for (auto& g : graphs_)
            if (g) (void)hipGraphExecDestroy(g);
        if (arena_ && !mem_.arena) device_free(arena_, size_t(arena_bytes_));
        if (h_flag_ && h_flag_ != mem_.host_flag) (void)hipHostFree(h_flag_);
        if (aux_) (void)hipStreamDestroy(aux_);
        for (auto& e : fork_)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : join_)
            if (e) (void)hipEventDestroy(e);
        if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    }

public:
    hipStream_t stream() const override { return stream_; }
    void set_overlap(bool on) override
    {
        if (on == overlap_) return;
        if (on && mem_.inplace) throw Error{ELA3D_E_INVALID, "this shot runs its launches in sequence"};
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (on) make_aux();
        overlap_ = on;
        tune_.overlap = on ? 1 : 0;
        drop_graphs();
    }
    int steps_done() const override { return host_step_; }

    void enqueue_phase(int phase) override
    {
        HIP_CHECK(hipSetDevice(device_));
        if (tune_.fused || dbuf_ || overlap_)
            throw Error{ELA3D_E_STATE, "stepping by phases needs the in-place two-pass schedule"};
        const int it = host_step_ + 1;
        const bool e = energy_step(it);
        if (phase == 0) {
            launch_stress(args_[0][1][0], stream_);
            launch_stress(args_[0][0][0], stream_);
        } else if (phase == 1) {
            launch_velocity(args_[0][1][1], e, stream_);
            launch_velocity(args_[0][0][1], e, stream_);
            if (e) launch_epart(d_epart_, nepart_, d_epart2_, epart_blocks(), stream_);
            launch_tail(tail_args(0), stream_);
            host_step_ = it;
        } else {
            throw Error{ELA3D_E_INVALID, "phase must be 0 (stress pass) or 1 (velocity pass and tail)"};
        }
        HIP_CHECK(hipGetLastError());
    }

    // fields that cross a z interface after `phase`, sent towards higher k (up) or lower k
    static int halo_fields(int phase, bool up, int (&q)[3])
    {
        if (phase == 0) {  // sigma_xz, sigma_yz feed the plane above; sigma_zz the plane below
            if (up) { q[0] = 7; q[1] = 8; return 2; }
            q[0] = 5;
            return 1;
        }
        if (up) { q[0] = 2; return 1; }  // vz feeds the plane above; vx, vy the plane below
        q[0] = 0; q[1] = 1;
        return 2;
    }

    size_t halo_bytes(int phase, int side) const override
    {
        int q[3];
        const int nf = halo_fields(phase, side == 1, q);
        return size_t(nf) * size_t(plan_.cfg.nx) * plan_.cfg.ny * sizeof(T);
    }

    // plane of the halo copy and its fields: pack sends from the boundary plane on
    // `side`; unpack writes the ghost plane on `side`, with the fields the
    // neighbour there sends towards this slab
    int halo_plane(int phase, int side, bool pack, T* (&f)[3], int& nf)
    {
        if (phase != 0 && phase != 1) throw Error{ELA3D_E_INVALID, "phase must be 0 or 1"};
        if (side != 0 && side != 1) throw Error{ELA3D_E_INVALID, "side must be 0 (below) or 1 (above)"};
        const bool has = side == 0 ? plan_.own_k0 > 0 : plan_.own_k1 < plan_.cfg.nz;
        if (!has) throw Error{ELA3D_E_RANGE, "no neighbour slab on that side"};
        const bool up = pack ? side == 1 : side == 0;
        int q[3];
        nf = halo_fields(phase, up, q);
        const int k = pack ? (side == 1 ? plan_.own_k1 - 1 : plan_.own_k0) : (side == 1 ? plan_.own_k1 : plan_.own_k0 - 1);
        for (int e = 0; e < nf; ++e) f[e] = d_f_[cur()][q[e]] + size_t(k) * size_t(nxyp_);
        return k;
    }

    void halo_pack(int phase, int side, void* buf) override
    {
        HIP_CHECK(hipSetDevice(device_));
        T* f[3];
        int nf = 0;
        halo_plane(phase, side, true, f, nf);
        launch_halo<T>(f, nf, long(plan_.cfg.nx) * plan_.cfg.ny, static_cast<T*>(buf), true, stream_);
        HIP_CHECK(hipGetLastError());
    }

    void halo_unpack(int phase, int side, const void* buf) override
    {
        HIP_CHECK(hipSetDevice(device_));
        T* f[3];
        int nf = 0;
        halo_plane(phase, side, false, f, nf);
        launch_halo<T>(f, nf, long(plan_.cfg.nx) * plan_.cfg.ny, static_cast<T*>(const_cast<void*>(buf)), false,
                       stream_);
        HIP_CHECK(hipGetLastError());
    }
    int64_t device_bytes() const override { return arena_bytes_; }

    // one step the production way: a replay of its captured graph (captured on
    // first use), or eager launches when graphs are off
    void graph_step()
    {
        const int it = host_step_ + 1;
        const bool e = energy_step(it);
        const int par = cur();
        if (plan_.cfg.use_graph) {
            hipGraphExec_t& g = graphs_[2 * par + (e ? 1 : 0)];
            if (!g) g = capture(e, par);
            HIP_CHECK(hipGraphLaunch(g, stream_));
        } else {
            launch_step(e, nullptr, par);
        }
        host_step_ = it;
    }

    void enqueue_steps(int n) override
    {
        HIP_CHECK(hipSetDevice(device_));
        for (int s = 0; s < n; ++s) graph_step();
        HIP_CHECK(hipGetLastError());
    }

    void step_plain(int n) override
    {
        const ela3d_config& c = plan_.cfg;
        HIP_CHECK(hipSetDevice(device_));
        std::vector<double> row(size_t(std::max(1, c.nrec)) * 3);
        for (int s = 0; s < n; ++s) {
            const int done = host_step_;
            const int it = done + 1;
            HIP_CHECK(hipMemcpyAsync(d_step_, &done, sizeof(int), hipMemcpyHostToDevice, stream_));
            HIP_CHECK(hipStreamSynchronize(stream_));
            launch_step(energy_step(it), nullptr, cur());
            HIP_CHECK(hipGetLastError());
            if (c.nrec > 0 && it % c.sample_every == 0 && it / c.sample_every - 1 < plan_.nsamples) {
                const size_t r = size_t(it / c.sample_every - 1) * c.nrec * 3;
                HIP_CHECK(hipMemcpyAsync(row.data(), d_seis_ + r, size_t(c.nrec) * 3 * sizeof(double),
                                         hipMemcpyDeviceToHost, stream_));
                HIP_CHECK(hipStreamSynchronize(stream_));
            }
            host_step_ = it;
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    bool batch_member(int block, BatchMember& out) override
    {
        if constexpr (!std::is_same<T, float>::value) {
            (void)block;
            (void)out;
            return false;
        } else {
            if (dbuf_ || plan_.cfg.energy_every > 0 || plan_.koff != 0 || plan_.own_k1 != plan_.cfg.nz) return false;
            Tuning t = tune_;
            t.block = block;
            t.vec = 1;
            t.merged = 1;
            t.fused = 0;
            t.overlap = 0;
            t.dbuf = 0;
            LaunchGeom slab, all;
            geometry(plan_, t, slab, all);
            if (all.nblocks > epart_cap_) return false;
            out.stress = step_args(all, d_epart_, 0, 0);
            out.velocity = step_args(all, d_epart_, 1, 0);
            out.tail = tail_args(0);
            out.nblocks = all.nblocks;
            return true;
        }
    }

    void enqueue_flag_readback(hipStream_t on) override
    {
        if (!h_flag_) {
            if (mem_.host_flag)
                h_flag_ = mem_.host_flag;
            else
                HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_flag_), sizeof(int), hipHostMallocDefault));
            *h_flag_ = 0;
        }
        HIP_CHECK(hipMemcpyAsync(h_flag_, d_flag_, sizeof(int), hipMemcpyDeviceToHost, on ? on : stream_));
    }

    bool unstable_seen() const override { return h_flag_ && *static_cast<volatile int*>(h_flag_) != 0; }

    int sync() override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        int flag = 0;
        HIP_CHECK(hipMemcpy(&flag, d_flag_, sizeof(int), hipMemcpyDeviceToHost));
        return flag ? ELA3D_E_UNSTABLE : ELA3D_OK;
    }

    void seismograms(double* out) override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        const size_t n = size_t(plan_.nsamples) * plan_.cfg.nrec * 3;
        if (n) HIP_CHECK(hipMemcpy(out, d_seis_, n * sizeof(double), hipMemcpyDeviceToHost));
    }

    void energy(double* out) override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        const size_t n = size_t(plan_.nenergy) * 2;
        if (n) HIP_CHECK(hipMemcpy(out, d_energy_, n * sizeof(double), hipMemcpyDeviceToHost));
    }

    void source_trace(double* out) override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        HIP_CHECK(hipMemcpy(out, d_srctrace_, (size_t(plan_.cfg.nstep) + 1) * 3 * sizeof(double),
                            hipMemcpyDeviceToHost));
    }

    void read_state(size_t offset, size_t bytes, void* host) override
    {
        if (offset > size_t(arena_bytes_) || bytes > size_t(arena_bytes_) - offset)
            throw Error{ELA3D_E_RANGE, "state range outside the allocation"};
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        HIP_CHECK(hipMemcpy(host, static_cast<char*>(arena_) + offset, bytes, hipMemcpyDeviceToHost));
    }

    void write_state(size_t offset, size_t bytes, const void* host) override
    {
        if (offset > size_t(arena_bytes_) || bytes > size_t(arena_bytes_) - offset)
            throw Error{ELA3D_E_RANGE, "state range outside the allocation"};
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        HIP_CHECK(hipMemcpy(static_cast<char*>(arena_) + offset, host, bytes, hipMemcpyHostToDevice));
    }

    void set_steps_done(int n) override { host_step_ = n; }

    void snapshot(int field, double* out) override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        const ela3d_config& c = plan_.cfg;
        const size_t cells = size_t(plan_.cells);
        if (field < 9) {
            std::vector<T> h(cells);
            copy_planes(h.data(), d_f_[cur()][field], sizeof(T), hipMemcpyDeviceToHost);
            const double inv = 1.0 / current_scale(field);
            for (size_t q = 0; q < cells; ++q) out[q] = double(h[q]) * inv;
            if (side_mode() && field >= 3 && field <= 5) {  // the source cell lives in the side slot
                float v = 0.f;
                HIP_CHECK(hipMemcpy(&v, d_side_ + (field - 3), sizeof(float), hipMemcpyDeviceToHost));
                out[src_cell()] = v;
            }
            return;
        }
        const int m = field - 9, axis = kMemAxis[m];
        const size_t nm = mem_count(axis);
        std::vector<T> h(nm);
        if (nm) HIP_CHECK(hipMemcpy(h.data(), d_m_[cur()][m], nm * sizeof(T), hipMemcpyDeviceToHost));
        std::fill(out, out + cells, 0.0);
        const double inv = 1.0 / current_scale(field);
        for_each_compact(axis, [&](size_t full, size_t comp) { out[full] = double(h[comp]) * inv; });
        (void)c;
    }

    void set_field(int field, const double* in) override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        const size_t cells = size_t(plan_.cells);
        if (field < 9) {
            std::vector<T> h(cells);
            const bool side = side_mode() && field >= 3 && field <= 5;
            const size_t sc_cell = side ? src_cell() : cells;  // kept out of the FP16 array and its scale
            double mx = 0.0;
            for (size_t q = 0; q < cells; ++q)
                if (q != sc_cell) mx = std::max(mx, std::fabs(in[q]));
            const double sc = host_scale(field, mx);
            for (size_t q = 0; q < cells; ++q) h[q] = q == sc_cell ? T(0) : T(in[q] * sc);
            if (side) {
                const float v = float(in[sc_cell]);
                HIP_CHECK(hipMemcpy(d_side_ + (field - 3), &v, sizeof(float), hipMemcpyHostToDevice));
            }
            for (int b = 0; b < (dbuf_ ? 2 : 1); ++b)  // both sets: cells a step never writes keep it
                copy_planes(d_f_[b][field], h.data(), sizeof(T), hipMemcpyHostToDevice);
            return;
        }
        const int m = field - 9, axis = kMemAxis[m];
        const size_t nm = mem_count(axis);
        std::vector<T> h(nm);
        std::vector<char> covered(cells, 0);
        double mx = 0.0;
        for_each_compact(axis, [&](size_t full, size_t) { mx = std::max(mx, std::fabs(in[full])); });
        const double sc = host_scale(field, mx);
        for_each_compact(axis, [&](size_t full, size_t comp) {
            h[comp] = T(in[full] * sc);
            covered[full] = 1;
        });
        for (size_t q = 0; q < cells; ++q)
            if (!covered[q] && in[q] != 0.0)
                throw Error{ELA3D_E_INVALID,
                            "memory variables are stored only inside the C-PML layers; "
                            "non-zero value outside them"};
        for (int b = 0; b < (dbuf_ ? 2 : 1); ++b)
            if (nm) HIP_CHECK(hipMemcpy(d_m_[b][m], h.data(), nm * sizeof(T), hipMemcpyHostToDevice));
    }

    void profile(int n, std::vector<ela3d_kernel_stats>& st, double& wall_ms, int every) override
    {
        HIP_CHECK(hipSetDevice(device_));
        static const char* split[kLaunches] = {"stress_cpml", "stress_interior", "velocity_cpml",
                                               "velocity_interior", "tail"};
        static const char* merged[kLaunches] = {"stress_cpml", "stress", "velocity_cpml", "velocity", "tail"};
        static const char* fused[kLaunches] = {"stress_cpml", "fused_interior", "velocity_cpml", "-", "tail"};
        const char* const* names = tune_.fused ? fused : tune_.merged ? merged : split;
        st.assign(kLaunches, ela3d_kernel_stats());
        for (int q = 0; q < kLaunches; ++q) {
            std::snprintf(st[q].name, sizeof(st[q].name), "%s", names[q]);
            st[q].launches = 0;
            st[q].total_ms = 0.0;
        }
        const bool m = tune_.merged != 0;  // the absorbing-layer launches are then empty
        st[0].alg_bytes_per_launch = m ? 0.0 : plan_.alg_stress_slab;
        st[1].alg_bytes_per_launch = plan_.alg_stress_interior + (m ? plan_.alg_stress_slab : 0.0);
        st[2].alg_bytes_per_launch = m ? 0.0 : plan_.alg_velocity_slab;
        st[3].alg_bytes_per_launch = plan_.alg_velocity_interior + (m ? plan_.alg_velocity_slab : 0.0);
        st[4].alg_bytes_per_launch = 0;
        if (tune_.fused) {
            st[1].alg_bytes_per_launch = plan_.alg_fused_interior();
            st[3].alg_bytes_per_launch = 0;
        }
        // step s of the run is instrumented (eager launches bracketed by events)
        // when s % every == 0; the others replay the step graphs
        every = std::max(1, every);
        const int chunk = 64;         // instrumented steps per host wait
        const int E = 2 * kLaunches;  // start and end event of every launch
        std::vector<hipEvent_t> ev(size_t(chunk) * E);
        for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
        hipEvent_t t0, t1;
        HIP_CHECK(hipEventCreate(&t0));
        HIP_CHECK(hipEventCreate(&t1));
        HIP_CHECK(hipStreamSynchronize(stream_));
        HIP_CHECK(hipEventRecord(t0, stream_));
        int done = 0;
        while (done < n) {
            int m = 0;  // instrumented steps of this chunk
            while (done < n && (m < chunk || done % every != 0)) {
                if (done % every == 0) {
                    const int it = host_step_ + 1;
                    launch_step(energy_step(it), &ev[size_t(m) * E], cur());
                    host_step_ = it;
                    ++m;
                } else {
                    graph_step();
                }
                ++done;
            }
            HIP_CHECK(hipGetLastError());
            if (m == 0) continue;
            HIP_CHECK(hipEventSynchronize(ev[size_t(m - 1) * E + E - 1]));
            for (int s = 0; s < m; ++s)
                for (int q = 0; q < kLaunches; ++q) {
                    float ms = 0.f;
                    HIP_CHECK(hipEventElapsedTime(&ms, ev[size_t(s) * E + 2 * q], ev[size_t(s) * E + 2 * q + 1]));
                    st[q].total_ms += ms;
                    st[q].launches += 1;
                }
        }
        HIP_CHECK(hipEventRecord(t1, stream_));
        HIP_CHECK(hipEventSynchronize(t1));
        float w = 0.f;
        HIP_CHECK(hipEventElapsedTime(&w, t0, t1));
        wall_ms = w;
        for (auto& e : ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    }

    void retune(const Tuning& t) override
    {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        if ((t.dbuf != 0 || t.fused != 0) && !dbuf_)
            throw Error{ELA3D_E_INVALID, "double buffering is fixed when the shot is created"};
        if (t.overlap != 0 && mem_.inplace) throw Error{ELA3D_E_INVALID, "this shot runs its launches in sequence"};
        if (t.fused) {
            const std::string why = fused_unsupported(plan_, t.tile);
            if (!why.empty()) throw Error{ELA3D_E_INVALID, why};
        }
        LaunchGeom g[2];
        geometry(plan_, t, g[1], g[0]);
        if (g[0].nblocks + g[1].nblocks > epart_cap_)
            throw Error{ELA3D_E_INVALID, "launch shape needs too many blocks"};
        tune_ = t;
        tune_.dbuf = dbuf_ ? 1 : 0;
        overlap_ = t.overlap != 0;
        if (overlap_) make_aux();
        geo_[0] = g[0];
        geo_[1] = g[1];
        nepart_ = g[0].nblocks + g[1].nblocks;
        set_args();
        drop_graphs();
    }

    void drop_graphs()
    {
        for (auto& x : graphs_)
            if (x) {
                (void)hipGraphExecDestroy(x);
                x = nullptr;
            }
    }

    void info(ela3d_info& o) override
    {
        std::memset(&o, 0, sizeof(o));
        o.cells = plan_.cells;
        o.device_bytes = arena_bytes_;
        o.alg_bytes_per_step = tune_.fused ? plan_.alg_step_fused() : plan_.alg_step();
        o.alg_bytes_stress = plan_.alg_stress_interior + plan_.alg_stress_slab;
        o.alg_bytes_velocity = plan_.alg_velocity_interior + plan_.alg_velocity_slab;
        o.pml_bytes_per_step = plan_.alg_pml;
        o.nsamples = plan_.nsamples;
        o.nenergy = plan_.nenergy;
        o.steps_done = host_step_;
        o.block = tune_.block;
        o.zseg = tune_.zseg_interior;
        o.zseg_slab = tune_.zseg_slab;
        o.vec = tune_.vec;
        o.block_slab = tune_.block_slab;
        o.vec_slab = tune_.vec_slab;
        o.overlap = overlap_ ? 1 : 0;
        o.merged = tune_.merged;
        o.dbuf = dbuf_ ? 1 : 0;
        o.fused = tune_.fused;
        o.tile = tune_.fused ? tune_.tile : 0;
        o.strips = tune_.fused ? geo_[0].box[0].nstrip : 0;  // the order in effect (0: up to 8 strips)
        for (int a = 0; a < 3; ++a) {
            o.region_lo[a] = plan_.ax[a].lo;
            o.region_hi[a] = plan_.ax[a].hi;
        }
    }

private:
    bool energy_step(int it) const
    {
        const int ee = plan_.cfg.energy_every;
        return ee > 0 && it % ee == 0 && (it / ee) <= plan_.nenergy;
    }

    size_t mem_count(int axis) const
    {
        const ela3d_config& c = plan_.cfg;
        const PmlAxis& A = plan_.ax[axis];
        if (axis == 0) return size_t(c.nz) * c.ny * A.width();
        if (axis == 1) return size_t(c.nz) * A.width() * c.nx;
        return size_t(A.width()) * c.ny * c.nx;
    }

    template <typename F>
    void for_each_compact(int axis, F fn) const
    {
        const ela3d_config& c = plan_.cfg;
        const PmlAxis& A = plan_.ax[axis];
        const int w = A.width();
        for (int k = 0; k < c.nz; ++k)
            for (int j = 0; j < c.ny; ++j)
                for (int i = 0; i < c.nx; ++i) {
                    size_t comp;
                    if (axis == 0) {
                        const int q = A.compact(i);
                        if (q < 0) continue;
                        comp = (size_t(k) * c.ny + j) * w + q;
                    } else if (axis == 1) {
                        const int q = A.compact(j);
                        if (q < 0) continue;
                        comp = (size_t(k) * w + q) * c.nx + i;
                    } else {
                        const int q = A.compact(k);
                        if (q < 0) continue;
                        comp = (size_t(q) * c.ny + j) * c.nx + i;
                    }
                    fn(size_t(i) + size_t(j) * c.nx + size_t(k) * c.nx * c.ny, comp);
                }
    }

    void allocate()
    {
        const ela3d_config& c = plan_.cfg;
        nxyp_ = plane_stride(c.nx, c.ny);
        guard_ = guard_elems(c.nx);
        nepart_ = geo_[0].nblocks + geo_[1].nblocks;
        const ArenaLayout L = arena_layout(plan_, int(sizeof(T)), nepart_, dbuf_);
        epart_cap_ = L.epart_cap;
        const size_t off = L.total;
        const size_t* o_f = L.f;
        const size_t* o_m = L.m;
        const size_t* o_model = L.model;
        const size_t* o_tab = L.tab;
        const size_t o_wav = L.wav, o_rec = L.rec, o_seis = L.seis, o_en = L.energy, o_ep = L.epart;
        const size_t o_ctl = L.ctl, o_scl = L.scl, o_slot = L.slot, o_src = L.srctrace;
        arena_bytes_ = int64_t(off);
        if (mem_.arena) {
            if (off > mem_.bytes)
                throw Error{ELA3D_E_RANGE, "shot needs " + std::to_string(off) + " bytes of device memory, " +
                                               std::to_string(mem_.bytes) + " available"};
            arena_ = mem_.arena;
        } else {
            arena_ = device_alloc(off);
        }
        char* b = static_cast<char*>(arena_);
        for (int q = 0; q < 9; ++q) {
            d_f_[0][q] = reinterpret_cast<T*>(b + o_f[q]) + guard_;
            d_f_[1][q] = reinterpret_cast<T*>(b + L.f2[q]) + guard_;
        }
        for (int q = 0; q < 18; ++q) {
            d_m_[0][q] = reinterpret_cast<T*>(b + o_m[q]);
            d_m_[1][q] = reinterpret_cast<T*>(b + L.m2[q]);
        }
        for (int q = 0; q < 3; ++q) d_model_[q] = reinterpret_cast<float*>(b + o_model[q]) + guard_;
        for (int q = 0; q < 3; ++q) d_tab_[q] = reinterpret_cast<double*>(b + o_tab[q]);
        d_wav_ = reinterpret_cast<double*>(b + o_wav);
        d_rec_ = reinterpret_cast<int*>(b + o_rec);
        d_seis_ = reinterpret_cast<double*>(b + o_seis);
        d_energy_ = reinterpret_cast<double*>(b + o_en);
        d_epart_ = reinterpret_cast<double*>(b + o_ep);
        d_epart2_ = reinterpret_cast<double*>(b + L.epart2);
        d_step_ = reinterpret_cast<int*>(b + o_ctl);
        d_flag_ = d_step_ + 1;
        d_scl_ = reinterpret_cast<float*>(b + o_scl);
        d_maxslot_ = reinterpret_cast<unsigned*>(b + o_slot);
        d_srctrace_ = reinterpret_cast<double*>(b + o_src);
        d_tabf_ = reinterpret_cast<float*>(b + L.tabf);
        d_side_ = reinterpret_cast<float*>(b + L.side);
        // zero everything: wavefields and memory variables (CPU:107-140), the guards
        // and plane padding (read as neighbours, never written), outputs
        if (mem_.arena) {  // memory the shot does not own may hold a previous user's data
            launch_zero(arena_, off, stream_);
            HIP_CHECK(hipGetLastError());
        } else {
            HIP_CHECK(hipMemsetAsync(arena_, 0, off, stream_));
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
        const std::vector<float> ones(kScaleTable, 1.0f);
        HIP_CHECK(hipMemcpy(d_scl_, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    // FP16 storage: scales before the first write, from the size each field can
    // reach near the source (one period of the peak injection; v and sigma related
    // by the impedance rho*cp, memory variables by 1/h), so that the first writes
    // fit the FP16 range.  The tail kernel then tracks the data every step.
    void initial_scales(const float* cp, const float* rho)
    {
        if (!std::is_same<T, half_t>::value) return;
        const ela3d_config& c = plan_.cfg;
        const size_t s = size_t(c.isrc) + size_t(c.jsrc) * c.nx + size_t(c.ksrc) * c.nx * c.ny;
        double peak = 0.0;
        for (double w : plan_.wavelet) peak = std::max(peak, std::fabs(w));
        const double r = rho[s], v = cp[s], imp = r * v;
        const double lobe = std::max(1.0, 1.0 / (c.f0 * c.dt));
        double sig, vel;
        if (c.src_kind == ELA3D_SRC_REFERENCE_SIGMA || c.src_kind == ELA3D_SRC_EXPLOSIVE) {
            sig = peak * c.dt * r * v * v * lobe;
            vel = sig / imp;
        } else {
            vel = peak * c.dt / r * lobe;
            sig = vel * imp;
        }
        const double h = std::min(c.dx, std::min(c.dy, c.dz));
        std::vector<float> tab(kScaleTable, 1.0f);
        for (int q = 0; q < 27; ++q) {
            const double est = q < 3 ? vel : (q < 9 ? sig : (q < 18 ? vel / h : sig / h));
            float sc = 1.0f;
            if (est > 0.0 && std::isfinite(est)) {
                int e = 0;
                std::frexp(est, &e);
                sc = std::ldexp(1.0f, std::min(100, std::max(-100, 12 - e)));
            }
            tab[q] = tab[54 + q] = sc;
            tab[27 + q] = tab[81 + q] = 1.0f / sc;
            tab[108 + q] = 0.0f;  // no previous maximum yet
        }
        HIP_CHECK(hipMemcpy(d_scl_, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    // host x-fastest nx*ny*nz  <->  device planes of stride nxyp_
    void copy_planes(void* dst, const void* src, size_t elem, hipMemcpyKind kind)
    {
        const ela3d_config& c = plan_.cfg;
        const size_t row = size_t(c.nx) * c.ny * elem, dev = size_t(nxyp_) * elem;
        const bool to_dev = kind == hipMemcpyHostToDevice;
        HIP_CHECK(hipMemcpy2D(dst, to_dev ? dev : row, src, to_dev ? row : dev, row, size_t(c.nz), kind));
    }

    void upload(const float* cp, const float* cs, const float* rho)
    {
        const ela3d_config& c = plan_.cfg;
        const float* src[3] = {cp, cs, rho};
        for (int q = 0; q < 3; ++q) copy_planes(d_model_[q], src[q], sizeof(float), hipMemcpyHostToDevice);
        for (int a = 0; a < 3; ++a) {
            const PmlAxis& A = plan_.ax[a];
            std::vector<double> t(6 * size_t(A.n));
            for (int i = 0; i < A.n; ++i) {  // the kernels multiply by 1/K (exact for K = 1)
                t[i] = A.a[i]; t[A.n + i] = A.b[i]; t[2 * A.n + i] = 1.0 / A.K[i];
                t[3 * A.n + i] = A.a_h[i]; t[4 * A.n + i] = A.b_h[i]; t[5 * A.n + i] = 1.0 / A.K_h[i];
            }
            HIP_CHECK(hipMemcpyAsync(d_tab_[a], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
            HIP_CHECK(hipStreamSynchronize(stream_));
            if (a < 2) {  // FP32 copies of the x and y tables, back to back
                const std::vector<float> f(t.begin(), t.end());
                HIP_CHECK(hipMemcpy(d_tabf_ + (a == 0 ? 0 : 6 * size_t(c.nx)), f.data(), f.size() * sizeof(float),
                                    hipMemcpyHostToDevice));
            }
        }
        if (!plan_.wavelet.empty())
            HIP_CHECK(hipMemcpyAsync(d_wav_, plan_.wavelet.data(), plan_.wavelet.size() * sizeof(double),
                                     hipMemcpyHostToDevice, stream_));
        if (!plan_.rec.empty())
            HIP_CHECK(hipMemcpyAsync(d_rec_, plan_.rec.data(), plan_.rec.size() * sizeof(int),
                                     hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        (void)c;
    }

    // Arguments of one launch for one pass (0 stress, 1 velocity) of a step that
    // starts from field set `par`: reads the current set (sigma of the velocity
    // pass from the set just written), writes the other set (the same set in place).
    StepArgs<T> step_args(const LaunchGeom& g, double* epart, int pass, int par) const
    {
        const ela3d_config& c = plan_.cfg;
        StepArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        // stencil kernels address fields and model from the allocation start
        // (guard_ elements before plane 0) so neighbour offsets never go negative
        const int in = par, out = dbuf_ ? 1 - par : par;
        for (int q = 0; q < 9; ++q) {
            const int src = (pass == 1 && q >= 3) ? out : in;
            a.fi[q] = d_f_[src][q] - guard_;
            a.fo[q] = d_f_[out][q] - guard_;
        }
        for (int q = 0; q < 18; ++q) {
            a.mi[q] = d_m_[in][q];
            a.mo[q] = d_m_[out][q];
        }
        if (g.mode == kFused) {  // one base per array, the sets at 32-bit element offsets
            const unsigned df = unsigned(d_f_[1][0] - d_f_[0][0]), dm = unsigned(d_m_[1][0] - d_m_[0][0]);
            for (int q = 0; q < 9; ++q) {
                a.fi[q] = d_f_[0][q] - guard_;
                a.fo[q] = nullptr;
            }
            for (int q = 0; q < 18; ++q) {
                a.mi[q] = d_m_[0][q];
                a.mo[q] = nullptr;
            }
            a.fdin = in ? df : 0u;
            a.fdout = out ? df : 0u;
            a.mdin = in ? dm : 0u;
            a.mdout = out ? dm : 0u;
        }
        if (std::is_same<T, half_t>::value && in != out)  // FP16 kernels store through fi/mi
            throw Error{ELA3D_E_INVALID, "FP16 storage updates in place"};
        a.cp = d_model_[0] - guard_;
        a.cs = d_model_[1] - guard_;
        a.rho = d_model_[2] - guard_;
        for (int q = 0; q < 3; ++q) a.tab[q] = d_tab_[q];
        a.tabf = d_tabf_;
        a.nx = c.nx; a.ny = c.ny; a.nz = c.nz; a.nxy = c.nx * c.ny;
        a.nxyp = int(nxyp_);
        a.g0 = int(guard_);
        for (int q = 0; q < 3; ++q) {
            a.lo[q] = plan_.ax[q].lo;
            a.hi[q] = plan_.ax[q].hi;
            a.w[q] = plan_.ax[q].width();
        }
        a.kmax_shear = plan_.kmax_shear;
        a.dirichlet = c.dirichlet ? 1 : 0;
        a.fs = (c.free_surface && plan_.koff == 0) ? 1 : 0;
        a.emargin = std::max(0, c.npoints_pml);
        a.ek0 = plan_.ek0;
        a.ek1 = plan_.ek1;
        a.dt = c.dt;
        a.idx = 1.0 / c.dx;
        a.idy = 1.0 / c.dy;
        a.idz = 1.0 / c.dz;
        a.scl = d_scl_;
        a.maxslot = d_maxslot_;
        a.side = side_mode() ? d_side_ : nullptr;
        a.sidx = unsigned(guard_) + unsigned(src_offset());
        a.epart = epart;
        for (int q = 0; q < g.nbox; ++q) a.box[q] = g.box[q];
        a.nbox = g.nbox;
        a.nblocks = g.nblocks;
        a.vec = g.vec;
        a.block = g.block;
        a.mode = g.mode;
        a.tile = g.tile;
        a.ldsreg = g.tile;
        return a;
    }

    // absorbing-layer launch first: its energy partials come first in block order
    void set_args()
    {
        for (int par = 0; par < 2; ++par)
            for (int pass = 0; pass < 2; ++pass) {
                args_[par][1][pass] = step_args(geo_[1], d_epart_, pass, par);
                args_[par][0][pass] = step_args(geo_[0], d_epart_ + 2 * size_t(geo_[1].nblocks), pass, par);
            }
    }

    // workgroups of the first energy-reduction level (0: the tail sums the
    // per-block partials itself)
    int epart_blocks() const
    {
        if (nepart_ <= kEpartDirect) return 0;
        return std::min(kEpartBlocks, (nepart_ + 1023) / 1024);
    }

    // tail of a step that starts from set `par` (works on the set that step writes)
    TailArgs<T> tail_args(int par) const
    {
        const ela3d_config& c = plan_.cfg;
        TailArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        for (int q = 0; q < 9; ++q) a.f[q] = d_f_[dbuf_ ? 1 - par : par][q];
        a.cp = d_model_[0];
        a.rho = d_model_[2];
        a.rec = d_rec_;
        a.nrec = c.nrec;
        a.seis = d_seis_;
        a.nsamples = plan_.nsamples;
        a.sample_every = c.sample_every;
        a.epart = epart_blocks() > 0 ? d_epart2_ : d_epart_;
        a.nepart = epart_blocks() > 0 ? epart_blocks() : nepart_;
        a.energy = d_energy_;
        a.nenergy = plan_.nenergy;
        a.energy_every = c.energy_every;
        a.evol = c.dx * c.dy * c.dz;
        a.wavelet = d_wav_;
        a.nwavelet = int(plan_.wavelet.size());
        a.src_kind = c.src_kind;
        a.nx = c.nx;
        a.nxy = int(nxyp_);  // plane stride of the device layout
        a.src = c.isrc + c.jsrc * c.nx + c.ksrc * int(nxyp_);
        const double th = c.angle_force * 3.141592653589793238462643 / 180.0;
        a.sin_t = std::sin(th);
        a.cos_t = std::cos(th);
        a.dt = c.dt;
        a.step = d_step_;
        a.flag = d_flag_;
        a.scl = d_scl_;
        a.maxslot = d_maxslot_;
        a.side = side_mode() ? d_side_ : nullptr;
        a.rescale_every = std::max(1, c.fp16_rescale_every);
        a.srctrace = d_srctrace_;
        a.nsrctrace = c.nstep + 1;
        return a;
    }

    // One time step: stress pass -> velocity pass -> tail (receivers, energy, source).
    // Within a pass the absorbing-layer launch runs on the auxiliary stream, forked
    // from the main stream, concurrently with the interior launch (they write
    // disjoint cells and read only what the previous pass wrote); the main stream
    // joins it before the next pass.  With overlap off both run on the main
    // stream in sequence.  With ev != nullptr, events ev[2q], ev[2q+1] bracket
    // launch q (stress layers, stress interior, velocity layers, velocity
    // interior, tail) on the stream it runs on.
    void launch_step(bool energy, hipEvent_t* ev, int par)
    {
        auto mark = [&](int q, hipStream_t s) {
            if (ev) HIP_CHECK(hipEventRecord(ev[q], s));
        };
        if (tune_.fused) {  // layers' stresses, the fused interior step, layers' velocities
            mark(0, stream_);
            launch_stress(args_[par][1][0], stream_);
            mark(1, stream_);
            mark(2, stream_);
            if constexpr (std::is_same<T, half_t>::value)
                throw Error{ELA3D_E_INVALID, "the fused schedule does not run FP16 storage"};
            else
                launch_fused(args_[par][0][0], energy, stream_);
            mark(3, stream_);
            mark(4, stream_);
            launch_velocity(args_[par][1][1], energy, stream_);
            mark(5, stream_);
            mark(6, stream_);
            mark(7, stream_);
            mark(8, stream_);
            if (energy) launch_epart(d_epart_, nepart_, d_epart2_, epart_blocks(), stream_);
            launch_tail(tail_args(par), stream_);
            mark(9, stream_);
            return;
        }
        hipStream_t side = overlap_ ? aux_ : stream_;
        for (int pass = 0; pass < 2; ++pass) {
            if (overlap_) {
                HIP_CHECK(hipEventRecord(fork_[pass], stream_));
                HIP_CHECK(hipStreamWaitEvent(aux_, fork_[pass], 0));
            }
            mark(4 * pass + 0, side);
            if (pass == 0)
                launch_stress(args_[par][1][0], side);
            else
                launch_velocity(args_[par][1][1], energy, side);
            mark(4 * pass + 1, side);
            mark(4 * pass + 2, stream_);
            if (pass == 0)
                launch_stress(args_[par][0][0], stream_);
            else
                launch_velocity(args_[par][0][1], energy, stream_);
            mark(4 * pass + 3, stream_);
            if (overlap_) {
                HIP_CHECK(hipEventRecord(join_[pass], aux_));
                HIP_CHECK(hipStreamWaitEvent(stream_, join_[pass], 0));
            }
        }
        mark(8, stream_);
        if (energy) launch_epart(d_epart_, nepart_, d_epart2_, epart_blocks(), stream_);
        launch_tail(tail_args(par), stream_);
        mark(9, stream_);
    }

    hipGraphExec_t capture(bool energy, int par)
    {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        launch_step(energy, nullptr, par);
        HIP_CHECK(hipStreamEndCapture(stream_, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_CHECK(e);
        return exec;
    }

    // second stream for the absorbing-layer launches, forked from and joined back
    // into the main stream every pass (default priority, like the main one);
    // made when an overlapped schedule is first asked for
    void make_aux()
    {
        if (aux_) return;
        HIP_CHECK(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
        for (auto& e : fork_) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto& e : join_) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }

    ShotMem mem_;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = true;
    hipStream_t aux_ = nullptr;       // absorbing-layer launches
    hipEvent_t fork_[2] = {nullptr, nullptr}, join_[2] = {nullptr, nullptr};
    bool overlap_ = false;
    Tuning tune_;
    // stress (absorbing layers, interior), velocity (same), tail
    static constexpr int kLaunches = 5;
    LaunchGeom geo_[2];   // [0] interior launch, [1] absorbing-layer launch
    StepArgs<T> args_[2][2][2];  // [field set of the step's start][launch][pass]
    void* arena_ = nullptr;
    int64_t arena_bytes_ = 0;
    T* d_f_[2][9] = {};   // two sets when double buffered; set 1 aliases set 0 in place
    T* d_m_[2][18] = {};
    bool dbuf_ = false;
    // set holding the current state: steps alternate sets when double buffered
    int cur() const { return dbuf_ ? (host_step_ & 1) : 0; }
    float* d_model_[3] = {};
    double* d_tab_[3] = {};
    double *d_wav_ = nullptr, *d_seis_ = nullptr, *d_energy_ = nullptr, *d_epart_ = nullptr;
    double* d_epart2_ = nullptr;  // first-level sums of the energy partials
    int *d_rec_ = nullptr, *d_step_ = nullptr, *d_flag_ = nullptr;
    int nepart_ = 0, epart_cap_ = 0;
    int64_t nxyp_ = 0, guard_ = 0;
    int host_step_ = 0;
    int* h_flag_ = nullptr;  // pinned mirror of d_flag_ (serving)
    double* d_srctrace_ = nullptr;   // [nstep+1][3]
    float* d_tabf_ = nullptr;        // FP32 x and y tables [6][nx], [6][ny]
    float* d_scl_ = nullptr;         // FP16 scale table (kScaleTable floats)
    float* d_side_ = nullptr;        // FP16 with a stress source: the source cell's normal stresses
    // FP16 storage with a stress source keeps the source cell's sigma_xx/yy/zz in d_side_
    bool side_mode() const
    {
        const int k = plan_.cfg.src_kind;
        return std::is_same<T, half_t>::value && (k == ELA3D_SRC_REFERENCE_SIGMA || k == ELA3D_SRC_EXPLOSIVE);
    }
    // index of the source cell in an x-fastest nx*ny*nz host array
    size_t src_cell() const
    {
        const ela3d_config& c = plan_.cfg;
        return size_t(c.isrc) + size_t(c.jsrc) * c.nx + size_t(c.ksrc) * c.nx * c.ny;
    }
    // element offset of the source cell from plane 0 of a field
    int64_t src_offset() const
    {
        const ela3d_config& c = plan_.cfg;
        return int64_t(c.isrc) + int64_t(c.jsrc) * c.nx + int64_t(c.ksrc) * nxyp_;
    }
    unsigned* d_maxslot_ = nullptr;  // [kMaxSlots][27]

    // scale of the data currently stored for field q (1 unless FP16 storage)
    double current_scale(int q)
    {
        if (!std::is_same<T, half_t>::value) return 1.0;
        float s = 1.f;
        HIP_CHECK(hipMemcpy(&s, d_scl_ + q, sizeof(float), hipMemcpyDeviceToHost));
        return s;
    }
    // FP16 storage: the scale field q is written at when the host sets it to values
    // whose largest magnitude is mx.  The current scale is kept while it holds mx in
    // [2^8, 2^13] (stored); otherwise the field moves to a scale that puts mx in
    // (2^11, 2^12], so no value overflows FP16, and the device table takes it as
    // both the current and the next scale.  mx becomes the field's previous maximum.
    double host_scale(int q, double mx)
    {
        if (!std::is_same<T, half_t>::value) return 1.0;
        double sc = current_scale(q);
        const double stored = mx * sc;
        if (mx > 0.0 && std::isfinite(mx) && (stored > 8192.0 || stored < 256.0)) {
            int e = 0;
            std::frexp(mx, &e);  // mx <= 2^e
            sc = std::ldexp(1.0, std::min(100, std::max(-100, 12 - e)));
            const float s = float(sc), inv = float(1.0 / sc);
            for (int row : {0, 54}) HIP_CHECK(hipMemcpy(d_scl_ + row + q, &s, sizeof(float), hipMemcpyHostToDevice));
            for (int row : {27, 81}) HIP_CHECK(hipMemcpy(d_scl_ + row + q, &inv, sizeof(float), hipMemcpyHostToDevice));
        }
        const float m = std::isfinite(mx) ? float(mx) : 0.f;
        HIP_CHECK(hipMemcpy(d_scl_ + 108 + q, &m, sizeof(float), hipMemcpyHostToDevice));
        return sc;
    }
    hipGraphExec_t graphs_[4] = {nullptr, nullptr, nullptr, nullptr};  // [set][energy]
};

std::unique_ptr<Shot> make_shot(ShotPlan plan, const float* cp, const float* cs, const float* rho,
                                hipStream_t ext, bool dbuf, const ShotMem& mem)
{
    switch (plan.cfg.storage) {
    case ELA3D_STORAGE_FP64:
        return std::unique_ptr<Shot>(new ShotImpl<double>(std::move(plan), cp, cs, rho, ext, dbuf, mem));
    case ELA3D_STORAGE_FP16:
        return std::unique_ptr<Shot>(new ShotImpl<half_t>(std::move(plan), cp, cs, rho, ext, dbuf, mem));
    default:
        return std::unique_ptr<Shot>(new ShotImpl<float>(std::move(plan), cp, cs, rho, ext, dbuf, mem));
    }
}

}  // namespace ela3d
