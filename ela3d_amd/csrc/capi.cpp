// capi.cpp -- extern "C" boundary: status codes, no exceptions across the ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "shot.hpp"

struct ela3d_shot {
    std::unique_ptr<ela3d::Shot> impl;
    void* halo_buf = nullptr;  // staging of ela3d_halo_exchange (this shot as the lower slab)
    size_t halo_cap = 0;
    ~ela3d_shot()
    {
        if (halo_buf) {
            impl.reset();  // its stream is drained first
            ela3d::device_free(halo_buf, halo_cap);
        }
    }
};

namespace ela3d {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace ela3d

using ela3d::Error;

template <typename F>
static int guarded(F fn)
{
    try {
        return fn();
    } catch (const Error& e) {
        ela3d::set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        ela3d::set_last_error("host allocation failed");
        return ELA3D_E_OOM;
    } catch (const std::exception& e) {
        ela3d::set_last_error(e.what());
        return ELA3D_E_INVALID;
    } catch (...) {
        ela3d::set_last_error("unknown error");
        return ELA3D_E_INVALID;
    }
}

static int need_device(int dev)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        ela3d::set_last_error("no HIP device available");
        return ELA3D_E_NODEVICE;
    }
    if (dev < 0 || dev >= n) {
        ela3d::set_last_error("device ordinal out of range");
        return ELA3D_E_INVALID;
    }
    return ELA3D_OK;
}

extern "C" {

void ela3d_config_default(ela3d_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->struct_size = sizeof(ela3d_config);
    c->abi_version = ELA3D_ABI_VERSION;
    // the homogeneous 101^3 cube workload: the reference's medium, spacing and C-PML
    // (CPU:41-78, 151-153) with a Ricker force source at f0 = 7 Hz; the reference's
    // own literals are Config.from_reference (Python) and reference=1 (CLI)
    c->nx = c->ny = c->nz = 101;
    c->dx = c->dy = c->dz = 10.0;
    c->dt = 1.6e-3;
    c->nstep = 500;
    c->npoints_pml = 10;
    c->npower = 2.0;
    c->k_max_pml = 1.0;
    c->alpha_max_pml = 0.0;  // pi*f0
    c->rcoef = 1e-3;
    c->vpml = 3000.0;
    c->f0 = 7.0;
    c->t0 = 0.0;             // 1.2/f0
    c->factor = 1e7;
    c->angle_force = 0.0;
    c->src_kind = ELA3D_SRC_FORCE;
    c->wavelet = ELA3D_WAVELET_RICKER;
    c->isrc = c->jsrc = c->ksrc = 50;
    c->nrec = 0;
    c->rec = nullptr;
    c->sample_every = 1;
    c->energy_every = 0;
    c->storage = ELA3D_STORAGE_FP32;
    c->compat_flags = 0;
    c->dirichlet = 1;
    c->fp16_rescale_every = 100;
    c->device = 0;
    c->use_graph = 1;
    c->block = 0;
    c->zseg = 0;
}

int ela3d_validate(const ela3d_config* cfg)
{
    return guarded([&] {
        if (!cfg) throw Error{ELA3D_E_INVALID, "null config"};
        ela3d::make_plan(*cfg);
        return ELA3D_OK;
    });
}

int ela3d_create(const ela3d_config* cfg, const float* cp, const float* cs, const float* rho,
                 ela3d_shot** out)
{
    return guarded([&] {
        if (!cfg || !cp || !cs || !rho || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        *out = nullptr;
        ela3d::ShotPlan plan = ela3d::make_plan(*cfg);
        ela3d::check_model(plan.cells, cp, cs, rho);
        int rc = need_device(plan.cfg.device);
        if (rc) return rc;
        auto s = new ela3d_shot;
        try {
            s->impl = ela3d::make_shot(std::move(plan), cp, cs, rho);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return ELA3D_OK;
    });
}

int ela3d_run_noalloc(const ela3d_config* cfg, const float* cp, const float* cs, const float* rho, double* seis,
                      size_t seis_cap, double* energy, size_t energy_cap, void* stream)
{
    return guarded([&] {
        if (!cfg || !cp || !cs || !rho) throw Error{ELA3D_E_INVALID, "null argument"};
        ela3d_config c = *cfg;
        c.use_graph = 0;  // graph instantiation may allocate device memory for kernel arguments
        ela3d::ShotPlan plan = ela3d::make_plan(c);
        ela3d::check_model(plan.cells, cp, cs, rho);
        const size_t nseis = size_t(plan.nsamples) * plan.cfg.nrec * 3, nen = size_t(plan.nenergy) * 2;
        if (seis_cap < nseis || (nseis && !seis)) throw Error{ELA3D_E_RANGE, "seismogram buffer too small"};
        if (energy_cap < nen || (nen && !energy)) throw Error{ELA3D_E_RANGE, "energy buffer too small"};
        int rc = need_device(plan.cfg.device);
        if (rc) return rc;
        HIP_CHECK(hipSetDevice(plan.cfg.device));
        ela3d::ShotMem mem;
        mem.arena = ela3d::noalloc_scratch(plan.cfg.device);
        mem.bytes = size_t(ELA3D_NOALLOC_BYTES);
        mem.inplace = true;
        auto shot = ela3d::make_shot(std::move(plan), cp, cs, rho, static_cast<hipStream_t>(stream), false, mem);
        shot->enqueue_steps(shot->plan().cfg.nstep);
        rc = shot->sync();
        if (nseis) shot->seismograms(seis);
        if (nen) shot->energy(energy);
        if (rc == ELA3D_E_UNSTABLE) ela3d::set_last_error("non-finite value reached a receiver or the energy");
        return rc;
    });
}

int ela3d_create_slab(const ela3d_config* cfg, int32_t k0, int32_t k1, const float* cp, const float* cs,
                      const float* rho, ela3d_shot** out, int32_t* receivers, size_t receivers_cap,
                      int32_t* nreceivers)
{
    return guarded([&] {
        if (!cfg || !cp || !cs || !rho || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        *out = nullptr;
        std::vector<int> recs;
        ela3d::ShotPlan plan = ela3d::make_slab_plan(*cfg, k0, k1, &recs);
        const int64_t nxy = int64_t(plan.cfg.nx) * plan.cfg.ny;
        ela3d::check_model(nxy * plan.nz_global, cp, cs, rho);
        if (nreceivers) *nreceivers = int32_t(recs.size());
        if (receivers) {
            if (receivers_cap < recs.size()) throw Error{ELA3D_E_RANGE, "receiver index buffer too small"};
            std::copy(recs.begin(), recs.end(), receivers);
        }
        int rc = need_device(plan.cfg.device);
        if (rc) return rc;
        const int64_t o = nxy * plan.koff;  // the slab's planes, ghosts included, of the global model
        ela3d::ShotMem mem;
        mem.inplace = true;
        auto s = new ela3d_shot;
        try {
            s->impl = ela3d::make_shot(std::move(plan), cp + o, cs + o, rho + o, nullptr, false, mem);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return ELA3D_OK;
    });
}

int ela3d_step_phase(ela3d_shot* s, int32_t phase)
{
    return guarded([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null shot"};
        s->impl->enqueue_phase(phase);
        return ELA3D_OK;
    });
}

int ela3d_halo_bytes(ela3d_shot* s, int32_t phase, int32_t side, size_t* bytes)
{
    return guarded([&] {
        if (!s || !bytes) throw Error{ELA3D_E_INVALID, "null argument"};
        if ((phase != 0 && phase != 1) || (side != 0 && side != 1)) throw Error{ELA3D_E_INVALID, "bad phase or side"};
        *bytes = s->impl->halo_bytes(phase, side);
        return ELA3D_OK;
    });
}

int ela3d_halo_pack(ela3d_shot* s, int32_t phase, int32_t side, void* dev_buf, size_t cap)
{
    return guarded([&] {
        if (!s || !dev_buf) throw Error{ELA3D_E_INVALID, "null argument"};
        if ((phase != 0 && phase != 1) || (side != 0 && side != 1)) throw Error{ELA3D_E_INVALID, "bad phase or side"};
        if (cap < s->impl->halo_bytes(phase, side)) throw Error{ELA3D_E_RANGE, "halo buffer too small"};
        s->impl->halo_pack(phase, side, dev_buf);
        return ELA3D_OK;
    });
}

int ela3d_halo_unpack(ela3d_shot* s, int32_t phase, int32_t side, const void* dev_buf, size_t n)
{
    return guarded([&] {
        if (!s || !dev_buf) throw Error{ELA3D_E_INVALID, "null argument"};
        if ((phase != 0 && phase != 1) || (side != 0 && side != 1)) throw Error{ELA3D_E_INVALID, "bad phase or side"};
        if (n < s->impl->halo_bytes(phase, 1 - side)) throw Error{ELA3D_E_RANGE, "halo buffer too small"};
        s->impl->halo_unpack(phase, side, dev_buf);
        return ELA3D_OK;
    });
}

int ela3d_halo_exchange(ela3d_shot* lower, ela3d_shot* upper, int32_t phase)
{
    return guarded([&] {
        if (!lower || !upper || lower == upper) throw Error{ELA3D_E_INVALID, "two distinct slabs needed"};
        if (phase != 0 && phase != 1) throw Error{ELA3D_E_INVALID, "phase must be 0 or 1"};
        const int dev = lower->impl->plan().cfg.device;
        if (upper->impl->plan().cfg.device != dev) throw Error{ELA3D_E_INVALID, "slabs on different devices"};
        const ela3d::ShotPlan &L = lower->impl->plan(), &U = upper->impl->plan();
        if (L.koff + L.own_k1 != U.koff + U.own_k0 || L.cfg.nx != U.cfg.nx || L.cfg.ny != U.cfg.ny ||
            L.cfg.storage != U.cfg.storage)
            throw Error{ELA3D_E_INVALID, "the slabs are not neighbours (lower's last plane + 1 = upper's first)"};
        HIP_CHECK(hipSetDevice(dev));
        const size_t up = lower->impl->halo_bytes(phase, 1), dn = upper->impl->halo_bytes(phase, 0);
        const size_t need = ((up + 255) & ~size_t(255)) + dn;
        if (lower->halo_cap < need) {
            if (lower->halo_buf) {
                HIP_CHECK(hipStreamSynchronize(lower->impl->stream()));
                ela3d::device_free(lower->halo_buf, lower->halo_cap);
                lower->halo_buf = nullptr;
                lower->halo_cap = 0;
            }
            lower->halo_buf = ela3d::device_alloc(need);
            lower->halo_cap = need;
        }
        char* bup = static_cast<char*>(lower->halo_buf);
        char* bdn = bup + ((up + 255) & ~size_t(255));
        lower->impl->halo_pack(phase, 1, bup);  // lower's top plane, sent up
        upper->impl->halo_pack(phase, 0, bdn);  // upper's bottom plane, sent down
        HIP_CHECK(hipStreamSynchronize(lower->impl->stream()));
        HIP_CHECK(hipStreamSynchronize(upper->impl->stream()));
        upper->impl->halo_unpack(phase, 0, bup);
        lower->impl->halo_unpack(phase, 1, bdn);
        HIP_CHECK(hipStreamSynchronize(lower->impl->stream()));
        HIP_CHECK(hipStreamSynchronize(upper->impl->stream()));
        return ELA3D_OK;
    });
}

int ela3d_set_device_limit(int64_t bytes)
{
    ela3d::set_device_limit(bytes);
    return ELA3D_OK;
}

int ela3d_step(ela3d_shot* s, int32_t n)
{
    return guarded([&] {
        if (!s || n < 0) throw Error{ELA3D_E_INVALID, "bad argument"};
        s->impl->enqueue_steps(n);
        return ELA3D_OK;
    });
}

int ela3d_sync(ela3d_shot* s)
{
    return guarded([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null shot"};
        int rc = s->impl->sync();
        if (rc == ELA3D_E_UNSTABLE) ela3d::set_last_error("non-finite value reached a receiver or the energy");
        return rc;
    });
}

int ela3d_run(ela3d_shot* s)
{
    return guarded([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null shot"};
        const int left = s->impl->plan().cfg.nstep - s->impl->steps_done();
        if (left > 0) s->impl->enqueue_steps(left);
        int rc = s->impl->sync();
        if (rc == ELA3D_E_UNSTABLE) ela3d::set_last_error("non-finite value reached a receiver or the energy");
        return rc;
    });
}

int ela3d_seismograms(ela3d_shot* s, double* out, size_t cap)
{
    return guarded([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null shot"};
        const auto& p = s->impl->plan();
        const size_t n = size_t(p.nsamples) * p.cfg.nrec * 3;
        if (cap < n || (n && !out)) throw Error{ELA3D_E_RANGE, "seismogram buffer too small"};
        s->impl->seismograms(out);
        return ELA3D_OK;
    });
}

int ela3d_energy(ela3d_shot* s, double* out, size_t cap)
{
    return guarded([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null shot"};
        const size_t n = size_t(s->impl->plan().nenergy) * 2;
        if (cap < n || (n && !out)) throw Error{ELA3D_E_RANGE, "energy buffer too small"};
        s->impl->energy(out);
        return ELA3D_OK;
    });
}

int ela3d_snapshot(ela3d_shot* s, int32_t field, double* out, size_t cap)
{
    return guarded([&] {
        if (!s || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        if (field < 0 || field >= ELA3D_NUM_FIELDS) throw Error{ELA3D_E_RANGE, "field id out of range"};
        if (cap < size_t(s->impl->plan().cells)) throw Error{ELA3D_E_RANGE, "snapshot buffer too small"};
        s->impl->snapshot(field, out);
        return ELA3D_OK;
    });
}

int ela3d_set_field(ela3d_shot* s, int32_t field, const double* in, size_t n)
{
    return guarded([&] {
        if (!s || !in) throw Error{ELA3D_E_INVALID, "null argument"};
        if (field < 0 || field >= ELA3D_NUM_FIELDS) throw Error{ELA3D_E_RANGE, "field id out of range"};
        if (n != size_t(s->impl->plan().cells)) throw Error{ELA3D_E_RANGE, "field size mismatch"};
        s->impl->set_field(field, in);
        return ELA3D_OK;
    });
}

int ela3d_source_trace(ela3d_shot* s, double* out, size_t cap)
{
    return guarded([&] {
        if (!s || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        if (cap < (size_t(s->impl->plan().cfg.nstep) + 1) * 3) throw Error{ELA3D_E_RANGE, "buffer too small"};
        s->impl->source_trace(out);
        return ELA3D_OK;
    });
}

namespace {

// Checkpoint file: this header, then the shot's device allocation byte for byte.
struct CheckpointHeader {
    char magic[8];
    uint32_t version;
    uint32_t storage;
    int32_t nx, ny, nz, nstep, nrec, steps_done;
    uint64_t state_bytes;
    uint64_t fingerprint;
};
const char kMagic[8] = {'E', 'L', 'A', '3', 'D', 'C', 'K', '1'};

// FNV-1a over every config value that shapes the device state
uint64_t fingerprint(const ela3d::ShotPlan& p)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* d, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(d);
        for (size_t q = 0; q < n; ++q) {
            h ^= b[q];
            h *= 1099511628211ull;
        }
    };
    const ela3d_config& c = p.cfg;
    const int32_t ints[] = {c.nx, c.ny, c.nz, c.nstep, c.npoints_pml, c.src_kind, c.wavelet, c.isrc, c.jsrc,
                            c.ksrc, c.nrec, c.sample_every, c.energy_every, c.storage, int32_t(c.compat_flags),
                            c.dirichlet, c.fp16_rescale_every};
    const double dbl[] = {c.dx, c.dy, c.dz, c.dt, c.npower, c.k_max_pml, c.alpha_max_pml, c.rcoef, c.vpml,
                          c.f0, c.t0, c.factor, c.angle_force};
    mix(ints, sizeof(ints));
    mix(dbl, sizeof(dbl));
    if (!p.rec.empty()) mix(p.rec.data(), p.rec.size() * sizeof(int32_t));
    if (c.free_surface) {  // (mixed only when set, so earlier checkpoints keep their fingerprint)
        const int32_t fs = 1;
        mix(&fs, sizeof(fs));
    }
    return h;
}

struct File {
    std::FILE* f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

const size_t kChunk = size_t(64) << 20;

}  // namespace

int ela3d_checkpoint_save(ela3d_shot* s, const char* path)
{
    return guarded([&] {
        if (!s || !path) throw Error{ELA3D_E_INVALID, "null argument"};
        const auto& p = s->impl->plan();
        CheckpointHeader h;
        std::memset(&h, 0, sizeof(h));
        std::memcpy(h.magic, kMagic, 8);
        h.version = 1;
        h.storage = uint32_t(p.cfg.storage);
        h.nx = p.cfg.nx; h.ny = p.cfg.ny; h.nz = p.cfg.nz;
        h.nstep = p.cfg.nstep;
        h.nrec = p.cfg.nrec;
        h.steps_done = s->impl->steps_done();
        h.state_bytes = uint64_t(s->impl->device_bytes());
        h.fingerprint = fingerprint(p);
        File f;
        f.f = std::fopen(path, "wb");
        if (!f.f) throw Error{ELA3D_E_IO, std::string("cannot open ") + path + " for writing"};
        if (std::fwrite(&h, sizeof(h), 1, f.f) != 1) throw Error{ELA3D_E_IO, "checkpoint header write failed"};
        std::vector<char> buf(std::min<size_t>(kChunk, h.state_bytes));
        for (size_t off = 0; off < h.state_bytes; off += buf.size()) {
            const size_t n = std::min<size_t>(buf.size(), h.state_bytes - off);
            s->impl->read_state(off, n, buf.data());
            if (std::fwrite(buf.data(), 1, n, f.f) != n) throw Error{ELA3D_E_IO, "checkpoint write failed"};
        }
        if (std::fclose(f.f) != 0) {
            f.f = nullptr;
            throw Error{ELA3D_E_IO, "checkpoint close failed"};
        }
        f.f = nullptr;
        return ELA3D_OK;
    });
}

int ela3d_checkpoint_load(ela3d_shot* s, const char* path)
{
    return guarded([&] {
        if (!s || !path) throw Error{ELA3D_E_INVALID, "null argument"};
        const auto& p = s->impl->plan();
        File f;
        f.f = std::fopen(path, "rb");
        if (!f.f) throw Error{ELA3D_E_IO, std::string("cannot open ") + path};
        CheckpointHeader h;
        if (std::fread(&h, sizeof(h), 1, f.f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1)
            throw Error{ELA3D_E_IO, "not an ela3d checkpoint"};
        if (h.fingerprint != fingerprint(p) || h.state_bytes != uint64_t(s->impl->device_bytes()) ||
            h.steps_done < 0 || h.steps_done > p.cfg.nstep)
            throw Error{ELA3D_E_INVALID, "checkpoint belongs to a different configuration"};
        // a short file is refused before any state is overwritten, so the shot is left as it was
        if (std::fseek(f.f, 0, SEEK_END) != 0 || std::ftell(f.f) != long(sizeof(h) + h.state_bytes) ||
            std::fseek(f.f, long(sizeof(h)), SEEK_SET) != 0)
            throw Error{ELA3D_E_IO, "checkpoint truncated or oversized"};
        std::vector<char> buf(std::min<size_t>(kChunk, h.state_bytes));
        for (size_t off = 0; off < h.state_bytes; off += buf.size()) {
            const size_t n = std::min<size_t>(buf.size(), h.state_bytes - off);
            if (std::fread(buf.data(), 1, n, f.f) != n) throw Error{ELA3D_E_IO, "checkpoint truncated"};
            s->impl->write_state(off, n, buf.data());
        }
        s->impl->set_steps_done(h.steps_done);
        return ELA3D_OK;
    });
}

int ela3d_info_get(ela3d_shot* s, ela3d_info* out)
{
    return guarded([&] {
        if (!s || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        s->impl->info(*out);
        return ELA3D_OK;
    });
}

int ela3d_profile(ela3d_shot* s, int32_t nsteps, ela3d_kernel_stats* out, int32_t cap,
                  int32_t* nkernels, double* wall_ms)
{
    return ela3d_profile_sampled(s, nsteps, 1, out, cap, nkernels, wall_ms);
}

int ela3d_profile_sampled(ela3d_shot* s, int32_t nsteps, int32_t every, ela3d_kernel_stats* out, int32_t cap,
                          int32_t* nkernels, double* wall_ms)
{
    return guarded([&] {
        if (!s || nsteps < 0 || every < 1) throw Error{ELA3D_E_INVALID, "bad argument"};
        std::vector<ela3d_kernel_stats> st;
        double w = 0.0;
        s->impl->profile(nsteps, st, w, every);
        if (nkernels) *nkernels = int32_t(st.size());
        if (wall_ms) *wall_ms = w;
        if (out) {
            if (cap < int32_t(st.size())) throw Error{ELA3D_E_RANGE, "kernel stats buffer too small"};
            for (size_t q = 0; q < st.size(); ++q) out[q] = st[q];
        }
        return ELA3D_OK;
    });
}

void ela3d_destroy(ela3d_shot* s) { delete s; }

int ela3d_plan(const ela3d_config* cfg, ela3d_info* out)
{
    return guarded([&] {
        if (!cfg || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        ela3d::ShotPlan p = ela3d::make_plan(*cfg);
        std::memset(out, 0, sizeof(*out));
        out->cells = p.cells;
        out->device_bytes = ela3d::device_bytes_estimate(p.cfg, &p);
        out->alg_bytes_per_step = p.alg_step();
        out->alg_bytes_stress = p.alg_stress_interior + p.alg_stress_slab;
        out->alg_bytes_velocity = p.alg_velocity_interior + p.alg_velocity_slab;
        out->pml_bytes_per_step = p.alg_pml;
        out->nsamples = p.nsamples;
        out->nenergy = p.nenergy;
        ela3d::Tuning t = ela3d::choose_tuning(p);
        out->block = t.block;
        out->zseg = t.zseg_interior;
        out->zseg_slab = t.zseg_slab;
        out->vec = t.vec;
        for (int a = 0; a < 3; ++a) {
            out->region_lo[a] = p.ax[a].lo;
            out->region_hi[a] = p.ax[a].hi;
        }
        return ELA3D_OK;
    });
}

int ela3d_pml_tables(const ela3d_config* cfg, int32_t axis, double* out, size_t cap)
{
    return guarded([&] {
        if (!cfg || !out || axis < 0 || axis > 2) throw Error{ELA3D_E_INVALID, "bad argument"};
        ela3d::ShotPlan p = ela3d::make_plan(*cfg);
        const ela3d::PmlAxis& A = p.ax[axis];
        const size_t n = size_t(A.n);
        if (cap < 10 * n) throw Error{ELA3D_E_RANGE, "table buffer too small"};
        const std::vector<double>* t[10] = {&A.d, &A.K, &A.alpha, &A.a, &A.b,
                                            &A.d_h, &A.K_h, &A.alpha_h, &A.a_h, &A.b_h};
        for (int q = 0; q < 10; ++q) std::memcpy(out + q * n, t[q]->data(), n * sizeof(double));
        return ELA3D_OK;
    });
}

const char* ela3d_strerror(int code)
{
    switch (code) {
    case ELA3D_OK: return "ok";
    case ELA3D_E_INVALID: return "invalid argument or config";
    case ELA3D_E_OOM: return "out of memory";
    case ELA3D_E_UNSTABLE: return "unstable: non-finite wavefield";
    case ELA3D_E_HIP: return "HIP runtime error";
    case ELA3D_E_NODEVICE: return "no GPU device";
    case ELA3D_E_RANGE: return "buffer too small or index out of range";
    case ELA3D_E_STATE: return "invalid state for this call";
    case ELA3D_E_BUDGET: return "exceeds the device-memory budget";
    case ELA3D_E_NOTFOUND: return "unknown ticket";
    case ELA3D_E_TIMEOUT: return "timed out";
    case ELA3D_E_IO: return "file error";
    default: return "unknown status";
    }
}

const char* ela3d_last_error(void) { return ela3d::g_last_error.c_str(); }

int ela3d_device_count(int32_t* out)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (out) *out = n;
    return ELA3D_OK;
}

int ela3d_copy_bench(int32_t device, int64_t bytes, int32_t reps, int32_t variant, double* ms_per_copy)
{
    return guarded([&] {
        int rc = need_device(device);
        if (rc) return rc;
        if (bytes < (variant == 8 ? 4 * ela3d::kStreamBlock : 16) || reps < 1 || !ms_per_copy || variant < 0 ||
            variant > 8 ||
            bytes > (int64_t(1) << 36))
            throw Error{ELA3D_E_INVALID, "bad argument"};
        HIP_CHECK(hipSetDevice(device));
        // released on every path, errors included
        struct Res {
            void* a = nullptr;
            void* b = nullptr;
            hipStream_t st = nullptr;
            hipEvent_t t0 = nullptr, t1 = nullptr;
            ~Res()
            {
                if (t0) (void)hipEventDestroy(t0);
                if (t1) (void)hipEventDestroy(t1);
                if (st) (void)hipStreamDestroy(st);
                if (a) (void)hipFree(a);
                if (b) (void)hipFree(b);
            }
        } R;
        HIP_CHECK(hipStreamCreateWithFlags(&R.st, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&R.t0));
        HIP_CHECK(hipEventCreate(&R.t1));
        std::function<void()> launch;
        if (variant >= 4) {  // multi-stream: every array holds `bytes`
            int nr = 0, nw = 0;
            ela3d::stream_counts(variant, nr, nw);
            // floats per array: a multiple of 4, and of the block for the interleaved variant
            const long q = variant == 8 ? 4 * ela3d::kStreamBlock : 16;
            const long n = long(bytes / q) * (q / 4);
            const size_t total = size_t(n) * 4 * size_t(nr + nw);
            HIP_CHECK(hipMalloc(&R.a, total));
            HIP_CHECK(hipMemsetAsync(R.a, 0, total, R.st));
            launch = [&, n] { ela3d::launch_streams(static_cast<float*>(R.a), n, variant, R.st); };
        } else {
            const long n4 = long(bytes / 16);
            HIP_CHECK(hipMalloc(&R.a, size_t(n4) * 16));
            HIP_CHECK(hipMalloc(&R.b, size_t(n4) * 16));
            HIP_CHECK(hipMemsetAsync(R.a, 0, size_t(n4) * 16, R.st));
            launch = [&, n4] { ela3d::launch_copy((const float4*)R.a, (float4*)R.b, n4, variant, R.st); };
        }
        launch();  // warm-up
        HIP_CHECK(hipEventRecord(R.t0, R.st));
        for (int r = 0; r < reps; ++r) launch();
        HIP_CHECK(hipEventRecord(R.t1, R.st));
        HIP_CHECK(hipEventSynchronize(R.t1));
        HIP_CHECK(hipGetLastError());
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, R.t0, R.t1));
        *ms_per_copy = ms / reps;
        return ELA3D_OK;
    });
}

}  // extern "C"
