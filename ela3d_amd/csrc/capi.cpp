// capi.cpp -- extern "C" boundary: status codes, no exceptions across the ABI.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "shot.hpp"

struct ela3d_shot {
    std::unique_ptr<ela3d::Shot> impl;
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
    // reference literals (CPU:19-78, 151-153), shrunk to a 101^3 grid
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
        const size_t cells = size_t(plan.cells);
        for (size_t q = 0; q < cells; ++q) {
            if (!(std::isfinite(cp[q]) && std::isfinite(cs[q]) && std::isfinite(rho[q])) ||
                !(cs[q] > 0.f) || !(rho[q] > 0.f) || !(cp[q] > cs[q]))
                throw Error{ELA3D_E_INVALID, "model cell " + std::to_string(q) +
                                                 " invalid (need finite cp > cs > 0, rho > 0)"};
        }
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
    return guarded([&] {
        if (!s || nsteps < 0) throw Error{ELA3D_E_INVALID, "bad argument"};
        std::vector<ela3d_kernel_stats> st;
        double w = 0.0;
        s->impl->profile(nsteps, st, w);
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
        if (bytes < 16 || reps < 1 || !ms_per_copy || variant < 0 || variant > 3 ||
            bytes > (int64_t(1) << 36))
            throw Error{ELA3D_E_INVALID, "bad argument"};
        HIP_CHECK(hipSetDevice(device));
        const long n4 = long(bytes / 16);
        void *a = nullptr, *b = nullptr;
        HIP_CHECK(hipMalloc(&a, size_t(n4) * 16));
        hipError_t e = hipMalloc(&b, size_t(n4) * 16);
        if (e != hipSuccess) {
            hipFree(a);
            HIP_CHECK(e);
        }
        hipStream_t st;
        hipEvent_t t0, t1;
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&t0));
        HIP_CHECK(hipEventCreate(&t1));
        HIP_CHECK(hipMemsetAsync(a, 0, size_t(n4) * 16, st));
        ela3d::launch_copy((const float4*)a, (float4*)b, n4, variant, st);  // warm-up
        HIP_CHECK(hipEventRecord(t0, st));
        for (int r = 0; r < reps; ++r) ela3d::launch_copy((const float4*)a, (float4*)b, n4, variant, st);
        HIP_CHECK(hipEventRecord(t1, st));
        HIP_CHECK(hipEventSynchronize(t1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
        *ms_per_copy = ms / reps;
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        hipStreamDestroy(st);
        hipFree(a);
        hipFree(b);
        return ELA3D_OK;
    });
}

}  // extern "C"
