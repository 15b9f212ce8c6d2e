// tune.cpp -- launch-shape autotuner and its process-wide cache.
//
// A launch shape is (threads per workgroup, z planes per workgroup in the
// interior box, z planes per workgroup in the absorbing-layer boxes).  The
// tuner times every candidate on a scratch shot of the same grid (source
// amplitude 0, so the traffic is identical and nothing is perturbed), keeps
// the fastest, and caches it keyed by (nx, ny, nz, storage, npoints_pml).
// Shots created with block = 0 and zseg = 0 take the cached shape.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "shot.hpp"

namespace ela3d {

typedef std::tuple<int, int, int, int, int> TuneKey;

static std::mutex& tune_mutex()
{
    static std::mutex m;
    return m;
}

static std::map<TuneKey, std::pair<Tuning, double>>& tune_cache()
{
    static std::map<TuneKey, std::pair<Tuning, double>> c;
    return c;
}

static TuneKey key_of(const ela3d_config& c)
{
    return TuneKey(c.nx, c.ny, c.nz, c.storage, c.npoints_pml);
}

bool cached_tuning(const ela3d_config& c, Tuning& out)
{
    std::lock_guard<std::mutex> g(tune_mutex());
    auto it = tune_cache().find(key_of(c));
    if (it == tune_cache().end()) return false;
    out = it->second.first;
    return true;
}

void store_tuning(const ela3d_config& c, const Tuning& t, double ms)
{
    std::lock_guard<std::mutex> g(tune_mutex());
    tune_cache()[key_of(c)] = std::make_pair(t, ms);
}

}  // namespace ela3d

using ela3d::Error;
using ela3d::Tuning;

extern "C" {

int ela3d_autotune(const ela3d_config* cfg, const float* cp, const float* cs, const float* rho,
                   int32_t steps, ela3d_tuning* out)
{
    try {
        if (!cfg || !cp || !cs || !rho || steps < 1) throw Error{ELA3D_E_INVALID, "bad argument"};
        ela3d_config c = *cfg;
        c.factor = 0.0;  // scratch shot: same traffic, no wavefield
        c.nrec = 0;
        c.rec = nullptr;
        c.use_graph = 0;
        c.nstep = std::max(c.nstep, steps * 64 + 16);
        c.block = 256;
        c.zseg = 0;
        ela3d::ShotPlan plan = ela3d::make_plan(c);
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= c.device) throw Error{ELA3D_E_NODEVICE, "no HIP device"};
        std::unique_ptr<ela3d::Shot> shot = ela3d::make_shot(plan, cp, cs, rho);
        static const int blocks[] = {64, 128, 256};
        static const int zsi[] = {2, 4, 8, 16, 32};
        static const int zss[] = {2, 4, 8};
        static const int vecs[] = {1, 2, 4};
        Tuning best;
        double best_ms = 1e300;
        int tried = 0;
        std::vector<ela3d_kernel_stats> st;
        for (int vv : vecs)
        for (int b : blocks)
            for (int zi : zsi)
                for (int zs : zss) {
                    Tuning t;
                    t.block = b;
                    t.zseg_interior = zi;
                    t.zseg_slab = zs;
                    t.vec = vv;
                    shot->retune(t);
                    double wall = 0.0;
                    shot->profile(1, st, wall);  // warm
                    shot->profile(steps, st, wall);
                    const double ms = (st[0].total_ms + st[1].total_ms) / steps;
                    ++tried;
                    if (std::getenv("ELA3D_TUNE_VERBOSE"))
                        std::fprintf(stderr, "[tune] vec %d block %3d zseg %3d/%d: stress %.4f velocity %.4f ms\n",
                                     vv, b, zi, zs, st[0].total_ms / steps, st[1].total_ms / steps);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = t;
                    }
                }
        ela3d::store_tuning(*cfg, best, best_ms);
        if (out) {
            out->block = best.block;
            out->zseg_interior = best.zseg_interior;
            out->zseg_slab = best.zseg_slab;
            out->vec = best.vec;
            out->ms_per_step = best_ms;
            out->candidates = tried;
        }
        return ELA3D_OK;
    } catch (const Error& e) {
        ela3d::set_last_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        ela3d::set_last_error(e.what());
        return ELA3D_E_INVALID;
    }
}

int ela3d_tuning_set(const ela3d_config* cfg, const ela3d_tuning* t)
{
    if (!cfg || !t || t->block <= 0 || t->zseg_interior <= 0 || t->zseg_slab <= 0) return ELA3D_E_INVALID;
    if (t->block != 64 && t->block != 128 && t->block != 256) return ELA3D_E_INVALID;
    Tuning u;
    u.block = t->block;
    u.zseg_interior = t->zseg_interior;
    u.zseg_slab = t->zseg_slab;
    if (t->vec != 1 && t->vec != 2 && t->vec != 4) return ELA3D_E_INVALID;
    u.vec = t->vec;
    ela3d::store_tuning(*cfg, u, t->ms_per_step);
    return ELA3D_OK;
}

int ela3d_tuning_get(const ela3d_config* cfg, ela3d_tuning* t)
{
    if (!cfg || !t) return ELA3D_E_INVALID;
    Tuning u;
    if (!ela3d::cached_tuning(*cfg, u)) return ELA3D_E_NOTFOUND;
    t->block = u.block;
    t->zseg_interior = u.zseg_interior;
    t->zseg_slab = u.zseg_slab;
    t->vec = u.vec;
    t->ms_per_step = 0.0;
    t->candidates = 0;
    return ELA3D_OK;
}

}  // extern "C"
