// core.hpp -- host-side shot description shared by the runtime, the server and the CLI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ela3d.h"

namespace ela3d {

// Error carried out of internal code; converted to a status code at the C ABI.
struct Error {
    int code;
    std::string msg;
};

void set_last_error(const std::string& msg);

// One axis of C-PML profile tables (CPU:158-361), FP64.  Index i is the 0-based
// grid index; "half" tables are evaluated at x = h*i + h/2.
struct PmlAxis {
    int n = 0;
    std::vector<double> d, K, alpha, a, b;
    std::vector<double> d_h, K_h, alpha_h, a_h, b_h;
    // compact region holding every index where a or a_h is non-zero: [0,lo) u [hi,n)
    int lo = 0, hi = 0;
    int width() const { return lo + (n - hi); }
    // compact index or -1
    int compact(int i) const { return i < lo ? i : (i >= hi ? lo + (i - hi) : -1); }
};

struct ShotPlan {
    ela3d_config cfg{};             // validated copy (rec pointer cleared)
    std::vector<int32_t> rec;       // [nrec][3]
    PmlAxis ax[3];
    std::vector<double> wavelet;    // s(t) at t=(it-1)dt, it=1..nstep (FP64)
    int64_t cells = 0;
    int nsamples = 0, nenergy = 0;
    int kmax_shear = 0;             // nz-2, or nz-1 with the ghost quirk
    double t0 = 0.0;
    // z-slab of a larger grid (make_slab_plan): cfg.nz counts the slab's planes
    // plus one ghost plane on each side that has a neighbour slab; local plane k
    // is global plane k + koff.  Only planes [own_k0, own_k1) are updated; the
    // ghost planes are written by halo exchange.  A whole grid has koff = 0,
    // own = [0, nz), nz_global = nz.
    int koff = 0, nz_global = 0, own_k0 = 0, own_k1 = 0;
    int ek0 = 0, ek1 = 0;           // planes [ek0, ek1) of the energy box (local indices)
    // algorithmic traffic (bytes) of the two-pass schedule, split by launch
    // (interior box; absorbing-layer boxes)
    double alg_stress_interior = 0, alg_stress_slab = 0;
    double alg_velocity_interior = 0, alg_velocity_slab = 0;
    double alg_pml = 0;
    int64_t interior_cells = 0;     // cells of the interior box
    double fused_saving = 0;        // bytes per interior cell the fused schedule does not move
    double alg_step() const {
        return alg_stress_interior + alg_stress_slab + alg_velocity_interior + alg_velocity_slab;
    }
    // fused interior step: every field read and written once, the model read once
    double alg_fused_interior() const {
        return alg_stress_interior + alg_velocity_interior - fused_saving * double(interior_cells);
    }
    double alg_step_fused() const { return alg_fused_interior() + alg_stress_slab + alg_velocity_slab; }
};

// An xy range [p0,p1) of the flattened plane (p = i + j*nx) times planes
// [k0,k1); mode 0 = interior (y and z outside the C-PML), 1 = absorbing layer.
struct BoxRange {
    int p0, p1, k0, k1, mode;
};
// The boxes of one pass: z-min and z-max absorbing layers over whole planes,
// then the y-min and y-max absorbing layers between them, then the interior,
// whose xy range is cut at multiples of 64 elements (rows between the y layer
// edge and that cut go to the y-layer boxes).  Empty boxes are dropped.
int split_boxes(const ShotPlan& p, BoxRange out[5]);
// Algorithmic bytes of one pass (0 stress, 1 velocity) over one box.
double box_alg_bytes(const ShotPlan& p, const BoxRange& b, int pass);

int storage_bytes(int storage);
// Validate and expand a config; throws Error.
ShotPlan make_plan(const ela3d_config& cfg);
// The plan of planes [k0, k1) of the grid `cfg` describes, as a slab with ghost
// planes (see ShotPlan::koff).  C-PML z tables are the global ones at global k;
// receivers outside the slab are dropped (slab_receivers lists the kept ones'
// global indices); the source is injected only by the slab that owns its plane.
ShotPlan make_slab_plan(const ela3d_config& cfg, int k0, int k1, std::vector<int>* slab_receivers = nullptr);
// Tables of one axis (0 x, 1 y, 2 z) for a validated config.
PmlAxis make_axis(const ela3d_config& cfg, int axis);
double wavelet_value(const ela3d_config& cfg, double t0, int it);
// Device bytes a shot of this config holds (used by admission control).
int64_t device_bytes_estimate(const ela3d_config& cfg, const ShotPlan* plan);
// Throws ELA3D_E_INVALID unless every cell has finite cp > cs > 0 and rho > 0.
void check_model(int64_t cells, const float* cp, const float* cs, const float* rho);
// exact size of the single device allocation a shot of this plan makes
int64_t shot_device_bytes(const ShotPlan& p);

}  // namespace ela3d
