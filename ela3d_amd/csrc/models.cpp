// models.cpp -- seeded synthetic earth models.
//
// Homogeneous: constant cp/cs/rho arrays, as the reference fills them (CPU:41-59).
// Layered anticline: nlayers flat-lying layers whose interfaces are pushed up by
// a Gaussian bump, z_i(x,y) = z0_i - A*exp(-((x-xc)^2 + (y-yc)^2)/w^2), with
// Vp rising linearly from vp_top to vp_bottom, +-noise uniform per cell from
// splitmix64, Vs = Vp/vp_vs, rho = gardner_a * Vp^gardner_b (Gardner's relation).
// z0_i are evenly spaced: z0_i = i * (nz-1)*h / nlayers for i = 1..nlayers-1.
#include <cmath>
#include <cstring>

#include "../../include/ela3d.h"

namespace {

// splitmix64 (Steele, Lea, Flood 2014), one draw per cell in x-fastest order
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform()  // [0,1) with 53 random bits
    {
        return double(next() >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

extern "C" {

int ela3d_model_homogeneous(int32_t nx, int32_t ny, int32_t nz, float vp, float vs, float rho,
                            float* cp, float* cs, float* rh)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !cp || !cs || !rh) return ELA3D_E_INVALID;
    const size_t n = size_t(nx) * ny * nz;
    for (size_t q = 0; q < n; ++q) {
        cp[q] = vp;
        cs[q] = vs;
        rh[q] = rho;
    }
    return ELA3D_OK;
}

void ela3d_layered_default(ela3d_layered_spec* s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->nx = 401;
    s->ny = 401;
    s->nz = 301;
    s->h = 10.0;
    s->nlayers = 6;
    s->vp_top = 2000.0;
    s->vp_bottom = 4500.0;
    s->noise = 0.03;
    s->seed = 20240611ull;
    s->bump_amp = 600.0;
    s->bump_width = 1200.0;
    s->vp_vs = 1.73;
    s->gardner_a = 310.0;
    s->gardner_b = 0.25;
}

int ela3d_model_layered(const ela3d_layered_spec* s, float* cp, float* cs, float* rh, double* vp_max)
{
    if (!s || !cp || !cs || !rh) return ELA3D_E_INVALID;
    if (s->nx <= 0 || s->ny <= 0 || s->nz <= 0 || s->nlayers < 1 || !(s->h > 0) || !(s->vp_vs > 1))
        return ELA3D_E_INVALID;
    const int L = s->nlayers;
    const double depth = (s->nz - 1) * s->h;
    const double xc = 0.5 * (s->nx - 1) * s->h, yc = 0.5 * (s->ny - 1) * s->h;
    const double w2 = s->bump_width * s->bump_width;
    SplitMix64 rng{s->seed};
    double vmax = 0.0;
    for (int k = 0; k < s->nz; ++k) {
        const double z = k * s->h;
        for (int j = 0; j < s->ny; ++j) {
            const double y = j * s->h;
            for (int i = 0; i < s->nx; ++i) {
                const double x = i * s->h;
                const double up = s->bump_amp * std::exp(-((x - xc) * (x - xc) + (y - yc) * (y - yc)) / w2);
                int layer = 0;  // count interfaces above this cell
                for (int q = 1; q < L; ++q) {
                    const double zq = q * depth / L - up;
                    if (z >= zq) layer = q;
                }
                const double vp0 = L > 1 ? s->vp_top + (s->vp_bottom - s->vp_top) * layer / (L - 1) : s->vp_top;
                const double u = rng.uniform();
                const double vp = vp0 * (1.0 + s->noise * (2.0 * u - 1.0));
                const size_t o = size_t(i) + size_t(j) * s->nx + size_t(k) * s->nx * s->ny;
                cp[o] = float(vp);
                cs[o] = float(vp / s->vp_vs);
                rh[o] = float(s->gardner_a * std::pow(vp, s->gardner_b));
                if (double(cp[o]) > vmax) vmax = cp[o];
            }
        }
    }
    if (vp_max) *vp_max = vmax;
    return ELA3D_OK;
}

}  // extern "C"
