// formats.cpp -- the reference's model and output files, and the MATLAB anticline.
//
// Host-only code: raw x-fastest model files (float64 at UNI:595-644, float32 at
// ZC:544-595; snapshots v?%05i.bin, CPU:685-722), text traces in the layout of
// rechorv?.txt (CPU:636-670) and recsrc?.txt (UNI:1202-1223), and a port of
// generate_3d_model_anticline.m.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "core.hpp"

using ela3d::Error;

namespace {

struct File {
    std::FILE* f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

template <typename F>
int guarded(F fn)
{
    try {
        return fn();
    } catch (const Error& e) {
        ela3d::set_last_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        ela3d::set_last_error(e.what());
        return ELA3D_E_INVALID;
    }
}

const int64_t kChunk = int64_t(1) << 22;  // values per read/write

}  // namespace

extern "C" {

int ela3d_read_raw(const char* path, int64_t n, int32_t bpv, float* out)
{
    return guarded([&] {
        if (!path || !out || n < 0 || (bpv != 4 && bpv != 8)) throw Error{ELA3D_E_INVALID, "bad argument"};
        File f;
        f.f = std::fopen(path, "rb");
        if (!f.f) throw Error{ELA3D_E_IO, std::string("cannot open ") + path};
        std::vector<double> d(bpv == 8 ? size_t(std::min(n, kChunk)) : 0);
        for (int64_t off = 0; off < n; off += kChunk) {
            const size_t m = size_t(std::min(kChunk, n - off));
            if (bpv == 4) {
                if (std::fread(out + off, 4, m, f.f) != m) throw Error{ELA3D_E_IO, std::string(path) + " is too short"};
            } else {
                if (std::fread(d.data(), 8, m, f.f) != m) throw Error{ELA3D_E_IO, std::string(path) + " is too short"};
                for (size_t q = 0; q < m; ++q) out[off + int64_t(q)] = float(d[q]);
            }
        }
        return ELA3D_OK;
    });
}

int ela3d_write_raw(const char* path, int64_t n, int32_t bpv, const float* in)
{
    return guarded([&] {
        if (!path || !in || n < 0 || (bpv != 4 && bpv != 8)) throw Error{ELA3D_E_INVALID, "bad argument"};
        File f;
        f.f = std::fopen(path, "wb");
        if (!f.f) throw Error{ELA3D_E_IO, std::string("cannot open ") + path + " for writing"};
        std::vector<double> d(bpv == 8 ? size_t(std::min(n, kChunk)) : 0);
        for (int64_t off = 0; off < n; off += kChunk) {
            const size_t m = size_t(std::min(kChunk, n - off));
            size_t w;
            if (bpv == 4) {
                w = std::fwrite(in + off, 4, m, f.f);
            } else {
                for (size_t q = 0; q < m; ++q) d[q] = double(in[off + int64_t(q)]);
                w = std::fwrite(d.data(), 8, m, f.f);
            }
            if (w != m) throw Error{ELA3D_E_IO, std::string("write to ") + path + " failed"};
        }
        if (std::fclose(f.f) != 0) {
            f.f = nullptr;
            throw Error{ELA3D_E_IO, std::string("close of ") + path + " failed"};
        }
        f.f = nullptr;
        return ELA3D_OK;
    });
}

int ela3d_write_traces(const char* path, const double* v, int64_t nrows, int64_t ncols, int32_t stride,
                       int32_t offset, int32_t append, int32_t final_newline)
{
    return guarded([&] {
        if (!path || (!v && nrows * ncols > 0) || nrows < 0 || ncols < 0 || stride < 1 || offset < 0 ||
            offset >= stride)
            throw Error{ELA3D_E_INVALID, "bad argument"};
        File f;
        f.f = std::fopen(path, append ? "a" : "w");
        if (!f.f) throw Error{ELA3D_E_IO, std::string("cannot open ") + path + " for writing"};
        char num[64];
        std::string line;
        for (int64_t r = 0; r < nrows; ++r) {
            line.clear();
            for (int64_t c = 0; c < ncols; ++c) {
                // the reference streams FP32 values with the default 6 significant digits
                const float x = float(v[(r * ncols + c) * stride + offset]);
                std::snprintf(num, sizeof(num), "%g ", double(x));
                line += num;
            }
            if (r + 1 < nrows || final_newline) line += '\n';
            if (std::fwrite(line.data(), 1, line.size(), f.f) != line.size())
                throw Error{ELA3D_E_IO, std::string("write to ") + path + " failed"};
        }
        if (std::fclose(f.f) != 0) {
            f.f = nullptr;
            throw Error{ELA3D_E_IO, std::string("close of ") + path + " failed"};
        }
        f.f = nullptr;
        return ELA3D_OK;
    });
}

// generate_3d_model_anticline.m on an n^3 grid (MAT:1-77).  1-based (i, j, k) as
// in the script; i runs along x (fastest), k along depth.  The surface is
// symmetric in i and j, so MATLAB's meshgrid orientation does not matter.
int ela3d_model_anticline(int32_t n, float* cp, float* cs, float* rh)
{
    return guarded([&] {
        if (n < 2 || n > 2048 || !cp || !cs || !rh) throw Error{ELA3D_E_INVALID, "bad argument"};
        const double s = double(n) / 201.0;  // the script's constants are for n = 201
        const double h1 = 121.0 * s, h2 = 161.0 * s, h3 = 201.0 * s, d = 100.0 * s, koil = 80.0 * s;
        // vp, vs, rho of the five media (MAT:72-76)
        static const float media[5][3] = {{3000.f, 1700.f, 2800.f},
                                          {3600.f, 1400.f, 2300.f},
                                          {3000.f, 1700.f, 2800.f},
                                          {3600.f, 1700.f, 2800.f},
                                          {3650.f, 1500.f, 2450.f}};
        const double pi = 3.141592653589793238462643;
        std::vector<double> sn(n);
        for (int i = 0; i < n; ++i) sn[i] = std::sin(pi * double(i) / double(n - 1));  // linspace(0, pi, n)
        const size_t nxy = size_t(n) * n;
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= n; ++i) {
                double z = ((sn[i - 1] + sn[j - 1]) / 2.0 - 0.5) / 0.5;
                if (z < 0.0) z = 0.0;
                const double s1 = h1 - z * d, s2 = h2 - z * d, s3 = h3 - z * d;
                for (int k = 1; k <= n; ++k) {
                    int m = 1;
                    if (k > s1) m = 2;
                    if (k > s2) m = (k < koil) ? 4 : 3;
                    if (k > s3) m = 5;
                    const size_t o = size_t(i - 1) + size_t(j - 1) * n + size_t(k - 1) * nxy;
                    cp[o] = media[m - 1][0];
                    cs[o] = media[m - 1][1];
                    rh[o] = media[m - 1][2];
                }
            }
        return ELA3D_OK;
    });
}

}  // extern "C"
