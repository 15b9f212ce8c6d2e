// ela3d_run -- command-line driver of the elastic time loop (replaces the
// reference program's main(), seismicCPML_3d_isotropic_simpelmodel.cpp:17-726).
//
// Every parameter of the reference's literal block is a key=value argument; the
// run goes through the C ABI of libela3d.so and writes the reference's output
// files (rechorv?.txt, v?%05i.bin, recsrc?.txt) plus energy.txt into out=DIR.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ela3d.h"

namespace {

const char* kUsage =
    "usage: ela3d_run [key=value ...]\n"
    "  grid/time : nx ny nz dx [dy dz] dt nstep            (NIMX.., DELTAX.., DELTAT, NSTEP)\n"
    "  C-PML     : npml npower kmax alphamax rcoef vpml\n"
    "  source    : f0 t0 factor angle isrc jsrc ksrc (0-based)\n"
    "              src=force|explosive|sigma|force3  wavelet=ricker|gaussian|dgaussian\n"
    "  receivers : rec=grid:NGATX,NGATY,K  (the reference's plane gather, CPU:593-602, 629-645)\n"
    "              rec=line:x,J,K,STEP | rec=line:y,I,K,STEP | rec=file:PATH (0-based 'i j k' lines)\n"
    "  model     : model=homogeneous [vp vs rho] | model=layered [seed] | model=anticline (cube)\n"
    "              model=raw:CP.bin,CS.bin,RHO.bin[:4|8]  (x-fastest float32 or float64)\n"
    "  run       : sampgat energy_every storage=fp32|fp16|fp64 compat=0..7|reference dirichlet=0|1\n"
    "              free_surface=0|1 (traction-free plane k = 0, no z-min absorbing layer)\n"
    "              it_output=N (v?%05i.bin) out=DIR checkpoint=PATH checkpoint_every=N resume=PATH\n"
    "              reference=1 (the reference's choices: sigma source, dGaussian wavelet, all\n"
    "              quirks, no Dirichlet faces, sampgat=2) device=N append=0|1\n";

struct Args {
    std::map<std::string, std::string> kv;
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    std::string str(const std::string& k, const std::string& d) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    double num(const std::string& k, double d) const
    {
        auto it = kv.find(k);
        if (it == kv.end()) return d;
        char* end = nullptr;
        const double v = std::strtod(it->second.c_str(), &end);
        if (!end || *end) {
            std::fprintf(stderr, "ela3d_run: %s=%s is not a number\n", k.c_str(), it->second.c_str());
            std::exit(2);
        }
        return v;
    }
    int integer(const std::string& k, int d) const { return int(std::lround(num(k, d))); }
};

std::vector<std::string> split(const std::string& s, char sep)
{
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t b = s.find(sep, a);
        out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return out;
}

[[noreturn]] void fail(const char* what, int rc)
{
    std::fprintf(stderr, "ela3d_run: %s: %s (%s)\n", what, ela3d_strerror(rc), ela3d_last_error());
    std::exit(1);
}

void check(int rc, const char* what)
{
    if (rc != ELA3D_OK) fail(what, rc);
}

std::vector<int32_t> receivers(const Args& a, const ela3d_config& c)
{
    std::vector<int32_t> r;
    const std::string spec = a.str("rec", "");
    if (spec.empty()) return r;
    const size_t colon = spec.find(':');
    const std::string kind = spec.substr(0, colon);
    const std::string rest = colon == std::string::npos ? "" : spec.substr(colon + 1);
    auto push = [&](int i, int j, int k) {
        r.push_back(i);
        r.push_back(j);
        r.push_back(k);
    };
    if (kind == "grid") {  // the reference's strided plane gather (CPU:629-645)
        const auto p = split(rest, ',');
        if (p.size() != 3) {
            std::fprintf(stderr, "ela3d_run: rec=grid:NGATX,NGATY,K\n");
            std::exit(2);
        }
        const int ngx = std::atoi(p[0].c_str()), ngy = std::atoi(p[1].c_str()), k = std::atoi(p[2].c_str());
        const int xlen = ngx > 0 ? c.nx / ngx : 0;
        const int ylen = ngy > 0 ? (c.ny - 2 * c.npoints_pml) / ngy : 0;
        if (xlen <= 0 || ylen <= 0) {  // CPU:631-634
            std::fprintf(stderr, "ela3d_run: more receivers than grid points\n");
            std::exit(2);
        }
        for (int j = 0; j < c.ny; j += ylen)
            for (int i = 0; i < c.nx; i += xlen) push(i, j, k);
    } else if (kind == "line") {
        const auto p = split(rest, ',');
        if (p.size() != 4 || (p[0] != "x" && p[0] != "y")) {
            std::fprintf(stderr, "ela3d_run: rec=line:x,J,K,STEP or rec=line:y,I,K,STEP\n");
            std::exit(2);
        }
        const int fixed = std::atoi(p[1].c_str()), k = std::atoi(p[2].c_str()), st = std::max(1, std::atoi(p[3].c_str()));
        if (p[0] == "x")
            for (int i = 0; i < c.nx; i += st) push(i, fixed, k);
        else
            for (int j = 0; j < c.ny; j += st) push(fixed, j, k);
    } else if (kind == "file") {
        std::FILE* f = std::fopen(rest.c_str(), "r");
        if (!f) {
            std::fprintf(stderr, "ela3d_run: cannot open %s\n", rest.c_str());
            std::exit(2);
        }
        int i, j, k;
        while (std::fscanf(f, "%d %d %d", &i, &j, &k) == 3) push(i, j, k);
        std::fclose(f);
    } else {
        std::fprintf(stderr, "ela3d_run: unknown receiver spec %s\n", spec.c_str());
        std::exit(2);
    }
    return r;
}

void write_snapshot(ela3d_shot* s, int field, const char* name, int it, const std::string& dir, size_t cells)
{
    std::vector<double> d(cells);
    check(ela3d_snapshot(s, field, d.data(), d.size()), "snapshot");
    std::vector<float> f(d.begin(), d.end());
    char fn[32];
    std::snprintf(fn, sizeof(fn), "%s%05i.bin", name, it);  // CPU:685-722
    const std::string path = dir + "/" + fn;
    check(ela3d_write_raw(path.c_str(), int64_t(cells), 4, f.data()), "snapshot write");
}

}  // namespace

int main(int argc, char** argv)
{
    Args a;
    for (int q = 1; q < argc; ++q) {
        const std::string s = argv[q];
        if (s == "-h" || s == "--help") {
            std::fputs(kUsage, stdout);
            return 0;
        }
        const size_t eq = s.find('=');
        if (eq == std::string::npos || eq == 0) {
            std::fprintf(stderr, "ela3d_run: expected key=value, got %s\n%s", s.c_str(), kUsage);
            return 2;
        }
        a.kv[s.substr(0, eq)] = s.substr(eq + 1);
    }
    const bool ref = a.integer("reference", 0) != 0;

    ela3d_config c;
    ela3d_config_default(&c);
    c.nx = a.integer("nx", c.nx);
    c.ny = a.integer("ny", c.ny);
    c.nz = a.integer("nz", c.nz);
    c.dx = a.num("dx", c.dx);
    c.dy = a.num("dy", c.dx);
    c.dz = a.num("dz", c.dx);
    c.dt = a.num("dt", c.dt);
    c.nstep = a.integer("nstep", c.nstep);
    c.npoints_pml = a.integer("npml", c.npoints_pml);
    c.npower = a.num("npower", c.npower);
    c.k_max_pml = a.num("kmax", c.k_max_pml);
    c.alpha_max_pml = a.num("alphamax", c.alpha_max_pml);
    c.rcoef = a.num("rcoef", c.rcoef);
    c.vpml = a.num("vpml", c.vpml);
    c.f0 = a.num("f0", c.f0);
    c.t0 = a.num("t0", c.t0);
    c.factor = a.num("factor", c.factor);
    c.angle_force = a.num("angle", c.angle_force);
    c.isrc = a.integer("isrc", c.nx / 2);
    c.jsrc = a.integer("jsrc", c.ny / 2);
    c.ksrc = a.integer("ksrc", c.nz / 2);
    if (ref) {  // CPU:64-72, 610, 620-622
        c.src_kind = ELA3D_SRC_REFERENCE_SIGMA;
        c.wavelet = ELA3D_WAVELET_DGAUSSIAN;
        c.compat_flags = ELA3D_COMPAT_REFERENCE;
        c.dirichlet = 0;
        c.sample_every = 2;
    }
    const std::string src = a.str("src", "");
    if (src == "force") c.src_kind = ELA3D_SRC_FORCE;
    else if (src == "explosive") c.src_kind = ELA3D_SRC_EXPLOSIVE;
    else if (src == "sigma") c.src_kind = ELA3D_SRC_REFERENCE_SIGMA;
    else if (src == "force3") c.src_kind = ELA3D_SRC_FORCE3;
    else if (!src.empty()) { std::fprintf(stderr, "ela3d_run: unknown src=%s\n", src.c_str()); return 2; }
    const std::string wav = a.str("wavelet", "");
    if (wav == "ricker") c.wavelet = ELA3D_WAVELET_RICKER;
    else if (wav == "gaussian") c.wavelet = ELA3D_WAVELET_GAUSSIAN;
    else if (wav == "dgaussian") c.wavelet = ELA3D_WAVELET_DGAUSSIAN;
    else if (!wav.empty()) { std::fprintf(stderr, "ela3d_run: unknown wavelet=%s\n", wav.c_str()); return 2; }
    c.sample_every = a.integer("sampgat", c.sample_every);
    c.energy_every = a.integer("energy_every", c.energy_every);
    const std::string st = a.str("storage", "fp32");
    c.storage = st == "fp16" ? ELA3D_STORAGE_FP16 : st == "fp64" ? ELA3D_STORAGE_FP64 : ELA3D_STORAGE_FP32;
    const std::string compat = a.str("compat", "");
    if (compat == "reference") c.compat_flags = ELA3D_COMPAT_REFERENCE;
    else if (!compat.empty()) c.compat_flags = uint32_t(std::atoi(compat.c_str()));
    c.dirichlet = a.integer("dirichlet", c.dirichlet);
    c.free_surface = a.integer("free_surface", 0);
    c.device = a.integer("device", 0);

    std::vector<int32_t> rec = receivers(a, c);
    c.nrec = int32_t(rec.size() / 3);
    c.rec = rec.empty() ? nullptr : rec.data();
    check(ela3d_validate(&c), "config");

    // model
    const size_t cells = size_t(c.nx) * c.ny * c.nz;
    std::vector<float> cp(cells), cs(cells), rho(cells);
    const std::string model = a.str("model", "homogeneous");
    if (model == "homogeneous") {
        const float vp = float(a.num("vp", 3300.0));
        check(ela3d_model_homogeneous(c.nx, c.ny, c.nz, vp, float(a.num("vs", vp / 1.732)),
                                      float(a.num("rho", 2800.0)), cp.data(), cs.data(), rho.data()),
              "model");
    } else if (model == "layered") {
        ela3d_layered_spec s;
        ela3d_layered_default(&s);
        s.nx = c.nx; s.ny = c.ny; s.nz = c.nz; s.h = c.dx;
        if (a.has("seed")) s.seed = uint64_t(std::strtoull(a.str("seed", "0").c_str(), nullptr, 10));
        double vmax = 0.0;
        check(ela3d_model_layered(&s, cp.data(), cs.data(), rho.data(), &vmax), "model");
    } else if (model == "anticline") {
        if (c.nx != c.ny || c.nx != c.nz) {
            std::fprintf(stderr, "ela3d_run: model=anticline needs nx = ny = nz\n");
            return 2;
        }
        check(ela3d_model_anticline(c.nx, cp.data(), cs.data(), rho.data()), "model");
    } else if (model.compare(0, 4, "raw:") == 0) {
        auto parts = split(model.substr(4), ':');
        const int bpv = parts.size() > 1 ? std::atoi(parts[1].c_str()) : 8;
        const auto files = split(parts[0], ',');
        if (files.size() != 3) {
            std::fprintf(stderr, "ela3d_run: model=raw:CP,CS,RHO[:4|8]\n");
            return 2;
        }
        check(ela3d_read_raw(files[0].c_str(), int64_t(cells), bpv, cp.data()), "read cp");
        check(ela3d_read_raw(files[1].c_str(), int64_t(cells), bpv, cs.data()), "read cs");
        check(ela3d_read_raw(files[2].c_str(), int64_t(cells), bpv, rho.data()), "read rho");
    } else {
        std::fprintf(stderr, "ela3d_run: unknown model=%s\n", model.c_str());
        return 2;
    }

    ela3d_shot* shot = nullptr;
    check(ela3d_create(&c, cp.data(), cs.data(), rho.data(), &shot), "create");
    std::vector<float>().swap(cp);
    std::vector<float>().swap(cs);
    std::vector<float>().swap(rho);
    if (a.has("resume")) check(ela3d_checkpoint_load(shot, a.str("resume", "").c_str()), "resume");
    ela3d_info info;
    check(ela3d_info_get(shot, &info), "info");

    const std::string out = a.str("out", "");
    const int it_output = a.integer("it_output", 0);
    const int ck_every = a.integer("checkpoint_every", 0);
    const std::string ck = a.str("checkpoint", "");
    const auto t0 = std::chrono::steady_clock::now();
    int done = info.steps_done;
    int rc = ELA3D_OK;
    while (done < c.nstep && rc == ELA3D_OK) {
        int next = c.nstep;
        if (it_output > 0) next = std::min(next, (done / it_output + 1) * it_output);
        if (ck_every > 0 && !ck.empty()) next = std::min(next, (done / ck_every + 1) * ck_every);
        check(ela3d_step(shot, next - done), "step");
        rc = ela3d_sync(shot);
        done = next;
        if (rc != ELA3D_OK) break;
        if (it_output > 0 && done % it_output == 0 && !out.empty()) {
            write_snapshot(shot, ELA3D_FIELD_VZ, "vz", done, out, cells);
            write_snapshot(shot, ELA3D_FIELD_VY, "vy", done, out, cells);
            write_snapshot(shot, ELA3D_FIELD_VX, "vx", done, out, cells);
        }
        if (ck_every > 0 && !ck.empty() && done % ck_every == 0) check(ela3d_checkpoint_save(shot, ck.c_str()), "checkpoint");
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != ELA3D_OK) {
        std::fprintf(stderr, "ela3d_run: stopped at step %d: %s\n", done, ela3d_strerror(rc));
    }
    if (!out.empty()) {
        const bool app = a.integer("append", 0) != 0;
        check(ela3d_info_get(shot, &info), "info");
        std::vector<double> seis(size_t(info.nsamples) * c.nrec * 3);
        if (!seis.empty()) {
            check(ela3d_seismograms(shot, seis.data(), seis.size()), "seismograms");
            const char* names[3] = {"rechorvx.txt", "rechorvy.txt", "rechorvz.txt"};
            for (int q = 0; q < 3; ++q)
                check(ela3d_write_traces((out + "/" + names[q]).c_str(), seis.data(), info.nsamples, c.nrec, 3, q,
                                         app ? 1 : 0, 1),
                      "rechor write");
        }
        std::vector<double> tr((size_t(c.nstep) + 1) * 3);
        check(ela3d_source_trace(shot, tr.data(), tr.size()), "source trace");
        const char* snames[3] = {"recsrcx.txt", "recsrcy.txt", "recsrcz.txt"};
        for (int q = 0; q < 3; ++q)
            check(ela3d_write_traces((out + "/" + snames[q]).c_str(), tr.data(), 1, c.nstep + 1, 3, q, 0, 0),
                  "recsrc write");
        if (info.nenergy > 0) {
            std::vector<double> en(size_t(info.nenergy) * 2);
            check(ela3d_energy(shot, en.data(), en.size()), "energy");
            std::FILE* f = std::fopen((out + "/energy.txt").c_str(), "w");
            if (!f) {
                std::fprintf(stderr, "ela3d_run: cannot write energy.txt\n");
                return 1;
            }
            for (int s = 0; s < info.nenergy; ++s)
                std::fprintf(f, "%d %.17g %.17g\n", (s + 1) * c.energy_every, en[2 * s], en[2 * s + 1]);
            std::fclose(f);
        }
    }
    const int ran = done - info.steps_done;
    std::printf("{\"cells\": %zu, \"nstep\": %d, \"steps_run\": %d, \"wall_s\": %.6f, "
                "\"mcell_updates_per_s\": %.2f, \"status\": \"%s\"}\n",
                cells, c.nstep, ran, wall, wall > 0 ? double(cells) * ran / wall / 1e6 : 0.0,
                rc == ELA3D_OK ? "ok" : ela3d_strerror(rc));
    ela3d_destroy(shot);
    return rc == ELA3D_OK ? 0 : 1;
}
