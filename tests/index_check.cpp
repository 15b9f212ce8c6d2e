// index_check.cpp -- CPU enumeration of every device address the stencil kernels
// form, checked against the sizes of the arrays a shot allocates.  Test
// infrastructure (run by tests/test_index_check.py), not part of the library.
//
// The launch geometry comes from the library itself (make_plan, geometry in
// libela3d.so) and every per-thread quantity from the functions the kernels call
// (geom.hpp: xcd_block, pick_box, tile_of, lane_p0, segment_planes, lane_geometry,
// cell_offset, layer_index, x/y/zmem_offset, fused_own, fused_cell).  What this
// file restates is only the set of neighbour offsets each kernel forms around a
// cell (kernels.hip stress_load/velocity_load and the queue set-up in
// stress_body/velocity_body; fused.hip fload/qinit and the velocity update).
//
// usage: index_check nx ny nz npml storage compat block zseg vec block_slab
//                    zseg_slab vec_slab merged fused tile [mutation [free_surface [strips]]]
// mutation 1: offsets formed from plane 0 instead of the allocation start (the
// wrap of unsigned neighbour offsets at the first row); 2: field arrays one plane
// short; 3: each fused tile one own cell short (a cell no tile owns).  Prints
// "accesses A violations V" and exits 0 (2 on bad arguments).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "core.hpp"
#include "geom.hpp"
#include "shot.hpp"

using namespace ela3d;

namespace {

struct Limits {
    long long field = 0;     // elements of one field or model array
    long long mem[3] = {};   // elements of one memory variable of each family
    int lds = 0;             // fused: LDS plane-buffer length (tile)
};

struct Count {
    long long acc = 0, bad = 0;
    void check(long long o, long long lim, const char* what)
    {
        ++acc;
        if (o < 0 || o >= lim) {
            if (bad < 8) std::printf("  out of bounds: %s offset %lld limit %lld\n", what, o, lim);
            ++bad;
        }
    }
};

int g_mutation = 0;

// element offset of (k, p) as the kernels form it (unsigned, from the allocation
// start), widened back to signed so a wrapped offset shows as out of bounds
long long off(int g0, int nxyp, int k, int p, long long d)
{
    const unsigned base = g_mutation == 1 ? cell_offset(0, nxyp, k, p) : cell_offset(g0, nxyp, k, p);
    const unsigned o = base + unsigned(d);
    return g_mutation == 1 ? (long long)o + g0 : (long long)o;  // plane-0 pointers sit g0 in
}

// The two-pass launches (interior, absorbing-layer or mixed): every workgroup,
// thread, VEC element and plane of its segment.
template <int V>
void split_launch(const ShotPlan& P, const LaunchGeom& G, const Limits& L, Count& C)
{
    const ela3d_config& c = P.cfg;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int nxyp = int(plane_stride(nx, ny)), g0 = int(guard_elems(nx));
    const int lo[3] = {P.ax[0].lo, P.ax[1].lo, P.ax[2].lo}, hi[3] = {P.ax[0].hi, P.ax[1].hi, P.ax[2].hi};
    const int w[3] = {P.ax[0].width(), P.ax[1].width(), P.ax[2].width()};
    for (int b = 0; b < G.nblocks; ++b) {
        const int lb = xcd_block(b, G.nblocks);
        const Box B = G.mode == kInterior ? G.box[0] : pick_box(G.box, G.nbox, lb);
        const bool slab = G.mode == kSlab || (G.mode == kMixed && B.mode == kSlab);
        int chunk, zs;
        tile_of(B, lb - B.blk0, chunk, zs);
        int kb, ke;
        segment_planes(B, zs, kb, ke);
        for (int t = 0; t < G.block; ++t) {
            const int p0 = lane_p0(B, chunk, G.block, t, V);
            if (p0 >= B.p1 || kb >= B.k1) continue;
            Lane<V> Ln;
            lane_geometry<V>(nx, lo, hi, slab, p0, B.p1, Ln);
            for (int e = 0; e < V; ++e) {
                // queue set-up: the segment's first plane and the plane below it
                C.check(off(g0, nxyp, kb, p0, e), L.field, "queue");
                if (kb >= 1) C.check(off(g0, nxyp, kb, p0, e - (long long)nxyp), L.field, "queue below");
            }
            for (int k = kb; k < ke; ++k) {
                const long long dU = k == nz - 1 ? 0 : nxyp;  // no plane above the top
                const long long nb[6] = {0, dU, 1, -1, nx, -nx};
                for (int e = 0; e < V; ++e)
                    for (long long d : nb) C.check(off(g0, nxyp, k, p0, d + e), L.field, "field");
                const int cz = slab ? layer_index(k, lo[2], hi[2]) : -1;
                for (int e = 0; e < V; ++e) {
                    if (Ln.cx[e] >= 0) C.check(xmem_offset(k, Ln.j[e], Ln.cx[e], ny, w[0]), L.mem[0], "x memory");
                    if (slab && Ln.cy[e] >= 0)
                        C.check(ymem_offset(k, Ln.cy[e], Ln.i[e], nx, w[1]), L.mem[1], "y memory");
                    if (slab && cz >= 0 && Ln.act[e])
                        C.check(zmem_offset(cz, Ln.j[e], Ln.i[e], nx, ny), L.mem[2], "z memory");
                }
            }
        }
    }
}

// The fused interior launch (fused.hip k_fused): one thread per region cell.
void fused_launch(const ShotPlan& P, const LaunchGeom& G, const Limits& L, Count& C)
{
    const ela3d_config& c = P.cfg;
    const int nx = c.nx, ny = c.ny;
    const int nxyp = int(plane_stride(nx, ny)), g0 = int(guard_elems(nx));
    const Box B = G.box[0];
    const int w0 = P.ax[0].width();
    const int own_full = fused_own(G.tile, nx);
    const int nzs = (B.k1 - B.k0 + B.zseg - 1) / B.zseg;
    std::vector<int> tiles(size_t(B.nchunk) * size_t(nzs), 0);  // blocks per (chunk, z segment)
    for (int b = 0; b < G.nblocks; ++b) {
        int chunk, zs;
        tile_of(B, xcd_block(b, G.nblocks), chunk, zs);
        if (chunk >= 0 && chunk < B.nchunk && zs >= 0 && zs < nzs) ++tiles[size_t(zs) * B.nchunk + chunk];
        int kb, ke;
        segment_planes(B, zs, kb, ke);
        const int q0 = B.p0 + chunk * own_full;
        const int n = std::min(own_full, B.p1 - q0);
        for (int t = 0; t < G.block; ++t) {
            const int cc = fused_cell(t, n, nx);
            const bool act = cc != kNoCell, own = t < n;
            const int p = q0 + (act ? cc : 0);
            const bool inb = act && p >= B.p0 && p < B.p1;
            const int j = p / nx, i = p - j * nx;
            const int cx = layer_index(i, P.ax[0].lo, P.ax[0].hi);
            const bool inx = cx >= 0;
            if (act) C.check(cc + nx, L.lds, "LDS plane");  // li = c + nx
            auto computes = [&](int m) {
                if (m < kb - 1 || m > ke || m < B.k0 || m >= B.k1) return false;
                return (m == kb - 1 || m == ke) ? own : inb;
            };
            auto loads = [&](int m) {  // fload at plane m: its own plane, or the safe offset (plane kb)
                const int pl = computes(m) ? m : kb;
                const int shift = m == kb - 1 ? 1 : (m == ke ? -1 : 0);  // dA, dB
                const long long nb[6] = {0, nxyp, 1, -1, nx, -nx};
                for (int s : {0, shift})
                    for (long long d : nb) C.check(off(g0, nxyp, pl + s, p, d), L.field, "fused field");
                if (computes(m) && inx)
                    C.check(xmem_offset(m, j, cx, ny, w0), L.mem[0], "fused x memory");
            };
            const int m0 = computes(kb - 1) ? kb - 1 : kb;
            if (computes(m0)) {  // qinit: the plane, the plane below, and plane kb for kb-1
                C.check(off(g0, nxyp, m0, p, 0), L.field, "fused queue");
                C.check(off(g0, nxyp, m0, p, -(long long)nxyp), L.field, "fused queue below");
                if (m0 == kb - 1) C.check(off(g0, nxyp, kb, p, 0), L.field, "fused queue kb");
            }
            for (int m = kb - 1; m <= ke; ++m) {
                loads(m);
                if (m > kb && own) {  // velocities of plane m-1: LDS neighbours and the stores
                    const int k = m - 1;
                    for (int d : {0, -1, 1, -nx, nx}) C.check(cc + nx + d, L.lds, "fused LDS read");
                    C.check(off(g0, nxyp, k, p, 0), L.field, "fused velocity");
                    if (inx) C.check(xmem_offset(k, j, cx, ny, w0), L.mem[0], "fused velocity memory");
                }
            }
        }
    }
    // every (chunk, z segment) is one block's (tile_of is a bijection for the strip
    // count in use)
    if (G.nblocks != B.nchunk * nzs) {
        std::printf("  %d blocks for %d tiles\n", G.nblocks, B.nchunk * nzs);
        ++C.bad;
    }
    for (size_t q = 0; q < tiles.size(); ++q)
        if (tiles[q] != 1) {
            if (C.bad < 8) std::printf("  tile %zu taken by %d blocks\n", q, tiles[q]);
            ++C.bad;
        }
    // every cell of the box is owned by exactly one tile, and every LDS cell an own
    // cell's velocity update uses is written by a region thread of its tile
    std::vector<int> owner(size_t(B.p1 - B.p0), 0);
    for (int ch = 0; ch < B.nchunk; ++ch) {
        const int q0 = B.p0 + ch * own_full;
        const int n = std::min(own_full, B.p1 - q0) - (g_mutation == 3 ? 1 : 0);
        std::vector<char> written(size_t(G.tile), 0);
        std::vector<long long> reads;
        for (int t = 0; t < G.block; ++t) {
            const int cc = fused_cell(t, n, nx);
            if (cc == kNoCell) continue;
            if (cc + nx >= 0 && cc + nx < G.tile) written[size_t(cc + nx)] = 1;
            if (t < n) {
                ++owner[size_t(q0 + cc - B.p0)];
                const int i = (q0 + cc) % nx;
                // sxx(i-1) is used where i >= 1, sxy/sxz(i+1) where i <= nx-2
                if (i >= 1) reads.push_back(cc + nx - 1);
                if (i <= nx - 2) reads.push_back(cc + nx + 1);
                reads.push_back(cc);           // j-1
                reads.push_back(cc + 2 * nx);  // j+1
            }
        }
        for (long long r : reads)
            if (r < 0 || r >= G.tile || !written[size_t(r)]) {
                if (C.bad < 8) std::printf("  LDS cell %lld used but not written (tile %d)\n", r, ch);
                ++C.bad;
            }
    }
    for (size_t q = 0; q < owner.size(); ++q)
        if (owner[q] != 1) {
            if (C.bad < 8) std::printf("  box cell %zu owned %d times\n", q, owner[q]);
            ++C.bad;
        }
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 16) {
        std::fprintf(stderr, "usage: index_check nx ny nz npml storage compat block zseg vec block_slab "
                             "zseg_slab vec_slab merged fused tile [mutation [free_surface [strips]]]\n");
        return 2;
    }
    int a[18] = {};
    for (int q = 1; q < argc && q <= 18; ++q) a[q - 1] = std::atoi(argv[q]);
    g_mutation = argc > 16 ? a[15] : 0;
    ela3d_config c;
    ela3d_config_default(&c);
    c.nx = a[0]; c.ny = a[1]; c.nz = a[2]; c.npoints_pml = a[3];
    c.storage = a[4]; c.compat_flags = a[5];
    c.isrc = c.nx / 2; c.jsrc = c.ny / 2; c.ksrc = c.nz / 2;
    c.nstep = 2; c.nrec = 0; c.rec = nullptr; c.energy_every = 0;
    c.free_surface = argc > 17 ? a[16] : 0;
    Tuning t;
    t.block = a[6]; t.zseg_interior = a[7]; t.vec = a[8];
    t.block_slab = a[9]; t.zseg_slab = a[10]; t.vec_slab = a[11];
    t.merged = a[12]; t.fused = a[13]; t.tile = a[14];
    t.strips = argc > 18 ? a[17] : 0;
    t.dbuf = t.fused;
    try {
        const ShotPlan P = make_plan(c);
        if (t.fused) {
            const std::string why = fused_unsupported(P, t.tile);
            if (!why.empty()) {
                std::printf("refused %s\n", why.c_str());
                return 0;
            }
        }
        LaunchGeom slab, inter;
        geometry(P, t, slab, inter);
        Limits L;
        L.field = plane_stride(c.nx, c.ny) * c.nz + 2 * guard_elems(c.nx);
        if (g_mutation == 2) L.field -= plane_stride(c.nx, c.ny);
        L.mem[0] = (long long)c.nz * c.ny * P.ax[0].width();
        L.mem[1] = (long long)c.nz * P.ax[1].width() * c.nx;
        L.mem[2] = (long long)P.ax[2].width() * c.ny * c.nx;
        L.lds = t.tile;
        Count C;
        for (const LaunchGeom* G : {&slab, &inter}) {
            if (G->nblocks == 0) continue;
            if (G->mode == kFused)
                fused_launch(P, *G, L, C);
            else if (G->vec == 4)
                split_launch<4>(P, *G, L, C);
            else if (G->vec == 2)
                split_launch<2>(P, *G, L, C);
            else
                split_launch<1>(P, *G, L, C);
        }
        std::printf("accesses %lld violations %lld\n", C.acc, C.bad);
    } catch (const Error& e) {
        std::printf("refused %s\n", e.msg.c_str());
    }
    return 0;
}
