// geom.hpp -- launch geometry shared by the stencil kernels and host-side checks:
// block renumbering over the XCDs, tile enumeration of a box, cell offsets, lane
// geometry, C-PML layer indices and memory-variable offsets, and the fused
// kernel's thread-to-cell map.  Every function is __host__ __device__,
// so a CPU program can enumerate exactly the addresses the kernels form.
#pragma once

#include "kernels.hpp"

namespace ela3d {

// Blocks are dispatched round-robin over the 8 XCDs; renumber so XCD x owns
// the contiguous logical range [x*q + min(x,r), ...).  A bijection for any n.
__host__ __device__ inline int xcd_block(int b, int n)
{
    const int X = 8;
    const int q = n / X, r = n % X;
    const int x = b % X, l = b / X;
    return x * q + (x < r ? x : r) + l;
}

// Box of logical block lb in a launch over nbox boxes (blk0 ascending).
__host__ __device__ inline Box pick_box(const Box* bx, int nbox, int lb)
{
    Box B = bx[0];
    if (nbox > 1 && lb >= bx[1].blk0) B = bx[1];
    if (nbox > 2 && lb >= bx[2].blk0) B = bx[2];
    if (nbox > 3 && lb >= bx[3].blk0) B = bx[3];
    if (nbox > 4 && lb >= bx[4].blk0) B = bx[4];
    return B;
}

// Tile (xy chunk, z segment) of a box's local block index, strip-major: the
// chunks are cut into B.nstrip strips (0: up to 8 strips of at least 8 chunks),
// and each strip's chunks are enumerated for every z segment before the next
// strip starts.
// With the contiguous per-XCD block ranges of xcd_block, each XCD then marches
// one xy strip through z (a narrow box, such as a y absorbing layer, is one
// strip: each XCD takes a run of its whole planes), so the plane above a z
// segment and the rows beside a chunk are re-read from that XCD's L2 rather
// than from HBM.  A bijection onto [0, nchunk) x [0, nzs) for local in
// [0, nchunk*nzs).
__host__ __device__ inline void tile_of(const Box& B, int local, int& chunk, int& zs)
{
    const int nzs = (B.k1 - B.k0 + B.zseg - 1) / B.zseg;
    int ns = B.nstrip;
    if (ns <= 0) {
        ns = B.nchunk / 8;
        ns = ns > 8 ? 8 : ns;
    }
    ns = ns < 1 ? 1 : (ns > B.nchunk ? B.nchunk : ns);
    const int w = B.nchunk / ns, r = B.nchunk % ns;  // strips 0..r-1 hold w+1 chunks
    const int cp = local / nzs;
    const int s = cp < r * (w + 1) ? cp / (w + 1) : r + (cp - r * (w + 1)) / w;
    const int ws = w + (s < r ? 1 : 0);
    const int c0 = s * w + (s < r ? s : r);
    const int off = local - c0 * nzs;
    chunk = c0 + off % ws;
    zs = off / ws;
}

// Element offset of cell (p, k) from the start of a field or model allocation.
// Field and model pointers handed to the stencil kernels point at the start of
// the allocation, g0 guard elements before plane 0, so every neighbour offset
// (o - 1, o - nx, o + nx + VEC - 1, ...) is a non-negative in-bounds 32-bit value.
__host__ __device__ inline unsigned cell_offset(int g0, int nxyp, int k, int p)
{
    return unsigned(g0) + unsigned(k) * unsigned(nxyp) + unsigned(p);
}

// Compact index of grid index c along an axis whose absorbing layers are
// [0, lo) and [hi, n): c below lo, lo + (c - hi) from hi on, -1 in between.
__host__ __device__ inline int layer_index(int c, int lo, int hi)
{
    return c < lo ? c : (c >= hi ? lo + (c - hi) : -1);
}

// Element offsets of the compact C-PML memory-variable families: x layers
// [nz][ny][wx] (cx = layer_index of i), y layers [nz][wy][nx] (cy of j), z
// layers [wz][ny][nx] (cz of k).
__host__ __device__ inline unsigned xmem_offset(int k, int j, int cx, int ny, int wx)
{
    return unsigned((k * ny + j) * wx + cx);
}
__host__ __device__ inline unsigned ymem_offset(int k, int cy, int i, int nx, int wy)
{
    return unsigned((k * wy + cy) * nx + i);
}
__host__ __device__ inline unsigned zmem_offset(int cz, int j, int i, int nx, int ny)
{
    return unsigned((cz * ny + j) * nx + i);
}

// Two-pass kernels: thread t of the workgroup on (chunk, z segment zs) of box B
// owns the VEC consecutive positions p0 .. p0+vec-1 of the flattened plane and
// marches planes [kb, ke); nothing to do when p0 >= B.p1.
__host__ __device__ inline int lane_p0(const Box& B, int chunk, int block, int t, int vec)
{
    return B.p0 + (chunk * block + t) * vec;
}
__host__ __device__ inline void segment_planes(const Box& B, int zs, int& kb, int& ke)
{
    kb = B.k0 + zs * B.zseg;
    ke = kb + B.zseg < B.k1 ? kb + B.zseg : B.k1;
}

// Per-lane geometry of the VEC elements p0..p0+VEC-1 of a box ending at p1: grid
// indices, whether each lies in the box, and its compact x (and, in an
// absorbing-layer box, y) layer index.  lo/hi: the compact C-PML regions.
template <int V>
struct Lane {
    int p0;
    int i[V], j[V];
    bool act[V];   // inside this box
    int cx[V], cy[V];
};

template <int V>
__host__ __device__ inline void lane_geometry(int nx, const int* lo, const int* hi, bool slab, int p0, int p1,
                                              Lane<V>& L)
{
    L.p0 = p0;
    int j = p0 / nx, i = p0 - j * nx;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        L.i[e] = i;
        L.j[e] = j;
        L.act[e] = (p0 + e) < p1;
        // elements past the box end own no memory variables (their j may exceed ny-1)
        L.cx[e] = L.act[e] ? layer_index(i, lo[0], hi[0]) : -1;
        L.cy[e] = (slab && L.act[e]) ? layer_index(j, lo[1], hi[1]) : -1;
        if (++i == nx) {
            i = 0;
            ++j;
        }
    }
}

// Fused interior step: a workgroup of `nt` threads holds one region of the
// flattened interior plane -- its `n` own cells [q0, q0+n) and one row of `nx`
// cells below and above them -- one cell per thread: threads [0, n) the own
// cells, then the row below (offsets -nx .. -1 from q0), then the row above
// (n .. n+nx-1).  Returns the region offset c of thread t (relative to q0), or
// kNoCell for a thread past the region of a short last tile.
constexpr int kNoCell = -(1 << 30);
__host__ __device__ inline int fused_cell(int t, int n, int nx)
{
    if (t < n) return t;
    if (t < n + nx) return t - n - nx;
    if (t < n + 2 * nx) return t - nx;
    return kNoCell;
}

// Own cells of a full fused tile of nt threads: the region minus two halo rows,
// rounded down to a multiple of 32 so that every tile starts on a 128-byte line
// (box starts are multiples of 64 cells) and each wave's 64 own cells span two
// lines, not three; the threads left over idle.
__host__ __device__ inline int fused_own(int nt, int nx) { return (nt - 2 * nx) / 32 * 32; }

}  // namespace ela3d
