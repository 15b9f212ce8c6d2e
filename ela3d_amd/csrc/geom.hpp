// geom.hpp -- launch geometry shared by the stencil kernels and host-side checks:
// block renumbering over the XCDs, tile enumeration of a box, cell offsets, and
// the fused kernel's thread-to-cell map.  Every function is __host__ __device__,
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

// Tile (xy chunk, z segment) of a box's local block index, strip-major: the
// chunks are cut into up to 8 strips of at least 8 chunks, and each strip's
// chunks are enumerated for every z segment before the next strip starts.
// With the contiguous per-XCD block ranges of xcd_block, each XCD then marches
// one xy strip through z (a narrow box, such as a y absorbing layer, is one
// strip: each XCD takes a run of its whole planes), so the plane above a z
// segment and the rows beside a chunk are re-read from that XCD's L2 rather
// than from HBM.  A bijection onto [0, nchunk) x [0, nzs) for local in
// [0, nchunk*nzs).
__host__ __device__ inline void tile_of(const Box& B, int local, int& chunk, int& zs)
{
    const int nzs = (B.k1 - B.k0 + B.zseg - 1) / B.zseg;
    int ns = B.nchunk / 8;
    ns = ns < 1 ? 1 : (ns > 8 ? 8 : ns);
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

// Own cells of a full fused tile of nt threads (region minus two halo rows).
__host__ __device__ inline int fused_own(int nt, int nx) { return nt - 2 * nx; }

}  // namespace ela3d
