// tail.hip -- per-step tail kernel (receivers, energy reduction, source, step
// counter) and the stream-copy kernel used to calibrate the HBM ceiling.
#include "device_util.hpp"

namespace ela3d {

// scale of the data written this step (FP16 storage; 1 otherwise)
template <typename T>
__device__ __forceinline__ float nsc(const TailArgs<T>& A, int q)
{
    if constexpr (std::is_same<T, half_t>::value) return A.scl[54 + q];
    return 1.f;
}
template <typename T>
__device__ __forceinline__ float nisc(const TailArgs<T>& A, int q)
{
    if constexpr (std::is_same<T, half_t>::value) return A.scl[81 + q];
    return 1.f;
}

// source injection into field q at cell o, recording the magnitude written
template <typename T>
__device__ __forceinline__ void inject(const TailArgs<T>& A, int q, int o, typename Arith<T>::R v, float* inj)
{
    stf(A.f[q], o, v, nsc<T>(A, q));
    inj[q] = fabsf((float)v);
}

// stress injection at the source cell: into the side slot where FP16 storage
// keeps the source cell's normal stresses there (see StepArgs::side)
template <typename T>
__device__ __forceinline__ void inject_sigma(const TailArgs<T>& A, int q, int o, typename Arith<T>::R add, float* inj)
{
    if constexpr (std::is_same<T, half_t>::value) {
        if (A.side) {
            A.side[q - 3] = A.side[q - 3] + (float)add;
            return;
        }
    }
    inject(A, q, o, ldf(A.f[q], o, nisc<T>(A, q)) + add, inj);
}

template <typename T>
__device__ __forceinline__ double sigma_at_source(const TailArgs<T>& A, int q, int o)
{
    if constexpr (std::is_same<T, half_t>::value) {
        if (A.side) return (double)A.side[q - 3];
    }
    return (double)ldf(A.f[q], o, nisc<T>(A, q));
}

// ---------------------------------------------------------------------------
// Per-step tail (one workgroup): receivers (before the source, as CPU:593-602
// precedes CPU:604-622), energy reduction, source injection, step counter.
// ---------------------------------------------------------------------------
constexpr int kTailThreads = 1024;  // wide enough to sum ~10^5 energy partials quickly

template <typename T>
__device__ __forceinline__ void tail_body(const TailArgs<T>& A)
{
    typedef typename Arith<T>::R R;
    __shared__ double red[2][kTailThreads / 64];
    __shared__ int bad;
    __shared__ float inj[9];  // |value| written by the source injection, per field
    const int tid = threadIdx.x;
    const int it = *A.step + 1;
    if (tid == 0) bad = 0;
    if (tid < 9) inj[tid] = 0.f;
    __syncthreads();
    if (A.nrec > 0 && it % A.sample_every == 0) {
        const int s = it / A.sample_every - 1;
        if (s < A.nsamples) {
            for (int r = tid; r < A.nrec; r += blockDim.x) {
                const int o = A.rec[3 * r] + A.rec[3 * r + 1] * A.nx + A.rec[3 * r + 2] * A.nxy;
                const double vx = (double)ldf(A.f[0], o, nisc<T>(A, 0));
                const double vy = (double)ldf(A.f[1], o, nisc<T>(A, 1));
                const double vz = (double)ldf(A.f[2], o, nisc<T>(A, 2));
                double* dst = A.seis + (size_t(s) * A.nrec + r) * 3;
                dst[0] = vx;
                dst[1] = vy;
                dst[2] = vz;
                if (!isfinite(vx) || !isfinite(vy) || !isfinite(vz)) bad = 1;
            }
        }
    }
    if (A.energy_every > 0 && it % A.energy_every == 0) {
        const int s = it / A.energy_every - 1;
        double a = 0.0, b = 0.0;
        for (int q = tid; q < A.nepart; q += blockDim.x) {
            a += A.epart[2 * q];
            b += A.epart[2 * q + 1];
        }
        a = wave_sum(a);
        b = wave_sum(b);
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = a;
            red[1][tid >> 6] = b;
        }
        __syncthreads();
        if (tid == 0 && s < A.nenergy) {
            double ea = 0.0, eb = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                ea += red[0][w];
                eb += red[1][w];
            }
            A.energy[2 * s] = ea * A.evol;
            A.energy[2 * s + 1] = eb * A.evol;
            if (!isfinite(ea) || !isfinite(eb)) bad = 1;
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (it - 1 < A.nwavelet) {
            const double st = A.wavelet[it - 1];
            const int o = A.src;
            const R DT = (R)A.dt;
            const R l2m = (R)A.rho[o] * (R)A.cp[o] * (R)A.cp[o];
            switch (A.src_kind) {
            case 0: {  // CPU:620-622
                const R fx = (R)(A.sin_t * st), fy = (R)(A.cos_t * st);
                inject_sigma(A, 3, o, fx * DT * l2m, inj);
                inject_sigma(A, 4, o, fx * DT * l2m, inj);
                inject_sigma(A, 5, o, fy * DT * l2m, inj);
                break;
            }
            case 1: {  // CPU:614-617
                const R fx = (R)(A.sin_t * st), fy = (R)(A.cos_t * st);
                const R rx = ((R)A.rho[o] + (R)A.rho[o - 1]) / (R)2;
                const R ry = ((R)A.rho[o] + (R)A.rho[o + A.nx]) / (R)2;
                inject(A, 0, o, ldf(A.f[0], o, nisc<T>(A, 0)) + fx * DT / rx, inj);
                inject(A, 1, o, ldf(A.f[1], o, nisc<T>(A, 1)) + fy * DT / ry, inj);
                break;
            }
            case 2: {
                const R s = (R)st;
                inject_sigma(A, 3, o, s * DT * l2m, inj);
                inject_sigma(A, 4, o, s * DT * l2m, inj);
                inject_sigma(A, 5, o, s * DT * l2m, inj);
                break;
            }
            default: {  // ZC:451-453
                const R s = (R)st;
                const R rx = ((R)A.rho[o] + (R)A.rho[o - 1]) / (R)2;
                const R ry = ((R)A.rho[o] + (R)A.rho[o + A.nx]) / (R)2;
                const R rz = ((R)A.rho[o] + (R)A.rho[o + A.nxy]) / (R)2;
                inject(A, 0, o, ldf(A.f[0], o, nisc<T>(A, 0)) + s * DT / rx, inj);
                inject(A, 1, o, ldf(A.f[1], o, nisc<T>(A, 1)) + s * DT / ry, inj);
                inject(A, 2, o, ldf(A.f[2], o, nisc<T>(A, 2)) + s * DT / rz, inj);
                break;
            }
            }
        }
        // the velocities and normal stresses at the source cell after injection, all
        // loaded at once: the stresses are the source trace (UNI:463-465), and every
        // step, whatever the receivers and energy cadence, a non-finite one marks the
        // shot unstable (a blow-up of the grid reaches the source cell within a few steps)
        const int o = A.src;
        const double sxx = sigma_at_source(A, 3, o), syy = sigma_at_source(A, 4, o), szz = sigma_at_source(A, 5, o);
        const double vx = ldf(A.f[0], o, nisc<T>(A, 0)), vy = ldf(A.f[1], o, nisc<T>(A, 1));
        const double vz = ldf(A.f[2], o, nisc<T>(A, 2));
        if (it < A.nsrctrace) {
            A.srctrace[3 * it] = sxx;
            A.srctrace[3 * it + 1] = syy;
            A.srctrace[3 * it + 2] = szz;
        }
        const bool fin = isfinite(sxx) && isfinite(syy) && isfinite(szz) && isfinite(vx) && isfinite(vy) && isfinite(vz);
        if (bad || !fin) *A.flag = 1;
        *A.step = it;
    }
    if constexpr (std::is_same<T, half_t>::value) {
        // FP16 storage: the data now sits at the "next" scales; make them current
        // and choose the scales of the next step from this step's maxima, every
        // rescale_every steps, or early when the stored maximum, grown once more
        // by its last step's growth g, would pass 2^13 (headroom below 65504), or
        // when it is below 2^8 (precision above the subnormals).
        __syncthreads();
        if (tid < 27) {
            const int q = tid;
            unsigned mb = 0;
            for (int s = 0; s < kMaxSlots; ++s) {
                mb = max(mb, A.maxslot[s * 27 + q]);
                A.maxslot[s * 27 + q] = 0u;
            }
            const float m = fmaxf(__uint_as_float(mb), q < 9 ? inj[q] : 0.f);
            float sc = A.scl[54 + q];
            A.scl[q] = sc;
            A.scl[27 + q] = A.scl[81 + q];
            const float mp = A.scl[108 + q];  // the field's maximum one step earlier
            A.scl[108 + q] = m;
            const float g = (mp > 0.f && m > mp) ? fminf(m / mp, 256.f) : 1.f;
            const float stored = m * sc;
            const bool due = (it % A.rescale_every == 0) || stored * g > 8192.f || stored < 256.f;
            if (due && m > 0.f && isfinite(m)) {
                int e, es;
                frexpf(m * g, &e);  // m * g <= 2^e
                frexpf(sc, &es);    // sc = 2^(es - 1)
                // stored maximum in (2^11, 2^12] / g, so one more step of growth g
                // lands it in (2^11, 2^12].  Where the growth is unknown or extreme
                // (a field that has just turned non-zero: a wavefront's leading edge
                // reaching a C-PML slab, whose maximum can grow a thousandfold in one
                // step), a scale rises at most 4x per step.
                const bool blind = !(mp > 0.f) || m > 256.f * mp;
                int x = min(100, max(-100, 12 - e));
                if (blind) x = min(x, es + 1);
                sc = ldexpf(1.f, x);
            }
            A.scl[54 + q] = sc;
            A.scl[81 + q] = 1.f / sc;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_tail(const TailArgs<T> A)
{
    tail_body(A);
}

// The tails of a batch of shots stepped together: one workgroup per shot, its
// arguments from device memory.
template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_tail_multi(const TailArgs<T>* __restrict__ args)
{
    tail_body(args[blockIdx.x]);
}

// HBM-ceiling calibration copies.  variant 0: grid-stride float4; 1: one
// float4 per thread, one pass; 2: four float4 per thread with nontemporal
// loads and stores; 3: one float per thread (the stencil kernels' access width).
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ s, float4* __restrict__ d, long n)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

__global__ void __launch_bounds__(256) k_copy1(const float4* __restrict__ s, float4* __restrict__ d, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy4nt(const f4v* __restrict__ s, f4v* __restrict__ d, long n)
{
    const long base = blockIdx.x * 1024L + threadIdx.x;
    f4v v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long i = base + 256L * u;
        if (i < n) v[u] = __builtin_nontemporal_load(s + i);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long i = base + 256L * u;
        if (i < n) __builtin_nontemporal_store(v[u], d + i);
    }
}

__global__ void __launch_bounds__(256) k_copyf(const float* __restrict__ s, float* __restrict__ d, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// Multi-stream calibration: NR arrays read and NW arrays written element-wise in
// one pass, W floats per lane -- the access pattern of a stencil pass without
// its neighbour loads (stress pass: 12 read, 6 written; velocity: 10 and 3;
// one fused step: 12 and 9).
struct StreamSet {
    const float* in[12];
    float* out[9];
    long n;  // floats per array
};

template <int NR, int NW, int W>
__global__ void __launch_bounds__(256) k_streams(const StreamSet S)
{
    typedef float vt __attribute__((ext_vector_type(W)));
    const long i = (blockIdx.x * (long)blockDim.x + threadIdx.x) * W;
    if (i + W > S.n) return;
    vt acc = 0.f;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        vt v;
        __builtin_memcpy(&v, S.in[r] + i, sizeof(vt));
        acc += v;
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const vt o = acc + (float)w;
        __builtin_memcpy(S.out[w] + i, &o, sizeof(vt));
    }
}

// The same 12-read / 6-write pattern with the arrays interleaved in blocks of B
// floats: the reads live in one region as [n/B][12][B], the writes as [n/B][6][B].
// Requires n % B == 0, so every address stays below 12n (reads) and 6n (writes).
template <int NR, int NW, int B>
__global__ void __launch_bounds__(256) k_streams_blocked(const float* __restrict__ in, float* __restrict__ out,
                                                         long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long blk = i / B, lane = i % B;
    float acc = 0.f;
#pragma unroll
    for (int r = 0; r < NR; ++r) acc += in[(blk * NR + r) * B + lane];
#pragma unroll
    for (int w = 0; w < NW; ++w) out[(blk * NW + w) * B + lane] = acc + (float)w;
}

int stream_counts(int variant, int& nr, int& nw)
{
    switch (variant) {
    case 8: nr = 12; nw = 6; return 1;
    case 4: nr = 12; nw = 6; return 1;
    case 5: nr = 10; nw = 3; return 1;
    case 6: nr = 12; nw = 9; return 1;
    case 7: nr = 12; nw = 6; return 4;
    default: nr = 1; nw = 1; return 4;
    }
}

void launch_streams(float* base, long n, int variant, hipStream_t s)
{
    int nr = 0, nw = 0;
    const int w = stream_counts(variant, nr, nw);
    StreamSet S;
    for (int r = 0; r < 12; ++r) S.in[r] = base + long(r < nr ? r : 0) * n;
    for (int q = 0; q < 9; ++q) S.out[q] = base + long(nr + (q < nw ? q : 0)) * n;
    S.n = n;
    const unsigned blocks = (unsigned)((n / w + 255) / 256);
    if (variant == 8) {
        hipLaunchKernelGGL((k_streams_blocked<12, 6, kStreamBlock>), dim3(blocks), dim3(256), 0, s, base,
                           base + 12 * n, n);
        return;
    }
    switch (variant) {
    case 4: hipLaunchKernelGGL((k_streams<12, 6, 1>), dim3(blocks), dim3(256), 0, s, S); break;
    case 5: hipLaunchKernelGGL((k_streams<10, 3, 1>), dim3(blocks), dim3(256), 0, s, S); break;
    case 6: hipLaunchKernelGGL((k_streams<12, 9, 1>), dim3(blocks), dim3(256), 0, s, S); break;
    default: hipLaunchKernelGGL((k_streams<12, 6, 4>), dim3(blocks), dim3(256), 0, s, S); break;
    }
}

// First level of the energy reduction: workgroup b sums the partial pairs
// [b*per, min((b+1)*per, n)) into out[b].  The assignment of partials to lanes
// and the reduction tree are fixed, so the sum is the same on every run.
constexpr int kEpartThreads = 256;

__global__ void __launch_bounds__(kEpartThreads) k_epart(const double* __restrict__ in, int n, int per,
                                                        double* __restrict__ out)
{
    __shared__ double red[2][kEpartThreads / 64];
    const int tid = threadIdx.x;
    const int q0 = blockIdx.x * per, q1 = min(q0 + per, n);
    double a = 0.0, b = 0.0;
    for (int q = q0 + tid; q < q1; q += kEpartThreads) {
        a += in[2 * q];
        b += in[2 * q + 1];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = a;
        red[1][tid >> 6] = b;
    }
    __syncthreads();
    if (tid == 0) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < kEpartThreads / 64; ++w) {
            x += red[0][w];
            y += red[1][w];
        }
        out[2 * blockIdx.x] = x;
        out[2 * blockIdx.x + 1] = y;
    }
}

void launch_epart(const double* in, int n, double* out, int nout, hipStream_t s)
{
    if (n <= 0 || nout <= 0 || nout > kEpartBlocks) return;
    const int per = (n + nout - 1) / nout;
    hipLaunchKernelGGL(k_epart, dim3(nout), dim3(kEpartThreads), 0, s, in, n, per, out);
}

// ---------------------------------------------------------------------------
template <typename T>
void launch_tail(const TailArgs<T>& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_tail<T>), dim3(1), dim3(kTailThreads), 0, s, a);
}

// Zero fill of a device range (256-byte aligned start, any length): one uint4 per
// lane, then the tail bytes.  A kernel rather than hipMemsetAsync so that a caller
// may capture it in a graph.
__global__ void __launch_bounds__(256) k_zero(uint4* __restrict__ d, long n16, unsigned char* __restrict__ tail,
                                              int ntail)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n16) d[i] = make_uint4(0u, 0u, 0u, 0u);
    if (i < ntail) tail[i] = 0;
}

// Halo planes of a z-slab: field plane <-> contiguous buffer, one element per lane.
template <typename T>
struct HaloPlanes {
    T* f[3];
};
template <typename T>
__global__ void __launch_bounds__(256) k_halo(const HaloPlanes<T> H, int nf, long n, T* __restrict__ buf, int pack)
{
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (e >= n || q >= nf) return;
    if (pack)
        buf[q * n + e] = H.f[q][e];
    else
        H.f[q][e] = buf[q * n + e];
}

template <typename T>
void launch_halo(T* const* fields, int nf, long n, T* buf, bool pack, hipStream_t s)
{
    if (nf <= 0 || nf > 3 || n <= 0) return;
    HaloPlanes<T> H{};
    for (int q = 0; q < nf; ++q) H.f[q] = fields[q];
    hipLaunchKernelGGL((k_halo<T>), dim3(unsigned((n + 255) / 256), unsigned(nf)), dim3(256), 0, s, H, nf, n, buf,
                       pack ? 1 : 0);
}
template void launch_halo<float>(float* const*, int, long, float*, bool, hipStream_t);
template void launch_halo<double>(double* const*, int, long, double*, bool, hipStream_t);
template void launch_halo<half_t>(half_t* const*, int, long, half_t*, bool, hipStream_t);

void launch_zero(void* p, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return;
    const long n16 = long(bytes / 16);
    const int ntail = int(bytes % 16);
    const long n = std::max<long>(n16, ntail);
    const long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<uint4*>(p), n16,
                       static_cast<unsigned char*>(p) + n16 * 16, ntail);
}

void launch_copy(const float4* src, float4* dst, long n4, int variant, hipStream_t s)
{
    if (variant == 1) {
        hipLaunchKernelGGL(k_copy1, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s, src, dst, n4);
    } else if (variant == 2) {
        hipLaunchKernelGGL(k_copy4nt, dim3((unsigned)((n4 + 1023) / 1024)), dim3(256), 0, s,
                           (const f4v*)src, (f4v*)dst, n4);
    } else if (variant == 3) {
        const long nf = 4 * n4;
        hipLaunchKernelGGL(k_copyf, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, (const float*)src,
                           (float*)dst, nf);
    } else {
        long blocks = (n4 + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n4);
    }
}

template <typename T>
void launch_tail_multi(const TailArgs<T>* args, int nshot, hipStream_t s)
{
    if (nshot > 0) hipLaunchKernelGGL((k_tail_multi<T>), dim3(nshot), dim3(kTailThreads), 0, s, args);
}
template void launch_tail_multi<float>(const TailArgs<float>*, int, hipStream_t);
template void launch_tail<float>(const TailArgs<float>&, hipStream_t);
template void launch_tail<double>(const TailArgs<double>&, hipStream_t);
template void launch_tail<half_t>(const TailArgs<half_t>&, hipStream_t);

}  // namespace ela3d
