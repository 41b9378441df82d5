// Mixed-precision BSE matvec Y = H X on the FP32 store with split-FP16 MFMA
// (gfx950 v_mfma_f32_16x16x32_f16, FP32 accumulation).
//
// Every FP32 operand is split into two scaled FP16 halves, a = s (a_hi + a_lo)
// with a_hi = f16(a/s) and a_lo = f16(a/s - a_hi), which keeps ~22 bits, and
// the product keeps the three leading terms:
//     a b ~ s t (a_hi b_hi + a_hi b_lo + a_lo b_hi)        (error ~2^-22 |a b|)
// Twelve FP16 MFMAs per complex 16x16x32 block therefore give FP32-class
// products at 4x the FP32 MFMA rate, so the kernel is bound by streaming the
// 8-byte complex FP32 store once per launch (8 N^2 bytes).
//
// H is Hermitian, so row i of H is the conjugate of stored column i: each
// wavefront streams 16 stored columns along k (the four lanes of a column read
// 64 contiguous bytes per load instruction), i.e. the A operand of output rows
// i is conj(H[k + i ld]).  X arrives
// pre-split in FP16 planes (re_hi, re_lo, im_hi, im_lo), column-major with k
// contiguous.  Each step covers 32 k: H slices are loaded into registers two
// steps ahead of their use and X slices two steps ahead, staged into a second
// LDS buffer at the end of the step, with one barrier per step.  Loads carry
// no bounds checks (the store's rows N..ld-1 are zero, indexes are clamped),
// so the compiler keeps the whole prefetch in flight across the MFMA phase.
#include "wtb_internal.h"

#include <cmath>
#include <mutex>
#include <type_traits>
#include <utility>

namespace wtb {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kKT = 32;        // k per step of the split-FP16 kernel (one MFMA depth)
// wavefronts per workgroup (16 rows each): 8 puts two wavefronts on each SIMD
// (a 256-register budget per wavefront), which keeps twice the H stream in
// flight of a 4-wavefront workgroup
constexpr int kWPB = 8;
// H register sets of the split-FP16 kernel: slices are loaded kHD - 1 steps
// ahead of their use (three sets fit the 256-register budget of two
// wavefronts per SIMD up to 128 columns)
constexpr int kHD = 3;

__device__ __forceinline__ f4 mfma16(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Order of k inside each 32-k block of the FP16 X planes.  MFMA lane group g
// (lanes 16g .. 16g+15) supplies 8 k values per 16x16x32 step; they are taken
// as k = 8q + 2g + t (q = 0..3, t = 0..1), so that the four lanes of one H
// column read 64 contiguous bytes with each load instruction, and the X
// planes store those 8 values contiguously (position 8g + 2q + t) so that the
// B operand stays a single 16-byte LDS read.  A and B use the same k set.
__device__ __forceinline__ long kperm(long k) {
    const long b = k & ~31L, r = k & 31;
    const long q = r >> 3, g = (r >> 1) & 3, t = r & 1;
    return b + 8 * g + 2 * q + t;
}

// X (N x ncol, complex FP64, column-major) -> four FP16 planes
// P[p][col][kperm(k)] (kpad rows, zero padded), scaled per column.
__global__ void split_x_kernel(long N, long kpad, int ncol, const double *__restrict__ X,
                               const float *__restrict__ xs, _Float16 *__restrict__ P) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int col = blockIdx.y;
    if (k >= kpad || col >= ncol) return;
    float re = 0.f, im = 0.f;
    if (k < N) {
        re = (float)X[2 * (k + (size_t)col * N)] * xs[col];
        im = (float)X[2 * (k + (size_t)col * N) + 1] * xs[col];
    }
    _Float16 rh = (_Float16)re, ih = (_Float16)im;
    _Float16 rl = (_Float16)(re - (float)rh), il = (_Float16)(im - (float)ih);
    const size_t plane = (size_t)ncol * kpad;
    const size_t o = kperm(k) + (size_t)col * kpad;
    P[o] = rh;
    P[plane + o] = rl;
    P[2 * plane + o] = ih;
    P[3 * plane + o] = il;
}

// Per-column scales of X: max |x| over the column (one workgroup per column)
// sets the power of two s with max |x| s < 2^13 (FP16 max 65504, room for the
// products); xs = s scales the split operand, ysc = hscale / s undoes both
// scalings on the output.
constexpr int kCS = 1024;
__global__ void __launch_bounds__(kCS)
col_scales_kernel(long N, const double *__restrict__ X, float hscale, float *__restrict__ xs,
                  float *__restrict__ ysc) {
    __shared__ float red[kCS / 64];
    const int col = blockIdx.x, tid = threadIdx.x;
    const double2 *x = reinterpret_cast<const double2 *>(X) + (size_t)col * N;
    float m = 0.f;
    // eight loads in flight per thread
    for (long k0 = 0; k0 < N; k0 += 8 * kCS) {
        double2 v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const long k = k0 + t * kCS + tid;
            v[t] = k < N ? x[k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) m = fmaxf(m, fmaxf(fabsf((float)v[t].x), fabsf((float)v[t].y)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < kCS / 64; ++w) m = fmaxf(m, red[w]);
        int e = 0;
        if (m > 0.f) frexpf(m, &e);          // m = f 2^e, 0.5 <= f < 1
        const float sc = ldexpf(1.0f, 13 - e);
        xs[col] = sc;
        ysc[col] = hscale / sc;
    }
}

// 8 complex FP32 values of a stored column -> their split FP16 halves, real
// and imaginary parts in separate operands (element e of an operand is the
// e-th value).  Per pair of floats (a0, a1) and power-of-two scale s:
//     hi = (f16(s a0), f16(s a1)),  lo = (f16(s a0 - hi.x), f16(s a1 - hi.y))
// by four v_fma_mix{lo,hi}_f16 (FP32 sources, the scale applied inside the
// fused operation, hi read back as FP16 through op_sel), so the split costs
// two VALU operations per float and no packing.
typedef unsigned u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split_pair(float a0, float a1, float s, unsigned &hi, unsigned &lo) {
    asm("v_fma_mixlo_f16 %0, %2, %4, 0\n\t"
        "v_fma_mixhi_f16 %0, %3, %4, 0\n\t"
        "v_fma_mixlo_f16 %1, %2, %4, -%0 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %1, %3, %4, -%0 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
        : "=&v"(hi), "=&v"(lo)
        : "v"(a0), "v"(a1), "v"(s));
}

struct Split {
    h8 rh, rl, ih, il;
};

__device__ __forceinline__ Split split8(const float4 *h, float hinv) {
    u4 rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        unsigned a, b;
        split_pair(h[q].x, h[q].z, hinv, a, b);
        rh[q] = a;
        rl[q] = b;
        split_pair(h[q].y, h[q].w, hinv, a, b);
        ih[q] = a;
        il[q] = b;
    }
    return Split{__builtin_bit_cast(h8, rh), __builtin_bit_cast(h8, rl), __builtin_bit_cast(h8, ih),
                 __builtin_bit_cast(h8, il)};
}

// sign flip of the eight FP16 values of an operand
__device__ __forceinline__ h8 neg8(h8 v) { return __builtin_bit_cast(h8, __builtin_bit_cast(u4, v) ^ 0x80008000u); }

template <int V>
using IC = std::integral_constant<int, V>;
template <class F, int... J>
__device__ __forceinline__ void run_steps(F &&f, long it, std::integer_sequence<int, J...>) {
    (f(it + J, IC<J>{}), ...);
}
template <class F, int... J>
__device__ __forceinline__ void run_tail(F &&f, long it, long niter, std::integer_sequence<int, J...>) {
    ((it + J < niter ? (f(it + J, IC<J>{}), 0) : 0), ...);
}

// WPB wavefronts per workgroup, each owning 16 output rows; the workgroup's X
// slice is staged once in LDS for all of them.  A step covers kKT = 32 k (one
// MFMA depth).  The H slice of step it + HD - 1 and the X slice of step it + 2
// are loaded into registers while step it computes (HD register sets for H,
// two for X), and the X slice of step it + 1 goes to the second LDS buffer
// at the end of the step, one barrier per step.
// The row element of H is conj(s) for the stored s = sr + i si, so
//     Re y = sr xr + si xi,   Im y = sr xi - si xr.
template <int NCB, int WPB, int HD>
__global__ void __launch_bounds__(64 * WPB)
hx_f16split_kernel(long N, long ld, int ncol, float hinv, const float2 *__restrict__ H,
                   const _Float16 *__restrict__ P, const float *__restrict__ ysc, double *__restrict__ Y) {
    constexpr int KT = kKT;
    constexpr int NC = NCB * 16;
    // halfs per staged column: the 96-byte row stride keeps the B-fragment
    // reads (16 lanes x 16 bytes at rows li, chunk blk) on distinct banks
    // (the counter reads 5x fewer bank-conflict cycles than an 80-byte stride)
    constexpr int XS = KT + 16;
    constexpr int QN = KT / 8;                   // 16-byte chunks per column and plane
    constexpr int NT = 64 * WPB;                 // threads per workgroup
    constexpr int XTOT = 4 * NC * QN;            // 16-byte X chunks per slice
    constexpr int XCH = (XTOT + NT - 1) / NT;    // per thread
    constexpr int U = HD == 3 ? 6 : 2;           // steps per unrolled round (register and LDS sets cycle)
    __shared__ __align__(16) _Float16 xs[2][4][NC][XS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int blk = lane >> 4, li = lane & 15;
    // Rows past N (last workgroup) and X columns past ncol read clamped, valid
    // addresses: their products land in outputs that are never written.
    // Rows N..ld-1 of H and of the X planes are zero, so every k slice of a
    // column is read whole without bounds checks.
    const long row = min((long)blockIdx.x * (16 * WPB) + wave * 16 + li, N - 1);
    const float2 *hcol = H + (size_t)row * ld + 2 * blk;
    const size_t plane = (size_t)ncol * ld;
    const long niter = (N + KT - 1) / KT;

    auto load_x = [&](long it, h8 (&r)[XCH]) {
        const long k0 = min(it, niter - 1) * KT;
#pragma unroll
        for (int t = 0; t < XCH; ++t) {
            const int c = tid + NT * t;
            if (XTOT % NT == 0 || c < XTOT) {
                const int q = c % QN, col = (c / QN) % NC, p = c / (QN * NC);
                r[t] = *reinterpret_cast<const h8 *>(P + p * plane + (size_t)min(col, ncol - 1) * ld + k0 + 8 * q);
            }
        }
    };
    auto store_x = [&](int buf, const h8 (&r)[XCH]) {
#pragma unroll
        for (int t = 0; t < XCH; ++t) {
            const int c = tid + NT * t;
            if (XTOT % NT == 0 || c < XTOT) {
                const int q = c % QN, col = (c / QN) % NC, p = c / (QN * NC);
                *reinterpret_cast<h8 *>(&xs[buf][p][col][8 * q]) = r[t];
            }
        }
    };
    // lane's A slice: k = k0 + 8 q + 2 blk + t (see kperm)
    auto load_h = [&](long it, float4 (&h)[4]) {
        const float4 *src = reinterpret_cast<const float4 *>(hcol + min(it, niter - 1) * KT);
#pragma unroll
        for (int q = 0; q < 4; ++q) h[q] = src[4 * q];
    };

    f4 yr[NCB], yi[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        yr[cb] = f4{0.f, 0.f, 0.f, 0.f};
        yi[cb] = f4{0.f, 0.f, 0.f, 0.f};
    }
    auto compute = [&](int buf, const float4 (&hc)[4]) {
        const int kk = 8 * blk;
        // every B fragment of the step is requested before the first MFMA, so
        // the LDS latency overlaps the earlier column blocks
        h8 bf[NCB][4];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int p = 0; p < 4; ++p) bf[cb][p] = *reinterpret_cast<const h8 *>(&xs[buf][p][cb * 16 + li][kk]);
        const Split a = split8(hc, hinv);
        const h8 nih = neg8(a.ih), nil = neg8(a.il);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const h8 xrh = bf[cb][0], xrl = bf[cb][1], xih = bf[cb][2], xil = bf[cb][3];
            // three split terms of each real product, two independent chains interleaved
            yr[cb] = mfma16(a.rh, xrh, yr[cb]);
            yi[cb] = mfma16(a.rh, xih, yi[cb]);
            yr[cb] = mfma16(a.rh, xrl, yr[cb]);
            yi[cb] = mfma16(a.rh, xil, yi[cb]);
            yr[cb] = mfma16(a.rl, xrh, yr[cb]);
            yi[cb] = mfma16(a.rl, xih, yi[cb]);
            yr[cb] = mfma16(a.ih, xih, yr[cb]);
            yi[cb] = mfma16(nih, xrh, yi[cb]);
            yr[cb] = mfma16(a.ih, xil, yr[cb]);
            yi[cb] = mfma16(nih, xrl, yi[cb]);
            yr[cb] = mfma16(a.il, xih, yr[cb]);
            yi[cb] = mfma16(nil, xrh, yi[cb]);
        }
    };

    h8 xr[2][XCH];
    float4 hb[HD][4];
    {
        h8 x0[XCH];
        load_x(0, x0);
        load_x(1, xr[1]);
#pragma unroll
        for (int d = 0; d < HD - 1; ++d) load_h(d, hb[d]);
        store_x(0, x0);
    }
    __syncthreads();
    // step it, j = it mod U: loads X(it + 2) into xr[j % 2] and H(it + HD - 1)
    // into hb[(j + HD - 1) % HD], computes from LDS buffer j % 2 and hb[j % HD],
    // then stages X(it + 1) (xr[(j + 1) % 2]) into LDS buffer (j + 1) % 2.
    // Prefetches past the last slice re-read it and are never used.
    // (the scheduling fences keep the compiler from sinking the prefetch
    // loads down to their first use, which would expose their latency)
    auto step = [&](long it, auto J) {
        constexpr int j = decltype(J)::value;
        load_x(it + 2, xr[j % 2]);
        load_h(it + HD - 1, hb[(j + HD - 1) % HD]);
        __builtin_amdgcn_sched_barrier(0);
        compute(j % 2, hb[j % HD]);
        __builtin_amdgcn_sched_barrier(0);
        store_x((j + 1) % 2, xr[(j + 1) % 2]);
        __syncthreads();
    };
    long it = 0;
    for (; it + U <= niter; it += U) run_steps(step, it, std::make_integer_sequence<int, U>{});
    run_tail(step, it, niter, std::make_integer_sequence<int, U - 1>{});

    // D layout (16x16 f32): register i of lane l -> row 4*(l/16)+i, column l%16
    const long r0 = (long)blockIdx.x * (16 * WPB) + wave * 16 + 4 * blk;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb * 16 + li;
        if (col >= ncol) continue;
        const float s = ysc[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long r = r0 + i;
            if (r < N) {
                Y[2 * (r + (size_t)col * N)] = (double)(yr[cb][i] * s);
                Y[2 * (r + (size_t)col * N) + 1] = (double)(yi[cb][i] * s);
            }
        }
    }
}

template <int WPB>
void launch_f16split(hipStream_t st, int ncb, long N, long ld, int nc, float hinv, const cf *H,
                     const _Float16 *P, const float *ysc, double *y) {
    const dim3 grid((unsigned)((N + 16 * WPB - 1) / (16 * WPB)));
    switch (ncb) {
#define WTB_F16_CASE(K)                                                                                         \
    case K:                                                                                                     \
        hipLaunchKernelGGL((hx_f16split_kernel<K, WPB, kHD>), grid, dim3(64 * WPB), 0, st, N, ld, nc, hinv, H, \
                           P, ysc, y);                                                                          \
        break;
        WTB_F16_CASE(1) WTB_F16_CASE(2) WTB_F16_CASE(3) WTB_F16_CASE(4)
        WTB_F16_CASE(5) WTB_F16_CASE(6) WTB_F16_CASE(7) WTB_F16_CASE(8)
#undef WTB_F16_CASE
    }
}

}  // namespace

// Y = H_offdiag X on the FP32 store with split-FP16 MFMA; the caller adds the
// FP64 diagonal.  hscale (set at assembly): power of two with
// max|H| / hscale in [2^12, 2^13).
namespace {

// Column blocks of up to `blk` columns (<= 128) through the split planes
// (4 blk ld FP16) and column scales (2 blk floats) given by the caller.
void f16split_blocks(Ctx &c, int ncol, const cd *X, cd *Y, int blk, _Float16 *planes, float *scales) {
    Bse &s = c.bse;
    const long N = s.N, ld = s.ld;
    for (int c0 = 0; c0 < ncol; c0 += blk) {
        const int nc = std::min(ncol - c0, blk);
        const double *x = (const double *)X + 2 * (size_t)c0 * N;
        double *y = (double *)Y + 2 * (size_t)c0 * N;
        hipLaunchKernelGGL(col_scales_kernel, dim3(nc), dim3(kCS), 0, c.stream, N, x, s.hscale, scales, scales + nc);
        hipLaunchKernelGGL(split_x_kernel, dim3((unsigned)((ld + 255) / 256), nc), dim3(256), 0, c.stream, N, ld,
                           nc, x, scales, planes);
        WTB_LAUNCH_CHECK();
        const int ncb = (nc + 15) / 16;
        c.kclock.begin(KernelClock::F16SPLIT, c.stream);
        launch_f16split<kWPB>(c.stream, ncb, N, ld, nc, 1.0f / s.hscale, s.h32.p, planes, scales + nc, y);
        c.kclock.end(KernelClock::F16SPLIT, c.stream, 8.0 * (double)ld * N, 8.0 * (double)N * N * nc);
        WTB_LAUNCH_CHECK();
    }
}

// Scratch of the allocation-free entry point wtb_bse_matvec_noalloc: the FP16
// split planes of X (4 ld bytes-pairs per column) in the first 32 MiB, the
// column scales (2 floats per column) after them.  Every call writes the
// planes and scales it reads, so nothing in it needs clearing; calls that
// overlap on one GPU share it and race.
constexpr size_t kScratchPlaneBytes = size_t(32) << 20;
constexpr size_t kScratchBytes = kScratchPlaneBytes + 4096;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

}  // namespace

void matvec_f16split(Ctx &c, int ncol, const cd *X, cd *Y) {
    const int blk = std::min(ncol, 128);
    c.ws_split.alloc((size_t)4 * blk * c.bse.ld);
    c.ws_colscale.alloc(2 * (size_t)blk);
    f16split_blocks(c, ncol, X, Y, 128, c.ws_split.p, c.ws_colscale.p);
}

// Device address of this GPU's copy of the scratch array, looked up once per
// device with that device current (the host-side name is only a stand-in).
void scratch_setup(Ctx &c) {
    if (c.device < 0 || c.device >= kMaxDevices) throw Error{WTB_ERR_ARG, "device ordinal out of range"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[c.device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error{WTB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch_ptr[c.device] = ptr;
    }
    c.scratch = g_scratch_ptr[c.device];
}

// The split-FP16 product through the scratch array: no device allocation.
// Up to 128 columns per launch, fewer when 128 columns' planes exceed 32 MiB
// (ld > 32768); ld > 4 Mi (one column's planes above 32 MiB) is refused.
void matvec_f16split_noalloc(Ctx &c, int ncol, const cd *X, cd *Y) {
    if (!c.scratch) throw Error{WTB_ERR_STATE, "allocation-free scratch not set up on this device"};
    const size_t per_col = 4 * (size_t)c.bse.ld * sizeof(_Float16);
    const size_t fit = kScratchPlaneBytes / per_col;
    if (fit < 1)
        throw Error{WTB_ERR_UNSUPPORTED, "N = " + std::to_string(c.bse.N) +
                                             " exceeds the allocation-free scratch (32 MiB of FP16 planes: ld <= 4194304)"};
    const int blk = (int)std::min<size_t>(128, fit);
    unsigned char *base = static_cast<unsigned char *>(c.scratch);
    f16split_blocks(c, ncol, X, Y, blk, reinterpret_cast<_Float16 *>(base),
                    reinterpret_cast<float *>(base + kScratchPlaneBytes));
}

}  // namespace wtb
