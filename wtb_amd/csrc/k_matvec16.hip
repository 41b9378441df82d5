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
// wavefront streams 16 stored columns along k (contiguous), i.e. the A
// operand of output rows i is conj(H[k + i N]).  X arrives pre-split in FP16
// planes (re_hi, re_lo, im_hi, im_lo), column-major with k contiguous, and is
// staged through LDS once per workgroup and K-step.
#include "wtb_internal.h"

#include <cmath>

namespace wtb {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 64;      // output rows per workgroup (16 per wavefront)
constexpr int kK = 32;         // MFMA K depth
constexpr int kColStride = 40; // halfs per staged column (32 + pad against bank conflicts)

__device__ __forceinline__ f4 mfma16(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// X (N x ncol, complex FP64, column-major) -> four FP16 planes
// P[p][col][k] (k contiguous, kpad rows), scaled per column by xs[col].
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
    const size_t o = k + (size_t)col * kpad;
    P[o] = rh;
    P[plane + o] = rl;
    P[2 * plane + o] = ih;
    P[3 * plane + o] = il;
}

// per-column max |x| (as float bits, atomicMax on non-negative floats)
__global__ void colmax_kernel(long N, int ncol, const double *__restrict__ X, unsigned *__restrict__ mx) {
    int col = blockIdx.y;
    float m = 0.f;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        m = fmaxf(m, fabsf((float)X[2 * (k + (size_t)col * N)]));
        m = fmaxf(m, fabsf((float)X[2 * (k + (size_t)col * N) + 1]));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(mx + col, __float_as_uint(m));
}

template <int NCB>
__global__ void __launch_bounds__(256)
hx_f16split_kernel(long N, long kpad, int ncol, float hscale_inv, const float *__restrict__ H,
                   const _Float16 *__restrict__ P, const float *__restrict__ ysc, double *__restrict__ Y) {
    constexpr int NC = NCB * 16;
    __shared__ __align__(16) _Float16 xs[4][NC][kColStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int blk = lane >> 4, li = lane & 15;
    const long row = (long)blockIdx.x * kRows + wave * 16 + li;   // output row = stored column
    const bool rok = row < N;
    const float2 *hcol = reinterpret_cast<const float2 *>(H) + (size_t)(rok ? row : 0) * N;
    const size_t plane = (size_t)ncol * kpad;

    f4 yr[NCB], yi[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        yr[cb] = f4{0.f, 0.f, 0.f, 0.f};
        yi[cb] = f4{0.f, 0.f, 0.f, 0.f};
    }

    for (long k0 = 0; k0 < N; k0 += kK) {
        // stage the split X slice: 4 planes x NC columns x 32 k (16 B per thread-load)
        for (int idx = tid; idx < 4 * NC * 4; idx += 256) {
            const int q = idx & 3;                 // which 8-k chunk
            const int col = (idx >> 2) % NC;
            const int p = (idx >> 2) / NC;
            h8 v = h8{0, 0, 0, 0, 0, 0, 0, 0};
            if (col < ncol) v = *reinterpret_cast<const h8 *>(P + p * plane + (size_t)col * kpad + k0 + 8 * q);
            *reinterpret_cast<h8 *>(&xs[p][col][8 * q]) = v;
        }
        // A operand: 8 consecutive k of stored column `row`, conjugated, split
        h8 arh, arl, aih, ail, naih, nail;
        {
            float re[8], im[8];
            const long kb = k0 + 8 * blk;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float2 h = make_float2(0.f, 0.f);
                if (rok && kb + e < N) h = hcol[kb + e];
                re[e] = h.x * hscale_inv;
                im[e] = -h.y * hscale_inv;         // row i of H = conj(column i)
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                _Float16 a = (_Float16)re[e], b = (_Float16)im[e];
                arh[e] = a;
                arl[e] = (_Float16)(re[e] - (float)a);
                aih[e] = b;
                ail[e] = (_Float16)(im[e] - (float)b);
                naih[e] = -aih[e];
                nail[e] = -ail[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const int col = cb * 16 + li;
            const h8 xrh = *reinterpret_cast<const h8 *>(&xs[0][col][8 * blk]);
            const h8 xrl = *reinterpret_cast<const h8 *>(&xs[1][col][8 * blk]);
            const h8 xih = *reinterpret_cast<const h8 *>(&xs[2][col][8 * blk]);
            const h8 xil = *reinterpret_cast<const h8 *>(&xs[3][col][8 * blk]);
            // Yr += Ar Xr - Ai Xi
            yr[cb] = mfma16(arh, xrh, yr[cb]);
            yr[cb] = mfma16(arh, xrl, yr[cb]);
            yr[cb] = mfma16(arl, xrh, yr[cb]);
            yr[cb] = mfma16(naih, xih, yr[cb]);
            yr[cb] = mfma16(naih, xil, yr[cb]);
            yr[cb] = mfma16(nail, xih, yr[cb]);
            // Yi += Ar Xi + Ai Xr
            yi[cb] = mfma16(arh, xih, yi[cb]);
            yi[cb] = mfma16(arh, xil, yi[cb]);
            yi[cb] = mfma16(arl, xih, yi[cb]);
            yi[cb] = mfma16(aih, xrh, yi[cb]);
            yi[cb] = mfma16(aih, xrl, yi[cb]);
            yi[cb] = mfma16(ail, xrh, yi[cb]);
        }
        __syncthreads();
    }
    // D layout (16x16 f32): register i of lane l -> row 4*(l/16)+i, column l%16
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb * 16 + li;
        if (col >= ncol) continue;
        const float s = ysc[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long r = (long)blockIdx.x * kRows + wave * 16 + 4 * blk + i;
            if (r < N) {
                Y[2 * (r + (size_t)col * N)] = (double)(yr[cb][i] * s);
                Y[2 * (r + (size_t)col * N) + 1] = (double)(yi[cb][i] * s);
            }
        }
    }
}

__global__ void col_scales_kernel(int ncol, const unsigned *__restrict__ mx, float hscale, float *__restrict__ xs,
                                  float *__restrict__ ysc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    float m = __uint_as_float(mx[c]);
    // scale so that max |x| s ~ 2^13 (FP16 max 65504, leave room for the lo part)
    int e = 0;
    if (m > 0.f) frexpf(m, &e);          // m = f 2^e, 0.5 <= f < 1
    float s = ldexpf(1.0f, 13 - e);
    xs[c] = s;
    ysc[c] = hscale / s;
}

}  // namespace

// Y = H X on the FP32 store with split-FP16 MFMA; H's zero diagonal is
// completed in FP64 by the caller.  hscale: power of two with max|H| / hscale
// around 2^13.
void matvec_f16split(Ctx &c, int ncol, const cd *X, cd *Y) {
    Bse &s = c.bse;
    const long N = s.N;
    const long kpad = (N + kK - 1) / kK * kK;
    for (int c0 = 0; c0 < ncol; c0 += 128) {
        const int nc = std::min(ncol - c0, 128);
        const double *x = (const double *)X + 2 * (size_t)c0 * N;
        double *y = (double *)Y + 2 * (size_t)c0 * N;
        c.ws_split.alloc((size_t)4 * nc * kpad);
        c.ws_colmax.alloc(nc);
        c.ws_colscale.alloc(2 * (size_t)nc);
        WTB_HIP_CHECK(hipMemsetAsync(c.ws_colmax.p, 0, sizeof(unsigned) * nc, c.stream));
        unsigned gx = (unsigned)std::min<long>((N + 255) / 256, 64);
        hipLaunchKernelGGL(colmax_kernel, dim3(gx, nc), dim3(256), 0, c.stream, N, nc, x, c.ws_colmax.p);
        hipLaunchKernelGGL(col_scales_kernel, dim3((nc + 63) / 64), dim3(64), 0, c.stream, nc, c.ws_colmax.p,
                           s.hscale, c.ws_colscale.p, c.ws_colscale.p + nc);
        hipLaunchKernelGGL(split_x_kernel, dim3((unsigned)((kpad + 255) / 256), nc), dim3(256), 0, c.stream, N, kpad,
                           nc, x, c.ws_colscale.p, c.ws_split.p);
        WTB_LAUNCH_CHECK();
        const int ncb = (nc + 15) / 16;
        dim3 grid((unsigned)((N + kRows - 1) / kRows));
        const float hinv = 1.0f / s.hscale;
        switch (ncb) {
#define WTB_F16_CASE(K)                                                                                        \
    case K:                                                                                                    \
        hipLaunchKernelGGL((hx_f16split_kernel<K>), grid, dim3(256), 0, c.stream, N, kpad, nc, hinv,           \
                           (const float *)s.h32.p, c.ws_split.p, c.ws_colscale.p + nc, y);                     \
        break;
            WTB_F16_CASE(1) WTB_F16_CASE(2) WTB_F16_CASE(3) WTB_F16_CASE(4)
            WTB_F16_CASE(5) WTB_F16_CASE(6) WTB_F16_CASE(7) WTB_F16_CASE(8)
#undef WTB_F16_CASE
        }
        WTB_LAUNCH_CHECK();
    }
}

}  // namespace wtb
