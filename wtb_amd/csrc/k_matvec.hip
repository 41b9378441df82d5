// Dense Hermitian BSE matvec Y = H X on MFMA (gfx950), the inner product of
// the block eigensolver (replaces PETSc's MatMult inside EPSSolve,
// src/module_slepc_solver.F90:74).
//
// H is N x N column-major complex (FP64 or FP32 store, leading dimension ld), X and Y are N x ncol
// column-major complex.  A 256-thread workgroup owns a 64-row slab of Y and up
// to 128 columns; each wavefront owns 16 rows.  Per K-step of 16 the X slice
// is staged in LDS as split real/imaginary planes and every wavefront streams
// its 16 x 16 block of H straight from HBM into the MFMA A operand.  Complex
// products use four real MFMAs:
//   Yr += Hr Xr - Hi Xi,   Yi += Hr Xi + Hi Xr.
// MFMA 16x16x4 operand layout (lane l of 64; D differs between f32 and f64):
//   A[l % 16][l / 16], B[l / 16][l % 16]; D f32: [4 (l / 16) + i][l % 16], D f64: [4 i + l / 16][l % 16].
// H is read exactly once per launch, so the kernel is MFMA-bound for
// ncol >~ 20 (FP64) and the bytes/flop ratio is 2/ncol (FP64).
#include "wtb_internal.h"

namespace wtb {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class T>
struct Mfma;
template <>
struct Mfma<double> {
    using V = d4;
    static __device__ __forceinline__ V fma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // f64 16x16x4 output: register i of lane l holds D[4 i + l/16][l % 16]
    static __device__ __forceinline__ int drow(int lane, int i) { return 4 * i + (lane >> 4); }
};
template <>
struct Mfma<float> {
    using V = f4;
    static __device__ __forceinline__ V fma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // f32 16x16x4 output: register i of lane l holds D[4 (l/16) + i][l % 16]
    static __device__ __forceinline__ int drow(int lane, int i) { return 4 * (lane >> 4) + i; }
};

constexpr int kRows = 64;
constexpr int kKT = 16;
constexpr int kMaxCB = 8;   // 128 columns per launch

// TH: storage type of H, TX: storage type of X and Y, TC: MFMA arithmetic type.
template <class TH, class TX, class TC, int NCB>
__global__ void __launch_bounds__(256)
zgemm_hx_kernel(long N, long ld, int ncol, const TH *__restrict__ H, const TX *__restrict__ X, TX *__restrict__ Y) {
    using M = Mfma<TC>;
    using V = typename M::V;
    constexpr int NC = NCB * 16;
    __shared__ TC xr[kKT][NC];
    __shared__ TC xi[kKT][NC];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const long row0 = (long)blockIdx.x * kRows + wave * 16;
    const long arow = row0 + (lane & 15);
    const int akk = lane >> 4;
    const bool rok = arow < N;

    V acc_r[NCB], acc_i[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        acc_r[cb] = V{0, 0, 0, 0};
        acc_i[cb] = V{0, 0, 0, 0};
    }

    for (long k0 = 0; k0 < N; k0 += kKT) {
        // stage X[k0 .. k0+15][0 .. NC) (complex, column-major) into LDS planes
        for (int idx = tid; idx < kKT * NC; idx += 256) {
            int kk = idx % kKT, col = idx / kKT;
            long kr = k0 + kk;
            TC re = 0, im = 0;
            if (kr < N && col < ncol) {
                re = (TC)X[2 * (kr + (long)col * N)];
                im = (TC)X[2 * (kr + (long)col * N) + 1];
            }
            xr[kk][col] = re;
            xi[kk][col] = im;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kKT / 4; ++s) {
            const long kc = k0 + 4 * s + akk;
            TC ar = 0, ai = 0;
            if (rok && kc < N) {
                const TH *hp = H + 2 * (arow + kc * ld);
                ar = (TC)hp[0];
                ai = (TC)hp[1];
            }
            const TC nai = -ai;
            const int bk = 4 * s + akk;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const TC br = xr[bk][cb * 16 + (lane & 15)];
                const TC bi = xi[bk][cb * 16 + (lane & 15)];
                acc_r[cb] = M::fma(ar, br, acc_r[cb]);
                acc_r[cb] = M::fma(nai, bi, acc_r[cb]);
                acc_i[cb] = M::fma(ar, bi, acc_i[cb]);
                acc_i[cb] = M::fma(ai, br, acc_i[cb]);
            }
        }
        __syncthreads();
    }
    // epilogue: lane holds rows 4*(lane/16)+i, column lane%16 of each block
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb * 16 + (lane & 15);
        if (col >= ncol) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long r = row0 + M::drow(lane, i);
            if (r < N) {
                Y[2 * (r + (long)col * N)] = (TX)acc_r[cb][i];
                Y[2 * (r + (long)col * N) + 1] = (TX)acc_i[cb][i];
            }
        }
    }
}

template <class TH, class TX, class TC>
void launch_hx(hipStream_t st, long N, long ld, int ncol, const TH *H, const TX *X, TX *Y) {
    for (int c0 = 0; c0 < ncol; c0 += kMaxCB * 16) {
        int nc = std::min(ncol - c0, kMaxCB * 16);
        int ncb = (nc + 15) / 16;
        dim3 grid((unsigned)((N + kRows - 1) / kRows));
        const TX *Xc = X + 2 * (size_t)c0 * N;
        TX *Yc = Y + 2 * (size_t)c0 * N;
        switch (ncb) {
#define WTB_HX_CASE(K) \
    case K: hipLaunchKernelGGL((zgemm_hx_kernel<TH, TX, TC, K>), grid, dim3(256), 0, st, N, ld, nc, H, Xc, Yc); break;
            WTB_HX_CASE(1) WTB_HX_CASE(2) WTB_HX_CASE(3) WTB_HX_CASE(4)
            WTB_HX_CASE(5) WTB_HX_CASE(6) WTB_HX_CASE(7) WTB_HX_CASE(8)
#undef WTB_HX_CASE
        }
        WTB_LAUNCH_CHECK();
    }
}

// Y += diag(d) X in FP64: the diagonal of the FP32 store is kept exactly.
__global__ void add_diag_kernel(long N, int ncol, const double *__restrict__ d, const double *__restrict__ X,
                                double *__restrict__ Y) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i >= N || c >= ncol) return;
    size_t o = 2 * (i + (size_t)c * N);
    Y[o] += d[i] * X[o];
    Y[o + 1] += d[i] * X[o + 1];
}

}  // namespace

// Y = H X with the arithmetic of `precision`:
//   WTB_FP64 : FP64 MFMA (on the FP64 store, or the FP32 store upcast in
//              registers -- the refinement product of the mixed path);
//   WTB_FP32 / WTB_MIXED : FP32 MFMA with FP32 accumulation on the FP32 store.
// X and Y are FP64 in every case; FP32 operands are rounded while staging.
void matvec(Ctx &c, int precision, int ncol, const cd *X, cd *Y) {
    Bse &s = c.bse;
    if (!s.valid) throw Error{WTB_ERR_STATE, "no BSE matrix built"};
    const long N = s.N;
    if (ncol <= 0) return;
    const double *x = (const double *)X;
    double *y = (double *)Y;
    if (s.precision == WTB_FP64) {
        launch_hx<double, double, double>(c.stream, N, s.ld, ncol, (const double *)s.h64.p, x, y);
    } else {
        if (precision == WTB_FP64)
            launch_hx<float, double, double>(c.stream, N, s.ld, ncol, (const float *)s.h32.p, x, y);
        else if (precision == WTB_MIXED)
            matvec_f16split(c, ncol, X, Y);
        else
            launch_hx<float, double, float>(c.stream, N, s.ld, ncol, (const float *)s.h32.p, x, y);
        for (int c0 = 0; c0 < ncol; c0 += 65535) {
            int nc = std::min(65535, ncol - c0);
            hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)((N + 255) / 256), nc), dim3(256), 0, c.stream, N, nc,
                               s.diag.p, x + 2 * (size_t)c0 * N, y + 2 * (size_t)c0 * N);
            WTB_LAUNCH_CHECK();
        }
    }
}

}  // namespace wtb
