// Dense Hermitian BSE matvec Y = H X on MFMA (gfx950), the inner product of
// the block eigensolver (replaces PETSc's MatMult inside EPSSolve,
// src/module_slepc_solver.F90:74), with FP64 or FP32 arithmetic.
//
// H is N x N column-major complex (FP64 or FP32 store, leading dimension ld,
// rows N..ld-1 zero), X and Y are N x ncol column-major complex FP64.  Since H
// is Hermitian, row i of H is the conjugate of stored column i, so a
// wavefront owning output rows i..i+15 streams stored columns i..i+15 along k
// (contiguous runs per lane).  A 256-thread workgroup owns a 64-row slab of Y
// and up to 128 columns.  Each iteration covers 16 k: the wavefront's next H
// slice is loaded into registers and the next X slice into a second LDS
// buffer (split real / imaginary planes) while the MFMAs of the current slice
// run, with one barrier per iteration.  Complex products use four real MFMAs
//   Yr += Hr Xr - Hi Xi,   Yi += Hr Xi + Hi Xr.
// MFMA 16x16x4 operand layout (lane l of 64; D differs between f32 and f64):
//   A[l % 16][l / 16], B[l / 16][l % 16]; D f32: [4 (l / 16) + i][l % 16], D f64: [4 i + l / 16][l % 16].
// Within an iteration lane group g = l / 16 takes k = 4 g + s for MFMA step
// s = 0..3 on both operands (any k order works if A and B agree), so each lane
// reads 4 consecutive k of its stored column.  FP64 MFMA (78.6 TFLOP/s) makes
// the kernel MFMA-bound for ncol >~ 8: 8 N^2 ncol flops per launch.
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
    static constexpr int kPad = 1;   // LDS row pad (elements): lane groups on distinct banks
};
template <>
struct Mfma<float> {
    using V = f4;
    static __device__ __forceinline__ V fma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // f32 16x16x4 output: register i of lane l holds D[4 (l/16) + i][l % 16]
    static __device__ __forceinline__ int drow(int lane, int i) { return 4 * (lane >> 4) + i; }
    static constexpr int kPad = 4;
};

// 4 consecutive complex values of a stored column (the lane's A slice)
template <class TH>
struct HSlice;
template <>
struct HSlice<double> {
    double2 v[4];
    __device__ __forceinline__ void load(const double *p) {
        const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = q[s];
    }
    __device__ __forceinline__ double re(int s) const { return v[s].x; }
    __device__ __forceinline__ double im(int s) const { return v[s].y; }
};
template <>
struct HSlice<float> {
    float4 v[2];
    __device__ __forceinline__ void load(const float *p) {
        const float4 *q = reinterpret_cast<const float4 *>(p);
        v[0] = q[0];
        v[1] = q[1];
    }
    __device__ __forceinline__ float re(int s) const { return (s & 1) ? v[s >> 1].z : v[s >> 1].x; }
    __device__ __forceinline__ float im(int s) const { return (s & 1) ? v[s >> 1].w : v[s >> 1].y; }
};

constexpr int kRows = 64;
constexpr int kKT = 16;
constexpr int kMaxCB = 8;   // 128 columns per launch

// TH: storage type of H, TC: MFMA arithmetic type; X and Y are complex FP64.
template <class TH, class TC, int NCB>
__global__ void __launch_bounds__(256)
zgemm_hx_kernel(long N, long ld, int ncol, const TH *__restrict__ H, const double *__restrict__ X,
                double *__restrict__ Y) {
    using M = Mfma<TC>;
    using V = typename M::V;
    constexpr int NC = NCB * 16;
    constexpr int LS = NC + M::kPad;
    __shared__ TC xs[2][2][kKT][LS];      // [buffer][re, im][k][column]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const long row0 = (long)blockIdx.x * kRows + wave * 16;
    // rows past N read a valid column; their outputs are never written
    const TH *hcol = H + 2 * ((size_t)min(row0 + li, N - 1) * ld + 4 * g);
    const long niter = (N + kKT - 1) / kKT;

    // X slice: 16 k x NC columns, NCB complex values per thread.  k past N and
    // columns past ncol read clamped, valid entries (H's rows N..ld-1 are
    // zero; extra columns feed outputs that are never written).
    auto load_x = [&](long it, double2 (&r)[NCB]) {
        const long k0 = min(it, niter - 1) * kKT;
#pragma unroll
        for (int t = 0; t < NCB; ++t) {
            const int c = tid + 256 * t;
            const int kk = c % kKT, col = c / kKT;
            const long k = min(k0 + kk, N - 1);
            r[t] = *reinterpret_cast<const double2 *>(X + 2 * (k + (size_t)min(col, ncol - 1) * N));
        }
    };
    auto store_x = [&](int buf, const double2 (&r)[NCB]) {
#pragma unroll
        for (int t = 0; t < NCB; ++t) {
            const int c = tid + 256 * t;
            const int kk = c % kKT, col = c / kKT;
            xs[buf][0][kk][col] = (TC)r[t].x;
            xs[buf][1][kk][col] = (TC)r[t].y;
        }
    };
    auto load_h = [&](long it, HSlice<TH> &h) { h.load(hcol + 2 * min(it, niter - 1) * kKT); };

    V acc_r[NCB], acc_i[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        acc_r[cb] = V{0, 0, 0, 0};
        acc_i[cb] = V{0, 0, 0, 0};
    }
    auto compute = [&](int buf, const HSlice<TH> &h) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const TC ar = (TC)h.re(s), ai = -(TC)h.im(s);     // conj of the stored column
            const TC nai = -ai;
            const int bk = 4 * g + s;
            TC br[NCB], bi[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                br[cb] = xs[buf][0][bk][cb * 16 + li];
                bi[cb] = xs[buf][1][bk][cb * 16 + li];
            }
            // dependent accumulations 2 NCB instructions apart
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                acc_r[cb] = M::fma(ar, br[cb], acc_r[cb]);
                acc_i[cb] = M::fma(ar, bi[cb], acc_i[cb]);
            }
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                acc_r[cb] = M::fma(nai, bi[cb], acc_r[cb]);
                acc_i[cb] = M::fma(ai, br[cb], acc_i[cb]);
            }
        }
    };

    double2 xr[NCB];
    HSlice<TH> ha, hb;
    load_x(0, xr);
    load_h(0, ha);
    store_x(0, xr);
    __syncthreads();
    // (the scheduling fences keep the compiler from sinking the prefetch
    // loads down to their first use, which would expose their latency)
    auto step = [&](long it, int buf, const HSlice<TH> &hc, HSlice<TH> &hn) {
        load_x(it + 1, xr);
        load_h(it + 1, hn);
        __builtin_amdgcn_sched_barrier(0);
        compute(buf, hc);
        __builtin_amdgcn_sched_barrier(0);
        store_x(buf ^ 1, xr);
        __syncthreads();
    };
    long it = 0;
    for (; it + 2 <= niter; it += 2) {
        step(it, 0, ha, hb);
        step(it + 1, 1, hb, ha);
    }
    if (it < niter) step(it, 0, ha, hb);

    // epilogue
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb * 16 + li;
        if (col >= ncol) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long r = row0 + M::drow(lane, i);
            if (r < N) {
                Y[2 * (r + (size_t)col * N)] = (double)acc_r[cb][i];
                Y[2 * (r + (size_t)col * N) + 1] = (double)acc_i[cb][i];
            }
        }
    }
}

template <class TH, class TC>
void launch_hx(hipStream_t st, long N, long ld, int ncol, const TH *H, const double *X, double *Y,
               KernelClock &clk, KernelClock::Kind kind) {
    for (int c0 = 0; c0 < ncol; c0 += kMaxCB * 16) {
        int nc = std::min(ncol - c0, kMaxCB * 16);
        int ncb = (nc + 15) / 16;
        dim3 grid((unsigned)((N + kRows - 1) / kRows));
        const double *Xc = X + 2 * (size_t)c0 * N;
        double *Yc = Y + 2 * (size_t)c0 * N;
        clk.begin(kind, st);
        switch (ncb) {
#define WTB_HX_CASE(K) \
    case K: hipLaunchKernelGGL((zgemm_hx_kernel<TH, TC, K>), grid, dim3(256), 0, st, N, ld, nc, H, Xc, Yc); break;
            WTB_HX_CASE(1) WTB_HX_CASE(2) WTB_HX_CASE(3) WTB_HX_CASE(4)
            WTB_HX_CASE(5) WTB_HX_CASE(6) WTB_HX_CASE(7) WTB_HX_CASE(8)
#undef WTB_HX_CASE
        }
        clk.end(kind, st, 2.0 * sizeof(TH) * (double)ld * N, 8.0 * (double)N * N * nc);
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

// the FP64 diagonal of a FP32 store (whose stored diagonal is zero)
void add_diag(Ctx &c, int ncol, const double *x, double *y) {
    const long N = c.bse.N;
    for (int c0 = 0; c0 < ncol; c0 += 65535) {
        int nc = std::min(65535, ncol - c0);
        hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)((N + 255) / 256), nc), dim3(256), 0, c.stream, N, nc,
                           c.bse.diag.p, x + 2 * (size_t)c0 * N, y + 2 * (size_t)c0 * N);
        WTB_LAUNCH_CHECK();
    }
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
        launch_hx<double, double>(c.stream, N, s.ld, ncol, (const double *)s.h64.p, x, y, c.kclock,
                                  KernelClock::F64_ON_F64);
    } else {
        if (precision == WTB_FP64)
            launch_hx<float, double>(c.stream, N, s.ld, ncol, (const float *)s.h32.p, x, y, c.kclock,
                                     KernelClock::F64_ON_F32);
        else if (precision == WTB_MIXED)
            matvec_f16split(c, ncol, X, Y);
        else if (precision == WTB_INT8X || precision == kInt8Short)
            matvec_int8(c, ncol, X, Y, precision == kInt8Short ? 4 : 5);
        else
            launch_hx<float, float>(c.stream, N, s.ld, ncol, (const float *)s.h32.p, x, y, c.kclock,
                                    KernelClock::F32_ON_F32);
        add_diag(c, ncol, x, y);
    }
}

// matvec's products without device allocation: the FP64 and FP32 kernels use
// none, the split-FP16 product takes its planes from the scratch array
void matvec_noalloc(Ctx &c, int precision, int ncol, const cd *X, cd *Y) {
    Bse &s = c.bse;
    if (!s.valid) throw Error{WTB_ERR_STATE, "no BSE matrix built"};
    if (ncol <= 0) return;
    const double *x = (const double *)X;
    double *y = (double *)Y;
    if (precision == WTB_INT8X) throw Error{WTB_ERR_UNSUPPORTED, "the int8-digit product needs a workspace"};
    if (s.precision == WTB_FP64) {
        launch_hx<double, double>(c.stream, s.N, s.ld, ncol, (const double *)s.h64.p, x, y, c.kclock,
                                  KernelClock::F64_ON_F64);
        return;
    }
    if (precision == WTB_FP64)
        launch_hx<float, double>(c.stream, s.N, s.ld, ncol, (const float *)s.h32.p, x, y, c.kclock,
                                 KernelClock::F64_ON_F32);
    else if (precision == WTB_MIXED)
        matvec_f16split_noalloc(c, ncol, X, Y);
    else
        launch_hx<float, float>(c.stream, s.N, s.ld, ncol, (const float *)s.h32.p, x, y, c.kclock,
                                KernelClock::F32_ON_F32);
    add_diag(c, ncol, x, y);
}

}  // namespace wtb
