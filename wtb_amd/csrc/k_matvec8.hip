// BSE matvec Y = H X on the FP32 store with exact int8 products (the Ozaki
// splitting: K. Ozaki, T. Ogita, S. Oishi, S. M. Rump, "Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59 (2012); with
// integer slices as in H. Ootomo, K. Ozaki, R. Yokota, "DGEMM on integer
// matrix multiplication unit", IJHPCA 2024), on gfx950's
// v_mfma_i32_16x16x64_i8.  This is the final residual-refinement product of
// the mixed-precision solver.
//
// Every operand becomes a 31-bit fixed-point integer split into four signed
// byte digits, V = sum_s d_s 2^(8s):
//   H:  V = rint(H * 2^17 / hscale)   (|H| / hscale < 2^13: one power of two for
//                                      the whole store, set at assembly)
//   X:  V = rint(X * xs_c)            (xs_c = 2^(30 - e_c), |X(:,c)| < 2^e_c)
// so |V| <= 2^30 and the top digit is within [-64, 64].  The byte digits
// come from (V + 0x808080) ^ 0x808080: adding 128 to each lower byte makes
// the carries of the signed-digit recoding, and the xor turns each biased byte
// u into the signed byte u - 128.  A digit product is exact in the int32
// accumulators.  Products of digit pairs (s_H, s_X) with s_H + s_X >= 2 are
// kept (13 of the 16), grouped by significance level L = 6 - s_H - s_X (five
// int32 accumulators per output component), and recombined in FP64:
//   Y = hscale 2^31 / xs_c  sum_L acc_L 2^(-8L).
// The error per term is then the rounding of V, 2^-31 max|H| |X_j| +
// 2^-31 |H_ij| max|X(:,c)|, plus the dropped pairs, below 2^-36 max|H|
// max|X(:,c)|: relative to a column of H X about 2^-31 (max|H| / rms|H|),
// which is the size of the FP32 store's own rounding (2^-24 of each element).
// The product is exact with respect to this fixed-point H and X; it is not an
// FP64 product, and it is compared with the FP64 MFMA product on the same
// store in the tests.
// A short form keeps four levels (the 10 pairs with s_H + s_X >= 3): the
// dropped pairs then reach 2^-28 max|H| max|X(:,c)| per term, about 2^3
// times the fixed-point rounding, for callers whose tolerance is far above
// it; it issues 10 instead of 13 digit-pair MFMA groups.  (Three levels, 6
// pairs, leave 2^-20 per term: a residual error near 5e-5 of |lambda| at
// N = 32400, above the solver's loosest tolerance.)
//
// Overflow: per k, a level accumulates two components times at most four
// digit pairs of |d_H d_X| <= 128 * 128 (64 * 128 with a top digit): 98304 in
// levels 3 and 4.  Each launch covers at most 21824 values of k (gridDim.y
// chunks of k, summed afterwards in a fixed order), so every int32 sum stays
// below 2^31 whatever the data.
//
// Layout and conjugation as in the split-FP16 kernel: row i of H is the
// conjugate of stored column i, so output row i uses A = (c_r, -c_i) of the
// stored column:  Y_r = c_r X_r + c_i X_i,  Y_i = c_r X_i + (-c_i) X_r.  The
// digits of c_r, c_i and -c_i are made in registers from the streamed store;
// X arrives as eight int8 planes (X_r and X_i, four digits each), k permuted
// within each 64-block to match the A operand's k set.
#include "wtb_internal.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>

namespace wtb {
namespace {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int kPl8 = 8;           // X planes: 2 components x 4 digits
constexpr long kChunkK8 = 21824;  // k per launch chunk (multiple of 64, 98304 * 21824 < 2^31)

__device__ __forceinline__ i4 mfma8(i4 a, i4 b, i4 c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// signed byte digits of an integer |v| <= 2^30, least significant byte first
__device__ __forceinline__ unsigned digit_word(int v) { return ((unsigned)v + 0x808080u) ^ 0x808080u; }

// 4 digit words (elements 0..3) -> 4 words, word s holding digit s of the
// four elements (byte e = element e).  v_perm_b32 selects bytes of
// {src0, src1}: selector values 0-3 pick src1's bytes, 4-7 src0's.
__device__ __forceinline__ void transpose4(unsigned w0, unsigned w1, unsigned w2, unsigned w3, int &d0, int &d1,
                                           int &d2, int &d3) {
    const unsigned t0 = __builtin_amdgcn_perm(w1, w0, 0x05010400u);   // w0.b0 w1.b0 w0.b1 w1.b1
    const unsigned t1 = __builtin_amdgcn_perm(w1, w0, 0x07030602u);   // w0.b2 w1.b2 w0.b3 w1.b3
    const unsigned t2 = __builtin_amdgcn_perm(w3, w2, 0x05010400u);
    const unsigned t3 = __builtin_amdgcn_perm(w3, w2, 0x07030602u);
    d0 = (int)__builtin_amdgcn_perm(t2, t0, 0x05040100u);
    d1 = (int)__builtin_amdgcn_perm(t2, t0, 0x07060302u);
    d2 = (int)__builtin_amdgcn_perm(t3, t1, 0x05040100u);
    d3 = (int)__builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// Per-column scales: max |x| over the column (one workgroup per column)
// gives e with max < 2^e; xs = 2^(30 - e), ysc = hscale 2^31 / xs.
constexpr int kCS8 = 1024;
__global__ void __launch_bounds__(kCS8)
col_scales8_kernel(long N, const double *__restrict__ X, double hscale, double *__restrict__ xs,
                   double *__restrict__ ysc) {
    __shared__ double red[kCS8 / 64];
    const int col = blockIdx.x, tid = threadIdx.x;
    const double2 *x = reinterpret_cast<const double2 *>(X) + (size_t)col * N;
    double m = 0.0;
    for (long k0 = 0; k0 < N; k0 += 8 * kCS8) {
        double2 v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const long k = k0 + t * kCS8 + tid;
            v[t] = k < N ? x[k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) m = fmax(m, fmax(fabs(v[t].x), fabs(v[t].y)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < kCS8 / 64; ++w) m = fmax(m, red[w]);
        int e = 0;
        if (m > 0.0) frexp(m, &e);               // m = f 2^e, 0.5 <= f < 1
        e = max(-900, min(900, e));
        xs[col] = ldexp(1.0, 30 - e);
        ysc[col] = ldexp(hscale, 1 + e);
    }
}

// X (N x ncol complex FP64) -> int8 planes P[comp * 4 + digit][col][pos]
// (ld rows, zero past N).  Inside each 64-block, k = 8q + 2g + t (q = 0..7,
// g = 0..3, t = 0..1) sits at position 16g + 2q + t: lane group g of the
// MFMA supplies the 16 values of k with that g (the four lanes of one H column
// then read 64 contiguous bytes per load), and its B operand is the 16
// contiguous bytes at 16g.  One thread per 4 consecutive positions.
__global__ void split_x8_kernel(long N, long ld, int ncol, const double *__restrict__ X,
                                const double *__restrict__ xs, int8_t *__restrict__ P) {
    const long pos4 = blockIdx.x * (long)blockDim.x + threadIdx.x;   // position / 4
    const int col = blockIdx.y;
    if (4 * pos4 >= ld) return;
    const double s = xs[col];
    unsigned wr[4], wi[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long pos = 4 * pos4 + j, b = pos & ~63L, r = pos & 63;
        const long g = r >> 4, e = r & 15, q = e >> 1, t = e & 1;
        const long k = b + 8 * q + 2 * g + t;
        double re = 0.0, im = 0.0;
        if (k < N) {
            re = X[2 * (k + (size_t)col * N)] * s;
            im = X[2 * (k + (size_t)col * N) + 1] * s;
        }
        wr[j] = digit_word(__double2int_rn(re));
        wi[j] = digit_word(__double2int_rn(im));
    }
    int d[8];
    transpose4(wr[0], wr[1], wr[2], wr[3], d[0], d[1], d[2], d[3]);
    transpose4(wi[0], wi[1], wi[2], wi[3], d[4], d[5], d[6], d[7]);
    const size_t plane = (size_t)ncol * ld;
#pragma unroll
    for (int p = 0; p < kPl8; ++p)
        *reinterpret_cast<int *>(P + p * plane + (size_t)col * ld + 4 * pos4) = d[p];
}

// RT row tiles of 16 per workgroup; WPR wavefronts share a row tile, each
// owning CPW blocks of 16 columns.  gridDim.y: chunks of kc_len values of k.
// One wavefront per SIMD (the 96 KB LDS tile allows one workgroup per CU):
// the MFMA accumulators take the AGPRs, the rest of the 512 registers hold
// the digit operands of two steps and the H prefetch.
// Yp: [chunk][col][N] complex FP64 partial products.
template <int CPW, int WPR, int RT, int NL>
__global__ void __launch_bounds__(64 * WPR * RT) __attribute__((amdgpu_waves_per_eu(1, 1)))
hx_int8_kernel(long N, long ld, int ncol, long kc_len, float hmul, const float2 *__restrict__ H,
               const int8_t *__restrict__ P, const double *__restrict__ ysc, double *__restrict__ Yp) {
    constexpr int NCB = CPW * WPR;
    constexpr int NC = 16 * NCB;
    constexpr int NT = 64 * WPR * RT;
    // bytes per staged (plane, column): the 96-byte row stride keeps the
    // B-fragment reads (16 lanes x 16 bytes at rows li, chunk blk) on
    // distinct LDS banks
    constexpr int XS = 64 + 32;
    constexpr int XCH = kPl8 * NC * 4 / NT;         // 16-byte X chunks per thread per step
    static_assert(kPl8 * NC * 4 % NT == 0, "X slice must split evenly over the workgroup");
    __shared__ __align__(16) int8_t xs[2][kPl8][NC][XS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int blk = lane >> 4, li = lane & 15;
    const int rt = wave / WPR, cp = wave % WPR;
    // rows past N read a clamped, valid row; X columns past ncol read the last
    // one: their products land in outputs that are never written
    const long row = min((long)blockIdx.x * (16 * RT) + rt * 16 + li, N - 1);
    const float2 *hcol = H + (size_t)row * ld + 2 * blk;
    const size_t plane = (size_t)ncol * ld;
    const long kb = (long)blockIdx.y * kc_len;
    const long niter = (min(ld, kb + kc_len) - kb) / 64;

    auto load_x = [&](long it, i4 (&r)[XCH]) {
        const long k0 = kb + min(it, niter - 1) * 64;
#pragma unroll
        for (int t = 0; t < XCH; ++t) {
            const int c = tid + NT * t;
            const int q = c & 3, col = (c >> 2) % NC, p = c / (4 * NC);
            r[t] = *reinterpret_cast<const i4 *>(P + p * plane + (size_t)min(col, ncol - 1) * ld + k0 + 16 * q);
        }
    };
    auto store_x = [&](int buf, const i4 (&r)[XCH]) {
#pragma unroll
        for (int t = 0; t < XCH; ++t) {
            const int c = tid + NT * t;
            const int q = c & 3, col = (c >> 2) % NC, p = c / (4 * NC);
            *reinterpret_cast<i4 *>(&xs[buf][p][col][16 * q]) = r[t];
        }
    };
    // lane's 16 complex values of a step: k = k0 + 8q + 2 blk + t, element 2q + t
    auto load_h = [&](long it, float4 (&h)[8]) {
        const float4 *src = reinterpret_cast<const float4 *>(hcol + kb + min(it, niter - 1) * 64);
#pragma unroll
        for (int q = 0; q < 8; ++q) h[q] = src[4 * q];
    };
    // digit operands of c_r, c_i and -c_i: a[c][s] holds digit s of the 16 elements
    auto make_digits = [&](const float4 (&h)[8], i4 (&a)[3][4]) {
        unsigned w[3][16];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int vr0 = __float2int_rn(h[q].x * hmul), vi0 = __float2int_rn(h[q].y * hmul);
            const int vr1 = __float2int_rn(h[q].z * hmul), vi1 = __float2int_rn(h[q].w * hmul);
            w[0][2 * q] = digit_word(vr0);
            w[0][2 * q + 1] = digit_word(vr1);
            w[1][2 * q] = digit_word(vi0);
            w[1][2 * q + 1] = digit_word(vi1);
            w[2][2 * q] = digit_word(-vi0);
            w[2][2 * q + 1] = digit_word(-vi1);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int d0, d1, d2, d3;
                transpose4(w[c][4 * j], w[c][4 * j + 1], w[c][4 * j + 2], w[c][4 * j + 3], d0, d1, d2, d3);
                a[c][0][j] = d0;
                a[c][1][j] = d1;
                a[c][2][j] = d2;
                a[c][3][j] = d3;
            }
    };

    i4 acc[NL][2][CPW];
#pragma unroll
    for (int L = 0; L < NL; ++L)
#pragma unroll
        for (int cb = 0; cb < CPW; ++cb) {
            acc[L][0][cb] = i4{0, 0, 0, 0};
            acc[L][1][cb] = i4{0, 0, 0, 0};
        }
    // X digit major: the B fragments of one X digit for every column block,
    // then every H digit paired with it; the column blocks innermost keep
    // 4 CPW independent accumulator chains between reuses
    auto compute = [&](int buf, const i4 (&a)[3][4]) {
#pragma unroll
        for (int sx = 0; sx < 4; ++sx) {
            i4 b[CPW][2];
#pragma unroll
            for (int cb = 0; cb < CPW; ++cb) {
                const int col = (cp * CPW + cb) * 16 + li;
#pragma unroll
                for (int c = 0; c < 2; ++c) b[cb][c] = *reinterpret_cast<const i4 *>(&xs[buf][4 * c + sx][col][16 * blk]);
            }
#pragma unroll
            for (int sh = 3; sh >= 0; --sh) {
                const int L = 6 - sh - sx;
                if (L >= NL) continue;
#pragma unroll
                for (int cb = 0; cb < CPW; ++cb) {
                    acc[L][0][cb] = mfma8(a[0][sh], b[cb][0], acc[L][0][cb]);   // c_r X_r
                    acc[L][1][cb] = mfma8(a[0][sh], b[cb][1], acc[L][1][cb]);   // c_r X_i
                }
#pragma unroll
                for (int cb = 0; cb < CPW; ++cb) {
                    acc[L][0][cb] = mfma8(a[1][sh], b[cb][1], acc[L][0][cb]);   // c_i X_i
                    acc[L][1][cb] = mfma8(a[2][sh], b[cb][0], acc[L][1][cb]);   // -c_i X_r
                }
            }
        }
    };

    // Step it runs its MFMAs on digits made in step it - 1 (from H fetched in
    // step it - 2): the next step's digits and the H two steps ahead are in
    // flight while the MFMAs run.
    i4 xr[XCH];
    float4 h[8];
    i4 a[3][4], an[3][4];
    load_x(0, xr);
    load_h(0, h);
    store_x(0, xr);
    make_digits(h, a);
    load_h(1, h);
    __syncthreads();
    for (long it = 0; it < niter; ++it) {
        const int buf = (int)(it & 1);
        load_x(it + 1, xr);
        make_digits(h, an);
        load_h(it + 2, h);
        compute(buf, a);
        store_x(buf ^ 1, xr);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2) a[c][s2] = an[c][s2];
        __syncthreads();
    }

    // D layout (16x16 i32): register i of lane l -> row 4*(l/16)+i, column l%16
    const long r0 = (long)blockIdx.x * (16 * RT) + rt * 16 + 4 * blk;
    double *yp = Yp + 2 * (size_t)blockIdx.y * ncol * N;
#pragma unroll
    for (int cb = 0; cb < CPW; ++cb) {
        const int col = (cp * CPW + cb) * 16 + li;
        if (col >= ncol) continue;
        const double s = ysc[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long r = r0 + i;
            if (r >= N) continue;
            double v[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double t = (double)acc[0][c][cb][i];
#pragma unroll
                for (int L = 1; L < NL; ++L) t += ldexp((double)acc[L][c][cb][i], -8 * L);
                v[c] = s * t;
            }
            yp[2 * (r + (size_t)col * N)] = v[0];
            yp[2 * (r + (size_t)col * N) + 1] = v[1];
        }
    }
}

// Y = sum over chunks of Yp, in chunk order
__global__ void sum_chunks8(size_t n, int nch, const double *__restrict__ Yp, double *__restrict__ Y) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = Yp[i];
    for (int c = 1; c < nch; ++c) s += Yp[i + c * n];
    Y[i] = s;
}

constexpr int kRT8 = 4, kWPR8 = 1, kCPW8 = 5;   // 4 wavefronts of 16 rows x up to 80 columns per workgroup
constexpr int kNL8 = 5;                          // significance levels kept (13 digit pairs)
constexpr int kNL8s = 4;                         // short form (10 digit pairs)

}  // namespace

// Y = H_offdiag X on the FP32 store with exact int8 digit products; the
// caller adds the FP64 diagonal.  levels: 5 (full) or 4 (short form);
// WTB_INT8_LEVELS=4 selects the short form for every call (tests).
void matvec_int8(Ctx &c, int ncol, const cd *X, cd *Y, int levels) {
    if (const char *e = std::getenv("WTB_INT8_LEVELS")) {
        if (std::atoi(e) == kNL8s) levels = kNL8s;
    }
    const bool shrt = levels == kNL8s;
    Bse &s = c.bse;
    const long N = s.N, ld = s.ld;
    constexpr int NCMAX = 16 * kCPW8 * kWPR8;
    // WTB_INT8_CHUNK (a multiple of 64 up to the overflow-safe length) shortens
    // the k chunks, so that small matrices exercise the chunked path in tests
    long kc = kChunkK8;
    if (const char *e = std::getenv("WTB_INT8_CHUNK")) {
        const long v = std::atol(e);
        if (v >= 64 && v % 64 == 0 && v < kChunkK8) kc = v;
    }
    const int nch = (int)((ld + kc - 1) / kc);
    for (int c0 = 0; c0 < ncol; c0 += NCMAX) {
        const int nc = std::min(ncol - c0, NCMAX);
        const double *x = (const double *)X + 2 * (size_t)c0 * N;
        double *y = (double *)Y + 2 * (size_t)c0 * N;
        c.ws_split.alloc((size_t)4 * nc * ld);     // 8 int8 planes = 4 FP16 planes of bytes
        c.ws_scale8.alloc(2 * (size_t)nc);
        hipLaunchKernelGGL(col_scales8_kernel, dim3(nc), dim3(kCS8), 0, c.stream, N, x, (double)s.hscale,
                           c.ws_scale8.p, c.ws_scale8.p + nc);
        int8_t *P = reinterpret_cast<int8_t *>(c.ws_split.p);
        hipLaunchKernelGGL(split_x8_kernel, dim3((unsigned)((ld / 4 + 255) / 256), nc), dim3(256), 0, c.stream, N, ld,
                           nc, x, c.ws_scale8.p, P);
        WTB_LAUNCH_CHECK();
        double *yp = y;
        if (nch > 1) {
            c.ws_ypart.alloc((size_t)nch * nc * N);
            yp = (double *)c.ws_ypart.p;
        }
        const dim3 grid((unsigned)((N + 16 * kRT8 - 1) / (16 * kRT8)), (unsigned)nch);
        c.kclock.begin(KernelClock::INT8X, c.stream);
        // 16-column blocks per wavefront: four up to 64 columns, five beyond
        const float hmul = (float)std::ldexp(1.0 / s.hscale, 17);
#define WTB_INT8_LAUNCH(CPW, NL)                                                                              \
    hipLaunchKernelGGL((hx_int8_kernel<CPW, kWPR8, kRT8, NL>), grid, dim3(64 * kWPR8 * kRT8), 0, c.stream, N, ld, \
                       nc, kc, hmul, (const float2 *)s.h32.p, P, c.ws_scale8.p + nc, yp)
        if (nc <= 64) {
            if (shrt) WTB_INT8_LAUNCH(4, kNL8s);
            else WTB_INT8_LAUNCH(4, kNL8);
        } else {
            if (shrt) WTB_INT8_LAUNCH(kCPW8, kNL8s);
            else WTB_INT8_LAUNCH(kCPW8, kNL8);
        }
#undef WTB_INT8_LAUNCH
        c.kclock.end(KernelClock::INT8X, c.stream, 8.0 * (double)ld * N, 8.0 * (double)N * N * nc);
        WTB_LAUNCH_CHECK();
        if (nch > 1) {
            const size_t n = 2 * (size_t)nc * N;
            hipLaunchKernelGGL(sum_chunks8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n, nch, yp, y);
            WTB_LAUNCH_CHECK();
        }
    }
}

}  // namespace wtb
