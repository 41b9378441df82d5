// Dense Hermitian eigensolver for the Rayleigh-Ritz step (FP64, n <= 1024).
//
//  1. Householder reduction to real symmetric tridiagonal form (lower, ZHETD2
//     semantics with ZLARFG reflectors) in one workgroup; the reflectors are
//     written to the lower triangle of A.  Up to n = 224 the lower triangle
//     is held in registers and LDS for the whole reduction; above, the
//     reduction streams the matrix from global memory in the blocked ZLATRD
//     scheme (trailing updates deferred over a panel of steps).
//  2. Eigenvalues of the tridiagonal matrix by Sturm-count multisection over
//     several workgroups; eigenvectors by inverse iteration with partial
//     pivoting, one thread per vector with its factors in LDS, then modified
//     Gram-Schmidt inside clusters of close eigenvalues (the scheme of LAPACK
//     DSTEBZ/DSTEIN).
//  3. Back-transformation Z = H(0) H(1) ... H(n-2) Y, independent per
//     column block, spread over many workgroups.
// Only the lowest `nvec` eigenpairs are formed.
#include "wtb_internal.h"

#include <cmath>

namespace wtb {
namespace {

constexpr int kT = 1024;
// largest n whose Hermitian matrix (n^2 complex) is staged in LDS by the
// Cholesky factorisation (with its work arrays it stays under the 160 KB of
// a CU)
constexpr int kCholLds = 96;
// Eigenvalues closer than this (relative to |T|) are treated as one cluster
// and their inverse-iteration vectors are re-orthogonalised.  Outside a
// cluster inverse iteration leaves an orthogonality error of about
// eps |T| / gap <= 2e-9.
constexpr double kClusterRel = 1e-7;

template <int NT = kT>
__device__ __forceinline__ double block_reduce(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    return s;
}

// ---------------------------------------------------------------------------
// 1. tridiagonalisation
// A (n x n, column-major, lda): full Hermitian matrix (the lower triangle is
// read), destroyed.  On exit A[r + i lda] (r >= i+2) holds v_i(r) of
// reflector i (v_i(i+1) = 1), tau[i] its scalar, d/e the real tridiagonal
// matrix T = Q^H A Q.
// Block sum with one barrier: alternate calls use the two halves of red, and
// every thread has read a half before any thread passes the next call's
// barrier, so the half is free again two calls later.
template <int NT = kT>
__device__ __forceinline__ double block_reduce_db(double v, double *red, int &half) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double *r = red + half * (NT / 64);
    if (lane == 0) r[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += r[i];
    half ^= 1;
    return s;
}

// Blocked form of the same reduction for n > 32 kRegNA (the LATRD scheme).  The
// trailing matrix in global memory is brought up to date only once per panel
// of nb steps; inside a panel the current matrix is A - V W^H - W V^H, with
// the panel's reflectors V and their companions W held in LDS.  Each step then
// reads the stale trailing matrix once (for y = A v, corrected by the panel
// terms) instead of reading it twice and writing it once.
// LDS: V, W (nb x n each, column j of the panel at [j n]), y (n), 1024
// partial sums.
constexpr int kMaxNb = 16;
__device__ __forceinline__ cd wave_sum(cd v) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    return v;
}

__global__ void __launch_bounds__(kT)
tridiag_blocked_kernel(int n, int nb, cd *__restrict__ A, int lda, cd *__restrict__ tau, double *__restrict__ d,
                       double *__restrict__ e) {
    extern __shared__ __align__(16) unsigned char smem[];
    cd *V = reinterpret_cast<cd *>(smem);
    cd *W = V + (size_t)nb * n;
    cd *y = W + (size_t)nb * n;
    cd *buf = y + n;
    __shared__ double red[2 * (kT / 64)];
    __shared__ cd s_c[2 * kMaxNb];      // W_j^H v, then V_j^H v
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int kW = kT / 64;
    int p0 = 0, half = 0;
    for (int i = 0; i < n - 1; ++i) {
        const int ip = i - p0, t0 = i + 1;
        // column i of the current matrix (rows i..n-1) into y; |.|^2 below the subdiagonal
        double xn2 = 0.0;
        for (int r = i + tid; r < n; r += kT) {
            cd a = A[r + (size_t)i * lda];
#pragma unroll 4
            for (int j = 0; j < ip; ++j) {
                const cd vr = V[j * n + r], wr = W[j * n + r], vi = V[j * n + i], wi = W[j * n + i];
                // a -= v_r conj(w_i) + w_r conj(v_i)
                a.x -= (vr.x * wi.x + vr.y * wi.y) + (wr.x * vi.x + wr.y * vi.y);
                a.y -= (vr.y * wi.x - vr.x * wi.y) + (wr.y * vi.x - wr.x * vi.y);
            }
            y[r] = a;
            if (r >= i + 2) xn2 += a.x * a.x + a.y * a.y;
        }
        xn2 = block_reduce_db(xn2, red, half);
        const cd alpha = y[t0];
        if (tid == 0) d[i] = y[i].x;
        cd *vp = V + (size_t)ip * n, *wp = W + (size_t)ip * n;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            // H(i) = I: v = e_(i+1), w = 0; the stale column below is cleared
            for (int r = t0 + tid; r < n; r += kT) {
                vp[r] = make_double2(r == t0 ? 1.0 : 0.0, 0.0);
                wp[r] = make_double2(0.0, 0.0);
                if (r > t0) A[r + (size_t)i * lda] = make_double2(0.0, 0.0);
            }
            if (tid == 0) {
                e[i] = alpha.x;
                tau[i] = make_double2(0.0, 0.0);
            }
            __syncthreads();
        } else {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            const double beta = alpha.x >= 0.0 ? -nrm : nrm;
            const cd t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            const double dd = (alpha.x - beta) * (alpha.x - beta) + alpha.y * alpha.y;
            const cd scal = make_double2((alpha.x - beta) / dd, -alpha.y / dd);   // 1/(alpha - beta)
            for (int r = t0 + tid; r < n; r += kT) {
                cd vr = make_double2(1.0, 0.0);
                if (r > t0) {
                    const cd a = y[r];
                    vr = make_double2(a.x * scal.x - a.y * scal.y, a.x * scal.y + a.y * scal.x);
                    A[r + (size_t)i * lda] = vr;      // keep the reflector
                }
                vp[r] = vr;
            }
            if (tid == 0) {
                e[i] = beta;
                tau[i] = t;
            }
            __syncthreads();
            // panel corrections W_j^H v and V_j^H v, one wavefront per product
            for (int k = wv; k < 2 * ip; k += kW) {
                const cd *B = k < ip ? W + (size_t)k * n : V + (size_t)(k - ip) * n;
                cd acc = make_double2(0.0, 0.0);
                for (int r = t0 + lane; r < n; r += 64) {
                    const cd b = B[r], vr = vp[r];
                    acc.x += b.x * vr.x + b.y * vr.y;
                    acc.y += b.x * vr.y - b.y * vr.x;
                }
                acc = wave_sum(acc);
                if (lane == 0) s_c[k < ip ? k : kMaxNb + (k - ip)] = acc;
            }
            // partial products of the stale trailing block with v: wavefront wv
            // takes row chunk q (64 rows from rb) and every G-th column
            const int rb = t0 & ~63, Qn = (n - rb + 63) / 64, G = kW / Qn;
            const int q = wv % Qn, g = wv / Qn;
            if (g < G) {
                const int r = rb + 64 * q + lane;
                const bool ok = r >= t0 && r < n;
                const cd *ar = A + (ok ? r : t0);
                cd acc = make_double2(0.0, 0.0);
                int c = t0 + g;
                for (; c + 7 * G < n; c += 8 * G) {
                    cd a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = ar[(size_t)(c + u * G) * lda];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const cd vc = vp[c + u * G];
                        acc.x += a[u].x * vc.x - a[u].y * vc.y;
                        acc.y += a[u].x * vc.y + a[u].y * vc.x;
                    }
                }
                for (; c < n; c += G) {
                    const cd a = ar[(size_t)c * lda], vc = vp[c];
                    acc.x += a.x * vc.x - a.y * vc.y;
                    acc.y += a.x * vc.y + a.y * vc.x;
                }
                buf[g * Qn * 64 + 64 * q + lane] = acc;
            }
            __syncthreads();
            // y = sum of partials - V (W^H v) - W (V^H v); mu = Re v^H y
            double mu = 0.0;
            for (int r = t0 + tid; r < n; r += kT) {
                cd s = make_double2(0.0, 0.0);
                for (int gg = 0; gg < G; ++gg) {
                    const cd b = buf[gg * Qn * 64 + (r - rb)];
                    s.x += b.x;
                    s.y += b.y;
                }
#pragma unroll 4
                for (int j = 0; j < ip; ++j) {
                    const cd vr = V[j * n + r], wr = W[j * n + r], cw = s_c[j], cv = s_c[kMaxNb + j];
                    s.x -= (vr.x * cw.x - vr.y * cw.y) + (wr.x * cv.x - wr.y * cv.y);
                    s.y -= (vr.x * cw.y + vr.y * cw.x) + (wr.x * cv.y + wr.y * cv.x);
                }
                y[r] = s;
                const cd vr = vp[r];
                mu += vr.x * s.x + vr.y * s.y;
            }
            mu = block_reduce_db(mu, red, half);
            // w = tau y - |tau|^2 mu / 2 v (same thread per row as above)
            const double t2 = t.x * t.x + t.y * t.y;
            for (int r = t0 + tid; r < n; r += kT) {
                const cd yr = y[r], vr = vp[r];
                wp[r] = make_double2(t.x * yr.x - t.y * yr.y - 0.5 * t2 * mu * vr.x,
                                     t.x * yr.y + t.y * yr.x - 0.5 * t2 * mu * vr.y);
            }
            __syncthreads();
        }
        // end of a panel: trailing block (rows, columns >= i+1) -= V W^H + W V^H;
        // each thread takes one row and four columns
        if (ip + 1 == nb || i == n - 2) {
            const int np = ip + 1, s = n - t0, c4 = (s + 3) / 4;
            for (int idx = tid; idx < s * c4; idx += kT) {
                const int r = t0 + idx % s, cb = t0 + 4 * (idx / s);
                cd a[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (cb + u < n) a[u] = A[r + (size_t)(cb + u) * lda];
                for (int j = 0; j < np; ++j) {
                    const cd vr = V[j * n + r], wr = W[j * n + r];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int c = min(cb + u, n - 1);
                        const cd vc = V[j * n + c], wc = W[j * n + c];
                        // a -= v_r conj(w_c) + w_r conj(v_c)
                        a[u].x -= (vr.x * wc.x + vr.y * wc.y) + (wr.x * vc.x + wr.y * vc.y);
                        a[u].y -= (vr.y * wc.x - vr.x * wc.y) + (wr.y * vc.x - wr.x * vc.y);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (cb + u < n) A[r + (size_t)(cb + u) * lda] = a[u];
            }
            __syncthreads();
            p0 = i + 1;
        }
    }
    if (tid == 0) {
        d[n - 1] = A[(n - 1) + (size_t)(n - 1) * lda].x;
        tau[n > 1 ? n - 1 : 0] = make_double2(0.0, 0.0);
    }
}

// Register-resident form of the same reduction for n <= 32 kRegNA: the lower
// triangle lives in the registers of 512 threads and in LDS, so a step reads
// no matrix data from memory and the pending rank-2 update shares one pass
// with the next product.
// Thread (p, q), p the row class mod 32 and q the column class mod 16 (a
// wavefront holds four p and all sixteen q, one p per 16-lane row), owns the
// elements (p + 32a, q + 16b):
//  - b < 2a, strictly below the diagonal 32 x 32 blocks: NA (NA - 1) values
//    in registers;
//  - b = 2a, 2a + 1, inside diagonal block a, lower half only: packed in LDS
//    (528 values a block).
// Step i starts from column i, fully updated and staged in LDS.  Every
// wavefront forms reflector i from it (a private copy, no barrier); every
// thread then applies the pending update (reflector i - 1) to its live
// elements and, in the same pass, adds their share of y = A v (A_rc v_c to row
// r and, off the diagonal, conj(A_rc) v_r to row c) and of v^H A v, while the
// owners of column i + 1 stage it.  Row sums reduce over the sixteen q lanes
// with DPP moves, column sums over the four p rows with the permlane swaps
// (two 16-column blocks at a time: each swap stage exchanges halves of two
// values, so one lane row ends with each of the four sums) and then over the
// eight wavefronts in LDS (barrier); each thread forms w = tau y - |tau|^2
// (v^H A v) / 2 v for its row and applies reflector i to the staged column
// i + 1 (barrier).  Elements whose 16-column block lies left of the trailing
// matrix, or wholly in the padding rows and columns n..32 NA - 1, are
// skipped.  Up to NA = 5 the row operands of a step stay in registers and the
// column operands of the next block are read while the current one is used.
constexpr int kRT = 512;
constexpr int kRegNA = 7;          // n <= 224

// The helpers below are chains of FMAs (one rounding per product): the pass
// over the live elements is bound by FP64 issue, 16 FMAs an element a step.
__device__ __forceinline__ void rank2_sub(cd &A, cd vr, cd wr, cd vc, cd wc) {
    // A -= v_r conj(w_c) + w_r conj(v_c)
    A.x = fma(-vr.x, wc.x, A.x);
    A.y = fma(-vr.y, wc.x, A.y);
    A.x = fma(-vr.y, wc.y, A.x);
    A.y = fma(vr.x, wc.y, A.y);
    A.x = fma(-wr.x, vc.x, A.x);
    A.y = fma(-wr.y, vc.x, A.y);
    A.x = fma(-wr.y, vc.y, A.x);
    A.y = fma(wr.x, vc.y, A.y);
}

__device__ __forceinline__ void cmac(cd &acc, cd a, cd b) {     // acc += a b
    acc.x = fma(a.x, b.x, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.y, b.x, acc.y);
}

__device__ __forceinline__ void cmac_conj(cd &acc, cd a, cd b) {   // acc += conj(a) b
    acc.x = fma(a.x, b.x, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(-a.y, b.x, acc.y);
}

// sum over the 16 lanes of a row, the same in every lane of it: lane pairs
// l ^ 1 and l ^ 2 (quad permutations), then l <-> 7 - l and l <-> 15 - l
// within the row (half mirror, mirror)
__device__ __forceinline__ double row_sum(double v) {
    v += dpp_double<0xB1>(v);
    v += dpp_double<0x4E>(v);
    v += dpp_double<0x141>(v);
    v += dpp_double<0x140>(v);
    return v;
}

// v[l] + v[l ^ 16] (permlane16 swap) or v[l] + v[l ^ 32] (permlane32 swap),
// added in the same order in both lanes of a pair
template <bool S32>
__device__ __forceinline__ double swap_sum(double v) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    auto L = S32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                 : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto H = S32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                 : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double a = __longlong_as_double(((long long)H[0] << 32) | L[0]);
    const double b = __longlong_as_double(((long long)H[1] << 32) | L[1]);
    return a + b;
}

__device__ __forceinline__ double rows_sum(double v) { return swap_sum<true>(swap_sum<false>(v)); }

// a <- rows (a0, b0, a2, b2), b <- rows (a1, b1, a3, b3) of the 16-lane rows
// (S32 false), or a <- (a0, a1, b0, b1), b <- (a2, a3, b2, b3) (S32 true)
template <bool S32>
__device__ __forceinline__ void swap_rows(double &a, double &b) {
    const long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
    auto L = S32 ? __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false)
                 : __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    auto H = S32 ? __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false)
                 : __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __longlong_as_double(((long long)H[0] << 32) | L[0]);
    b = __longlong_as_double(((long long)H[1] << 32) | L[1]);
}

// sums over the four lane rows of four values at once: lane row j ends with
// the sum of value j, added in the order of rows_sum ((r0 + r1) + (r2 + r3))
__device__ __forceinline__ double rows_sum4(double v0, double v1, double v2, double v3) {
    swap_rows<false>(v0, v1);
    swap_rows<false>(v2, v3);
    double s = v0 + v1, t = v2 + v3;
    swap_rows<true>(s, t);
    return s + t;
}
__device__ __forceinline__ double wave_sum_d(double v) { return rows_sum(row_sum(v)); }

// 1/q from the hardware reciprocal refined by two Newton steps: within an ulp
// of the divided form at a fraction of an IEEE division's latency, for the
// sequential chains (reflectors, Sturm counts, pivots).
__device__ __forceinline__ double recip(double q) {
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    return fma(r, fma(-q, r, 1.0), r);
}

// 1/q to ~2^-48 relative: the estimate (2^-24.4 measured) and one Newton step.
// For the tridiagonal recurrences (Sturm counts, the inverse iteration's LU),
// where it perturbs each step like a relative change of e_j^2 of ~20 eps and
// halves the chain of dependent operations per step.
__device__ __forceinline__ double recip1(double q) {
    const double r = __builtin_amdgcn_rcp(q);
    return fma(r, fma(-q, r, 1.0), r);
}

// 1/sqrt(s) (s > 0) from the hardware estimate refined by two Newton steps
__device__ __forceinline__ double rsqrt_nr(double s) {
    double r = __builtin_amdgcn_rsq(s);
    const double h = 0.5 * s;
    r = fma(r, fma(-h * r, r, 0.5), r);
    return fma(r, fma(-h * r, r, 0.5), r);
}

template <int NA>
__global__ void __launch_bounds__(kRT)
tridiag_reg_kernel(int n, cd *__restrict__ Ag, int lda, cd *__restrict__ tau, double *__restrict__ d,
                   double *__restrict__ e) {
    constexpr int NR = 32 * NA, NDB = 528, NW = kRT / 64, NK = NR / 64 + (NR % 64 != 0);
    constexpr int NREG = NA > 1 ? NA * (NA - 1) : 1;
    __shared__ cd D[NA * NDB];     // packed lower halves of the diagonal blocks
    __shared__ cd colbuf[NR];      // column i, up to date
    __shared__ cd colpre[NR];      // column i + 1 before reflector i
    __shared__ cd vsh[2][NR];      // reflectors (pending, current), written by wavefront 0
    __shared__ cd wsh[2][NR];      // their companions w
    __shared__ cd vw[NW][NR];      // the current reflector, one copy per wavefront
    __shared__ cd rowpart[NR];     // row sums of y
    __shared__ double colx[NW][NR], coly[NW][NR];   // column sums of y, one row per wavefront
    __shared__ double mured[NW];
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
    const int q = l & 15, p = 4 * wv + (l >> 4);
    const int prow = p * (p + 1) / 2;
    cd *vme = vw[wv];

    cd R[NREG];
#pragma unroll
    for (int a = 1; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < 2 * a; ++b) {
            const int r = p + 32 * a, c = q + 16 * b;
            R[a * (a - 1) + b] = (r < n && c < n) ? Ag[r + (size_t)c * lda] : make_double2(0.0, 0.0);
        }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int qq = q + 16 * h, r = 32 * a + p, c = 32 * a + qq;
            if (qq <= p) {
                cd x = (r < n && c < n) ? Ag[r + (size_t)c * lda] : make_double2(0.0, 0.0);
                if (qq == p) x.y = 0.0;
                D[a * NDB + prow + qq] = x;
            }
        }
    for (int r = tid; r < 2 * NR; r += kRT) {
        vsh[r / NR][r % NR] = make_double2(0.0, 0.0);
        wsh[r / NR][r % NR] = make_double2(0.0, 0.0);
    }
    for (int r = tid; r < NR; r += kRT) {
        cd x = r < n ? Ag[r] : make_double2(0.0, 0.0);
        if (r == 0) x.y = 0.0;
        colbuf[r] = x;
    }
    __syncthreads();
    for (int i = 0; i < n - 1; ++i) {
        const int pn = i & 1;
        const cd *vo = vsh[pn ^ 1], *wo = wsh[pn ^ 1];
        // reflector i from the staged column, formed by every wavefront
        cd t = make_double2(0.0, 0.0);
        {
            cd x[NK];
            double xn2 = 0.0;
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const int r = l + 64 * k;
                x[k] = r < NR ? colbuf[r] : make_double2(0.0, 0.0);
                if (r >= i + 2 && r < n) xn2 = fma(x[k].x, x[k].x, fma(x[k].y, x[k].y, xn2));
            }
            xn2 = wave_sum_d(xn2);
            const cd alpha = colbuf[i + 1];
            const bool ident = xn2 == 0.0 && alpha.y == 0.0;    // H(i) = I
            double beta = alpha.x;
            cd scal = make_double2(0.0, 0.0);
            if (!ident) {
                // sqrt and the two reciprocals by Newton-refined hardware
                // estimates: they sit on the step's critical path
                // (the norm gets one residual correction: its rounding sets
                // how far the reflector is from unitary)
                const double s2 = alpha.x * alpha.x + alpha.y * alpha.y + xn2;
                const double rs = rsqrt_nr(s2), g = s2 * rs;
                const double nrm = fma(fma(-g, g, s2), 0.5 * rs, g);
                beta = alpha.x >= 0.0 ? -nrm : nrm;
                const double ib = recip(beta);
                t = make_double2((beta - alpha.x) * ib, -alpha.y * ib);
                const double idd = recip((alpha.x - beta) * (alpha.x - beta) + alpha.y * alpha.y);
                scal = make_double2((alpha.x - beta) * idd, -alpha.y * idd);   // 1/(alpha - beta)
            }
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const int r = l + 64 * k;
                if (r < NR) {
                    cd v = make_double2(0.0, 0.0);
                    if (!ident && r > i && r < n)
                        v = r == i + 1 ? make_double2(1.0, 0.0)
                                       : make_double2(x[k].x * scal.x - x[k].y * scal.y,
                                                      x[k].x * scal.y + x[k].y * scal.x);
                    vme[r] = v;
                    if (wv == 0) {
                        vsh[pn][r] = v;
                        if (r >= i + 2 && r < n) Ag[r + (size_t)i * lda] = v;   // kept for the back-transformation
                    }
                }
            }
            if (wv == 0 && l == 0) {
                d[i] = colbuf[i].x;
                e[i] = beta;
                tau[i] = t;
            }
        }
        __builtin_amdgcn_wave_barrier();
        // pending update and y = A v in one pass over the live elements; the
        // owners of column i + 1 stage it
        const int bnx = (i + 1) >> 4, qnx = (i + 1) & 15;
        const int b0 = bnx;                       // first live 16-column block
        cd racc[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) racc[a] = make_double2(0.0, 0.0);
        double mup = 0.0;
        // row operands of the step: pending reflector, its w, current reflector
        constexpr bool HOIST = NA <= 5, AHEAD = HOIST;
        cd rvo[HOIST ? NA : 1], rwo[HOIST ? NA : 1], rvn[HOIST ? NA : 1];
        if constexpr (HOIST) {
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int r = p + 32 * a;
                rvo[a] = vo[r];
                rwo[a] = wo[r];
                rvn[a] = vme[r];
            }
        }
        auto row_vo = [&](int a) { if constexpr (HOIST) return rvo[a]; else return vo[p + 32 * a]; };
        auto row_wo = [&](int a) { if constexpr (HOIST) return rwo[a]; else return wo[p + 32 * a]; };
        auto row_vn = [&](int a) { if constexpr (HOIST) return rvn[a]; else return vme[p + 32 * a]; };
        // column operands and diagonal element of the next live block (read
        // one block ahead when the registers allow)
        cd nvo, nwo, nvn, nD;
        if constexpr (AHEAD) {
            const int c = q + 16 * b0, qq = q + 16 * (b0 & 1);
            nvo = vo[c];
            nwo = wo[c];
            nvn = vme[c];
            nD = D[(b0 >> 1) * NDB + prow + min(qq, p)];
        }
        cd cpair[2];     // column partials of blocks b - 1 and b
#pragma unroll
        for (int b = 0; b < 2 * NA; ++b) {
            cpair[b & 1] = make_double2(0.0, 0.0);
            if (b >= b0 && 16 * b < n) {
                const int c = q + 16 * b;
                if constexpr (!AHEAD) {
                    nvo = vo[c];
                    nwo = wo[c];
                    nvn = vme[c];
                    nD = D[(b >> 1) * NDB + prow + min(q + 16 * (b & 1), p)];
                }
                const cd voc = nvo, woc = nwo, vnc = nvn;
                // the diagonal-block element is read by every lane (clamped to
                // a valid element) ahead of the lanes' branch
                cd Ad = nD;
                if (AHEAD && b + 1 < 2 * NA) {
                    const int qn = q + 16 * ((b + 1) & 1);
                    nvo = vo[c + 16];
                    nwo = wo[c + 16];
                    nvn = vme[c + 16];
                    nD = D[((b + 1) >> 1) * NDB + prow + min(qn, p)];
                }
                cd cacc = make_double2(0.0, 0.0);
                const int qq = q + 16 * (b & 1), rd = 32 * (b >> 1) + p;
                if (qq <= p) {
                    rank2_sub(Ad, row_vo(b >> 1), row_wo(b >> 1), voc, woc);
                    D[(b >> 1) * NDB + prow + qq] = Ad;
                    cmac(racc[b >> 1], Ad, vnc);
                    if (qq < p) cmac_conj(cacc, Ad, row_vn(b >> 1));
                }
#pragma unroll
                for (int a = b / 2 + 1; a < NA; ++a) {
                    if (32 * a >= n) break;
                    cd &A = R[a * (a - 1) + b];
                    rank2_sub(A, row_vo(a), row_wo(a), voc, woc);
                    cmac(racc[a], A, vnc);
                    cmac_conj(cacc, A, row_vn(a));
                }
                if (b == bnx && q == qnx) {
                    if (qq <= p) colpre[rd] = Ad;
#pragma unroll
                    for (int a = b / 2 + 1; a < NA; ++a) colpre[p + 32 * a] = R[a * (a - 1) + b];
                }
                mup = fma(vnc.x, cacc.x, fma(vnc.y, cacc.y, mup));
                cpair[b & 1] = cacc;
            }
            if ((b & 1) && b >= b0 && 16 * (b - 1) < n) {
                // lane row j: x (j = 0) and y (j = 1) of block b - 1, x and y
                // (j = 2, 3) of block b
                const double sum = rows_sum4(cpair[0].x, cpair[0].y, cpair[1].x, cpair[1].y);
                const int j = l >> 4;
                ((j & 1) ? coly : colx)[wv][q + 16 * (b - 1 + (j >> 1))] = sum;
            }
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (32 * a + 31 <= i) continue;      // rows <= i are not needed
            if (32 * a >= n) break;
            const int r = p + 32 * a;
            const cd vnr = row_vn(a);
            cd s = racc[a];
            mup = fma(vnr.x, s.x, fma(vnr.y, s.y, mup));
            s.x = row_sum(s.x);
            s.y = row_sum(s.y);
            if (q == 0) rowpart[r] = s;
        }
        mup = wave_sum_d(mup);
        if (l == 0) mured[wv] = mup;
        __syncthreads();
        // w = tau y - |tau|^2 mu / 2 v (mu = v^H A v) for each row; column
        // i + 1 (rows > i) after reflector i, with w_(i+1) formed by every thread
        {
            double mu = 0.0;
#pragma unroll
            for (int k = 0; k < NW; ++k) mu += mured[k];
            const double hm = 0.5 * (t.x * t.x + t.y * t.y) * mu;
            auto w_of = [&](int r) {
                cd y = rowpart[r];
#pragma unroll
                for (int k = 0; k < NW; ++k) {
                    y.x += colx[k][r];
                    y.y += coly[k][r];
                }
                const cd v = vme[r];
                return make_double2(t.x * y.x - t.y * y.y - hm * v.x, t.x * y.y + t.y * y.x - hm * v.y);
            };
            const int j = i + 1;
            const cd wj = w_of(j), vj = vme[j];
            for (int r = tid; r < NR; r += kRT) {
                cd w = make_double2(0.0, 0.0), cb = make_double2(0.0, 0.0);
                if (r > i && r < n) {
                    w = w_of(r);
                    cb = colpre[r];
                    rank2_sub(cb, vme[r], w, vj, wj);
                    if (r == j) cb.y = 0.0;
                }
                wsh[pn][r] = w;
                colbuf[r] = cb;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[n - 1] = colbuf[n - 1].x;
        tau[n - 1] = make_double2(0.0, 0.0);
    }
}

// ---------------------------------------------------------------------------
// 2. tridiagonal eigenpairs
// Number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots),
// with 1/q from recip1().

__device__ int sturm_count(int n, const double *d, const double *e2, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    // operands of 8 steps are read from LDS together, off the recurrence chain
    constexpr int kS = 8;
    for (int j0 = 1; j0 < n; j0 += kS) {
        double db[kS], eb[kS];
#pragma unroll
        for (int t = 0; t < kS; ++t) {
            const int j = min(j0 + t, n - 1);
            db[t] = d[j] - x;
            eb[t] = e2[j - 1];
        }
#pragma unroll
        for (int t = 0; t < kS; ++t) {
            if (j0 + t < n) {
                q = fma(-eb[t], recip1(q), db[t]);
                if (fabs(q) < pivmin) q = -pivmin;
                cnt += q < 0.0;
            }
        }
    }
    return cnt;
}

// d, e (e[n-1] = 0) and e^2 into LDS, and the Gershgorin bounds of T:
// bounds = {lo, hi, |T|}.  Every kernel below derives |T| the same way.
__device__ void stage_tridiag(int n, const double *__restrict__ d, const double *__restrict__ e, double *sd,
                              double *se, double *se2, double *bounds, double *red, int nt) {
    const int tid = threadIdx.x;
    double lo = 1e300, hi = -1e300, tn = 0.0;
    for (int j = tid; j < n; j += nt) {
        const double dj = d[j], ej = (j < n - 1) ? e[j] : 0.0, em = j > 0 ? fabs(e[j - 1]) : 0.0;
        sd[j] = dj;
        se[j] = ej;
        se2[j] = ej * ej;
        const double r = em + fabs(ej);
        lo = fmin(lo, dj - r);
        hi = fmax(hi, dj + r);
        tn = fmax(tn, fabs(dj) + r);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
        tn = fmax(tn, __shfl_xor(tn, off, 64));
    }
    const int lane = tid & 63, w = tid >> 6, nw = nt / 64;
    if (lane == 0) {
        red[3 * w] = lo;
        red[3 * w + 1] = hi;
        red[3 * w + 2] = tn;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < nw; ++i) {
            lo = fmin(lo, red[3 * i]);
            hi = fmax(hi, red[3 * i + 1]);
            tn = fmax(tn, red[3 * i + 2]);
        }
        const double pad = 2.2e-16 * tn * 4.0 * n + 1e-300;
        bounds[0] = lo - pad;
        bounds[1] = hi + pad;
        bounds[2] = tn;
    }
    __syncthreads();
}

// Eigenvalues W[0..nvec) (ascending) by multisection: one workgroup per
// eigenvalue, its kBis threads evaluating the Sturm count at kBis interior
// points of the bracket per round, so the bracket shrinks (kBis + 1)-fold per
// round.  The new bracket is the largest point with count <= k and the
// smallest with count > k (min/max reductions, robust to counts that are not
// monotone in the last bits).
constexpr int kBis = 256;
__global__ void __launch_bounds__(kBis)
bisect_kernel(int n, int nvec, const double *__restrict__ d, const double *__restrict__ e, double *__restrict__ W) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *sd = reinterpret_cast<double *>(smem);
    double *se = sd + n, *se2 = se + n;
    __shared__ double red[3 * kBis / 64];
    __shared__ double bounds[3];
    __shared__ double s_lo[kBis / 64], s_hi[kBis / 64];
    stage_tridiag(n, d, e, sd, se, se2, bounds, red, kBis);
    const double tnorm = bounds[2];
    const double pivmin = fmax(1e-300, 1e-300 * tnorm);
    const int tid = threadIdx.x, kk = blockIdx.x;
    double lo = bounds[0], hi = bounds[1];
    for (int round = 0; round < 64; ++round) {
        const double x = lo + (double)(tid + 1) * ((hi - lo) / (double)(kBis + 1));
        const int cnt = sturm_count(n, sd, se2, x, pivmin);
        double nlo = cnt > kk ? lo : x, nhi = cnt > kk ? x : hi;
        nlo = fmax(nlo, lo);
        nhi = fmin(nhi, hi);
        for (int off = 32; off > 0; off >>= 1) {
            nlo = fmax(nlo, __shfl_xor(nlo, off, 64));
            nhi = fmin(nhi, __shfl_xor(nhi, off, 64));
        }
        if ((tid & 63) == 0) {
            s_lo[tid >> 6] = nlo;
            s_hi[tid >> 6] = nhi;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kBis / 64; ++w) {
            nlo = fmax(nlo, s_lo[w]);
            nhi = fmin(nhi, s_hi[w]);
        }
        __syncthreads();
        if (nlo > nhi) nlo = nhi = 0.5 * (nlo + nhi);   // rounding at the last bits
        const bool done = (nhi - nlo <= 2.0 * 2.2e-16 * fmax(fabs(nlo), fabs(nhi)) + pivmin) ||
                          (nlo == lo && nhi == hi);
        lo = nlo;
        hi = nhi;
        if (done) break;     // uniform: every thread holds the same bracket
    }
    if (tid == 0 && kk < nvec) W[kk] = 0.5 * (lo + hi);
}

// Inverse iteration steps: with the eigenvalue from bisection (to ~2 eps |l|)
// the first step leaves an error of ~eps |T| / gap times the start vector's
// weight ratio, the second squares it; vectors of a cluster are made
// orthogonal afterwards (cluster_gs_kernel).
constexpr int kInvIter = 2;
// Inverse iteration (kInvIter steps from a fixed start vector) with partial
// pivoting, one thread per eigenvector and `vpw` vectors per workgroup.  The
// LU factors and the iterate of a vector live in LDS, interleaved across the
// workgroup's vectors (element j of local vector k at [j vpw + k]), so each
// step of the sequential sweeps costs an LDS access instead of a global one.
// A repeated eigenvalue is shifted apart by multiples of 10 eps |T| so that
// the vectors of a cluster start apart.  Y (n x nvec, column-major, real).
__global__ void __launch_bounds__(64)
inviter_kernel(int n, int nvec, int vpw, const double *__restrict__ d, const double *__restrict__ e,
               const double *__restrict__ W, double *__restrict__ Y) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *sd = reinterpret_cast<double *>(smem);
    double *se = sd + n, *se2 = se + n;
    const size_t sn = (size_t)n * vpw;
    double *u = se2 + n;          // reciprocal of the U diagonal
    double *u1 = u + sn;          // first superdiagonal
    double *u2 = u1 + sn;         // second superdiagonal
    double *l = u2 + sn;          // multipliers
    double *pv = l + sn;          // 1 where rows j, j+1 were swapped
    double *z = pv + sn;          // iterate
    __shared__ double red[3];
    __shared__ double bounds[3];
    stage_tridiag(n, d, e, sd, se, se2, bounds, red, 64);
    const double tnorm = bounds[2];
    const double eps = 2.2e-16;
    const int kl = threadIdx.x, k = blockIdx.x * vpw + kl;
    if (kl >= vpw || k >= nvec) return;
    double lam = W[k];
    int kk = k;
    while (kk > 0 && fabs(W[kk - 1] - W[k]) < kClusterRel * tnorm) --kk;
    lam += (k - kk) * (eps * tnorm * 10.0);
    // LU of T - lam I: rows may swap with the next (branch-free: the pivot row
    // is chosen by selects, one reciprocal per step)
    double a_cur = sd[0] - lam, b_cur = se[0];
    for (int j = 0; j < n - 1; ++j) {
        const size_t o = (size_t)j * vpw + kl;
        const double c_j = se[j];                 // subdiagonal element below row j
        const double a_nxt = sd[j + 1] - lam, b_nxt = se[j + 1];
        if (a_cur == 0.0 && c_j == 0.0) a_cur = eps * tnorm + 1e-300;
        const bool sw = fabs(a_cur) < fabs(c_j);
        const double piv = sw ? c_j : a_cur, oth = sw ? a_cur : c_j;
        const double ip = recip1(piv), m = oth * ip;
        u[o] = ip;
        u1[o] = sw ? a_nxt : b_cur;
        u2[o] = sw ? b_nxt : 0.0;
        l[o] = m;
        pv[o] = sw ? 1.0 : 0.0;
        const double na = sw ? b_cur - m * a_nxt : a_nxt - m * b_cur;
        b_cur = sw ? -m * b_nxt : b_nxt;
        a_cur = na;
    }
    if (a_cur == 0.0) a_cur = eps * tnorm + 1e-300;
    {
        const size_t o = (size_t)(n - 1) * vpw + kl;
        u[o] = recip1(a_cur); u1[o] = 0.0; u2[o] = 0.0;
    }
    // generic start vector (distinct per vector)
    for (int j = 0; j < n; ++j)
        z[(size_t)j * vpw + kl] = 1.0 + 1e-4 * (double)(((unsigned)j * 7919u + (unsigned)k * 104729u) & 1023u);
    constexpr int kB = 8;
    double scale = 1.0;     // the iterate is z * scale
    for (int iter = 0; iter < kInvIter; ++iter) {
        // forward: apply the row operations (L^-1 P) to z, carrying z[j] in a register
        double zj = z[kl] * scale;
        for (int j0 = 0; j0 < n - 1; j0 += kB) {
            double lb[kB], pb[kB], zb[kB];
#pragma unroll
            for (int t = 0; t < kB; ++t) {
                const size_t o = (size_t)min(j0 + t, n - 2) * vpw + kl;
                lb[t] = l[o];
                pb[t] = pv[o];
                zb[t] = z[o + vpw];
            }
#pragma unroll
            for (int t = 0; t < kB; ++t) {
                if (j0 + t < n - 1) {
                    double zn = zb[t] * scale;
                    const bool s = pb[t] != 0.0;
                    const double lo_ = s ? zn : zj, hi_ = s ? zj : zn;
                    z[(size_t)(j0 + t) * vpw + kl] = lo_;
                    zj = hi_ - lb[t] * lo_;
                }
            }
        }
        // back substitution with U (upper, two superdiagonals)
        double z2 = 0.0, z1 = zj * u[(size_t)(n - 1) * vpw + kl];   // z[j+2], z[j+1]
        z[(size_t)(n - 1) * vpw + kl] = z1;
        double nn = z1 * z1;
        for (int j1 = n - 2; j1 >= 0; j1 -= kB) {
            double zb[kB], ub[kB], u1b[kB], u2b[kB];
#pragma unroll
            for (int t = 0; t < kB; ++t) {
                const size_t o = (size_t)max(j1 - t, 0) * vpw + kl;
                zb[t] = z[o];
                ub[t] = u[o];
                u1b[t] = u1[o];
                u2b[t] = u2[o];
            }
#pragma unroll
            for (int t = 0; t < kB; ++t) {
                if (j1 - t >= 0) {
                    const double zj0 = (zb[t] - u1b[t] * z1 - u2b[t] * z2) * ub[t];
                    z[(size_t)(j1 - t) * vpw + kl] = zj0;
                    nn += zj0 * zj0;
                    z2 = z1;
                    z1 = zj0;
                }
            }
        }
        scale = 1.0 / sqrt(nn);
    }
    double *yk = Y + (size_t)k * n;
    for (int j = 0; j < n; ++j) yk[j] = z[(size_t)j * vpw + kl] * scale;
}

// Modified Gram-Schmidt inside clusters of close eigenvalues (sequential
// over vectors, parallel over the inner products).  Outside a cluster
// inverse iteration leaves an orthogonality error of about eps |T| / gap.
__global__ void __launch_bounds__(kT)
cluster_gs_kernel(int n, int nvec, const double *__restrict__ d, const double *__restrict__ e,
                  const double *__restrict__ W, double *__restrict__ Y) {
    __shared__ double red[3 * kT / 64];
    __shared__ double bounds[3];
    __shared__ int s_any;
    const int tid = threadIdx.x;
    // |T| as the other kernels see it (no staging needed: bounds only)
    {
        double tn = 0.0;
        for (int j = tid; j < n; j += kT) {
            const double ej = (j < n - 1) ? e[j] : 0.0, em = j > 0 ? fabs(e[j - 1]) : 0.0;
            tn = fmax(tn, fabs(d[j]) + em + fabs(ej));
        }
        for (int off = 32; off > 0; off >>= 1) tn = fmax(tn, __shfl_xor(tn, off, 64));
        if ((tid & 63) == 0) red[tid >> 6] = tn;
        if (tid == 0) s_any = 0;
        __syncthreads();
        if (tid == 0) {
            for (int i = 1; i < kT / 64; ++i) tn = fmax(tn, red[i]);
            bounds[2] = tn;
        }
        __syncthreads();
    }
    const double tnorm = bounds[2];
    for (int k = 1 + tid; k < nvec; k += kT)
        if (fabs(W[k - 1] - W[k]) < kClusterRel * tnorm) s_any = 1;
    __syncthreads();
    if (!s_any) return;
    for (int k = 1; k < nvec; ++k) {
        int k0 = k;
        while (k0 > 0 && fabs(W[k0 - 1] - W[k0]) < kClusterRel * tnorm) --k0;
        if (k0 == k) continue;     // not in a cluster with a previous vector
        double *zk = Y + (size_t)k * n;
        for (int q = k0; q < k; ++q) {
            const double *zq = Y + (size_t)q * n;
            double dot = 0.0;
            for (int j = tid; j < n; j += kT) dot += zq[j] * zk[j];
            dot = block_reduce(dot, red);
            for (int j = tid; j < n; j += kT) zk[j] -= dot * zq[j];
            __syncthreads();
        }
        double nn = 0.0;
        for (int j = tid; j < n; j += kT) nn += zk[j] * zk[j];
        nn = block_reduce(nn, red);
        const double inv = 1.0 / sqrt(nn);
        for (int j = tid; j < n; j += kT) zk[j] *= inv;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 3. Z = H(0) ... H(n-2) Y for a block of columns per workgroup
constexpr int kBT = 256;
__global__ void __launch_bounds__(kBT)
backtransform_kernel(int n, int nvec, const cd *__restrict__ A, int lda, const cd *__restrict__ tau,
                     const double *__restrict__ Y, cd *__restrict__ Z, int ldz) {
    // one wavefront per column, 4 columns per workgroup
    const int lane = threadIdx.x & 63;
    const int col = blockIdx.x * (kBT / 64) + (threadIdx.x >> 6);
    if (col >= nvec) return;
    cd *z = Z + (size_t)col * ldz;
    for (int r = lane; r < n; r += 64) z[r] = make_double2(Y[(size_t)col * n + r], 0.0);
    for (int i = n - 2; i >= 0; --i) {
        const cd t = tau[i];
        if (t.x == 0.0 && t.y == 0.0) continue;
        // w = v^H z over rows i+1..n-1 (v(i+1) = 1)
        double wr = 0.0, wi = 0.0;
        for (int r = i + 1 + lane; r < n; r += 64) {
            cd vr = (r == i + 1) ? make_double2(1.0, 0.0) : A[r + (size_t)i * lda];
            cd zr = z[r];
            wr += vr.x * zr.x + vr.y * zr.y;
            wi += vr.x * zr.y - vr.y * zr.x;
        }
        for (int off = 32; off > 0; off >>= 1) {
            wr += __shfl_xor(wr, off, 64);
            wi += __shfl_xor(wi, off, 64);
        }
        // z -= tau v w
        const cd tw = make_double2(t.x * wr - t.y * wi, t.x * wi + t.y * wr);
        for (int r = i + 1 + lane; r < n; r += 64) {
            cd vr = (r == i + 1) ? make_double2(1.0, 0.0) : A[r + (size_t)i * lda];
            cd zr = z[r];
            zr.x -= vr.x * tw.x - vr.y * tw.y;
            zr.y -= vr.x * tw.y + vr.y * tw.x;
            z[r] = zr;
        }
    }
}

// The same product with the column in registers (lane l holds rows l + 64 s,
// s < S, n <= 64 S) and the reflectors staged in LDS kBTChunk at a time,
// shared by the four column wavefronts of the workgroup: the chain of n - 1
// reflectors then waits on global memory once per chunk, not twice per
// reflector.
constexpr int kBTChunk = 16;
template <int S>
__global__ void __launch_bounds__(kBT)
backtransform_reg_kernel(int n, int nvec, const cd *__restrict__ A, int lda, const cd *__restrict__ tau,
                         const double *__restrict__ Y, cd *__restrict__ Z, int ldz) {
    extern __shared__ __align__(16) unsigned char smem[];
    cd *vs = reinterpret_cast<cd *>(smem);      // [kBTChunk][n]: v of reflector i, zero above row i + 1
    __shared__ cd ts[kBTChunk];
    const int tid = threadIdx.x, lane = tid & 63;
    const int col = blockIdx.x * (kBT / 64) + (tid >> 6);
    const bool active = col < nvec;             // every thread stages; inactive ones compute nothing
    cd z[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int r = lane + 64 * s;
        z[s] = make_double2((active && r < n) ? Y[(size_t)col * n + r] : 0.0, 0.0);
    }
    // reflector q of a chunk (i = i1 - q), rows r = tid + kBT u (S rows per 64
    // lanes: n <= 64 S, so u < U); the next chunk's loads are in flight while
    // the current one is applied
    constexpr int U = (S + 3) / 4;
    cd vn[kBTChunk][U];
    cd tn = make_double2(0.0, 0.0);
    auto load_chunk = [&](int i1) {
        const int cnt = min(kBTChunk, i1 + 1);
        if (tid < cnt) tn = tau[i1 - tid];
#pragma unroll
        for (int q = 0; q < kBTChunk; ++q) {
            const int i = i1 - q;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = tid + kBT * u;
                cd x = make_double2(0.0, 0.0);
                if (q < cnt) {
                    if (r == i + 1) x = make_double2(1.0, 0.0);
                    else if (r > i + 1 && r < n) x = A[r + (size_t)i * lda];
                }
                vn[q][u] = x;
            }
        }
    };
    if (n >= 2) load_chunk(n - 2);
    for (int i1 = n - 2; i1 >= 0; i1 -= kBTChunk) {
        const int cnt = min(kBTChunk, i1 + 1);     // reflectors i1, i1 - 1, ..., i1 - cnt + 1
        __syncthreads();                           // the previous chunk is consumed
#pragma unroll
        for (int q = 0; q < kBTChunk; ++q)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = tid + kBT * u;
                if (q < cnt && r < n) vs[(size_t)q * n + r] = vn[q][u];
            }
        if (tid < cnt) ts[tid] = tn;
        __syncthreads();
        if (i1 - kBTChunk >= 0) load_chunk(i1 - kBTChunk);
        if (!active) continue;
        for (int q = 0; q < cnt; ++q) {
            const cd t = ts[q];
            if (t.x == 0.0 && t.y == 0.0) continue;
            const cd *v = vs + (size_t)q * n;
            // w = v^H z
            double wr = 0.0, wi = 0.0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int r = lane + 64 * s;
                if (r < n) {
                    const cd vr = v[r];
                    wr += vr.x * z[s].x + vr.y * z[s].y;
                    wi += vr.x * z[s].y - vr.y * z[s].x;
                }
            }
            wr = wave_sum_dpp(wr);
            wi = wave_sum_dpp(wi);
            // z -= tau v w
            const cd tw = make_double2(t.x * wr - t.y * wi, t.x * wi + t.y * wr);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int r = lane + 64 * s;
                if (r < n) {
                    const cd vr = v[r];
                    z[s].x -= vr.x * tw.x - vr.y * tw.y;
                    z[s].y -= vr.x * tw.y + vr.y * tw.x;
                }
            }
        }
    }
    if (active) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int r = lane + 64 * s;
            if (r < n) Z[(size_t)col * ldz + r] = z[s];
        }
    }
}

// Upper Cholesky factor of a Hermitian positive definite G (n x n, ldg), in
// place in the upper triangle: G = R^H R.  A pivot below rel * (original
// diagonal) stops the factorisation and sets *flag = column + 1.
//
// n <= kCholLds: G is staged in LDS and factorised by one 256-thread
// workgroup with one barrier per column.  Step j updates the trailing upper
// triangle with the unscaled pivot row, g_pq -= conj(g_jp) g_jq / g_jj, and
// leaves row j unscaled; the rows are divided by sqrt(g_jj) at the end.
// Each thread owns the same upper-triangle elements for the whole
// factorisation: element e = tid + kCT u, the elements numbered row by row
// from the last row up (row and column decoded once).  The trailing triangle
// of step j (rows > j) is then the first (n - 1 - j)(n - j) / 2 elements, so
// a step stops at the first slot past it.
constexpr int kCT = 256;
constexpr int kCholU = (kCholLds * (kCholLds + 1) / 2 + kCT - 1) / kCT;   // elements per thread
__global__ void __launch_bounds__(kCT)
cholesky_lds_kernel(int n, cd *__restrict__ Gg, int ldgg, double rel, int *__restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *d0 = reinterpret_cast<double *>(smem);
    cd *G = reinterpret_cast<cd *>(d0 + ((n + 1) & ~1));
    // odd leading dimension: row j of consecutive columns falls in different banks
    const int ln = n | 1;
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < n * n; idx += kCT) {
        const int p = idx % n, q = idx / n;
        if (p <= q) G[p + q * ln] = Gg[p + (size_t)q * ldgg];
    }
    for (int j = tid; j < n; j += kCT) d0[j] = Gg[j + (size_t)j * ldgg].x;
    if (tid == 0) {
        s_fail = 0;
        *flag = 0;
    }
    __syncthreads();
    // owned elements: row | column << 8; element e lies in row n - 1 - r with
    // r (r + 1) / 2 <= e < (r + 1) (r + 2) / 2
    int own[kCholU];
#pragma unroll
    for (int u = 0; u < kCholU; ++u) {
        const int e = tid + kCT * u;
        int r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
        while ((r + 1) * (r + 2) / 2 <= e) ++r;
        while (r * (r + 1) / 2 > e) --r;
        const int P = n - 1 - r;
        own[u] = P | ((P + e - r * (r + 1) / 2) << 8);
    }
    for (int j = 0; j < n; ++j) {
        const double piv = G[j + j * ln].x;
        if (!(piv > rel * d0[j])) {
            if (tid == 0) {
                s_fail = j + 1;
                *flag = j + 1;
            }
            break;      // every thread reads the same pivot and leaves together
        }
        const double ip = recip(piv);      // off the division's latency: the step's chain
        const int live = (n - 1 - j) * (n - j) / 2;      // elements of rows > j
#pragma unroll
        for (int u = 0; u < kCholU; ++u) {
            if (kCT * u >= live) break;
            const int P = own[u] & 255, Q = own[u] >> 8;
            if (tid + kCT * u < live) {
                const cd a = G[j + P * ln], b = G[j + Q * ln];
                cd g = G[P + Q * ln];
                // g -= conj(a) b / piv
                g.x -= (a.x * b.x + a.y * b.y) * ip;
                g.y -= (a.x * b.y - a.y * b.x) * ip;
                G[P + Q * ln] = g;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (s_fail) return;
    // smallest pivot relative to its original diagonal (flag[2..3] as a double)
    if (tid < 64) {
        double m = 1e300;
        for (int j = tid; j < n; j += 64) m = fmin(m, G[j + j * ln].x / d0[j]);
        for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
        if (tid == 0) *reinterpret_cast<double *>(flag + 2) = m;
    }
    // R: row j = (unscaled row j) / sqrt(g_jj), diagonal sqrt(g_jj)
    for (int idx = tid; idx < n * n; idx += kCT) {
        const int p = idx % n, q = idx / n;
        if (p > q) continue;
        const double r = sqrt(G[p + p * ln].x);
        cd g = G[p + q * ln];
        if (p == q) g = make_double2(r, 0.0);
        else g = make_double2(g.x / r, g.y / r);
        Gg[p + (size_t)q * ldgg] = g;
    }
}

// n > kCholLds: right-looking factorisation in global memory, one
// 1024-thread workgroup.
__global__ void __launch_bounds__(kT)
cholesky_kernel(int n, cd *__restrict__ G, int ldg, double rel, int *__restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *d0 = reinterpret_cast<double *>(smem);
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    double ratio = 1e300;      // smallest pivot relative to its diagonal (thread 0)
    for (int j = tid; j < n; j += kT) d0[j] = G[j + (size_t)j * ldg].x;
    if (tid == 0) {
        s_fail = 0;
        *flag = 0;
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const double piv = G[j + (size_t)j * ldg].x;
        if (!(piv > rel * d0[j])) {
            if (tid == 0) {
                s_fail = j + 1;
                *flag = j + 1;
            }
        }
        __syncthreads();
        if (s_fail) break;
        const double r = sqrt(piv), inv = 1.0 / r;
        for (int k = j + 1 + tid; k < n; k += kT) {
            cd g = G[j + (size_t)k * ldg];
            G[j + (size_t)k * ldg] = make_double2(g.x * inv, g.y * inv);
        }
        __syncthreads();
        if (tid == 0) G[j + (size_t)j * ldg] = make_double2(r, 0.0);
        if (tid == 0) ratio = fmin(ratio, piv / d0[j]);
        const int s = n - j - 1;
        for (int idx = tid; idx < s * s; idx += kT) {
            const int p = j + 1 + idx % s, q = j + 1 + idx / s;
            if (p > q) continue;
            cd a = G[j + (size_t)p * ldg], b = G[j + (size_t)q * ldg];
            cd g = G[p + (size_t)q * ldg];
            // g -= conj(a) b
            g.x -= a.x * b.x + a.y * b.y;
            g.y -= a.x * b.y - a.y * b.x;
            G[p + (size_t)q * ldg] = g;
        }
        __syncthreads();
    }
    if (tid == 0 && !s_fail) *reinterpret_cast<double *>(flag + 2) = ratio;
}

}  // namespace

// Returns 0 on success, or 1 + the first failing column.  With check false the
// pivot test is not read back (no host synchronisation) and 0 is returned.
// With check and min_ratio, the smallest pivot relative to its diagonal entry
// of G comes back in the same round trip (on success).
int cholesky_upper(Ctx &c, int n, cd *G, int ldg, double rel, bool check, double *min_ratio) {
    if (n > kT * 8) throw Error{WTB_ERR_UNSUPPORTED, "Cholesky size"};
    c.ws_flag.alloc(4);     // [0] failing column + 1, [2..3] the pivot ratio (double)
    const size_t d0 = (size_t)((n + 1) & ~1) * sizeof(double);
    if (n <= kCholLds)
        hipLaunchKernelGGL(cholesky_lds_kernel, dim3(1), dim3(kCT), d0 + (size_t)(n | 1) * n * sizeof(cd), c.stream, n, G,
                           ldg, rel, c.ws_flag.p);
    else
        hipLaunchKernelGGL(cholesky_kernel, dim3(1), dim3(kT), d0, c.stream, n, G, ldg, rel, c.ws_flag.p);
    WTB_LAUNCH_CHECK();
    if (!check) return 0;
    alignas(8) int hf[4] = {0, 0, 0, 0};
    WTB_HIP_CHECK(hipMemcpyAsync(hf, c.ws_flag.p, sizeof(hf), hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (min_ratio) *min_ratio = hf[0] ? 0.0 : *reinterpret_cast<const double *>(hf + 2);
    return hf[0];
}

// cholesky_upper without a host round trip: the status words (failing column
// + 1, then the smallest pivot ratio as a double) are copied to the pinned
// host words hflag[0..3] and ev is recorded after the copy.  Returns the
// device status word, which a conditional triangular solve reads.
const int *cholesky_upper_async(Ctx &c, int n, cd *G, int ldg, double rel, int *hflag, hipEvent_t ev) {
    cholesky_upper(c, n, G, ldg, rel, false);
    WTB_HIP_CHECK(hipMemcpyAsync(hflag, c.ws_flag.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipEventRecord(ev, c.stream));
    return c.ws_flag.p;
}

// Lowest nvec eigenpairs of the Hermitian matrix A (n x n, lda; destroyed).
void heev_lowest(Ctx &c, int n, int nvec, cd *A, int lda, double *W, cd *Z, int ldz) {
    if (n <= 0 || nvec <= 0) return;
    if (n > kT) throw Error{WTB_ERR_UNSUPPORTED, "dense eigenproblem larger than 1024"};
    if (nvec > n) nvec = n;
    c.ws_heev_c.alloc((size_t)n);
    c.ws_heev_d.alloc((size_t)2 * n + (size_t)n * nvec);
    cd *tau = c.ws_heev_c.p;
    double *d = c.ws_heev_d.p, *e = d + n, *Y = e + n;
    if (n <= 32 * kRegNA) {
        switch ((n + 31) / 32) {
#define WTB_TRID_CASE(K)                                                                                  \
    case K:                                                                                               \
        hipLaunchKernelGGL(tridiag_reg_kernel<K>, dim3(1), dim3(kRT), 0, c.stream, n, A, lda, tau, d, e); \
        break;
            WTB_TRID_CASE(1) WTB_TRID_CASE(2) WTB_TRID_CASE(3) WTB_TRID_CASE(4)
            WTB_TRID_CASE(5) WTB_TRID_CASE(6) WTB_TRID_CASE(7)
#undef WTB_TRID_CASE
        }
    } else {
        // panel width: V and W (2 nb n), y (n) and the partial sums (kT) in 156 KB
        const size_t fixed = ((size_t)n + kT) * sizeof(cd);
        const int nb = (int)std::min<size_t>(kMaxNb, (156 * 1024 - fixed) / (2 * (size_t)n * sizeof(cd)));
        hipLaunchKernelGGL(tridiag_blocked_kernel, dim3(1), dim3(kT), fixed + 2 * (size_t)nb * n * sizeof(cd),
                           c.stream, n, nb, A, lda, tau, d, e);
    }
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(bisect_kernel, dim3(nvec), dim3(kBis), 3 * (size_t)n * sizeof(double),
                       c.stream, n, nvec, d, e, W);
    WTB_LAUNCH_CHECK();
    // vectors per workgroup: d, e, e^2 and six n-vectors each in 150 KB
    const int vpw = (int)std::max<size_t>(1, std::min<size_t>(16, (150 * 1024 / sizeof(double) - 3 * (size_t)n) /
                                                                        (6 * (size_t)n)));
    hipLaunchKernelGGL(inviter_kernel, dim3((nvec + vpw - 1) / vpw), dim3(64),
                       (3 * (size_t)n + 6 * (size_t)n * vpw) * sizeof(double), c.stream, n, nvec, vpw, d, e, W, Y);
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(cluster_gs_kernel, dim3(1), dim3(kT), 0, c.stream, n, nvec, d, e, W, Y);
    WTB_LAUNCH_CHECK();
    const dim3 gbt((nvec + kBT / 64 - 1) / (kBT / 64));
    const size_t shbt = (size_t)kBTChunk * n * sizeof(cd);
    switch ((n + 63) / 64) {
#define WTB_BT_CASE(K)                                                                                          \
    case K:                                                                                                     \
        hipLaunchKernelGGL(backtransform_reg_kernel<K>, gbt, dim3(kBT), shbt, c.stream, n, nvec, A, lda, tau, Y, \
                           Z, ldz);                                                                             \
        break;
        WTB_BT_CASE(1) WTB_BT_CASE(2) WTB_BT_CASE(3) WTB_BT_CASE(4)
        WTB_BT_CASE(5) WTB_BT_CASE(6) WTB_BT_CASE(7) WTB_BT_CASE(8)
#undef WTB_BT_CASE
    default:
        hipLaunchKernelGGL(backtransform_kernel, gbt, dim3(kBT), 0, c.stream, n, nvec, A, lda, tau, Y, Z, ldz);
    }
    WTB_LAUNCH_CHECK();
}

}  // namespace wtb
