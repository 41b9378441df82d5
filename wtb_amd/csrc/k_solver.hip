// Lowest eigenpairs of the Hermitian BSE matrix on the GPU (replaces
// slepc_solver, src/module_slepc_solver.F90:12-97: Krylov-Schur for the nev
// smallest-magnitude pairs, tol 1e-8, max_it 200).
//
// Method: block LOBPCG with soft locking and the exact-diagonal
// preconditioner W = (diag(H) - l)^-1 R.  The BSE matrix is strongly
// diagonally dominant (diag = E_c - E_v), so the block converges in a handful
// of iterations and the cost is dominated by the H*X products (k_matvec).
// The Rayleigh-Ritz step is done in FP64 on the basis S = [X W P]: W and P
// are made orthogonal to what precedes them (two classical Gram-Schmidt
// passes) and orthonormal (CholQR2), so the projected problem S^H H S is a
// standard Hermitian eigenproblem solved on the GPU (tridiagonalisation,
// bisection, inverse iteration); a rank-deficient W falls back to SVQB on the
// Gram matrix.  The Ritz block is formed with tall-skinny GEMMs.
// Mixed precision: the products run on the FP32 store; a final refinement
// applies H to the converged block with exact int8-digit products recombined
// in FP64, redoes the Rayleigh-Ritz step and reports FP64 residuals.
#include "wtb_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace wtb {

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

namespace {

__device__ __forceinline__ cd cmul(cd a, cd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cconj(cd a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cscale(double s, cd a) { return make_double2(s * a.x, s * a.y); }

__device__ __forceinline__ void rr_pair(int m, int t, int i, int &p, int &q) {
    int mm = m - 1, x, y;
    if (i == 0) { x = t % mm; y = mm; }
    else { x = (t + i) % mm; y = (t - i + mm) % mm; }
    p = min(x, y);
    q = max(x, y);
}

constexpr int kJT = 1024;

__device__ double block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = (blockDim.x + 63) / 64;
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
}

// Parallel cyclic Jacobi on one workgroup.  A (m x m, column-major, lda) is
// Hermitian and is diagonalised in place; V (m x m, ldv) accumulates the
// rotations.  Round-robin pairing: each step applies m/2 disjoint rotations,
// updating every 2x2 block (a, b) as J_a^H A_ab J_b in one pass.
__global__ void __launch_bounds__(kJT)
jacobi_kernel(int m, cd *__restrict__ A, int lda, cd *__restrict__ V, int ldv, double *__restrict__ W,
              int *__restrict__ perm, int max_sweeps) {
    __shared__ double rc[kJT], rs[kJT];
    __shared__ cd rph[kJT];
    __shared__ int rp[kJT], rq[kJT];
    __shared__ double red[kJT / 64];
    const int tid = threadIdx.x;
    const int me = (m + 1) & ~1, np = me / 2;
    for (int idx = tid; idx < m * m; idx += kJT) {
        int r = idx % m, s = idx / m;
        V[r + (size_t)s * ldv] = make_double2(r == s ? 1.0 : 0.0, 0.0);
    }
    double fro = 0.0;
    for (int idx = tid; idx < m * m; idx += kJT) {
        cd a = A[idx % m + (size_t)(idx / m) * lda];
        fro += a.x * a.x + a.y * a.y;
    }
    fro = block_sum(fro, red);
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        double off = 0.0;
        for (int idx = tid; idx < m * m; idx += kJT) {
            int r = idx % m, s = idx / m;
            if (r != s) {
                cd a = A[r + (size_t)s * lda];
                off += a.x * a.x + a.y * a.y;
            }
        }
        off = block_sum(off, red);
        if (off <= 1e-28 * fro || off < 1e-280) break;
        for (int t = 0; t < me - 1; ++t) {
            for (int i = tid; i < np; i += kJT) {
                int p, q;
                rr_pair(me, t, i, p, q);
                double c = 1.0, s = 0.0;
                cd e = make_double2(1.0, 0.0);
                if (q < m) {
                    cd b = A[p + (size_t)q * lda];
                    double ab = sqrt(b.x * b.x + b.y * b.y);
                    if (ab > 0.0) {
                        double theta = (A[q + (size_t)q * lda].x - A[p + (size_t)p * lda].x) / (2.0 * ab);
                        double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                        e = make_double2(b.x / ab, b.y / ab);
                    }
                }
                rp[i] = p; rq[i] = q; rc[i] = c; rs[i] = s; rph[i] = e;
            }
            __syncthreads();
            // 2x2 block updates of A (upper block triangle, mirrored)
            const int nblk = np * (np + 1) / 2;
            for (int task = tid; task < nblk; task += kJT) {
                // map task -> (a, b) with a <= b
                int b = (int)((sqrt(8.0 * task + 1.0) - 1.0) / 2.0);
                while (b * (b + 1) / 2 > task) --b;
                while ((b + 1) * (b + 2) / 2 <= task) ++b;
                int a = task - b * (b + 1) / 2;
                int pa = rp[a], qa = rq[a], pb = rp[b], qb = rq[b];
                bool va = qa < m, vb = qb < m;
                // gather (missing dummy index -> zero row/col)
                cd x00 = A[pa + (size_t)pb * lda];
                cd x10 = va ? A[qa + (size_t)pb * lda] : make_double2(0, 0);
                cd x01 = vb ? A[pa + (size_t)qb * lda] : make_double2(0, 0);
                cd x11 = (va && vb) ? A[qa + (size_t)qb * lda] : make_double2(0, 0);
                // right multiply by J_b = [[c, s], [-s conj(e), c conj(e)]]
                double cb = rc[b], sb = rs[b];
                cd eb = cconj(rph[b]);
                cd y00 = csub(cscale(cb, x00), cscale(sb, cmul(eb, x01)));
                cd y01 = cadd(cscale(sb, x00), cscale(cb, cmul(eb, x01)));
                cd y10 = csub(cscale(cb, x10), cscale(sb, cmul(eb, x11)));
                cd y11 = cadd(cscale(sb, x10), cscale(cb, cmul(eb, x11)));
                // left multiply by J_a^H = [[c, -s e], [s, c e]]
                double ca = rc[a], sa = rs[a];
                cd ea = rph[a];
                cd z00 = csub(cscale(ca, y00), cscale(sa, cmul(ea, y10)));
                cd z01 = csub(cscale(ca, y01), cscale(sa, cmul(ea, y11)));
                cd z10 = cadd(cscale(sa, y00), cscale(ca, cmul(ea, y10)));
                cd z11 = cadd(cscale(sa, y01), cscale(ca, cmul(ea, y11)));
                if (a == b) {
                    z00.y = 0.0; z11.y = 0.0;
                    z01 = make_double2(0, 0); z10 = make_double2(0, 0);
                    A[pa + (size_t)pa * lda] = z00;
                    if (va) {
                        A[qa + (size_t)qa * lda] = z11;
                        A[pa + (size_t)qa * lda] = z01;
                        A[qa + (size_t)pa * lda] = z10;
                    }
                } else {
                    A[pa + (size_t)pb * lda] = z00;
                    A[pb + (size_t)pa * lda] = cconj(z00);
                    if (va) { A[qa + (size_t)pb * lda] = z10; A[pb + (size_t)qa * lda] = cconj(z10); }
                    if (vb) { A[pa + (size_t)qb * lda] = z01; A[qb + (size_t)pa * lda] = cconj(z01); }
                    if (va && vb) { A[qa + (size_t)qb * lda] = z11; A[qb + (size_t)qa * lda] = cconj(z11); }
                }
            }
            // V <- V J (column pairs)
            for (int task = tid; task < m * np; task += kJT) {
                int r = task % m, a = task / m;
                int p = rp[a], q = rq[a];
                if (q >= m) continue;
                double c = rc[a], s = rs[a];
                cd ec = cconj(rph[a]);
                cd x = V[r + (size_t)p * ldv], y = V[r + (size_t)q * ldv];
                cd ey = cmul(ec, y);
                V[r + (size_t)p * ldv] = csub(cscale(c, x), cscale(s, ey));
                V[r + (size_t)q * ldv] = cadd(cscale(s, x), cscale(c, ey));
            }
            __syncthreads();
        }
    }
    // ascending order
    for (int j = tid; j < m; j += kJT) {
        double w = A[j + (size_t)j * lda].x;
        int rank = 0;
        for (int i = 0; i < m; ++i) {
            double u = A[i + (size_t)i * lda].x;
            rank += (u < w) || (u == w && i < j);
        }
        perm[rank] = j;
        W[rank] = w;
    }
}

__global__ void gather_cols(int m, int ncol, const int *__restrict__ perm, const cd *__restrict__ V, int ldv,
                            cd *__restrict__ Z, int ldz) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (r < m && j < ncol) Z[r + (size_t)j * ldz] = V[r + (size_t)perm[j] * ldv];
}

__global__ void hermitize(int m, cd *A, int lda) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int s = blockIdx.y;
    if (r >= m || s >= m || r > s) return;
    cd a = A[r + (size_t)s * lda], b = A[s + (size_t)r * lda];
    cd h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    if (r == s) h.y = 0.0;
    A[r + (size_t)s * lda] = h;
    A[s + (size_t)r * lda] = cconj(h);
}

// p[0..n) = v
__global__ void fill_value(int n, double v, double *__restrict__ p) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

__global__ void set_units(long N, int b, const int *__restrict__ idx, cd *__restrict__ X) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= N) return;
    X[i + (size_t)j * N] = make_double2(i == idx[j] ? 1.0 : 0.0, 0.0);
}

// H X for the unit block X = [e_idx[0] ... e_idx[b-1]]: column j of H is
// stored column j (the store is exactly Hermitian), so AX[:, c] is stored
// column idx[c] converted to FP64, plus the FP64 diagonal at row idx[c] when
// the store keeps a zero diagonal (diag != null).  Exact: no product rounding.
template <class T>
__global__ void unit_images(long N, long ld, int b, const int *__restrict__ idx, const T *__restrict__ H,
                            const double *__restrict__ diag, cd *__restrict__ AX) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= N || c >= b) return;
    const long j = idx[c];
    const T *h = H + 2 * (i + (size_t)j * ld);
    cd v = make_double2((double)h[0], (double)h[1]);
    if (diag && i == j) v.x += diag[i];
    AX[i + (size_t)c * N] = v;
}

// norm2[j] += sum_i |AX_ij - lam_j X_ij|^2   (lam may be null: plain column norms of AX)
// Column norms^2 of AX - lam X (or of AX when lam is null), one workgroup per
// column (a fixed reduction order, no atomics).
constexpr int kNT = 1024;
__global__ void __launch_bounds__(kNT)
resid_norms(long N, const cd *__restrict__ X, const cd *__restrict__ AX, const double *__restrict__ lam,
            double *__restrict__ norm2) {
    __shared__ double red[kNT / 64];
    const int j = blockIdx.x;
    const cd *ax = AX + (size_t)j * N;
    const cd *x = X + (size_t)j * N;
    const double l = lam ? lam[j] : 0.0;
    double s = 0.0;
    for (long i = threadIdx.x; i < N; i += kNT) {
        cd a = ax[i];
        if (lam) {
            const cd xi = x[i];
            a.x -= l * xi.x;
            a.y -= l * xi.y;
        }
        s += a.x * a.x + a.y * a.y;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kNT / 64; ++w) t += red[w];
        norm2[j] = t;
    }
}

__global__ void precondition(long N, const cd *__restrict__ X, const cd *__restrict__ AX,
                             const double *__restrict__ lam, const double *__restrict__ diag,
                             const int *__restrict__ act, const double *__restrict__ sc, cd *__restrict__ W,
                             double floor_) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (i >= N) return;
    int j = act[a];
    double l = lam[j];
    cd x = X[i + (size_t)j * N], ax = AX[i + (size_t)j * N];
    double den = diag[i] - l;
    if (fabs(den) < floor_) den = den < 0 ? -floor_ : floor_;
    double f = sc[a] / den;
    W[i + (size_t)a * N] = make_double2(f * (ax.x - l * x.x), f * (ax.y - l * x.y));
}

__global__ void gather_scale(long N, const int *__restrict__ act, const double *__restrict__ sc,
                             const cd *__restrict__ P, const cd *__restrict__ AP, cd *__restrict__ Po,
                             cd *__restrict__ APo) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (i >= N) return;
    int j = act[a];
    double f = sc ? sc[a] : 1.0;
    cd p = P[i + (size_t)j * N], ap = AP[i + (size_t)j * N];
    Po[i + (size_t)a * N] = cscale(f, p);
    APo[i + (size_t)a * N] = cscale(f, ap);
}

// W[:, a] /= sqrt(norm2[a]) (columns of zero norm are left alone)
__global__ void unit_cols(long N, const double *__restrict__ norm2, cd *__restrict__ W) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (i >= N) return;
    const double n2 = norm2[a];
    if (n2 > 0.0) W[i + (size_t)a * N] = cscale(1.0 / sqrt(n2), W[i + (size_t)a * N]);
}

// T = U[:, off:off+r] diag(d^-1/2)
__global__ void svqb_t(int m, int r, int off, const cd *__restrict__ U, int ldu, const double *__restrict__ d,
                       cd *__restrict__ T, int ldt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= m || j >= r) return;
    T[i + (size_t)j * ldt] = cscale(1.0 / sqrt(d[off + j]), U[i + (size_t)(off + j) * ldu]);
}

const rocblas_double_complex kOne = {1.0, 0.0};

inline const rocblas_double_complex *rb(const cd *p) { return reinterpret_cast<const rocblas_double_complex *>(p); }
inline rocblas_double_complex *rb(cd *p) { return reinterpret_cast<rocblas_double_complex *>(p); }

// C (m x n) = A^H B, A: K x m, B: K x n (column-major, ld = K)
// sum of nsplit partial m x n products (ld m, stride m n) into C (ldc)
// ---------------------------------------------------------------------------
// Tall-skinny Gram products C = A^H B (K >> m, n) on FP64 MFMA
// (v_mfma_f64_16x16x4f64; operand lane l holds row/column l % 16 at k = l / 16,
// result register r of lane l holds element (4 r + l / 16, l % 16)).
// A workgroup owns an MT x NT block of C for one K chunk: each K-step of A
// and B is staged in LDS once and shared by the four wavefronts, which split
// the 16 x 16 sub-tiles; the chunk partials are summed by reduce_chunks in a
// fixed order (no atomics, so the result does not depend on scheduling).
typedef double d4 __attribute__((ext_vector_type(4)));

// HERM (m == n, square blocks): only the upper triangle of C is wanted.
// Only the upper blocks are launched, and a diagonal block forms only
// its tiles with ti <= tj (15 of 25 at 80 x 80), dealt to the wavefronts in
// that order; reduce_chunks then mirrors the upper triangle.
template <int MT, int NT, bool HERM>
__global__ void __launch_bounds__(256)
gram_block_kernel(int m, int n, int bx, long K, long kc, const cd *__restrict__ A, int lda, const cd *__restrict__ B,
                  int ldb, cd *__restrict__ part) {
    constexpr int KS = 16, TI = MT / 16, TJ = NT / 16, NTILE = TI * TJ, TPW = (NTILE + 3) / 4;
    constexpr int AE = KS * MT / 256, BE = KS * NT / 256;       // staged elements per thread
    static_assert(KS * MT % 256 == 0 && KS * NT % 256 == 0, "block must split evenly over the workgroup");
    static_assert(!HERM || MT == NT, "Hermitian blocks are square");
    __shared__ cd sa[KS][MT + 1], sb[KS][NT + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // output block (bi, bj) of blockIdx.x: column-major over the bx x by
    // blocks, or over the upper ones (bi <= bj) of a Hermitian product
    int bi, bj;
    if (HERM) {
        const int t = blockIdx.x;
        bj = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
        while ((bj + 1) * (bj + 2) / 2 <= t) ++bj;
        while (bj * (bj + 1) / 2 > t) --bj;
        bi = t - bj * (bj + 1) / 2;
    } else {
        bi = blockIdx.x % bx;
        bj = blockIdx.x / bx;
    }
    const int i0 = bi * MT, j0 = bj * NT;
    const bool tri = HERM && bi == bj;
    const int ntile = tri ? TI * (TI + 1) / 2 : NTILE;
    // tile t of the wavefront -> (ti, tj): column-major over all tiles, or
    // over the upper ones (ti <= tj) of a diagonal block
    int tis[TPW], tjs[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = w + 4 * t;
        if (tri) {
            int tj = 0;
            while ((tj + 1) * (tj + 2) / 2 <= tile) ++tj;
            tjs[t] = tj;
            tis[t] = tile - tj * (tj + 1) / 2;
        } else {
            tis[t] = tile % TI;
            tjs[t] = tile / TI;
        }
    }
    const long k0 = (long)blockIdx.y * kc, k1 = min(K, k0 + kc);
    d4 cre[TPW], cim[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        cre[t] = d4{0.0, 0.0, 0.0, 0.0};
        cim[t] = d4{0.0, 0.0, 0.0, 0.0};
    }
    // columns past m (n) read the last valid one; rows past the chunk read zero
    cd ra[AE], rb[BE];
    auto load = [&](long s) {
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            const int idx = tid + 256 * e, k = idx % KS, i = idx / KS;
            const long kk = s + k;
            ra[e] = kk < k1 ? A[kk + (size_t)min(i0 + i, m - 1) * lda] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            const int idx = tid + 256 * e, k = idx % KS, j = idx / KS;
            const long kk = s + k;
            rb[e] = kk < k1 ? B[kk + (size_t)min(j0 + j, n - 1) * ldb] : make_double2(0.0, 0.0);
        }
    };
    load(k0);
    for (long s = k0; s < k1; s += KS) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            const int idx = tid + 256 * e;
            sa[idx % KS][idx / KS] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            const int idx = tid + 256 * e;
            sb[idx % KS][idx / KS] = rb[e];
        }
        __syncthreads();
        if (s + KS < k1) load(s + KS);      // next K-step in flight during the MFMAs
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
            const int k = 4 * q + (lane >> 4);
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int tile = w + 4 * t;
                if (tile < ntile) {
                    const int ti = tis[t], tj = tjs[t];
                    const cd a = sa[k][16 * ti + (lane & 15)], b = sb[k][16 * tj + (lane & 15)];
                    // conj(a) b
                    cre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, cre[t], 0, 0, 0);
                    cre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, cre[t], 0, 0, 0);
                    cim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, cim[t], 0, 0, 0);
                    cim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.x, cim[t], 0, 0, 0);
                }
            }
        }
    }
    cd *p = part + (size_t)blockIdx.y * m * n;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = w + 4 * t;
        if (tile >= ntile) continue;
        const int ti = tis[t], tj = tjs[t];
        const int jj = j0 + 16 * tj + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ii = i0 + 16 * ti + 4 * r + (lane >> 4);
            if (ii < m && jj < n) p[ii + (size_t)jj * m] = make_double2(cre[t][r], cim[t][r]);
        }
    }
}

// Single-block form (C of at most 80 x 80, the whole product one block; upper
// tiles only when HERM): a wavefront owns NS tiles (w + 4 t, repeating tile 0
// past the last one, which is then not stored), and per 4-deep k slice the
// first product of every tile is issued, then the second, ..., so consecutive
// MFMAs never share an accumulator.  The partial sums per tile are those of
// gram_block_kernel, in the same order.
template <int MT, bool HERM>
__global__ void __launch_bounds__(256)
gram_one_kernel(int m, int n, long K, long kc, const cd *__restrict__ A, int lda, const cd *__restrict__ B, int ldb,
                cd *__restrict__ part) {
    constexpr int KS = 16, TI = MT / 16, NTILE = HERM ? TI * (TI + 1) / 2 : TI * TI, NS = (NTILE + 3) / 4;
    constexpr int AE = KS * MT / 256;
    static_assert(KS * MT % 256 == 0, "block must split evenly over the workgroup");
    __shared__ cd sa[KS][MT + 1], sb[KS][MT + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int tis[NS], tjs[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        int tile = w + 4 * t;
        if (tile >= NTILE) tile = 0;
        if (HERM) {
            int tj = 0;
            while ((tj + 1) * (tj + 2) / 2 <= tile) ++tj;
            tjs[t] = tj;
            tis[t] = tile - tj * (tj + 1) / 2;
        } else {
            tis[t] = tile % TI;
            tjs[t] = tile / TI;
        }
    }
    const long k0 = (long)blockIdx.y * kc, k1 = min(K, k0 + kc);
    // columns past m (n) read the last valid one; rows past the chunk read zero
    cd ra[AE], rb[AE];
    auto load = [&](long s) {
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            const int idx = tid + 256 * e, k = idx % KS, i = idx / KS;
            const long kk = s + k;
            ra[e] = kk < k1 ? A[kk + (size_t)min(i, m - 1) * lda] : make_double2(0.0, 0.0);
            rb[e] = kk < k1 ? B[kk + (size_t)min(i, n - 1) * ldb] : make_double2(0.0, 0.0);
        }
    };
    d4 cre[NS], cim[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        cre[t] = d4{0.0, 0.0, 0.0, 0.0};
        cim[t] = d4{0.0, 0.0, 0.0, 0.0};
    }
    load(k0);
    for (long s = k0; s < k1; s += KS) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            const int idx = tid + 256 * e;
            sa[idx % KS][idx / KS] = ra[e];
            sb[idx % KS][idx / KS] = rb[e];
        }
        __syncthreads();
        if (s + KS < k1) load(s + KS);      // next K-step in flight during the MFMAs
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
            const int k = 4 * q + (lane >> 4);
            cd a[NS], b[NS];
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                a[t] = sa[k][16 * tis[t] + (lane & 15)];
                b[t] = sb[k][16 * tjs[t] + (lane & 15)];
            }
            // conj(a) b
#pragma unroll
            for (int t = 0; t < NS; ++t) cre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t].x, b[t].x, cre[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NS; ++t) cim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t].x, b[t].y, cim[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NS; ++t) cre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t].y, b[t].y, cre[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NS; ++t)
                cim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[t].y, b[t].x, cim[t], 0, 0, 0);
        }
    }
    cd *p = part + (size_t)blockIdx.y * m * n;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        if (w + 4 * t >= NTILE) continue;
        const int jj = 16 * tjs[t] + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ii = 16 * tis[t] + 4 * r + (lane >> 4);
            if (ii < m && jj < n) p[ii + (size_t)jj * m] = make_double2(cre[t][r], cim[t][r]);
        }
    }
}

// C[i + j ldc] = sum over chunks of part[s][i + j m]: lanes take 64 consecutive
// elements, the kRedG wavefronts every kRedG-th chunk with eight loads in
// flight, then an LDS sum in a fixed order.  Sixteen wavefronts per block keep
// enough loads in flight to stream the partials near the HBM rate.
constexpr int kRedG = 16;
__global__ void __launch_bounds__(64 * kRedG)
reduce_chunks(int m, int n, int nch, const cd *__restrict__ part, cd *__restrict__ C, int ldc, int herm) {
    __shared__ cd red[kRedG][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const long mn = (long)m * n, e = (long)blockIdx.x * 64 + lane;
    const long ec = e < mn ? e : mn - 1;
    cd acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = make_double2(0.0, 0.0);
    for (int s0 = g; s0 < nch; s0 += 8 * kRedG) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int s = s0 + kRedG * u;
            if (s < nch) {
                const cd v = part[(size_t)s * mn + ec];
                acc[u].x += v.x;
                acc[u].y += v.y;
            }
        }
    }
    cd t = acc[0];
#pragma unroll
    for (int u = 1; u < 8; ++u) {
        t.x += acc[u].x;
        t.y += acc[u].y;
    }
    red[g][lane] = t;
    __syncthreads();
    if (g == 0 && e < mn) {
        cd r = red[0][lane];
        for (int q = 1; q < kRedG; ++q) {
            r.x += red[q][lane].x;
            r.y += red[q][lane].y;
        }
        const long i = e % m, j = e / m;
        if (!herm) {
            C[i + (size_t)j * ldc] = r;
        } else if (i <= j) {
            // herm: the upper triangle (the only part formed) and its mirror
            if (i == j) r.y = 0.0;
            C[i + (size_t)j * ldc] = r;
            C[j + (size_t)i * ldc] = make_double2(r.x, -r.y);
        }
    }
}

// C (m x n) = A^H B for tall A (K x m) and B (K x n).  Output blocks of up to
// 80 x 80; K is split so that about 256 workgroups run.
// herm (m == n): the Hermitian part of A^H B from its upper triangle (exact
// for B = A; for an approximately Hermitian A^H B it replaces the mean of C
// and C^H by the upper triangle mirrored).
void gram(Ctx &c, int m, int n, long K, const cd *A, int lda, const cd *B, int ldb, cd *C, int ldc,
          bool herm = false) {
    if (m <= 0 || n <= 0) return;
    if (herm && m != n) throw Error{WTB_ERR_STATE, "Hermitian Gram product of a non-square block"};
    const int bs = std::min(80, std::max((m + 15) / 16, (n + 15) / 16) * 16);
    const int bx = (m + bs - 1) / bs, by = (n + bs - 1) / bs;
    // output blocks (the upper ones only when Hermitian) times K chunks: about
    // two workgroups per CU
    const int nblk = herm ? bx * (bx + 1) / 2 : bx * by;
    // one block: about one workgroup per CU (the single-block kernel holds
    // more accumulators); several blocks: about two per CU
    int nsp = std::max(1, (nblk == 1 ? 256 : 512) / nblk);
    long kc = (K + nsp - 1) / nsp;
    kc = std::max<long>(16, (kc + 15) / 16 * 16);
    nsp = (int)((K + kc - 1) / kc);
    c.ws_gram.alloc((size_t)nsp * m * n);
    const dim3 g(nblk, nsp);
    if (nblk == 1) {
        switch (bs) {
#define WTB_GRAM1_CASE(S)                                                                                         \
    case S:                                                                                                       \
        if (herm)                                                                                                 \
            hipLaunchKernelGGL((gram_one_kernel<S, true>), g, dim3(256), 0, c.stream, m, n, K, kc, A, lda, B, ldb, \
                               c.ws_gram.p);                                                                      \
        else                                                                                                      \
            hipLaunchKernelGGL((gram_one_kernel<S, false>), g, dim3(256), 0, c.stream, m, n, K, kc, A, lda, B, ldb, \
                               c.ws_gram.p);                                                                      \
        break;
            WTB_GRAM1_CASE(16) WTB_GRAM1_CASE(32) WTB_GRAM1_CASE(48) WTB_GRAM1_CASE(64) WTB_GRAM1_CASE(80)
#undef WTB_GRAM1_CASE
        }
    } else switch (bs) {
#define WTB_GRAM_CASE(S)                                                                                       \
    case S:                                                                                                    \
        if (herm)                                                                                              \
            hipLaunchKernelGGL((gram_block_kernel<S, S, true>), g, dim3(256), 0, c.stream, m, n, bx, K, kc, A, \
                               lda, B, ldb, c.ws_gram.p);                                                      \
        else                                                                                                   \
            hipLaunchKernelGGL((gram_block_kernel<S, S, false>), g, dim3(256), 0, c.stream, m, n, bx, K, kc, A, \
                               lda, B, ldb, c.ws_gram.p);                                                      \
        break;
        WTB_GRAM_CASE(16) WTB_GRAM_CASE(32) WTB_GRAM_CASE(48) WTB_GRAM_CASE(64) WTB_GRAM_CASE(80)
#undef WTB_GRAM_CASE
    }
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_chunks, dim3((unsigned)(((long)m * n + 63) / 64)), dim3(64 * kRedG), 0, c.stream, m, n, nsp,
                       c.ws_gram.p, C, ldc, herm ? 1 : 0);
    WTB_LAUNCH_CHECK();
}

// B <- B R^-1 in place for tall B (N x nb, ldb) and upper triangular R
// (nb x nb, ldr): each row of B solves y R = x by forward substitution, 16
// columns at a time.  R is staged in LDS, packed by columns (column j at
// j (j + 1) / 2), with the reciprocals of its diagonal.  gridDim.y = 2
// applies the same R to B2.
// Four lanes per row.  Per 16-column block J: the update x_J -= y_k R[k, J] by
// the solved columns is split over the lanes by k mod 4 and summed with lane
// shuffles in a fixed order; the substitution inside the block then runs in
// all four lanes alike, and lane q writes the columns j = q mod 4.  The
// update of lane q reads only columns it wrote itself (k = q mod 4), so its
// own program order makes them visible.
constexpr int kTr4 = 256;
__global__ void __launch_bounds__(kTr4)
trsm_ru4_kernel(long N, int nb, const cd *__restrict__ R, int ldr, cd *__restrict__ B, cd *__restrict__ B2,
                long ldb, const int *__restrict__ skip) {
    constexpr int JB = 16;
    if (skip && *skip) return;      // the factorisation failed: B stays as it is
    extern __shared__ __align__(16) unsigned char smem[];
    cd *rp = reinterpret_cast<cd *>(smem);                       // packed upper R
    double *rinv = reinterpret_cast<double *>(rp + (size_t)nb * (nb + 1) / 2);
    {   // stage R: eight loads in flight per thread
        const int tot = nb * nb;
        for (int base = 0; base < tot; base += 8 * kTr4) {
            cd v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int idx = min(base + t * kTr4 + (int)threadIdx.x, tot - 1);
                v[t] = R[idx % nb + (size_t)(idx / nb) * ldr];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int idx = base + t * kTr4 + threadIdx.x, i = idx % nb, j = idx / nb;
                if (idx < tot && i <= j) {
                    rp[(size_t)j * (j + 1) / 2 + i] = v[t];
                    if (i == j) rinv[j] = 1.0 / v[t].x;
                }
            }
        }
    }
    __syncthreads();
    cd *M = blockIdx.y ? B2 : B;
    const int q = threadIdx.x & 3;
    const long row = (long)blockIdx.x * (kTr4 / 4) + (threadIdx.x >> 2);
    const bool own = row < N;
    cd *x = M + (own ? row : N - 1);      // rows past N read a valid row and write nothing
    for (int j0 = 0; j0 < nb; j0 += JB) {
        cd acc[JB];
#pragma unroll
        for (int u = 0; u < JB; ++u) acc[u] = make_double2(0.0, 0.0);
        // this lane's share of the update: solved columns k = 4 t + q < j0,
        // read back eight at a time
        const int nt = j0 / 4;
        for (int t0 = 0; t0 < nt; t0 += 8) {
            cd yk[8];
#pragma unroll
            for (int tt = 0; tt < 8; ++tt)
                yk[tt] = x[(size_t)(4 * min(t0 + tt, nt - 1) + q) * ldb];
#pragma unroll
            for (int tt = 0; tt < 8; ++tt) {
                if (t0 + tt < nt) {
                    const int k = 4 * (t0 + tt) + q;
#pragma unroll
                    for (int u = 0; u < JB; ++u) {
                        const int j = min(j0 + u, nb - 1);
                        const cd r = rp[(size_t)j * (j + 1) / 2 + k];
                        acc[u].x += yk[tt].x * r.x - yk[tt].y * r.y;
                        acc[u].y += yk[tt].x * r.y + yk[tt].y * r.x;
                    }
                }
            }
        }
        // sum over the four lanes of the row (xor 1, then xor 2: the same
        // order in every lane, so all four hold identical sums)
#pragma unroll
        for (int u = 0; u < JB; ++u) {
#pragma unroll
            for (int o = 1; o <= 2; o <<= 1) {
                acc[u].x += __shfl_xor(acc[u].x, o, 64);
                acc[u].y += __shfl_xor(acc[u].y, o, 64);
            }
        }
        cd y[JB];
#pragma unroll
        for (int u = 0; u < JB; ++u) {
            const int j = j0 + u;
            y[u] = make_double2(0.0, 0.0);
            if (j < nb) {
                const cd xv = x[(size_t)j * ldb];
                y[u] = make_double2(xv.x - acc[u].x, xv.y - acc[u].y);
            }
        }
        // substitution inside the block (every lane)
#pragma unroll
        for (int u = 0; u < JB; ++u) {
            const int j = j0 + u;
            if (j < nb) {
#pragma unroll
                for (int w = 0; w < u; ++w) {
                    const cd r = rp[(size_t)j * (j + 1) / 2 + j0 + w];
                    y[u].x -= y[w].x * r.x - y[w].y * r.y;
                    y[u].y -= y[w].x * r.y + y[w].y * r.x;
                }
                const double ri = rinv[j];
                y[u].x *= ri;
                y[u].y *= ri;
            }
        }
#pragma unroll
        for (int s4 = 0; s4 < JB / 4; ++s4) {
            const cd v0 = y[4 * s4], v1 = y[4 * s4 + 1], v2 = y[4 * s4 + 2], v3 = y[4 * s4 + 3];
            const cd v = q == 0 ? v0 : (q == 1 ? v1 : (q == 2 ? v2 : v3));
            const int j = j0 + 4 * s4 + q;
            if (own && j < nb) x[(size_t)j * ldb] = v;
        }
    }
}

// skip (device, may be null): a nonzero value leaves B and B2 unchanged.
void trsm_right_upper(Ctx &c, long N, int nb, const cd *R, int ldr, cd *B, cd *B2, const int *skip = nullptr) {
    const size_t sh = (size_t)nb * (nb + 1) / 2 * sizeof(cd) + (size_t)nb * sizeof(double);
    if (sh > 150 * 1024) {      // R too large to stage: library triangular solve
        if (skip) throw Error{WTB_ERR_STATE, "conditional triangular solve of a large factor"};
        for (cd *M : {B, B2})
            if (M)
                WTB_BLAS_CHECK(rocblas_ztrsm(c.blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                             rocblas_diagonal_non_unit, (rocblas_int)N, nb, &kOne, rb(R), ldr, rb(M),
                                             (rocblas_int)N));
        return;
    }
    hipLaunchKernelGGL(trsm_ru4_kernel, dim3((unsigned)((N + kTr4 / 4 - 1) / (kTr4 / 4)), B2 ? 2 : 1), dim3(kTr4),
                       sh, c.stream, N, nb, R, ldr, B, B2, N, skip);
    WTB_LAUNCH_CHECK();
}

// C (M x n) = A (M x k) B (k x n)
// ---------------------------------------------------------------------------
// Tall products C_s = A_s B (or C_s -= A_s B) for s = 0, 1 with one small
// coefficient block B (k x n, n <= 80) shared by both: the Ritz and
// projection updates of the block solver, where the pair (X, H X) takes the
// same combination.  FP64 MFMA (v_mfma_f64_16x16x4f64, layout as in the Gram
// kernel above).  A workgroup of four wavefronts owns 64 rows; each wavefront
// forms its 16 rows for all 16-column tiles of C.  B is staged in LDS 16 rows
// of k at a time (the next 16 rows and the next A fragments are in flight in
// registers during the MFMAs); A fragments come straight from global memory.
// A complex tile step is three real MFMAs (Gauss's product: ar br, ai bi and
// (ar + ai)(br + bi), combined at the end), issued phase by phase over the
// tiles so consecutive MFMAs never share an accumulator.
// gridDim.y = number of products.  A and C must not overlap.
constexpr int kTgKB = 16;

template <int NCB>
__global__ void __launch_bounds__(256, 2)
tall_gemm_kernel(long M, int k, int n, const cd *__restrict__ A0, long lda, cd *C0, const cd *__restrict__ A1,
                 cd *C1, long ldc, const cd *__restrict__ B, int ldb, int sub) {
    constexpr int KB = kTgKB, NC = 16 * NCB, BE = (KB * NC + 255) / 256;
    __shared__ cd sb[2][KB][NC + 1];
    const cd *A = blockIdx.y ? A1 : A0;
    cd *C = blockIdx.y ? C1 : C0;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long r0 = (long)blockIdx.x * 64 + 16 * w;
    const long ra = min(r0 + li, M - 1);      // rows past M read the last one (never written)
    // three real products per complex tile step (Gauss): t1 = ar br,
    // t2 = ai bi, t3 = (ar + ai)(br + bi); re = t1 - t2, im = t3 - t1 - t2
    d4 c1[NCB], c2[NCB], c3[NCB];
#pragma unroll
    for (int t = 0; t < NCB; ++t) {
        c1[t] = d4{0.0, 0.0, 0.0, 0.0};
        c2[t] = d4{0.0, 0.0, 0.0, 0.0};
        c3[t] = d4{0.0, 0.0, 0.0, 0.0};
    }
    const int nch = (k + KB - 1) / KB;
    cd af[KB / 4], an[KB / 4], bq[BE];
    auto load_a = [&](int ch, cd (&f)[KB / 4]) {
#pragma unroll
        for (int s = 0; s < KB / 4; ++s) {
            const int kk = ch * KB + 4 * s + lk;
            f[s] = kk < k ? A[ra + (size_t)kk * lda] : make_double2(0.0, 0.0);
        }
    };
    auto load_b = [&](int ch) {
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            const int idx = tid + 256 * e, kq = idx % KB, j = idx / KB, kk = ch * KB + kq;
            bq[e] = (idx < KB * NC && kk < k && j < n) ? B[kk + (size_t)j * ldb] : make_double2(0.0, 0.0);
        }
    };
    auto store_b = [&](int buf) {
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            const int idx = tid + 256 * e;
            if (idx < KB * NC) sb[buf][idx % KB][idx / KB] = bq[e];
        }
    };
    load_a(0, af);
    load_b(0);
    store_b(0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nch) {
            load_a(ch + 1, an);
            load_b(ch + 1);
        }
#pragma unroll
        for (int s = 0; s < KB / 4; ++s) {
            cd b[NCB];
#pragma unroll
            for (int t = 0; t < NCB; ++t) b[t] = sb[buf][4 * s + lk][16 * t + li];
            const double ar = af[s].x, ai = af[s].y, as = ar + ai;
#pragma unroll
            for (int t = 0; t < NCB; ++t) c1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b[t].x, c1[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NCB; ++t) c2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, b[t].y, c2[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NCB; ++t)
                c3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(as, b[t].x + b[t].y, c3[t], 0, 0, 0);
        }
        if (ch + 1 < nch) {
            store_b(buf ^ 1);
#pragma unroll
            for (int s = 0; s < KB / 4; ++s) af[s] = an[s];
        }
        __syncthreads();
    }
    // result register q of lane l: row 4 q + l / 16, column l % 16 of the tile
#pragma unroll
    for (int t = 0; t < NCB; ++t) {
        const int j = 16 * t + li;
        if (j >= n) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long r = r0 + 4 * q + lk;
            if (r >= M) continue;
            cd *cp = C + r + (size_t)j * ldc;
            const double re = c1[t][q] - c2[t][q], im = c3[t][q] - c1[t][q] - c2[t][q];
            if (sub) {
                const cd v = *cp;
                *cp = make_double2(v.x - re, v.y - im);
            } else {
                *cp = make_double2(re, im);
            }
        }
    }
}

// C0 (=|-=) A0 B and, when A1 is given, C1 (=|-=) A1 B: M x n outputs, A_s
// M x k (lda), C_s (ldc), B k x n (ldb).  Tall products with n <= 80 run on
// tall_gemm_kernel; wider or short ones on rocBLAS ZGEMM.
void gemm_pair(Ctx &c, long M, int n, int k, const cd *A0, const cd *A1, long lda, const cd *B, int ldb, cd *C0,
               cd *C1, long ldc, bool sub) {
    if (M <= 0 || n <= 0) return;
    if (n <= 80 && M >= 1024) {
        const dim3 g((unsigned)((M + 63) / 64), A1 ? 2 : 1);
        switch ((n + 15) / 16) {
#define WTB_TG_CASE(K)                                                                                          \
    case K:                                                                                                     \
        hipLaunchKernelGGL(tall_gemm_kernel<K>, g, dim3(256), 0, c.stream, M, k, n, A0, lda, C0, A1, C1, ldc, B, \
                           ldb, sub ? 1 : 0);                                                                   \
        break;
            WTB_TG_CASE(1) WTB_TG_CASE(2) WTB_TG_CASE(3) WTB_TG_CASE(4) WTB_TG_CASE(5)
#undef WTB_TG_CASE
        }
        WTB_LAUNCH_CHECK();
        return;
    }
    const rocblas_double_complex one = {1.0, 0.0}, mone = {-1.0, 0.0}, zero = {0.0, 0.0};
    for (int s = 0; s < (A1 ? 2 : 1); ++s)
        WTB_BLAS_CHECK(rocblas_zgemm(c.blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)M, n, k,
                                     sub ? &mone : &one, rb(s ? A1 : A0), (rocblas_int)lda, rb(B), ldb,
                                     sub ? &one : &zero, rb(s ? C1 : C0), (rocblas_int)ldc));
}

// C (M x n) = A (M x k) B (k x n)
void gemm_nn(Ctx &c, long M, int n, int k, const cd *A, long lda, const cd *B, int ldb, cd *C, long ldc) {
    gemm_pair(c, M, n, k, A, nullptr, lda, B, ldb, C, nullptr, ldc, false);
}

// C (M x n) -= A (M x k) B (k x n)
void gemm_nn_sub(Ctx &c, long M, int n, int k, const cd *A, long lda, const cd *B, int ldb, cd *C, long ldc) {
    gemm_pair(c, M, n, k, A, nullptr, lda, B, ldb, C, nullptr, ldc, true);
}

}  // namespace

void small_heev(Ctx &c, int m, cd *A, int lda, double *W, cd *Z, int ldz) {
    if (m <= 0) return;
    if (m > 2 * kJT) throw Error{WTB_ERR_UNSUPPORTED, "dense eigenproblem larger than 2048"};
    DBuf<cd> &V = c.ws_c;
    DBuf<int> &perm = c.ws_i;
    V.alloc((size_t)m * m);
    perm.alloc(m);
    hipLaunchKernelGGL(jacobi_kernel, dim3(1), dim3(kJT), 0, c.stream, m, A, lda, V.p, m, W, perm.p, 60);
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(gather_cols, dim3((m + 255) / 256, m), dim3(256), 0, c.stream, m, m, perm.p, V.p, m, Z, ldz);
    WTB_LAUNCH_CHECK();
}

namespace {

struct Timer {
    hipStream_t st;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    void begin() {
        hipEvent_t a, b;
        WTB_HIP_CHECK(hipEventCreate(&a));
        WTB_HIP_CHECK(hipEventCreate(&b));
        WTB_HIP_CHECK(hipEventRecord(a, st));
        ev.push_back({a, b});
    }
    void end() { WTB_HIP_CHECK(hipEventRecord(ev.back().second, st)); }
    double total() {
        double t = 0;
        for (auto &p : ev) t += elapsed_ms(p.first, p.second);
        return t;
    }
    ~Timer() {
        for (auto &p : ev) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
};

struct Lobpcg {
    Ctx &c;
    long N;
    int b, nev, m3;
    DBuf<cd> &S, &AS, &T1, &T2, &P, &AP;   // N x cols
    DBuf<cd> &G, &A, &U, &T, &Tmp, &Cm;  // small m3 x m3
    DBuf<double> &D, &lam, &lamr, &norm2, &sc;
    DBuf<int> &idx, &act;
    // matvec time (the only per-segment events on the stream: each event
    // marker between two launches idles the queue for a few microseconds)
    Timer tmv;
    hipEvent_t ev_chol = nullptr;      // the first CholQR factor's status has reached the host
    hipEvent_t ev_kept = nullptr;      // the projected W's column norms have reached the host
    std::vector<double> h_lam;

    Lobpcg(Ctx &cx, int nev_, int b_)
        : c(cx), N(cx.bse.N), b(b_), nev(nev_), m3(3 * b_), S(cx.lws.S), AS(cx.lws.AS), T1(cx.lws.T1), T2(cx.lws.T2),
          P(cx.lws.P), AP(cx.lws.AP), G(cx.lws.G), A(cx.lws.A), U(cx.lws.U), T(cx.lws.T), Tmp(cx.lws.Tmp),
          Cm(cx.lws.Cm), D(cx.lws.D), lam(cx.lws.lam), lamr(cx.lws.lamr), norm2(cx.lws.norm2), sc(cx.lws.sc),
          idx(cx.lws.idx), act(cx.lws.act) {
        c.lws.hd.alloc(std::max<size_t>((size_t)N, 2 * (size_t)m3));
        c.lws.hi.alloc((size_t)m3);
        c.lws.hflag.alloc(4);
        WTB_HIP_CHECK(hipEventCreateWithFlags(&ev_chol, hipEventDisableTiming));
        WTB_HIP_CHECK(hipEventCreateWithFlags(&ev_kept, hipEventDisableTiming));
        c.lws.hkept.alloc((size_t)m3);
        tmv.st = c.stream;
        S.alloc((size_t)N * m3);
        AS.alloc((size_t)N * m3);
        T1.alloc((size_t)N * m3);       // trade places with S and AS
        T2.alloc((size_t)N * m3);
        P.alloc((size_t)N * b);
        AP.alloc((size_t)N * b);
        G.alloc((size_t)m3 * m3);
        A.alloc((size_t)m3 * m3);
        U.alloc((size_t)m3 * m3);
        T.alloc((size_t)m3 * m3);
        Tmp.alloc((size_t)m3 * m3);
        Cm.alloc((size_t)m3 * m3);
        D.alloc(m3);
        lam.alloc(m3);
        lamr.alloc(m3);
        norm2.alloc(m3);
        sc.alloc(m3);
        idx.alloc(m3);
        act.alloc(m3);
    }

    ~Lobpcg() {
        if (ev_chol) (void)hipEventDestroy(ev_chol);
        if (ev_kept) (void)hipEventDestroy(ev_kept);
    }
    Lobpcg(const Lobpcg &) = delete;
    Lobpcg &operator=(const Lobpcg &) = delete;

    dim3 gN(int cols) const { return dim3((unsigned)((N + 255) / 256), (unsigned)cols); }

    void apply(int precision, int ncol, const cd *X, cd *Y) {
        tmv.begin();
        matvec(c, precision, ncol, X, Y);
        tmv.end();
    }

    // column norms^2 of (AX - lam X) or of AX when lam == nullptr; with lam,
    // the Ritz values of the last Rayleigh-Ritz step come back in the same
    // round trip (h_lam)
    std::vector<double> norms(int ncol, const cd *X, const cd *AX, const double *lamd) {
        hipLaunchKernelGGL(resid_norms, dim3(ncol), dim3(kNT), 0, c.stream, N, X, AX, lamd, norm2.p);
        WTB_LAUNCH_CHECK();
        double *hp = c.lws.hd.p;
        WTB_HIP_CHECK(hipMemcpyAsync(hp, norm2.p, sizeof(double) * ncol, hipMemcpyDeviceToHost, c.stream));
        if (lamd)
            WTB_HIP_CHECK(hipMemcpyAsync(hp + ncol, lamd, sizeof(double) * ncol, hipMemcpyDeviceToHost, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
        if (lamd) h_lam.assign(hp + ncol, hp + 2 * ncol);
        return std::vector<double>(hp, hp + ncol);
    }

    // B -= Q (Q^H B) for Q (N x q) with orthonormal columns, B (N x nb); the
    // same combination is applied to B2 -= Q2 (Q^H B) when Q2 is given (keeps
    // B2 = H B).  FP64 classical Gram-Schmidt.
    void project_out(int q, const cd *Q, const cd *Q2, int nb, cd *B, cd *B2) {
        gram(c, q, nb, N, Q, (int)N, B, (int)N, Tmp.p, m3);
        gemm_pair(c, N, nb, q, Q, Q2, N, Tmp.p, m3, B, B2, N, true);
    }

    // scale columns of B (and B2 with the same factors) to unit norm, on the device
    void normalize(int nb, cd *B, cd *B2) {
        hipLaunchKernelGGL(resid_norms, dim3(nb), dim3(kNT), 0, c.stream, N, nullptr, B, nullptr, norm2.p);
        if (B2) hipLaunchKernelGGL(unit_cols, gN(nb), dim3(256), 0, c.stream, N, norm2.p, B2);
        hipLaunchKernelGGL(unit_cols, gN(nb), dim3(256), 0, c.stream, N, norm2.p, B);
        WTB_LAUNCH_CHECK();
    }

    double drop_rel = 1e-12;   // SVQB cut relative to the largest Gram eigenvalue

    // Rayleigh-Ritz on S[:, :m] (first b columns become the new X, the
    // remainder combination the new P when withP).
    // B (N x nb) <- B R^-1 with G = B^H B = R^H R (and B2 <- B2 R^-1), two
    // passes (CholQR2).  False when a pivot falls below piv_rel: B is then
    // numerically rank deficient and left half-processed.
    // The first pass leaves B R^-1 orthonormal to about eps cond(G), and with
    // unit columns cond(G) follows the smallest pivot ratio rho (the squared
    // distance of a column from the span of the ones before it): about
    // 0.2 eps / rho^2 was measured (5e-12 at rho = 1.5e-3, 2e-10 at 3.6e-4).
    // With loose set (a tolerance of 1e-6 or above, or reduced-precision
    // products whose errors are far larger) the second pass is skipped when
    // rho >= one_pass_loose (orthonormal to ~2e-9 at worst), otherwise when
    // rho >= one_pass_tight.  The Rayleigh-Ritz step then sees residual
    // components of that size inside the basis, far below the tolerance.
    double piv_rel = 1e-12;
    double one_pass_loose = 1e-4, one_pass_tight = 0.05;
    bool verbose = false;
    bool cholqr(int nb, cd *B, cd *B2, double piv, bool loose = false) {
        const double one_pass = loose ? one_pass_loose : one_pass_tight;
        for (int pass = 0; pass < 2; ++pass) {
            gram(c, nb, nb, N, B, (int)N, B, (int)N, Tmp.p, m3, true);
            // the second pass factors I + O(eps cond^2) and cannot fail once the
            // first has passed: only the first pass's pivot test is read back
            double ratio = 0.0;
            int fail = cholesky_upper(c, nb, Tmp.p, m3, piv, pass == 0, &ratio);
            if (fail) {
                return false;
            }
            trsm_right_upper(c, N, nb, Tmp.p, m3, B, B2);
            if (pass == 0) {
                if (verbose) std::fprintf(stderr, "[wtb lobpcg] CholQR of %d columns: smallest pivot ratio %.3e\n", nb, ratio);
                if (ratio >= one_pass) {
                    if (verbose) {
                        // the orthonormality the single pass left
                        gram(c, nb, nb, N, B, (int)N, B, (int)N, Tmp.p, m3, true);
                        std::vector<cd> g((size_t)m3 * nb);
                        WTB_HIP_CHECK(hipMemcpyAsync(g.data(), Tmp.p, sizeof(cd) * g.size(), hipMemcpyDeviceToHost,
                                                     c.stream));
                        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
                        double e = 0;
                        for (int q = 0; q < nb; ++q)
                            for (int p = 0; p <= q; ++p) {
                                const cd v = g[p + (size_t)q * m3];
                                e = std::max(e, std::hypot(v.x - (p == q ? 1.0 : 0.0), v.y));
                            }
                        std::fprintf(stderr, "[wtb lobpcg] one CholQR pass: max |B^H B - I| = %.3e\n", e);
                    }
                    break;
                }
            }
        }
        return true;
    }

    // CholQR of W without a host round trip before its product: the first
    // pass (Gram, factor, W R^-1 unless the factor failed, when W stays as it
    // is) is queued here; its status reaches the pinned host words while the
    // product that follows runs, and cholqr_finish reads it.  Needs a factor
    // the triangular-solve kernel can stage (cholqr_async_ok).
    bool cholqr_async_ok(int nb) const { return (size_t)nb * (nb + 1) / 2 * sizeof(cd) + nb * sizeof(double) <= 150 * 1024; }
    void cholqr_begin(int nb, cd *B, double piv) {
        gram(c, nb, nb, N, B, (int)N, B, (int)N, Tmp.p, m3, true);
        const int *df = cholesky_upper_async(c, nb, Tmp.p, m3, piv, c.lws.hflag.p, ev_chol);
        trsm_right_upper(c, N, nb, Tmp.p, m3, B, nullptr, df);
    }
    // False when the first factor failed (B numerically rank deficient, B and
    // B2 untouched).  Otherwise the second pass, when the pivot ratio asks
    // for it, is applied to B and to its product B2 = H B alike (R^-1 of the
    // second pass is I + O(eps cond): the product's error is not amplified).
    bool cholqr_finish(int nb, cd *B, cd *B2, bool loose) {
        WTB_HIP_CHECK(hipEventSynchronize(ev_chol));
        const int *hf = c.lws.hflag.p;
        if (hf[0]) return false;
        double ratio;
        std::memcpy(&ratio, hf + 2, sizeof(double));
        const double one_pass = loose ? one_pass_loose : one_pass_tight;
        if (verbose) std::fprintf(stderr, "[wtb lobpcg] CholQR of %d columns: smallest pivot ratio %.3e\n", nb, ratio);
        if (ratio >= one_pass) {
            if (verbose) {
                // the orthonormality the single pass left
                gram(c, nb, nb, N, B, (int)N, B, (int)N, Tmp.p, m3, true);
                std::vector<cd> g((size_t)m3 * nb);
                WTB_HIP_CHECK(
                    hipMemcpyAsync(g.data(), Tmp.p, sizeof(cd) * g.size(), hipMemcpyDeviceToHost, c.stream));
                WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
                double e = 0;
                for (int q = 0; q < nb; ++q)
                    for (int p = 0; p <= q; ++p) {
                        const cd v = g[p + (size_t)q * m3];
                        e = std::max(e, std::hypot(v.x - (p == q ? 1.0 : 0.0), v.y));
                    }
                std::fprintf(stderr, "[wtb lobpcg] one CholQR pass: max |B^H B - I| = %.3e\n", e);
            }
            return true;
        }
        gram(c, nb, nb, N, B, (int)N, B, (int)N, Tmp.p, m3, true);
        cholesky_upper(c, nb, Tmp.p, m3, piv_rel, false);
        trsm_right_upper(c, N, nb, Tmp.p, m3, B, B2);
        return true;
    }

    // Rayleigh-Ritz on S[:, :m].  With an orthonormal basis the projected
    // problem is a standard Hermitian eigenproblem; otherwise (orth == false)
    // the Gram matrix is diagonalised first and nearly dependent directions
    // are dropped (SVQB).  The first b columns become the new X, the
    // remainder combination the new P when withP.
    // Ritz pairs of the m columns of S: the lowest nout (default b) replace
    // X = S[:, :nout] and AX (and P, AP when withP).
    void rayleigh_ritz(int m, bool withP, bool orth = true, int nout = 0) {
        const int nb = nout > 0 ? nout : b;
        if (withP && nb != b) throw Error{WTB_ERR_STATE, "partial Rayleigh-Ritz with P"};
        gram(c, m, m, N, S.p, (int)N, AS.p, (int)N, A.p, m3, true);
        if (orth) {
            heev_lowest(c, m, nb, A.p, m3, lamr.p, Cm.p, m3);
        } else {
            gram(c, m, m, N, S.p, (int)N, S.p, (int)N, G.p, m3, true);
            small_heev(c, m, G.p, m3, D.p, U.p, m3);
            std::vector<double> hd(m);
            WTB_HIP_CHECK(hipMemcpyAsync(hd.data(), D.p, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
            WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
            const double dmax = hd[m - 1];
            int off = 0;
            while (off < m && hd[off] <= drop_rel * dmax) ++off;
            const int r = m - off;
            if (r < nb) throw Error{WTB_ERR_NOCONV, "Rayleigh-Ritz basis lost rank"};
            hipLaunchKernelGGL(svqb_t, dim3((m + 63) / 64, r), dim3(64), 0, c.stream, m, r, off, U.p, m3, D.p, T.p,
                               m3);
            WTB_LAUNCH_CHECK();
            // A' = T^H A T  (r x r) into G
            gemm_nn(c, m, r, m, A.p, m3, T.p, m3, Tmp.p, m3);
            gram(c, r, r, m, T.p, m3, Tmp.p, m3, G.p, m3);
            hipLaunchKernelGGL(hermitize, dim3((r + 63) / 64, r), dim3(64), 0, c.stream, r, G.p, m3);
            WTB_LAUNCH_CHECK();
            heev_lowest(c, r, nb, G.p, m3, lamr.p, U.p, m3);
            // C = T Y[:, :b]  (m x b)
            gemm_nn(c, m, nb, r, T.p, m3, U.p, m3, Cm.p, m3);
        }
        WTB_HIP_CHECK(hipMemcpyAsync(lam.p, lamr.p, sizeof(double) * nb, hipMemcpyDeviceToDevice, c.stream));
        // new X, AX (and P, AP)
        // P and H P come straight out of the products (they read only S and AS);
        // the new X and H X are formed in T1 and T2, which then trade places
        // with S and AS: the rest of S is rebuilt every iteration.  A partial update
        // (nb < b) copies its columns back instead.
        if (withP && m > b)
            gemm_pair(c, N, b, m - b, S.p + (size_t)b * N, AS.p + (size_t)b * N, N, Cm.p + b, m3, P.p, AP.p, N,
                      false);
        gemm_pair(c, N, nb, m, S.p, AS.p, N, Cm.p, m3, T1.p, T2.p, N, false);
        if (nb == b) {
            swap_bufs(S, T1);
            swap_bufs(AS, T2);
        } else {
            WTB_HIP_CHECK(hipMemcpyAsync(S.p, T1.p, sizeof(cd) * N * nb, hipMemcpyDeviceToDevice, c.stream));
            WTB_HIP_CHECK(hipMemcpyAsync(AS.p, T2.p, sizeof(cd) * N * nb, hipMemcpyDeviceToDevice, c.stream));
        }
    }
};

}  // namespace

// Squared norm a unit column must keep through one Gram-Schmidt projection
// against X for that pass to be final (orthogonal to X to ~eps 1e4).
constexpr double kReorth = 1e-8;

void lobpcg(Ctx &c, int nev, double tol, int max_it, int precision, wtb_report *rep) {
    const long N = c.bse.N;
    if (nev < 1 || nev > N) throw Error{WTB_ERR_ARG, "nev out of range"};
    // block: nev plus guard vectors.  At nev = 64 a 72-column block takes as
    // many iterations as an 80-column one and keeps the Rayleigh-Ritz problem
    // at 216 instead of 240.
    // Guard columns beyond nev: 12.5 % (at least 8) up to nev = 64, 25 % above.
    // Large nev reaches into the dense continuum of unbound states (the
    // reference's MoS2 example at nev = 100: 11 iterations with 25 guard
    // columns, 107 with 12).  WTB_LOBPCG_GUARD overrides the count.
    int guard = nev <= 64 ? std::max(8, nev / 8) : nev / 4;
    if (const char *g = std::getenv("WTB_LOBPCG_GUARD")) guard = std::max(1, std::atoi(g));
    int b = (int)std::min<long>(N, nev + guard);
    const char *venv = std::getenv("WTB_VERBOSE");
    const bool verbose = venv && venv[0] && venv[0] != '0';
    if (N <= 1024 || 3L * b >= N) {
        // small problem: full dense diagonalisation (the reference's mkl_solver
        // path, module_bse.F90:2614-2642)
        hipEvent_t a0, a1;
        WTB_HIP_CHECK(hipEventCreate(&a0));
        WTB_HIP_CHECK(hipEventCreate(&a1));
        WTB_HIP_CHECK(hipEventRecord(a0, c.stream));
        c.ws_dense.alloc((size_t)N * N);
        DBuf<cd> Z;
        DBuf<double> W;
        Z.alloc((size_t)N * N);
        W.alloc(N);
        bse_block_to_c64(c, 0, N, 0, N, c.ws_dense.p, N);
        heev_lowest(c, (int)N, nev, c.ws_dense.p, (int)N, W.p, Z.p, (int)N);
        Solution &s = c.sol;
        s.evals.alloc(nev);
        s.evecs.alloc((size_t)N * nev);
        WTB_HIP_CHECK(hipMemcpyAsync(s.evals.p, W.p, sizeof(double) * nev, hipMemcpyDeviceToDevice, c.stream));
        WTB_HIP_CHECK(hipMemcpyAsync(s.evecs.p, Z.p, sizeof(cd) * N * nev, hipMemcpyDeviceToDevice, c.stream));
        WTB_HIP_CHECK(hipEventRecord(a1, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
        s.nev = nev;
        s.valid = true;
        if (rep) {
            *rep = wtb_report{};
            rep->nconv = nev;
            rep->t_solve_ms = elapsed_ms(a0, a1);
            rep->t_rr_ms = rep->t_solve_ms;
        }
        (void)hipEventDestroy(a0);
        (void)hipEventDestroy(a1);
        return;
    }
    // per-launch device time of the matvec kernels, for the report
    struct ClockGuard {
        KernelClock &k;
        ~ClockGuard() {
            k.on = false;
            k.reset();
        }
    } clock_guard{c.kclock};
    c.kclock.reset();
    c.kclock.on = rep != nullptr;
    hipEvent_t t0, t1, tr;
    WTB_HIP_CHECK(hipEventCreate(&t0));
    WTB_HIP_CHECK(hipEventCreate(&t1));
    WTB_HIP_CHECK(hipEventCreate(&tr));
    WTB_HIP_CHECK(hipEventRecord(t0, c.stream));
    Lobpcg L(c, nev, b);
    L.verbose = verbose;
    // WTB_CHOLQR_ONE_PASS overrides the pivot ratio above which CholQR stops
    // after one pass (above 1: always two passes)
    if (const char *e = std::getenv("WTB_CHOLQR_ONE_PASS")) L.one_pass_loose = L.one_pass_tight = std::atof(e);
    // pivot floor of W's CholQR factor (relative to the diagonal); a factor
    // below it sends the step to the SVQB Rayleigh-Ritz.  WTB_CHOLQR_PIV
    // raises it (tests of that fallback)
    double piv_w = 1e-12;
    if (const char *e = std::getenv("WTB_CHOLQR_PIV")) piv_w = std::atof(e);
    // squared norm below which a projected column takes the second
    // Gram-Schmidt pass (WTB_REORTH overrides it; above 1: always)
    double reorth = kReorth;
    if (const char *e = std::getenv("WTB_REORTH")) reorth = std::atof(e);
    // With FP32 products the projected matrix carries absolute errors of
    // ~1e-7 |H|; directions whose Gram eigenvalue is below 1e-6 of the largest
    // would amplify that noise beyond the spectral resolution, so they are
    // dropped from the basis.
    L.drop_rel = (precision == WTB_FP64) ? 1e-12 : 1e-4;
    // P block thresholds of the FP64 phase (the accurate phase after a
    // reduced-precision one sets its own; the reduced-precision phase has no P):
    // CholQR pivot floor and the smallest projected norm a P column keeps
    double piv_p = 1e-12;
    double p_keep = 1e-6;
    // initial block: unit vectors on the b smallest diagonal entries
    const double *d = c.lws.hd.p;
    WTB_HIP_CHECK(hipMemcpyAsync(c.lws.hd.p, c.bse.diag.p, sizeof(double) * N, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::partial_sort(order.begin(), order.begin() + b, order.end(),
                      [&](int x, int y) { return d[x] < d[y] || (d[x] == d[y] && x < y); });
    std::copy(order.begin(), order.begin() + b, c.lws.hi.p);
    WTB_HIP_CHECK(hipMemcpyAsync(L.idx.p, c.lws.hi.p, sizeof(int) * b, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(set_units, L.gN(b), dim3(256), 0, c.stream, N, b, L.idx.p, L.S.p);
    WTB_LAUNCH_CHECK();
    const int mv_prec = precision;   // FP64 / split-FP16 (mixed) / FP32 products
    long mvcols = 0;
    // the image of the unit starting block is read off the store (its columns)
    if (c.bse.precision == WTB_FP64)
        hipLaunchKernelGGL(unit_images<double>, L.gN(b), dim3(256), 0, c.stream, N, c.bse.ld, b, L.idx.p,
                           (const double *)c.bse.h64.p, (const double *)nullptr, L.AS.p);
    else
        hipLaunchKernelGGL(unit_images<float>, L.gN(b), dim3(256), 0, c.stream, N, c.bse.ld, b, L.idx.p,
                           (const float *)c.bse.h32.p, (const double *)c.bse.diag.p, L.AS.p);
    WTB_LAUNCH_CHECK();
    // The unit block starts as it is: its Rayleigh quotients are the diagonal
    // entries (ascending), and the first Rayleigh-Ritz step on [X W] mixes it.
    {
        std::vector<double> lam0(b);
        for (int j = 0; j < b; ++j) lam0[j] = d[order[j]];
        std::copy(lam0.begin(), lam0.end(), c.lws.hd.p);
        WTB_HIP_CHECK(hipMemcpyAsync(L.lam.p, c.lws.hd.p, sizeof(double) * b, hipMemcpyHostToDevice, c.stream));
    }
    // FP32 products cannot resolve residuals much below ~1e-6 relative.  The
    // reduced-precision phase stops at 5e-6 (or the caller's tol if larger):
    // a Ritz value with relative residual r is within r |l| of an eigenvalue
    // (within (r |l|)^2 / gap once it is isolated).  Its last step, or the
    // refinement below, then applies an accurate product to H X.  When the
    // caller's tol is below what that reaches, the iteration continues on the
    // nev wanted columns with the accurate product (int8 digits, FP64
    // recombination: ~3e-11 relative) until tol is met; every pair counts as
    // converged only at the caller's tol, in every precision.
    // The preconditioner (diag - l)^-1 is clamped at |diag - l| >= prec_floor
    // (eV).  Ritz values inside the band of transition energies (resonant
    // states) meet diagonal entries within meV; 0.1 eV keeps the correction
    // from collapsing onto those few entries (MoS2 at nev = 100: converged in
    // 11 iterations against stalling at 2e-8 with 0.01 eV), and leaves the
    // bound states of the benchmark workloads unchanged.  WTB_PREC_FLOOR
    // overrides it.
    double prec_floor = 1e-1;
    if (const char *f = std::getenv("WTB_PREC_FLOOR")) prec_floor = std::atof(f);
    bool haveP = false;
    int it = 0, nconv = 0;
    std::vector<double> rel(b);
    // The accurate product of the refinement (int8 digits recombined in FP64,
    // or FP64 MFMA arithmetic with WTB_REFINE=fp64).
    const char *renv = std::getenv("WTB_REFINE");
    const int acc = (renv && std::strcmp(renv, "fp64") == 0) ? WTB_FP64 : WTB_INT8X;
    // The accurate step that ends the reduced-precision phase: with a
    // tolerance of 1e-6 or above the int8 product's short form (10 instead of
    // 13 digit-pair groups; its dropped pairs add ~2^-28 max|H| max|X| per
    // term) certifies the residuals and fixes the energies to far below the
    // tolerance; tighter tolerances, and the accurate phase that may follow,
    // take the full product.
    const int acc_step = (acc == WTB_INT8X && tol >= 1e-6) ? kInt8Short : acc;
    // one LOBPCG phase with products of precision mvp, until the first nev
    // pairs have relative residual < tol_ph, a reduced-precision stall, or
    // the iteration budget.
    // useP: the previous step's direction P joins the basis.  With the
    // reduced-precision products it does not: H P would be carried by
    // combinations, and after its projection against [X W] every column
    // fell below the threshold that bounds the amplified product error
    // (on the bench workload none survived in any iteration), so the
    // phase runs as a block Davidson iteration on [X W].
    // With merge, the reduced-precision phase ends in an accurate step once at
    // most 80 - b columns are still active and every wanted residual is below
    // 1e-4: [X W] (b + active columns) takes one accurate product, which
    // replaces both the reduced-precision H W of that step and the separate
    // refinement of H X.  merged reports whether this happened.
    bool merged = false;
    auto run_phase = [&](int mvp, double tol_ph, bool stall_guard, bool merge, bool useP) {
        double prev_best = 1e300;
        int stall = 0;
        const int bb = L.b;
        for (; it < max_it; ++it) {
            std::vector<double> n2 = L.norms(bb, L.S.p, L.AS.p, L.lam.p);
            nconv = 0;
            std::vector<int> active;
            // the guard columns (j >= nev) get a search direction in the first
            // iteration only: afterwards they follow the Rayleigh-Ritz rotations
            // of the block, which keeps their part in the wanted pairs' convergence
            // (the same iteration counts on the bench workload) with fewer columns
            // in every product and a smaller Rayleigh-Ritz problem
            for (int j = 0; j < bb; ++j) {
                rel[j] = std::sqrt(n2[j]) / std::max(std::fabs(L.h_lam[j]), 1e-300);
                if (j < nev && rel[j] < tol_ph) ++nconv;
                if (rel[j] >= tol_ph && (j < nev || it == 0)) active.push_back(j);
            }
            double worst = 0;
            for (int j = 0; j < nev; ++j) worst = std::max(worst, rel[j]);
            if (verbose) {
                std::fprintf(stderr, "[wtb lobpcg] prec=%d it=%d m_active=%zu nconv=%d/%d max_rel_res=%.3e lam0=%.9f\n",
                             mvp, it, active.size(), nconv, nev, worst, L.h_lam[0]);
                std::fprintf(stderr, "[wtb lobpcg] active:");
                for (int j : active) std::fprintf(stderr, " %d", j);
                std::fprintf(stderr, "\n");
            }
            if (nconv == nev) break;
            // stagnation guard for products with a residual floor: stop when
            // three iterations near it gain less than 10 %
            if (stall_guard && worst < 1e-4 && worst > 0.9 * prev_best) {
                if (++stall >= 3) break;
            } else {
                stall = 0;
            }
            // restart without the P block when the residual jumps up tenfold
            if (worst > 10.0 * prev_best) haveP = false;
            prev_best = std::min(prev_best, worst);
            const int na = (int)active.size();
            // (the pinned index buffer is rewritten only after the next norms() sync)
            std::copy(active.begin(), active.end(), c.lws.hi.p);
            WTB_HIP_CHECK(hipMemcpyAsync(L.act.p, c.lws.hi.p, sizeof(int) * na, hipMemcpyHostToDevice, c.stream));
            // W = T R (unit scale first, normalised below)
            hipLaunchKernelGGL(fill_value, dim3(1), dim3(256), 0, c.stream, na, 1.0, L.sc.p);
            cd *W = L.S.p + (size_t)bb * N;
            cd *AW = L.AS.p + (size_t)bb * N;
            hipLaunchKernelGGL(precondition, L.gN(na), dim3(256), 0, c.stream, N, L.S.p, L.AS.p, L.lam.p,
                               c.bse.diag.p, L.act.p, L.sc.p, W, prec_floor);
            WTB_LAUNCH_CHECK();
            // W orthogonal to X (classical Gram-Schmidt in FP64), unit columns,
            // then its image H W.  One pass leaves X^H w at ~eps |w0| / |w1| for
            // a unit column w0 and its projection w1: the second pass runs only
            // when some column kept less than kReorth of its squared norm
            // (then ~1e-12 or worse; CGS2 brings it back to eps).
            // The projected W's column norms travel to the host behind the
            // projection and are read after the product: the rare second pass
            // then transforms W and H W alike (H X known, H W linear in W).
            L.normalize(na, W, nullptr);
            L.project_out(bb, L.S.p, nullptr, na, W, nullptr);
            hipLaunchKernelGGL(resid_norms, dim3(na), dim3(kNT), 0, c.stream, N, nullptr, W, nullptr, L.norm2.p);
            WTB_LAUNCH_CHECK();
            WTB_HIP_CHECK(hipMemcpyAsync(c.lws.hkept.p, L.norm2.p, sizeof(double) * na, hipMemcpyDeviceToHost,
                                         c.stream));
            WTB_HIP_CHECK(hipEventRecord(L.ev_kept, c.stream));
            hipLaunchKernelGGL(unit_cols, L.gN(na), dim3(256), 0, c.stream, N, L.norm2.p, W);
            WTB_LAUNCH_CHECK();
            // orthonormal W (CholQR2, or one pass when well conditioned); a
            // rank-deficient W falls back to SVQB.  H W is formed from the
            // once-orthonormalised W (no error amplification); the first pass's
            // status is read while the product runs, and a second pass, when
            // needed, transforms W and H W alike.
            const bool reduced = mvp != WTB_FP64 && mvp != WTB_INT8X && mvp != kInt8Short;
            const bool loose = reduced || tol_ph >= 1e-6;
            const bool async_qr = L.cholqr_async_ok(na);
            bool orth = true;
            if (async_qr) L.cholqr_begin(na, W, piv_w);
            else orth = L.cholqr(na, W, nullptr, piv_w, loose);
            const bool last = merge && bb + na <= 80 && worst < 1e-4;
            if (last) {
                L.apply(acc_step, bb + na, L.S.p, L.AS.p);   // X and W are contiguous in S
                mvcols += bb + na;
            } else {
                L.apply(mvp, na, W, AW);
                mvcols += na;
            }
            if (async_qr) orth = L.cholqr_finish(na, W, AW, loose);
            {
                WTB_HIP_CHECK(hipEventSynchronize(L.ev_kept));
                const double *pn = c.lws.hkept.p;
                const double kept = *std::min_element(pn, pn + na);
                if (verbose) std::fprintf(stderr, "[wtb lobpcg] projection kept a squared norm of %.3e at least\n", kept);
                if (kept < reorth) {
                    // second Gram-Schmidt pass on W and H W (H X = AS[:, :bb]), then
                    // their CholQR again
                    L.project_out(bb, L.S.p, L.AS.p, na, W, AW);
                    L.normalize(na, W, AW);
                    orth = L.cholqr(na, W, AW, piv_w, loose);
                }
            }
            int m = bb + na;
            if (haveP && useP) {
                // P orthogonal to X and W; H P follows the same combinations
                cd *Pd = L.S.p + (size_t)(bb + na) * N;
                cd *APd = L.AS.p + (size_t)(bb + na) * N;
                hipLaunchKernelGGL(gather_scale, L.gN(na), dim3(256), 0, c.stream, N, L.act.p, nullptr, L.P.p,
                                   L.AP.p, Pd, APd);
                WTB_LAUNCH_CHECK();
                L.normalize(na, Pd, APd);
                for (int pass = 0; pass < 2; ++pass) {
                    if (orth) {
                        // W = S[:, b:b+na] follows X with [X W] orthonormal: one
                        // projection against both (P and H P read once per pass)
                        L.project_out(bb + na, L.S.p, L.AS.p, na, Pd, APd);
                    } else {
                        L.project_out(bb, L.S.p, L.AS.p, na, Pd, APd);
                        L.project_out(na, W, AW, na, Pd, APd);
                    }
                }
                // H P is carried by linear combinations, not recomputed: a column
                // whose projection kept only a fraction f of its norm has its
                // product error scaled by 1/f on renormalisation.  Columns below
                // p_keep are dropped (the threshold follows the product precision).
                std::vector<double> pn = L.norms(na, nullptr, Pd, nullptr);
                int np = 0;
                for (int a = 0; a < na; ++a) {
                    if (std::sqrt(pn[a]) < p_keep) continue;
                    if (np != a) {
                        WTB_HIP_CHECK(hipMemcpyAsync(Pd + (size_t)np * N, Pd + (size_t)a * N, sizeof(cd) * N,
                                                     hipMemcpyDeviceToDevice, c.stream));
                        WTB_HIP_CHECK(hipMemcpyAsync(APd + (size_t)np * N, APd + (size_t)a * N, sizeof(cd) * N,
                                                     hipMemcpyDeviceToDevice, c.stream));
                    }
                    ++np;
                }
                if (np > 0) {
                    L.normalize(np, Pd, APd);
                    // a P block that has become dependent on [X W] is skipped this step
                    // H P carries product errors: R^-1 may amplify them only mildly
                    if (!orth || L.cholqr(np, Pd, APd, piv_p)) m += np;
                }
            }
            L.rayleigh_ritz(m, useP, orth);
            haveP = useP;
            if (last) {
                merged = true;
                ++it;
                break;
            }
        }
    };
    run_phase(mv_prec, precision == WTB_FP64 ? tol : std::max(tol, 5e-6), precision != WTB_FP64,
              precision != WTB_FP64, precision == WTB_FP64);
    WTB_HIP_CHECK(hipEventRecord(tr, c.stream));
    double max_rel = 0;
    auto final_check = [&]() {
        std::vector<double> n2 = L.norms(nev, L.S.p, L.AS.p, L.lam.p);
        nconv = 0;
        max_rel = 0;
        for (int j = 0; j < nev; ++j) {
            const double r = std::sqrt(n2[j]) / std::max(std::fabs(L.h_lam[j]), 1e-300);
            max_rel = std::max(max_rel, r);
            if (r < tol) ++nconv;
        }
    };
    if (precision != WTB_FP64) {
        // residual refinement: unless the last step already took the accurate
        // product, H X on the nev wanted columns with the accurate product,
        // then an FP64 Rayleigh-Ritz step on them
        if (!merged) {
            L.apply(acc_step, nev, L.S.p, L.AS.p);
            mvcols += nev;
            L.rayleigh_ritz(nev, false, true, nev);
        }
        final_check();
        if (nconv < nev && it < max_it) {
            // tol below what one refinement step reaches: iterate on the nev
            // columns with the accurate product (its product errors are
            // ~1e-10 of the reduced-precision ones: FP64-like thresholds)
            L.b = nev;
            haveP = false;
            L.drop_rel = 1e-10;
            piv_p = 1e-10;
            p_keep = 1e-5;
            ++it;
            run_phase(acc, tol, true, false, true);
            final_check();
        }
    } else {
        final_check();
    }
    WTB_HIP_CHECK(hipEventRecord(t1, c.stream));
    // results
    Solution &s = c.sol;
    s.evals.alloc(nev);
    s.evecs.alloc((size_t)N * nev);
    WTB_HIP_CHECK(hipMemcpyAsync(s.evals.p, L.lam.p, sizeof(double) * nev, hipMemcpyDeviceToDevice, c.stream));
    WTB_HIP_CHECK(hipMemcpyAsync(s.evecs.p, L.S.p, sizeof(cd) * N * nev, hipMemcpyDeviceToDevice, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    s.nev = nev;
    s.valid = true;
    if (rep) {
        rep->iterations = it;
        rep->nconv = nconv;
        rep->matvec_columns = mvcols;
        rep->max_rel_residual = max_rel;
        rep->t_solve_ms = elapsed_ms(t0, t1);
        rep->t_refine_ms = precision != WTB_FP64 ? elapsed_ms(tr, t1) : 0.0;
        rep->t_matvec_ms = L.tmv.total();
        rep->t_rr_ms = std::max(0.0, rep->t_solve_ms - rep->t_matvec_ms);
        const KernelClock &k = c.kclock;
        const KernelClock::Kind main = c.bse.precision == WTB_FP64 ? KernelClock::F64_ON_F64
                                       : precision == WTB_MIXED    ? KernelClock::F16SPLIT
                                       : precision == WTB_FP32     ? KernelClock::F32_ON_F32
                                                                   : KernelClock::F64_ON_F32;
        rep->t_kernel_ms = k.ms(main);
        rep->kernel_launches = (int)k.slot[main].ev.size();
        rep->kernel_bytes = k.slot[main].bytes;
        rep->kernel_flops = k.slot[main].flops;
        if (main != KernelClock::F64_ON_F32 && main != KernelClock::F64_ON_F64) {
            const KernelClock::Kind rk =
                k.slot[KernelClock::INT8X].ev.empty() ? KernelClock::F64_ON_F32 : KernelClock::INT8X;
            rep->t_refine_kernel_ms = k.ms(rk);
            rep->refine_launches = (int)k.slot[rk].ev.size();
            rep->refine_flops = k.slot[rk].flops;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(tr);
}

// Tall-skinny products for other files (the adjoint's linear solves).
void gram_tall(Ctx &c, int m, int n, long K, const cd *A, int lda, const cd *B, int ldb, cd *C, int ldc) {
    gram(c, m, n, K, A, lda, B, ldb, C, ldc);
}
void gemm_tall_sub(Ctx &c, long M, int n, int k, const cd *A, long lda, const cd *B, int ldb, cd *C, long ldc) {
    gemm_nn_sub(c, M, n, k, A, lda, B, ldb, C, ldc);
}
void gemm_tall(Ctx &c, long M, int n, int k, const cd *A, long lda, const cd *B, int ldb, cd *C, long ldc) {
    gemm_nn(c, M, n, k, A, lda, B, ldb, C, ldc);
}

}  // namespace wtb
