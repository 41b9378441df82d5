// Dense Hermitian eigensolver for the Rayleigh-Ritz step (FP64, n <= 1024).
//
//  1. Householder reduction to real symmetric tridiagonal form (unblocked,
//     lower, ZHETD2 semantics with ZLARFG reflectors), one 1024-thread
//     workgroup; the reflectors stay in the lower triangle of A.
//  2. Eigenvalues of the tridiagonal matrix by Sturm-count bisection, one
//     thread per eigenvalue; eigenvectors by inverse iteration with partial
//     pivoting, one thread per vector, then modified Gram-Schmidt inside
//     clusters of close eigenvalues (the scheme of LAPACK DSTEBZ/DSTEIN).
//  3. Back-transformation Z = H(0) H(1) ... H(n-2) Y, independent per
//     column block, spread over many workgroups.
// Only the lowest `nvec` eigenpairs are formed.
#include "wtb_internal.h"

#include <cmath>

namespace wtb {
namespace {

constexpr int kT = 1024;
// largest n whose Hermitian matrix (n^2 complex) is staged in LDS by the
// tridiagonalisation and the Cholesky factorisation (with their work arrays
// it stays under the 160 KB of a CU)
constexpr int kTridLds = 90;
constexpr int kCholLds = 96;
// Eigenvalues closer than this (relative to |T|) are treated as one cluster
// and their inverse-iteration vectors are re-orthogonalised.  Outside a
// cluster inverse iteration leaves an orthogonality error of about
// eps |T| / gap <= 2e-9.
constexpr double kClusterRel = 1e-7;

__device__ __forceinline__ double block_reduce(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kT / 64; ++i) s += red[i];
    return s;
}

// ---------------------------------------------------------------------------
// 1. tridiagonalisation
// A (n x n, column-major, lda): full Hermitian matrix, destroyed.  On exit
// A[r + i lda] (r >= i+2) holds v_i(r) of reflector i (v_i(i+1) = 1), tau[i]
// its scalar, d/e the real tridiagonal matrix T = Q^H A Q.
// L: the whole matrix is staged in LDS (n <= kTridLds), otherwise the
// trailing matrix is read from and written to global memory at every step.
template <bool L>
__global__ void __launch_bounds__(kT)
tridiag_kernel(int n, cd *__restrict__ Ag, int ldg, cd *__restrict__ tau, double *__restrict__ d,
               double *__restrict__ e) {
    extern __shared__ __align__(16) unsigned char smem[];
    cd *v = reinterpret_cast<cd *>(smem);     // n
    cd *y = v + n;                            // n
    cd *part = y + n;                         // kT partial sums
    cd *M = part + kT;                        // n x n when L
    __shared__ double red[kT / 64];
    __shared__ cd s_tau;
    const int tid = threadIdx.x;
    cd *A = L ? M : Ag;
    const int lda = L ? n : ldg;
    if (L) {
        for (int idx = tid; idx < n * n; idx += kT) M[idx] = Ag[(idx % n) + (size_t)(idx / n) * ldg];
        __syncthreads();
    }
    for (int i = 0; i < n - 1; ++i) {
        const int s = n - i - 1;                               // trailing size
        // column norm below the subdiagonal
        double xn2 = 0.0;
        for (int r = i + 2 + tid; r < n; r += kT) {
            cd a = A[r + (size_t)i * lda];
            xn2 += a.x * a.x + a.y * a.y;
        }
        xn2 = block_reduce(xn2, red);
        const cd alpha = A[(i + 1) + (size_t)i * lda];
        d[i] = A[i + (size_t)i * lda].x;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            // H(i) = I
            if (tid == 0) {
                e[i] = alpha.x;
                tau[i] = make_double2(0.0, 0.0);
            }
            __syncthreads();
            continue;
        }
        const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        const double beta = alpha.x >= 0.0 ? -nrm : nrm;
        const cd t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
        const double dd = (alpha.x - beta) * (alpha.x - beta) + alpha.y * alpha.y;
        const cd scal = make_double2((alpha.x - beta) / dd, -alpha.y / dd);   // 1/(alpha - beta)
        for (int r = i + 1 + tid; r < n; r += kT) {
            cd vr;
            if (r == i + 1) {
                vr = make_double2(1.0, 0.0);
            } else {
                cd a = A[r + (size_t)i * lda];
                vr = make_double2(a.x * scal.x - a.y * scal.y, a.x * scal.y + a.y * scal.x);
                A[r + (size_t)i * lda] = vr;      // keep the reflector
            }
            v[r] = vr;
        }
        if (tid == 0) {
            e[i] = beta;
            tau[i] = t;
            s_tau = t;
        }
        __syncthreads();
        // y = A_trail v  (rows i+1..n-1), threads split the columns in 4 groups
        {
            const int g = tid / 256, lr = tid % 256;
            for (int r0 = 0; r0 < s; r0 += 256) {
                const int r = i + 1 + r0 + lr;
                cd acc = make_double2(0.0, 0.0);
                if (r < n) {
                    for (int c = i + 1 + g; c < n; c += 4) {
                        cd a = A[r + (size_t)c * lda], vc = v[c];
                        acc.x += a.x * vc.x - a.y * vc.y;
                        acc.y += a.x * vc.y + a.y * vc.x;
                    }
                }
                part[tid] = acc;
                __syncthreads();
                if (g == 0 && r < n) {
                    cd sum = part[lr];
                    for (int q = 1; q < 4; ++q) {
                        sum.x += part[q * 256 + lr].x;
                        sum.y += part[q * 256 + lr].y;
                    }
                    y[r] = sum;
                }
                __syncthreads();
            }
        }
        // mu = Re(v^H y), w = tau y - |tau|^2 mu / 2 v
        double mu = 0.0;
        for (int r = i + 1 + tid; r < n; r += kT) mu += v[r].x * y[r].x + v[r].y * y[r].y;
        mu = block_reduce(mu, red);
        const double t2 = t.x * t.x + t.y * t.y;
        for (int r = i + 1 + tid; r < n; r += kT) {
            cd yr = y[r], vr = v[r];
            y[r] = make_double2(t.x * yr.x - t.y * yr.y - 0.5 * t2 * mu * vr.x,
                                t.x * yr.y + t.y * yr.x - 0.5 * t2 * mu * vr.y);
        }
        __syncthreads();
        // A_trail -= v w^H + w v^H
        for (int idx = tid; idx < s * s; idx += kT) {
            const int r = i + 1 + idx % s, c = i + 1 + idx / s;
            cd vr = v[r], wr = y[r], vc = v[c], wc = y[c];
            cd a = A[r + (size_t)c * lda];
            // v_r conj(w_c) + w_r conj(v_c)
            a.x -= (vr.x * wc.x + vr.y * wc.y) + (wr.x * vc.x + wr.y * vc.y);
            a.y -= (vr.y * wc.x - vr.x * wc.y) + (wr.y * vc.x - wr.x * vc.y);
            A[r + (size_t)c * lda] = a;
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[n - 1] = A[(n - 1) + (size_t)(n - 1) * lda].x;
        if (n >= 1) tau[n > 1 ? n - 1 : 0] = make_double2(0.0, 0.0);
    }
    if (L) {
        // the reflectors (strict lower part) are read by the back-transformation
        for (int idx = tid; idx < n * n; idx += kT) Ag[(idx % n) + (size_t)(idx / n) * ldg] = M[idx];
    }
}

// ---------------------------------------------------------------------------
// 2. tridiagonal eigenpairs
// Number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots).
// 1/q comes from the hardware reciprocal refined by two Newton steps, which
// keeps the pivots to within an ulp of the divided form at a fraction of the
// latency of an IEEE division on the sequential chain.
__device__ int sturm_count(int n, const double *d, const double *e2, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int j = 1; j < n; ++j) {
        double r = __builtin_amdgcn_rcp(q);
        r = fma(r, fma(-q, r, 1.0), r);
        r = fma(r, fma(-q, r, 1.0), r);
        q = fma(-e2[j - 1], r, d[j] - x);
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// Y (n x nvec, column-major, real) = eigenvectors of T for the nvec smallest
// eigenvalues W (ascending).  work: 6 n doubles per vector.
__global__ void __launch_bounds__(kT)
tridiag_eig_kernel(int n, int nvec, const double *__restrict__ d, const double *__restrict__ e,
                   double *__restrict__ W, double *__restrict__ Y, double *__restrict__ work) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *sd = reinterpret_cast<double *>(smem);   // n
    double *se = sd + n;                              // n (e, e[n-1] = 0)
    double *se2 = se + n;                             // n (e^2)
    double *sw = se2 + n;                             // n eigenvalues
    __shared__ double red[kT / 64];
    __shared__ double s_bounds[3];
    const int tid = threadIdx.x;
    for (int j = tid; j < n; j += kT) {
        sd[j] = d[j];
        double ej = (j < n - 1) ? e[j] : 0.0;
        se[j] = ej;
        se2[j] = ej * ej;
    }
    __syncthreads();
    if (tid == 0) {
        // Gershgorin interval
        double lo = 1e300, hi = -1e300, tn = 0.0;
        for (int j = 0; j < n; ++j) {
            double r = (j > 0 ? fabs(se[j - 1]) : 0.0) + (j < n - 1 ? fabs(se[j]) : 0.0);
            lo = fmin(lo, sd[j] - r);
            hi = fmax(hi, sd[j] + r);
            tn = fmax(tn, fabs(sd[j]) + r);
        }
        double pad = 2.2e-16 * tn * 4.0 * n + 1e-300;
        s_bounds[0] = lo - pad;
        s_bounds[1] = hi + pad;
        s_bounds[2] = tn;
    }
    __syncthreads();
    const double tnorm = s_bounds[2];
    const double pivmin = fmax(1e-300, 1e-300 * tnorm);
    // multisection: eigenvalue k (0-based) is bracketed by [lo, hi] with
    // count(lo) <= k < count(hi).  A group of G threads owns one eigenvalue and
    // evaluates the Sturm count at G interior points per round, so the
    // bracket shrinks (G + 1)-fold per round instead of 2-fold.
    {
        __shared__ double xs[kT];
        __shared__ int cs[kT];
        const int G = max(1, min(32, kT / max(nvec, 1)));
        const int per = kT / G;
        for (int k0 = 0; k0 < nvec; k0 += per) {
            const int grp = tid / G, g = tid % G, kk = k0 + grp;
            const bool own = grp < per && kk < nvec;
            double lo = s_bounds[0], hi = s_bounds[1];
            bool done = !own;
            while (__syncthreads_or(!done)) {
                double x = 0.0;
                int cnt = 0;
                if (!done) {
                    x = lo + (double)(g + 1) * ((hi - lo) / (double)(G + 1));
                    cnt = sturm_count(n, sd, se2, x, pivmin);
                }
                xs[tid] = x;
                cs[tid] = cnt;
                __syncthreads();
                if (!done) {
                    double nlo = lo, nhi = hi;
                    for (int t = 0; t < G; ++t) {
                        const double xt = xs[grp * G + t];
                        if (cs[grp * G + t] > kk) nhi = fmin(nhi, xt);
                        else nlo = fmax(nlo, xt);
                    }
                    if (nlo > nhi) nlo = nhi = 0.5 * (nlo + nhi);   // rounding at the last bits
                    done = (nhi - nlo <= 2.0 * 2.2e-16 * fmax(fabs(nlo), fabs(nhi)) + pivmin) ||
                           (nlo == lo && nhi == hi);
                    lo = nlo;
                    hi = nhi;
                }
            }
            if (own && g == 0) {
                sw[kk] = 0.5 * (lo + hi);
                W[kk] = sw[kk];
            }
        }
        __syncthreads();
    }
    // inverse iteration, one thread per eigenvector.  The per-vector work
    // arrays are interleaved across threads (element j of vector k at
    // [j * nvec + k]), so each step of the sequential sweeps is one coalesced
    // access for the whole wavefront.
    const double eps = 2.2e-16;
    for (int k = tid; k < nvec; k += kT) {
        const size_t sn = (size_t)nvec;
        double *__restrict__ u = work;                    // U diagonal
        double *__restrict__ u1 = u + (size_t)n * sn;     // first superdiagonal
        double *__restrict__ u2 = u1 + (size_t)n * sn;    // second superdiagonal
        double *__restrict__ l = u2 + (size_t)n * sn;     // multipliers
        double *__restrict__ pv = l + (size_t)n * sn;     // 1 where rows j, j+1 were swapped
        double *__restrict__ z = pv + (size_t)n * sn;     // iterate
        double lam = sw[k];
        // perturb a repeated eigenvalue slightly so that clustered vectors start apart
        double sep = eps * tnorm * 10.0;
        int kk = k;
        while (kk > 0 && fabs(sw[kk - 1] - sw[k]) < kClusterRel * tnorm) --kk;
        lam += (k - kk) * sep;
        // LU with partial pivoting of T - lam I (tridiagonal): rows may swap
        // with the next; the diagonal of U is stored as its reciprocal
        double a_cur = sd[0] - lam, b_cur = se[0];
        for (int j = 0; j < n - 1; ++j) {
            const size_t o = (size_t)j * sn + k;
            double c_j = se[j];                 // subdiagonal element below row j
            double a_nxt = sd[j + 1] - lam, b_nxt = se[j + 1];
            if (fabs(a_cur) >= fabs(c_j)) {
                if (a_cur == 0.0) a_cur = eps * tnorm + 1e-300;
                double m = c_j / a_cur;
                u[o] = 1.0 / a_cur; u1[o] = b_cur; u2[o] = 0.0;
                l[o] = m; pv[o] = 0.0;
                a_cur = a_nxt - m * b_cur;
                b_cur = b_nxt;
            } else {
                double m = a_cur / c_j;
                u[o] = 1.0 / c_j; u1[o] = a_nxt; u2[o] = b_nxt;
                l[o] = m; pv[o] = 1.0;
                a_cur = b_cur - m * a_nxt;
                b_cur = -m * b_nxt;
            }
        }
        if (a_cur == 0.0) a_cur = eps * tnorm + 1e-300;
        {
            const size_t o = (size_t)(n - 1) * sn + k;
            u[o] = 1.0 / a_cur; u1[o] = 0.0; u2[o] = 0.0;
        }
        for (int j = 0; j < n; ++j) z[(size_t)j * sn + k] = 1.0 + 0.0001 * ((j * 7919 + k * 104729) % 1000) / 1000.0;
        // The sweeps are sequential in j: each block of kB steps loads its
        // operands first, so one memory latency is paid per block.
        constexpr int kB = 8;
        for (int iter = 0; iter < 3; ++iter) {
            // forward: apply the row operations (L^-1 P) to z, carrying z[j] in a register
            double zj = z[k];
            for (int j0 = 0; j0 < n - 1; j0 += kB) {
                double lb[kB], pb[kB], zb[kB];
#pragma unroll
                for (int t = 0; t < kB; ++t) {
                    const int j = min(j0 + t, n - 2);
                    const size_t o = (size_t)j * sn + k;
                    lb[t] = l[o];
                    pb[t] = pv[o];
                    zb[t] = z[o + sn];
                }
#pragma unroll
                for (int t = 0; t < kB; ++t) {
                    if (j0 + t < n - 1) {
                        double zn = zb[t];
                        if (pb[t] != 0.0) {
                            const double tt = zj; zj = zn; zn = tt;
                        }
                        zn -= lb[t] * zj;
                        z[(size_t)(j0 + t) * sn + k] = zj;
                        zj = zn;
                    }
                }
            }
            // back substitution with U (upper, two superdiagonals)
            double z2 = 0.0, z1 = zj * u[(size_t)(n - 1) * sn + k];   // z[j+2], z[j+1]
            z[(size_t)(n - 1) * sn + k] = z1;
            double nn = z1 * z1;
            for (int j1 = n - 2; j1 >= 0; j1 -= kB) {
                double zb[kB], ub[kB], u1b[kB], u2b[kB];
#pragma unroll
                for (int t = 0; t < kB; ++t) {
                    const int j = max(j1 - t, 0);
                    const size_t o = (size_t)j * sn + k;
                    zb[t] = z[o];
                    ub[t] = u[o];
                    u1b[t] = u1[o];
                    u2b[t] = u2[o];
                }
#pragma unroll
                for (int t = 0; t < kB; ++t) {
                    if (j1 - t >= 0) {
                        const double zj0 = (zb[t] - u1b[t] * z1 - u2b[t] * z2) * ub[t];
                        z[(size_t)(j1 - t) * sn + k] = zj0;
                        nn += zj0 * zj0;
                        z2 = z1;
                        z1 = zj0;
                    }
                }
            }
            nn = 1.0 / sqrt(nn);
            for (int j = 0; j < n; ++j) z[(size_t)j * sn + k] *= nn;
        }
        double *yk = Y + (size_t)k * n;
        for (int j = 0; j < n; ++j) yk[j] = z[(size_t)j * sn + k];
    }
    __syncthreads();
    // modified Gram-Schmidt inside clusters (sequential over vectors, parallel
    // over the inner products)
    for (int k = 1; k < nvec; ++k) {
        int k0 = k;
        while (k0 > 0 && fabs(sw[k0 - 1] - sw[k0]) < kClusterRel * tnorm) --k0;
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

// Upper Cholesky factor of a Hermitian positive definite G (n x n, ldg), in
// place in the upper triangle: G = R^H R.  A pivot below rel * (original
// diagonal) stops the factorisation and sets *flag = column + 1.  L: G is
// staged in LDS for the factorisation (n <= kCholLds).
template <bool L>
__global__ void __launch_bounds__(kT)
cholesky_kernel(int n, cd *__restrict__ Gg, int ldgg, double rel, int *__restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *d0 = reinterpret_cast<double *>(smem);
    cd *M = reinterpret_cast<cd *>(d0 + ((n + 1) & ~1));
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    cd *G = L ? M : Gg;
    const int ldg = L ? n : ldgg;
    if (L)
        for (int idx = tid; idx < n * n; idx += kT) M[idx] = Gg[(idx % n) + (size_t)(idx / n) * ldgg];
    __syncthreads();
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
    if (L && !s_fail) {
        for (int idx = tid; idx < n * n; idx += kT) {
            const int p = idx % n, q = idx / n;
            if (p <= q) Gg[p + (size_t)q * ldgg] = M[idx];
        }
    }
}

}  // namespace

// Returns 0 on success, or 1 + the first failing column.
int cholesky_upper(Ctx &c, int n, cd *G, int ldg, double rel) {
    if (n > kT * 8) throw Error{WTB_ERR_UNSUPPORTED, "Cholesky size"};
    c.ws_flag.alloc(1);
    const size_t d0 = (size_t)((n + 1) & ~1) * sizeof(double);
    if (n <= kCholLds)
        hipLaunchKernelGGL(cholesky_kernel<true>, dim3(1), dim3(kT), d0 + (size_t)n * n * sizeof(cd), c.stream, n, G,
                           ldg, rel, c.ws_flag.p);
    else
        hipLaunchKernelGGL(cholesky_kernel<false>, dim3(1), dim3(kT), d0, c.stream, n, G, ldg, rel, c.ws_flag.p);
    WTB_LAUNCH_CHECK();
    int flag = 0;
    WTB_HIP_CHECK(hipMemcpyAsync(&flag, c.ws_flag.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    return flag;
}

// Lowest nvec eigenpairs of the Hermitian matrix A (n x n, lda; destroyed).
void heev_lowest(Ctx &c, int n, int nvec, cd *A, int lda, double *W, cd *Z, int ldz) {
    if (n <= 0 || nvec <= 0) return;
    if (n > kT) throw Error{WTB_ERR_UNSUPPORTED, "dense eigenproblem larger than 1024"};
    if (nvec > n) nvec = n;
    c.ws_heev_c.alloc((size_t)n);
    c.ws_heev_d.alloc((size_t)2 * n + (size_t)n * nvec + (size_t)6 * n * nvec);
    cd *tau = c.ws_heev_c.p;
    double *d = c.ws_heev_d.p, *e = d + n, *Y = e + n, *work = Y + (size_t)n * nvec;
    if (n <= kTridLds)
        hipLaunchKernelGGL(tridiag_kernel<true>, dim3(1), dim3(kT), (2 * (size_t)n + kT + (size_t)n * n) * sizeof(cd),
                           c.stream, n, A, lda, tau, d, e);
    else
        hipLaunchKernelGGL(tridiag_kernel<false>, dim3(1), dim3(kT), (2 * (size_t)n + kT) * sizeof(cd), c.stream, n, A, lda,
                       tau, d, e);
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(tridiag_eig_kernel, dim3(1), dim3(kT), 4 * (size_t)n * sizeof(double), c.stream, n, nvec,
                       d, e, W, Y, work);
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(backtransform_kernel, dim3((nvec + 3) / 4), dim3(kBT), 0, c.stream, n, nvec, A, lda, tau,
                       Y, Z, ldz);
    WTB_LAUNCH_CHECK();
}

}  // namespace wtb
