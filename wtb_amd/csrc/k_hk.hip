// Batched H(k) = sum_R e^{ik.R} H(R)/deg(R) - E_F build and Hermitian
// eigensolve for every k in one launch (replaces the eigsys k-loop of
// bsesolver, src/module_bse.F90:2230-2245 and src/module_hamiltonian.F90:200-320).
//
// One 64-lane workgroup (a single wavefront) per k-point, norb <= 32:
//   1. Fourier sum: each lane owns up to 16 matrix entries and accumulates
//      over R in the reference's order; phases are staged in LDS per chunk.
//   2. Parallel cyclic Jacobi (round-robin pairing, norb/2 rotations per step)
//      with the matrix and the eigenvector accumulator in LDS.
//   3. Optional gauge fix: each eigenvector is rotated so that Q^H z is real,
//      Q being the unblocked lower Householder reduction of H(k) (LAPACK
//      ZHETD2 with ZLARFG reflectors).  This reproduces the phases returned by
//      ZHEEV('V','L') up to a real sign.
//   4. nocp (first band above 0), scissor, band selection.
#include "wtb_internal.h"

namespace wtb {
namespace {

__device__ __forceinline__ cd cmul(cd a, cd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cconjmul(cd a, cd b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }  // conj(a) b
__device__ __forceinline__ cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cscale(double s, cd a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ cd cconj(cd a) { return make_double2(a.x, -a.y); }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Round-robin pairing (circle method) of m (even) indices; step t in [0, m-2].
__device__ __forceinline__ void rr_pair(int m, int t, int i, int &p, int &q) {
    int mm = m - 1;
    int x, y;
    if (i == 0) {
        x = t % mm;
        y = mm;
    } else {
        x = (t + i) % mm;
        y = (t - i + mm) % mm;
    }
    p = min(x, y);
    q = max(x, y);
}

constexpr int kLanes = 64;

__global__ void __launch_bounds__(kLanes)
hk_eig_kernel(int nk, int n, int nvec, int nv, int nc, int flags, double efermi, double scissor,
              const double *__restrict__ kpt, const cd *__restrict__ hop,
              const double *__restrict__ rcart, const double *__restrict__ invdeg,
              double *__restrict__ eall, cd *__restrict__ uall, int *__restrict__ nocp_out,
              double *__restrict__ esel, cd *__restrict__ usel, int *__restrict__ status) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    const int nn = n * n;
    cd *A = reinterpret_cast<cd *>(smem);            // [r*n + s]
    cd *V = A + nn;                                  // eigenvectors, [o*n + j]
    cd *H0 = V + nn;                                 // copy of H(k) for the gauge fix
    cd *ph = H0 + nn;                                // kLanes phases
    cd *vb = ph + kLanes;                            // 32
    cd *yb = vb + kMaxOrb;                           // 32
    cd *rph = yb + kMaxOrb;                          // 16 rotation phases
    double *rc = reinterpret_cast<double *>(rph + kMaxOrb / 2);
    double *rs = rc + kMaxOrb / 2;
    double *wv = rs + kMaxOrb / 2;                   // 32 eigenvalues
    int *rp = reinterpret_cast<int *>(wv + kMaxOrb);
    int *rq = rp + kMaxOrb / 2;
    int *perm = rq + kMaxOrb / 2;

    const bool compat = flags & WTB_FLAG_COMPAT_F32;
    const double kx = kpt[3 * k], ky = kpt[3 * k + 1], kz = kpt[3 * k + 2];

    // ---- 1. Fourier sum ----
    cd acc[kMaxOrb * kMaxOrb / kLanes];
#pragma unroll
    for (int s = 0; s < kMaxOrb * kMaxOrb / kLanes; ++s) acc[s] = make_double2(0.0, 0.0);
    for (int r0 = 0; r0 < nvec; r0 += kLanes) {
        int r = r0 + tid;
        if (r < nvec) {
            double x = kdotr(kx, ky, kz, rcart + 3 * r);
            if (compat) x = (double)(float)x;
            double sn, cs;
            sincos(x, &sn, &cs);
            ph[tid] = make_double2(cs * invdeg[r], sn * invdeg[r]);
        }
        __syncthreads();
        int rend = min(kLanes, nvec - r0);
        for (int rr = 0; rr < rend; ++rr) {
            cd e = ph[rr];
            const cd *hr = hop + (size_t)(r0 + rr) * nn;
#pragma unroll
            for (int s = 0; s < kMaxOrb * kMaxOrb / kLanes; ++s) {
                int idx = tid + s * kLanes;
                if (idx < nn) acc[s] = cadd(acc[s], cmul(e, hr[idx]));
            }
        }
        __syncthreads();
    }
    const double ef = compat ? (double)(float)efermi : efermi;
#pragma unroll
    for (int s = 0; s < kMaxOrb * kMaxOrb / kLanes; ++s) {
        int idx = tid + s * kLanes;
        if (idx < nn) {
            cd h = acc[s];
            if (idx / n == idx % n) h.x -= ef;
            A[idx] = h;
            H0[idx] = h;
            V[idx] = make_double2((idx / n == idx % n) ? 1.0 : 0.0, 0.0);
        }
    }
    __syncthreads();

    // ---- 2. parallel cyclic Jacobi ----
    const int m = (n + 1) & ~1;
    const int npair = m / 2;
    double fro = 0.0;
    for (int idx = tid; idx < nn; idx += kLanes) fro += A[idx].x * A[idx].x + A[idx].y * A[idx].y;
    fro = wave_sum(fro);
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0.0;
        for (int idx = tid; idx < nn; idx += kLanes)
            if (idx / n != idx % n) off += A[idx].x * A[idx].x + A[idx].y * A[idx].y;
        off = wave_sum(off);
        if (off <= 1e-30 * fro || off < 1e-300) break;
        for (int t = 0; t < m - 1; ++t) {
            if (tid < npair) {
                int p, q;
                rr_pair(m, t, tid, p, q);
                double c = 1.0, s = 0.0;
                cd e = make_double2(1.0, 0.0);
                if (q < n) {
                    cd b = A[p * n + q];
                    double ab = sqrt(b.x * b.x + b.y * b.y);
                    if (ab > 0.0) {
                        double theta = (A[q * n + q].x - A[p * n + p].x) / (2.0 * ab);
                        double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                        e = make_double2(b.x / ab, b.y / ab);
                    }
                }
                rp[tid] = p; rq[tid] = q; rc[tid] = c; rs[tid] = s; rph[tid] = e;
            }
            __syncthreads();
            // column pass: A <- A J, V <- V J
            for (int task = tid; task < n * npair; task += kLanes) {
                int r = task % n, a = task / n;
                int p = rp[a], q = rq[a];
                if (q >= n) continue;
                double c = rc[a], s = rs[a];
                cd ec = cconj(rph[a]);
                cd x = A[r * n + p], y = A[r * n + q];
                cd ey = cmul(ec, y);
                A[r * n + p] = csub(cscale(c, x), cscale(s, ey));
                A[r * n + q] = cadd(cscale(s, x), cscale(c, ey));
                x = V[r * n + p]; y = V[r * n + q];
                ey = cmul(ec, y);
                V[r * n + p] = csub(cscale(c, x), cscale(s, ey));
                V[r * n + q] = cadd(cscale(s, x), cscale(c, ey));
            }
            __syncthreads();
            // row pass: A <- J^H A
            for (int task = tid; task < n * npair; task += kLanes) {
                int sc = task % n, a = task / n;
                int p = rp[a], q = rq[a];
                if (q >= n) continue;
                double c = rc[a], s = rs[a];
                cd e = rph[a];
                cd x = A[p * n + sc], y = A[q * n + sc];
                cd ey = cmul(e, y);
                cd np = csub(cscale(c, x), cscale(s, ey));
                cd nq = cadd(cscale(s, x), cscale(c, ey));
                if (sc == p) { np.y = 0.0; nq = make_double2(0.0, 0.0); }
                if (sc == q) { np = make_double2(0.0, 0.0); nq.y = 0.0; }
                A[p * n + sc] = np;
                A[q * n + sc] = nq;
            }
            __syncthreads();
        }
    }

    // ---- sort ascending ----
    if (tid < n) wv[tid] = A[tid * n + tid].x;
    __syncthreads();
    if (tid < n) {
        double w = wv[tid];
        int rank = 0;
        for (int i = 0; i < n; ++i) {
            double u = wv[i];
            rank += (u < w) || (u == w && i < tid);
        }
        perm[rank] = tid;
    }
    __syncthreads();

    // ---- 3. LAPACK phase gauge ----
    if (flags & WTB_FLAG_GAUGE_LAPACK) {
        cd *Q = A;  // reuse
        for (int idx = tid; idx < nn; idx += kLanes)
            Q[idx] = make_double2((idx / n == idx % n) ? 1.0 : 0.0, 0.0);
        __syncthreads();
        for (int i = 0; i < n - 1; ++i) {
            double xn2 = 0.0;
            if (tid >= i + 2 && tid < n) {
                cd h = H0[tid * n + i];
                xn2 = h.x * h.x + h.y * h.y;
            }
            xn2 = wave_sum(xn2);
            cd alpha = H0[(i + 1) * n + i];
            if (xn2 == 0.0 && alpha.y == 0.0) {
                __syncthreads();
                continue;
            }
            double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            double beta = alpha.x >= 0.0 ? -nrm : nrm;
            cd tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            cd den = make_double2(alpha.x - beta, alpha.y);          // 1/(alpha - beta)
            double dd = den.x * den.x + den.y * den.y;
            cd scal = make_double2(den.x / dd, -den.y / dd);
            if (tid < n) {
                cd v = make_double2(0.0, 0.0);
                if (tid == i + 1) v = make_double2(1.0, 0.0);
                else if (tid > i + 1) v = cmul(H0[tid * n + i], scal);
                vb[tid] = v;
            }
            __syncthreads();
            if (tid < n) {
                cd y = make_double2(0.0, 0.0);
                for (int c = i + 1; c < n; ++c) y = cadd(y, cmul(H0[tid * n + c], vb[c]));
                yb[tid] = y;
            }
            __syncthreads();
            double mu = 0.0;
            if (tid > i && tid < n) mu = vb[tid].x * yb[tid].x + vb[tid].y * yb[tid].y;
            mu = wave_sum(mu);
            double t2 = tau.x * tau.x + tau.y * tau.y;
            for (int idx = tid; idx < nn; idx += kLanes) {
                int r = idx / n, c = idx % n;
                cd vr = vb[r], vc = vb[c], yr = yb[r], yc = yb[c];
                // - tau y_r conj(v_c) - conj(tau) v_r conj(y_c) + |tau|^2 mu v_r conj(v_c)
                cd t1 = cmul(tau, cmul(yr, cconj(vc)));
                cd t2c = cmul(cconj(tau), cmul(vr, cconj(yc)));
                cd t3 = cscale(t2 * mu, cmul(vr, cconj(vc)));
                H0[idx] = cadd(csub(csub(H0[idx], t1), t2c), t3);
            }
            if (tid < n) {
                cd s = make_double2(0.0, 0.0);
                for (int c = i + 1; c < n; ++c) s = cadd(s, cmul(Q[tid * n + c], vb[c]));
                cd ts = cmul(tau, s);
                for (int c = i + 1; c < n; ++c) Q[tid * n + c] = csub(Q[tid * n + c], cmul(ts, cconj(vb[c])));
            }
            __syncthreads();
        }
        // rotate eigenvector columns (thread j owns column j)
        if (tid < n) {
            double best = -1.0;
            cd bph = make_double2(1.0, 0.0);
            for (int mm2 = 0; mm2 < n; ++mm2) {
                cd s = make_double2(0.0, 0.0);
                for (int r = 0; r < n; ++r) s = cadd(s, cconjmul(Q[r * n + mm2], V[r * n + tid]));
                double a2 = s.x * s.x + s.y * s.y;
                if (a2 > best) { best = a2; bph = s; }
            }
            double ab = sqrt(best);
            cd f = make_double2(bph.x / ab, -bph.y / ab);
            for (int r = 0; r < n; ++r) V[r * n + tid] = cmul(V[r * n + tid], f);
        }
        __syncthreads();
    }

    // ---- 4. nocp, scissor, outputs ----
    __shared__ int s_nocp;
    if (tid == 0) {
        int nocp = -1;
        for (int j = 0; j < n; ++j)
            if (wv[perm[j]] > 0.0) { nocp = j; break; }
        s_nocp = nocp;
        nocp_out[k] = nocp;
        if (nocp < 0) atomicOr(status, 1);
        else if (nv > nocp || nc > n - nocp) atomicOr(status, 2);
    }
    __syncthreads();
    const int nocp = s_nocp;
    const int nb = nv + nc;
    if (tid < n) {
        double e = wv[perm[tid]];
        if (nocp >= 0 && tid >= nocp) e += scissor;
        eall[(size_t)k * n + tid] = e;
        int j = tid - (nocp - nv);
        if (nocp >= 0 && j >= 0 && j < nb) esel[(size_t)k * nb + j] = e;
    }
    for (int idx = tid; idx < nn; idx += kLanes) {
        int o = idx / n, j = idx % n;
        cd z = V[o * n + perm[j]];
        uall[(size_t)k * nn + idx] = z;
        int js = j - (nocp - nv);
        if (nocp >= 0 && js >= 0 && js < nb) usel[((size_t)o * nb + js) * nk + k] = z;
    }
}

__global__ void to_f32_kernel(const cd *__restrict__ in, cf *__restrict__ out, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

}  // namespace

void launch_hk_eig(Ctx &c, int nv, int nc) {
    Model &m = c.model;
    Bands &b = c.bands;
    if (m.norb <= 0) throw Error{WTB_ERR_STATE, "no TB model set"};
    if (c.nk <= 0) throw Error{WTB_ERR_STATE, "no k-points set"};
    if (m.norb > kMaxOrb) throw Error{WTB_ERR_UNSUPPORTED, "norb > 32 not supported by the batched eigensolver"};
    if (nv < 1 || nc < 1 || nv + nc > m.norb) throw Error{WTB_ERR_ARG, "invalid nv/nc"};
    const int n = m.norb, nk = c.nk, nb = nv + nc;
    b.valid = false;
    b.nk = nk; b.nv = nv; b.nc = nc; b.nb = nb;
    b.eall.alloc((size_t)nk * n);
    b.uall.alloc((size_t)nk * n * n);
    b.nocp.alloc(nk);
    b.esel.alloc((size_t)nk * nb);
    b.usel.alloc((size_t)nk * nb * n);
    b.usel32.alloc((size_t)nk * nb * n);
    b.status.alloc(1);
    WTB_HIP_CHECK(hipMemsetAsync(b.status.p, 0, sizeof(int), c.stream));
    size_t shmem = 3 * (size_t)n * n * sizeof(cd) + (kLanes + 2 * kMaxOrb + kMaxOrb / 2) * sizeof(cd) +
                   (kMaxOrb + kMaxOrb) * sizeof(double) + (kMaxOrb + kMaxOrb) * sizeof(int);
    hipLaunchKernelGGL(hk_eig_kernel, dim3(nk), dim3(kLanes), shmem, c.stream, nk, n, m.nvec, nv, nc,
                       m.flags, m.efermi, m.scissor, c.kpt.p, m.hop.p, m.rcart.p, m.invdeg.p,
                       b.eall.p, b.uall.p, b.nocp.p, b.esel.p, b.usel.p, b.status.p);
    WTB_LAUNCH_CHECK();
    size_t tot = (size_t)nk * nb * n;
    hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream,
                       b.usel.p, b.usel32.p, tot);
    WTB_LAUNCH_CHECK();
    int st = 0;
    WTB_HIP_CHECK(hipMemcpyAsync(&st, b.status.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (st & 1) throw Error{WTB_ERR_NOCP, "no eigenvalue above the Fermi level at some k (nocp undefined)"};
    if (st & 2) throw Error{WTB_ERR_NOCP, "Number of valence/conduction states higher than occupied/unoccupied states"};
    b.valid = true;
}

}  // namespace wtb
