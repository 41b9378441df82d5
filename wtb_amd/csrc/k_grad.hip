// Adjoint of the exciton energies: gradients of L = sum_n w_n E_n over the
// states of the last solve with respect to the Coulomb parameters (r0,
// EDIEL_T, EDIEL, EDIEL_B, EDIEL_Z, W_COUL, LC), the scissor shift, E_F and
// every H(R) entry.  The reference has no counterpart.
//
// Hellmann-Feynman: dE_n = x_n^H dH x_n.  In block form the (k,k') block of H
// is V(k-k') (M^c (x) M^v) with M^c = U_c(k)^H U_c(k'), M^v likewise (conjugated
// in the textbook convention), so
//   dL = sum_{k,k'} dV(k-k') s(k,k')
//      + 2 Re sum_{k,j} dU_j(k)^H Ubar_j(k) + sum_{k,j} Ebar_j(k) dE_j(k)
// with s(k,k') = sum_n w_n x_n(k)^H (M^c (x) M^v) x_n(k').  Nothing of size N^2
// is formed: a pair kernel rebuilds the overlaps and V (with exact parameter
// derivatives from forward-mode dual numbers) per (k,k') tile and contracts
// them with the exciton vectors into the scalar s and the small blocks
//   G^c(k,k')[c,c'] = V sum_n w_n sum_{vv'} conj(x_n[kcv]) M^v[v,v'] x_n[k'c'v']
//   G^v(k,k')[v,v'] = V sum_n w_n sum_{cc'} conj(x_n[kcv]) M^c[c,c'] x_n[k'c'v'],
// which rocBLAS ZGEMMs fold into Ubar_c(k) = sum_{k'c'} G^c(k,k')[c,c'] U_c'(k').
// Per k the TB eigenvector adjoint
//   Hbar(k) = Herm( U D U^H ),  D[m][m] = Ebar_m,  D[m][l] = 2 conj(u_l^H ubar_m) / (l_m - l_l)
// (pairs closer than a degeneracy tolerance are skipped: L is invariant under
// rotations inside a degenerate pair of selected bands) and the adjoint of the
// Fourier sum g(R) = sum_k Hbar(k) e^{-i k.R} / deg(R) give dL/dH(R).
#include "coulomb.h"
#include "wtb_internal.h"

#include <cmath>
#include <cstring>

namespace wtb {
namespace {

constexpr int kNP = 7;   // r0, EDIEL_T, EDIEL, EDIEL_B, EDIEL_Z, W_COUL, LC

// forward-mode dual number over the Coulomb parameters
struct Dual {
    double v;
    double d[kNP];
};
__device__ inline Dual dc(double x) {
    Dual r;
    r.v = x;
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = 0.0;
    return r;
}
__device__ inline Dual dvar(double x, int i) {
    Dual r = dc(x);
    r.d[i] = 1.0;
    return r;
}
__device__ inline Dual operator+(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
__device__ inline Dual operator-(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
__device__ inline Dual operator*(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
__device__ inline Dual operator/(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v / b.v;
    const double ib2 = 1.0 / (b.v * b.v);
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) * ib2;
    return r;
}
__device__ inline Dual operator*(double s, const Dual &a) {
    Dual r;
    r.v = s * a.v;
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = s * a.d[i];
    return r;
}
__device__ inline Dual operator+(double s, const Dual &a) {
    Dual r = a;
    r.v += s;
    return r;
}
__device__ inline Dual operator-(double s, const Dual &a) { return s + (-1.0 * a); }
__device__ inline Dual dexp(const Dual &a) {
    Dual r;
    r.v = exp(a.v);
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = r.v * a.d[i];
    return r;
}
__device__ inline Dual dsqrt(const Dual &a) {
    Dual r;
    r.v = sqrt(a.v);
    const double h = 0.5 / r.v;
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = h * a.d[i];
    return r;
}
__device__ inline Dual dcos(const Dual &a) {
    Dual r;
    r.v = cos(a.v);
    const double s = -sin(a.v);
#pragma unroll
    for (int i = 0; i < kNP; ++i) r.d[i] = s * a.d[i];
    return r;
}

// V(k - k') and its derivatives with respect to the raw parameters, the same
// formulas as coulomb_v with the parameter-dependent constants recomputed
// (module_coulomb.F90:8-387).  Potentials without screening parameters
// (V3D, V2DT, V0DT) have zero derivatives.
__device__ Dual coulomb_dual(const CoulombDev &c, double dx, double dy, double dz) {
    const double q = sqrt(dx * dx + dy * dy + dz * dz);
    const Dual r0 = dvar(c.r0, 0), et = dvar(c.et, 1), epar = dvar(c.epar, 2), eb = dvar(c.eb, 3);
    const Dual ez = dvar(c.ez, 4), w = dvar(c.w, 5), lc = dvar(c.lc, 6);
    switch (c.kind) {
    case WTB_V2DK: {
        const Dual r0v = ((epar - dc(1.0)) * lc) / (et + eb);
        const Dual ed = 0.5 * (et + eb);
        const Dual pref = (c.vbz2d * c.cic) * (dc(1.0) / ed);
        if (q < c.ktol) {
            const Dual auxi = (2.0 * c.pi) * r0v * dc(1.0 / c.v2dk_a0sg);
            return (c.v2dk_a0sg / (2.0 * c.pi)) * (pref * ((double)1.76f + auxi));
        }
        return pref / (q * (1.0 + q * r0v));
    }
    case WTB_V3DL:
        if (q < c.ktol) return dc(0.0);
        return (c.vbz3d * c.cic / (q * q)) * (dc(1.0) / epar);
    case WTB_V2DT2: {
        if (q < c.ktol) return dc(0.0);
        const double qxy = sqrt(dx * dx + dy * dy);
        return ((c.vbz3d * c.cic) / (q * q)) * (1.0 - dexp(-0.5 * qxy * lc) * dcos(0.5 * dz * lc));
    }
    case WTB_V2DOH:
        if (q < c.ktol) return dc(0.0);
        return (c.vbz2d * c.cic / q) * (dexp(-q * w) / ez);
    case WTB_V2DRK: {
        if (q < c.ktol) return dc(0.0);
        const Dual eta = dsqrt(epar / ez), kappa = dsqrt(epar * ez);
        const Dual pb = (eb - kappa) / (eb + kappa), pt = (et - kappa) / (et + kappa);
        const Dual a1 = (1.0 - pb * pt * dexp(-2.0 * q * eta * lc)) * kappa;
        const Dual e1 = dexp(-q * eta * lc);
        const Dual a2 = (1.0 - pt * e1) * (1.0 - pb * e1);
        const Dual ew = a1 / a2 + q * r0 * dexp(-q * w);
        return ((c.vbz2d * c.cic) / q) * (dexp(-q * w) / ew);
    }
    default:
        return dc(coulomb_v(c, dx, dy, dz));
    }
}

__device__ inline cd cmul(cd a, cd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline cd cmulc(cd a, cd b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }  // conj(a) b
__device__ inline cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline cd cscale(double s, cd a) { return make_double2(s * a.x, s * a.y); }

constexpr int kPT = 256;
constexpr int kMaxStates = 64;

// One workgroup per column k' of the pair grid, threads over the rows k.
// G: (NC*NC + NV*NV) matrices nk x nk column-major, matrix p at p * nk^2,
// element (k, k') at k + k' nk.
template <int NC, int NV>
__global__ void __launch_bounds__(kPT)
grad_pair_kernel(int nk, int norb, int nst, int textbook, int accumulate, CoulombDev coul,
                 const double *__restrict__ kpt, const cd *__restrict__ usel, const cd *__restrict__ X, long N,
                 const int *__restrict__ sidx, const double *__restrict__ sw, cd *__restrict__ G,
                 double *__restrict__ dcoul) {
    constexpr int NB = NC + NV, NCV = NC * NV;
    extern __shared__ __align__(16) unsigned char smem[];
    cd *ub = reinterpret_cast<cd *>(smem);           // U(k')[o][j]
    cd *xb = ub + (size_t)norb * NB;                 // x_s(k')[i]
    __shared__ double red[kNP][kPT / 64];
    const int kp = blockIdx.x, tid = threadIdx.x;
    for (int idx = tid; idx < norb * NB; idx += kPT) {
        const int o = idx / NB, j = idx % NB;
        ub[idx] = usel[((size_t)o * NB + j) * nk + kp];
    }
    for (int idx = tid; idx < nst * NCV; idx += kPT) {
        const int s = idx / NCV, i = idx % NCV;
        xb[idx] = X[(size_t)kp * NCV + i + (size_t)sidx[s] * N];
    }
    __syncthreads();
    double dacc[kNP];
#pragma unroll
    for (int p = 0; p < kNP; ++p) dacc[p] = 0.0;
    const size_t nk2 = (size_t)nk * nk;
    for (int k = tid; k < nk; k += kPT) {
        cd mc[NC][NC], mv[NV][NV];
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) mc[a][b] = make_double2(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < NV; ++b) mv[a][b] = make_double2(0.0, 0.0);
        for (int o = 0; o < norb; ++o) {
            cd uk[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) uk[j] = usel[((size_t)o * NB + j) * nk + k];
            const cd *up = ub + (size_t)o * NB;
#pragma unroll
            for (int a = 0; a < NC; ++a)
#pragma unroll
                for (int b = 0; b < NC; ++b) mc[a][b] = cadd(mc[a][b], cmulc(uk[NV + a], up[NV + b]));
#pragma unroll
            for (int a = 0; a < NV; ++a)
#pragma unroll
                for (int b = 0; b < NV; ++b) mv[a][b] = cadd(mv[a][b], cmulc(uk[a], up[b]));
        }
        if (textbook) {
#pragma unroll
            for (int a = 0; a < NV; ++a)
#pragma unroll
                for (int b = 0; b < NV; ++b) mv[a][b].y = -mv[a][b].y;
        }
        const double dx = kpt[3 * k] - kpt[3 * kp], dy = kpt[3 * k + 1] - kpt[3 * kp + 1];
        const double dz = kpt[3 * k + 2] - kpt[3 * kp + 2];
        const double V = coulomb_v(coul, dx, dy, dz);
        const Dual dV = coulomb_dual(coul, dx, dy, dz);
        cd gc[NC][NC], gv[NV][NV];
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) gc[a][b] = make_double2(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < NV; ++b) gv[a][b] = make_double2(0.0, 0.0);
        for (int s = 0; s < nst; ++s) {
            cd xa[NCV];
            const cd *xr = X + (size_t)k * NCV + (size_t)sidx[s] * N;
#pragma unroll
            for (int i = 0; i < NCV; ++i) xa[i] = xr[i];
            const cd *xc = xb + (size_t)s * NCV;
            const double ws = sw[s];
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int c2 = 0; c2 < NC; ++c2)
#pragma unroll
                    for (int v = 0; v < NV; ++v)
#pragma unroll
                        for (int v2 = 0; v2 < NV; ++v2) {
                            const cd t = cscale(ws, cmulc(xa[c * NV + v], xc[c2 * NV + v2]));
                            gc[c][c2] = cadd(gc[c][c2], cmul(t, mv[v][v2]));
                            gv[v][v2] = cadd(gv[v][v2], cmul(t, mc[c][c2]));
                        }
        }
        // s(k,k') = sum_cc' M^c[c,c'] gc[c,c']
        double sre = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int c2 = 0; c2 < NC; ++c2) sre += cmul(mc[c][c2], gc[c][c2]).x;
#pragma unroll
        for (int p = 0; p < kNP; ++p) dacc[p] += dV.d[p] * sre;
        const size_t col = (size_t)kp * nk + k;
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int c2 = 0; c2 < NC; ++c2) {
                cd g = cscale(V, gc[c][c2]);
                cd *dst = G + (size_t)(c * NC + c2) * nk2 + col;
                if (accumulate) g = cadd(g, *dst);
                *dst = g;
            }
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int v2 = 0; v2 < NV; ++v2) {
                cd g = cscale(V, gv[v][v2]);
                cd *dst = G + (size_t)(NC * NC + v * NV + v2) * nk2 + col;
                if (accumulate) g = cadd(g, *dst);
                *dst = g;
            }
    }
    // block reduction of the Coulomb-parameter derivatives
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int p = 0; p < kNP; ++p) {
        double v = dacc[p];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[p][wv] = v;
    }
    __syncthreads();
    // this workgroup's partial sums in its own slot (a later state chunk
    // adds to it): the fixed-order sum over k' follows in fixed_sum_kernel,
    // so the result does not depend on the order the workgroups run in
    if (tid < kNP) {
        double v = 0.0;
        for (int i = 0; i < kPT / 64; ++i) v += red[tid][i];
        double *slot = dcoul + (size_t)kp * kNP + tid;
        *slot = accumulate ? *slot + v : v;
    }
}

// out[q] = sum_i part[i nq + q], i = 0..nparts-1, one workgroup per q: each
// thread sums a fixed stride of the partials in order, then a fixed tree.
__global__ void __launch_bounds__(256)
fixed_sum_kernel(int nparts, int nq, const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double red[256];
    const int q = blockIdx.x, tid = threadIdx.x;
    double v = 0.0;
    for (int i = tid; i < nparts; i += 256) v += part[(size_t)i * nq + q];
    red[tid] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
    }
    if (tid == 0) out[q] = red[0];
}

__device__ inline void atomic_max_nonneg(double *addr, double v) {
    unsigned long long *a = reinterpret_cast<unsigned long long *>(addr);
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// TB eigenvector adjoint, one wavefront per k.  W: [(j n + o) nk + k] (Ubar of
// selected band j).  acc: [2] max |Im u_m^H ubar_m|; [3 + 2k] and [4 + 2k]
// this k's parts of dL/dscissor and dL/dE_F (summed by fixed_sum_kernel).
// ebx, ubx (optional): further dL/dE_j(k) [k nb + j] and Ubar contributions in
// the layout of W (from the dipole adjoint).
__global__ void __launch_bounds__(64)
tb_adjoint_kernel(int nk, int n, int nv, int nc, int nst, double scissor, double degtol,
                  const double *__restrict__ eall, const cd *__restrict__ uall, const int *__restrict__ nocpk,
                  const cd *__restrict__ W, const cd *__restrict__ X, long N, const int *__restrict__ sidx,
                  const double *__restrict__ sw, const double *__restrict__ ebx, const cd *__restrict__ ubx,
                  cd *__restrict__ Hbar, double *__restrict__ acc) {
    constexpr int MO = kMaxOrb, MB = 8;
    __shared__ cd U[MO * MO];        // U[o * n + l]
    __shared__ cd Ub[MO * MB];       // Ubar[o * nb + j]
    __shared__ cd Cm[MO * MB];       // C[l * nb + j] = u_l^H ubar_j
    __shared__ cd T[MB * MO];        // T[j * n + b] = (D U^H)[j][b]
    __shared__ cd Gm[MO * MO];
    __shared__ double lam[MO], lbar[MB];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int nb = nc + nv, nn = n * n, ncv = nc * nv;
    const int nocp = nocpk[k];
    for (int idx = tid; idx < nn; idx += 64) U[idx] = uall[(size_t)k * nn + idx];
    for (int l = tid; l < n; l += 64) lam[l] = eall[(size_t)k * n + l] - (l >= nocp ? scissor : 0.0);
    for (int idx = tid; idx < n * nb; idx += 64) {
        const int o = idx / nb, j = idx % nb;
        const size_t wi = ((size_t)j * n + o) * nk + k;
        Ub[idx] = W ? W[wi] : make_double2(0.0, 0.0);
        if (ubx) Ub[idx] = cadd(Ub[idx], ubx[wi]);
    }
    if (tid < nb) {
        // dL/dE_j: +sum over v for a conduction band, -sum over c for a valence band
        double s = 0.0;
        for (int st = 0; st < nst; ++st) {
            const cd *xr = X + (size_t)k * ncv + (size_t)sidx[st] * N;
            double t = 0.0;
            if (tid >= nv) {
                const int c = tid - nv;
                for (int v = 0; v < nv; ++v) t += xr[c * nv + v].x * xr[c * nv + v].x + xr[c * nv + v].y * xr[c * nv + v].y;
            } else {
                for (int c = 0; c < nc; ++c) t -= xr[c * nv + tid].x * xr[c * nv + tid].x + xr[c * nv + tid].y * xr[c * nv + tid].y;
            }
            s += sw[st] * t;
        }
        if (ebx) s += ebx[(size_t)k * nb + tid];
        lbar[tid] = s;
    }
    __syncthreads();
    for (int idx = tid; idx < n * nb; idx += 64) {
        const int l = idx / nb, j = idx % nb;
        cd s = make_double2(0.0, 0.0);
        for (int o = 0; o < n; ++o) s = cadd(s, cmulc(U[o * n + l], Ub[o * nb + j]));
        Cm[idx] = s;
    }
    __syncthreads();
    for (int idx = tid; idx < nb * n; idx += 64) {
        // T[j][b] = sum_l D[j][l] conj(U[b][l]),  m = selected band of j
        const int j = idx / n, b = idx % n, m = nocp - nv + j;
        cd s = make_double2(0.0, 0.0);
        for (int l = 0; l < n; ++l) {
            cd d;
            if (l == m) {
                d = make_double2(lbar[j], 0.0);
            } else {
                const double dl = lam[m] - lam[l];
                if (fabs(dl) < degtol) continue;
                const cd cc = Cm[l * nb + j];
                d = make_double2(2.0 * cc.x / dl, -2.0 * cc.y / dl);
            }
            const cd ubl = U[b * n + l];
            s = cadd(s, cmul(d, make_double2(ubl.x, -ubl.y)));
        }
        T[idx] = s;
    }
    __syncthreads();
    for (int idx = tid; idx < nn; idx += 64) {
        const int a = idx / n, b = idx % n;
        cd s = make_double2(0.0, 0.0);
        for (int j = 0; j < nb; ++j) s = cadd(s, cmul(U[a * n + (nocp - nv + j)], T[j * n + b]));
        Gm[idx] = s;
    }
    __syncthreads();
    double tr = 0.0;
    for (int idx = tid; idx < nn; idx += 64) {
        const int a = idx / n, b = idx % n;
        const cd g1 = Gm[a * n + b], g2 = Gm[b * n + a];
        const cd h = make_double2(0.5 * (g1.x + g2.x), 0.5 * (g1.y - g2.y));   // Hermitian part
        Hbar[(size_t)k * nn + idx] = h;
        if (a == b) tr += h.x;
    }
    for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off, 64);
    if (tid == 0) {
        double ds = 0.0, gi = 0.0;
        for (int j = nv; j < nb; ++j) ds += lbar[j];
        for (int j = 0; j < nb; ++j) gi = fmax(gi, fabs(Cm[(nocp - nv + j) * nb + j].y));
        acc[3 + 2 * (size_t)k] = ds;          // per-k partials, summed in a fixed order
        acc[4 + 2 * (size_t)k] = -tr;
        atomic_max_nonneg(acc + 2, gi);        // a maximum: independent of the order
    }
}

// g(R)_ab = sum_k Hbar(k)_ab e^{-i k.R} / deg(R), with the phase of the
// forward Fourier sum (float32-rounded in compat mode).  Output in the
// reference layout (nvec, norb, norb), column-major.
// With Dbar (the adjoint of the dipole matrix dH/dk_alpha(k) = sum_R i R_alpha
// e^{ik.R} H(R)/deg), the term sum_k Dbar(k) (-i R_alpha) e^{-ik.R}/deg is added,
// R_alpha rounded as in the dipole kernel.
__global__ void fourier_adjoint_kernel(int nk, int n, int nvec, int compat, const double *__restrict__ kpt,
                                       const double *__restrict__ rcart, const double *__restrict__ invdeg,
                                       const cd *__restrict__ Hbar, const cd *__restrict__ Dbar, int alpha,
                                       double *__restrict__ gre, double *__restrict__ gim) {
    const int nn = n * n;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nvec * nn) return;
    const int r = (int)(t / nn), idx = (int)(t % nn), a = idx / n, b = idx % n;
    double Ra = Dbar ? rcart[3 * r + alpha] : 0.0;
    if (compat) Ra = (double)(float)Ra;
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < nk; ++k) {
        double x = kdotr(kpt[3 * k], kpt[3 * k + 1], kpt[3 * k + 2], rcart + 3 * r);
        if (compat) x = (double)(float)x;
        double sn, cs;
        sincos(x, &sn, &cs);
        cd h = Hbar[(size_t)k * nn + idx];
        if (Dbar) {
            // + Dbar (-i R_a) = + R_a (Dbar.y, -Dbar.x)
            const cd db = Dbar[(size_t)k * nn + idx];
            h.x += Ra * db.y;
            h.y -= Ra * db.x;
        }
        ar += h.x * cs + h.y * sn;
        ai += h.y * cs - h.x * sn;
    }
    const size_t o = (size_t)r + (size_t)nvec * ((size_t)a + (size_t)n * b);
    gre[o] = ar * invdeg[r];
    gim[o] = ai * invdeg[r];
}

template <int NC, int NV>
void launch_pair(Ctx &c, const CoulombDev &cd_, const cd *X, int nst, int accumulate, const int *sidx,
                 const double *sw, cd *G, double *dcoul) {
    const int nk = c.bands.nk, norb = c.model.norb;
    const size_t shmem = ((size_t)norb * (NC + NV) + (size_t)nst * NC * NV) * sizeof(cd);
    hipLaunchKernelGGL((grad_pair_kernel<NC, NV>), dim3(nk), dim3(kPT), shmem, c.stream, nk, norb, nst,
                       c.bse.convention == WTB_CONV_TEXTBOOK ? 1 : 0, accumulate, cd_, c.kpt.p, c.bands.usel.p, X,
                       c.bse.N, sidx, sw, G, dcoul);
    WTB_LAUNCH_CHECK();
}

void dispatch_pair(Ctx &c, const CoulombDev &cd_, const cd *X, int nst, int accumulate, const int *sidx,
                   const double *sw, cd *G, double *dcoul) {
    const int nc = c.bands.nc, nv = c.bands.nv;
#define WTB_PAIR_CASE(A, B) \
    if (nc == A && nv == B) return launch_pair<A, B>(c, cd_, X, nst, accumulate, sidx, sw, G, dcoul);
    WTB_PAIR_CASE(1, 1) WTB_PAIR_CASE(1, 2) WTB_PAIR_CASE(2, 1) WTB_PAIR_CASE(2, 2)
    WTB_PAIR_CASE(1, 3) WTB_PAIR_CASE(3, 1) WTB_PAIR_CASE(2, 3) WTB_PAIR_CASE(3, 2) WTB_PAIR_CASE(3, 3)
    WTB_PAIR_CASE(1, 4) WTB_PAIR_CASE(4, 1) WTB_PAIR_CASE(2, 4) WTB_PAIR_CASE(4, 2) WTB_PAIR_CASE(3, 4)
    WTB_PAIR_CASE(4, 3) WTB_PAIR_CASE(4, 4)
#undef WTB_PAIR_CASE
    throw Error{WTB_ERR_UNSUPPORTED, "nc, nv must be in 1..4"};
}

}  // namespace

// dL for L = sum_s w_s x_s^H H x_s over the columns of X (N x nx, ld N),
// plus the optional dipole-adjoint terms (GradExtra).  grad_energies is the
// case X = the solved eigenvectors.
void grad_core(Ctx &c, const cd *X, int nx, const double *w, const GradExtra *ex, wtb_grad *out, double *g_re,
               double *g_im) {
    if (!c.bse.valid || !c.bands.valid) throw Error{WTB_ERR_STATE, "wtb_bse_build must run first"};
    if (!w || !out) throw Error{WTB_ERR_ARG, "null weights or output"};
    Bands &b = c.bands;
    const int nk = b.nk, nc = b.nc, nv = b.nv, nb = nc + nv, n = c.model.norb;
    if (nb > 8) throw Error{WTB_ERR_UNSUPPORTED, "nc + nv > 8 in the adjoint"};
    std::vector<int> idx;
    std::vector<double> ws;
    for (int s = 0; s < nx; ++s)
        if (w[s] != 0.0) {
            if (!std::isfinite(w[s])) throw Error{WTB_ERR_ARG, "non-finite weight"};
            idx.push_back(s);
            ws.push_back(w[s]);
        }
    std::memset(out, 0, sizeof(*out));
    const int npq = nc * nc + nv * nv;
    DBuf<cd> G, Wb, Hbar;
    DBuf<double> scal, gdev, dpart, apart;
    DBuf<int> didx;
    DBuf<double> dw;
    scal.alloc(kNP + 3);
    WTB_HIP_CHECK(hipMemsetAsync(scal.p, 0, sizeof(double) * (kNP + 3), c.stream));
    // per-workgroup partials of the Coulomb derivatives (one row per k') and
    // of the scissor / E_F derivatives (two per k) plus the gauge maximum
    dpart.alloc((size_t)nk * kNP);
    apart.alloc(3 + 2 * (size_t)nk);
    WTB_HIP_CHECK(hipMemsetAsync(dpart.p, 0, sizeof(double) * nk * kNP, c.stream));
    WTB_HIP_CHECK(hipMemsetAsync(apart.p, 0, sizeof(double) * (3 + 2 * (size_t)nk), c.stream));
    Hbar.alloc((size_t)nk * n * n);
    const bool extra = ex && (ex->ebar || ex->ubar);
    if (idx.empty() && !extra) {
        WTB_HIP_CHECK(hipMemsetAsync(Hbar.p, 0, sizeof(cd) * nk * n * n, c.stream));
    } else {
        if (!idx.empty()) {
            didx.alloc(idx.size());
            dw.alloc(ws.size());
            WTB_HIP_CHECK(
                hipMemcpyAsync(didx.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c.stream));
            WTB_HIP_CHECK(hipMemcpyAsync(dw.p, ws.data(), sizeof(double) * ws.size(), hipMemcpyHostToDevice, c.stream));
            const CoulombDev cd_ = make_coulomb(c.bse.coul, c.model.rlat, c.ngrid);
            G.alloc((size_t)npq * nk * nk);
            for (size_t s0 = 0; s0 < idx.size(); s0 += kMaxStates) {
                const int ns = (int)std::min<size_t>(kMaxStates, idx.size() - s0);
                dispatch_pair(c, cd_, X, ns, s0 > 0, didx.p + s0, dw.p + s0, G.p, dpart.p);
            }
            // Ubar_j(k) = sum_{j'} A_{jj'} U_{j'}  (textbook valence: sum_{j'} A_{j'j}^T U_{j'})
            Wb.alloc((size_t)nb * n * nk);
            const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
            const int ldu = nb * nk;
            const bool textbook = c.bse.convention == WTB_CONV_TEXTBOOK;
            auto mat = [&](int p) { return (const rocblas_double_complex *)(G.p + (size_t)p * nk * nk); };
            auto ucol = [&](int j) { return (const rocblas_double_complex *)(b.usel.p + (size_t)j * nk); };
            for (int j = 0; j < nb; ++j) {
                rocblas_double_complex *Wj = (rocblas_double_complex *)(Wb.p + (size_t)j * n * nk);
                const bool cond = j >= nv;
                const int m = cond ? nc : nv;
                for (int j2 = 0; j2 < m; ++j2) {
                    int p;
                    rocblas_operation op = rocblas_operation_none;
                    if (cond) {
                        p = (j - nv) * nc + j2;
                    } else if (!textbook) {
                        p = nc * nc + j * nv + j2;
                    } else {
                        p = nc * nc + j2 * nv + j;
                        op = rocblas_operation_transpose;
                    }
                    const int src = cond ? nv + j2 : j2;
                    WTB_BLAS_CHECK(rocblas_zgemm(c.blas, op, rocblas_operation_none, nk, n, nk, &one, mat(p), nk,
                                                 ucol(src), ldu, j2 == 0 ? &zero : &one, Wj, nk));
                }
            }
        }
        hipLaunchKernelGGL(tb_adjoint_kernel, dim3(nk), dim3(64), 0, c.stream, nk, n, nv, nc, (int)idx.size(),
                           c.model.scissor, 1e-8, b.eall.p, b.uall.p, b.nocp.p, idx.empty() ? nullptr : Wb.p, X,
                           c.bse.N, didx.p, dw.p, ex ? ex->ebar : nullptr, ex ? ex->ubar : nullptr, Hbar.p,
                           apart.p);
        WTB_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(fixed_sum_kernel, dim3(kNP), dim3(256), 0, c.stream, nk, kNP, dpart.p, scal.p);
    hipLaunchKernelGGL(fixed_sum_kernel, dim3(2), dim3(256), 0, c.stream, nk, 2, apart.p + 3, scal.p + kNP);
    WTB_HIP_CHECK(hipMemcpyAsync(scal.p + kNP + 2, apart.p + 2, sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    WTB_LAUNCH_CHECK();
    if (g_re || g_im) {
        const int nvec = c.model.nvec;
        const size_t tot = (size_t)nvec * n * n;
        gdev.alloc(2 * tot);
        hipLaunchKernelGGL(fourier_adjoint_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, nk, n,
                           nvec, (c.model.flags & WTB_FLAG_COMPAT_F32) ? 1 : 0, c.kpt.p, c.model.rcart.p,
                           c.model.invdeg.p, Hbar.p, ex ? ex->dbar : nullptr, ex ? ex->alpha : 0, gdev.p,
                           gdev.p + tot);
        WTB_LAUNCH_CHECK();
        if (g_re) WTB_HIP_CHECK(hipMemcpyAsync(g_re, gdev.p, sizeof(double) * tot, hipMemcpyDeviceToHost, c.stream));
        if (g_im)
            WTB_HIP_CHECK(hipMemcpyAsync(g_im, gdev.p + tot, sizeof(double) * tot, hipMemcpyDeviceToHost, c.stream));
    }
    double h[kNP + 3];
    WTB_HIP_CHECK(hipMemcpyAsync(h, scal.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int p = 0; p < kNP; ++p) out->coulomb[p] = h[p];
    out->scissor = h[kNP];
    out->efermi = h[kNP + 1];
    out->gauge_residual = h[kNP + 2];
}

void grad_energies(Ctx &c, const double *w, wtb_grad *out, double *g_re, double *g_im) {
    if (!c.sol.valid) throw Error{WTB_ERR_STATE, "wtb_bse_solve must run first"};
    grad_core(c, c.sol.evecs.p, c.sol.nev, w, nullptr, out, g_re, g_im);
}

}  // namespace wtb
