// Optical dipoles and exciton oscillator strengths.
//
// Dipoles (replaces the optsp/hlm loop of bsesolver, src/module_bse.F90:2317-2331,
// src/module_optics.F90:12-167): the reference rebuilds dH/dk for every
// transition; here one 64-lane workgroup per k builds
//   dH/dk_a(k) = sum_R (i R_a) e^{ik.R} H(R)/deg(R)
// once for a = x, y, z in LDS and contracts it with all nc*nv band pairs:
//   d_a(k,c,v) = <c|dH/dk_a|v> * Re(1/(E_c - E_v + i eta)).
// Oscillator strengths (dielbsev/dielbsep, src/module_bse.F90:171-245):
//   act(n) = Re[(sum_j X_jn d1_j) conj(sum_j X_jn d2_j)]   (no conjugate on X);
// a store built with a temperature weights the second sum by f_v - f_c of
// state j (dielbseptemp, src/module_bse.F90:784-810).
#include "wtb_internal.h"

namespace wtb {
namespace {

__device__ __forceinline__ cd cmul(cd a, cd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }

constexpr int kL = 64;

__global__ void __launch_bounds__(kL)
dipole_kernel(int nk, int n, int nvec, int nv, int nc, int compat, int bz, double eta, double isq2,
              const double *__restrict__ kpt, const cd *__restrict__ hop, const double *__restrict__ rcart,
              const double *__restrict__ invdeg, const double *__restrict__ esel, const cd *__restrict__ usel,
              cd *__restrict__ d, long N) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int k = blockIdx.x, tid = threadIdx.x, nn = n * n, nb = nc + nv;
    cd *D = reinterpret_cast<cd *>(smem);       // 3 * nn
    cd *U = D + 3 * nn;                         // nb * n, [j][o]
    cd *ph = U + nb * n;                        // nvec
    const double kx = kpt[3 * k], ky = kpt[3 * k + 1], kz = kpt[3 * k + 2];
    for (int r = tid; r < nvec; r += kL) {
        double x = kdotr(kx, ky, kz, rcart + 3 * r);   // same value as the H(k) build
        if (compat) x = (double)(float)x;
        double sn, cs;
        sincos(x, &sn, &cs);
        ph[r] = make_double2(cs * invdeg[r], sn * invdeg[r]);
    }
    for (int idx = tid; idx < nb * n; idx += kL) {
        int j = idx / n, o = idx % n;
        U[idx] = usel[((size_t)o * nb + j) * nk + k];
    }
    __syncthreads();
    for (int e = tid; e < nn; e += kL) {
        cd ax = make_double2(0, 0), ay = ax, az = ax;
        for (int r = 0; r < nvec; ++r) {
            cd f = cmul(ph[r], hop[(size_t)r * nn + e]);
            double Rx = rcart[3 * r], Ry = rcart[3 * r + 1], Rz = rcart[3 * r + 2];
            if (compat) { Rx = (double)(float)Rx; Ry = (double)(float)Ry; Rz = (double)(float)Rz; }
            // (i R) f = R * (-f.y, f.x)
            ax.x -= Rx * f.y; ax.y += Rx * f.x;
            ay.x -= Ry * f.y; ay.y += Ry * f.x;
            az.x -= Rz * f.y; az.y += Rz * f.x;
        }
        D[e] = ax;
        D[nn + e] = ay;
        D[2 * nn + e] = az;
    }
    __syncthreads();
    const int ncv = nc * nv;
    for (int t = tid; t < ncv; t += kL) {
        const int c = t / nv, v = t % nv;
        const cd *uc = U + (nv + c) * n;
        const cd *uv = U + v * n;
        const double de = esel[(size_t)k * nb + nv + c] - esel[(size_t)k * nb + v];
        const double factor = de / (de * de + eta * eta);   // Re(1/(de + i eta))
        const long i = (long)k * ncv + t;
        cd r[3];
        for (int a = 0; a < 3; ++a) {
            const cd *Dm = D + a * nn;
            cd s = make_double2(0, 0);
            for (int o = 0; o < n; ++o) {
                cd row = make_double2(0, 0);
                for (int p = 0; p < n; ++p) row = cadd(row, cmul(Dm[o * n + p], uv[p]));
                s.x += uc[o].x * row.x + uc[o].y * row.y;   // conj(uc[o]) * row
                s.y += uc[o].x * row.y - uc[o].y * row.x;
            }
            if (bz) {   // <c|dH/dk_a|v> / (de + i eta) (spoptpolbz, src/module_bse.F90:4676-4682)
                const double den = de * de + eta * eta;
                r[a] = make_double2((s.x * de + s.y * eta) / den, (s.y * de - s.x * eta) / den);
            } else {
                r[a] = make_double2(factor * s.x, factor * s.y);
            }
            d[(size_t)a * N + i] = r[a];
        }
        // circular components (x +- i y)/sqrt(2)
        d[3 * N + i] = make_double2((r[0].x - r[1].y) * isq2, (r[0].y + r[1].x) * isq2);
        d[4 * N + i] = make_double2((r[0].x + r[1].y) * isq2, (r[0].y - r[1].x) * isq2);
    }
}

// per k: sum over the nc nv transitions of |d_a|^2, a = x, y, z, sp, sm, in
// the reference's c-outer / v-inner order; act is (nk, 5) column-major
__global__ void bz_sum_kernel(int nk, int ncv, long N, const cd *__restrict__ d, double *__restrict__ act) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= 5L * nk) return;
    const int a = (int)(t / nk), k = (int)(t % nk);
    double s = 0.0;
    for (int i = 0; i < ncv; ++i) {
        const cd x = d[(size_t)a * N + (size_t)k * ncv + i];
        s = s + (x.x * x.x + x.y * x.y);
    }
    act[(size_t)a * nk + k] = s;
}

__global__ void __launch_bounds__(256)
oscillator_kernel(long N, const cd *__restrict__ X, const cd *__restrict__ d1, const cd *__restrict__ d2,
                  const double *__restrict__ fdeh, double *__restrict__ act) {
    __shared__ double red[4][4];
    const int n = blockIdx.x;
    const cd *x = X + (size_t)n * N;
    double s1r = 0, s1i = 0, s2r = 0, s2i = 0;
    for (long j = threadIdx.x; j < N; j += 256) {
        cd xv = x[j], a = d1[j], b = d2[j];
        if (fdeh) {     // temperature BSE: the second sum carries f_v - f_c (dielbseptemp)
            b.x *= fdeh[j];
            b.y *= fdeh[j];
        }
        s1r += xv.x * a.x - xv.y * a.y;
        s1i += xv.x * a.y + xv.y * a.x;
        s2r += xv.x * b.x - xv.y * b.y;
        s2i += xv.x * b.y + xv.y * b.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s1r += __shfl_xor(s1r, off, 64);
        s1i += __shfl_xor(s1i, off, 64);
        s2r += __shfl_xor(s2r, off, 64);
        s2i += __shfl_xor(s2i, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[w][0] = s1r; red[w][1] = s1i; red[w][2] = s2r; red[w][3] = s2i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0, e = 0;
        for (int i = 0; i < 4; ++i) { a += red[i][0]; b += red[i][1]; c += red[i][2]; e += red[i][3]; }
        act[n] = a * c + b * e;   // Re[(a + ib) conj(c + ie)]
    }
}

}  // namespace

void launch_optics(Ctx &c, double cshift) {
    Bands &b = c.bands;
    if (!b.valid) throw Error{WTB_ERR_STATE, "wtb_hk_eig must run before wtb_optics"};
    const int n = c.model.norb, nk = b.nk;
    const long N = (long)nk * b.nc * b.nv;
    const bool compat = c.model.flags & WTB_FLAG_COMPAT_F32;
    const double eta = compat ? (double)(float)cshift : cshift;
    const double isq2 = compat ? (double)(1.0f / sqrtf(2.0f)) : 1.0 / std::sqrt(2.0);
    size_t shmem = (3 * (size_t)n * n + (size_t)b.nb * n + c.model.nvec) * sizeof(cd);
    if (shmem > 160 * 1024) throw Error{WTB_ERR_UNSUPPORTED, "too many R vectors for the dipole kernel"};
    c.dip.alloc(5 * (size_t)N);
    hipLaunchKernelGGL(dipole_kernel, dim3(nk), dim3(kL), shmem, c.stream, nk, n, c.model.nvec, b.nv, b.nc,
                       (int)compat, 0, eta, isq2, c.kpt.p, c.model.hop.p, c.model.rcart.p, c.model.invdeg.p,
                       b.esel.p, b.usel.p, c.dip.p, N);
    WTB_LAUNCH_CHECK();
    c.dip_valid = true;
}

void launch_optics_bz(Ctx &c, double sme, double *act_host) {
    Bands &b = c.bands;
    if (!b.valid) throw Error{WTB_ERR_STATE, "wtb_hk_eig must run before wtb_optics_bz"};
    if (!act_host) throw Error{WTB_ERR_ARG, "null output"};
    const int n = c.model.norb, nk = b.nk;
    const long N = (long)nk * b.nc * b.nv;
    // spoptpolbz's norm = 1.0/sqrt(2.) is a default-real expression
    const double isq2 = (double)(1.0f / sqrtf(2.0f));
    size_t shmem = (3 * (size_t)n * n + (size_t)b.nb * n + c.model.nvec) * sizeof(cd);
    if (shmem > 160 * 1024) throw Error{WTB_ERR_UNSUPPORTED, "too many R vectors for the dipole kernel"};
    const bool compat = c.model.flags & WTB_FLAG_COMPAT_F32;
    DBuf<cd> d;
    DBuf<double> act;
    d.alloc(5 * (size_t)N);
    act.alloc(5 * (size_t)nk);
    hipLaunchKernelGGL(dipole_kernel, dim3(nk), dim3(kL), shmem, c.stream, nk, n, c.model.nvec, b.nv, b.nc,
                       (int)compat, 1, sme, isq2, c.kpt.p, c.model.hop.p, c.model.rcart.p, c.model.invdeg.p,
                       b.esel.p, b.usel.p, d.p, N);
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(bz_sum_kernel, dim3((unsigned)((5L * nk + 255) / 256)), dim3(256), 0, c.stream, nk,
                       b.nc * b.nv, N, d.p, act.p);
    WTB_LAUNCH_CHECK();
    WTB_HIP_CHECK(hipMemcpyAsync(act_host, act.p, sizeof(double) * 5 * nk, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
}

void launch_oscillator(Ctx &c, int nev, int a, int b, double *act_dev) {
    const long N = c.bse.N;
    const double *f = c.bse.temperature >= 0.0 ? c.bse.fdeh.p : nullptr;
    hipLaunchKernelGGL(oscillator_kernel, dim3(nev), dim3(256), 0, c.stream, N, c.sol.evecs.p,
                       c.dip.p + (size_t)a * N, c.dip.p + (size_t)b * N, f, act_dev);
    WTB_LAUNCH_CHECK();
}

}  // namespace wtb
