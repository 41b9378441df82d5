// Gradient of a spectrum loss over the solved exciton states (no reference
// counterpart; the forward spectrum follows optdiel, src/module_optics.F90:171-241,
// with a fixed Lorentzian width).
//
//   eps2(w) = C sum_n f_n g / ((w - E_n)^2 + g^2),   C = 180.90 / (N_k Vol)
//   f_n = |a_n|^2,  a_n = sum_i x_n[i] d_i          (reference convention,
//                                                    src/module_bse.F90:231-238)
//              or   a_n = sum_i conj(x_n[i]) d_i    (textbook convention)
//   L = sum_j (eps2(w_j) - t_j)^2
// with d the dipoles of one Cartesian component (the wtb_optics kernel).
//
//   dL = sum_n Ebar_n dE_n + sum_n fbar_n df_n,   df_n = 2 Re(conj(a_n) da_n).
// (1) sum_n Ebar_n dE_n is the Hellmann-Feynman adjoint with weights Ebar.
// (2) Through the eigenvectors: sum_n 2 Re(dx_n^H g_n) with
//       g_n = fbar_n conj(a_n) d (textbook) or fbar_n a_n conj(d) (reference),
//     and dx_n = sum_{m != n} x_m (x_m^H dH x_n) / (E_n - E_m), this is
//       sum_n 2 Re(x_n^H dH l_n),   l_n = (E_n - H)^+ g_n on the complement of x_n:
//     the other solved states explicitly, the rest of the spectrum by
//     preconditioned CG on (H - E_n) restricted to the complement of all
//     solved states (positive definite when E_n lies below the unsolved
//     spectrum).  By polarisation
//       2 Re(x^H dH l) = [(x + t l)^H dH (x + t l) - x^H dH x - t^2 l^H dH l] / t,
//     three quadratic forms for the same adjoint (t = 1 / |l|).
// (3) Through the dipoles: 2 Re sum_i conj(h_i) dd_i with
//       h = sum_n fbar_n a_n x_n (textbook) or sum_n fbar_n a_n conj(x_n) (reference),
//     and d_i = F(dE) u_c^H D_alpha(k) u_v, F(x) = x / (x^2 + eta^2): Ubar and
//     Ebar of the selected TB bands and the adjoint Dbar of D_alpha(k), which
//     the TB eigenvector and Fourier adjoints carry to H(R).
// The reference convention's strengths depend on the eigenvector gauge (a
// phase change of a conduction band multiplies a term by e^{-2i theta}), so
// there only the Coulomb-parameter and scissor gradients (which leave the TB
// bands unchanged) are defined.
#include "wtb_internal.h"

#include <cmath>
#include <algorithm>
#include <vector>

namespace wtb {
namespace {

__device__ __forceinline__ cd cmul(cd a, cd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cconj(cd a) { return make_double2(a.x, -a.y); }

constexpr int kRT = 256;
__device__ __forceinline__ cd block_sum_c(cd v, cd *red) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    cd s = make_double2(0.0, 0.0);
    for (int w = 0; w < kRT / 64; ++w) {
        s.x += red[w].x;
        s.y += red[w].y;
    }
    return s;
}

// a_n = sum_i op(x_n[i]) d_i, op = conj when conjx
__global__ void __launch_bounds__(kRT)
amplitude_kernel(long N, const cd *__restrict__ X, const cd *__restrict__ d, int conjx, cd *__restrict__ a) {
    __shared__ cd red[kRT / 64];
    const int n = blockIdx.x;
    const cd *x = X + (size_t)n * N;
    cd s = make_double2(0.0, 0.0);
    for (long i = threadIdx.x; i < N; i += kRT) {
        const cd xv = conjx ? cconj(x[i]) : x[i];
        const cd p = cmul(xv, d[i]);
        s.x += p.x;
        s.y += p.y;
    }
    s = block_sum_c(s, red);
    if (threadIdx.x == 0) a[n] = s;
}

// G[:, n] = coef_n op(d), op = conj when conjd
__global__ void scaled_dipole_kernel(long N, const cd *__restrict__ d, int conjd, const cd *__restrict__ coef,
                                     cd *__restrict__ G) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= N) return;
    const cd dv = conjd ? cconj(d[i]) : d[i];
    G[i + (size_t)n * N] = cmul(coef[n], dv);
}

// h = sum_n coef_n op(x_n), op = conj when conjx
__global__ void combine_kernel(long N, int nev, const cd *__restrict__ X, int conjx, const cd *__restrict__ coef,
                               cd *__restrict__ h) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= N) return;
    cd s = make_double2(0.0, 0.0);
    for (int n = 0; n < nev; ++n) {
        const cd xv = conjx ? cconj(X[i + (size_t)n * N]) : X[i + (size_t)n * N];
        const cd p = cmul(coef[n], xv);
        s.x += p.x;
        s.y += p.y;
    }
    h[i] = s;
}

// out[n] = Re sum_i conj(A[i, n]) B[i, n]
__global__ void __launch_bounds__(kRT)
coldot_kernel(long N, const cd *__restrict__ A, const cd *__restrict__ B, double *__restrict__ out) {
    __shared__ cd red[kRT / 64];
    const int n = blockIdx.x;
    const cd *a = A + (size_t)n * N, *b = B + (size_t)n * N;
    double s = 0.0;
    for (long i = threadIdx.x; i < N; i += kRT) s += a[i].x * b[i].x + a[i].y * b[i].y;
    cd v = block_sum_c(make_double2(s, 0.0), red);
    if (threadIdx.x == 0) out[n] = v.x;
}

// Q[:, n] -= E_n P[:, n]
__global__ void shift_kernel(long N, const double *__restrict__ E, const cd *__restrict__ P, cd *__restrict__ Q) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= N) return;
    const size_t o = i + (size_t)n * N;
    Q[o].x -= E[n] * P[o].x;
    Q[o].y -= E[n] * P[o].y;
}

// Z[:, n] = R[:, n] / max(|diag - E_n|, 0.05): positive definite, as CG needs
__global__ void jacobi_prec_kernel(long N, const double *__restrict__ diag, const double *__restrict__ E,
                                   const cd *__restrict__ R, cd *__restrict__ Z) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= N) return;
    const double den = fmax(fabs(diag[i] - E[n]), 0.05);
    const size_t o = i + (size_t)n * N;
    Z[o] = make_double2(R[o].x / den, R[o].y / den);
}

// B -= A over n elements
__global__ void sub_kernel(size_t n, const cd *__restrict__ A, cd *__restrict__ B) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    B[i].x -= A[i].x;
    B[i].y -= A[i].y;
}

// Y += alpha P, R -= alpha Q  (per column)
__global__ void cg_step_kernel(long N, const double *__restrict__ alpha, const cd *__restrict__ P,
                               const cd *__restrict__ Q, cd *__restrict__ Y, cd *__restrict__ R) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= N) return;
    const size_t o = i + (size_t)n * N;
    const double a = alpha[n];
    Y[o].x += a * P[o].x;
    Y[o].y += a * P[o].y;
    R[o].x -= a * Q[o].x;
    R[o].y -= a * Q[o].y;
}

// P = Z + beta P
__global__ void cg_dir_kernel(long N, const double *__restrict__ beta, const cd *__restrict__ Z, cd *__restrict__ P) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= N) return;
    const size_t o = i + (size_t)n * N;
    const double b = beta[n];
    P[o] = make_double2(Z[o].x + b * P[o].x, Z[o].y + b * P[o].y);
}

// Xall = [X, X + t L, L]
__global__ void polar_kernel(long N, int nev, const cd *__restrict__ X, const cd *__restrict__ L,
                             const double *__restrict__ t, cd *__restrict__ Xall) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= N) return;
    const size_t o = i + (size_t)n * N;
    const cd x = X[o], l = L[o];
    Xall[o] = x;
    Xall[o + (size_t)nev * N] = make_double2(x.x + t[n] * l.x, x.y + t[n] * l.y);
    Xall[o + 2 * (size_t)nev * N] = l;
}

// Dipole adjoint, one 64-lane workgroup per k.  For the band pairs (c, v) of
// k with d = F(dE) u_c^H D u_v (D = dH/dk_alpha) and the weight h:
//   Ebar_c += 2 Re(conj(h) s) F'(dE), Ebar_v -= the same   (s = u_c^H D u_v)
//   Ubar_c += conj(h) F D u_v,        Ubar_v += h F D^H u_c
//   Dbar   += 2 h F u_c u_v^H
constexpr int kDL = 64;
__global__ void __launch_bounds__(kDL)
dipole_adjoint_kernel(int nk, int n, int nvec, int nv, int nc, int compat, double eta, int alpha,
                      const double *__restrict__ kpt, const cd *__restrict__ hop, const double *__restrict__ rcart,
                      const double *__restrict__ invdeg, const double *__restrict__ esel,
                      const cd *__restrict__ usel, const cd *__restrict__ h, double *__restrict__ ebar,
                      cd *__restrict__ ubar, cd *__restrict__ dbar) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int k = blockIdx.x, tid = threadIdx.x, nn = n * n, nb = nc + nv, ncv = nc * nv;
    cd *D = reinterpret_cast<cd *>(smem);       // nn, [o * n + p]
    cd *U = D + nn;                             // nb * n, [j][o]
    cd *Du = U + nb * n;                        // nv * n: D u_v
    cd *DHu = Du + nv * n;                      // nc * n: D^H u_c
    cd *ph = DHu + nc * n;                      // nvec
    cd *hk = ph + nvec;                         // ncv: h of this k
    double *F = reinterpret_cast<double *>(hk + ncv);   // ncv
    double *Fp = F + ncv;                                // ncv
    cd *S = reinterpret_cast<cd *>(Fp + ncv);            // ncv: u_c^H D u_v
    const double kx = kpt[3 * k], ky = kpt[3 * k + 1], kz = kpt[3 * k + 2];
    for (int r = tid; r < nvec; r += kDL) {
        double x = kdotr(kx, ky, kz, rcart + 3 * r);    // the phase of the dipole kernel
        if (compat) x = (double)(float)x;
        double sn, cs;
        sincos(x, &sn, &cs);
        ph[r] = make_double2(cs * invdeg[r], sn * invdeg[r]);
    }
    for (int idx = tid; idx < nb * n; idx += kDL) {
        const int j = idx / n, o = idx % n;
        U[idx] = usel[((size_t)o * nb + j) * nk + k];
    }
    for (int t = tid; t < ncv; t += kDL) {
        hk[t] = h[(size_t)k * ncv + t];
        const int c = t / nv, v = t % nv;
        const double de = esel[(size_t)k * nb + nv + c] - esel[(size_t)k * nb + v];
        const double den = de * de + eta * eta;
        F[t] = de / den;
        Fp[t] = (eta * eta - de * de) / (den * den);
    }
    __syncthreads();
    for (int e = tid; e < nn; e += kDL) {
        cd a = make_double2(0.0, 0.0);
        for (int r = 0; r < nvec; ++r) {
            const cd f = cmul(ph[r], hop[(size_t)r * nn + e]);
            double Ra = rcart[3 * r + alpha];
            if (compat) Ra = (double)(float)Ra;
            // (i R) f
            a.x -= Ra * f.y;
            a.y += Ra * f.x;
        }
        D[e] = a;
    }
    __syncthreads();
    for (int idx = tid; idx < nv * n; idx += kDL) {
        const int v = idx / n, o = idx % n;
        cd s = make_double2(0.0, 0.0);
        for (int p = 0; p < n; ++p) {
            const cd m = cmul(D[o * n + p], U[v * n + p]);
            s.x += m.x;
            s.y += m.y;
        }
        Du[idx] = s;
    }
    for (int idx = tid; idx < nc * n; idx += kDL) {
        const int c = idx / n, o = idx % n;
        cd s = make_double2(0.0, 0.0);
        for (int p = 0; p < n; ++p) {
            const cd m = cmul(cconj(D[p * n + o]), U[(nv + c) * n + p]);
            s.x += m.x;
            s.y += m.y;
        }
        DHu[idx] = s;
    }
    __syncthreads();
    for (int t = tid; t < ncv; t += kDL) {
        const int c = t / nv, v = t % nv;
        cd s = make_double2(0.0, 0.0);
        for (int o = 0; o < n; ++o) {
            const cd m = cmul(cconj(U[(nv + c) * n + o]), Du[v * n + o]);
            s.x += m.x;
            s.y += m.y;
        }
        S[t] = s;
    }
    __syncthreads();
    if (tid < nb) {
        double s = 0.0;
        if (tid >= nv) {
            const int c = tid - nv;
            for (int v = 0; v < nv; ++v) {
                const int t = c * nv + v;
                const cd hs = cmul(cconj(hk[t]), S[t]);
                s += 2.0 * hs.x * Fp[t];
            }
        } else {
            for (int c = 0; c < nc; ++c) {
                const int t = c * nv + tid;
                const cd hs = cmul(cconj(hk[t]), S[t]);
                s -= 2.0 * hs.x * Fp[t];
            }
        }
        ebar[(size_t)k * nb + tid] = s;
    }
    for (int idx = tid; idx < nb * n; idx += kDL) {
        const int j = idx / n, o = idx % n;
        cd s = make_double2(0.0, 0.0);
        if (j >= nv) {
            const int c = j - nv;
            for (int v = 0; v < nv; ++v) {
                const int t = c * nv + v;
                const cd m = cmul(cconj(hk[t]), Du[v * n + o]);
                s.x += F[t] * m.x;
                s.y += F[t] * m.y;
            }
        } else {
            for (int c = 0; c < nc; ++c) {
                const int t = c * nv + j;
                const cd m = cmul(hk[t], DHu[c * n + o]);
                s.x += F[t] * m.x;
                s.y += F[t] * m.y;
            }
        }
        ubar[((size_t)j * n + o) * nk + k] = s;
    }
    for (int e = tid; e < nn; e += kDL) {
        const int a = e / n, b = e % n;
        cd s = make_double2(0.0, 0.0);
        for (int t = 0; t < ncv; ++t) {
            const int c = t / nv, v = t % nv;
            const cd m = cmul(hk[t], cmul(U[(nv + c) * n + a], cconj(U[v * n + b])));
            s.x += 2.0 * F[t] * m.x;
            s.y += 2.0 * F[t] * m.y;
        }
        dbar[(size_t)k * nn + e] = s;
    }
}

}  // namespace

void grad_spectrum(Ctx &c, int comp, double cshift, double gamma, int nw, const double *omega, const double *target,
                   double *loss, double *eps2, wtb_grad *out, double *g_re, double *g_im) {
    if (!c.sol.valid || !c.bse.valid || !c.bands.valid) throw Error{WTB_ERR_STATE, "wtb_bse_solve must run first"};
    if (comp < 0 || comp > 2) throw Error{WTB_ERR_ARG, "component must be x, y or z"};
    if (nw < 1 || !omega || !target || !out || !(gamma > 0.0)) throw Error{WTB_ERR_ARG, "bad spectrum arguments"};
    const bool textbook = c.bse.convention == WTB_CONV_TEXTBOOK;
    if (!textbook && (g_re || g_im))
        throw Error{WTB_ERR_UNSUPPORTED,
                    "H(R) gradients of a spectrum need the gauge-invariant textbook convention"};
    Bands &b = c.bands;
    const int nev = c.sol.nev, nk = b.nk, n = c.model.norb, nb = b.nc + b.nv;
    const long N = c.bse.N;
    const bool compat = c.model.flags & WTB_FLAG_COMPAT_F32;
    launch_optics(c, cshift);
    const cd *d = c.dip.p + (size_t)comp * N;
    const cd *X = c.sol.evecs.p;
    // amplitudes and energies
    DBuf<cd> a_dev;
    a_dev.alloc(nev);
    hipLaunchKernelGGL(amplitude_kernel, dim3(nev), dim3(kRT), 0, c.stream, N, X, d, textbook ? 1 : 0, a_dev.p);
    WTB_LAUNCH_CHECK();
    std::vector<cd> a(nev);
    std::vector<double> E(nev);
    WTB_HIP_CHECK(hipMemcpyAsync(a.data(), a_dev.p, sizeof(cd) * nev, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipMemcpyAsync(E.data(), c.sol.evals.p, sizeof(double) * nev, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    // spectrum, loss, Ebar_n and fbar_n
    const double *a1 = c.model.rlat, *a2 = c.model.rlat + 3, *a3 = c.model.rlat + 6;
    const double cr[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
    const double vol = std::fabs(cr[0] * a3[0] + cr[1] * a3[1] + cr[2] * a3[2]);
    const double cpre = (compat ? (double)180.90f : 180.90) / ((double)c.ngrid[0] * c.ngrid[1] * c.ngrid[2] * vol);
    std::vector<double> f(nev), ebar(nev, 0.0), fbar(nev, 0.0);
    for (int m = 0; m < nev; ++m) f[m] = a[m].x * a[m].x + a[m].y * a[m].y;
    double L = 0.0;
    for (int j = 0; j < nw; ++j) {
        double e2 = 0.0;
        for (int m = 0; m < nev; ++m) {
            const double x = omega[j] - E[m];
            e2 += f[m] * gamma / (x * x + gamma * gamma);
        }
        e2 *= cpre;
        if (eps2) eps2[j] = e2;
        const double r = e2 - target[j];
        L += r * r;
        for (int m = 0; m < nev; ++m) {
            const double x = omega[j] - E[m], den = x * x + gamma * gamma;
            fbar[m] += 2.0 * r * cpre * gamma / den;
            ebar[m] += 2.0 * r * cpre * f[m] * 2.0 * x * gamma / (den * den);   // d/dE_n of 1/den
        }
    }
    if (loss) *loss = L;
    // g_n and h coefficients
    std::vector<cd> cg(nev), chh(nev);
    for (int m = 0; m < nev; ++m) {
        chh[m] = make_double2(fbar[m] * a[m].x, fbar[m] * a[m].y);
        cg[m] = textbook ? make_double2(fbar[m] * a[m].x, -fbar[m] * a[m].y) : chh[m];
    }
    DBuf<cd> coef, G, Lm, Y, R, Z, P, Q, Bm, Cx, Tm, hvec, Xall;
    DBuf<double> Ed, dots, dots2;
    coef.alloc(2 * (size_t)nev);
    Ed.alloc(nev);
    dots.alloc(nev);
    dots2.alloc(nev);
    WTB_HIP_CHECK(hipMemcpyAsync(coef.p, cg.data(), sizeof(cd) * nev, hipMemcpyHostToDevice, c.stream));
    WTB_HIP_CHECK(hipMemcpyAsync(coef.p + nev, chh.data(), sizeof(cd) * nev, hipMemcpyHostToDevice, c.stream));
    WTB_HIP_CHECK(hipMemcpyAsync(Ed.p, E.data(), sizeof(double) * nev, hipMemcpyHostToDevice, c.stream));
    const size_t NE = (size_t)N * nev;
    G.alloc(NE);
    Lm.alloc(NE);
    Y.alloc(NE);
    R.alloc(NE);
    Z.alloc(NE);
    P.alloc(NE);
    Q.alloc(NE);
    Bm.alloc((size_t)nev * nev);
    Cx.alloc((size_t)nev * nev);
    Tm.alloc((size_t)nev * nev);
    const dim3 gN((unsigned)((N + 255) / 256), nev);
    hipLaunchKernelGGL(scaled_dipole_kernel, gN, dim3(256), 0, c.stream, N, d, textbook ? 0 : 1, coef.p, G.p);
    WTB_LAUNCH_CHECK();
    // explicit part over the other solved states: l_n += sum_{m != n} x_m (x_m^H g_n) / (E_n - E_m)
    gram_tall(c, nev, nev, N, X, (int)N, G.p, (int)N, Bm.p, nev);
    std::vector<cd> hb((size_t)nev * nev), hc((size_t)nev * nev);
    WTB_HIP_CHECK(hipMemcpyAsync(hb.data(), Bm.p, sizeof(cd) * nev * nev, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int nn_ = 0; nn_ < nev; ++nn_)
        for (int m = 0; m < nev; ++m) {
            const double de = E[nn_] - E[m];
            cd v = make_double2(0.0, 0.0);
            if (m != nn_ && std::fabs(de) > 1e-10 * std::max(1.0, std::fabs(E[nn_])))
                v = make_double2(hb[m + (size_t)nn_ * nev].x / de, hb[m + (size_t)nn_ * nev].y / de);
            hc[m + (size_t)nn_ * nev] = v;
        }
    WTB_HIP_CHECK(hipMemcpyAsync(Cx.p, hc.data(), sizeof(cd) * nev * nev, hipMemcpyHostToDevice, c.stream));
    gemm_tall(c, N, nev, nev, X, N, Cx.p, nev, Lm.p, N);
    // complement: (H - E_n) y = -P g_n, y in the complement of the solved states (PCG)
    auto project = [&](cd *V) {
        gram_tall(c, nev, nev, N, X, (int)N, V, (int)N, Tm.p, nev);
        gemm_tall_sub(c, N, nev, nev, X, N, Tm.p, nev, V, N);
    };
    auto coldots = [&](const cd *A, const cd *B2, std::vector<double> &h) {
        hipLaunchKernelGGL(coldot_kernel, dim3(nev), dim3(kRT), 0, c.stream, N, A, B2, dots.p);
        WTB_LAUNCH_CHECK();
        h.resize(nev);
        WTB_HIP_CHECK(hipMemcpyAsync(h.data(), dots.p, sizeof(double) * nev, hipMemcpyDeviceToHost, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    };
    // (H - E_n) y = P g_n on the complement: l_n = explicit part - y
    WTB_HIP_CHECK(hipMemcpyAsync(R.p, G.p, sizeof(cd) * NE, hipMemcpyDeviceToDevice, c.stream));
    project(R.p);
    WTB_HIP_CHECK(hipMemsetAsync(Y.p, 0, sizeof(cd) * NE, c.stream));
    // host copies of the per-column scalars stay alive until the next stream sync
    std::vector<double> bnorm, rr, rz, rzn, pq, al(nev), be(nev);
    coldots(R.p, R.p, bnorm);
    hipLaunchKernelGGL(jacobi_prec_kernel, gN, dim3(256), 0, c.stream, N, c.bse.diag.p, Ed.p, R.p, Z.p);
    WTB_LAUNCH_CHECK();
    project(Z.p);
    WTB_HIP_CHECK(hipMemcpyAsync(P.p, Z.p, sizeof(cd) * NE, hipMemcpyDeviceToDevice, c.stream));
    coldots(R.p, Z.p, rz);
    std::vector<char> done(nev, 0);
    for (int m = 0; m < nev; ++m) done[m] = bnorm[m] == 0.0;
    const double tol = 1e-10;
    for (int it = 0; it < 4000; ++it) {
        bool all = true;
        for (int m = 0; m < nev; ++m) all = all && done[m];
        if (all) break;
        matvec(c, WTB_FP64, nev, P.p, Q.p);
        hipLaunchKernelGGL(shift_kernel, gN, dim3(256), 0, c.stream, N, Ed.p, P.p, Q.p);
        WTB_LAUNCH_CHECK();
        project(Q.p);
        coldots(P.p, Q.p, pq);
        for (int m = 0; m < nev; ++m) {
            if (!done[m] && !(pq[m] > 0.0))
                throw Error{WTB_ERR_NOCONV, "spectrum adjoint: H - E_n is not positive on the complement of the "
                                            "solved states (are they the lowest, with a gap above?)"};
            al[m] = done[m] ? 0.0 : rz[m] / pq[m];
        }
        WTB_HIP_CHECK(hipMemcpyAsync(dots2.p, al.data(), sizeof(double) * nev, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(cg_step_kernel, gN, dim3(256), 0, c.stream, N, dots2.p, P.p, Q.p, Y.p, R.p);
        WTB_LAUNCH_CHECK();
        coldots(R.p, R.p, rr);
        for (int m = 0; m < nev; ++m)
            if (!done[m] && rr[m] <= tol * tol * bnorm[m]) done[m] = 1;
        hipLaunchKernelGGL(jacobi_prec_kernel, gN, dim3(256), 0, c.stream, N, c.bse.diag.p, Ed.p, R.p, Z.p);
        WTB_LAUNCH_CHECK();
        project(Z.p);
        coldots(R.p, Z.p, rzn);
        for (int m = 0; m < nev; ++m) be[m] = (done[m] || rz[m] == 0.0) ? 0.0 : rzn[m] / rz[m];
        rz = rzn;
        WTB_HIP_CHECK(hipMemcpyAsync(dots2.p, be.data(), sizeof(double) * nev, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(cg_dir_kernel, gN, dim3(256), 0, c.stream, N, dots2.p, Z.p, P.p);
        WTB_LAUNCH_CHECK();
    }
    for (int m = 0; m < nev; ++m)
        if (!done[m]) throw Error{WTB_ERR_NOCONV, "spectrum adjoint: complement solve did not converge"};
    hipLaunchKernelGGL(sub_kernel, dim3((unsigned)((NE + 255) / 256)), dim3(256), 0, c.stream, NE, Y.p, Lm.p);
    WTB_LAUNCH_CHECK();
    // polarisation: columns [x, x + t l, l] with weights [Ebar - 1/t, 1/t, -t]
    std::vector<double> ln2;
    coldots(Lm.p, Lm.p, ln2);
    std::vector<double> t(nev), w(3 * (size_t)nev);
    for (int m = 0; m < nev; ++m) {
        t[m] = ln2[m] > 0.0 ? 1.0 / std::sqrt(ln2[m]) : 1.0;
        w[m] = ebar[m] - 1.0 / t[m];
        w[nev + m] = 1.0 / t[m];
        w[2 * nev + m] = -t[m];
        if (ln2[m] == 0.0) {    // no eigenvector term for this state
            w[m] = ebar[m];
            w[nev + m] = 0.0;
            w[2 * nev + m] = 0.0;
        }
    }
    WTB_HIP_CHECK(hipMemcpyAsync(dots2.p, t.data(), sizeof(double) * nev, hipMemcpyHostToDevice, c.stream));
    Xall.alloc(3 * NE);
    hipLaunchKernelGGL(polar_kernel, gN, dim3(256), 0, c.stream, N, nev, X, Lm.p, dots2.p, Xall.p);
    WTB_LAUNCH_CHECK();
    // dipole adjoint
    hvec.alloc(N);
    hipLaunchKernelGGL(combine_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.stream, N, nev, X,
                       textbook ? 0 : 1, coef.p + nev, hvec.p);
    WTB_LAUNCH_CHECK();
    DBuf<double> eb;
    DBuf<cd> ub, db;
    eb.alloc((size_t)nk * nb);
    ub.alloc((size_t)nb * n * nk);
    db.alloc((size_t)nk * n * n);
    const int ncv = b.nc * b.nv, nvec = c.model.nvec;
    const size_t shmem = ((size_t)n * n + (size_t)nb * n + (size_t)(b.nv + b.nc) * n + nvec + ncv) * sizeof(cd) +
                         2 * (size_t)ncv * sizeof(double) + (size_t)ncv * sizeof(cd);
    if (shmem > 160 * 1024) throw Error{WTB_ERR_UNSUPPORTED, "too many R vectors for the dipole adjoint"};
    const double eta = compat ? (double)(float)cshift : cshift;
    hipLaunchKernelGGL(dipole_adjoint_kernel, dim3(nk), dim3(kDL), shmem, c.stream, nk, n, nvec, b.nv, b.nc,
                       compat ? 1 : 0, eta, comp, c.kpt.p, c.model.hop.p, c.model.rcart.p, c.model.invdeg.p,
                       b.esel.p, b.usel.p, hvec.p, eb.p, ub.p, db.p);
    WTB_LAUNCH_CHECK();
    GradExtra ex;
    ex.ebar = eb.p;
    ex.ubar = ub.p;
    ex.dbar = db.p;
    ex.alpha = comp;
    grad_core(c, Xall.p, 3 * nev, w.data(), &ex, out, g_re, g_im);
}

}  // namespace wtb
