// Density of states (dostool, src/module_tools.F90:390-651) on the bands of
// wtb_hk_eig.  Every state (k, band n) carries four weights:
//   tdos = 1;
//   spin up / down (spincont, :41-99): sum of |z_o|^2 over the first / second
//     half of the orbitals (SP, and SOC with DFT=V), over the odd / even
//     orbitals (other SOC), zero for NP;
//   layer weight sum_o w_o |z_o|^2 (layercont, :153-175).
// Each curve is (1/nk) sum_states weight * gaussian(E - E_state, sigma)
// (endos :103-149; gaussian, module_math.F90:330-339, whose pi is the
// single-precision acos(-1.)) on the grid E_i = e0 + (e1 - e0)(i - 1)/(ndos - 1)
// with e0 = min(0, lowest band), e1 = max(0, highest band) (dostool :465-490).
// One workgroup per energy point sums over all states; the state weights are
// formed once from the eigenvectors.
#include "wtb_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace wtb {
namespace {

constexpr int kDosThreads = 256;

__global__ void dos_weights_kernel(int nk, int n, int spin_mode, const double *__restrict__ ow,
                                   const cd *__restrict__ uall, double4 *__restrict__ w) {
    const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= (long)nk * n) return;
    const int k = (int)(s / n), j = (int)(s % n), h = n / 2;
    const cd *z = uall + (size_t)k * n * n;
    double up = 0.0, dn = 0.0, lw = 0.0;
    for (int o = 0; o < n; ++o) {
        const cd c = z[(size_t)o * n + j];
        const double p = c.x * c.x + c.y * c.y;
        lw += ow ? ow[o] * p : p;
        if (spin_mode == 1) {
            if (o < h) up += p;
            else if (o < 2 * h) dn += p;
        } else if (spin_mode == 2 && o < 2 * h) {
            if ((o & 1) == 0) up += p;
            else dn += p;
        }
    }
    w[s] = make_double4(1.0, up, dn, lw);
}

__global__ void __launch_bounds__(kDosThreads)
dos_kernel(long ns, double inv_nk, double e0, double e1, double ndos, double sigma,
           const double *__restrict__ e, const double4 *__restrict__ w, double *__restrict__ en,
           double *__restrict__ out, int ld) {
    __shared__ double4 red[kDosThreads];
    const int i = blockIdx.x;
    const double x = e0 + ((e1 - e0) * (double)i) / (ndos - 1.0);
    const double pi = (double)3.14159265358979323846f;
    const double norm = 1.0 / (sigma * std::sqrt(2.0 * pi)), deno = 2.0 * sigma * sigma;
    double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
    for (long s = threadIdx.x; s < ns; s += kDosThreads) {
        const double d = x - e[s];
        const double g = norm * std::exp(-(d * d) / deno);
        const double4 ws = w[s];
        acc.x += ws.x * g;
        acc.y += ws.y * g;
        acc.z += ws.z * g;
        acc.w += ws.w * g;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kDosThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            double4 a = red[threadIdx.x], b = red[threadIdx.x + off];
            red[threadIdx.x] = make_double4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double4 r = red[0];
        en[i] = x;
        out[i] = r.x * inv_nk;
        out[i + ld] = r.y * inv_nk;
        out[i + 2 * (size_t)ld] = r.z * inv_nk;
        out[i + 3 * (size_t)ld] = r.w * inv_nk;
    }
}

}  // namespace

void launch_dos(Ctx &c, int ndos, double sigma, const double *orbweight, int spin_mode, double *en_host,
                double *dos_host) {
    Bands &b = c.bands;
    if (!b.valid) throw Error{WTB_ERR_STATE, "wtb_hk_eig must run before wtb_dos"};
    if (ndos < 2) throw Error{WTB_ERR_ARG, "ndos must be at least 2"};
    if (!(sigma > 0.0)) throw Error{WTB_ERR_ARG, "sigma must be positive"};
    if (spin_mode < 0 || spin_mode > 2) throw Error{WTB_ERR_ARG, "spin_mode must be 0, 1 or 2"};
    const int nk = b.nk, n = c.model.norb;
    const long ns = (long)nk * n;
    // energy window: dostool starts both ends at 0 and widens them to the bands
    std::vector<double> eh(ns);
    WTB_HIP_CHECK(hipMemcpyAsync(eh.data(), b.eall.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    double e0 = 0.0, e1 = 0.0;
    for (int k = 0; k < nk; ++k) {
        e0 = std::min(e0, eh[(size_t)k * n]);
        e1 = std::max(e1, eh[(size_t)k * n + n - 1]);
    }
    DBuf<double4> w;
    DBuf<double> ow, en, out;
    w.alloc(ns);
    en.alloc(ndos);
    out.alloc((size_t)4 * ndos);
    if (orbweight) {
        ow.alloc(n);
        WTB_HIP_CHECK(hipMemcpyAsync(ow.p, orbweight, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    }
    hipLaunchKernelGGL(dos_weights_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c.stream, nk, n,
                       spin_mode, orbweight ? (const double *)ow.p : nullptr, b.uall.p, w.p);
    WTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(dos_kernel, dim3((unsigned)ndos), dim3(kDosThreads), 0, c.stream, ns, 1.0 / (double)nk, e0,
                       e1, (double)ndos, sigma, b.eall.p, w.p, en.p, out.p, ndos);
    WTB_LAUNCH_CHECK();
    if (en_host)
        WTB_HIP_CHECK(hipMemcpyAsync(en_host, en.p, sizeof(double) * ndos, hipMemcpyDeviceToHost, c.stream));
    if (dos_host)
        WTB_HIP_CHECK(hipMemcpyAsync(dos_host, out.p, sizeof(double) * 4 * ndos, hipMemcpyDeviceToHost, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace wtb
