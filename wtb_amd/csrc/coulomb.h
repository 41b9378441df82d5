// Reciprocal-space Coulomb potentials V(k - k') of the reference
// (src/module_coulomb.F90), evaluated on the device.  q = |k - k'| is the
// plain Cartesian norm of the unfolded difference (modvec,
// src/module_math.F90:221-225); below KTOL the reference's regularised value
// is used.  Constants are the reference's single-precision literals:
// cic = -(0.0904756)*10**3 and pi = acos(-1.) in real(4).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/wtb_hip.h"

namespace wtb {

struct CoulombDev {
    int kind;
    double ktol, r0, et, epar, eb, ez, w, lc;
    // precomputed per model / mesh
    double vbz2d;     // 1/(N_k * |a1 x a2|)
    double vbz3d;     // 1/(N_k * |(a1 x a2).a3|)
    double cic;
    // V2DRK
    double eta, kappa, pb, pt;
    // V2DK
    double v2dk_r0, v2dk_ed, v2dk_q0, v2dk_pref;
    double v2dk_a0sg;  // a0 sqrt(n1 n2), for the parameter derivatives
    double pi;         // float32 acos(-1.)
    // V2DT / V0DT
    double rc, cr;
};

// Modified Bessel functions K0(x), K1(x) (x > 0) from the integral
// representation K_n(x) = int_0^inf exp(-x cosh t) cosh(n t) dt, summed by the
// trapezoidal rule with step 1/5.  The integrand is analytic in the strip
// |Im t| < pi/2, so the rule converges like exp(-pi^2 / h) ~ 1e-21 relative;
// the sum stops where x cosh t passes 745 (exp underflows there).
__host__ __device__ inline void bessel_k01(double x, double &k0, double &k1) {
    const double h = 0.2;
    const double e0 = exp(-x);
    double s0 = 0.5 * e0, s1 = 0.5 * e0;
    for (int j = 1; j < 1024; ++j) {
        const double ch = cosh(h * (double)j), a = x * ch;
        if (a > 745.0) break;
        const double e = exp(-a);
        s0 += e;
        s1 += e * ch;
    }
    k0 = h * s0;
    k1 = h * s1;
}

// int r J0(g r) ln r dr by the reference's 100-point trapezoid on
// r_i = (i - 1)/99 rmax + 1e-4 (integralJ0pot, src/module_math.F90:429-449;
// eta is a single-precision literal there)
__host__ __device__ inline double j0_log_integral(double rmax, double g) {
    const int n = 100;
    const double eta = (double)1e-4f;
    const double r1 = (0.0 / (n - 1.0)) * rmax + eta, r2 = (1.0 / (n - 1.0)) * rmax + eta;
    const double h = r2 - r1;
    double res = 0.0, fprev = r1 * j0(r1 * g) * log(r1);
    for (int i = 2; i <= n; ++i) {
        const double r = (((double)i - 1.0) / (n - 1.0)) * rmax + eta;
        const double f = r * j0(r * g) * log(r);
        res = res + (h / 2.0) * (fprev + f);
        fprev = f;
    }
    return res;
}

__host__ __device__ inline double coulomb_v(const CoulombDev &c, double dx, double dy, double dz) {
    const double q = sqrt(dx * dx + dy * dy + dz * dz);
    switch (c.kind) {
    case WTB_V2DK:   // v2dk, module_coulomb.F90:8-60
        if (q < c.ktol) return c.v2dk_q0;
        return c.v2dk_pref * (1.0 / (q * (1 + c.v2dk_r0 * q)));
    case WTB_V3D:    // vcoul, module_coulomb.F90:64-94
        if (q < c.ktol) return 0.0;
        return c.vbz3d * c.cic * (1.0 / (q * q));
    case WTB_V3DL:   // v3diel, module_coulomb.F90:98-131
        if (q < c.ktol) return 0.0;
        return c.vbz3d * (c.cic / c.epar) * (1.0 / (q * q));
    case WTB_V2DT: { // v2dt, module_coulomb.F90:135-193
        const double gz = fabs(dz), gpar = sqrt(dx * dx + dy * dy);
        if (gpar < c.ktol && gz < c.ktol) return (c.vbz3d * c.cic) * (-0.5 * c.rc * c.rc);
        if (gpar < c.ktol)
            return (c.vbz3d * c.cic) * (1.0 / (q * q)) *
                   (1.0 - cos(gz * c.rc) - (gz * c.rc * sin(gz * c.rc)));
        const double a1 = gz / gpar, a2 = gpar * c.rc, a3 = gz * c.rc;
        return (c.vbz3d * c.cic) * (1.0 / (q * q)) * (1.0 + exp(-a2) * (a1 * sin(a3) - cos(a3)));
    }
    case WTB_V0DT:   // v0dt, module_coulomb.F90:197-238
        if (q < c.ktol) return (c.vbz3d * c.cic) * 0.5 * c.cr * c.cr;
        return (c.vbz3d * c.cic) * (1.0 / (q * q)) * (1.0 - cos(c.cr * q));
    case WTB_V1DT: { // v1dt, module_coulomb.F90:391-442: q_x against |q_yz|, length lc
        const double qx = fabs(dx), qyz = sqrt(dy * dy + dz * dz), pre = c.vbz3d * c.cic;
        if (qx < c.ktol && qyz >= c.ktol) return pre * j0_log_integral(c.lc, qyz);
        if (qx < c.ktol) return pre * ((c.lc * c.lc) / 16.0) * (2.0 * log(c.lc) - 1.0);
        double k0, k1;
        bessel_k01(qx * c.lc, k0, k1);
        const double c1 = qyz * c.lc * j1(qyz * c.lc) * k0;
        const double c2 = qx * c.lc * j0(qyz * c.lc) * k1;
        return pre * (1.0 + c1 + c2);
    }
    case WTB_V2DT2: {// v2dt2, module_coulomb.F90:243-282
        if (q < c.ktol) return 0.0;
        const double qxy = sqrt(dx * dx + dy * dy);
        return (c.vbz3d * c.cic) * (1.0 / (q * q)) * (1.0 - exp(-0.5 * qxy * c.lc) * cos(0.5 * c.lc * dz));
    }
    case WTB_V2DRK: {// v2drk, module_coulomb.F90:286-350
        if (q < c.ktol) return 0.0;
        const double a1 = (1.0 - (c.pb * c.pt * exp(-2.0 * q * c.eta * c.lc))) * c.kappa;
        const double e1 = exp(-c.eta * q * c.lc);
        const double a2 = (1.0 - (c.pt * e1)) * (1.0 - (c.pb * e1));
        const double ew = (a1 / a2) + c.r0 * q * exp(-q * c.w);
        return (c.vbz2d * c.cic) * exp(-q * c.w) * (1.0 / ew) * (1.0 / q);
    }
    case WTB_V2DOH:  // v2dohono, module_coulomb.F90:354-387
        if (q < c.ktol) return 0.0;
        return (c.vbz2d * c.cic) * exp(-c.w * q) * (1.0 / (c.ez * q));
    default:
        return 0.0;
    }
}

// Host-side preparation of the constants (shared by the assembly and adjoint).
CoulombDev make_coulomb(const wtb_coulomb &in, const double *rlat, const int *ngrid);

}  // namespace wtb
