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
