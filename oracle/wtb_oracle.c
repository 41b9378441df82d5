/*
 * CPU FP64 oracle for the WTB exciton (BSE) path -- TEST INFRASTRUCTURE ONLY.
 * Only tests, the smoke check and the benchmark's CPU-baseline leg use it.
 *
 * Each function names the reference routine it restates (file:line in the
 * WTB sources).  Where the reference evaluates an expression in default
 * (single) precision, the OR_COMPAT_F32 flag reproduces that rounding.
 */
/* j0, j1 of the C library are POSIX (X/Open) functions */
#define _XOPEN_SOURCE 700
#include "wtb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static inline double f32(double x) { return (double)(float)x; }

/* acos(-1.) evaluated in real(4): module_coulomb.F90:15, module_kmesh.F90:392 */
double or_pi_f32(void) { return (double)acosf(-1.0f); }

/* cic = -(0.0904756)*10**(3) evaluated in real(4): module_coulomb.F90:15 */
double or_cic(void) { return (double)(-(0.0904756f * 1000.0f)); }

static void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* recvec, module_kmesh.F90:391-407 */
void or_recvec(const double *rlat, double *blat) {
    double v23[3], v31[3], v12[3];
    cross3(rlat + 3, rlat + 6, v23);
    cross3(rlat + 6, rlat, v31);
    cross3(rlat, rlat + 3, v12);
    double vol = fabs(rlat[0] * v23[0] + rlat[1] * v23[1] + rlat[2] * v23[2]);
    double f = (2.0 * or_pi_f32()) / vol;
    for (int c = 0; c < 3; ++c) {
        blat[c] = f * v23[c];
        blat[3 + c] = f * v31[c];
        blat[6 + c] = f * v12[c];
    }
}

/* monhkhorst_pack, module_kmesh.F90:338-386: i outer, l inner, no folding */
void or_mp_mesh(int n1, int n2, int n3, const double *shift, const double *rlat, double *kpt) {
    double b[9], ks[3];
    or_recvec(rlat, b);
    for (int c = 0; c < 3; ++c) ks[c] = b[c] * shift[0] + b[3 + c] * shift[1] + b[6 + c] * shift[2];
    int cnt = 0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int l = 0; l < n3; ++l) {
                for (int c = 0; c < 3; ++c)
                    kpt[3 * cnt + c] = (b[c] / n1) * i + (b[3 + c] / n2) * j + (b[6 + c] / n3) * l + ks[c];
                ++cnt;
            }
}

/* vcell2D / vcell3D, module_math.F90:143-162 */
double or_vcell2d(const double *rlat) {
    double v[3];
    cross3(rlat, rlat + 3, v);
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

double or_vcell3d(const double *rlat) {
    double v[3];
    cross3(rlat, rlat + 3, v);
    return fabs(v[0] * rlat[6] + v[1] * rlat[7] + v[2] * rlat[8]);
}

/* Cartesian R of lattice vector r, summed in the reference's order */
static void rcart(const or_model *m, int r, double *R) {
    const int *n = m->rvec + 3 * r;
    for (int c = 0; c < 3; ++c)
        R[c] = (n[0] * m->rlat[c] + n[1] * m->rlat[3 + c]) + n[2] * m->rlat[6 + c];
}

static double kphase(const or_model *m, const double *k, const double *R) {
    double x = (k[0] * R[0] + k[1] * R[1]) + k[2] * R[2];
    return (m->flags & OR_COMPAT_F32) ? f32(x) : x;
}

/* H(k) = sum_R e^{i k.R} H(R)/deg(R) - E_F, eigsys + fermilvl,
 * module_hamiltonian.F90:241-254,480-488 */
void or_hk(const or_model *m, const double *k, double complex *h) {
    int n = m->norb;
    memset(h, 0, sizeof(double complex) * n * n);
    for (int r = 0; r < m->nvec; ++r) {
        double R[3];
        rcart(m, r, R);
        double ph = kphase(m, k, R);
        double complex e = (cos(ph) + I * sin(ph)) * (1.0 / (double)m->deg[r]);
        const double complex *hr = m->hop + (size_t)r * n * n;
        for (int i = 0; i < n * n; ++i) h[i] += e * hr[i];
    }
    double ef = (m->flags & OR_COMPAT_F32) ? f32(m->efermi) : m->efermi;
    for (int i = 0; i < n; ++i) h[i * n + i] -= ef;
}

/* dH/dk_alpha = sum_R (i R_alpha) e^{i k.R} H(R)/deg(R), hlm, module_optics.F90:12-108.
 * The onsite diagonal of H(R=0) is dropped as in the reference (it carries R=0 anyway). */
void or_dhk(const or_model *m, const double *k, double complex *hx, double complex *hy,
            double complex *hz) {
    int n = m->norb;
    double complex *out[3] = {hx, hy, hz};
    for (int a = 0; a < 3; ++a) memset(out[a], 0, sizeof(double complex) * n * n);
    for (int r = 0; r < m->nvec; ++r) {
        double R[3];
        rcart(m, r, R);
        double ph = kphase(m, k, R);
        double complex e = (cos(ph) + I * sin(ph)) * (1.0 / (double)m->deg[r]);
        const int *nr = m->rvec + 3 * r;
        int onsite = (nr[0] == 0 && nr[1] == 0 && nr[2] == 0);
        const double complex *hr = m->hop + (size_t)r * n * n;
        for (int a = 0; a < 3; ++a) {
            double Ra = (m->flags & OR_COMPAT_F32) ? f32(R[a]) : R[a];
            double complex f = (I * Ra) * e;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    if (onsite && i == j) continue;
                    out[a][i * n + j] += f * hr[i * n + j];
                }
        }
    }
}

/* Cyclic two-sided Jacobi for a Hermitian matrix (textbook algorithm).
 * On return w holds eigenvalues ascending and z[m*n + j] component m of
 * eigenvector j.  a is destroyed.  Returns the number of sweeps or -1. */
int or_heev_jacobi(int n, double complex *a, double *w, double complex *z) {
    for (int i = 0; i < n * n; ++i) z[i] = 0;
    for (int i = 0; i < n; ++i) z[i * n + i] = 1;
    double norm = 0;
    for (int i = 0; i < n * n; ++i) norm += creal(a[i] * conj(a[i]));
    int sweep;
    for (sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += creal(a[p * n + q] * conj(a[p * n + q]));
        if (off <= 1e-30 * norm || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double complex b = a[p * n + q];
                double ab = cabs(b);
                if (ab == 0.0) continue;
                double complex ph = b / ab;               /* e^{i phi} */
                double app = creal(a[p * n + p]), aqq = creal(a[q * n + q]);
                double theta = (aqq - app) / (2.0 * ab);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                double complex phc = conj(ph);
                /* J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]]; A <- J^H A J, Z <- Z J */
                for (int r = 0; r < n; ++r) {
                    double complex arp = a[r * n + p], arq = a[r * n + q];
                    a[r * n + p] = c * arp - s * phc * arq;
                    a[r * n + q] = s * arp + c * phc * arq;
                    double complex zrp = z[r * n + p], zrq = z[r * n + q];
                    z[r * n + p] = c * zrp - s * phc * zrq;
                    z[r * n + q] = s * zrp + c * phc * zrq;
                }
                for (int r = 0; r < n; ++r) {
                    double complex apr = a[p * n + r], aqr = a[q * n + r];
                    a[p * n + r] = c * apr - s * ph * aqr;
                    a[q * n + r] = s * apr + c * ph * aqr;
                }
                a[p * n + q] = 0;
                a[q * n + p] = 0;
                a[p * n + p] = creal(a[p * n + p]);
                a[q * n + q] = creal(a[q * n + q]);
            }
    }
    for (int i = 0; i < n; ++i) w[i] = creal(a[i * n + i]);
    /* ascending order (stable selection) */
    for (int i = 0; i < n - 1; ++i) {
        int mi = i;
        for (int j = i + 1; j < n; ++j)
            if (w[j] < w[mi]) mi = j;
        if (mi != i) {
            double t = w[i]; w[i] = w[mi]; w[mi] = t;
            for (int r = 0; r < n; ++r) {
                double complex tz = z[r * n + i];
                z[r * n + i] = z[r * n + mi];
                z[r * n + mi] = tz;
            }
        }
    }
    return sweep < 60 ? sweep : -1;
}

/* Q = H(1) H(2) ... H(n-1) of the unblocked lower Householder reduction
 * (LAPACK ZHETD2 with UPLO='L', reflectors generated as in ZLARFG), i.e. the
 * orthogonal factor that ZHEEV('V','L') applies before its real tridiagonal QL. */
void or_householder_q(int n, const double complex *h, double complex *q) {
    double complex *a = (double complex *)malloc(sizeof(double complex) * n * n);
    double complex *v = (double complex *)malloc(sizeof(double complex) * n);
    double complex *y = (double complex *)malloc(sizeof(double complex) * n);
    memcpy(a, h, sizeof(double complex) * n * n);
    for (int i = 0; i < n * n; ++i) q[i] = 0;
    for (int i = 0; i < n; ++i) q[i * n + i] = 1;
    for (int i = 0; i < n - 1; ++i) {
        double complex alpha = a[(i + 1) * n + i];
        double xn2 = 0;
        for (int r = i + 2; r < n; ++r) xn2 += creal(a[r * n + i] * conj(a[r * n + i]));
        double alphr = creal(alpha), alphi = cimag(alpha);
        if (xn2 == 0.0 && alphi == 0.0) continue;               /* tau = 0 */
        double nrm = sqrt(alphr * alphr + alphi * alphi + xn2);
        double beta = alphr >= 0 ? -nrm : nrm;
        double complex tau = (beta - alphr) / beta - I * (alphi / beta);
        double complex scal = 1.0 / (alpha - beta);
        for (int r = 0; r < n; ++r) v[r] = 0;
        v[i + 1] = 1;
        for (int r = i + 2; r < n; ++r) v[r] = a[r * n + i] * scal;
        /* A <- H^H A H = A - tau y v^H - conj(tau) v y^H + |tau|^2 mu v v^H, y = A v */
        double mu = 0;
        for (int r = 0; r < n; ++r) {
            double complex s = 0;
            for (int c = i + 1; c < n; ++c) s += a[r * n + c] * v[c];
            y[r] = s;
        }
        for (int r = i + 1; r < n; ++r) mu += creal(conj(v[r]) * y[r]);
        double t2 = creal(tau * conj(tau));
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c)
                a[r * n + c] += -tau * y[r] * conj(v[c]) - conj(tau) * v[r] * conj(y[c])
                                + t2 * mu * v[r] * conj(v[c]);
        /* Q <- Q H(i) */
        for (int r = 0; r < n; ++r) {
            double complex s = 0;
            for (int c = i + 1; c < n; ++c) s += q[r * n + c] * v[c];
            for (int c = i + 1; c < n; ++c) q[r * n + c] -= tau * s * conj(v[c]);
        }
    }
    free(a); free(v); free(y);
}

/* Rotate each eigenvector so that Q^H z is real with its largest component
 * positive: the phase gauge of ZHEEV('V','L') (which returns Q times the real
 * eigenvectors of the real tridiagonal matrix), up to the real sign. */
void or_gauge_lapack(int n, const double complex *h, double complex *z) {
    double complex *q = (double complex *)malloc(sizeof(double complex) * n * n);
    or_householder_q(n, h, q);
    for (int j = 0; j < n; ++j) {
        double best = -1;
        double complex ph = 1;
        for (int m = 0; m < n; ++m) {
            double complex s = 0;
            for (int r = 0; r < n; ++r) s += conj(q[r * n + m]) * z[r * n + j];
            double a2 = creal(s * conj(s));
            if (a2 > best) { best = a2; ph = s; }
        }
        double complex f = conj(ph) / cabs(ph);
        for (int r = 0; r < n; ++r) z[r * n + j] *= f;
    }
    free(q);
}

/* eigsys, module_hamiltonian.F90:200-320: ascending eigenpairs of H(k),
 * nocp = (index of first E > 0) - 1, scissor added to bands above nocp.
 * nocp is left untouched when no eigenvalue is positive (as the reference). */
int or_eigsys(const or_model *m, const double *k, double *e, double complex *z, int *nocp) {
    int n = m->norb;
    double complex *h = (double complex *)malloc(sizeof(double complex) * n * n);
    double complex *a = (double complex *)malloc(sizeof(double complex) * n * n);
    or_hk(m, k, h);
    memcpy(a, h, sizeof(double complex) * n * n);
    int sw = or_heev_jacobi(n, a, e, z);
    if (m->flags & OR_GAUGE_LAPACK) or_gauge_lapack(n, h, z);
    for (int i = 0; i < n; ++i)
        if (e[i] > 0.0) { *nocp = i; break; }
    if (*nocp >= 0)
        for (int i = *nocp; i < n; ++i) e[i] += m->scissor;
    free(h); free(a);
    return sw < 0 ? -3 : 0;
}

/* k loop + band selection of bsesolver, module_bse.F90:2230-2245, with the
 * quantumnumbers2 checks (module_math.F90:362-368). */
int or_hk_eig(const or_model *m, int nk, const double *kpt, int nv, int nc,
              double *eig, double complex *vec, int *nocp, double *eall, double complex *zall) {
    int n = m->norb, nb = nc + nv, status = 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nk; ++k) {
        double *e = (double *)malloc(sizeof(double) * n);
        double complex *z = (double complex *)malloc(sizeof(double complex) * n * n);
        int nocc = -1;
        int st = or_eigsys(m, kpt + 3 * k, e, z, &nocc);
        if (st) status = st;
        nocp[k] = nocc;
        if (nocc < 0 || nv > nocc || nc > n - nocc) {
            status = -4;
        } else {
            for (int j = 0; j < nb; ++j) {
                int band = nocc - nv + j;
                eig[(size_t)k * nb + j] = e[band];
                for (int o = 0; o < n; ++o) vec[((size_t)k * nb + j) * n + o] = z[o * n + band];
            }
        }
        if (eall) memcpy(eall + (size_t)k * n, e, sizeof(double) * n);
        if (zall) memcpy(zall + (size_t)k * n * n, z, sizeof(double complex) * n * n);
        free(e); free(z);
    }
    return status;
}

/* --------------------------- Coulomb potentials --------------------------- */

static double vnorm(const double *k1, const double *k2) {
    double dx = k1[0] - k2[0], dy = k1[1] - k2[1], dz = k1[2] - k2[2];
    return sqrt(dx * dx + dy * dy + dz * dz);                  /* modvec, module_math.F90:221 */
}

/* K0(x), K1(x), x > 0 (the reference's calck0/calck1, Cody's rational
 * approximations, module_math.F90:453+): here from K_n(x) = int_0^inf
 * exp(-x cosh t) cosh(n t) dt by the trapezoidal rule with step 1/20, which
 * converges geometrically for this analytic integrand; pinned against
 * scipy.special.k0/k1 in the tests. */
void or_bessel_k01(double x, double *k0, double *k1) {
    const double h = 0.05;
    double s0 = 0.5 * exp(-x), s1 = s0;
    for (int j = 1; j < 100000; ++j) {
        double ch = cosh(h * j), a = x * ch;
        if (a > 745.0) break;
        s0 += exp(-a);
        s1 += exp(-a) * ch;
    }
    *k0 = h * s0;
    *k1 = h * s1;
}

/* integralJ0pot, module_math.F90:429-449: 100-point trapezoid of r J0(g r) ln r
 * on r_i = (i - 1)/99 rmax + 1e-4 (a default-real literal) */
static double or_j0_log_integral(double rmax, double g) {
    double r[100], f[100], res = 0.0;
    for (int i = 0; i < 100; ++i) {
        r[i] = ((double)i / 99.0) * rmax + (double)1e-4f;
        f[i] = r[i] * j0(r[i] * g) * log(r[i]);
    }
    double h = r[1] - r[0];
    for (int i = 0; i < 99; ++i) res = res + (h / 2.0) * (f[i] + f[i + 1]);
    return res;
}

double or_coulomb_v(const or_coulomb *c, const double *k1, const double *k2,
                    const int *ngrid, const double *rlat) {
    const double cic = or_cic(), pi = or_pi_f32();
    double nk = (double)(ngrid[0] * ngrid[1] * ngrid[2]);
    double q = vnorm(k1, k2);
    double dk[3] = {k1[0] - k2[0], k1[1] - k2[1], k1[2] - k2[2]};
    switch (c->kind) {
    case OR_V2DK: { /* v2dk, module_coulomb.F90:8-60 */
        double a0 = 0.5 * (sqrt(rlat[0] * rlat[0] + rlat[1] * rlat[1] + rlat[2] * rlat[2])
                           + sqrt(rlat[3] * rlat[3] + rlat[4] * rlat[4] + rlat[5] * rlat[5]));
        double vc = or_vcell2d(rlat);
        double r0 = ((c->ediel[1] - 1.0) * c->lc) / (c->ediel[0] + c->ediel[2]);
        double vbz = 1.0 / (nk * vc);
        double g = (double)(ngrid[0] * ngrid[1]);
        double auxi = (2.0 * pi * r0) / (a0 * sqrt(g));
        double ed = (c->ediel[0] + c->ediel[2]) / 2.0;
        if (q < c->ktol)
            return vbz * (cic / ed) * (a0 * sqrt(g) / (2.0 * pi)) * ((double)1.76f + auxi * 1.0);
        return vbz * (cic / ed) * (1.0 / (q * (1 + r0 * q)));
    }
    case OR_V3D: /* vcoul, module_coulomb.F90:64-94 */
    case OR_V3DL: { /* v3diel, module_coulomb.F90:98-131 */
        double vbz = 1.0 / (nk * or_vcell3d(rlat));
        double ed = (c->kind == OR_V3D) ? 1.0 : c->ediel[1];
        if (q < c->ktol) return 0.0;
        return vbz * (cic / ed) * (1.0 / (q * q));
    }
    case OR_V2DT: { /* v2dt, module_coulomb.F90:135-193 */
        double vbz = 1.0 / (nk * or_vcell3d(rlat));
        double gz = fabs(dk[2]), gpar = sqrt(dk[0] * dk[0] + dk[1] * dk[1]);
        double rc = 0.5 * rlat[8];
        if (gpar < c->ktol && gz < c->ktol) return (vbz * cic) * (-0.5 * rc * rc);
        if (gpar < c->ktol)
            return (vbz * cic) * (1.0 / (q * q)) * (1.0 - cos(gz * rc) - (gz * rc * sin(gz * rc)));
        double a1 = gz / gpar, a2 = gpar * rc, a3 = gz * rc;
        return (vbz * cic) * (1.0 / (q * q)) * (1.0 + exp(-a2) * (a1 * sin(a3) - cos(a3)));
    }
    case OR_V0DT: { /* v0dt, module_coulomb.F90:197-238 */
        double vbz = 1.0 / (nk * or_vcell3d(rlat));
        double s1 = sqrt(rlat[0] * rlat[0] + rlat[1] * rlat[1] + rlat[2] * rlat[2]);
        double s2 = sqrt(rlat[3] * rlat[3] + rlat[4] * rlat[4] + rlat[5] * rlat[5]);
        double s3 = sqrt(rlat[6] * rlat[6] + rlat[7] * rlat[7] + rlat[8] * rlat[8]);
        double cr = 0.5 * fmin(s1, fmin(s2, s3));
        if (q < c->ktol) return (vbz * cic) * 0.5 * cr * cr;
        return (vbz * cic) * (1.0 / (q * q)) * (1.0 - cos(cr * q));
    }
    case OR_V2DT2: { /* v2dt2, module_coulomb.F90:243-282 */
        double vbz = 1.0 / (nk * or_vcell3d(rlat));
        double qxy = sqrt(dk[0] * dk[0] + dk[1] * dk[1]);
        if (q < c->ktol) return 0.0;
        return (vbz * cic) * (1.0 / (q * q)) * (1.0 - exp(-0.5 * qxy * c->lc) * cos(0.5 * c->lc * dk[2]));
    }
    case OR_V2DRK: { /* v2drk, module_coulomb.F90:286-350 */
        double vbz = 1.0 / (nk * or_vcell2d(rlat));
        double epar = c->ediel[1], et = c->ediel[0], eb = c->ediel[2];
        double eta = sqrt(epar / c->ez), kappa = sqrt(epar * c->ez);
        double pb = (eb - kappa) / (eb + kappa), pt = (et - kappa) / (et + kappa);
        if (q < c->ktol) return 0.0;
        double a1 = (1.0 - (pb * pt * exp(-2.0 * q * eta * c->lc))) * kappa;
        double a2 = (1.0 - (pt * exp(-eta * q * c->lc))) * (1.0 - (pb * exp(-eta * q * c->lc)));
        double a3 = c->r0 * q * exp(-q * c->w);
        double ew = (a1 / a2) + a3;
        return (vbz * cic) * exp(-q * c->w) * (1.0 / ew) * (1.0 / q);
    }
    case OR_V2DOH: { /* v2dohono, module_coulomb.F90:354-387 */
        double vbz = 1.0 / (nk * or_vcell2d(rlat));
        if (q < c->ktol) return 0.0;
        return (vbz * cic) * exp(-c->w * q) * (1.0 / (c->ez * q));
    }
    case OR_V1DT: { /* v1dt, module_coulomb.F90:391-442 */
        double vbz = 1.0 / (nk * or_vcell3d(rlat));
        double qx = fabs(dk[0]), qyz = sqrt(dk[1] * dk[1] + dk[2] * dk[2]);
        if (qx < c->ktol && qyz >= c->ktol) return (vbz * cic) * or_j0_log_integral(c->lc, qyz);
        if (qx < c->ktol) return (vbz * cic) * ((c->lc * c->lc) / 16.0) * (2.0 * log(c->lc) - 1.0);
        double kk0, kk1;
        or_bessel_k01(qx * c->lc, &kk0, &kk1);
        double c1 = qyz * c->lc * j1(qyz * c->lc) * kk0;
        double c2 = qx * c->lc * j0(qyz * c->lc) * kk1;
        return (vbz * cic) * (1.0 + c1 + c2);
    }
    default:
        return NAN;
    }
}

/* -------------------------------- assembly -------------------------------- */

/* matrizelbse + assembly loop, module_bse.F90:24-83,2392-2415.  h is row-major
 * N x N; the reference computes j >= i and mirrors the conjugate. */
int or_bse_matrix(int nk, int nc, int nv, int norb, const double *kpt, const double *eig,
                  const double complex *vec, const or_coulomb *c, const int *ngrid,
                  const double *rlat, int textbook, double complex *h) {
    int nb = nc + nv;
    long N = (long)nk * nc * nv;
    if (c->kind < OR_V2DK || c->kind > OR_V1DT) return -7;
#pragma omp parallel for schedule(dynamic, 8)
    for (long i = 0; i < N; ++i) {
        int ki = (int)(i / (nc * nv)), ci = (int)((i / nv) % nc), vi = (int)(i % nv);
        const double complex *cvi = vec + ((size_t)ki * nb + nv + ci) * norb;
        const double complex *vvi = vec + ((size_t)ki * nb + vi) * norb;
        for (long j = i; j < N; ++j) {
            int kj = (int)(j / (nc * nv)), cj = (int)((j / nv) % nc), vj = (int)(j % nv);
            double v = or_coulomb_v(c, kpt + 3 * ki, kpt + 3 * kj, ngrid, rlat);
            double complex val;
            if (i == j) {
                val = (eig[(size_t)ki * nb + nv + ci] - eig[(size_t)ki * nb + vi]) + v;
            } else {
                const double complex *cvj = vec + ((size_t)kj * nb + nv + cj) * norb;
                const double complex *vvj = vec + ((size_t)kj * nb + vj) * norb;
                double complex sc = 0, sv = 0;
                for (int o = 0; o < norb; ++o) sc += conj(cvi[o]) * cvj[o];
                if (textbook)
                    for (int o = 0; o < norb; ++o) sv += conj(vvj[o]) * vvi[o];
                else
                    for (int o = 0; o < norb; ++o) sv += conj(vvi[o]) * vvj[o];
                val = v * sc * sv;
            }
            h[i * N + j] = val;
            h[j * N + i] = conj(val);
        }
    }
    return 0;
}

/* --------------------------------- optics --------------------------------- */

/* optsp loop of bsesolver, module_bse.F90:2317-2331, module_optics.F90:113-167:
 * d_alpha(i) = <c|dH/dk_alpha|v> * Re(1/(E_c - E_v + i*eta)); eta is rounded
 * to real(4) by cmplx(0.0,sme).  d[a*N + i] for a = x, y, z, sp, sm. */
int or_dipoles(const or_model *m, int nk, const double *kpt, int nc, int nv,
               const double *eig, const double complex *vec, double cshift,
               double complex *d) {
    int n = m->norb, nb = nc + nv;
    long N = (long)nk * nc * nv;
    int compat = m->flags & OR_COMPAT_F32;
    double eta = compat ? f32(cshift) : cshift;
    double isq2 = compat ? (double)(1.0f / sqrtf(2.0f)) : 1.0 / sqrt(2.0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nk; ++k) {
        double complex *hd = (double complex *)malloc(sizeof(double complex) * 3 * n * n);
        double complex *t = (double complex *)malloc(sizeof(double complex) * n);
        or_dhk(m, kpt + 3 * k, hd, hd + n * n, hd + 2 * n * n);
        for (int c = 0; c < nc; ++c)
            for (int v = 0; v < nv; ++v) {
                long i = (long)k * nc * nv + c * nv + v;
                const double complex *cv = vec + ((size_t)k * nb + nv + c) * n;
                const double complex *vv = vec + ((size_t)k * nb + v) * n;
                double complex esp = (eig[(size_t)k * nb + nv + c] - eig[(size_t)k * nb + v]) + I * eta;
                double factor = creal(1.0 / esp);
                double complex r[3];
                for (int a = 0; a < 3; ++a) {
                    const double complex *hm = hd + a * n * n;
                    for (int o = 0; o < n; ++o) {
                        double complex s = 0;
                        for (int p = 0; p < n; ++p) s += hm[o * n + p] * vv[p];
                        t[o] = s;
                    }
                    double complex s = 0;
                    for (int o = 0; o < n; ++o) s += conj(cv[o]) * t[o];
                    r[a] = factor * s;
                    d[a * N + i] = r[a];
                }
                d[3 * N + i] = (r[0] + I * r[1]) * isq2;
                d[4 * N + i] = (r[0] - I * r[1]) * isq2;
            }
        free(hd); free(t);
    }
    return 0;
}

/* dielbsev / dielbsep, module_bse.F90:171-245: act(n) = Re[(sum_j X_jn d1_j)
 * conj(sum_j X_jn d2_j)] with no conjugate on X; x is column-major n x nev. */
void or_oscillator(int n, int nev, const double complex *x, const double complex *d1,
                   const double complex *d2, double *act) {
    for (int i = 0; i < n; ++i) act[i] = 0;
#pragma omp parallel for
    for (int e = 0; e < nev; ++e) {
        double complex s1 = 0, s2 = 0;
        for (int j = 0; j < n; ++j) {
            s1 += x[(size_t)e * n + j] * d1[j];
            s2 += x[(size_t)e * n + j] * d2[j];
        }
        act[e] = creal(s1 * conj(s2));
    }
}

/* avgsme, module_optics.F90:313-335 (the differences are summed one by one) */
static double avgsme(int n, const double *ex) {
    double s = 0.0;
    for (int i = 1; i < n; ++i) s = s + ex[i] - ex[i - 1];
    return 10.0 * (s / (double)n);
}

/* optdiel, module_optics.F90:171-255 */
void or_optdiel(int rtype, double vc, int n, const int *ngrid, double elux,
                const double *ex, const double *fosc, double sme, double *re, double *im) {
    double baux = (double)(float)(ngrid[0] * ngrid[1] * ngrid[2]) * vc;
    double caux = (double)180.90f / (baux * 1.0);
    double delta = (rtype == 1) ? 1.0 : 0.0;
    double smef = avgsme(n, ex) + sme;
    double rp = 0, ip = 0;
    for (int i = 0; i < n; ++i) {
        double eaux = (elux - ex[i]) * (elux - ex[i]) + smef * smef;
        rp += fosc[i] * ((ex[i] - elux) / eaux);
        ip += fosc[i] * (smef / eaux);
    }
    *im = ip * caux;
    *re = delta + rp * caux;
}

/* imagrenorm, module_optics.F90:338-401; dielf is row-major nelux x 3 */
void or_imagrenorm(int n, int nelux, const double *ex, double *dielf) {
    (void)n;
    double aux = ex[0] - (double)0.15f;
    if (aux <= 0.0) return;
    int ib = 0;
    for (int i = 0; i < nelux; ++i) {
        if (dielf[3 * i] <= aux) ib = i;
        else break;
    }
    if (ib + 1 >= nelux) {
        /* every grid point lies below E_1 - 0.15: the reference zeroes them all
           (its read of point ib+1 is past the grid and never used) */
        for (int i = 0; i < nelux; ++i) dielf[3 * i + 2] = 0.0;
        return;
    }
    double aux3 = dielf[3 * (ib + 1) + 2];
    for (int i = 0; i <= ib; ++i) dielf[3 * i + 2] = 0.0;
    for (int i = ib + 1; i < nelux; ++i) {
        dielf[3 * i + 2] -= aux3;
        if (dielf[3 * i + 2] < 0.0) dielf[3 * i + 2] = 0.0;
    }
}

/* bsedielraw energy loop for one tensor component, module_bse.F90:936-1009 */
void or_spectrum(int n, const double *ex, const double *fosc, int diagonal, const int *ngrid,
                 double vc, double e0, double e1, int npts, double sme, int renorm, double *out) {
    for (int j = 0; j < npts; ++j) {
        double elux = e0 + ((e1 - e0) * j) / ((double)npts - 1.0);
        double re, im;
        or_optdiel(diagonal ? 1 : 0, vc, n, ngrid, elux, ex, fosc, sme, &re, &im);
        out[3 * j] = elux;
        out[3 * j + 1] = re;
        out[3 * j + 2] = im;
    }
    if (renorm) or_imagrenorm(n, npts, ex, out);
}

/* Rows i0 .. i0+ni-1 of the BSE matrix (row-major ni x N), the same element
 * formula as or_bse_matrix; used to time a bounded sample of the assembly. */
int or_bse_rows(int nk, int nc, int nv, int norb, const double *kpt, const double *eig,
                const double complex *vec, const or_coulomb *c, const int *ngrid,
                const double *rlat, long i0, long ni, double complex *h) {
    int nb = nc + nv;
    long N = (long)nk * nc * nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (long ii = 0; ii < ni; ++ii) {
        long i = i0 + ii;
        int ki = (int)(i / (nc * nv)), ci = (int)((i / nv) % nc), vi = (int)(i % nv);
        const double complex *cvi = vec + ((size_t)ki * nb + nv + ci) * norb;
        const double complex *vvi = vec + ((size_t)ki * nb + vi) * norb;
        for (long j = 0; j < N; ++j) {
            int kj = (int)(j / (nc * nv)), cj = (int)((j / nv) % nc), vj = (int)(j % nv);
            double v = or_coulomb_v(c, kpt + 3 * ki, kpt + 3 * kj, ngrid, rlat);
            double complex val;
            if (i == j) {
                val = (eig[(size_t)ki * nb + nv + ci] - eig[(size_t)ki * nb + vi]) + v;
            } else {
                const double complex *cvj = vec + ((size_t)kj * nb + nv + cj) * norb;
                const double complex *vvj = vec + ((size_t)kj * nb + vj) * norb;
                double complex sc = 0, sv = 0;
                for (int o = 0; o < norb; ++o) sc += conj(cvi[o]) * cvj[o];
                for (int o = 0; o < norb; ++o) sv += conj(vvi[o]) * vvj[o];
                val = v * sc * sv;
            }
            h[ii * N + j] = val;
        }
    }
    return 0;
}

/* Y (ni x ncol, row-major) = H_rows (ni x N, row-major) X (N x ncol, row-major):
 * the dense FP64 matvec of a CPU eigensolver, OpenMP over rows. */
void or_matvec_rows(long ni, long N, int ncol, const double complex *h, const double complex *x,
                    double complex *y) {
    /* explicit real arithmetic: C99 complex '*' would go through the
       NaN-checking library multiply */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < ni; ++i) {
        double *yi = (double *)(y + i * ncol);
        for (int c = 0; c < 2 * ncol; ++c) yi[c] = 0.0;
        const double *hi = (const double *)(h + i * N);
        for (long j = 0; j < N; ++j) {
            const double hr = hi[2 * j], hm = hi[2 * j + 1];
            const double *xj = (const double *)(x + j * ncol);
            for (int c = 0; c < ncol; ++c) {
                const double xr = xj[2 * c], xm = xj[2 * c + 1];
                yi[2 * c] += hr * xr - hm * xm;
                yi[2 * c + 1] += hr * xm + hm * xr;
            }
        }
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------- Krylov-Schur (CPU baseline) --------------------- */

/* y = h x for one vector (row-major N x N h), four independent partial sums
 * per row so the loop is bound by the stream of h, not by one add chain. */
static void mv1(long N, const double complex *h, const double complex *x, double complex *y) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        const double *hi = (const double *)(h + i * N);
        const double *xx = (const double *)x;
        double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
        long j = 0;
        for (; j + 4 <= N; j += 4)
            for (int u = 0; u < 4; ++u) {
                const double hr = hi[2 * (j + u)], hm = hi[2 * (j + u) + 1];
                const double xr = xx[2 * (j + u)], xm = xx[2 * (j + u) + 1];
                ar[u] += hr * xr - hm * xm;
                ai[u] += hr * xm + hm * xr;
            }
        for (; j < N; ++j) {
            const double hr = hi[2 * j], hm = hi[2 * j + 1], xr = xx[2 * j], xm = xx[2 * j + 1];
            ar[0] += hr * xr - hm * xm;
            ai[0] += hr * xm + hm * xr;
        }
        y[i] = ((ar[0] + ar[1]) + (ar[2] + ar[3])) + I * ((ai[0] + ai[1]) + (ai[2] + ai[3]));
    }
}

/* c[0..m) = V[:, 0..m)^H w, V column-major (ldv = N) */
static void proj(long N, int m, const double complex *V, const double complex *w, double complex *c) {
#pragma omp parallel for schedule(static)
    for (int q = 0; q < m; ++q) {
        const double complex *v = V + (size_t)q * N;
        double sr = 0, si = 0;
        for (long i = 0; i < N; ++i) {
            const double vr = creal(v[i]), vi = cimag(v[i]), wr = creal(w[i]), wi = cimag(w[i]);
            sr += vr * wr + vi * wi;
            si += vr * wi - vi * wr;
        }
        c[q] = sr + I * si;
    }
}

/* w -= V[:, 0..m) c */
static void sub(long N, int m, const double complex *V, const double complex *c, double complex *w) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        double complex s = 0;
        for (int q = 0; q < m; ++q) s += V[(size_t)q * N + i] * c[q];
        w[i] -= s;
    }
}

static double nrm(long N, const double complex *w) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < N; ++i) s += creal(w[i]) * creal(w[i]) + cimag(w[i]) * cimag(w[i]);
    return sqrt(s);
}

/* The nev eigenpairs of least |E| of a dense Hermitian matrix h (row-major
 * N x N) by the Krylov-Schur method as slepc_solver configures it
 * (module_slepc_solver.F90:29-74): EPS_HEP, EPS_SMALLEST_MAGNITUDE, ncv =
 * nev + mpd (mpd = nev below 1000, nev/5 above), relative tolerance tol on
 * ||H x - l x|| / |l| (SLEPc's EPS_CONV_REL), max_it restarts.  For a
 * Hermitian matrix Krylov-Schur is thick-restart Lanczos: the basis is kept
 * orthonormal by two classical Gram-Schmidt passes per step, the projected
 * matrix is diagonalised at every restart (Jacobi), its Ritz pairs are sorted
 * by |theta|, the residual of pair i is |beta s_last,i|, and a restart keeps
 * the converged pairs plus half of the others (SLEPc's default keep = 0.5)
 * with the last Lanczos vector coupled to them in arrowhead form.  The start
 * vector is a seeded random vector.
 * Outputs: e (nev) and x (N x nev column-major) in order of |E|; *its
 * restarts; *nmv matrix-vector products.  Returns the number of converged
 * pairs (nev on success). */
int or_krylov_schur(long N, const double complex *h, int nev, double tol, int max_it, unsigned seed,
                    double *e, double complex *x, int *its, long *nmv) {
    const int mpd = nev < 1000 ? nev : nev / 5;
    int ncv = nev + mpd;
    if (ncv > N) ncv = (int)N;
    double complex *V = (double complex *)malloc(sizeof(double complex) * (size_t)N * (ncv + 1));
    double complex *T = (double complex *)calloc((size_t)ncv * ncv, sizeof(double complex));
    double complex *A = (double complex *)malloc(sizeof(double complex) * (size_t)ncv * ncv);
    double complex *S = (double complex *)malloc(sizeof(double complex) * (size_t)ncv * ncv);
    double complex *c1 = (double complex *)malloc(sizeof(double complex) * (ncv + 1));
    double complex *c2 = (double complex *)malloc(sizeof(double complex) * (ncv + 1));
    double complex *Vn = (double complex *)malloc(sizeof(double complex) * (size_t)N * ncv);
    double *th = (double *)malloc(sizeof(double) * ncv);
    int *ord = (int *)malloc(sizeof(int) * ncv);
    /* seeded start vector (xorshift) */
    unsigned s = seed ? seed : 1u;
    for (long i = 0; i < N; ++i) {
        s ^= s << 13; s ^= s >> 17; s ^= s << 5;
        double a = (double)s / 4294967296.0 - 0.5;
        s ^= s << 13; s ^= s >> 17; s ^= s << 5;
        double b = (double)s / 4294967296.0 - 0.5;
        V[i] = a + I * b;
    }
    double n0 = nrm(N, V);
    for (long i = 0; i < N; ++i) V[i] /= n0;
    int k = 0, nconv = 0, it;
    long mv = 0;
    double beta = 0;
    for (it = 0; it < max_it; ++it) {
        /* expand the basis from column k to ncv; T[:, j] from the projections */
        for (int j = k; j < ncv; ++j) {
            double complex *w = V + (size_t)(j + 1) * N;
            mv1(N, h, V + (size_t)j * N, w);
            ++mv;
            proj(N, j + 1, V, w, c1);
            sub(N, j + 1, V, c1, w);
            proj(N, j + 1, V, w, c2);
            sub(N, j + 1, V, c2, w);
            for (int q = 0; q <= j; ++q) c1[q] += c2[q];
            /* column j of the projected matrix; rows below k of the kept part
               are the arrowhead, set at the restart */
            for (int q = (j == k ? k : j - 1); q <= j; ++q) {
                T[q * ncv + j] = (q == j) ? creal(c1[q]) : c1[q];
                T[j * ncv + q] = conj(T[q * ncv + j]);
            }
            if (j == k)
                for (int q = 0; q < k; ++q) {      /* arrowhead row from the kept pairs */
                    T[q * ncv + j] = c1[q];
                    T[j * ncv + q] = conj(c1[q]);
                }
            beta = nrm(N, w);
            if (beta == 0.0) beta = 1e-300;
            for (long i = 0; i < N; ++i) w[i] /= beta;
            if (j + 1 < ncv) {
                T[j * ncv + j + 1] = beta;
                T[(j + 1) * ncv + j] = beta;
            }
        }
        /* Ritz pairs, ordered by |theta| (EPS_SMALLEST_MAGNITUDE) */
        memcpy(A, T, sizeof(double complex) * (size_t)ncv * ncv);
        or_heev_jacobi(ncv, A, th, S);
        for (int i = 0; i < ncv; ++i) ord[i] = i;
        for (int i = 1; i < ncv; ++i) {          /* insertion sort by |theta| (stable) */
            int t = ord[i], q = i;
            while (q > 0 && fabs(th[ord[q - 1]]) > fabs(th[t])) { ord[q] = ord[q - 1]; --q; }
            ord[q] = t;
        }
        nconv = 0;
        for (int i = 0; i < ncv && nconv < nev; ++i) {
            const double res = fabs(beta) * cabs(S[(size_t)(ncv - 1) * ncv + ord[i]]);
            if (res < tol * fabs(th[ord[i]])) ++nconv;
            else break;
        }
        if (nconv >= nev || it == max_it - 1) break;
        /* restart: keep the converged pairs and half of the rest */
        int keep = nconv + (int)(0.5 * (ncv - nconv));
        if (keep >= ncv) keep = ncv - 1;
        if (keep < 1) keep = 1;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; ++i)
            for (int q = 0; q < keep; ++q) {
                double complex acc = 0;
                for (int r = 0; r < ncv; ++r) acc += V[(size_t)r * N + i] * S[(size_t)r * ncv + ord[q]];
                Vn[(size_t)q * N + i] = acc;
            }
        memcpy(V, Vn, sizeof(double complex) * (size_t)N * keep);
        memmove(V + (size_t)keep * N, V + (size_t)ncv * N, sizeof(double complex) * N);
        memset(T, 0, sizeof(double complex) * (size_t)ncv * ncv);
        for (int q = 0; q < keep; ++q) T[q * ncv + q] = th[ord[q]];
        /* the coupling of the kept pairs to the continuation vector enters
           through the projections of its first product (arrowhead row) */
        k = keep;
    }
    const int nout = nconv < nev ? nconv : nev;
    for (int q = 0; q < nev; ++q) {
        e[q] = th[ord[q]];
        if (x) {
#pragma omp parallel for schedule(static)
            for (long i = 0; i < N; ++i) {
                double complex acc = 0;
                for (int r = 0; r < ncv; ++r) acc += V[(size_t)r * N + i] * S[(size_t)r * ncv + ord[q]];
                x[(size_t)q * N + i] = acc;
            }
        }
    }
    *its = it + 1;
    *nmv = mv;
    free(V); free(T); free(A); free(S); free(c1); free(c2); free(Vn); free(th); free(ord);
    return nout;
}
