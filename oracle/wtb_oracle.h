/*
 * CPU FP64 oracle for the WTB exciton (BSE) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This restates the reference algorithm in plain C
 * so that tests, the smoke check and the CPU-baseline leg of the benchmark can
 * check the HIP library against it.  Nothing in the product library links,
 * loads or calls this code.
 *
 * Conventions (all row-major C arrays, complex = interleaved double pairs):
 *   rlat[3*i + c]      component c of lattice vector a_{i+1}  (reference rlat(i,c))
 *   rvec[3*r + c]      lattice coordinate c of R_r
 *   hop[(r*norb + m)*norb + n] = H_mn(R_r)  (complex)
 *   kpt[3*k + c]       Cartesian k-point (1/Angstrom)
 *   H(k) and eigenvector matrices: a[m*n + j] is row m, column j.
 *   Selected bands per k (reference eigv/vector, module_bse.F90:2230-2245):
 *     eig[k*nb + j], vec[(k*nb + j)*norb + o], j = 0..nv-1 valence
 *     (bands nocp-nv .. nocp-1), j = nv..nv+nc-1 conduction.
 *   BSE basis (quantumnumbers2, module_math.F90:352-379):
 *     i = k*nc*nv + c*nv + v.
 */
#ifndef WTB_ORACLE_H
#define WTB_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* flags */
#define OR_COMPAT_F32   1   /* float32 rounding where the reference uses default-kind cmplx() */
#define OR_GAUGE_LAPACK 2   /* rotate eigenvectors into the ZHEEV('V','L') phase gauge */

/* Coulomb potential kinds (module_bse.F90:49-72) */
#define OR_V2DK  1
#define OR_V3D   2
#define OR_V3DL  3
#define OR_V2DT  4
#define OR_V2DT2 5
#define OR_V2DOH 6
#define OR_V2DRK 7
#define OR_V0DT  8
#define OR_V1DT  9

typedef struct {
    int kind;
    double r0;
    double ediel[3];   /* EDIEL_T, EDIEL, EDIEL_B */
    double ez, w, lc, ktol;
} or_coulomb;

typedef struct {
    int norb, nvec;
    const int *rvec;
    const int *deg;
    const double complex *hop;
    double rlat[9];
    double efermi, scissor;
    int flags;
} or_model;

/* constants as the reference evaluates them (single-precision literals) */
double or_pi_f32(void);
double or_cic(void);
void or_bessel_k01(double x, double *k0, double *k1);

void or_recvec(const double *rlat, double *blat);
void or_mp_mesh(int n1, int n2, int n3, const double *shift, const double *rlat, double *kpt);
double or_vcell2d(const double *rlat);
double or_vcell3d(const double *rlat);

void or_hk(const or_model *m, const double *k, double complex *h);
void or_dhk(const or_model *m, const double *k, double complex *hx, double complex *hy,
            double complex *hz);

int or_heev_jacobi(int n, double complex *a, double *w, double complex *z);
void or_householder_q(int n, const double complex *a, double complex *q);
void or_gauge_lapack(int n, const double complex *h, double complex *z);

int or_eigsys(const or_model *m, const double *k, double *e, double complex *z, int *nocp);
int or_hk_eig(const or_model *m, int nk, const double *kpt, int nv, int nc,
              double *eig, double complex *vec, int *nocp, double *eall, double complex *zall);

double or_coulomb_v(const or_coulomb *c, const double *k1, const double *k2,
                    const int *ngrid, const double *rlat);

int or_bse_matrix(int nk, int nc, int nv, int norb, const double *kpt, const double *eig,
                  const double complex *vec, const or_coulomb *c, const int *ngrid,
                  const double *rlat, int textbook, double complex *h);

int or_dipoles(const or_model *m, int nk, const double *kpt, int nc, int nv,
               const double *eig, const double complex *vec, double cshift,
               double complex *d);

void or_oscillator(int n, int nev, const double complex *x, const double complex *d1,
                   const double complex *d2, double *act);

void or_optdiel(int rtype, double vc, int n, const int *ngrid, double elux,
                const double *ex, const double *fosc, double sme, double *re, double *im);
void or_imagrenorm(int n, int nelux, const double *ex, double *dielf);
void or_spectrum(int n, const double *ex, const double *fosc, int diagonal, const int *ngrid,
                 double vc, double e0, double e1, int npts, double sme, int renorm,
                 double *out /* npts x 3 */);

int or_bse_rows(int nk, int nc, int nv, int norb, const double *kpt, const double *eig,
                const double complex *vec, const or_coulomb *c, const int *ngrid,
                const double *rlat, long i0, long ni, double complex *h);
void or_matvec_rows(long ni, long N, int ncol, const double complex *h, const double complex *x,
                    double complex *y);
int or_num_threads(void);

/* Krylov-Schur (thick-restart Lanczos) for the nev eigenpairs of least |E| of
 * the dense Hermitian h (row-major N x N), with slepc_solver's settings
 * (module_slepc_solver.F90:29-74): ncv = nev + mpd, relative tolerance tol,
 * max_it restarts.  e (nev), x (N x nev column-major, may be NULL) in order
 * of |E|.  Returns the converged count; its = restarts, nmv = products. */
int or_krylov_schur(long N, const double complex *h, int nev, double tol, int max_it, unsigned seed,
                    double *e, double complex *x, int *its, long *nmv);

#ifdef __cplusplus
}
#endif
#endif
