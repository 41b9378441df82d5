/*
 * wtb_hip.h -- C-ABI of the MI355X (gfx950) exciton / Bethe-Salpeter path.
 *
 * The functions replace the hot calls of the reference's bsesolver
 * (src/module_bse.F90:2023-2594).  Host arrays use the reference's Fortran
 * layouts (column-major, 1-based in Fortran, complex(8) as interleaved double
 * pairs), so a Fortran caller binds them with ISO_C_BINDING and passes its
 * own arrays unchanged.  Device state is owned by the opaque context and
 * stays resident between calls: the BSE matrix never crosses PCIe.
 *
 * Every call returns 0 on success or a negative WTB_ERR_* code; the message
 * of the last failure is available from wtb_last_error().  One context per
 * host thread; calls on a context are serialised on its HIP stream.
 */
#ifndef WTB_HIP_H
#define WTB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define WTB_OK               0
#define WTB_ERR_ARG         -1   /* bad argument or shape                   */
#define WTB_ERR_HIP         -2   /* HIP runtime / kernel launch failure      */
#define WTB_ERR_NOCONV      -3   /* eigensolver did not converge             */
#define WTB_ERR_NOCP        -4   /* band selection impossible (E_F / nv / nc) */
#define WTB_ERR_NOMEM       -5   /* device allocation failed                 */
#define WTB_ERR_STATE       -6   /* call out of order (e.g. solve before build) */
#define WTB_ERR_UNSUPPORTED -7   /* option not available on this path        */

/* model flags */
#define WTB_FLAG_COMPAT_F32   1  /* round k.R, R_alpha, E_F, eta, 1/sqrt2 to float32 as the reference */
#define WTB_FLAG_GAUGE_LAPACK 2  /* eigenvector phases of ZHEEV('V','L') (up to a real sign) */

/* Coulomb potentials (COULOMB_POT card, module_bse.F90:49-72) */
#define WTB_V2DK  1
#define WTB_V3D   2
#define WTB_V3DL  3
#define WTB_V2DT  4
#define WTB_V2DT2 5
#define WTB_V2DOH 6
#define WTB_V2DRK 7
#define WTB_V0DT  8
#define WTB_V1DT  9              /* build and solve; the gradients return WTB_ERR_UNSUPPORTED */

/* precision of the BSE kernel store and the Krylov matvecs */
#define WTB_FP64   0             /* complex FP64 store, FP64 MFMA matvec                     */
#define WTB_MIXED  1             /* complex FP32 store (FP64 diagonal), split-FP16 MFMA matvec
                                    with FP32 accumulation, FP64 orthogonalisation /
                                    Rayleigh-Ritz / final residual refinement               */
#define WTB_FP32   2             /* complex FP32 store, FP32 MFMA matvec (no refinement)     */
#define WTB_INT8X  3             /* matvec only (wtb_bse_matvec on the FP32 store): exact int8
                                    digit products with FP64 recombination, ~2^-28 relative to
                                    max|H| max|X|; the refinement product of WTB_MIXED        */

/* kernel overlap convention */
#define WTB_CONV_REFERENCE 0     /* V <c_i|c_j><v_i|v_j>, module_bse.F90:77-81 */
#define WTB_CONV_TEXTBOOK  1     /* V <c_i|c_j><v_j|v_i>                        */

/* optical components for wtb_oscillator */
#define WTB_X  0
#define WTB_Y  1
#define WTB_Z  2
#define WTB_SP 3
#define WTB_SM 4

typedef struct wtb_ctx wtb_ctx;

/* Coulomb parameters: input cards COULOMB_POT, R_0, EDIEL_T/EDIEL/EDIEL_B,
 * EDIEL_Z, W_COUL, LC, KTOL (module_input.F90:55-122). */
typedef struct {
    int kind;
    double r0;
    double ediel[3];   /* EDIEL_T, EDIEL, EDIEL_B */
    double ez, w, lc, ktol;
} wtb_coulomb;

/* Filled by wtb_bse_solve. */
typedef struct {
    int iterations;          /* outer block iterations                       */
    int nconv;               /* converged pairs among the nev requested       */
    long matvec_columns;     /* columns pushed through H*X (all precisions)   */
    double max_rel_residual; /* max_j ||H x_j - l_j x_j|| / |l_j| (FP64)       */
    double t_solve_ms;       /* eigensolver wall time on the stream           */
    double t_refine_ms;      /* FP64 refinement part of t_solve_ms            */
    double t_matvec_ms;      /* time inside H*X kernels                       */
    double t_rr_ms;          /* the rest: orthogonalisation, Rayleigh-Ritz,
                                residual norms (t_solve_ms - t_matvec_ms)      */
    /* The solve's dominant matvec kernel (split-FP16 MFMA for WTB_MIXED, FP64
     * MFMA for WTB_FP64, FP32 MFMA for WTB_FP32), timed with HIP events on the
     * library stream around each launch: */
    double t_kernel_ms;      /* summed device time of its launches            */
    int kernel_launches;     /* number of launches                            */
    double kernel_bytes;     /* algorithmic HBM bytes: the store, once a launch */
    double kernel_flops;     /* 8 N^2 per column of X                         */
    /* The FP64-arithmetic refinement kernel of the mixed path (zero for FP64): */
    double t_refine_kernel_ms;
    int refine_launches;
    double refine_flops;
} wtb_report;

/* Context. device: HIP ordinal.  Replaces the implicit module state of
 * module_bse.F90:9-15 (hbse, Ws, HE). */
int wtb_create(wtb_ctx **ctx, int device);
int wtb_destroy(wtb_ctx *ctx);
const char *wtb_last_error(const wtb_ctx *ctx);

/* TB model, replaces hamiltonian_input_read (module_hamiltonian.F90:35-194).
 *   rvec  : int    rvec(nvec,3)           lattice coordinates of R
 *   deg   : int    ffactor(nvec)          Wigner-Seitz degeneracies
 *   hop_re, hop_im : double hopmatrices(nvec,norb,norb), ihopmatrices(...)
 *   rlat  : double rlat(3,3), rlat(i,:) = a_i  (Angstrom)
 *   efermi, scissor : eV;  flags : WTB_FLAG_* */
int wtb_set_tb_model(wtb_ctx *ctx, int norb, int nvec, const int *rvec, const int *deg,
                     const double *hop_re, const double *hop_im, const double *rlat,
                     double efermi, double scissor, int flags);

/* k-points, replaces monhkhorst_pack's output (module_kmesh.F90:338-386).
 *   kpt : double kpt(nk,3) Cartesian (1/Angstrom);  ngrid : the MP grid
 *   (its product normalises the Coulomb kernel, module_coulomb.F90:316). */
int wtb_set_kpoints(wtb_ctx *ctx, int nk, const double *kpt, const int *ngrid);

/* Batched H(k) build + Hermitian eigensolve for every k, band selection and
 * scissor: replaces the eigsys loop of bsesolver (module_bse.F90:2230-2245,
 * module_hamiltonian.F90:200-320).  Outputs (each may be NULL):
 *   eig  : double eigv(nk, nc+nv)
 *   vec  : complex(8) vector(nk, nc+nv, norb)
 *   nocp : int nocpk(nk) */
int wtb_hk_eig(wtb_ctx *ctx, int nv, int nc, double *eig, double *vec, int *nocp);

/* Finite-Q exciton (the BSE_BND=T path: bsebnds, module_bse.F90:1206-1566).
 * Replaces the eigsys loop over the k+Q mesh (:1398-1417): qpt(nk,3)
 * column-major is the monhkhorst_packq mesh (module_kmesh.F90:281-334) of the
 * nk points set by wtb_set_kpoints, q[3] the Cartesian Q (the reference's
 * q(2:4)).  The following wtb_bse_build then forms matrizelbsekq
 * (module_bse.F90:452-569): conduction bands from k+Q, valence bands from k,
 * V(k_i - k_j) of the k mesh and, when |Q| != 0, the exchange term
 * -V(Q) <c_i|v_i><v_j|c_j> (with a temperature: matrizelbsekqtemp :664-780).
 * nv, nc must equal those of the last wtb_hk_eig.  eig/vec/nocp (optional)
 * as for wtb_hk_eig, of the k+Q mesh.  q == NULL switches finite Q off.
 * wtb_set_tb_model and wtb_set_kpoints switch it off too.  On a finite-Q
 * store wtb_oscillator and the gradients return WTB_ERR_UNSUPPORTED. */
int wtb_hk_eig_q(wtb_ctx *ctx, const double *q, const double *qpt, int nv, int nc, double *eig, double *vec,
                 int *nocp);

/* Optical activity over the Brillouin zone (OPT_BZ=T: spoptpolbz,
 * module_bse.F90:4532-4797): for every k of the mesh, act(nk,5) column-major
 * = sum over the nc*nv band pairs of |d_a|^2, a = x, y, z, sp, sm, with
 * d_a = <c|dH/dk_a|v> / (E_c - E_v + i sme) (optspbz, module_optics.F90:257-311)
 * and sp, sm = (d_x +- i d_y)/sqrt(2).  The bands are those of wtb_hk_eig. */
int wtb_optics_bz(wtb_ctx *ctx, double sme, double *act);

/* Density of states (dostool, module_tools.F90:390-651) of all bands of the
 * last wtb_hk_eig: en(ndos) is the reference's grid from min(0, lowest band)
 * to max(0, highest band); dos(ndos,4) column-major holds tdos, spin-up,
 * spin-down and the orbital-weighted (layercont, :153-175) density, each
 * (1/nk) sum over states of weight * gaussian(E - E_n(k), sigma) (endos
 * :103-149), before the spin factor of NP models.  orbweight(norb) may be
 * NULL (all weights 1, as when ORB_W= names no file).  spin_mode (spincont
 * :41-99): 0 none (NP), 1 first/second half of the orbitals (SP, and SOC
 * with DFT=V), 2 odd/even orbitals (other SOC). */
int wtb_dos(wtb_ctx *ctx, int ndos, double sigma, const double *orbweight, int spin_mode, double *en,
            double *dos);

/* Full spectrum of every H(k): energies(nk,norb) (scissor applied) and
 * eigenvectors(nk,norb,norb) with ev(k,o,n) = component o of band n. */
int wtb_get_bands(wtb_ctx *ctx, double *energies, double *vectors);

/* BSE kernel assembly on the device: replaces matrizelbse + the assembly
 * loop (module_bse.F90:24-83,2392-2415).  precision: WTB_FP64/MIXED/FP32;
 * convention: WTB_CONV_*.  The N x N matrix (N = nk*nc*nv) stays on the GPU. */
int wtb_bse_build(wtb_ctx *ctx, const wtb_coulomb *coulomb, int precision, int convention);

/* Temperature BSE (the BSET=T path: bsesolvertemp, module_bse.F90:2649-3263).
 * temp_k >= 0 makes the following wtb_bse_build multiply every kernel
 * element H(i,j), i <= j, by f(i) = f_v - f_c of state i (Fermi level 0;
 * matrizelbsetemp :580-660 and fermidisteh, module_math.F90:410-421, which
 * takes T = 0 as 1e-8 K) and H(j,i) = conj(H(i,j)); the diagonal becomes
 * E_c - E_v + V(0) f(i).  wtb_oscillator then weights its second dipole sum
 * by f (dielbseptemp :784-810).  temp_k < 0 (the default) switches it off.
 * The gap correction of the TA= card (gapcortemp/gapcortemp2, module_math.F90:
 * 383-404) is a scissor shift, passed through the model's scissor.  A store
 * built at another temperature is dropped.  The gradient entry points return
 * WTB_ERR_UNSUPPORTED on a temperature store. */
int wtb_set_temperature(wtb_ctx *ctx, double temp_k);

/* Lowest nev eigenpairs of the BSE matrix: replaces slepc_solver(HX,EX,HE,NX,NE)
 * (module_slepc_solver.F90:12-97) with nev explicit.  tol: relative residual
 * (SLEPc default 1e-8); max_it: outer iterations (200).
 *   E : double EX(nev), ascending;  X : complex(8) HE(N,nev) or NULL;
 *   rep : may be NULL.
 * tol bounds every returned pair's relative residual |Hx - Ex| / |E| in every
 * precision, measured with the FP64 product (FP64 store) or the exact
 * int8-digit product recombined in FP64 (FP32 store).  On an FP32 store the
 * reduced-precision iterations stop at max(tol, 5e-6) and the solver goes on
 * with the accurate product while a residual is above tol.  A solve that
 * does not reach tol within max_it returns WTB_ERR_NOCONV; rep then holds
 * nconv and the achieved max_rel_residual. */
int wtb_bse_solve(wtb_ctx *ctx, int nev, double tol, int max_it, double *E, double *X,
                  wtb_report *rep);

/* Optical dipoles of every transition: replaces the optsp/hlm loop
 * (module_bse.F90:2317-2331, module_optics.F90:12-167).
 *   d : complex(8) d(N,5) with columns hrx, hry, hrz, hrsp, hrsm, or NULL. */
int wtb_optics(wtb_ctx *ctx, double cshift, double *d);

/* Exciton optical activity: replaces dielbsev (comp_a == comp_b) and
 * dielbsep (module_bse.F90:171-245).  act : double act(N); rows beyond nev
 * are zero.  Needs wtb_bse_solve and wtb_optics first. */
int wtb_oscillator(wtb_ctx *ctx, int nev, int comp_a, int comp_b, double *act);

/* Test and integration helpers.
 *   wtb_bse_get_block : copy H(i0:i0+ni-1, j0:j0+nj-1) (0-based) as complex(8)
 *                       column-major ni x nj.
 *   wtb_bse_matvec    : Y = H X for host X, Y complex(8) (N, ncol), with the
 *                       kernel of the given precision (WTB_FP64 / WTB_MIXED /
 *                       WTB_FP32).
 *   wtb_bse_diag      : the diagonal of H, E_c - E_v + V(q=0) (N). */
int wtb_bse_get_block(wtb_ctx *ctx, long i0, long ni, long j0, long nj, double *out);
int wtb_bse_matvec(wtb_ctx *ctx, int ncol, const double *X, double *Y, int precision);
int wtb_bse_diag(wtb_ctx *ctx, double *diag);

/* Y = H X like wtb_bse_matvec, without any device allocation: X and Y are
 * DEVICE arrays (complex(8) (N, ncol) column-major, checked to be device
 * memory), precision WTB_MIXED, WTB_FP32 or WTB_FP64 (WTB_INT8X is refused).
 * The split-FP16 product of WTB_MIXED takes its FP16 planes and column
 * scales from a static device array of 32 MiB + 4 KiB that each GPU holds
 * until the process ends; up to 128 columns per launch, fewer above
 * N = 32768, and N above 4194304 is refused.  Every call writes the part of
 * the array it reads; calls that overlap on one GPU share the array and
 * race.  No counterpart in the reference: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
int wtb_bse_matvec_noalloc(wtb_ctx *ctx, int ncol, const double *X, double *Y, int precision);

/* Device time in ms of the kernels of the last wtb_hk_eig, wtb_bse_build,
 * wtb_optics and wtb_oscillator calls (ms[0..3]), measured with HIP events on
 * the context stream.  No counterpart in the reference, whose MSG timers
 * (mpi_p.F90:134-160) print per-stage CPU time; its callers are benchmarks
 * and roofline reports. */
int wtb_stage_times(wtb_ctx *ctx, double *ms);

/* Gradients of L = sum_n w[n] E_n over the nev states of the last
 * wtb_bse_solve (Hellmann-Feynman contractions x^H dH x evaluated per (k,k')
 * tile, chained through the TB eigenvector adjoint).  No counterpart in the
 * reference; its callers are fitting loops over the screening parameters,
 * the scissor shift and the TB hoppings.
 *   w            : double w(nev); states of a degenerate level need equal weights
 *   g            : scalar gradients
 *   hop_re_grad, hop_im_grad : double (nvec,norb,norb), dL/dRe H(R)_mn and
 *                  dL/dIm H(R)_mn with each entry an independent parameter of
 *                  H(k) = sum_R e^{ik.R} H(R)/deg, or NULL */
typedef struct {
    double coulomb[7];       /* dL/d r0, EDIEL_T, EDIEL, EDIEL_B, EDIEL_Z, W_COUL, LC */
    double scissor;          /* dL/d scissor                                          */
    double efermi;           /* dL/d E_F (zero while the band selection is fixed)      */
    double gauge_residual;   /* max |Im u_m^H ubar_m| over the selected bands: 0 for a
                                loss invariant under TB eigenvector phases             */
} wtb_grad;
int wtb_grad_energies(wtb_ctx *ctx, const double *w, wtb_grad *g, double *hop_re_grad, double *hop_im_grad);

/* Gradients of the spectrum loss L = sum_j (eps2(omega[j]) - target[j])^2 over
 * the nev states of the last wtb_bse_solve, with
 *   eps2(w) = 180.90 / (N_k V_cell) sum_n f_n gamma / ((w - E_n)^2 + gamma^2)
 * (the Lorentzian of optdiel, module_optics.F90:171-241, at a fixed width) and
 * f_n = |sum_i x_n[i] d_i|^2 (reference convention, dielbsev,
 * module_bse.F90:231-238) or |x_n^H d|^2 (textbook convention), d the dipoles
 * of Cartesian component comp (0..2) with broadening cshift (wtb_optics).
 * Eigenvector derivatives come from the other solved states plus a
 * preconditioned CG solve on the complement of them, so the solved states
 * must be the lowest nev with a gap above.  No counterpart in the reference;
 * its callers are fits of the screening, the scissor and the hoppings to a
 * measured absorption spectrum.  In the reference convention the strengths
 * depend on the TB eigenvector phases, so only the Coulomb and scissor
 * gradients are defined there (hop_re_grad / hop_im_grad must be NULL).
 *   omega, target : double (nw);  loss : L;  eps2 : double (nw) or NULL */
int wtb_grad_spectrum(wtb_ctx *ctx, int comp, double cshift, double gamma, int nw, const double *omega,
                      const double *target, double *loss, double *eps2, wtb_grad *g, double *hop_re_grad,
                      double *hop_im_grad);

/* Drop-in for slepc_solver(HX,EX,HE,NX,NE) (module_slepc_solver.F90:12-97):
 * uploads the caller's full Hermitian host matrix HX(NX,NX) (FP64, device 0)
 * and returns the NE eigenpairs of smallest magnitude (EPS_SMALLEST_MAGNITUDE,
 * :65), tolerance 1e-8, EX(NE) ascending in |E| (ties in ascending E),
 * HE(NX,NE).  Up to NX = 1024 all eigenpairs are computed and the set is
 * chosen among them, so an indefinite spectrum gets its interior pairs.
 * Above, the block solver's lowest algebraic pairs are returned, which are
 * the smallest-magnitude set for a non-negative spectrum (bound excitons);
 * a negative lowest eigenvalue gives WTB_ERR_UNSUPPORTED. */
int wtb_slepc_solver_compat(const double *HX, double *EX, double *HE, int NX, int NE);

#ifdef __cplusplus
}
#endif
#endif
