// Internal declarations of the gfx950 exciton library.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wtb_hip.h"

namespace wtb {

using cd = double2;   // complex FP64 (re, im)
using cf = float2;    // complex FP32

constexpr int kMaxOrb = 32;   // H(k) eigensolver limit (one wavefront per k)
constexpr int kMaxNcv = 16;   // nc*nv limit of the assembly kernel

// Status with message, raised through the C-ABI.
struct Error {
    int code;
    std::string msg;
};

#define WTB_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::wtb::Error{WTB_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define WTB_BLAS_CHECK(expr)                                                             \
    do {                                                                                 \
        rocblas_status s_ = (expr);                                                      \
        if (s_ != rocblas_status_success)                                                \
            throw ::wtb::Error{WTB_ERR_HIP, std::string(#expr) + ": " +                  \
                                                rocblas_status_to_string(s_)};           \
    } while (0)

#define WTB_LAUNCH_CHECK() WTB_HIP_CHECK(hipGetLastError())

// Device buffer owned by the context.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (count <= n && p) return;
        free();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{WTB_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e)};
        }
        n = count;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { free(); }
};

// Pinned host buffer for the solver's small device <-> host transfers: a
// copy from or to it is a plain DMA, queued without a host-side stage.
template <class T>
struct HBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (count <= n && p) return;
        free();
        if (count == 0) return;
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{WTB_ERR_NOMEM, "hipHostMalloc of " + std::to_string(count * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e)};
        }
        n = count;
    }
    void free() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    ~HBuf() { free(); }
};

// Block eigensolver workspace, kept across solves (the Krylov block alone is
// 6 N b complex values; allocating it per solve costs ~0.3 ms of hipMalloc).
// exchange two buffers (the solver's basis and its scratch trade places)
template <class T>
inline void swap_bufs(DBuf<T> &a, DBuf<T> &b) {
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

struct SolverWs {
    DBuf<cd> S, AS, T1, T2, P, AP;  // N x cols
    DBuf<cd> G, A, U, T, Tmp, Cm;   // small (3b) x (3b)
    DBuf<double> D, lam, lamr, norm2, sc;
    DBuf<int> idx, act;
    HBuf<double> hd;                // pinned: diagonal, residual norms, Ritz values
    HBuf<int> hi;                   // pinned: active column indices
    HBuf<int> hflag;                // pinned: status words of the CholQR factor
    HBuf<double> hkept;             // pinned: column norms of the projected search block
};

struct Model {
    int norb = 0, nvec = 0;
    double rlat[9] = {0};        // rows a_i
    double efermi = 0, scissor = 0;
    int flags = 0;
    DBuf<cd> hop;                // [r][m][n]
    DBuf<double> rcart;          // [r][3] Cartesian R, reference summation order
    DBuf<double> invdeg;         // [r] 1/deg
    DBuf<int> onsite;            // [r] 1 if R == 0
};

struct Bands {
    int nk = 0, nv = 0, nc = 0, nb = 0;
    DBuf<double> eall;   // [k][n] all bands, scissor applied
    DBuf<cd> uall;       // [k][o][n] component o of band n
    DBuf<cd> hk;         // [k][r*n + s] H(k) of the eigensolve (before the Fermi shift)
    DBuf<int> nocp;      // [k]
    DBuf<double> esel;   // [k][j] selected bands
    DBuf<cd> usel;       // [(o*nb + j)*nk + k]  (k fastest, as the reference's vector(k,j,o))
    DBuf<cf> usel32;     // FP32 copy for the mixed-precision assembly
    DBuf<double> zrot;   // [k][o][j] real QL rotations of the batched eigensolver (fallback path)
    DBuf<double> tri;    // [k][3][32] diagonal, off-diagonal and eigenvalues of T(k)
    DBuf<double2> rot;   // [k][cap] QL rotation log (cos, sin) of the batched eigensolver
    DBuf<unsigned char> rix;  // [k][cap] rotated column pair (i, i + 1) of each logged rotation
    DBuf<int> nrot;      // [k] logged rotations
    DBuf<int> status;    // device error flag
    bool valid = false;
};

// Leading dimension of the BSE store: N rounded up to a multiple of 64.  Rows
// N..ld-1 of every column are zero, so the matvec streams whole 64-row tiles
// of each column without bounds checks and every column starts 512-byte aligned.
inline long store_ld(long N) { return (N + 63) / 64 * 64; }

struct Bse {
    long N = 0;
    long ld = 0;         // store_ld(N)
    int precision = -1;
    int convention = 0;
    wtb_coulomb coul{};
    DBuf<cd> h64;        // ld x N column-major (FP64 store)
    DBuf<cf> h32;        // ld x N column-major (FP32 store)
    DBuf<double> diag;   // E_c - E_v
    DBuf<unsigned> hmax; // max |Re|, |Im| of the FP32 store (float bits)
    float hscale = 1.f;  // power of two: max |H| / hscale ~ 2^13 for the FP16 split
    double temperature = -1.0;  // of the build (< 0: none)
    DBuf<double> fdeh;   // f_v - f_c of every state when built with a temperature
    bool finite_q = false;      // built on the k+Q conduction bands
    DBuf<cd> xa;         // exchange amplitudes <c(k+Q)|v(k)> of every state (Q != 0)
    DBuf<double> vqx;    // V(Q) of the exchange term
    bool valid = false;
};

// Finite-Q exciton (bsebnds, src/module_bse.F90:1206-1566): conduction bands
// on the k+Q mesh, valence bands on the k mesh
struct FiniteQ {
    bool on = false;
    double q[3] = {0, 0, 0};   // Cartesian Q (the reference's q(2:4))
    DBuf<double> kq;     // [k][3] k+Q mesh
    Bands bq;            // eigensolve on the k+Q mesh
    DBuf<double> esel;   // [k][j]: j < nv from the k mesh, conduction from k+Q
    DBuf<cd> usel;       // layout of Bands::usel
    DBuf<cf> usel32;
};

struct Solution {
    int nev = 0;
    DBuf<double> evals;  // nev
    DBuf<cd> evecs;      // N x nev column-major
    bool valid = false;
};

// Device time of the matvec kernels proper: while `on`, every launch of a
// matvec kernel is bracketed by HIP events on the context stream and its
// algorithmic bytes (the store, read once) and flops (8 N^2 per column) are
// added to the slot of its kind.
struct KernelClock {
    enum Kind { F16SPLIT = 0, F64_ON_F64 = 1, F64_ON_F32 = 2, F32_ON_F32 = 3, INT8X = 4, NKIND = 5 };
    struct Slot {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
        double bytes = 0, flops = 0;
    };
    bool on = false;
    Slot slot[NKIND];
    void begin(Kind k, hipStream_t st) {
        if (!on) return;
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
        (void)hipEventRecord(a, st);
        slot[k].ev.push_back({a, b});
    }
    void end(Kind k, hipStream_t st, double bytes, double flops) {
        if (!on || slot[k].ev.empty()) return;
        (void)hipEventRecord(slot[k].ev.back().second, st);
        slot[k].bytes += bytes;
        slot[k].flops += flops;
    }
    // summed milliseconds of a slot (the stream must have completed)
    double ms(Kind k) const {
        double t = 0;
        for (auto &p : slot[k].ev) {
            float m = 0.f;
            (void)hipEventElapsedTime(&m, p.first, p.second);
            t += m;
        }
        return t;
    }
    void reset() {
        for (auto &s : slot) {
            for (auto &p : s.ev) {
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
            s.ev.clear();
            s.bytes = s.flops = 0;
        }
    }
    ~KernelClock() { reset(); }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    std::string err;
    Model model;
    int nk = 0;
    int ngrid[3] = {1, 1, 1};
    DBuf<double> kpt;    // [k][3]
    // Regular mesh: kpt[k] = kpt[0] + c_1 g_1 + c_2 g_2 + c_3 g_3 with integer
    // c_j in [0, ngrid_j) for every k (checked when the k-points are set).
    // Then V(k - k') depends on c(k) - c(k') only: the assembly reads it from
    // a table over the offsets, kmesh_off[k] = c_1 S_1 + c_2 S_2 + c_3 with
    // S_2 = 2 n_3 - 1, S_1 = (2 n_2 - 1) S_2.
    bool kmesh = false;
    double kgen[9] = {0};       // g_1, g_2, g_3
    DBuf<int> kmesh_off;        // [k]
    DBuf<double> vtab;          // [offset] V of the last assembly's potential
    Bands bands;
    FiniteQ fq;
    Bse bse;
    Solution sol;
    DBuf<cd> dip;        // [a][N], a = x, y, z, sp, sm
    bool dip_valid = false;
    // solver workspace
    DBuf<cd> ws_c;       // Jacobi rotation accumulator
    DBuf<int> ws_i;      // Jacobi eigenvalue permutation
    DBuf<cd> ws_dense;   // dense FP64 copy for small problems
    DBuf<cd> ws_heev_c;  // Householder scalars
    DBuf<double> ws_heev_d;  // tridiagonal matrix, its eigenvectors, inverse-iteration work
    DBuf<int> ws_flag;   // device status word of the small factorisations
    DBuf<cd> ws_gram;    // split-K partial Gram products
    DBuf<_Float16> ws_split;   // FP16 split planes of X for the mixed matvec
    DBuf<float> ws_colscale;   // per-column split / output scales
    DBuf<double> ws_scale8;    // per-column scales of the int8-digit matvec
    DBuf<cd> ws_ypart;         // k-chunk partial products of the int8-digit matvec
    SolverWs lws;
    KernelClock kclock;
    double temperature = -1.0;  // wtb_set_temperature: < 0 off (kelvin otherwise)
    // device time (ms) of the last wtb_hk_eig, wtb_bse_build, wtb_optics and
    // wtb_oscillator, from HIP events on the stream around their kernels
    double stage_ms[4] = {0, 0, 0, 0};
    // device address of this GPU's scratch array of the allocation-free
    // matvec (set by scratch_setup in wtb_create)
    void *scratch = nullptr;
};

// Brackets the kernels of one entry point with HIP events on the context
// stream; stop() waits for the second event and stores the elapsed time.
struct StageTimer {
    Ctx &c;
    int slot;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(Ctx &cx, int s) : c(cx), slot(s) {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        (void)hipEventRecord(a, c.stream);
    }
    void stop() {
        if (!a || !b) return;
        (void)hipEventRecord(b, c.stream);
        if (hipEventSynchronize(b) != hipSuccess) return;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, a, b) == hipSuccess) c.stage_ms[slot] = ms;
    }
    ~StageTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// value of lane i (uniform i) of a per-lane double
__device__ __forceinline__ double wave_lane(double v, int i) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, i);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), i);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v moved across lanes by the DPP pattern CTRL (both 32-bit halves); a lane
// whose source lane is disabled reads zero (bound_ctrl), so no old value has
// to be set up
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes, returned in every lane, without LDS round trips:
// DPP quad_perm [1,0,3,2] and [2,3,0,1] give each lane its quad's sum,
// row_half_mirror (lane i <- 7 - i) the sum over 8 and row_mirror (i <- 15 - i)
// over 16; the four row sums are then read with readlane and added in a
// fixed order, so every lane holds the same value.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_double<0xB1>(v);
    v += dpp_double<0x4E>(v);
    v += dpp_double<0x141>(v);
    v += dpp_double<0x140>(v);
    return (wave_lane(v, 0) + wave_lane(v, 16)) + (wave_lane(v, 32) + wave_lane(v, 48));
}

// k.R = (kx Rx + ky Ry) + kz Rz, evaluated without FMA contraction in the
// reference's order.  On a Monkhorst-Pack mesh k.R can land exactly on a
// float32 rounding midpoint, so the sign of the double rounding error decides
// the float32 phase of the compat mode: every evaluation must use the same
// plain IEEE operations.
__device__ __forceinline__ double kdotr(double kx, double ky, double kz, const double *R) {
#pragma clang fp contract(off)
    const double a = kx * R[0];
    const double b = ky * R[1];
    const double c = kz * R[2];
    return (a + b) + c;
}

// ---- kernels (launch wrappers) ----
void launch_hk_eig(Ctx &c, int nv, int nc);
void hk_eig_on(Ctx &c, const double *kpt, int nk, int nv, int nc, Bands &b);
void launch_bse_build(Ctx &c, const wtb_coulomb &coul, int precision, int convention);
void launch_optics(Ctx &c, double cshift);
void launch_optics_bz(Ctx &c, double sme, double *act_host);
void launch_dos(Ctx &c, int ndos, double sigma, const double *orbweight, int spin_mode, double *en_host,
                double *dos_host);
void launch_oscillator(Ctx &c, int nev, int a, int b, double *act_dev);

// Y = H X, X and Y column-major (ldx = ldy = N), complex FP64 host-visible
// device buffers.  precision selects the kernel.
void matvec(Ctx &c, int precision, int ncol, const cd *X, cd *Y);

// Y = H_offdiag X on the FP32 store with split-FP16 MFMA (mixed precision).
void matvec_f16split(Ctx &c, int ncol, const cd *X, cd *Y);

// The same product with its split planes and column scales in the file-scope
// scratch array of k_matvec16.hip (no device allocation); scratch_setup looks
// up that array's device address for the context's device.
void matvec_f16split_noalloc(Ctx &c, int ncol, const cd *X, cd *Y);
void scratch_setup(Ctx &c);

// Y = H X without device allocation (split-FP16 products through the scratch
// array for WTB_MIXED); X, Y device arrays
void matvec_noalloc(Ctx &c, int precision, int ncol, const cd *X, cd *Y);

// Y = H_offdiag X on the FP32 store with exact int8 digit products (the
// refinement product of the mixed-precision solver).
void matvec_int8(Ctx &c, int ncol, const cd *X, cd *Y, int levels = 5);
// solver-internal product id (never accepted at the C-ABI): the int8-digit
// product in its short form (four significance levels)
constexpr int kInt8Short = 103;

// Hermitian eigensolver for small dense matrices on the GPU (FP64 Jacobi).
// A (m x m column-major, lda) is destroyed; W receives ascending eigenvalues,
// Z (ldz) the eigenvectors.  Work arrays are allocated from the context.
void small_heev(Ctx &c, int m, cd *A, int lda, double *W, cd *Z, int ldz);

// Lowest nvec eigenpairs of a dense Hermitian matrix (n <= 1024) by
// Householder tridiagonalisation, bisection and inverse iteration.  A is
// destroyed; W ascending; Z (n x nvec, ldz).
void heev_lowest(Ctx &c, int n, int nvec, cd *A, int lda, double *W, cd *Z, int ldz);

// In-place upper Cholesky factor G = R^H R (n x n, ldg).  Returns 0, or
// 1 + the first column whose pivot falls below rel * its original diagonal.
int cholesky_upper(Ctx &c, int n, cd *G, int ldg, double rel, bool check = true, double *min_ratio = nullptr);
const int *cholesky_upper_async(Ctx &c, int n, cd *G, int ldg, double rel, int *hflag, hipEvent_t ev);
// Regular-mesh check of the k-points k ([k][3], host) against c.ngrid: sets
// c.kmesh, c.kgen and c.kmesh_off.
void detect_kmesh(Ctx &c, const std::vector<double> &k);

// Copy H(i0:i0+ni, j0:j0+nj) of the stored matrix into out (ni x nj, ldo) as FP64.
// gradients of sum_n w_n E_n (k_grad.hip)
// Extra inputs of the quadratic-form adjoint from the dipole adjoint:
// dL/dE_j(k) [k nb + j], Ubar in the layout [(j norb + o) nk + k], and the
// adjoint Dbar(k) [k norb^2 + a norb + b] of dH/dk_alpha (any may be null).
struct GradExtra {
    const double *ebar = nullptr;
    const cd *ubar = nullptr;
    const cd *dbar = nullptr;
    int alpha = 0;
};
void grad_core(Ctx &c, const cd *X, int nx, const double *w, const GradExtra *ex, wtb_grad *out, double *g_re,
               double *g_im);
void grad_energies(Ctx &c, const double *w, wtb_grad *out, double *g_re, double *g_im);
void bse_block_to_c64(Ctx &c, long i0, long ni, long j0, long nj, cd *out, long ldo);

// Block eigensolver driver (LOBPCG with soft locking, diagonal preconditioner).
void lobpcg(Ctx &c, int nev, double tol, int max_it, int precision, wtb_report *rep);
// C (m x n) = A^H B for tall A (K x m), B (K x n), FP64 MFMA
void gram_tall(Ctx &c, int m, int n, long K, const cd *A, int lda, const cd *B, int ldb, cd *C, int ldc);
// C (M x n) -= A (M x k) B (k x n);  C = A B
void gemm_tall_sub(Ctx &c, long M, int n, int k, const cd *A, long lda, const cd *B, int ldb, cd *C, long ldc);
void gemm_tall(Ctx &c, long M, int n, int k, const cd *A, long lda, const cd *B, int ldb, cd *C, long ldc);
// gradient of the spectrum loss sum_j (eps2(w_j) - t_j)^2 (k_spec.hip)
void grad_spectrum(Ctx &c, int comp, double cshift, double gamma, int nw, const double *omega, const double *target,
                   double *loss, double *eps2, wtb_grad *out, double *g_re, double *g_im);

float elapsed_ms(hipEvent_t a, hipEvent_t b);

}  // namespace wtb
