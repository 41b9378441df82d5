// BSE kernel assembly on the device (replaces matrizelbse and the assembly
// loop of bsesolver, src/module_bse.F90:24-83,2392-2415).
//
// In block form the (k,k') block of H is V(k,k') (M^c kron M^v) with
//   M^c[c,c'] = sum_o conj(U_c(k)[o,c]) U_c(k')[o,c']       (nc x nc)
//   M^v[v,v'] = sum_o conj(U_v(k)[o,v]) U_v(k')[o,v']       (nv x nv)
// (reference convention; the textbook one uses conj(M^v)), and the diagonal
// is E_c - E_v + V(k,k).  Basis index i = k*nc*nv + c*nv + v.
//
// H is Hermitian, so only the lower triangle of square tiles of TK x TK
// k-points is computed: one 256-thread workgroup per tile (bk, bk'), bk' <= bk.
//   phase 1: U(k') of the tile is staged in LDS; every (k,k') pair gets V
//            (FP64) and its two overlap blocks (FP64 or FP32), which then
//            replace the staged U(k') in LDS.  On the FP32 store the overlap
//            blocks of the whole tile are two complex products
//            U_c(k)^H U_c(k') and U_v(k)^H U_v(k') on FP32 MFMA (U(k) staged
//            too); the FP64 store forms them with FMAs per pair.
//   phase 2: the TK*nc*nv square tile is written column by column (each store
//            covering consecutive rows of one column) and, off the diagonal,
//            its conjugate transpose into the mirrored tile: the stored
//            matrix is exactly Hermitian by construction.  Off-diagonal tiles
//            form their overlaps with FMAs; the diagonal tiles, which hold
//            both triangles of their block, use uncontracted products whose
//            rounding is symmetric under k <-> k'.
// Half the pair products of a full sweep, so the kernel runs near the HBM
// write rate of the 8 N^2 (FP32) or 16 N^2 (FP64) byte store.
#include "coulomb.h"
#include "wtb_internal.h"

#include <cmath>
#include <cstring>

namespace wtb {

CoulombDev make_coulomb(const wtb_coulomb &in, const double *rlat, const int *ngrid) {
    CoulombDev c{};
    c.kind = in.kind;
    c.ktol = in.ktol;
    c.r0 = in.r0;
    c.et = in.ediel[0];
    c.epar = in.ediel[1];
    c.eb = in.ediel[2];
    c.ez = in.ez;
    c.w = in.w;
    c.lc = in.lc;
    c.cic = (double)(-(0.0904756f * 1000.0f));
    const double pi = (double)acosf(-1.0f);
    const double nk = (double)(ngrid[0] * ngrid[1] * ngrid[2]);
    const double *a1 = rlat, *a2 = rlat + 3, *a3 = rlat + 6;
    double v[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
    const double vc2 = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double vc3 = std::fabs(v[0] * a3[0] + v[1] * a3[1] + v[2] * a3[2]);
    c.vbz2d = 1.0 / (nk * vc2);
    c.vbz3d = 1.0 / (nk * vc3);
    // V2DRK
    c.eta = std::sqrt(c.epar / c.ez);
    c.kappa = std::sqrt(c.epar * c.ez);
    c.pb = (c.eb - c.kappa) / (c.eb + c.kappa);
    c.pt = (c.et - c.kappa) / (c.et + c.kappa);
    // V2DK (its own r0 from the dielectric constants, module_coulomb.F90:37)
    {
        const double s1 = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
        const double s2 = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
        const double a0 = 0.5 * (s1 + s2);
        const double r0 = ((c.epar - 1.0) * c.lc) / (c.et + c.eb);
        const double g = (double)(ngrid[0] * ngrid[1]);
        const double auxi = (2.0 * pi * r0) / (a0 * std::sqrt(g));
        const double ed = (c.et + c.eb) / 2.0;
        c.v2dk_r0 = r0;
        c.v2dk_ed = ed;
        c.v2dk_pref = c.vbz2d * (c.cic / ed);
        c.v2dk_q0 = c.vbz2d * (c.cic / ed) * (a0 * std::sqrt(g) / (2.0 * pi)) * ((double)1.76f + auxi * 1.0);
        c.v2dk_a0sg = a0 * std::sqrt(g);
    }
    c.pi = pi;
    c.rc = 0.5 * a3[2];
    {
        const double s1 = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
        const double s2 = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
        const double s3 = std::sqrt(a3[0] * a3[0] + a3[1] * a3[1] + a3[2] * a3[2]);
        c.cr = 0.5 * std::fmin(s1, std::fmin(s2, s3));
    }
    return c;
}

namespace {

constexpr int kThreads = 256;

template <class R>
struct CT;
template <>
struct CT<double> { using C = double2; };
template <>
struct CT<float> { using C = float2; };

template <class C>
__device__ __forceinline__ C mk(double a, double b) { return C{(decltype(C{}.x))a, (decltype(C{}.x))b}; }

__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// t += conj(a) b.  SYM: no contraction, so the sums for (a, b) and for (b, a)
// are exact conjugates of each other (a diagonal tile holds both triangles
// of its block); otherwise two FMAs per component.
template <bool SYM, class C>
__device__ __forceinline__ void cmac(C &t, C a, C b) {
    if constexpr (SYM) {
#pragma clang fp contract(off)
        t.x += a.x * b.x + a.y * b.y;
        t.y += a.x * b.y - a.y * b.x;
    } else {
        t.x = fma_r(a.y, b.y, fma_r(a.x, b.x, t.x));
        t.y = fma_r(-a.y, b.x, fma_r(a.x, b.y, t.y));
    }
}

// V a b without contraction: the value for the pair (k', k) of a diagonal tile
// is then the exact conjugate of the one for (k, k').
template <class C, class R>
__device__ __forceinline__ C pair_value(R vq, C a, C b) {
#pragma clang fp contract(off)
    C v;
    v.x = vq * (a.x * b.x - a.y * b.y);
    v.y = vq * (a.x * b.y + a.y * b.x);
    return v;
}

// Overlap blocks M^c, M^v of thread's S pairs (k, k'): U(k) from global,
// OB orbitals at a time, U(k') from the staged copy in LDS.
template <bool SYM, class C, int NC, int NV, int S, int TK>
__device__ __forceinline__ void pair_products(int k, int nk, int norb, const C *__restrict__ usel,
                                              const C *uks, int tid, C (&mc)[S][NC * NC],
                                              C (&mv)[S][NV * NV]) {
    constexpr int NB = NC + NV;
    constexpr int OB = sizeof(C) == 8 ? 8 : 4;
    for (int o0 = 0; o0 < norb; o0 += OB) {
        C ukb[OB][NB];
#pragma unroll
        for (int ob = 0; ob < OB; ++ob)
#pragma unroll
            for (int j = 0; j < NB; ++j)
                ukb[ob][j] = (o0 + ob < norb) ? usel[((size_t)(o0 + ob) * NB + j) * nk + k] : mk<C>(0.0, 0.0);
#pragma unroll
        for (int ob = 0; ob < OB; ++ob) {
            const int o = o0 + ob;
            if (o >= norb) break;
            const C *uk = ukb[ob];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int kpl = tid / TK + (kThreads / TK) * s;
                const C *up = uks + (size_t)kpl * NB * norb + o;
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int c2 = 0; c2 < NC; ++c2) cmac<SYM>(mc[s][c * NC + c2], uk[NV + c], up[(NV + c2) * norb]);
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int v2 = 0; v2 < NV; ++v2) cmac<SYM>(mv[s][v * NV + v2], uk[v], up[v2 * norb]);
            }
        }
    }
}

// Exchange term of the finite-Q kernel (matrizelbsekq, src/module_bse.F90:536-567):
// -V(Q) <c_I|v_I> <v_J|c_J> = -V(Q) a_I conj(a_J) with a_I = <c(k_I + Q)|v(k_I)>,
// the same expression for either triangle (its conjugate is the swapped term).
// Uncontracted, so that (J, I) of a diagonal tile is the exact conjugate of (I, J).
template <class C>
__device__ __forceinline__ void sub_exchange(C &val, const cd *__restrict__ xa, double vx, long I, long J) {
#pragma clang fp contract(off)
    const cd ai = xa[I], aj = xa[J];
    const double re = ai.x * aj.x + ai.y * aj.y, im = ai.y * aj.x - ai.x * aj.y;
    val.x -= (decltype(val.x))(vx * re);
    val.y -= (decltype(val.y))(vx * im);
}

typedef float mf4 __attribute__((ext_vector_type(4)));

// Phase 1 of the FP32 tile on FP32 MFMA (v_mfma_f32_16x16x4f32): the overlap
// blocks of all pairs of the tile are the complex products
//   Mc = U_c(k)^H U_c(k')   ((TK nc) x (TK nc), depth norb)
//   Mv = U_v(k)^H U_v(k')   ((TK nv) x (TK nv))
// with columns (k, c) of U_c(k) at kl nc + c.  Each 16 x 16 output tile keeps
// four real accumulators, Ar^T Br, Ai^T Bi, Ar^T Bi and Ai^T Br, combined at
// the end (Re = rr + ii, Im = ri - ir): on a diagonal tile (U(k') = U(k))
// element (j, i) then sums the same products in the same order as (i, j), so
// the two are exact conjugates and the store stays exactly Hermitian.
// Operand lane l holds row (column) l % 16 at depth l / 16; result register r
// of lane l holds element (4 (l / 16) + r, l % 16).  The four wavefronts take
// every fourth output tile; the results are written into the pair planes.
template <int NC, int NV, int TK>
__device__ __forceinline__ void pair_products_mfma(int norb, const float2 *uk, const float2 *ukp, int tid,
                                                   float2 *pmc, float2 *pmv, int npp) {
    constexpr int NB = NC + NV;
    constexpr int TC = TK * NC / 16, TV = TK * NV / 16;     // 16-wide tiles per side
    constexpr int NT = TC * TC + TV * TV;
    constexpr int W = kThreads / 64, MAXT = (NT + W - 1) / W;
    const int lane = tid & 63, wave = tid >> 6, li = lane & 15, lk = lane >> 4;
    mf4 re[MAXT], im[MAXT];
#pragma unroll
    for (int u = 0; u < MAXT; ++u) {
        const int t = wave + W * u;
        if (t >= NT) break;
        const bool isc = t < TC * TC;
        const int tt = isc ? t : t - TC * TC, side = isc ? TC : TV, nb = isc ? NC : NV, j0 = isc ? NV : 0;
        const int ti = tt / side, tj = tt % side;
        // row i of A^H (column i of U(k)) and column j of B (column j of U(k'))
        const int i = 16 * ti + li, j = 16 * tj + li;
        const float2 *pa = uk + (size_t)(i / nb) * NB * norb + (size_t)(j0 + i % nb) * norb;
        const float2 *pb = ukp + (size_t)(j / nb) * NB * norb + (size_t)(j0 + j % nb) * norb;
        mf4 rr = {0.f, 0.f, 0.f, 0.f}, ii = rr, ri = rr, ir = rr;
        for (int o0 = 0; o0 < norb; o0 += 4) {
            const int o = o0 + lk;
            const float2 a = o < norb ? pa[o] : make_float2(0.f, 0.f);
            const float2 b = o < norb ? pb[o] : make_float2(0.f, 0.f);
            rr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, rr, 0, 0, 0);
            ii = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, ii, 0, 0, 0);
            ri = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, ri, 0, 0, 0);
            ir = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, ir, 0, 0, 0);
        }
        re[u] = rr + ii;
        im[u] = ri - ir;
    }
    __syncthreads();    // the pair planes overwrite the staged U
#pragma unroll
    for (int u = 0; u < MAXT; ++u) {
        const int t = wave + W * u;
        if (t >= NT) break;
        const bool isc = t < TC * TC;
        const int tt = isc ? t : t - TC * TC, side = isc ? TC : TV, nb = isc ? NC : NV;
        const int ti = tt / side, tj = tt % side;
        const int j = 16 * tj + li, kpl = j / nb, c2 = j % nb;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * ti + 4 * lk + r, kl = i / nb, c = i % nb;
            const int p = kpl * (TK + 1) + kl;
            float2 *plane = isc ? pmc + (size_t)(c * NC + c2) * npp : pmv + (size_t)(c * NV + c2) * npp;
            plane[p] = make_float2(re[u][r], im[u][r]);
        }
    }
}

// Pair slots of one plane of pair data in LDS: pair (kl, kpl) sits at
// kpl (TK + 1) + kl, so that lanes running over kl (row writes) and lanes
// running over kpl (mirror writes) both hit distinct banks.
template <int TK>
constexpr int pair_slots() { return TK * (TK + 1) + 1; }

// V(k - k') of a pair: from the table over mesh offsets on a regular mesh
// (vtab non-null), else from the k-points.
__device__ __forceinline__ double pair_potential(const CoulombDev &coul, const double *kpt, const double *vtab,
                                                 const int *koff, long vcenter, int k, int kp) {
    if (vtab) return vtab[vcenter + koff[k] - koff[kp]];
    return coulomb_v(coul, kpt[3 * k] - kpt[3 * kp], kpt[3 * k + 1] - kpt[3 * kp + 1], kpt[3 * k + 2] - kpt[3 * kp + 2]);
}

// V over the offsets of a regular mesh: entry t = (d_1 + n_1 - 1) S_1 +
// (d_2 + n_2 - 1) S_2 + d_3 + n_3 - 1 holds V(d_1 g_1 + d_2 g_2 + d_3 g_3).
struct KGen {
    double g[9];
};
__global__ void vtab_kernel(CoulombDev coul, int n1, int n2, int n3, KGen kg, double *__restrict__ vtab) {
    const long S2 = 2L * n3 - 1, S1 = (2L * n2 - 1) * S2, T = (2L * n1 - 1) * S1;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= T) return;
    const long d1 = t / S1 - (n1 - 1), r = t % S1, d2 = r / S2 - (n2 - 1), d3 = r % S2 - (n3 - 1);
    const double *g = kg.g;
    const double dx = (double)d1 * g[0] + (double)d2 * g[3] + (double)d3 * g[6];
    const double dy = (double)d1 * g[1] + (double)d2 * g[4] + (double)d3 * g[7];
    const double dz = (double)d1 * g[2] + (double)d2 * g[5] + (double)d3 * g[8];
    vtab[t] = coulomb_v(coul, dx, dy, dz);
}

// R: arithmetic/store precision of the overlaps and the matrix; NC, NV:
// bands; TK: k-points per side of the square tile (TK nc nv rows and columns).
// Workgroup t covers the tile (bk, bkp) of the lower block triangle,
// t = bk (bk + 1) / 2 + bkp with bkp <= bk.
// fdeh (temperature BSE, else null): every element H(I, J) off the diagonal
// is multiplied by fdeh[min(I, J)] and the diagonal V(0) term by fdeh[I]
// (matrizelbsetemp fills j >= i with the factor of row i,
// src/module_bse.F90:580-660,3062-3070); the mirror of an element carries the
// same factor, so the store stays exactly Hermitian.
// xa (finite Q != 0, else null): the exchange term above, vqx[0] = V(Q), added
// before the factor (matrizelbsekqtemp, src/module_bse.F90:664-780).
// EXT: compiled with the temperature and exchange terms; the plain kernel
// (EXT = false) carries neither, so their registers do not cost it occupancy.
template <class R, int NC, int NV, int TK, bool EXT>
__global__ void __launch_bounds__(kThreads)
bse_build_kernel(int nk, int norb, long ld, int textbook, CoulombDev coul,
                 const double *__restrict__ kpt, const double *__restrict__ vtab, const int *__restrict__ koff,
                 long vcenter, const double *__restrict__ esel,
                 const typename CT<R>::C *__restrict__ usel, typename CT<R>::C *__restrict__ H,
                 unsigned *__restrict__ hmax, const double *__restrict__ fdeh, const cd *__restrict__ xa,
                 const double *__restrict__ vqx) {
    using C = typename CT<R>::C;
    constexpr int NB = NC + NV, NCV = NC * NV, NC2 = NC * NC, NV2 = NV * NV;
    const double vx = (EXT && xa) ? vqx[0] : 0.0;
    constexpr int S = TK * TK / kThreads;              // pairs per thread
    constexpr int NPP = pair_slots<TK>();
    static_assert(kThreads % TK == 0 && S >= 1, "tile must split evenly over the workgroup");
    extern __shared__ __align__(16) unsigned char smem[];
    const long N = (long)nk * NCV;
    const long t = blockIdx.x;
    int bk = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)bk * (bk + 1) / 2 > t) --bk;
    while ((long)(bk + 1) * (bk + 2) / 2 <= t) ++bk;
    const int bkp = (int)(t - (long)bk * (bk + 1) / 2);
    const bool dtile = bk == bkp;
    const int k0 = bk * TK, kp0 = bkp * TK;
    // LDS: the pair planes (NC2 of M^c, NV2 of M^v) share their space with
    // the staged U(k') [kpl][j][o] read before them; then pair V (R).
    const size_t pairs = (size_t)(NC2 + NV2) * NPP, stage = (size_t)TK * NB * norb;
    C *pmc = reinterpret_cast<C *>(smem);
    C *pmv = pmc + (size_t)NC2 * NPP;
    C *uks = pmc;
    constexpr bool MF = sizeof(R) == sizeof(float);   // phase 1 on FP32 MFMA
    const size_t stage_all = (MF ? 2 : 1) * stage;
    R *pv = reinterpret_cast<R *>(pmc + (pairs > stage_all ? pairs : stage_all));
    C *ukk = uks + stage;                               // U(k) of the tile's rows (MFMA path)
    const int tid = threadIdx.x;

    // stage U(k') for the tile's k' columns (and U(k) for its k rows on the
    // MFMA path), eight loads in flight per thread
    {
        constexpr int kL = 8;
        const int tot = TK * NB * norb, tot2 = (MF ? 2 : 1) * tot;
        for (int base = 0; base < tot2; base += kL * kThreads) {
            C v[kL];
#pragma unroll
            for (int u = 0; u < kL; ++u) {
                const int idx0 = base + u * kThreads + tid;
                const bool second = idx0 >= tot;
                const int idx = second ? idx0 - tot : idx0;
                const int kpl = idx / (NB * norb), rem = idx % (NB * norb), j = rem / norb, o = rem % norb;
                const int kp = (second ? k0 : kp0) + kpl;
                v[u] = (idx0 < tot2 && kp < nk) ? usel[((size_t)o * NB + j) * nk + kp] : mk<C>(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < kL; ++u) {
                const int idx = base + u * kThreads + tid;
                if (idx < tot2) uks[idx] = v[u];
            }
        }
    }
    __syncthreads();

    if constexpr (MF) {
        pair_products_mfma<NC, NV, TK>(norb, reinterpret_cast<const float2 *>(ukk),
                                       reinterpret_cast<const float2 *>(uks), tid, reinterpret_cast<float2 *>(pmc),
                                       reinterpret_cast<float2 *>(pmv), NPP);
        // Coulomb potential of every pair
        const int kl = tid % TK, k = k0 + kl;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int kpl = tid / TK + (kThreads / TK) * s;
            const int kp = kp0 + kpl;
            double vq = 0.0;
            if (k < nk && kp < nk)
                vq = pair_potential(coul, kpt, vtab, koff, vcenter, k, kp);
            pv[kpl * (TK + 1) + kl] = (R)vq;
        }
    } else {
    // phase 1: pair products.  Thread -> kl = tid % TK, kpl = tid / TK + (kThreads / TK) s
    {
        const int kl = tid % TK;
        const int k = k0 + kl;
        C mc[S][NC2], mv[S][NV2];
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int e = 0; e < NC2; ++e) mc[s][e] = mk<C>(0.0, 0.0);
#pragma unroll
            for (int e = 0; e < NV2; ++e) mv[s][e] = mk<C>(0.0, 0.0);
        }
        if (k < nk) {
            if (dtile) pair_products<true, C, NC, NV, S, TK>(k, nk, norb, usel, uks, tid, mc, mv);
            else pair_products<false, C, NC, NV, S, TK>(k, nk, norb, usel, uks, tid, mc, mv);
        }
        __syncthreads();    // the pair products overwrite the staged U
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int kpl = tid / TK + (kThreads / TK) * s;
            const int p = kpl * (TK + 1) + kl;
            const int kp = kp0 + kpl;
            double vq = 0.0;
            if (k < nk && kp < nk)
                vq = pair_potential(coul, kpt, vtab, koff, vcenter, k, kp);
            pv[p] = (R)vq;
#pragma unroll
            for (int e = 0; e < NC2; ++e) pmc[(size_t)e * NPP + p] = mc[s][e];
#pragma unroll
            for (int e = 0; e < NV2; ++e) pmv[(size_t)e * NPP + p] = mv[s][e];
        }
    }
    }
    __syncthreads();

    // phase 2: the tile, column by column.  Thread -> (row il, k' kpl): the
    // nc + nv overlap factors and V are read once from LDS for the nc nv
    // outputs of the pair; each store covers consecutive rows of one column.
    float vmax = 0.f;   // largest |Re|, |Im| written (scales the FP16 split of the matvec)
    constexpr int ROWS = TK * NCV;
    const long I0 = (long)k0 * NCV, J0 = (long)kp0 * NCV;
    for (int idx = tid; idx < ROWS * TK; idx += kThreads) {
        const int il = idx % ROWS, kpl = idx / ROWS;
        const long I = I0 + il, Jb = J0 + (long)kpl * NCV;
        if (I >= N || Jb >= N) continue;       // rows N..ld-1 are cleared separately
        const int kl = il / NCV, ci = (il % NCV) / NV, vi = il % NV;
        const int p = kpl * (TK + 1) + kl;
        const R vq = pv[p];
        C a[NC], b[NV];
#pragma unroll
        for (int cj = 0; cj < NC; ++cj) a[cj] = pmc[(size_t)(ci * NC + cj) * NPP + p];
#pragma unroll
        for (int vj = 0; vj < NV; ++vj) {
            b[vj] = pmv[(size_t)(vi * NV + vj) * NPP + p];
            if (textbook) b[vj].y = -b[vj].y;
        }
#pragma unroll
        for (int cj = 0; cj < NC; ++cj)
#pragma unroll
            for (int vj = 0; vj < NV; ++vj) {
                const long J = Jb + cj * NV + vj;
                C val;
                if (I == J) {
                    // The FP32 store keeps a zero diagonal: E_c - E_v + V(0)
                    // lives in the FP64 diagonal array and is applied in FP64
                    // by the matvec, so FP32 rounding only touches the Coulomb part.
                    if (sizeof(R) == sizeof(float)) {
                        val = mk<C>(0.0, 0.0);
                    } else {
                        const int k = k0 + kl;
                        double de = esel[(size_t)k * NB + NV + ci] - esel[(size_t)k * NB + vi];
                        double w = (double)vq;
                        if (EXT && xa) w -= vx * (xa[I].x * xa[I].x + xa[I].y * xa[I].y);
                        val = mk<C>((EXT && fdeh) ? de + w * fdeh[I] : de + w, 0.0);
                    }
                } else {
                    val = pair_value(vq, a[cj], b[vj]);
                    if (EXT && xa) sub_exchange(val, xa, vx, I, J);
                    if (EXT && fdeh) {
                        const R f = (R)fdeh[I < J ? I : J];
                        val.x *= f;
                        val.y *= f;
                    }
                }
                H[I + J * ld] = val;
                vmax = fmaxf(vmax, fmaxf(fabsf((float)val.x), fabsf((float)val.y)));
            }
    }
    // the mirror of an off-diagonal tile: rows J (k'), columns I (k), the
    // conjugates of the values above, so the store is exactly Hermitian.
    // Thread -> (k' kpl, column il): the nc nv rows of the pair are
    // consecutive, and consecutive threads continue them down the column.
    if (!dtile) {
        for (int idx = tid; idx < ROWS * TK; idx += kThreads) {
            const int kpl = idx % TK, il = idx / TK;
            const long I = I0 + il, Jb = J0 + (long)kpl * NCV;
            if (I >= N || Jb >= N) continue;
            const int kl = il / NCV, ci = (il % NCV) / NV, vi = il % NV;
            const int p = kpl * (TK + 1) + kl;
            const R vq = pv[p];
            C a[NC], b[NV];
#pragma unroll
            for (int cj = 0; cj < NC; ++cj) a[cj] = pmc[(size_t)(ci * NC + cj) * NPP + p];
#pragma unroll
            for (int vj = 0; vj < NV; ++vj) {
                b[vj] = pmv[(size_t)(vi * NV + vj) * NPP + p];
                if (textbook) b[vj].y = -b[vj].y;
            }
            C *col = H + (size_t)I * ld + Jb;
#pragma unroll
            for (int cj = 0; cj < NC; ++cj)
#pragma unroll
                for (int vj = 0; vj < NV; ++vj) {
                    C val = pair_value(vq, a[cj], b[vj]);
                    if (EXT && xa) sub_exchange(val, xa, vx, I, Jb + cj * NV + vj);
                    if (EXT && fdeh) {     // J < I off the diagonal tiles: the factor of J
                        const R f = (R)fdeh[Jb + cj * NV + vj];
                        val.x *= f;
                        val.y *= f;
                    }
                    col[cj * NV + vj] = C{val.x, -val.y};
                }
        }
    }
    if (hmax) {
        for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
        if ((threadIdx.x & 63) == 0) atomicMax(hmax, __float_as_uint(vmax));
    }
}

// rows N..ld-1 of every stored column are zero (the matvecs read whole
// k slices without bounds checks)
template <class C>
__global__ void zero_pad_rows_kernel(long N, long ld, C *__restrict__ H) {
    const long pad = ld - N;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= pad * N) return;
    H[N + i % pad + (i / pad) * ld] = C{0, 0};
}

template <class R, int NC, int NV>
void launch_tile(Ctx &c, const CoulombDev &cd_, int textbook, const double *esel, const typename CT<R>::C *usel,
                 typename CT<R>::C *H) {
    using C = typename CT<R>::C;
    constexpr int per = NC * NC + NV * NV;
    // square tiles of 32 k-points where their pair planes fit in 72 KB (two
    // workgroups per CU), else of 16
    constexpr int TK = (size_t)pair_slots<32>() * per * sizeof(C) <= 72 * 1024 ? 32 : 16;
    const int nk = c.bands.nk, norb = c.model.norb;
    const long N = (long)nk * NC * NV, ld = c.bse.ld;
    const size_t pairs = (size_t)pair_slots<TK>() * per,
                 stage = (sizeof(R) == sizeof(float) ? 2 : 1) * (size_t)TK * (NC + NV) * norb;
    const size_t shmem = std::max(pairs, stage) * sizeof(C) + (size_t)pair_slots<TK>() * sizeof(R) + 16;
    if (shmem > 160 * 1024) throw Error{WTB_ERR_UNSUPPORTED, "assembly tile does not fit in LDS"};
    const long nblk = (nk + TK - 1) / TK;
    const long ntile = nblk * (nblk + 1) / 2;
    unsigned *hm = sizeof(R) == sizeof(float) ? c.bse.hmax.p : nullptr;
    // the potential table of a regular mesh (filled by launch_bse_build)
    const double *vt = c.kmesh ? c.vtab.p : nullptr;
    const long vc = (long)(c.ngrid[0] - 1) * (2L * c.ngrid[1] - 1) * (2L * c.ngrid[2] - 1) +
                    (long)(c.ngrid[1] - 1) * (2L * c.ngrid[2] - 1) + (c.ngrid[2] - 1);
    const double *fd = c.bse.temperature >= 0.0 ? c.bse.fdeh.p : nullptr;
    const cd *xa = c.bse.xa.n ? c.bse.xa.p : nullptr;
    if (fd || xa)
        hipLaunchKernelGGL((bse_build_kernel<R, NC, NV, TK, true>), dim3((unsigned)ntile), dim3(kThreads), shmem,
                           c.stream, nk, norb, ld, textbook, cd_, c.kpt.p, vt, c.kmesh_off.p, vc, esel, usel, H, hm, fd, xa,
                           c.bse.vqx.p);
    else
        hipLaunchKernelGGL((bse_build_kernel<R, NC, NV, TK, false>), dim3((unsigned)ntile), dim3(kThreads), shmem,
                           c.stream, nk, norb, ld, textbook, cd_, c.kpt.p, vt, c.kmesh_off.p, vc, esel, usel, H, hm, nullptr,
                           nullptr,
                           nullptr);
    WTB_LAUNCH_CHECK();
    if (ld > N) {
        const long tot = (ld - N) * N;
        hipLaunchKernelGGL(zero_pad_rows_kernel<C>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, N,
                           ld, H);
        WTB_LAUNCH_CHECK();
    }
}

template <class R>
void dispatch_tile(Ctx &c, const CoulombDev &cd_, int textbook, const double *esel, const typename CT<R>::C *usel,
                   typename CT<R>::C *H) {
    const int nc = c.bands.nc, nv = c.bands.nv;
#define WTB_TILE_CASE(A, B) \
    if (nc == A && nv == B) return launch_tile<R, A, B>(c, cd_, textbook, esel, usel, H);
    WTB_TILE_CASE(1, 1) WTB_TILE_CASE(1, 2) WTB_TILE_CASE(2, 1) WTB_TILE_CASE(2, 2)
    WTB_TILE_CASE(1, 3) WTB_TILE_CASE(3, 1) WTB_TILE_CASE(2, 3) WTB_TILE_CASE(3, 2) WTB_TILE_CASE(3, 3)
    WTB_TILE_CASE(1, 4) WTB_TILE_CASE(4, 1) WTB_TILE_CASE(2, 4) WTB_TILE_CASE(4, 2) WTB_TILE_CASE(3, 4)
    WTB_TILE_CASE(4, 3) WTB_TILE_CASE(4, 4)
#undef WTB_TILE_CASE
    throw Error{WTB_ERR_UNSUPPORTED, "nc, nv must be in 1..4 on the assembly kernel"};
}

__global__ void bse_diag_kernel(int nk, int nc, int nv, CoulombDev coul, const double *__restrict__ esel,
                                const double *__restrict__ fdeh, const cd *__restrict__ xa,
                                const double *__restrict__ vqx, double *__restrict__ diag) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long N = (long)nk * nc * nv;
    if (i >= N) return;
    int k = (int)(i / (nc * nv)), c = (int)((i / nv) % nc), v = (int)(i % nv);
    int nb = nc + nv;
    double w = coulomb_v(coul, 0.0, 0.0, 0.0);
    if (xa) w -= vqx[0] * (xa[i].x * xa[i].x + xa[i].y * xa[i].y);
    diag[i] = esel[(size_t)k * nb + nv + c] - esel[(size_t)k * nb + v] + (fdeh ? w * fdeh[i] : w);
}

// Finite-Q band set: slot j < nv of every k from the k mesh, slots nv.. from
// the k+Q mesh (bsebnds pairs vector(k, v) with vectorq(k, c),
// src/module_bse.F90:1436-1442); energies [k][j], vectors [(o nb + j) nk + k].
__global__ void fq_splice_kernel(int nk, int nb, int nv, int norb, const double *__restrict__ ek,
                                 const double *__restrict__ eq, const cd *__restrict__ uk,
                                 const cd *__restrict__ uq, double *__restrict__ e, cd *__restrict__ u,
                                 cf *__restrict__ u32) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long nu = (long)nk * nb * norb;
    if (i < nu) {
        const int j = (int)((i / nk) % nb);
        const cd x = j < nv ? uk[i] : uq[i];
        u[i] = x;
        u32[i] = make_float2((float)x.x, (float)x.y);
    }
    if (i < (long)nk * nb) e[i] = (i % nb) < nv ? ek[i] : eq[i];
}

// a_I = <c(k+Q)|v(k)> = sum_o conj(c_o) v_o of every state I = k nc nv + c nv + v
// (vcv of matrizelbsekq, src/module_bse.F90:551-555), and V(Q) (vcoulq: the
// potential at q(2:4) - 0, :478-533)
__global__ void fq_exchange_kernel(int nk, int nc, int nv, int norb, CoulombDev coul, double qx, double qy,
                                   double qz, const cd *__restrict__ u, cd *__restrict__ xa,
                                   double *__restrict__ vqx) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long N = (long)nk * nc * nv;
    if (i == 0) vqx[0] = coulomb_v(coul, qx, qy, qz);
    if (i >= N) return;
    const int k = (int)(i / (nc * nv)), c = (int)((i / nv) % nc), v = (int)(i % nv);
    const int nb = nc + nv;
    double re = 0.0, im = 0.0;
    for (int o = 0; o < norb; ++o) {
        const cd cc = u[((size_t)o * nb + nv + c) * nk + k], vv = u[((size_t)o * nb + v) * nk + k];
        re += cc.x * vv.x + cc.y * vv.y;
        im += cc.x * vv.y - cc.y * vv.x;
    }
    xa[i] = make_double2(re, im);
}

// f_v - f_c of every state at temperature temp (fermidisteh,
// src/module_math.F90:410-421): Fermi level 0, the reference's single-precision
// Boltzmann constant, and T = 0 taken as 1e-8 K (also a single-precision literal).
__global__ void fdeh_kernel(int nk, int nc, int nv, double temp, const double *__restrict__ esel,
                            double *__restrict__ fdeh) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long N = (long)nk * nc * nv;
    if (i >= N) return;
    int k = (int)(i / (nc * nv)), c = (int)((i / nv) % nc), v = (int)(i % nv);
    int nb = nc + nv;
    const double ec = esel[(size_t)k * nb + nv + c], ev = esel[(size_t)k * nb + v];
    const double kb = (double)8.617330350e-5f;
    const double t = temp == 0.0 ? (double)1e-8f : temp;
    const double beta = 1.0 / (kb * t);
    const double fv = 1.0 / (exp(ev * beta) + 1.0), fc = 1.0 / (exp(ec * beta) + 1.0);
    fdeh[i] = fv - fc;
}

}  // namespace

void detect_kmesh(Ctx &c, const std::vector<double> &k) {
    c.kmesh = false;
    const long n1 = c.ngrid[0], n2 = c.ngrid[1], n3 = c.ngrid[2], nk = c.nk;
    if (n1 * n2 * n3 != nk) return;
    const long T = (2 * n1 - 1) * (2 * n2 - 1) * (2 * n3 - 1);
    if (T > (1L << 24)) return;
    // the last coordinate running fastest, or the first
    for (int order = 0; order < 2; ++order) {
        const auto idx = [&](long c1, long c2, long c3) {
            return order == 0 ? (c1 * n2 + c2) * n3 + c3 : (c3 * n2 + c2) * n1 + c1;
        };
        double g[9] = {0};
        double gmax = 0.0;
        const long unit[3] = {n1 > 1 ? idx(1, 0, 0) : -1, n2 > 1 ? idx(0, 1, 0) : -1, n3 > 1 ? idx(0, 0, 1) : -1};
        for (int j = 0; j < 3; ++j)
            if (unit[j] >= 0)
                for (int d = 0; d < 3; ++d) {
                    g[3 * j + d] = k[3 * unit[j] + d] - k[d];
                    gmax = std::max(gmax, std::fabs(g[3 * j + d]));
                }
        const double tol = 1e-9 * std::max(gmax, 1e-300);
        bool ok = gmax > 0.0;
        for (long c1 = 0; c1 < n1 && ok; ++c1)
            for (long c2 = 0; c2 < n2 && ok; ++c2)
                for (long c3 = 0; c3 < n3 && ok; ++c3) {
                    const long kk = idx(c1, c2, c3);
                    for (int d = 0; d < 3; ++d) {
                        const double p = k[d] + (double)c1 * g[d] + (double)c2 * g[3 + d] + (double)c3 * g[6 + d];
                        if (!(std::fabs(p - k[3 * kk + d]) <= tol)) ok = false;
                    }
                }
        if (!ok) continue;
        const long S2 = 2 * n3 - 1, S1 = (2 * n2 - 1) * S2;
        std::vector<int> off(nk);
        for (long c1 = 0; c1 < n1; ++c1)
            for (long c2 = 0; c2 < n2; ++c2)
                for (long c3 = 0; c3 < n3; ++c3) off[idx(c1, c2, c3)] = (int)(c1 * S1 + c2 * S2 + c3);
        c.kmesh_off.alloc(nk);
        WTB_HIP_CHECK(hipMemcpy(c.kmesh_off.p, off.data(), sizeof(int) * nk, hipMemcpyHostToDevice));
        for (int i = 0; i < 9; ++i) c.kgen[i] = g[i];
        c.kmesh = true;
        return;
    }
}

void launch_bse_build(Ctx &c, const wtb_coulomb &coul, int precision, int convention) {
    Bands &b = c.bands;
    if (!b.valid) throw Error{WTB_ERR_STATE, "wtb_hk_eig must run before wtb_bse_build"};
    if (coul.kind < WTB_V2DK || coul.kind > WTB_V1DT) throw Error{WTB_ERR_ARG, "Wrong Coulomb Potential"};
    if (precision != WTB_FP64 && precision != WTB_MIXED && precision != WTB_FP32)
        throw Error{WTB_ERR_ARG, "unknown precision"};
    if (b.nc > 4 || b.nv > 4) throw Error{WTB_ERR_UNSUPPORTED, "nc, nv <= 4 on the assembly kernel"};
    const int nk = b.nk, nc = b.nc, nv = b.nv;
    const long N = (long)nk * nc * nv;
    CoulombDev cd_ = make_coulomb(coul, c.model.rlat, c.ngrid);
    if (c.kmesh) {
        const long T = (2L * c.ngrid[0] - 1) * (2L * c.ngrid[1] - 1) * (2L * c.ngrid[2] - 1);
        c.vtab.alloc(T);
        KGen kg;
        for (int i = 0; i < 9; ++i) kg.g[i] = c.kgen[i];
        hipLaunchKernelGGL(vtab_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, c.stream, cd_, c.ngrid[0],
                           c.ngrid[1], c.ngrid[2], kg, c.vtab.p);
        WTB_LAUNCH_CHECK();
    }
    Bse &s = c.bse;
    s.valid = false;
    c.sol.valid = false;
    const bool f64 = precision == WTB_FP64;
    // free the store of the other precision first: 288 GB is plenty, but
    // there is no reason to keep both.
    if (f64) s.h32.free(); else s.h64.free();
    s.ld = store_ld(N);
    if (f64) s.h64.alloc((size_t)s.ld * N); else s.h32.alloc((size_t)s.ld * N);
    s.diag.alloc(N);
    s.hmax.alloc(1);
    WTB_HIP_CHECK(hipMemsetAsync(s.hmax.p, 0, sizeof(unsigned), c.stream));
    // finite Q: conduction bands from the k+Q mesh, exchange when |Q| != 0
    // (bsebnds, src/module_bse.F90:1398-1447; modq test of matrizelbsekq :535)
    const double *esel = b.esel.p;
    const cd *u64 = b.usel.p;
    const cf *u32 = b.usel32.p;
    s.finite_q = c.fq.on;
    s.xa.free();
    s.vqx.free();
    if (c.fq.on) {
        FiniteQ &f = c.fq;
        if (!f.bq.valid || f.bq.nk != nk || f.bq.nv != nv || f.bq.nc != nc)
            throw Error{WTB_ERR_STATE, "wtb_hk_eig_q must run with the nv, nc of wtb_hk_eig before wtb_bse_build"};
        const int norb = c.model.norb, nb = nc + nv;
        f.esel.alloc((size_t)nk * nb);
        f.usel.alloc((size_t)nk * nb * norb);
        f.usel32.alloc((size_t)nk * nb * norb);
        const long nu = (long)nk * nb * norb;
        hipLaunchKernelGGL(fq_splice_kernel, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, c.stream, nk, nb, nv,
                           norb, b.esel.p, f.bq.esel.p, b.usel.p, f.bq.usel.p, f.esel.p, f.usel.p, f.usel32.p);
        WTB_LAUNCH_CHECK();
        esel = f.esel.p;
        u64 = f.usel.p;
        u32 = f.usel32.p;
        const double modq = std::sqrt(f.q[0] * f.q[0] + f.q[1] * f.q[1] + f.q[2] * f.q[2]);
        if (modq != 0.0) {
            s.xa.alloc(N);
            s.vqx.alloc(1);
            hipLaunchKernelGGL(fq_exchange_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.stream, nk, nc,
                               nv, norb, cd_, f.q[0], f.q[1], f.q[2], f.usel.p, s.xa.p, s.vqx.p);
            WTB_LAUNCH_CHECK();
        }
    }
    s.temperature = c.temperature;
    if (s.temperature >= 0.0) {
        s.fdeh.alloc(N);
        hipLaunchKernelGGL(fdeh_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.stream, nk, nc, nv,
                           s.temperature, esel, s.fdeh.p);
        WTB_LAUNCH_CHECK();
    }
    if (f64) dispatch_tile<double>(c, cd_, convention == WTB_CONV_TEXTBOOK, esel, u64, s.h64.p);
    else dispatch_tile<float>(c, cd_, convention == WTB_CONV_TEXTBOOK, esel, u32, s.h32.p);
    hipLaunchKernelGGL(bse_diag_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.stream, nk, nc, nv,
                       cd_, esel, s.temperature >= 0.0 ? (const double *)s.fdeh.p : nullptr,
                       s.xa.n ? (const cd *)s.xa.p : nullptr, s.vqx.p, s.diag.p);
    WTB_LAUNCH_CHECK();
    if (!f64) {
        // power-of-two scale for the FP16 split: max |H| / hscale in [2^12, 2^13)
        unsigned bits = 0;
        WTB_HIP_CHECK(hipMemcpyAsync(&bits, s.hmax.p, sizeof(unsigned), hipMemcpyDeviceToHost, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
        float hm;
        std::memcpy(&hm, &bits, sizeof(float));
        int e = 0;
        if (hm > 0.f) std::frexp(hm, &e);
        s.hscale = std::ldexp(1.0f, e - 13);
    }
    s.N = N;
    s.precision = precision;
    s.convention = convention;
    s.coul = coul;
    s.valid = true;
}

}  // namespace wtb
