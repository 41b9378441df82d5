// BSE kernel assembly on the device (replaces matrizelbse and the assembly
// loop of bsesolver, src/module_bse.F90:24-83,2392-2415).
//
// In block form the (k,k') block of H is V(k,k') (M^c kron M^v) with
//   M^c[c,c'] = sum_o conj(U_c(k)[o,c]) U_c(k')[o,c']       (nc x nc)
//   M^v[v,v'] = sum_o conj(U_v(k)[o,v]) U_v(k')[o,v']       (nv x nv)
// (reference convention; the textbook one uses conj(M^v)), and the diagonal
// is E_c - E_v + V(k,k).  Basis index i = k*nc*nv + c*nv + v.
//
// One 256-thread workgroup per tile of TK=64 k-points x TKP k'-points:
//   phase 1: U(k') of the tile is staged in LDS; every (k,k') pair gets V
//            (FP64) and its two overlap blocks (FP64 or FP32), which then
//            replace the staged U(k') in LDS.
//   phase 2: the 64*nc*nv x TKP*nc*nv output tile is written column by column,
//            each store covering 256 consecutive rows of one column
//            (2 KB FP32 / 4 KB FP64 contiguous): the stage is HBM-write bound.
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

constexpr int kTK = 64;
constexpr int kThreads = 256;

template <class R>
struct CT;
template <>
struct CT<double> { using C = double2; };
template <>
struct CT<float> { using C = float2; };

template <class C>
__device__ __forceinline__ C mk(double a, double b) { return C{(decltype(C{}.x))a, (decltype(C{}.x))b}; }

// R: arithmetic/store precision of the overlaps and the matrix; NC, NV: bands;
// TK k-points per tile (TK nc nv rows), S: k' pairs per thread (tile width
// TKP = S kThreads / TK k'-points).
template <class R, int NC, int NV, int S, int TK = kTK>
__global__ void __launch_bounds__(kThreads)
bse_build_kernel(int nk, int norb, long ld, int textbook, CoulombDev coul,
                 const double *__restrict__ kpt, const double *__restrict__ esel,
                 const typename CT<R>::C *__restrict__ usel, typename CT<R>::C *__restrict__ H,
                 unsigned *__restrict__ hmax) {
    // No FMA contraction: every product and sum is then rounded the same way
    // for the pair (k,k') as for (k',k) with the roles of the factors swapped,
    // so the stored matrix is exactly Hermitian, H[j][i] == conj(H[i][j]) bit
    // for bit (the matvec reads row i as the conjugate of stored column i).
#pragma clang fp contract(off)
    using C = typename CT<R>::C;
    float vmax = 0.f;   // largest |Re|, |Im| written (scales the FP16 split of the matvec)
    constexpr int NB = NC + NV, NCV = NC * NV, NC2 = NC * NC, NV2 = NV * NV;
    constexpr int TKP = S * kThreads / TK;
    extern __shared__ __align__(16) unsigned char smem[];
    const long N = (long)nk * NCV;
    const int k0 = blockIdx.x * TK, kp0 = blockIdx.y * TKP;
    constexpr int NPAIR = TK * TKP;
    // LDS layout: a union of {pair Mc, pair Mv} (written after phase 1) and
    // the staged U(k') [kpl][j][o] (read in phase 1), then pair V (R).
    const size_t pairs = (size_t)NPAIR * (NC2 + NV2), stage = (size_t)TKP * NB * norb;
    C *pmc = reinterpret_cast<C *>(smem);
    C *pmv = pmc + (size_t)NPAIR * NC2;
    C *uks = pmc;
    R *pv = reinterpret_cast<R *>(pmc + (pairs > stage ? pairs : stage));
    const int tid = threadIdx.x;

    // stage U(k') for the tile's k' columns, eight loads in flight per thread
    {
        constexpr int kL = 8;
        const int tot = TKP * NB * norb;
        for (int base = 0; base < tot; base += kL * kThreads) {
            C v[kL];
#pragma unroll
            for (int t = 0; t < kL; ++t) {
                const int idx = base + t * kThreads + tid;
                const int kpl = idx / (NB * norb), rem = idx % (NB * norb), j = rem / norb, o = rem % norb;
                const int kp = kp0 + kpl;
                v[t] = (idx < tot && kp < nk) ? usel[((size_t)o * NB + j) * nk + kp] : mk<C>(0.0, 0.0);
            }
#pragma unroll
            for (int t = 0; t < kL; ++t) {
                const int idx = base + t * kThreads + tid;
                if (idx < tot) uks[idx] = v[t];
            }
        }
    }
    __syncthreads();

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
            // U(k) of OB orbitals is loaded before their products are formed
            constexpr int OB = sizeof(R) == sizeof(float) ? 8 : 4;
            for (int o0 = 0; o0 < norb; o0 += OB) {
                C ukb[OB][NB];
#pragma unroll
                for (int ob = 0; ob < OB; ++ob)
#pragma unroll
                    for (int j = 0; j < NB; ++j)
                        ukb[ob][j] = (o0 + ob < norb) ? usel[((size_t)(o0 + ob) * NB + j) * nk + k]
                                                      : mk<C>(0.0, 0.0);
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
                        for (int c = 0; c < NC; ++c) {
                            const C a = uk[NV + c];
#pragma unroll
                            for (int c2 = 0; c2 < NC; ++c2) {
                                const C b = up[(NV + c2) * norb];
                                C &t = mc[s][c * NC + c2];
                                t.x += a.x * b.x + a.y * b.y;
                                t.y += a.x * b.y - a.y * b.x;
                            }
                        }
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            const C a = uk[v];
#pragma unroll
                            for (int v2 = 0; v2 < NV; ++v2) {
                                const C b = up[v2 * norb];
                                C &t = mv[s][v * NV + v2];
                                t.x += a.x * b.x + a.y * b.y;
                                t.y += a.x * b.y - a.y * b.x;
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();    // the pair products overwrite the staged U
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int kpl = tid / TK + (kThreads / TK) * s;
            const int p = kpl * TK + kl;
            const int kp = kp0 + kpl;
            double vq = 0.0;
            if (k < nk && kp < nk)
                vq = coulomb_v(coul, kpt[3 * k] - kpt[3 * kp], kpt[3 * k + 1] - kpt[3 * kp + 1],
                               kpt[3 * k + 2] - kpt[3 * kp + 2]);
            pv[p] = (R)vq;
#pragma unroll
            for (int e = 0; e < NC2; ++e) pmc[(size_t)p * NC2 + e] = mc[s][e];
#pragma unroll
            for (int e = 0; e < NV2; ++e) pmv[(size_t)p * NV2 + e] = mv[s][e];
        }
    }
    __syncthreads();

    // phase 2: coalesced column writes.  Thread row il (k, c, v) takes one
    // pair (k, k') at a time: the nc + nv overlap factors and V are read once
    // from LDS for the nc nv outputs of the pair, each store covering 256
    // consecutive rows of one column.
    constexpr int ROWS = TK * NCV;
    const long I0 = (long)k0 * NCV, J0 = (long)kp0 * NCV;
    for (int il = tid; il < ROWS; il += kThreads) {
        const long I = I0 + il;
        if (I >= ld) break;     // rows N..ld-1 come out zero (k >= nk has zero V and overlaps)
        const int kl = il / NCV, ci = (il % NCV) / NV, vi = il % NV;
        for (int kpl = 0; kpl < TKP; ++kpl) {
            const long Jb = J0 + (long)kpl * NCV;
            if (Jb >= N) break;
            const int p = kpl * TK + kl;
            const R vq = pv[p];
            C a[NC], b[NV];
#pragma unroll
            for (int cj = 0; cj < NC; ++cj) a[cj] = pmc[(size_t)p * NC2 + ci * NC + cj];
#pragma unroll
            for (int vj = 0; vj < NV; ++vj) {
                b[vj] = pmv[(size_t)p * NV2 + vi * NV + vj];
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
                            val = mk<C>(de + (double)vq, 0.0);
                        }
                    } else {
                        val.x = vq * (a[cj].x * b[vj].x - a[cj].y * b[vj].y);
                        val.y = vq * (a[cj].x * b[vj].y + a[cj].y * b[vj].x);
                    }
                    H[I + J * ld] = val;
                    vmax = fmaxf(vmax, fmaxf(fabsf((float)val.x), fabsf((float)val.y)));
                }
        }
    }
    if (hmax) {
        for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
        if ((threadIdx.x & 63) == 0) atomicMax(hmax, __float_as_uint(vmax));
    }
}

template <class R, int NC, int NV>
void launch_tile(Ctx &c, const CoulombDev &cd_, int textbook, const typename CT<R>::C *usel,
                 typename CT<R>::C *H) {
    using C = typename CT<R>::C;
    constexpr int per = NC * NC + NV * NV;
    // k' pairs per thread: the pair products of a tile stay within 64 KB
    constexpr size_t kPairBytes = 64 * 1024;
    constexpr int S0 = (int)(kPairBytes / ((size_t)kThreads * per * sizeof(C)));
    constexpr int S = S0 >= 4 ? 4 : (S0 >= 2 ? 2 : 1);
    constexpr int TK = kTK, TKP = S * kThreads / TK;
    const int nk = c.bands.nk, norb = c.model.norb;
    const size_t pairs = (size_t)TK * TKP * per, ukp = (size_t)TKP * (NC + NV) * norb;
    const size_t shmem = std::max(pairs, ukp) * sizeof(C) + (size_t)TK * TKP * sizeof(R) + 16;
    if (shmem > 160 * 1024) throw Error{WTB_ERR_UNSUPPORTED, "assembly tile does not fit in LDS"};
    dim3 grid((nk + TK - 1) / TK, (nk + TKP - 1) / TKP);
    unsigned *hm = sizeof(R) == sizeof(float) ? c.bse.hmax.p : nullptr;
    hipLaunchKernelGGL((bse_build_kernel<R, NC, NV, S, TK>), grid, dim3(kThreads), shmem, c.stream, nk, norb,
                       c.bse.ld, textbook, cd_, c.kpt.p, c.bands.esel.p, usel, H, hm);
    WTB_LAUNCH_CHECK();
}

template <class R>
void dispatch_tile(Ctx &c, const CoulombDev &cd_, int textbook, const typename CT<R>::C *usel,
                   typename CT<R>::C *H) {
    const int nc = c.bands.nc, nv = c.bands.nv;
#define WTB_TILE_CASE(A, B) \
    if (nc == A && nv == B) return launch_tile<R, A, B>(c, cd_, textbook, usel, H);
    WTB_TILE_CASE(1, 1) WTB_TILE_CASE(1, 2) WTB_TILE_CASE(2, 1) WTB_TILE_CASE(2, 2)
    WTB_TILE_CASE(1, 3) WTB_TILE_CASE(3, 1) WTB_TILE_CASE(2, 3) WTB_TILE_CASE(3, 2) WTB_TILE_CASE(3, 3)
    WTB_TILE_CASE(1, 4) WTB_TILE_CASE(4, 1) WTB_TILE_CASE(2, 4) WTB_TILE_CASE(4, 2) WTB_TILE_CASE(3, 4)
    WTB_TILE_CASE(4, 3) WTB_TILE_CASE(4, 4)
#undef WTB_TILE_CASE
    throw Error{WTB_ERR_UNSUPPORTED, "nc, nv must be in 1..4 on the assembly kernel"};
}

__global__ void bse_diag_kernel(int nk, int nc, int nv, CoulombDev coul, const double *__restrict__ esel,
                                double *__restrict__ diag) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long N = (long)nk * nc * nv;
    if (i >= N) return;
    int k = (int)(i / (nc * nv)), c = (int)((i / nv) % nc), v = (int)(i % nv);
    int nb = nc + nv;
    diag[i] = esel[(size_t)k * nb + nv + c] - esel[(size_t)k * nb + v] + coulomb_v(coul, 0.0, 0.0, 0.0);
}

}  // namespace

void launch_bse_build(Ctx &c, const wtb_coulomb &coul, int precision, int convention) {
    Bands &b = c.bands;
    if (!b.valid) throw Error{WTB_ERR_STATE, "wtb_hk_eig must run before wtb_bse_build"};
    if (coul.kind == WTB_V1DT) throw Error{WTB_ERR_UNSUPPORTED, "V1DT Coulomb potential is not available"};
    if (coul.kind < WTB_V2DK || coul.kind > WTB_V0DT) throw Error{WTB_ERR_ARG, "Wrong Coulomb Potential"};
    if (precision != WTB_FP64 && precision != WTB_MIXED && precision != WTB_FP32)
        throw Error{WTB_ERR_ARG, "unknown precision"};
    if (b.nc > 4 || b.nv > 4) throw Error{WTB_ERR_UNSUPPORTED, "nc, nv <= 4 on the assembly kernel"};
    const int nk = b.nk, nc = b.nc, nv = b.nv;
    const long N = (long)nk * nc * nv;
    CoulombDev cd_ = make_coulomb(coul, c.model.rlat, c.ngrid);
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
    if (f64) dispatch_tile<double>(c, cd_, convention == WTB_CONV_TEXTBOOK, b.usel.p, s.h64.p);
    else dispatch_tile<float>(c, cd_, convention == WTB_CONV_TEXTBOOK, b.usel32.p, s.h32.p);
    hipLaunchKernelGGL(bse_diag_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.stream, nk, nc, nv,
                       cd_, b.esel.p, s.diag.p);
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
