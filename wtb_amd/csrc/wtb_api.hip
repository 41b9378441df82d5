// C-ABI entry points (include/wtb_hip.h).  Every entry point converts
// exceptions into a WTB_ERR_* status and keeps the message in the context.
#include "wtb_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

using namespace wtb;

struct wtb_ctx {
    Ctx c;
};

namespace {

// Every host<->device copy of the API runs on the context stream and waits
// for it: the stream is non-blocking, so a null-stream hipMemcpy would not
// order after the kernels queued on it.
void copy(Ctx &c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    WTB_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, c.stream));
    WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
}

template <class F>
int guarded(wtb_ctx *ctx, F &&f) {
    if (!ctx) return WTB_ERR_ARG;
    try {
        hipError_t e = hipSetDevice(ctx->c.device);
        if (e != hipSuccess) throw Error{WTB_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)};
        f(ctx->c);
        ctx->c.err.clear();
        return WTB_OK;
    } catch (const Error &e) {
        ctx->c.err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->c.err = "host allocation failed";
        return WTB_ERR_NOMEM;
    } catch (const std::exception &e) {
        ctx->c.err = e.what();
        return WTB_ERR_HIP;
    }
}

__global__ void block_to_c64_f64(long ld, long i0, long ni, long j0, long nj, const cd *__restrict__ H,
                                 cd *__restrict__ out, long ldo) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long j = blockIdx.y;
    if (i < ni && j < nj) out[i + j * ldo] = H[(i0 + i) + (j0 + j) * ld];
}

// the FP32 store has a zero diagonal; the FP64 diagonal array completes it
__global__ void block_to_c64_f32(long ld, long i0, long ni, long j0, long nj, const cf *__restrict__ H,
                                 const double *__restrict__ diag, cd *__restrict__ out, long ldo) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long j = blockIdx.y;
    if (i < ni && j < nj) {
        cf h = H[(i0 + i) + (j0 + j) * ld];
        double re = h.x;
        if (i0 + i == j0 + j) re = diag[i0 + i];
        out[i + j * ldo] = make_double2(re, h.y);
    }
}

}  // namespace

namespace wtb {
void bse_block_to_c64(Ctx &c, long i0, long ni, long j0, long nj, cd *out, long ldo) {
    const long ld = c.bse.ld;
    if (nj > 65535) {
        for (long j = 0; j < nj; j += 65535)
            bse_block_to_c64(c, i0, ni, j0 + j, std::min(65535L, nj - j), out + j * ldo, ldo);
        return;
    }
    dim3 grid((unsigned)((ni + 255) / 256), (unsigned)nj);
    if (c.bse.precision == WTB_FP64)
        hipLaunchKernelGGL(block_to_c64_f64, grid, dim3(256), 0, c.stream, ld, i0, ni, j0, nj, c.bse.h64.p, out, ldo);
    else
        hipLaunchKernelGGL(block_to_c64_f32, grid, dim3(256), 0, c.stream, ld, i0, ni, j0, nj, c.bse.h32.p,
                           c.bse.diag.p, out, ldo);
    WTB_LAUNCH_CHECK();
}
}  // namespace wtb

extern "C" {

int wtb_create(wtb_ctx **out, int device) {
    if (!out) return WTB_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return WTB_ERR_ARG;
    wtb_ctx *ctx = new (std::nothrow) wtb_ctx;
    if (!ctx) return WTB_ERR_NOMEM;
    ctx->c.device = device;
    int st = guarded(ctx, [](Ctx &c) {
        WTB_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        WTB_BLAS_CHECK(rocblas_create_handle(&c.blas));
        WTB_BLAS_CHECK(rocblas_set_stream(c.blas, c.stream));
        scratch_setup(c);
    });
    if (st != WTB_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return WTB_OK;
}

int wtb_destroy(wtb_ctx *ctx) {
    if (!ctx) return WTB_OK;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    if (ctx->c.blas) rocblas_destroy_handle(ctx->c.blas);
    hipStream_t st = ctx->c.stream;
    delete ctx;  // frees device buffers
    if (st) (void)hipStreamDestroy(st);
    return WTB_OK;
}

const char *wtb_last_error(const wtb_ctx *ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int wtb_set_tb_model(wtb_ctx *ctx, int norb, int nvec, const int *rvec, const int *deg, const double *hop_re,
                     const double *hop_im, const double *rlat, double efermi, double scissor, int flags) {
    return guarded(ctx, [&](Ctx &c) {
        if (norb <= 0 || nvec <= 0 || !rvec || !deg || !hop_re || !hop_im || !rlat)
            throw Error{WTB_ERR_ARG, "invalid TB model arguments"};
        if (norb > kMaxOrb) throw Error{WTB_ERR_UNSUPPORTED, "norb > 32"};
        Model &m = c.model;
        // rlat(i,j) column-major -> rows a_i
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m.rlat[3 * i + j] = rlat[i + 3 * j];
        std::vector<cd> hop((size_t)nvec * norb * norb);
        std::vector<double> rc(3 * (size_t)nvec), inv(nvec);
        std::vector<int> ons(nvec);
        for (int r = 0; r < nvec; ++r) {
            if (deg[r] <= 0) throw Error{WTB_ERR_ARG, "degeneracy must be positive"};
            int n0 = rvec[r], n1 = rvec[r + nvec], n2 = rvec[r + 2 * nvec];
            for (int cc = 0; cc < 3; ++cc)
                rc[3 * r + cc] = ((double)n0 * m.rlat[cc] + (double)n1 * m.rlat[3 + cc]) + (double)n2 * m.rlat[6 + cc];
            inv[r] = 1.0 / (double)deg[r];
            ons[r] = (n0 == 0 && n1 == 0 && n2 == 0);
            for (int a = 0; a < norb; ++a)
                for (int b = 0; b < norb; ++b) {
                    // hopmatrices(r, a, b), column-major (nvec, norb, norb)
                    size_t src = r + (size_t)nvec * (a + (size_t)norb * b);
                    double re = hop_re[src], im = hop_im[src];
                    if (!std::isfinite(re) || !std::isfinite(im)) throw Error{WTB_ERR_ARG, "non-finite hopping"};
                    hop[((size_t)r * norb + a) * norb + b] = make_double2(re, im);
                }
        }
        m.hop.alloc(hop.size());
        m.rcart.alloc(rc.size());
        m.invdeg.alloc(nvec);
        m.onsite.alloc(nvec);
        copy(c, m.hop.p, hop.data(), sizeof(cd) * hop.size(), hipMemcpyHostToDevice);
        copy(c, m.rcart.p, rc.data(), sizeof(double) * rc.size(), hipMemcpyHostToDevice);
        copy(c, m.invdeg.p, inv.data(), sizeof(double) * nvec, hipMemcpyHostToDevice);
        copy(c, m.onsite.p, ons.data(), sizeof(int) * nvec, hipMemcpyHostToDevice);
        m.norb = norb;
        m.nvec = nvec;
        m.efermi = efermi;
        m.scissor = scissor;
        m.flags = flags;
        c.bands.valid = false;
        c.fq.on = false;
        c.fq.bq.valid = false;
        c.bse.valid = false;
        c.sol.valid = false;
        c.dip_valid = false;
    });
}

int wtb_set_kpoints(wtb_ctx *ctx, int nk, const double *kpt, const int *ngrid) {
    return guarded(ctx, [&](Ctx &c) {
        if (nk <= 0 || !kpt || !ngrid) throw Error{WTB_ERR_ARG, "invalid k-points"};
        if (ngrid[0] <= 0 || ngrid[1] <= 0 || ngrid[2] <= 0) throw Error{WTB_ERR_ARG, "invalid grid"};
        std::vector<double> k(3 * (size_t)nk);
        for (int i = 0; i < nk; ++i)
            for (int cc = 0; cc < 3; ++cc) k[3 * i + cc] = kpt[i + (size_t)nk * cc];
        c.kpt.alloc(k.size());
        copy(c, c.kpt.p, k.data(), sizeof(double) * k.size(), hipMemcpyHostToDevice);
        c.nk = nk;
        for (int i = 0; i < 3; ++i) c.ngrid[i] = ngrid[i];
        detect_kmesh(c, k);
        c.bands.valid = false;
        c.fq.on = false;
        c.fq.bq.valid = false;
        c.bse.valid = false;
        c.sol.valid = false;
        c.dip_valid = false;
    });
}

int wtb_hk_eig(wtb_ctx *ctx, int nv, int nc, double *eig, double *vec, int *nocp) {
    return guarded(ctx, [&](Ctx &c) {
        StageTimer tm(c, 0);
        launch_hk_eig(c, nv, nc);
        tm.stop();
        c.bse.valid = false;
        c.sol.valid = false;
        c.dip_valid = false;
        Bands &b = c.bands;
        const int nk = b.nk, nb = b.nb, n = c.model.norb;
        if (eig) {
            std::vector<double> e((size_t)nk * nb);
            copy(c, e.data(), b.esel.p, sizeof(double) * e.size(), hipMemcpyDeviceToHost);
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < nb; ++j) eig[k + (size_t)nk * j] = e[(size_t)k * nb + j];
        }
        if (vec) {
            // device usel[(o*nb + j)*nk + k] is exactly Fortran vector(nk, nb, norb)
            copy(c, vec, b.usel.p, sizeof(cd) * (size_t)nk * nb * n, hipMemcpyDeviceToHost);
        }
        if (nocp) copy(c, nocp, b.nocp.p, sizeof(int) * nk, hipMemcpyDeviceToHost);
    });
}

int wtb_hk_eig_q(wtb_ctx *ctx, const double *q, const double *qpt, int nv, int nc, double *eig, double *vec,
                 int *nocp) {
    return guarded(ctx, [&](Ctx &c) {
        FiniteQ &f = c.fq;
        c.bse.valid = false;
        c.sol.valid = false;
        if (!q) {
            f.on = false;
            f.bq.valid = false;
            return;
        }
        if (!qpt) throw Error{WTB_ERR_ARG, "null k+Q mesh"};
        if (c.nk <= 0) throw Error{WTB_ERR_STATE, "no k-points set"};
        for (int i = 0; i < 3; ++i)
            if (!std::isfinite(q[i])) throw Error{WTB_ERR_ARG, "non-finite Q"};
        const int nk = c.nk;
        std::vector<double> k(3 * (size_t)nk);
        for (int i = 0; i < nk; ++i)
            for (int cc = 0; cc < 3; ++cc) k[3 * i + cc] = qpt[i + (size_t)nk * cc];
        f.kq.alloc(k.size());
        copy(c, f.kq.p, k.data(), sizeof(double) * k.size(), hipMemcpyHostToDevice);
        f.on = false;
        hk_eig_on(c, f.kq.p, nk, nv, nc, f.bq);
        for (int i = 0; i < 3; ++i) f.q[i] = q[i];
        f.on = true;
        Bands &b = f.bq;
        const int nb = b.nb, n = c.model.norb;
        if (eig) {
            std::vector<double> e((size_t)nk * nb);
            copy(c, e.data(), b.esel.p, sizeof(double) * e.size(), hipMemcpyDeviceToHost);
            for (int kk = 0; kk < nk; ++kk)
                for (int j = 0; j < nb; ++j) eig[kk + (size_t)nk * j] = e[(size_t)kk * nb + j];
        }
        if (vec) copy(c, vec, b.usel.p, sizeof(cd) * (size_t)nk * nb * n, hipMemcpyDeviceToHost);
        if (nocp) copy(c, nocp, b.nocp.p, sizeof(int) * nk, hipMemcpyDeviceToHost);
    });
}

int wtb_optics_bz(wtb_ctx *ctx, double sme, double *act) {
    return guarded(ctx, [&](Ctx &c) { launch_optics_bz(c, sme, act); });
}

int wtb_dos(wtb_ctx *ctx, int ndos, double sigma, const double *orbweight, int spin_mode, double *en,
            double *dos) {
    return guarded(ctx, [&](Ctx &c) { launch_dos(c, ndos, sigma, orbweight, spin_mode, en, dos); });
}

int wtb_get_bands(wtb_ctx *ctx, double *energies, double *vectors) {
    return guarded(ctx, [&](Ctx &c) {
        Bands &b = c.bands;
        if (!b.valid) throw Error{WTB_ERR_STATE, "wtb_hk_eig has not run"};
        const int nk = b.nk, n = c.model.norb;
        if (energies) {
            std::vector<double> e((size_t)nk * n);
            copy(c, e.data(), b.eall.p, sizeof(double) * e.size(), hipMemcpyDeviceToHost);
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < n; ++j) energies[k + (size_t)nk * j] = e[(size_t)k * n + j];
        }
        if (vectors) {
            std::vector<cd> u((size_t)nk * n * n);
            copy(c, u.data(), b.uall.p, sizeof(cd) * u.size(), hipMemcpyDeviceToHost);
            cd *out = reinterpret_cast<cd *>(vectors);
            for (int k = 0; k < nk; ++k)
                for (int o = 0; o < n; ++o)
                    for (int j = 0; j < n; ++j)
                        out[k + (size_t)nk * (o + (size_t)n * j)] = u[((size_t)k * n + o) * n + j];
        }
    });
}

int wtb_bse_build(wtb_ctx *ctx, const wtb_coulomb *coulomb, int precision, int convention) {
    return guarded(ctx, [&](Ctx &c) {
        if (!coulomb) throw Error{WTB_ERR_ARG, "null Coulomb parameters"};
        if (convention != WTB_CONV_REFERENCE && convention != WTB_CONV_TEXTBOOK)
            throw Error{WTB_ERR_ARG, "unknown overlap convention"};
        StageTimer tm(c, 1);
        launch_bse_build(c, *coulomb, precision, convention);
        tm.stop();
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int wtb_set_temperature(wtb_ctx *ctx, double temp_k) {
    return guarded(ctx, [&](Ctx &c) {
        if (!(temp_k == temp_k)) throw Error{WTB_ERR_ARG, "temperature is NaN"};
        c.temperature = temp_k < 0.0 ? -1.0 : temp_k;
        // a store built at another temperature is stale
        if (c.bse.valid && c.bse.temperature != c.temperature) {
            c.bse.valid = false;
            c.sol.valid = false;
        }
    });
}

int wtb_bse_solve(wtb_ctx *ctx, int nev, double tol, int max_it, double *E, double *X, wtb_report *rep) {
    return guarded(ctx, [&](Ctx &c) {
        if (!c.bse.valid) throw Error{WTB_ERR_STATE, "wtb_bse_build must run before wtb_bse_solve"};
        if (nev <= 0 || nev > c.bse.N) throw Error{WTB_ERR_ARG, "nev out of range"};
        if (!(tol > 0.0) || max_it <= 0) throw Error{WTB_ERR_ARG, "tol and max_it must be positive"};
        wtb_report r{};
        lobpcg(c, nev, tol, max_it, c.bse.precision, &r);
        if (rep) *rep = r;
        if (E) copy(c, E, c.sol.evals.p, sizeof(double) * nev, hipMemcpyDeviceToHost);
        if (X) copy(c, X, c.sol.evecs.p, sizeof(cd) * (size_t)c.bse.N * nev, hipMemcpyDeviceToHost);
        if (r.nconv < nev)
            throw Error{WTB_ERR_NOCONV, "eigensolver: " + std::to_string(r.nconv) + " of " + std::to_string(nev) +
                                            " pairs converged"};
    });
}

int wtb_optics(wtb_ctx *ctx, double cshift, double *d) {
    return guarded(ctx, [&](Ctx &c) {
        StageTimer tm(c, 2);
        launch_optics(c, cshift);
        tm.stop();
        if (d) {
            const long N = (long)c.bands.nk * c.bands.nc * c.bands.nv;
            copy(c, d, c.dip.p, sizeof(cd) * 5 * (size_t)N, hipMemcpyDeviceToHost);
        }
    });
}

int wtb_oscillator(wtb_ctx *ctx, int nev, int comp_a, int comp_b, double *act) {
    return guarded(ctx, [&](Ctx &c) {
        if (!c.sol.valid) throw Error{WTB_ERR_STATE, "wtb_bse_solve must run before wtb_oscillator"};
        if (!c.dip_valid) throw Error{WTB_ERR_STATE, "wtb_optics must run before wtb_oscillator"};
        if (nev <= 0 || nev > c.sol.nev) throw Error{WTB_ERR_ARG, "nev exceeds the solved pairs"};
        if (comp_a < 0 || comp_a > 4 || comp_b < 0 || comp_b > 4) throw Error{WTB_ERR_ARG, "component"};
        if (!act) throw Error{WTB_ERR_ARG, "null output"};
        if (c.bse.finite_q) throw Error{WTB_ERR_UNSUPPORTED, "oscillator strengths of a finite-Q store"};
        const long N = c.bse.N;
        DBuf<double> a;
        a.alloc(nev);
        StageTimer tm(c, 3);
        launch_oscillator(c, nev, comp_a, comp_b, a.p);
        tm.stop();
        std::fill(act, act + N, 0.0);
        WTB_HIP_CHECK(hipMemcpyAsync(act, a.p, sizeof(double) * nev, hipMemcpyDeviceToHost, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int wtb_bse_get_block(wtb_ctx *ctx, long i0, long ni, long j0, long nj, double *out) {
    return guarded(ctx, [&](Ctx &c) {
        if (!c.bse.valid) throw Error{WTB_ERR_STATE, "no BSE matrix"};
        const long N = c.bse.N;
        if (i0 < 0 || j0 < 0 || ni <= 0 || nj <= 0 || i0 + ni > N || j0 + nj > N || !out)
            throw Error{WTB_ERR_ARG, "block out of range"};
        DBuf<cd> tmp;
        tmp.alloc((size_t)ni * nj);
        bse_block_to_c64(c, i0, ni, j0, nj, tmp.p, ni);
        WTB_HIP_CHECK(hipMemcpyAsync(out, tmp.p, sizeof(cd) * ni * nj, hipMemcpyDeviceToHost, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int wtb_bse_matvec(wtb_ctx *ctx, int ncol, const double *X, double *Y, int precision) {
    return guarded(ctx, [&](Ctx &c) {
        if (!c.bse.valid) throw Error{WTB_ERR_STATE, "no BSE matrix"};
        if (ncol <= 0 || !X || !Y) throw Error{WTB_ERR_ARG, "invalid matvec arguments"};
        if (precision != WTB_FP64 && precision != WTB_MIXED && precision != WTB_FP32 && precision != WTB_INT8X)
            throw Error{WTB_ERR_ARG, "unknown precision"};
        if (precision == WTB_INT8X && c.bse.precision == WTB_FP64)
            throw Error{WTB_ERR_UNSUPPORTED, "the int8-digit product runs on the FP32 store"};
        const long N = c.bse.N;
        DBuf<cd> x, y;
        x.alloc((size_t)N * ncol);
        y.alloc((size_t)N * ncol);
        WTB_HIP_CHECK(hipMemcpyAsync(x.p, X, sizeof(cd) * N * ncol, hipMemcpyHostToDevice, c.stream));
        matvec(c, precision, ncol, x.p, y.p);
        WTB_HIP_CHECK(hipMemcpyAsync(Y, y.p, sizeof(cd) * N * ncol, hipMemcpyDeviceToHost, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int wtb_bse_matvec_noalloc(wtb_ctx *ctx, int ncol, const double *X, double *Y, int precision) {
    return guarded(ctx, [&](Ctx &c) {
        if (!c.bse.valid) throw Error{WTB_ERR_STATE, "no BSE matrix"};
        if (ncol <= 0 || !X || !Y) throw Error{WTB_ERR_ARG, "invalid matvec arguments"};
        if (precision != WTB_FP64 && precision != WTB_MIXED && precision != WTB_FP32)
            throw Error{WTB_ERR_ARG, "precision must be WTB_FP64, WTB_MIXED or WTB_FP32"};
        // X and Y must be device memory: a host pointer would fault the kernels
        for (const void *p : {static_cast<const void *>(X), static_cast<const void *>(Y)}) {
            hipPointerAttribute_t at;
            hipError_t e = hipPointerGetAttributes(&at, p);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error{WTB_ERR_ARG, "X and Y must be device pointers"};
            }
            if (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)
                throw Error{WTB_ERR_ARG, "X and Y must be device pointers"};
        }
        matvec_noalloc(c, precision, ncol, reinterpret_cast<const cd *>(X), reinterpret_cast<cd *>(Y));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int wtb_stage_times(wtb_ctx *ctx, double *ms) {
    return guarded(ctx, [&](Ctx &c) {
        if (!ms) throw Error{WTB_ERR_ARG, "null output"};
        for (int i = 0; i < 4; ++i) ms[i] = c.stage_ms[i];
    });
}

int wtb_bse_diag(wtb_ctx *ctx, double *diag) {
    return guarded(ctx, [&](Ctx &c) {
        if (!c.bse.valid) throw Error{WTB_ERR_STATE, "no BSE matrix"};
        if (!diag) throw Error{WTB_ERR_ARG, "null output"};
        copy(c, diag, c.bse.diag.p, sizeof(double) * c.bse.N, hipMemcpyDeviceToHost);
    });
}

int wtb_grad_energies(wtb_ctx *ctx, const double *w, wtb_grad *g, double *hop_re_grad, double *hop_im_grad) {
    return guarded(ctx, [&](Ctx &c) {
        if (c.bse.temperature >= 0.0)
            throw Error{WTB_ERR_UNSUPPORTED, "gradients of the temperature BSE are not available"};
        if (c.bse.finite_q) throw Error{WTB_ERR_UNSUPPORTED, "gradients of the finite-Q BSE are not available"};
        if (c.bse.coul.kind == WTB_V1DT) throw Error{WTB_ERR_UNSUPPORTED, "gradients with V1DT are not available"};
        grad_energies(c, w, g, hop_re_grad, hop_im_grad);
    });
}

int wtb_grad_spectrum(wtb_ctx *ctx, int comp, double cshift, double gamma, int nw, const double *omega,
                      const double *target, double *loss, double *eps2, wtb_grad *g, double *hop_re_grad,
                      double *hop_im_grad) {
    return guarded(ctx, [&](Ctx &c) {
        if (c.bse.temperature >= 0.0)
            throw Error{WTB_ERR_UNSUPPORTED, "gradients of the temperature BSE are not available"};
        if (c.bse.finite_q) throw Error{WTB_ERR_UNSUPPORTED, "gradients of the finite-Q BSE are not available"};
        if (c.bse.coul.kind == WTB_V1DT) throw Error{WTB_ERR_UNSUPPORTED, "gradients with V1DT are not available"};
        grad_spectrum(c, comp, cshift, gamma, nw, omega, target, loss, eps2, g, hop_re_grad, hop_im_grad);
    });
}

namespace {
__global__ void upload_diag(long N, long ld, const cd *__restrict__ H, double *__restrict__ d) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < N) d[i] = H[i + i * ld].x;
}
}  // namespace

int wtb_slepc_solver_compat(const double *HX, double *EX, double *HE, int NX, int NE) {
    if (!HX || !EX || NX <= 0 || NE <= 0 || NE > NX) return WTB_ERR_ARG;
    wtb_ctx *ctx = nullptr;
    int st = wtb_create(&ctx, 0);
    if (st) return st;
    st = guarded(ctx, [&](Ctx &c) {
        const long N = NX, ld = store_ld(N);
        c.bse.h64.alloc((size_t)ld * N);
        c.bse.diag.alloc(N);
        // the host matrix has leading dimension N; the store pads it to ld with zero rows
        WTB_HIP_CHECK(hipMemsetAsync(c.bse.h64.p, 0, sizeof(cd) * ld * N, c.stream));
        WTB_HIP_CHECK(hipMemcpy2DAsync(c.bse.h64.p, sizeof(cd) * ld, HX, sizeof(cd) * N, sizeof(cd) * N, N,
                                       hipMemcpyHostToDevice, c.stream));
        WTB_HIP_CHECK(hipStreamSynchronize(c.stream));
        hipLaunchKernelGGL(upload_diag, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.stream, N, ld,
                           c.bse.h64.p, c.bse.diag.p);
        WTB_LAUNCH_CHECK();
        c.bse.N = N;
        c.bse.ld = ld;
        c.bse.precision = WTB_FP64;
        c.bse.valid = true;
        // EPS_SMALLEST_MAGNITUDE (module_slepc_solver.F90:65): the NE eigenvalues
        // of least |E|.  Up to 1024 every eigenpair comes from the dense solver
        // and the set is chosen from all of them, so an indefinite spectrum
        // gets its interior pairs.  Above, the block solver returns the lowest
        // algebraic pairs: the same set whenever the spectrum is non-negative
        // (the BSE case, bound excitons above zero); a negative lowest
        // eigenvalue is refused rather than answered with the wrong set.
        std::vector<double> e;
        std::vector<cd> x;
        int nget = NE;
        if (N <= 1024) {
            nget = (int)N;
            DBuf<cd> A, Z;
            DBuf<double> W;
            A.alloc((size_t)N * N);
            Z.alloc((size_t)N * N);
            W.alloc(N);
            bse_block_to_c64(c, 0, N, 0, N, A.p, N);
            heev_lowest(c, (int)N, (int)N, A.p, (int)N, W.p, Z.p, (int)N);
            e.resize(N);
            x.resize((size_t)N * N);
            copy(c, e.data(), W.p, sizeof(double) * N, hipMemcpyDeviceToHost);
            copy(c, x.data(), Z.p, sizeof(cd) * x.size(), hipMemcpyDeviceToHost);
        } else {
            wtb_report r{};
            lobpcg(c, NE, 1e-8, 200, WTB_FP64, &r);
            e.resize(NE);
            x.resize((size_t)N * NE);
            copy(c, e.data(), c.sol.evals.p, sizeof(double) * NE, hipMemcpyDeviceToHost);
            copy(c, x.data(), c.sol.evecs.p, sizeof(cd) * x.size(), hipMemcpyDeviceToHost);
            // a Ritz value bounds the lowest eigenvalue from above, so a
            // negative one proves the matrix indefinite, converged or not
            if (e[0] < 0.0)
                throw Error{WTB_ERR_UNSUPPORTED, "indefinite matrix above N = 1024: the smallest-magnitude "
                                                 "pairs are interior eigenvalues"};
            if (r.nconv < NE) throw Error{WTB_ERR_NOCONV, "not converged"};
        }
        // ascending |E|, ties by position (stable)
        std::vector<int> ord(nget);
        for (int i = 0; i < nget; ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return std::fabs(e[a]) < std::fabs(e[b]); });
        for (int i = 0; i < NE; ++i) {
            EX[i] = e[ord[i]];
            if (HE) std::memcpy(HE + 2 * (size_t)N * i, &x[(size_t)N * ord[i]], sizeof(cd) * N);
        }
    });
    wtb_destroy(ctx);
    return st;
}

}  // extern "C"
