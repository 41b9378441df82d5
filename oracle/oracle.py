"""ctypes binding of the CPU FP64 oracle (``oracle/libwtb_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``, never by the product package.
Array conventions are those documented in ``wtb_oracle.h`` (row-major,
complex128 interleaved).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libwtb_oracle.so")
_lib = None

COMPAT_F32 = 1
GAUGE_LAPACK = 2
COULOMB_KINDS = {"V2DK": 1, "V3D": 2, "V3DL": 3, "V2DT": 4, "V2DT2": 5, "V2DOH": 6,
                 "V2DRK": 7, "V0DT": 8, "V1DT": 9}


class _Coulomb(C.Structure):
    _fields_ = [("kind", C.c_int), ("r0", C.c_double), ("ediel", C.c_double * 3),
                ("ez", C.c_double), ("w", C.c_double), ("lc", C.c_double), ("ktol", C.c_double)]


class _Model(C.Structure):
    _fields_ = [("norb", C.c_int), ("nvec", C.c_int), ("rvec", C.c_void_p), ("deg", C.c_void_p),
                ("hop", C.c_void_p), ("rlat", C.c_double * 9), ("efermi", C.c_double),
                ("scissor", C.c_double), ("flags", C.c_int)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        d, i, p = C.c_double, C.c_int, C.c_void_p
        _lib.or_pi_f32.restype = d
        _lib.or_cic.restype = d
        _lib.or_vcell2d.restype = d
        _lib.or_vcell3d.restype = d
        _lib.or_coulomb_v.restype = d
        _lib.or_coulomb_v.argtypes = [p, p, p, p, p]
        _lib.or_heev_jacobi.restype = i
        _lib.or_hk_eig.restype = i
        _lib.or_bse_matrix.restype = i
        _lib.or_dipoles.restype = i
        _lib.or_optdiel.argtypes = [i, d, i, p, d, p, p, d, p, p]
        _lib.or_spectrum.argtypes = [i, p, p, i, p, d, d, d, i, d, i, p]
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


class Model:
    """Keeps the numpy buffers of a TB model alive for the C struct."""

    def __init__(self, tb, flags=COMPAT_F32 | GAUGE_LAPACK):
        self.rvec = np.ascontiguousarray(tb.rvec, np.int32)
        self.deg = np.ascontiguousarray(tb.deg, np.int32)
        self.hop = np.ascontiguousarray(tb.hop, np.complex128)
        self.s = _Model(tb.norb, tb.nvec, _p(self.rvec).value, _p(self.deg).value,
                        _p(self.hop).value, (C.c_double * 9)(*np.asarray(tb.rlat, float).ravel()),
                        float(tb.efermi), float(tb.scissor), int(flags))
        self.norb = tb.norb
        self.rlat = np.ascontiguousarray(tb.rlat, np.float64)

    @property
    def ref(self):
        return C.byref(self.s)


def coulomb(kind="V2DRK", r0=1.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001):
    return _Coulomb(COULOMB_KINDS[kind] if isinstance(kind, str) else int(kind), float(r0),
                    (C.c_double * 3)(*ediel), float(ez), float(w), float(lc), float(ktol))


def pi_f32():
    return lib().or_pi_f32()


def cic():
    return lib().or_cic()


def mp_mesh(n1, n2, n3, rlat, shift=(0.0, 0.0, 0.0)):
    out = np.zeros((n1 * n2 * n3, 3))
    sh = np.asarray(shift, float)
    lib().or_mp_mesh(n1, n2, n3, _p(sh), _p(np.ascontiguousarray(rlat, float)), _p(out))
    return out


def recvec(rlat):
    out = np.zeros((3, 3))
    lib().or_recvec(_p(np.ascontiguousarray(rlat, float)), _p(out))
    return out


def hk(model: Model, k):
    out = np.zeros((model.norb, model.norb), np.complex128)
    lib().or_hk(model.ref, _p(np.ascontiguousarray(k, float)), _p(out))
    return out


def dhk(model: Model, k):
    n = model.norb
    out = np.zeros((3, n, n), np.complex128)
    lib().or_dhk(model.ref, _p(np.ascontiguousarray(k, float)), _p(out[0]), _p(out[1]), _p(out[2]))
    return out


def heev_jacobi(a):
    a = np.array(a, np.complex128, order="C")
    n = a.shape[0]
    w = np.zeros(n)
    z = np.zeros((n, n), np.complex128)
    st = lib().or_heev_jacobi(n, _p(a), _p(w), _p(z))
    if st < 0:
        raise RuntimeError("oracle Jacobi did not converge")
    return w, z


def householder_q(h):
    h = np.ascontiguousarray(h, np.complex128)
    q = np.zeros_like(h)
    lib().or_householder_q(h.shape[0], _p(h), _p(q))
    return q


def gauge_lapack(h, z):
    h = np.ascontiguousarray(h, np.complex128)
    z = np.array(z, np.complex128, order="C")
    lib().or_gauge_lapack(h.shape[0], _p(h), _p(z))
    return z


def hk_eig(model: Model, kpt, nv, nc, full=False):
    kpt = np.ascontiguousarray(kpt, float)
    nk, n, nb = kpt.shape[0], model.norb, nc + nv
    eig = np.zeros((nk, nb))
    vec = np.zeros((nk, nb, n), np.complex128)
    nocp = np.zeros(nk, np.int32)
    eall = np.zeros((nk, n)) if full else None
    zall = np.zeros((nk, n, n), np.complex128) if full else None
    st = lib().or_hk_eig(model.ref, nk, _p(kpt), nv, nc, _p(eig), _p(vec), _p(nocp),
                         _p(eall) if full else None, _p(zall) if full else None)
    res = dict(eig=eig, vec=vec, nocp=nocp, status=st)
    if full:
        res.update(eall=eall, zall=zall)
    return res


def bessel_k01(x):
    """(K0(x), K1(x)) of the oracle's quadrature (used by V1DT)."""
    a, b = C.c_double(), C.c_double()
    lib().or_bessel_k01(C.c_double(float(x)), C.byref(a), C.byref(b))
    return a.value, b.value


def coulomb_v(c, k1, k2, ngrid, rlat):
    return lib().or_coulomb_v(C.byref(c), _p(np.ascontiguousarray(k1, float)),
                              _p(np.ascontiguousarray(k2, float)),
                              _p(np.ascontiguousarray(ngrid, np.int32)),
                              _p(np.ascontiguousarray(rlat, float)))


def bse_matrix(kpt, eig, vec, c, ngrid, rlat, nc, nv, textbook=False):
    nk, nb, norb = vec.shape
    N = nk * nc * nv
    h = np.zeros((N, N), np.complex128)
    st = lib().or_bse_matrix(nk, nc, nv, norb, _p(np.ascontiguousarray(kpt, float)),
                             _p(np.ascontiguousarray(eig, float)),
                             _p(np.ascontiguousarray(vec, np.complex128)), C.byref(c),
                             _p(np.ascontiguousarray(ngrid, np.int32)),
                             _p(np.ascontiguousarray(rlat, float)), int(textbook), _p(h))
    if st:
        raise RuntimeError(f"oracle bse_matrix status {st}")
    return h


def fermidisteh(ec, ev, temp):
    """f_v - f_c with the Fermi level at 0 (fermidisteh, src/module_math.F90:410-421):
    the Boltzmann constant is the single-precision literal 8.617330350E-5 and
    T = 0 is replaced by the single-precision 1E-8."""
    kb = float(np.float32(8.617330350e-5))
    t = float(np.float32(1e-8)) if temp == 0.0 else float(temp)
    beta = 1.0 / (kb * t)
    with np.errstate(over="ignore"):
        return 1.0 / (np.exp(np.asarray(ev) * beta) + 1.0) - 1.0 / (np.exp(np.asarray(ec) * beta) + 1.0)


def state_factors(eig, nc, nv, temp):
    """fdeh(i) of every BSE state i = k nc nv + c nv + v (bsesolvertemp,
    src/module_bse.F90:2984-2988): f_v - f_c of its band pair."""
    eig = np.asarray(eig, float)
    nk = eig.shape[0]
    ec = np.repeat(eig[:, nv:nv + nc], nv, axis=1).reshape(nk * nc * nv)
    ev = np.tile(eig[:, :nv], (1, nc)).reshape(nk * nc * nv)
    return fermidisteh(ec, ev, temp), ec - ev


def bse_matrix_temp(kpt, eig, vec, c, ngrid, rlat, nc, nv, temp, textbook=False):
    """Temperature BSE matrix (matrizelbsetemp, src/module_bse.F90:580-660, filled
    for j >= i and read through its upper triangle by ZHEEV('U'),
    :3057-3087): H(i,j) = f(i) K(i,j) for i < j with K the plain kernel of
    bse_matrix, H(j,i) = conj(H(i,j)), and H(i,i) = (E_c - E_v) + V(0) f(i)."""
    K = bse_matrix(kpt, eig, vec, c, ngrid, rlat, nc, nv, textbook=textbook)
    f, de = state_factors(eig, nc, nv, temp)
    n = K.shape[0]
    idx = np.arange(n)
    H = K * f[np.minimum.outer(idx, idx)]
    v0 = np.diag(K).real - de
    H[idx, idx] = de + v0 * f
    return H


def splice_q(eig, vec, eigq, vecq, nv):
    """Bands of the finite-Q basis (bsebnds, src/module_bse.F90:1436-1442):
    valence slots j < nv from the k mesh, conduction slots from the k+Q mesh."""
    e = np.array(eig, float)
    u = np.array(vec, np.complex128)
    e[:, nv:] = np.asarray(eigq)[:, nv:]
    u[:, nv:, :] = np.asarray(vecq)[:, nv:, :]
    return e, u


def exchange_amplitudes(vec, nc, nv):
    """a_i = <c(k+Q)|v(k)> = sum_o conj(c_o) v_o of every state
    i = k nc nv + c nv + v of the spliced bands (vcv of matrizelbsekq,
    src/module_bse.F90:551-555)."""
    u = np.asarray(vec, np.complex128)
    nk = u.shape[0]
    cb = u[:, nv:nv + nc, :]                   # (nk, nc, norb)
    vb = u[:, :nv, :]                          # (nk, nv, norb)
    return np.einsum("kco,kvo->kcv", cb.conj(), vb).reshape(nk * nc * nv)


def bse_matrix_q(kpt, eig, vec, eigq, vecq, c, ngrid, rlat, nc, nv, q, temp=None, exchange=True,
                 textbook=False):
    """Finite-Q BSE matrix (matrizelbsekq, src/module_bse.F90:452-569, filled
    in bsebnds :1444-1454).  Conduction bands at k+Q, valence at k, the direct
    term V(k_i - k_j) <c_i|c_j><v_i|v_j> of bse_matrix on those bands and, for
    |Q| != 0, the exchange -V(Q) <c_i|v_i><v_j|c_j> on every element.  With
    temp: matrizelbsekqtemp (:664-780), the whole kernel part times f of the
    row state, read through the upper triangle as bse_matrix_temp."""
    e, u = splice_q(eig, vec, eigq, vecq, nv)
    K = bse_matrix(kpt, e, u, c, ngrid, rlat, nc, nv, textbook=textbook)
    qv = np.asarray(q, float)
    if exchange and math.sqrt(qv[0] ** 2 + qv[1] ** 2 + qv[2] ** 2) != 0.0:
        a = exchange_amplitudes(u, nc, nv)
        vq = coulomb_v(c, qv, np.zeros(3), np.asarray(ngrid, np.int32), rlat)
        K = K - vq * np.outer(a, a.conj())
    if temp is None:
        return K
    f, de = state_factors(e, nc, nv, temp)
    n = K.shape[0]
    idx = np.arange(n)
    H = K * f[np.minimum.outer(idx, idx)]
    H[idx, idx] = de + (np.diag(K).real - de) * f
    return H


def optics_bz(model: Model, kpt, eig, vec, nc, nv, sme):
    """spoptpolbz (src/module_bse.F90:4532-4797): per k the sums over the band
    pairs (c outer, v inner) of |d_a|^2 with d_a = <c|dH/dk_a|v> / (E_c - E_v
    + i sme) (optspbz, src/module_optics.F90:257-311) and sp, sm = (d_x +-
    i d_y) / sqrt(2) with the default-real 1/sqrt(2.).  Returns (nk, 5)."""
    nk = kpt.shape[0]
    norm = float(np.float32(1.0) / np.sqrt(np.float32(2.0)))
    act = np.zeros((nk, 5))
    for k in range(nk):
        h = dhk(model, kpt[k])
        for c in range(nc):
            uc = vec[k, nv + c]
            for v in range(nv):
                uv = vec[k, v]
                den = complex(eig[k, nv + c] - eig[k, v], sme)
                dx, dy, dz = (uc.conj() @ h[a] @ uv / den for a in range(3))
                sp, sm = norm * (dx + 1j * dy), norm * (dx - 1j * dy)
                act[k] += [abs(dx) ** 2, abs(dy) ** 2, abs(dz) ** 2, abs(sp) ** 2, abs(sm) ** 2]
    return act


def dos(eall, zall, sigma, ndos, orbweight=None, spin_mode=0):
    """dostool (src/module_tools.F90:390-651): the grid from min(0, lowest
    band) to max(0, highest band) and, per grid energy, (1/nk) sum over all
    states of weight * gaussian (endos :103-149, gaussian
    src/module_math.F90:330-339 with the single-precision pi); weights 1,
    spin up / down (spincont :41-99; spin_mode 1 halves, 2 odd/even) and
    sum_o w_o |z_o|^2 (layercont :153-175).  zall[k, o, n] is component o of
    band n.  Returns en (ndos,), d (ndos, 4)."""
    eall = np.asarray(eall, float)
    nk, n = eall.shape
    e0 = min(0.0, float(eall[:, 0].min()))
    e1 = max(0.0, float(eall[:, -1].max()))
    en = e0 + ((e1 - e0) * np.arange(ndos)) / (ndos - 1.0)
    p = np.abs(np.asarray(zall)) ** 2
    w = np.ones(n) if orbweight is None else np.asarray(orbweight, float)
    h = n // 2
    up = np.zeros((nk, n))
    dn = np.zeros((nk, n))
    if spin_mode == 1:
        up, dn = p[:, :h, :].sum(1), p[:, h:2 * h, :].sum(1)
    elif spin_mode == 2:
        up, dn = p[:, 0:2 * h:2, :].sum(1), p[:, 1:2 * h:2, :].sum(1)
    lw = np.einsum("kon,o->kn", p, w)
    pi32 = float(np.float32(np.pi))
    g = (1.0 / (sigma * math.sqrt(2.0 * pi32))) * np.exp(-((en[:, None] - eall.ravel()[None, :]) ** 2)
                                                       / (2.0 * sigma * sigma))
    wts = np.stack([np.ones(nk * n), up.ravel(), dn.ravel(), lw.ravel()], 1)
    return en, (g @ wts) / nk


def dipoles(model: Model, kpt, eig, vec, nc, nv, cshift):
    nk = kpt.shape[0]
    N = nk * nc * nv
    d = np.zeros((5, N), np.complex128)
    lib().or_dipoles(model.ref, nk, _p(np.ascontiguousarray(kpt, float)), nc, nv,
                     _p(np.ascontiguousarray(eig, float)),
                     _p(np.ascontiguousarray(vec, np.complex128)), C.c_double(cshift), _p(d))
    return d


def oscillator(x, d1, d2):
    """x: (N, nev) exciton amplitudes (column n = state n)."""
    x = np.asfortranarray(x, np.complex128)
    n, nev = x.shape
    act = np.zeros(n)
    lib().or_oscillator(n, nev, _p(x), _p(np.ascontiguousarray(d1, np.complex128)),
                        _p(np.ascontiguousarray(d2, np.complex128)), _p(act))
    return act


def bse_rows(kpt, eig, vec, c, ngrid, rlat, nc, nv, i0, ni):
    nk, nb, norb = vec.shape
    N = nk * nc * nv
    h = np.zeros((ni, N), np.complex128)
    lib().or_bse_rows(nk, nc, nv, norb, _p(np.ascontiguousarray(kpt, float)),
                      _p(np.ascontiguousarray(eig, float)), _p(np.ascontiguousarray(vec, np.complex128)),
                      C.byref(c), _p(np.ascontiguousarray(ngrid, np.int32)),
                      _p(np.ascontiguousarray(rlat, float)), C.c_long(i0), C.c_long(ni), _p(h))
    return h


def matvec_rows(h, x):
    h = np.ascontiguousarray(h, np.complex128)
    x = np.ascontiguousarray(x, np.complex128)
    y = np.zeros((h.shape[0], x.shape[1]), np.complex128)
    lib().or_matvec_rows(C.c_long(h.shape[0]), C.c_long(h.shape[1]), x.shape[1], _p(h), _p(x), _p(y))
    return y


def krylov_schur(h, nev, tol=1e-8, max_it=200, seed=1, vectors=True):
    """Krylov-Schur for the nev least-|E| pairs of the row-major Hermitian h
    (slepc_solver's settings): (E, X or None, restarts, matvecs, nconv)."""
    h = np.ascontiguousarray(h, np.complex128)
    N = h.shape[0]
    e = np.zeros(nev)
    x = np.zeros((nev, N), np.complex128) if vectors else None
    its = C.c_int(0)
    nmv = C.c_long(0)
    nconv = lib().or_krylov_schur(C.c_long(N), _p(h), int(nev), C.c_double(tol), int(max_it), C.c_uint(seed),
                                  _p(e), _p(x) if vectors else None, C.byref(its), C.byref(nmv))
    return e, (x.T if vectors else None), its.value, nmv.value, nconv


def num_threads():
    return lib().or_num_threads()


def spectrum(ex, fosc, diagonal, ngrid, vc, e0, e1, npts, sme, renorm=True):
    ex = np.ascontiguousarray(ex, float)
    fosc = np.ascontiguousarray(fosc, float)
    out = np.zeros((npts, 3))
    lib().or_spectrum(len(ex), _p(ex), _p(fosc), int(diagonal),
                      _p(np.ascontiguousarray(ngrid, np.int32)), vc, e0, e1, npts, sme,
                      int(renorm), _p(out))
    return out
