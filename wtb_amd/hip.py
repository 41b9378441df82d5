"""ctypes binding of the gfx950 exciton library ``libwtb_hip.so``.

The library is the product path: there is no CPU fallback.  Loading fails
loudly when the shared object is missing.  Arrays cross the boundary in the
reference's Fortran layouts (see ``include/wtb_hip.h``); this module converts
to and from numpy.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwtb_hip.so")

WTB_OK = 0
ERRORS = {-1: "WTB_ERR_ARG", -2: "WTB_ERR_HIP", -3: "WTB_ERR_NOCONV", -4: "WTB_ERR_NOCP",
          -5: "WTB_ERR_NOMEM", -6: "WTB_ERR_STATE", -7: "WTB_ERR_UNSUPPORTED"}
FLAG_COMPAT_F32 = 1
FLAG_GAUGE_LAPACK = 2
FP64, MIXED, FP32, INT8X = 0, 1, 2, 3
CONV_REFERENCE, CONV_TEXTBOOK = 0, 1
COULOMB = {"V2DK": 1, "V3D": 2, "V3DL": 3, "V2DT": 4, "V2DT2": 5, "V2DOH": 6, "V2DRK": 7,
           "V0DT": 8, "V1DT": 9}
COMPONENTS = {"x": 0, "y": 1, "z": 2, "sp": 3, "sm": 4}

EXPORTS = ["wtb_create", "wtb_destroy", "wtb_last_error", "wtb_set_tb_model", "wtb_set_kpoints",
           "wtb_hk_eig", "wtb_get_bands", "wtb_bse_build", "wtb_bse_solve", "wtb_optics",
           "wtb_oscillator", "wtb_bse_get_block", "wtb_bse_matvec", "wtb_bse_diag",
           "wtb_slepc_solver_compat", "wtb_grad_energies", "wtb_grad_spectrum", "wtb_set_temperature",
           "wtb_hk_eig_q", "wtb_dos", "wtb_optics_bz", "wtb_stage_times", "wtb_bse_matvec_noalloc"]


class WTBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class GradScalars(C.Structure):
    _fields_ = [("coulomb", C.c_double * 7), ("scissor", C.c_double), ("efermi", C.c_double),
                ("gauge_residual", C.c_double)]


COULOMB_PARAMS = ("r0", "ediel_t", "ediel", "ediel_b", "ez", "w", "lc")


class CoulombParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("r0", C.c_double), ("ediel", C.c_double * 3),
                ("ez", C.c_double), ("w", C.c_double), ("lc", C.c_double), ("ktol", C.c_double)]


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("nconv", C.c_int), ("matvec_columns", C.c_long),
                ("max_rel_residual", C.c_double), ("t_solve_ms", C.c_double),
                ("t_refine_ms", C.c_double), ("t_matvec_ms", C.c_double), ("t_rr_ms", C.c_double),
                ("t_kernel_ms", C.c_double), ("kernel_launches", C.c_int), ("kernel_bytes", C.c_double),
                ("kernel_flops", C.c_double), ("t_refine_kernel_ms", C.c_double),
                ("refine_launches", C.c_int), ("refine_flops", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def load():
    """Load libwtb_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build the HIP library first "
                          "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    P, I, D, L = C.c_void_p, C.c_int, C.c_double, C.c_long
    sig = {
        "wtb_create": [C.POINTER(P), I],
        "wtb_destroy": [P],
        "wtb_set_tb_model": [P, I, I, P, P, P, P, P, D, D, I],
        "wtb_set_kpoints": [P, I, P, P],
        "wtb_hk_eig": [P, I, I, P, P, P],
        "wtb_hk_eig_q": [P, P, P, I, I, P, P, P],
        "wtb_dos": [P, I, D, P, I, P, P],
        "wtb_optics_bz": [P, D, P],
        "wtb_get_bands": [P, P, P],
        "wtb_bse_build": [P, C.POINTER(CoulombParams), I, I],
        "wtb_bse_solve": [P, I, D, I, P, P, C.POINTER(Report)],
        "wtb_optics": [P, D, P],
        "wtb_oscillator": [P, I, I, I, P],
        "wtb_bse_get_block": [P, L, L, L, L, P],
        "wtb_bse_matvec": [P, I, P, P, I],
        "wtb_bse_diag": [P, P],
        "wtb_slepc_solver_compat": [P, P, P, I, I],
        "wtb_grad_energies": [P, P, C.POINTER(GradScalars), P, P],
        "wtb_grad_spectrum": [P, I, D, D, I, P, P, P, P, C.POINTER(GradScalars), P, P],
        "wtb_set_temperature": [P, D],
        "wtb_stage_times": [P, P],
        "wtb_bse_matvec_noalloc": [P, I, P, P, I],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = I
    lib.wtb_last_error.argtypes = [P]
    lib.wtb_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def coulomb_params(kind="V2DRK", r0=1.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001):
    k = COULOMB[kind] if isinstance(kind, str) else int(kind)
    return CoulombParams(k, float(r0), (C.c_double * 3)(*[float(x) for x in ediel]), float(ez),
                         float(w), float(lc), float(ktol))


class Context:
    """One device context (opaque ``wtb_ctx``) with its own HIP stream."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        st = self.lib.wtb_create(C.byref(self.h), device)
        if st != WTB_OK:
            raise WTBError(st, f"wtb_create(device={device}) failed")
        self.N = 0
        self.nk = 0
        self.norb = 0

    def close(self):
        if self.h:
            self.lib.wtb_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != WTB_OK:
            raise WTBError(st, self.lib.wtb_last_error(self.h).decode())

    # ---- boundary calls ----
    def set_tb_model(self, model, flags=FLAG_COMPAT_F32 | FLAG_GAUGE_LAPACK):
        hop = np.asarray(model.hop, np.complex128)
        rvec = np.asfortranarray(model.rvec, np.int32)
        deg = np.ascontiguousarray(model.deg, np.int32)
        hre = np.asfortranarray(hop.real)
        him = np.asfortranarray(hop.imag)
        rlat = np.asfortranarray(model.rlat, np.float64)
        self._check(self.lib.wtb_set_tb_model(self.h, model.norb, model.nvec, _ptr(rvec), _ptr(deg),
                                              _ptr(hre), _ptr(him), _ptr(rlat), float(model.efermi),
                                              float(model.scissor), int(flags)))
        self.norb = model.norb
        self.nvec = model.nvec

    def set_kpoints(self, kpt, ngrid):
        k = np.asfortranarray(kpt, np.float64)
        g = np.ascontiguousarray(ngrid, np.int32)
        self._check(self.lib.wtb_set_kpoints(self.h, k.shape[0], _ptr(k), _ptr(g)))
        self.nk = k.shape[0]

    def hk_eig(self, nv, nc, host=True):
        """Bands of every k.  host=False keeps them on the device only (they
        feed the assembly there) and returns None."""
        nb = nv + nc
        self.N = self.nk * nc * nv
        if not host:
            self._check(self.lib.wtb_hk_eig(self.h, nv, nc, None, None, None))
            return None
        eig = np.zeros((self.nk, nb), order="F")
        vec = np.zeros((self.nk, nb, self.norb), np.complex128, order="F")
        nocp = np.zeros(self.nk, np.int32)
        self._check(self.lib.wtb_hk_eig(self.h, nv, nc, _ptr(eig), _ptr(vec), _ptr(nocp)))
        return eig, vec, nocp

    def hk_eig_q(self, q, qpt, nv, nc):
        """Finite-Q exciton (BSE_BND=T): bands on the k+Q mesh qpt (nk, 3) for
        the conduction states of the following bse_build, which adds the
        exchange term when |Q| != 0.  q is the Cartesian Q (3,); q None
        switches finite Q off.  Returns the k+Q (eig, vec, nocp)."""
        if q is None:
            self._check(self.lib.wtb_hk_eig_q(self.h, None, None, nv, nc, None, None, None))
            return None
        qv = np.ascontiguousarray(q, np.float64).reshape(3)
        k = np.asfortranarray(qpt, np.float64)
        if k.shape != (self.nk, 3):
            raise ValueError("qpt must be (nk, 3)")
        nb = nv + nc
        eig = np.zeros((self.nk, nb), order="F")
        vec = np.zeros((self.nk, nb, self.norb), np.complex128, order="F")
        nocp = np.zeros(self.nk, np.int32)
        self._check(self.lib.wtb_hk_eig_q(self.h, _ptr(qv), _ptr(k), nv, nc, _ptr(eig), _ptr(vec), _ptr(nocp)))
        return eig, vec, nocp

    def optics_bz(self, sme):
        """Per-k optical activity (OPT_BZ=T): (nk, 5) sums of |d_a|^2 for
        a = x, y, z, sp, sm over the selected band pairs."""
        act = np.zeros((self.nk, 5), order="F")
        self._check(self.lib.wtb_optics_bz(self.h, float(sme), _ptr(act)))
        return act

    def dos(self, ndos, sigma, orbweight=None, spin_mode=0):
        """Density of states of the bands of the last hk_eig (dostool):
        returns (en (ndos,), dos (ndos, 4): tdos, up, down, weighted)."""
        en = np.zeros(ndos)
        d = np.zeros((ndos, 4), order="F")
        ow = None if orbweight is None else np.ascontiguousarray(orbweight, np.float64)
        self._check(self.lib.wtb_dos(self.h, int(ndos), float(sigma), None if ow is None else _ptr(ow),
                                     int(spin_mode), _ptr(en), _ptr(d)))
        return en, d

    def get_bands(self):
        e = np.zeros((self.nk, self.norb), order="F")
        v = np.zeros((self.nk, self.norb, self.norb), np.complex128, order="F")
        self._check(self.lib.wtb_get_bands(self.h, _ptr(e), _ptr(v)))
        return e, v

    def bse_build(self, coulomb, precision=FP64, convention=CONV_REFERENCE):
        if isinstance(coulomb, dict):
            coulomb = coulomb_params(**coulomb)
        self._check(self.lib.wtb_bse_build(self.h, C.byref(coulomb), int(precision), int(convention)))

    def set_temperature(self, temp_k):
        """Temperature BSE (BSET=T) for the following bse_build: kernel
        elements H(i,j), i <= j, times f_v - f_c of state i, and the oscillator
        strengths weighted accordingly; temp_k < 0 (None) switches it off."""
        self._check(self.lib.wtb_set_temperature(self.h, -1.0 if temp_k is None else float(temp_k)))

    def bse_solve(self, nev, tol=1e-8, max_it=200, vectors=True, check=True):
        """Lowest nev pairs.  With check=False a WTB_ERR_NOCONV status is
        reported through rep['nconv'] instead of raising."""
        E = np.zeros(nev)
        X = np.zeros((self.N, nev), np.complex128, order="F") if vectors else None
        rep = Report()
        st = self.lib.wtb_bse_solve(self.h, nev, float(tol), int(max_it), _ptr(E), _ptr(X), C.byref(rep))
        if not (st == -3 and not check):
            self._check(st)
        return E, X, rep.as_dict()

    def optics(self, cshift):
        d = np.zeros((self.N, 5), np.complex128, order="F")
        self._check(self.lib.wtb_optics(self.h, float(cshift), _ptr(d)))
        return d

    def oscillator(self, nev, a="x", b=None):
        b = a if b is None else b
        act = np.zeros(self.N)
        self._check(self.lib.wtb_oscillator(self.h, nev, COMPONENTS[a], COMPONENTS[b], _ptr(act)))
        return act

    def get_block(self, i0, ni, j0, nj):
        out = np.zeros((ni, nj), np.complex128, order="F")
        self._check(self.lib.wtb_bse_get_block(self.h, i0, ni, j0, nj, _ptr(out)))
        return out

    def matvec(self, X, precision=FP64):
        X = np.asfortranarray(X, np.complex128)
        if X.ndim == 1:
            X = X[:, None]
        Y = np.zeros_like(X, order="F")
        self._check(self.lib.wtb_bse_matvec(self.h, X.shape[1], _ptr(X), _ptr(Y), int(precision)))
        return Y

    def stage_times(self):
        """device ms of the last hk_eig, bse_build, optics and oscillator calls"""
        ms = np.zeros(4)
        self._check(self.lib.wtb_stage_times(self.h, _ptr(ms)))
        return dict(zip(("hk_eig", "bse_build", "optics", "oscillator"), ms.tolist()))

    def matvec_noalloc(self, x_dev, y_dev, ncol, precision=MIXED):
        """Y = H X on device arrays (addresses of complex128 (N, ncol)
        column-major buffers) without any device allocation by the library"""
        self._check(self.lib.wtb_bse_matvec_noalloc(self.h, int(ncol), C.c_void_p(x_dev), C.c_void_p(y_dev),
                                                     int(precision)))

    def diag(self):
        d = np.zeros(self.N)
        self._check(self.lib.wtb_bse_diag(self.h, _ptr(d)))
        return d

    def grad_energies(self, w, hop=True):
        """Gradient of L = sum_n w[n] E_n over the states of the last solve:
        {'coulomb': {name: dL/dparam}, 'scissor', 'efermi', 'gauge_residual',
        'hop': complex (nvec, norb, norb) dL/dRe H(R) + i dL/dIm H(R)}."""
        w = np.ascontiguousarray(w, np.float64)
        g = GradScalars()
        gre = gim = None
        if hop:
            gre = np.zeros((self.nvec, self.norb, self.norb), order="F")
            gim = np.zeros((self.nvec, self.norb, self.norb), order="F")
        self._check(self.lib.wtb_grad_energies(self.h, _ptr(w), C.byref(g),
                                               _ptr(gre) if hop else None, _ptr(gim) if hop else None))
        out = {"coulomb": dict(zip(COULOMB_PARAMS, list(g.coulomb))), "scissor": g.scissor,
               "efermi": g.efermi, "gauge_residual": g.gauge_residual}
        if hop:
            out["hop"] = gre + 1j * gim
        return out

    def grad_spectrum(self, omega, target, comp="x", cshift=0.05, gamma=0.05, hop=True):
        """Spectrum loss L = sum_j (eps2(omega_j) - target_j)^2 over the states
        of the last solve and its gradient (the keys of grad_energies plus
        'loss' and 'eps2')."""
        omega = np.ascontiguousarray(omega, np.float64)
        target = np.ascontiguousarray(target, np.float64)
        if omega.shape != target.shape or omega.ndim != 1:
            raise ValueError("omega and target must be 1-D arrays of one length")
        g = GradScalars()
        loss = C.c_double(0.0)
        eps2 = np.zeros(len(omega))
        gre = gim = None
        if hop:
            gre = np.zeros((self.nvec, self.norb, self.norb), order="F")
            gim = np.zeros((self.nvec, self.norb, self.norb), order="F")
        self._check(self.lib.wtb_grad_spectrum(self.h, COMPONENTS[comp], cshift, gamma, len(omega), _ptr(omega),
                                               _ptr(target), C.byref(loss), _ptr(eps2), C.byref(g),
                                               _ptr(gre) if hop else None, _ptr(gim) if hop else None))
        out = {"loss": loss.value, "eps2": eps2, "coulomb": dict(zip(COULOMB_PARAMS, list(g.coulomb))),
               "scissor": g.scissor, "efermi": g.efermi, "gauge_residual": g.gauge_residual}
        if hop:
            out["hop"] = gre + 1j * gim
        return out


def slepc_solver_compat(HX, NE):
    """Drop-in of slepc_solver(HX,EX,HE,NX,NE) on a host matrix."""
    lib = load()
    HX = np.asfortranarray(HX, np.complex128)
    NX = HX.shape[0]
    EX = np.zeros(NE)
    HE = np.zeros((NX, NE), np.complex128, order="F")
    st = lib.wtb_slepc_solver_compat(_ptr(HX), _ptr(EX), _ptr(HE), NX, NE)
    if st != WTB_OK:
        raise WTBError(st, "wtb_slepc_solver_compat failed")
    return EX, HE


@dataclass
class Workload:
    """A prepared exciton problem (model + mesh + physics parameters)."""
    model: object
    kpt: np.ndarray
    ngrid: tuple
    nv: int
    nc: int
    coulomb: dict
    cshift: float = 0.05
