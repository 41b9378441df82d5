"""CPU oracle checks: every oracle function is pinned by an independent route
(closed forms, LAPACK through scipy, invariants, brute force)."""
import math

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import hex_model
from oracle import oracle as O
from wtb_amd import tbmodel as T


def rand_herm(rng, n):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a + a.conj().T


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 22, 31, 32])
def test_jacobi_matches_lapack(rng, n):
    a = rand_herm(rng, n)
    w, z = O.heev_jacobi(a)
    assert np.allclose(w, np.linalg.eigvalsh(a), atol=1e-12 * max(1, np.abs(w).max()))
    assert np.allclose(a @ z, z * w, atol=1e-11)
    assert np.allclose(z.conj().T @ z, np.eye(n), atol=1e-12)


@pytest.mark.parametrize("n", [2, 3, 7, 22])
def test_householder_q_reduces_to_real_tridiagonal(rng, n):
    a = rand_herm(rng, n)
    q = O.householder_q(a)
    t = q.conj().T @ a @ q
    assert np.allclose(q.conj().T @ q, np.eye(n), atol=1e-12)
    assert np.abs(np.triu(t, 2)).max(initial=0) < 1e-12
    assert np.abs(np.diag(t, -1).imag).max(initial=0) < 1e-12
    # the same reduction as LAPACK zhetrd('L'): identical tridiagonal
    _, d, e, _, info = lapack.zhetrd(a, lower=1)
    assert info == 0
    assert np.allclose(np.diag(t).real, d, atol=1e-11)
    assert np.allclose(np.diag(t, -1).real, e, atol=1e-11)


@pytest.mark.parametrize("n", [2, 5, 22])
def test_gauge_matches_zheev_up_to_real_sign(rng, n):
    a = rand_herm(rng, n)
    w, z = O.heev_jacobi(a)
    z = O.gauge_lapack(a, z)
    w2, v2, info = lapack.zheev(a, compute_v=1, lower=1)
    assert info == 0
    ov = np.sum(z.conj() * v2, axis=0)
    assert np.allclose(np.abs(ov.real), 1.0, atol=1e-10)
    assert np.abs(ov.imag).max() < 1e-10


def test_constants():
    assert O.pi_f32() == float(np.float32(np.pi))
    assert O.cic() == float(-(np.float32(0.0904756) * np.float32(1000.0)))
    assert O.cic() == -90.47559356689453


def test_recvec_and_mesh():
    m = T.honeycomb_two_band()
    b = O.recvec(m.rlat)
    assert np.allclose(b @ m.rlat.T, 2 * O.pi_f32() * np.eye(3), atol=1e-12)
    k = O.mp_mesh(4, 3, 2, m.rlat, shift=(0.1, 0.0, 0.0))
    assert k.shape == (24, 3)
    # i outer, l inner: the second point steps along b3
    assert np.allclose(k[1] - k[0], b[2] / 2)
    assert np.allclose(k[2] - k[0], b[1] / 3)
    assert np.allclose(k[6] - k[0], b[0] / 4)
    assert np.allclose(k[0], 0.1 * b[0])
    assert np.allclose(O.mp_mesh(24, 24, 1, m.rlat), T.mp_mesh(24, 24, 1, m.rlat))


def test_honeycomb_closed_form():
    m = T.honeycomb_two_band()
    k = O.mp_mesh(24, 24, 1, m.rlat)
    r = O.hk_eig(O.Model(m, flags=0), k, 1, 1)
    a1, a2 = m.rlat[0], m.rlat[1]
    f = 1 + np.exp(-1j * k @ a1) + np.exp(-1j * k @ a2)
    e = np.sqrt(4.0 + 2.8 ** 2 * np.abs(f) ** 2)
    assert r["status"] == 0
    assert np.all(r["nocp"] == 1)
    assert np.abs(r["eig"][:, 1] - e).max() < 1e-12
    assert np.abs(r["eig"][:, 0] + e).max() < 1e-12
    # gap 4 eV at K = (b1 - b2)/3 lies on the 24x24 mesh
    assert abs(np.min(r["eig"][:, 1] - r["eig"][:, 0]) - 4.0) < 1e-9


def test_f32_phase_compat_changes_hk_at_float_level():
    m = hex_model((8, 8, 1))
    k = O.mp_mesh(8, 8, 1, m.rlat)[13]
    h64 = O.hk(O.Model(m, flags=0), k)
    h32 = O.hk(O.Model(m, flags=O.COMPAT_F32), k)
    d = np.abs(h64 - h32).max()
    assert 0 < d < 1e-4
    # the float32 phase is exactly what the compat build evaluates
    rc = m.rvec @ m.rlat
    ph = np.float32(k @ rc.T).astype(np.float64)
    href = np.einsum("r,rmn->mn", np.exp(1j * ph) / m.deg, m.hop) - float(np.float32(m.efermi)) * np.eye(m.norb)
    assert np.abs(href - h32).max() < 1e-12


def test_eigsys_scissor_and_band_selection():
    m = hex_model((6, 6, 1))
    m.scissor = 0.25
    k = O.mp_mesh(6, 6, 1, m.rlat)
    r = O.hk_eig(O.Model(m, flags=0), k, 2, 2, full=True)
    assert np.all(r["nocp"] == 11)
    w = np.linalg.eigvalsh(m.hk(k))
    w[:, 11:] += 0.25
    assert np.abs(r["eall"] - w).max() < 1e-10
    assert np.abs(r["eig"] - w[:, 9:13]).max() < 1e-10
    # selected vectors are eigenvectors of H(k) for the selected bands
    h = m.hk(k)
    for kk in (0, 7, 35):
        for j in range(4):
            v = r["vec"][kk, j]
            e = r["eig"][kk, j] - (0.25 if j >= 2 else 0)
            assert np.abs(h[kk] @ v - e * v).max() < 1e-10


def test_band_selection_errors():
    m = hex_model((4, 4, 1))
    k = O.mp_mesh(4, 4, 1, m.rlat)
    assert O.hk_eig(O.Model(m), k, 12, 1)["status"] == -4   # nv > nocp
    m2 = m.copy()
    m2.efermi = 1e3                                         # no band above E_F
    assert O.hk_eig(O.Model(m2), k, 1, 1)["status"] == -4


# ----------------------------- Coulomb ------------------------------------

def _v(kind, q, rlat, ngrid=(10, 10, 1), **kw):
    c = O.coulomb(kind, **kw)
    return O.coulomb_v(c, np.array([q, 0.0, 0.0]), np.zeros(3), np.array(ngrid, np.int32), rlat)


def test_v2drk_isotropic_closed_form():
    m = T.honeycomb_two_band()
    area = np.linalg.norm(np.cross(m.rlat[0], m.rlat[1]))
    nk = 100
    for q in (0.002, 0.05, 0.7):
        for r0, eps, w in ((10.0, 1.0, 0.0), (40.0, 2.5, 0.3)):
            v = _v("V2DRK", q, m.rlat, r0=r0, ediel=(eps, eps, eps), ez=eps, w=w, lc=7.0)
            # p_t = p_b = 0 when all permittivities agree: eps_w = eps + r0 q e^{-q w}
            ref = O.cic() * math.exp(-q * w) / (nk * area * q * (eps + r0 * q * math.exp(-q * w)))
            assert abs(v - ref) <= 1e-12 * abs(ref)
    assert _v("V2DRK", 5e-4, m.rlat, r0=10.0) == 0.0


def test_v3d_and_v3dl_closed_form():
    rlat = np.eye(3) * 4.0
    for q in (0.01, 0.3):
        ref = O.cic() / (10 * 10 * 1 * 64.0 * q * q)
        assert abs(_v("V3D", q, rlat) - ref) < 1e-12 * abs(ref)
        assert abs(_v("V3DL", q, rlat, ediel=(1.0, 6.0, 1.0)) - ref / 6.0) < 1e-12 * abs(ref)
    assert _v("V3DL", 1e-4, rlat, ediel=(1.0, 6.0, 1.0)) == 0.0


def test_v0dt_regularisation_is_the_q_to_zero_limit():
    rlat = np.eye(3) * 4.0
    v0 = _v("V0DT", 1e-4, rlat)
    vq = _v("V0DT", 2e-3, rlat)      # (1 - cos(cr q))/q^2 -> cr^2/2
    assert abs(vq - v0) < 1e-5 * abs(v0)


def test_v2dk_and_v2dt2_limits():
    m = T.honeycomb_two_band()
    area = np.linalg.norm(np.cross(m.rlat[0], m.rlat[1]))
    # V2DK with EDIEL=1 has r0 = 0: pure 2D Coulomb cic/(N A q ed)
    q = 0.2
    ref = O.cic() / (100 * area * q * 1.0)
    assert abs(_v("V2DK", q, m.rlat) - ref) < 1e-12 * abs(ref)
    # V2DT2 -> (1 - e^{-q lc/2})/q^2 / (N V) for in-plane q
    vol = area * 20.0
    lc = 8.0
    ref = O.cic() / (100 * vol) / q ** 2 * (1 - math.exp(-0.5 * q * lc))
    assert abs(_v("V2DT2", q, m.rlat, lc=lc) - ref) < 1e-12 * abs(ref)


# ----------------------------- assembly -----------------------------------

def _small_problem(mesh=(6, 6, 1), nv=2, nc=2, flags=O.COMPAT_F32 | O.GAUGE_LAPACK):
    m = hex_model(mesh)
    k = O.mp_mesh(*mesh, m.rlat)
    r = O.hk_eig(O.Model(m, flags=flags), k, nv, nc)
    assert r["status"] == 0
    return m, k, r


COUL = dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)


def _block_form(k, r, c, ngrid, rlat, nv, nc, textbook=False):
    """Independent route: V(k,k') (M^c kron M^v) with numpy einsum."""
    vec, eig = r["vec"], r["eig"]
    nk = k.shape[0]
    V = np.array([[O.coulomb_v(c, k[i], k[j], np.array(ngrid, np.int32), rlat) for j in range(nk)]
                  for i in range(nk)])
    uc, uv = vec[:, nv:], vec[:, :nv]
    Mc = np.einsum("kco,pdo->kcpd", uc.conj(), uc)
    Mv = np.einsum("kvo,pwo->kvpw", uv.conj(), uv)
    if textbook:
        Mv = Mv.conj()
    H = np.einsum("kp,kcpd,kvpw->kcvpdw", V, Mc, Mv).reshape(nk * nc * nv, -1)
    de = (eig[:, nv:, None] - eig[:, None, :nv]).reshape(-1)
    H[np.diag_indices_from(H)] = de + np.diag(V).repeat(nc * nv)
    return H


@pytest.mark.parametrize("textbook", [False, True])
def test_bse_matrix_block_form_and_hermitian(textbook):
    m, k, r = _small_problem()
    c = O.coulomb(**COUL)
    H = O.bse_matrix(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2, textbook=textbook)
    assert np.abs(H - H.conj().T).max() == 0.0
    Hb = _block_form(k, r, c, (6, 6, 1), m.rlat, 2, 2, textbook)
    assert np.abs(H - Hb).max() < 1e-14


def test_bse_noninteracting_limit():
    m, k, r = _small_problem()
    c = O.coulomb(**dict(COUL, kind="V2DRK"))
    H = O.bse_matrix(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2)
    D = np.diag(H).real
    c0 = O.coulomb(**dict(COUL, ktol=1e9))     # every q below KTOL -> V = 0
    H0 = O.bse_matrix(k, r["eig"], r["vec"], c0, (6, 6, 1), m.rlat, 2, 2)
    assert np.abs(H0 - np.diag(D)).max() == 0.0
    assert np.allclose(np.linalg.eigvalsh(H0), np.sort(D))


def test_bse_energies_gauge_invariant_both_conventions(rng):
    m, k, r = _small_problem()
    c = O.coulomb(**COUL)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, r["vec"].shape[:2]))
    vec2 = r["vec"] * ph[:, :, None]
    for tb in (False, True):
        w1 = np.linalg.eigvalsh(O.bse_matrix(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2, tb))
        w2 = np.linalg.eigvalsh(O.bse_matrix(k, r["eig"], vec2, c, (6, 6, 1), m.rlat, 2, 2, tb))
        assert np.abs(w1 - w2).max() < 1e-12
    wr = np.linalg.eigvalsh(O.bse_matrix(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2, False))
    wt = np.linalg.eigvalsh(O.bse_matrix(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2, True))
    assert np.abs(wr - wt).max() > 1e-6      # the two conventions are different physics


# ------------------------- temperature BSE (BSET=T) -----------------------

def test_fermidisteh_closed_forms():
    kb = float(np.float32(8.617330350e-5))
    e = np.array([0.05, 0.3, 0.9, 2.0])
    for T in (300.0, 3000.0, 20000.0):
        # symmetric pair (E_v = -E_c): f_v - f_c = tanh(E_c / (2 kB T))
        assert np.allclose(O.fermidisteh(e, -e, T), np.tanh(e / (2 * kb * T)), rtol=0, atol=1e-15)
    # T = 0 is taken as 1e-8 K: fully occupied valence, empty conduction
    assert np.array_equal(O.fermidisteh(e, -e, 0.0), np.ones(4))
    # both bands on the same side of the Fermi level: no transition weight at T -> 0
    assert np.array_equal(O.fermidisteh(np.array([0.5]), np.array([0.2]), 0.0), [0.0])
    # very high T: occupations tend to 1/2 and the factor to 0
    assert np.abs(O.fermidisteh(e, -e, 1e9)).max() < 1e-4


def test_temperature_gap_shift_closed_forms():
    """TA= VE / BE / FA gap shifts (gapcortemp, gapcortemp2) at points where
    coth and the Bose factor take known values."""
    kb = float(np.float32(8.617330350e-5))
    s, ph = 4e-4, 0.025
    # coth(x) = 2 at x = ln(3) / 2: the VE shift is -S <hw>
    t = ph / (2 * kb * (math.log(3.0) / 2))
    assert abs(T.temperature_gap_shift("VE", s, ph, t) + s * ph) < 1e-15
    # T -> 0: coth -> 1 and no shift; high T, x = <hw> / 2 kB T small:
    # coth x = 1/x + x/3 + O(x^3), the shift -2 S kB T + S <hw> - S <hw> x / 3
    assert T.temperature_gap_shift("VE", s, ph, 1.0) == 0.0
    hi = 1e5
    x = ph / (2 * kb * hi)
    assert abs(T.temperature_gap_shift("VE", s, ph, hi) - (-2 * s * kb * hi + s * ph - s * ph * x / 3)) < 1e-13
    # Bose factor 1 at exp(<hw> / T) = 2
    assert abs(T.temperature_gap_shift("BE", s, ph, ph / math.log(2.0)) + 2 * s) < 1e-15
    assert T.temperature_gap_shift("FA", s, ph, 300.0) == 0.0
    assert T.temperature_gap_shift("XX", s, ph, 300.0) == 0.0


def test_bse_matrix_temp_fill_rule():
    """The vectorised f[min(i, j)] rule against a loop restatement of the
    reference's fill (matrizelbsetemp for j >= i with the factor of row i, the
    lower triangle read as the conjugate)."""
    m, k, r = _small_problem()
    c = O.coulomb(**COUL)
    T = 8000.0
    K = O.bse_matrix(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2)
    H = O.bse_matrix_temp(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2, T)
    n = K.shape[0]
    nk = r["eig"].shape[0]
    L = np.zeros_like(K)
    for i in range(n):
        kk, cc, vv = i // 4, (i // 2) % 2, i % 2
        ec, ev = r["eig"][kk, 2 + cc], r["eig"][kk, vv]
        fi = float(O.fermidisteh(ec, ev, T))
        for j in range(i, n):
            if i == j:
                L[i, i] = (ec - ev) + (K[i, i].real - (ec - ev)) * fi
            else:
                L[i, j] = K[i, j] * fi
                L[j, i] = np.conj(L[i, j])
    assert nk * 4 == n
    assert np.abs(H - L).max() < 1e-15 * np.abs(K).max()
    assert np.array_equal(H, H.conj().T)
    f, _ = O.state_factors(r["eig"], 2, 2, T)
    assert 0.3 < f.min() < 0.99           # a visible effect at this temperature
    # T -> 0: every factor is 1 and the plain kernel comes back
    H0 = O.bse_matrix_temp(k, r["eig"], r["vec"], c, (6, 6, 1), m.rlat, 2, 2, 0.0)
    assert np.abs(H0 - K).max() < 1e-14 * np.abs(K).max()


# ----------------------------- finite Q -----------------------------------

def _q_problem(qfrac, mesh=(3, 3, 1), nv=2, nc=2, flags=O.COMPAT_F32 | O.GAUGE_LAPACK, pi=T.PI_F32):
    m = hex_model(mesh)
    k = T.mp_mesh(*mesh, m.rlat, pi=pi)
    q = np.asarray(qfrac, float) @ T.recip_vectors(m.rlat, pi)
    kq = T.mp_mesh_q(*mesh, m.rlat, q, pi=pi)
    om = O.Model(m, flags=flags)
    r, rq = O.hk_eig(om, k, nv, nc), O.hk_eig(om, kq, nv, nc)
    assert r["status"] == 0 and rq["status"] == 0
    return m, k, q, r, rq


def test_mp_mesh_q_is_the_shifted_mesh():
    m = hex_model((4, 4, 1))
    q = np.array([0.031, -0.017, 0.0])
    kq = T.mp_mesh_q(4, 4, 1, m.rlat, q, shift=(0.1, 0.0, 0.0))
    k = T.mp_mesh(4, 4, 1, m.rlat, shift=(0.1, 0.0, 0.0))
    assert np.abs(kq - (k + q)).max() < 1e-15
    assert np.array_equal(T.mp_mesh_q(4, 4, 1, m.rlat, np.zeros(3)), T.mp_mesh(4, 4, 1, m.rlat))


@pytest.mark.parametrize("temp", [None, 9000.0])
def test_bse_matrix_q_against_the_element_formula(temp):
    """The finite-Q matrix (splice + direct kernel + rank-one exchange) against
    a loop over matrizelbsekq's (and matrizelbsekqtemp's) element formula for
    j >= i, with the lower triangle its conjugate."""
    mesh, nv, nc = (3, 3, 1), 2, 2
    m, k, q, r, rq = _q_problem((0.11, -0.07, 0.0), mesh, nv, nc)
    c = O.coulomb(**COUL)
    H = O.bse_matrix_q(k, r["eig"], r["vec"], rq["eig"], rq["vec"], c, mesh, m.rlat, nc, nv, q, temp=temp)
    g = np.array(mesh, np.int32)
    vq = O.coulomb_v(c, q, np.zeros(3), g, m.rlat)
    assert vq != 0.0
    N = k.shape[0] * nc * nv
    L = np.zeros((N, N), np.complex128)
    for i in range(N):
        ki, ci, vi = i // (nc * nv), (i // nv) % nc, i % nv
        c1, v1 = rq["vec"][ki, nv + ci], r["vec"][ki, vi]
        ec1, ev1 = rq["eig"][ki, nv + ci], r["eig"][ki, vi]
        f1 = 1.0 if temp is None else O.fermidisteh(ec1, ev1, temp)
        for j in range(i, N):
            kj, cj, vj = j // (nc * nv), (j // nv) % nc, j % nv
            c2, v2 = rq["vec"][kj, nv + cj], r["vec"][kj, vj]
            vk = O.coulomb_v(c, k[ki], k[kj], g, m.rlat)
            vcv = np.sum(c1.conj() * v1)
            vvc = np.sum(v2.conj() * c2)
            if i == j:
                L[i, i] = (ec1 - ev1) + (vk - vq * vcv * vvc) * f1
            else:
                vc = np.sum(c1.conj() * c2)
                vv = np.sum(v1.conj() * v2)
                L[i, j] = (vk * vc * vv - vq * vcv * vvc) * f1
                L[j, i] = np.conj(L[i, j])
    assert np.abs(H - L).max() < 1e-13 * np.abs(L).max()
    # the exchange is visible at this Q
    H0 = O.bse_matrix_q(k, r["eig"], r["vec"], rq["eig"], rq["vec"], c, mesh, m.rlat, nc, nv, q, temp=temp,
                        exchange=False)
    assert np.abs(H - H0).max() > 1e-6 * np.abs(H).max()


def test_exchange_is_a_positive_rank_one_update():
    """-V(Q) a a^H with V(Q) < 0: one positive singular direction, and the
    spectra interlace, w0[i] <= w1[i] <= w0[i + 1]."""
    mesh = (3, 3, 1)
    m, k, q, r, rq = _q_problem((0.08, 0.05, 0.0), mesh)
    c = O.coulomb(**COUL)
    args = (k, r["eig"], r["vec"], rq["eig"], rq["vec"], c, mesh, m.rlat, 2, 2, q)
    H1, H0 = O.bse_matrix_q(*args), O.bse_matrix_q(*args, exchange=False)
    X = H1 - H0
    s = np.linalg.svd(X, compute_uv=False)
    assert s[1] < 1e-12 * s[0]
    assert np.linalg.eigvalsh(X)[-1] > 0 and np.linalg.eigvalsh(X)[0] > -1e-12 * s[0]
    w0, w1 = np.linalg.eigvalsh(H0), np.linalg.eigvalsh(H1)
    eps = 1e-10 * np.abs(w0).max()
    assert np.all(w0 <= w1 + eps)
    assert np.all(w1[:-1] <= w0[1:] + eps)


def test_finite_q_at_zero_and_at_a_reciprocal_vector():
    """Q = 0 is the plain kernel; Q = G (H(k + G) = H(k) with FP64 phases and
    an exact 2 pi) gives the plain spectrum, the exchange amplitudes
    <c(k)|v(k)> being zero."""
    mesh = (3, 3, 1)
    c = O.coulomb(**COUL)
    m, k, q, r, rq = _q_problem((0.0, 0.0, 0.0), mesh)
    K = O.bse_matrix(k, r["eig"], r["vec"], c, mesh, m.rlat, 2, 2)
    assert np.array_equal(O.bse_matrix_q(k, r["eig"], r["vec"], rq["eig"], rq["vec"], c, mesh, m.rlat, 2, 2, q), K)
    m, k, q, r, rq = _q_problem((1.0, 0.0, 0.0), mesh, flags=O.GAUGE_LAPACK, pi=math.pi)
    HG = O.bse_matrix_q(k, r["eig"], r["vec"], rq["eig"], rq["vec"], c, mesh, m.rlat, 2, 2, q)
    K = O.bse_matrix(k, r["eig"], r["vec"], c, mesh, m.rlat, 2, 2)
    assert np.abs(np.linalg.eigvalsh(HG) - np.linalg.eigvalsh(K)).max() < 1e-10
    assert np.abs(O.exchange_amplitudes(O.splice_q(r["eig"], r["vec"], rq["eig"], rq["vec"], 2)[1], 2, 2)).max() < 1e-12


# ------------------------------ optics ------------------------------------

def test_dhk_matches_finite_difference():
    m = hex_model((6, 6, 1))
    M = O.Model(m, flags=0)
    k = np.array([0.13, -0.21, 0.0])
    d = O.dhk(M, k)
    h = 1e-5
    for a in range(2):
        e = np.zeros(3)
        e[a] = h
        fd = (O.hk(M, k + e) - O.hk(M, k - e)) / (2 * h)
        assert np.abs(d[a] - fd).max() < 1e-6 * np.abs(d[a]).max()


def test_dipoles_definition_and_circular():
    m, k, r = _small_problem(flags=0)
    M = O.Model(m, flags=0)
    d = O.dipoles(M, k, r["eig"], r["vec"], 2, 2, 0.05)
    kk, c, v = 5, 1, 0
    i = kk * 4 + c * 2 + v
    dh = O.dhk(M, k[kk])
    de = r["eig"][kk, 2 + c] - r["eig"][kk, v]
    fac = (1 / (de + 0.05j)).real
    for a in range(3):
        ref = r["vec"][kk, 2 + c].conj() @ dh[a] @ r["vec"][kk, v] * fac
        assert abs(d[a, i] - ref) < 1e-12
    assert np.allclose(d[3], (d[0] + 1j * d[1]) / math.sqrt(2))
    assert np.allclose(d[4], (d[0] - 1j * d[1]) / math.sqrt(2))


def test_oscillator_phase_rule(rng):
    """|sum_j X_j d_j|^2 has no conjugate on X: a conduction phase e^{i t}
    multiplies term j by e^{-2 i t} (reference convention)."""
    n, nev = 40, 3
    x = rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev))
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = O.oscillator(x, d, d)
    assert np.allclose(a[:nev], np.abs(x.T @ d) ** 2)
    assert np.all(a[nev:] == 0)
    # real sign flips of a basis state leave it unchanged
    s = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    assert np.allclose(O.oscillator(x * s[:, None], d * s, d * s), a)
    t = 0.3
    x2, d2 = x.copy(), d.copy()
    x2[0] *= np.exp(-1j * t)
    d2[0] *= np.exp(-1j * t)
    b = O.oscillator(x2, d2, d2)
    ref = np.abs(x.T @ d + (np.exp(-2j * t) - 1) * x[0] * d[0]) ** 2
    assert np.allclose(b[:nev], ref)


# ----------------------------- spectrum -----------------------------------

def test_single_line_lorentzian():
    rlat = np.eye(3) * 3.0
    vc = 27.0
    out = O.spectrum(np.array([2.0]), np.array([1.5]), True, np.array([4, 4, 1], np.int32), vc,
                     1.0, 3.0, 201, 0.05, renorm=False)
    caux = float(np.float32(180.90)) / (16 * vc)
    w = out[:, 0]
    assert np.allclose(w, np.linspace(1.0, 3.0, 201))
    assert np.allclose(out[:, 2], caux * 1.5 * 0.05 / ((w - 2.0) ** 2 + 0.05 ** 2))
    assert np.allclose(out[:, 1], 1.0 + caux * 1.5 * (2.0 - w) / ((w - 2.0) ** 2 + 0.05 ** 2))


def test_avgsme_telescopes_and_renorm():
    ex = np.array([1.0, 1.5, 2.5])
    f = np.array([1.0, 0.0, 2.0])
    out = O.spectrum(ex, f, False, np.array([2, 2, 1], np.int32), 10.0, 0.0, 4.0, 41, 0.1, renorm=False)
    gam = 0.1 + 10 * (2.5 - 1.0) / 3
    caux = float(np.float32(180.90)) / (4 * 10.0)
    w = out[:, 0]
    ref = caux * sum(fi * gam / ((w - ei) ** 2 + gam ** 2) for ei, fi in zip(ex, f))
    assert np.allclose(out[:, 2], ref)
    assert np.allclose(out[:, 1], caux * sum(fi * (ei - w) / ((w - ei) ** 2 + gam ** 2) for ei, fi in zip(ex, f)))
    rn = O.spectrum(ex, f, False, np.array([2, 2, 1], np.int32), 10.0, 0.0, 4.0, 41, 0.1, renorm=True)
    cut = 1.0 - float(np.float32(0.15))
    ib = np.max(np.nonzero(w <= cut)[0])
    assert np.all(rn[:ib + 1, 2] == 0)
    assert np.allclose(rn[ib + 1:, 2], np.maximum(out[ib + 1:, 2] - out[ib + 1, 2], 0))


# --------------------------- TB file I/O -----------------------------------

def test_tb_file_roundtrip(tmp_path):
    m = hex_model((4, 4, 1))
    p = tmp_path / "tb.dat"
    T.write_tb_file(m, str(p))
    m2 = T.read_tb_file(str(p))
    assert m2.systype == m.systype and m2.norb == m.norb and m2.nvec == m.nvec
    assert np.array_equal(m2.rvec, m.rvec) and np.array_equal(m2.deg, m.deg)
    assert np.abs(m2.hop - m.hop).max() < 1e-15
    assert abs(m2.efermi - m.efermi) < 1e-15
    assert m.hermiticity_error() < 1e-15


# ----------------------------- DOS ----------------------------------------

def test_dos_flat_bands_and_weight_identities():
    """Two flat bands at -D/2 and +D/2 (onsite only): the grid runs exactly
    from -D/2 to D/2 and its end points carry norm (1 + exp(-D^2 / 2 s^2));
    with unit weights on all orbitals the weighted density is the total, and
    up + down is the total for either spin ordering."""
    D, s = 2.0, 0.3
    m = T.honeycomb_two_band(delta=D, t=0.0)
    k = O.mp_mesh(4, 4, 1, m.rlat)
    r = O.hk_eig(O.Model(m, flags=0), k, 1, 1, full=True)
    en, d = O.dos(r["eall"], r["zall"], s, 41)
    assert en[0] == -D / 2 and abs(en[-1] - D / 2) < 1e-15
    norm = 1.0 / (s * math.sqrt(2.0 * float(np.float32(np.pi))))
    assert abs(d[0, 0] - norm * (1.0 + math.exp(-D * D / (2 * s * s)))) < 1e-14
    assert abs(d[20, 0] - 2.0 * norm * math.exp(-(D / 2) ** 2 / (2 * s * s))) < 1e-14
    mh = hex_model((4, 4, 1))
    kh = O.mp_mesh(4, 4, 1, mh.rlat)
    rh = O.hk_eig(O.Model(mh, flags=0), kh, 2, 2, full=True)
    for mode in (1, 2):
        en, d = O.dos(rh["eall"], rh["zall"], 0.1, 301, spin_mode=mode)
        assert np.abs(d[:, 1] + d[:, 2] - d[:, 0]).max() < 1e-12 * d[:, 0].max()
        assert np.abs(d[:, 3] - d[:, 0]).max() < 1e-12 * d[:, 0].max()
    # weights of the first half of the orbitals: the spin-up density of the halves ordering
    w = np.r_[np.ones(11), np.zeros(11)]
    en, d = O.dos(rh["eall"], rh["zall"], 0.1, 301, orbweight=w, spin_mode=1)
    assert np.abs(d[:, 3] - d[:, 1]).max() < 1e-12 * d[:, 0].max()
    # the integral over the window: sum over states of the Gaussian mass
    # inside [e0, e1] (erf), against the trapezoid of the curve
    en, d = O.dos(rh["eall"], rh["zall"], 0.05, 4001)
    z0 = (en[0] - rh["eall"].ravel()) / (0.05 * math.sqrt(2.0))
    z1 = (en[-1] - rh["eall"].ravel()) / (0.05 * math.sqrt(2.0))
    from scipy.special import erf
    mass = 0.5 * (erf(z1) - erf(z0)).sum() / 16
    assert 20.0 < mass < 22.0
    assert abs(np.trapezoid(d[:, 0], en) - mass) < 1e-3 * mass


# ----------------------------- V1DT ---------------------------------------

def test_bessel_k01_against_scipy():
    from scipy.special import k0, k1
    for x in (1e-3, 0.01, 0.1, 0.5, 1.0, 3.0, 10.0, 40.0, 200.0):
        a, b = O.bessel_k01(x)
        assert abs(a - k0(x)) <= 1e-14 * k0(x), x
        assert abs(b - k1(x)) <= 1e-14 * k1(x), x


def test_v1dt_branches():
    """v1dt's three branches: the q -> 0 closed form, the J0 log integral for
    q_x = 0 (a 100-point trapezoid, checked with scipy's J0), and the Bessel
    form 1 + q_yz lc J1 K0 + |q_x| lc J0 K1 with scipy's special functions."""
    from scipy.special import j0, j1, k0, k1
    rlat = np.eye(3) * np.array([3.0, 20.0, 20.0])
    g = np.array([10, 1, 1], np.int32)
    lc = 6.0
    c = O.coulomb("V1DT", lc=lc, ktol=0.001)
    pre = O.cic() / (10 * 3.0 * 400.0)
    z = np.zeros(3)
    assert abs(O.coulomb_v(c, z, z, g, rlat) - pre * (lc * lc / 16.0) * (2 * math.log(lc) - 1)) < 1e-12 * abs(pre)
    qyz = 0.3
    r = np.arange(100) / 99.0 * lc + float(np.float32(1e-4))
    f = r * j0(r * qyz) * np.log(r)
    ref = pre * np.sum((r[1] - r[0]) / 2.0 * (f[:-1] + f[1:]))
    assert abs(O.coulomb_v(c, np.array([0.0, qyz, 0.0]), z, g, rlat) - ref) < 1e-12 * abs(ref)
    for qx, qy in ((0.2, 0.0), (0.05, 0.4), (1.3, 0.7)):
        a, b = qx * lc, qy * lc
        ref = pre * (1.0 + b * j1(b) * k0(a) + a * j0(b) * k1(a))
        got = O.coulomb_v(c, np.array([qx, qy, 0.0]), z, g, rlat)
        assert abs(got - ref) < 1e-12 * abs(ref), (qx, qy)


def test_optics_bz_against_the_dipoles():
    """The per-k activity sums: (E_c - E_v)^2 + eta^2 times |d_a|^2 of
    optics_bz equals |<c|dH/dk_a|v>|^2, which is dipoles' |d_a|^2 divided by
    Re(1/(E_c - E_v + i eta))^2; sp + sm = x + y (the circular pair is unitary)."""
    m = hex_model((3, 3, 1))
    k = O.mp_mesh(3, 3, 1, m.rlat)
    om = O.Model(m, flags=O.GAUGE_LAPACK)
    r = O.hk_eig(om, k, 1, 1)
    eta = 0.07
    act = O.optics_bz(om, k, r["eig"], r["vec"], 1, 1, eta)
    d = O.dipoles(om, k, r["eig"], r["vec"], 1, 1, eta)
    de = r["eig"][:, 1] - r["eig"][:, 0]
    re = de / (de * de + eta * eta)
    for a in range(3):
        raw = np.abs(d[a]) ** 2 / re ** 2
        assert np.allclose(act[:, a] * (de * de + eta * eta), raw, rtol=1e-10)
    norm2 = float(np.float32(1.0) / np.sqrt(np.float32(2.0))) ** 2
    assert np.allclose(act[:, 3] + act[:, 4], 2 * norm2 * (act[:, 0] + act[:, 1]), rtol=1e-12)


# ------------------- truncated and screened potentials --------------------
# Each branch is checked against the Fourier integral of the potential it
# represents, evaluated by quadrature, or against another potential it
# reduces to.  The lattices give |a1| != |a2| != |a3| and a tilted a3, so a
# wrong cell length, cutoff or cell volume fails.

def _vq(c, dk, rlat, ngrid):
    return O.coulomb_v(c, np.asarray(dk, float), np.zeros(3), np.array(ngrid, np.int32), rlat)


def test_v2dt_is_the_slab_truncated_coulomb_integral():
    """v2dt (module_coulomb.F90:135-193) is the 3-D transform of 1/r cut to
    |z| < rc = rlat(3,3)/2 without the 4 pi: (1/g) int_0^rc cos(gz z) e^{-g z} dz
    for in-plane g > 0; for g = 0 the transform's g-independent part
    -int_0^rc z cos(gz z) dz (the 1/g term of a charged sheet removed), which
    is -rc^2/2 at gz = 0."""
    from scipy.integrate import quad
    rlat = np.array([[3.1, 0.0, 0.0], [-1.2, 2.9, 0.0], [0.4, 0.3, 18.0]])
    ngrid = (7, 5, 1)
    pre = O.cic() / (35 * abs(np.linalg.det(rlat)))
    rc = 9.0
    c = O.coulomb("V2DT", ktol=0.001)
    for gx, gy, gz in ((0.3, 0.0, 0.0), (0.05, 0.02, 0.4), (0.7, -0.4, 0.9), (0.01, 0.0, 2.5)):
        g = math.hypot(gx, gy)
        f, _ = quad(lambda z: math.cos(gz * z) * math.exp(-g * z), 0.0, rc, limit=400, epsabs=0, epsrel=1e-12)
        ref = pre * f / g
        got = _vq(c, (gx, gy, gz), rlat, ngrid)
        got_neg = _vq(c, (gx, gy, -gz), rlat, ngrid)      # |gz| enters
        assert abs(got - ref) < 1e-10 * abs(ref), (gx, gy, gz)
        assert got_neg == got
    for gz in (0.05, 0.8, 3.0):
        f, _ = quad(lambda z: z * math.cos(gz * z), 0.0, rc, limit=400, epsabs=0, epsrel=1e-12)
        ref = -pre * f
        assert abs(_vq(c, (0.0, 0.0, gz), rlat, ngrid) - ref) < 1e-10 * abs(ref), gz
    assert abs(_vq(c, (0.0, 0.0, 0.0), rlat, ngrid) - (-pre * rc * rc / 2)) < 1e-13 * abs(pre * rc * rc)


def test_v0dt_is_the_sphere_truncated_coulomb_integral():
    """v0dt (module_coulomb.F90:197-238): the transform of 1/r cut to a sphere
    of radius cr = min |a_i| / 2 without the 4 pi, int_0^cr sin(q r)/q dr,
    and cr^2/2 = int_0^cr r dr for q below KTOL."""
    from scipy.integrate import quad
    rlat = np.array([[4.2, 0.0, 0.0], [0.0, 3.0, 0.5], [0.3, 0.0, 5.0]])
    ngrid = (4, 4, 4)
    pre = O.cic() / (64 * abs(np.linalg.det(rlat)))
    cr = 0.5 * min(np.linalg.norm(rlat, axis=1))
    c = O.coulomb("V0DT", ktol=0.001)
    for dk in ((0.3, 0.0, 0.0), (0.2, -0.5, 0.7), (1.9, 0.4, 0.0)):
        q = np.linalg.norm(dk)
        f, _ = quad(lambda r: math.sin(q * r) / q, 0.0, cr, epsabs=0, epsrel=1e-13)
        assert abs(_vq(c, dk, rlat, ngrid) - pre * f) < 1e-11 * abs(pre * f), dk
    assert abs(_vq(c, (2e-4, 0, 0), rlat, ngrid) - pre * cr * cr / 2) < 1e-13 * abs(pre * cr * cr)


def test_v2dohono_reduces_to_rytova_keldysh_without_polarisability():
    """v2dohono (module_coulomb.F90:354-387) is the Rytova-Keldysh potential
    with r0 = 0 in a uniform medium of permittivity EDIEL_Z at height W_COUL;
    the two functions are independent restatements.  Scaling: 1/EDIEL_Z and
    e^{-q W_COUL}; zero below KTOL."""
    m = T.honeycomb_two_band()
    rlat = m.rlat.copy()
    rlat[1] = [-1.0, 2.4, 0.0]                                   # |a1| != |a2|
    ngrid = (9, 6, 1)
    for ez, w in ((1.0, 0.0), (2.5, 0.0), (3.0, 0.8)):
        oh = O.coulomb("V2DOH", ez=ez, w=w, ediel=(7.0, 5.0, 3.0), lc=4.0, r0=11.0)
        rk = O.coulomb("V2DRK", r0=0.0, ediel=(ez, ez, ez), ez=ez, w=w, lc=4.0)
        for dk in ((0.01, 0.0, 0.0), (0.2, 0.3, 0.0), (1.1, -0.6, 0.0)):
            a, b = _vq(oh, dk, rlat, ngrid), _vq(rk, dk, rlat, ngrid)
            assert abs(a - b) < 1e-13 * abs(b), (ez, w, dk)
    base = _vq(O.coulomb("V2DOH", ez=1.0, w=0.0), (0.4, 0.0, 0.0), rlat, ngrid)
    assert abs(_vq(O.coulomb("V2DOH", ez=4.0, w=0.0), (0.4, 0.0, 0.0), rlat, ngrid) - base / 4) < 1e-14 * abs(base)
    assert abs(_vq(O.coulomb("V2DOH", ez=1.0, w=2.0), (0.4, 0.0, 0.0), rlat, ngrid)
               - base * math.exp(-0.8)) < 1e-14 * abs(base)
    assert _vq(O.coulomb("V2DOH", ez=1.0, w=0.0), (5e-4, 0.0, 0.0), rlat, ngrid) == 0.0


def test_v2dk_keldysh_form_and_small_q_value():
    """v2dk (module_coulomb.F90:8-60).  Above KTOL it is the Keldysh form
    cic / (N A ed q (1 + r0 q)) with ed = (EDIEL_T + EDIEL_B)/2 and
    r0 = (EDIEL - 1) LC / (EDIEL_T + EDIEL_B): the Rytova-Keldysh function in a
    uniform medium ed with polarisability length ed r0 and W_COUL = 0.  Below
    KTOL the regularised value vbz (cic/ed) (1.76 a0 sqrt(n1 n2)/(2 pi) + r0),
    a0 the mean of |a1| and |a2|, 1.76 and pi single precision: its intercept
    at EDIEL = 1 (r0 = 0) and its slope in LC pin the two terms."""
    rlat = np.array([[3.0, 0.0, 0.0], [-1.3, 2.5, 0.0], [0.0, 0.0, 15.0]])
    ngrid = (8, 6, 1)
    area = np.linalg.norm(np.cross(rlat[0], rlat[1]))
    et, eps, eb, lc = 2.0, 5.0, 4.0, 6.5
    ed = 0.5 * (et + eb)
    r0 = (eps - 1.0) * lc / (et + eb)
    k = O.coulomb("V2DK", ediel=(et, eps, eb), lc=lc)
    rk = O.coulomb("V2DRK", r0=ed * r0, ediel=(ed, ed, ed), ez=ed, w=0.0, lc=lc)
    for dk in ((0.002, 0.0, 0.0), (0.3, 0.1, 0.0), (1.4, -0.9, 0.0)):
        a, b = _vq(k, dk, rlat, ngrid), _vq(rk, dk, rlat, ngrid)
        assert abs(a - b) < 1e-13 * abs(b), dk
    vbz = 1.0 / (48 * area)
    a0 = 0.5 * (np.linalg.norm(rlat[0]) + np.linalg.norm(rlat[1]))
    small = (5e-4, 0.0, 0.0)
    v1 = _vq(O.coulomb("V2DK", ediel=(et, 1.0, eb), lc=lc), small, rlat, ngrid)
    ref1 = vbz * (O.cic() / ed) * float(np.float32(1.76)) * a0 * math.sqrt(48.0) / (2 * O.pi_f32())
    assert abs(v1 - ref1) < 1e-13 * abs(ref1)
    va = _vq(O.coulomb("V2DK", ediel=(et, eps, eb), lc=lc), small, rlat, ngrid)
    vb = _vq(O.coulomb("V2DK", ediel=(et, eps, eb), lc=2 * lc), small, rlat, ngrid)
    slope = vbz * (O.cic() / ed) * (eps - 1.0) / (et + eb)
    assert abs((vb - va) / lc - slope) < 1e-11 * abs(slope)
    assert abs((va - v1) - slope * lc) < 1e-11 * abs(slope * lc)


# ------------------- the reference's own TB input --------------------------

def test_reference_mos2_file_read_and_hermitian():
    """examples/example-3/tb_mos2.dat (copied as tests/golden/tb_mos2.dat): the
    reader gives SOC, norb = 22, nvec = 147 with six R of degeneracy 2, and
    H(-R) = H(R)^dagger for every R, so every H(k) is Hermitian; the oracle's
    H(k) (with 1/deg) equals the model's own Fourier sum and the bands leave
    E_F in the gap at every k (nocp = 14)."""
    from golden.make_bse_golden import MOS2_FILE
    m = T.read_tb_file(MOS2_FILE)
    assert m.systype == "SOC" and m.norb == 22 and m.nvec == 147
    assert abs(m.efermi - (-1.2008)) < 1e-12 and m.scissor == 0.0
    assert np.count_nonzero(m.deg == 2) == 6 and np.count_nonzero(m.deg == 1) == 141
    assert tuple(m.rvec[0]) == (-8, -4, 0) and tuple(m.rvec[-1]) == (8, 4, 0)
    where = {tuple(r): i for i, r in enumerate(m.rvec)}
    for i, r in enumerate(m.rvec):
        j = where[tuple(-r)]
        assert np.array_equal(m.hop[j], m.hop[i].conj().T)
        assert m.deg[i] == m.deg[j]
    om = O.Model(m, flags=0)
    k = O.mp_mesh(5, 5, 1, m.rlat)
    for kk in (0, 7, 24):
        h = O.hk(om, k[kk])
        ph = np.exp(1j * (m.rvec @ m.rlat) @ k[kk])
        href = np.einsum("r,rmn->mn", ph / m.deg, m.hop) - m.efermi * np.eye(22)
        assert np.abs(h - href).max() < 1e-12
        assert np.abs(h - h.conj().T).max() < 1e-13
    r = O.hk_eig(O.Model(m), k, 2, 2)
    assert r["status"] == 0 and np.all(r["nocp"] == 14)


def test_sp_tb_file_merges_up_and_down_blocks(tmp_path):
    """An SP file (module_hamiltonian.F90:63-162) holds an up block and a down
    block, each with its own R list; the reader appends the down list to the
    up list (nvec = nvec_up + nvec_down) with the up hoppings in the first
    norb orbitals and the down hoppings in the last, so H(k) is the block
    diagonal of the two spin Hamiltonians.  Written here from two halves of a
    seeded model (the down block with its R list reversed) and read back."""
    base = hex_model((3, 3, 1))
    n = 6
    up = base.hop[:, :n, :n].copy()
    dn = base.hop[::-1, n:2 * n, n:2 * n].copy()
    rv_dn = base.rvec[::-1].copy()
    p = tmp_path / "tb_sp.dat"
    with open(p, "w") as f:
        f.write("SP\n   0.10\n   0.25\n")
        for row in base.rlat:
            f.write(" ".join(f"{x:.7f}" for x in row) + "\n")
        for blk, rv, label in ((up, base.rvec, "up"), (dn, rv_dn, "down")):
            f.write(f" {label} block\n{n}\n{base.nvec}\n")
            f.write(" ".join(str(int(x)) for x in base.deg) + "\n")
            for i in range(base.nvec):
                for b in range(n):
                    for a in range(n):
                        v = blk[i, a, b]
                        r = rv[i]
                        f.write(f"{r[0]} {r[1]} {r[2]} {a + 1} {b + 1} {v.real:.17e} {v.imag:.17e}\n")
    m = T.read_tb_file(str(p))
    nv = base.nvec
    assert m.systype == "SP" and m.norb == 2 * n and m.nvec == 2 * nv
    assert m.scissor == 0.1 and m.efermi == 0.25
    assert np.array_equal(m.rvec[:nv], base.rvec) and np.array_equal(m.rvec[nv:], rv_dn)
    assert np.array_equal(m.hop[:nv, :n, :n], up) and np.array_equal(m.hop[nv:, n:, n:], dn)
    assert np.all(m.hop[:nv, n:, :] == 0) and np.all(m.hop[:nv, :, n:] == 0)
    assert np.all(m.hop[nv:, :n, :] == 0) and np.all(m.hop[nv:, :, :n] == 0)
    # H(k) is block diagonal: the up and down spin Hamiltonians
    k = O.mp_mesh(3, 3, 1, m.rlat)[4]
    h = O.hk(O.Model(m, flags=0), k)
    ph = np.exp(1j * (base.rvec @ m.rlat) @ k)
    hu = np.einsum("r,rmn->mn", ph, up) - 0.25 * np.eye(n)
    hd = np.einsum("r,rmn->mn", ph[::-1], dn) - 0.25 * np.eye(n)
    assert np.abs(h[:n, :n] - hu).max() < 1e-12 and np.abs(h[n:, n:] - hd).max() < 1e-12
    assert np.all(h[:n, n:] == 0)


def test_krylov_schur_against_lapack():
    """The CPU baseline's Krylov-Schur (slepc_solver's settings) finds the
    nev least-|E| pairs of a positive BSE-like Hermitian matrix: LAPACK eigh,
    residuals below the relative tolerance, orthonormal vectors.  (On an
    indefinite matrix the least-|E| pairs are interior and converge slowly
    without a spectral transformation, as in SLEPc; the baseline solves only
    positive exciton spectra.)"""
    rng = np.random.default_rng(11)
    n = 400
    a = 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    h = 0.5 * (a + a.conj().T)
    h[np.diag_indices(n)] = np.sort(rng.uniform(1.0, 8.0, n))
    w = np.linalg.eigvalsh(h)
    e, x, its, nmv, nconv = O.krylov_schur(h, 12, tol=1e-8, max_it=200)
    assert nconv == 12
    assert np.abs(e - w[:12]).max() < 1e-10
    r = np.linalg.norm(h @ x - x * e, axis=0) / np.abs(e)
    assert r.max() < 1e-8
    assert np.allclose(x.conj().T @ x, np.eye(12), atol=1e-10)
