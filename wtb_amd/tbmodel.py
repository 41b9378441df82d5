"""Tight-binding model container, reference-format TB file I/O and the seeded
synthetic model generators used for the exciton workloads.

The TB file layout is the one read by ``hamiltonian_input_read``
(src/module_hamiltonian.F90:35-194):

    line 1   systype (NP | SOC | SP)
    line 2   scissor shift (eV)
    line 3   Fermi energy (eV)
    line 4-6 lattice vectors a1, a2, a3 as rows (Angstrom)
    then a wannier90 ``_hr.dat`` body: comment, norb, nvec, nvec degeneracies,
    and nvec*norb*norb lines ``R1 R2 R3 m n Re Im``.  For every R the column
    index n runs outer and the row index m inner, and the two printed orbital
    indices are ignored by the reader (module_hamiltonian.F90:166-172).

``hop[i, m, n]`` is H_mn(R_i) in eV, the reference's
``hopmatrices(i,m,n) + i*ihopmatrices(i,m,n)``.  ``rlat[i, :]`` is lattice
vector a_{i+1}, the reference's ``rlat(i,:)``.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

__all__ = [
    "TBModel", "read_tb_file", "write_tb_file", "mp_mesh", "mp_mesh_q", "recip_vectors",
    "temperature_gap_shift",
    "honeycomb_two_band", "random_hexagonal", "random_cubic", "workload",
    "WORKLOADS",
]

# float32 value of acos(-1.) as used by every module of the reference.
PI_F32 = float(np.float32(math.pi))


@dataclasses.dataclass
class TBModel:
    systype: str
    scissor: float
    efermi: float
    rlat: np.ndarray          # (3, 3) float64, rows are a1, a2, a3
    rvec: np.ndarray          # (nvec, 3) int32, lattice coordinates of R
    deg: np.ndarray           # (nvec,) int32, Wigner-Seitz degeneracy
    hop: np.ndarray           # (nvec, norb, norb) complex128, H_mn(R)
    comment: str = "synthetic model"
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def norb(self) -> int:
        return int(self.hop.shape[1])

    @property
    def nvec(self) -> int:
        return int(self.hop.shape[0])

    def copy(self) -> "TBModel":
        return TBModel(self.systype, self.scissor, self.efermi,
                       self.rlat.copy(), self.rvec.copy(), self.deg.copy(),
                       self.hop.copy(), self.comment, dict(self.meta))

    def hk(self, kpts: np.ndarray) -> np.ndarray:
        """H(k) - E_F for Cartesian k (nk, 3), plain FP64 (no float32 phase)."""
        rcart = self.rvec.astype(np.float64) @ self.rlat
        ph = np.exp(1j * (np.asarray(kpts, np.float64) @ rcart.T))
        h = np.einsum("kr,rmn->kmn", ph / self.deg[None, :], self.hop)
        h -= self.efermi * np.eye(self.norb)[None]
        return h

    def hermiticity_error(self) -> float:
        """max |H(R) - H(-R)^dagger| over all R whose partner -R is present."""
        index = {tuple(r): i for i, r in enumerate(self.rvec.tolist())}
        err = 0.0
        for i, r in enumerate(self.rvec.tolist()):
            j = index.get((-r[0], -r[1], -r[2]))
            if j is None:
                continue
            err = max(err, float(np.max(np.abs(self.hop[i] - self.hop[j].conj().T))))
        return err


def read_tb_file(path: str) -> TBModel:
    """Parse a reference-format TB file (NP/SOC single block; SP two blocks
    merged block-diagonally as in module_hamiltonian.F90:63-162)."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    pos = 0

    def next_tokens():
        nonlocal pos
        while pos < len(lines) and not lines[pos].strip():
            pos += 1
        toks = lines[pos].split("!")[0].split()
        pos += 1
        return toks

    systype = next_tokens()[0].strip("'\"")
    scissor = float(next_tokens()[0])
    efermi = float(next_tokens()[0])
    rlat = np.array([[float(x) for x in next_tokens()[:3]] for _ in range(3)])

    def read_block():
        nonlocal pos
        comment = lines[pos].strip()
        pos += 1
        norb = int(next_tokens()[0])
        nvec = int(next_tokens()[0])
        deg = []
        while len(deg) < nvec:
            deg.extend(int(x) for x in next_tokens())
        body = np.loadtxt(lines[pos:pos + nvec * norb * norb], dtype=np.float64, ndmin=2)
        pos += nvec * norb * norb
        body = body.reshape(nvec, norb, norb, 7)        # [R][n (col)][m (row)]
        rvec = body[:, 0, 0, :3].astype(np.int32)
        hop = (body[..., 5] + 1j * body[..., 6]).transpose(0, 2, 1)  # -> [R][m][n]
        return comment, np.asarray(deg, np.int32), rvec, np.ascontiguousarray(hop)

    if systype == "SP":
        cu, du, ru, hu = read_block()
        cd, dd, rd, hd = read_block()
        nu, nd = hu.shape[1], hd.shape[1]
        hop = np.zeros((hu.shape[0] + hd.shape[0], nu + nd, nu + nd), np.complex128)
        hop[:hu.shape[0], :nu, :nu] = hu
        hop[hu.shape[0]:, nu:, nu:] = hd
        return TBModel(systype, scissor, efermi, rlat, np.vstack([ru, rd]),
                       np.concatenate([du, dd]), hop, cu)
    comment, deg, rvec, hop = read_block()
    return TBModel(systype, scissor, efermi, rlat, rvec, deg, hop, comment)


def write_tb_file(model: TBModel, path: str) -> None:
    """Write ``model`` in the reference TB format (NP/SOC single block)."""
    out = [f" {model.systype}", f"   {model.scissor!r}                  !scissors operator",
           f"   {model.efermi!r}                  !efermi"]
    for a in model.rlat:
        out.append("  " + "  ".join(f"{x:.16e}" for x in a))
    out.append(f" {model.comment}")
    out.append(f"  {model.norb}")
    out.append(f"  {model.nvec}")
    for s in range(0, model.nvec, 15):
        out.append("".join(f"{d:5d}" for d in model.deg[s:s + 15]))
    norb = model.norb
    for i in range(model.nvec):
        r = model.rvec[i]
        for n in range(norb):
            for m in range(norb):
                h = model.hop[i, m, n]
                out.append(f"{r[0]:5d}{r[1]:5d}{r[2]:5d}{m + 1:5d}{n + 1:5d} "
                           f"{h.real:.17e} {h.imag:.17e}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def recip_vectors(rlat: np.ndarray, pi: float = PI_F32) -> np.ndarray:
    """Reciprocal vectors b_i (rows), b_i . a_j = 2 pi delta_ij, with the
    reference's float32 pi (src/module_kmesh.F90:391-407)."""
    a1, a2, a3 = np.asarray(rlat, np.float64)
    v23, v31, v12 = np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)
    vol = abs(float(a1 @ v23))
    return np.array([v23, v31, v12]) * ((2.0 * pi) / vol)


def mp_mesh(n1: int, n2: int, n3: int, rlat: np.ndarray,
            shift=(0.0, 0.0, 0.0), pi: float = PI_F32) -> np.ndarray:
    """Gamma-centred Monkhorst-Pack mesh in Cartesian 1/Angstrom, loop order
    i (outer), j, l (inner), unfolded (src/module_kmesh.F90:338-386)."""
    b = recip_vectors(rlat, pi)
    kshift = b.T @ np.asarray(shift, np.float64)
    i, j, l = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    k = (np.outer(i.ravel(), b[0] / n1) + np.outer(j.ravel(), b[1] / n2)
         + np.outer(l.ravel(), b[2] / n3))
    return k + kshift[None, :]


def mp_mesh_q(n1: int, n2: int, n3: int, rlat: np.ndarray, q,
              shift=(0.0, 0.0, 0.0), pi: float = PI_F32) -> np.ndarray:
    """The k+Q mesh of the finite-Q exciton, q the Cartesian Q: the points of
    mp_mesh in the same order, summed as q + i b1/n1 + j b2/n2 + l b3/n3 +
    shift (src/module_kmesh.F90:281-334)."""
    b = recip_vectors(rlat, pi)
    kshift = b.T @ np.asarray(shift, np.float64)
    i, j, l = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    k = np.asarray(q, np.float64)[None, :] + np.outer(i.ravel(), b[0] / n1)
    k = k + np.outer(j.ravel(), b[1] / n2)
    k = k + np.outer(l.ravel(), b[2] / n3)
    return k + kshift[None, :]


KB_F32 = float(np.float32(8.617330350e-5))   # eV/K, a default-real literal in the reference


def temperature_gap_shift(ta: str, st: float, phavg: float, temp: float) -> float:
    """Gap shift that the temperature BSE adds to the scissor (bsesolvertemp,
    src/module_bse.F90:2857-2875), in eV.

    TA "VE": -S <hw> (coth(<hw> / 2 kB T) - 1) (gapcortemp,
    src/module_math.F90:383-391); TA "BE": -2 S / (exp(<hw> / T) - 1)
    (gapcortemp2, :395-404, with T in the exponent as written there); any
    other TA ("FA"): 0."""
    if ta == "VE":
        return -st * phavg * (1.0 / math.tanh(phavg / (2.0 * KB_F32 * temp)) - 1.0)
    if ta == "BE":
        return -(2.0 * st / (math.exp(phavg / temp) - 1.0))
    return 0.0


# --------------------------------------------------------------------------
# synthetic generators
# --------------------------------------------------------------------------

def honeycomb_two_band(a: float = 2.5, delta: float = 4.0, t: float = -2.8,
                       c: float = 20.0) -> TBModel:
    """Two-orbital gapped honeycomb model: onsite +delta/2 (A), -delta/2 (B),
    nearest-neighbour A-B hopping t on three bonds, E_F = 0, no spin.

    H_AB(k) = t (1 + e^{-i k.a1} + e^{-i k.a2}), so
    E(k) = +- sqrt(delta^2/4 + t^2 |1 + e^{-i k.a1} + e^{-i k.a2}|^2).
    """
    rlat = np.array([[a, 0.0, 0.0], [-0.5 * a, 0.5 * math.sqrt(3.0) * a, 0.0], [0.0, 0.0, c]])
    rvec = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [1, 0, 0], [0, 1, 0]], np.int32)
    hop = np.zeros((5, 2, 2), np.complex128)
    hop[0] = [[delta / 2, t], [t, -delta / 2]]
    hop[1, 0, 1] = t          # A <- B in cell -a1
    hop[2, 0, 1] = t          # A <- B in cell -a2
    hop[3, 1, 0] = t          # partners H(-R) = H(R)^dagger
    hop[4, 1, 0] = t
    return TBModel("NP", 0.0, 0.0, rlat, rvec, np.ones(5, np.int32), hop,
                   "two-orbital gapped honeycomb",
                   {"kind": "honeycomb", "a": a, "delta": delta, "t": t, "c": c})


def _shell_vectors(rlat: np.ndarray, dims: int, rmax: float, a: float):
    """Integer lattice vectors with |R| <= rmax*a, ordered by (|R|, n1, n2, n3)."""
    n = int(math.ceil(rmax)) + 2
    rng3 = range(-n, n + 1) if dims == 3 else range(0, 1)
    out = []
    for n1 in range(-n, n + 1):
        for n2 in range(-n, n + 1):
            for n3 in rng3:
                r = n1 * rlat[0] + n2 * rlat[1] + n3 * rlat[2]
                d = float(np.linalg.norm(r)) / a
                if d <= rmax + 1e-9:
                    out.append((round(d, 9), n1, n2, n3))
    out.sort()
    return np.array([o[1:] for o in out], np.int32), np.array([o[0] for o in out])


def _random_hermitian_set(rng, rvec, dist, norb, amp, decay):
    index = {tuple(r): i for i, r in enumerate(rvec.tolist())}
    hop = np.zeros((len(rvec), norb, norb), np.complex128)
    done = set()
    for i, r in enumerate(rvec.tolist()):
        if i in done:
            continue
        scale = amp * math.exp(-dist[i] / decay)
        g = scale * (rng.standard_normal((norb, norb))
                     + 1j * rng.standard_normal((norb, norb))) / math.sqrt(2.0)
        j = index[(-r[0], -r[1], -r[2])]
        if j == i:
            hop[i] = 0.5 * (g + g.conj().T)
        else:
            hop[i] = g
            hop[j] = g.conj().T
        done.update((i, j))
    return hop


def _band_edges(model: TBModel, mesh, nocc: int):
    k = mp_mesh(*mesh, model.rlat)
    e = np.linalg.eigvalsh(model.hk(k) + model.efermi * np.eye(model.norb)[None])
    return float(e[:, nocc - 1].max()), float(e[:, nocc].min()), float((e[:, nocc] - e[:, nocc - 1]).min())


def _split_for_gap(base: TBModel, nocc: int, gap: float, mesh) -> TBModel:
    """Add onsite -s (first nocc orbitals) / +s (rest) and bisect s until the
    global gap between band nocc and nocc+1 on ``mesh`` equals ``gap``;
    E_F is put mid-gap so that nocp == nocc at every k of the mesh."""
    i0 = int(np.flatnonzero((base.rvec == 0).all(axis=1))[0])
    split = np.concatenate([-np.ones(nocc), np.ones(base.norb - nocc)])

    def with_split(s):
        m = base.copy()
        m.hop[i0] += np.diag(s * split) * base.deg[i0]
        return m

    def global_gap(s):
        vmax, cmin, _ = _band_edges(with_split(s), mesh, nocc)
        return cmin - vmax

    lo, hi = 0.0, 1.0
    while global_gap(hi) < gap:
        hi *= 2.0
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if global_gap(mid) < gap:
            lo = mid
        else:
            hi = mid
    m = with_split(hi)
    vmax, cmin, dgap = _band_edges(m, mesh, nocc)
    m.efermi = 0.5 * (vmax + cmin)
    m.meta.update({"split": hi, "global_gap": cmin - vmax, "direct_gap": dgap,
                   "gap_mesh": list(mesh)})
    return m


def random_hexagonal(seed: int = 2026, norb: int = 22, a: float = 3.2,
                     shells: float = 3.0, amp: float = 1.5, decay: float = 1.5,
                     gap: float = 1.8, c: float = 20.0,
                     gap_mesh=(48, 48, 1)) -> TBModel:
    """Seeded random 2D hexagonal TB model (spinful, 'SOC' file type).

    H(R) for |R| <= shells*a (37 vectors for shells=3): complex Gaussian
    entries of amplitude amp*exp(-|R|/(a*decay)) with |R| in units of a,
    H(-R) = H(R)^dagger, H(0) Hermitian.  An onsite split opens a global gap
    ``gap`` between bands norb/2 and norb/2+1 on ``gap_mesh``.
    """
    rlat = np.array([[a, 0.0, 0.0], [-0.5 * a, 0.5 * math.sqrt(3.0) * a, 0.0], [0.0, 0.0, c]])
    rvec, dist = _shell_vectors(rlat, 2, shells, a)
    rng = np.random.default_rng(seed)
    hop = _random_hermitian_set(rng, rvec, dist, norb, amp, decay)
    base = TBModel("SOC", 0.0, 0.0, rlat, rvec, np.ones(len(rvec), np.int32), hop,
                   f"random hexagonal seed={seed}",
                   {"kind": "random_hexagonal", "seed": seed, "norb": norb, "a": a,
                    "shells": shells, "amp": amp, "decay": decay, "c": c})
    return _split_for_gap(base, norb // 2, gap, gap_mesh)


def random_cubic(seed: int = 2027, norb: int = 16, a: float = 4.0,
                 shells: float = 2.0, amp: float = 1.5, decay: float = 1.2,
                 gap: float = 1.2, gap_mesh=(16, 16, 16)) -> TBModel:
    """Seeded random 3D simple-cubic TB model (33 R vectors for shells=2)."""
    rlat = np.eye(3) * a
    rvec, dist = _shell_vectors(rlat, 3, shells, a)
    rng = np.random.default_rng(seed)
    hop = _random_hermitian_set(rng, rvec, dist, norb, amp, decay)
    base = TBModel("SOC", 0.0, 0.0, rlat, rvec, np.ones(len(rvec), np.int32), hop,
                   f"random cubic seed={seed}",
                   {"kind": "random_cubic", "seed": seed, "norb": norb, "a": a,
                    "shells": shells, "amp": amp, "decay": decay})
    return _split_for_gap(base, norb // 2, gap, gap_mesh)


# Workload table: the reference's exciton cases on synthetic models.
WORKLOADS = {
    "honeycomb-24": dict(model="honeycomb", mesh=(24, 24, 1), nv=1, nc=1,
                         coulomb=dict(kind="V2DRK", r0=10.0, ediel=(1.0, 1.0, 1.0), ez=1.0,
                                      w=0.0, lc=1.0, ktol=0.001),
                         nev=576, cshift=0.05, spectrum=(3.0, 7.0, 1000)),
    "hex22-60": dict(model="hex22", mesh=(60, 60, 1), nv=2, nc=2,
                     coulomb=dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0,
                                  w=0.0, lc=1.0, ktol=0.001),
                     nev=64, cshift=0.05, spectrum=(1.0, 5.0, 1000)),
    "hex22-90": dict(model="hex22", mesh=(90, 90, 1), nv=2, nc=2,
                     coulomb=dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0,
                                  w=0.0, lc=1.0, ktol=0.001),
                     nev=64, cshift=0.05, spectrum=(1.0, 5.0, 1000)),
    "cubic16-16": dict(model="cubic16", mesh=(16, 16, 16), nv=3, nc=3,
                       coulomb=dict(kind="V3DL", r0=1.0, ediel=(1.0, 6.0, 1.0), ez=1.0,
                                    w=0.0, lc=1.0, ktol=0.001),
                       nev=128, cshift=0.05, spectrum=(0.5, 6.0, 2000)),
    "hex22-48": dict(model="hex22", mesh=(48, 48, 1), nv=2, nc=2,
                     coulomb=dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0,
                                  w=0.0, lc=1.0, ktol=0.001),
                     nev=8, cshift=0.05, spectrum=(1.0, 5.0, 1000)),
}

_MODEL_CACHE: dict = {}


def workload(name: str, seed: Optional[int] = None):
    """(model, spec) for a named workload; models are cached per (kind, seed)."""
    spec = dict(WORKLOADS[name])
    kind = spec["model"]
    key = (kind, seed)
    if key not in _MODEL_CACHE:
        if kind == "honeycomb":
            _MODEL_CACHE[key] = honeycomb_two_band()
        elif kind == "hex22":
            _MODEL_CACHE[key] = random_hexagonal(seed=2026 if seed is None else seed)
        elif kind == "cubic16":
            _MODEL_CACHE[key] = random_cubic(seed=2027 if seed is None else seed)
        else:
            raise KeyError(kind)
    return _MODEL_CACHE[key].copy(), spec
