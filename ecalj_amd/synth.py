"""Seeded synthetic inputs for the QSGW chi0 -> W -> Sigma_c path.

ecalj obtains these inputs from its DFT/GW front end (eigenvectors cphi,
radial integrals ppb, Coulomb matrix v(q), tetrahedron weights).  Here they
are drawn from fixed distributions so that every run, on any machine, sees
the same numbers:

* energies: occupied U[-12, 0] eV, unoccupied U[0.5, 60] eV, sorted, E_F = 0;
* eigenvectors: complex Gaussian nb x nbt, QR-orthonormalised;
* ppb: real Gaussian, shape nb x nb x nI (real, as S/m_zmel.f90:38), with
  variance 1/(nI*Omega), Omega = alat^3/4 the fcc cell volume.  This makes
  sum_I |zmel(I,n,m)|^2 ~ 1/Omega, the value for cell-normalised product
  basis and Bloch functions; the literal variance 1/nb gives nI/nb instead
  (600x larger at nI=600), which puts epsilon at ~1e4 and Sigma_c at ~1e4 eV;
* v(q) = Q diag(4 pi/|q+G|^2) Q^H with Q unitary (QR of a complex Gaussian),
  G the first nI integer-lattice vectors sorted by shell, q = (0.1,0.2,0.3) 2pi/a;
* histogram weights: each (occ n at k, unocc m at k+q) pair spreads pi*w_k
  uniformly over the window [0.98 D, 1.02 D], D = e_m - e_n in Hartree,
  split over the histogram bins it overlaps.  This stands in for the
  tetrahedron weights whw (S/tetwt5.f90:335, which carry the same pi/vol
  normalisation); the bins are ecalj's own mesh (S/m_freq.f90:95-108).

Draw order is fixed: energies -> eigenvectors -> ppb -> v(q), from
numpy PCG64(seed).  Arrays are Fortran-ordered complex128 / float64.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

RYDBERG_EV = 13.605693122994
HARTREE_EV = 2.0 * RYDBERG_EV


@dataclass(frozen=True)
class Config:
    name: str
    nk: int          # internal k-points (nqbz)
    nb: int          # eigenvector basis size (ndima)
    no: int          # occupied bands
    nu: int          # unoccupied bands
    ni: int          # mixed product basis size (ngb)
    nf: int          # imaginary frequencies (niw), Gauss-Legendre supported sizes only
    seed: int
    alat: float = 10.68
    ntq: int = 0     # bands whose Sigma is computed; 0 -> 2*no
    esmr: float = 0.003   # Ry
    ua: float = 1.0
    nspin: int = 1
    hist_ratio: float = 1.03
    hist_dw: float = 1e-5
    window: float = 0.02
    natom: int = 1   # atoms of the MT product basis: ppb in natom blocks of nb/natom x nb/natom x ni/natom
    nu_chi0: int = 0  # unoccupied bands in the chi0 band sum (0 = all); nband_chi0, S/m_readgwinput.f90:34-35
    # ppb variance: "cell" natom^2/(ni*Omega) (cell-normalised product basis, the
    # default), or "spec" 1/nb (the workloads' literal N(0, 1/nb) statement)
    ppb_var: str = "cell"
    note: str = ""

    @property
    def nbt(self) -> int:
        return self.no + self.nu

    @property
    def nq(self) -> int:
        return self.ntq if self.ntq else 2 * self.no

    @property
    def nv_list(self):
        return [self.nb // self.natom] * self.natom

    @property
    def mdim_list(self):
        return [self.ni // self.natom] * self.natom

    def scaled(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


# The workloads named by the metric, plus small shapes for parity tests.
CONFIGS = {
    # Si-like, CPU-sized
    "C1": Config("C1", nk=8, nb=100, no=4, nu=60, ni=200, nf=12, seed=20261203, alat=10.26),
    # GaAs-like, FP64 GPU baseline form
    "C2": Config("C2", nk=64, nb=180, no=9, nu=150, ni=600, nf=16, seed=20261204),
    # Mixed precision; nb raised from 300 to 320 so that QR can give nbt = 318
    # orthonormal eigenvectors (300 < 318 is inconsistent).
    "C3": Config("C3", nk=64, nb=320, no=18, nu=300, ni=1200, nf=20, seed=20261205,
                 note="nb=320 (spec says 300 < nbt=318)"),
    # Learned chi0 tail: chi0 truncated at 200 of 800 unoccupied bands; 8 atoms
    # of 104 eigenvector rows and 150 product-basis functions each (nb = 832 >= nbt)
    "C5": Config("C5", nk=8, nb=832, no=8, nu=800, ni=1200, nf=20, seed=20261207, natom=8, nu_chi0=200,
                 note="nk, no and the atom blocks are not fixed by the spec; chosen here"),
    # Batched uneven workload, one GPU's share: 27 of the 216 k-points (6x6x6),
    # per-k windows of U{8..24} occupied and U{150..400} unoccupied bands, nI per
    # q-point from U[800, 1600] (c4_qpoints); nb = 8 atoms x 54 >= 424 bands
    "C4": Config("C4", nk=27, nb=432, no=24, nu=400, ni=1200, nf=20, seed=20261206, natom=8, ntq=16,
                 note="per-GPU share (27 k); nI drawn per q-point"),
    # parity-test shapes
    "T1": Config("T1", nk=2, nb=12, no=2, nu=5, ni=16, nf=6, seed=11),
    "T2": Config("T2", nk=3, nb=24, no=3, nu=13, ni=40, nf=10, seed=12),
    "T3": Config("T3", nk=4, nb=40, no=4, nu=30, ni=72, nf=12, seed=13),
}


def _qr_unitary(rng, n, m):
    a = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    q, r = np.linalg.qr(a)
    # fix the phase so the draw is deterministic across LAPACK builds
    d = np.diag(r)
    q = q * (d / np.abs(d))[None, :].conj()
    return np.asfortranarray(q)


def _energies(rng, no, nu):
    eo = np.sort(rng.uniform(-12.0, 0.0, no))
    eu = np.sort(rng.uniform(0.5, 60.0, nu))
    return np.concatenate([eo, eu])


def gvectors(ni: int) -> np.ndarray:
    """First ni integer vectors of a cubic lattice, sorted by (|n|^2, n)."""
    r = 1
    while (2 * r + 1) ** 3 < 2 * ni + 64:
        r += 1
    rng = np.arange(-r - 1, r + 2)
    g = np.array(np.meshgrid(rng, rng, rng, indexing="ij")).reshape(3, -1).T
    key = (g ** 2).sum(1)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], key))
    return g[order[:ni]]


@dataclass
class QPoint:
    """All inputs of one q-point of the chi0 -> W -> Sigma_c path."""
    cfg: Config
    # chi0 side: bands at k (occupied used) and k+q (unoccupied used)
    e_k: np.ndarray        # nbt x nk  (eV)
    e_kq: np.ndarray       # nbt x nk
    c_k: np.ndarray        # nb x nbt x nk
    c_kq: np.ndarray       # nb x nbt x nk
    # Sigma side: end states at q, middle states at q-k
    e_q: np.ndarray        # nbt
    e_qmk: np.ndarray      # nbt x nk
    c_q: np.ndarray        # nb x nbt
    c_qmk: np.ndarray      # nb x nbt x nk
    ppb: np.ndarray        # nb x nb x ni (real); natom > 1: the atom blocks nv x nv x mdim, concatenated
    vcoud: np.ndarray      # ni
    zcousq: np.ndarray     # ni x ni
    wk: np.ndarray         # nk BZ weights
    extra: dict = field(default_factory=dict)

    @property
    def vcousq(self):
        return np.sqrt(self.vcoud)


def make_qpoint(cfg: Config) -> QPoint:
    rng = np.random.Generator(np.random.PCG64(cfg.seed))
    nk, nb, nbt, ni = cfg.nk, cfg.nb, cfg.nbt, cfg.ni
    if nb < nbt:
        raise ValueError(f"{cfg.name}: nb={nb} < nbt={nbt}; QR cannot give nbt orthonormal vectors")
    if nb % cfg.natom or ni % cfg.natom:
        raise ValueError(f"{cfg.name}: nb and ni must split evenly over {cfg.natom} atoms")
    # energies
    e_k = np.stack([_energies(rng, cfg.no, cfg.nu) for _ in range(nk)], axis=1)
    e_kq = np.stack([_energies(rng, cfg.no, cfg.nu) for _ in range(nk)], axis=1)
    e_q = _energies(rng, cfg.no, cfg.nu)
    e_qmk = np.stack([_energies(rng, cfg.no, cfg.nu) for _ in range(nk)], axis=1)
    # eigenvectors
    c_k = np.stack([_qr_unitary(rng, nb, nbt) for _ in range(nk)], axis=2)
    c_kq = np.stack([_qr_unitary(rng, nb, nbt) for _ in range(nk)], axis=2)
    c_q = _qr_unitary(rng, nb, nbt)
    c_qmk = np.stack([_qr_unitary(rng, nb, nbt) for _ in range(nk)], axis=2)
    # ppb: real Gaussian with variance natom^2/(ni*Omega): an eigenvector keeps
    # 1/natom of its norm on each atom, so sum_I |zmel|^2 ~ 1/Omega for any natom
    omega_cell = cfg.alat ** 3 / 4.0
    if cfg.ppb_var not in ("cell", "spec"):
        raise ValueError(f"ppb_var {cfg.ppb_var!r}: 'cell' or 'spec'")
    if cfg.natom == 1:
        sd = 1.0 / math.sqrt(nb) if cfg.ppb_var == "spec" else 1.0 / math.sqrt(ni * omega_cell)
        ppb = np.asfortranarray(rng.standard_normal((nb, nb, ni)) * sd)
    else:
        nv, md = nb // cfg.natom, ni // cfg.natom
        sd = 1.0 / math.sqrt(nb) if cfg.ppb_var == "spec" else cfg.natom / math.sqrt(ni * omega_cell)
        ppb = rng.standard_normal(cfg.natom * nv * nv * md) * sd
    # v(q)
    q = np.array([0.1, 0.2, 0.3])
    g = gvectors(ni)
    tpioa = 2.0 * math.pi / cfg.alat
    qg2 = (((q[None, :] + g) * tpioa) ** 2).sum(1)
    vcoud = 4.0 * math.pi / qg2
    zcousq = _qr_unitary(rng, ni, ni)
    wk = np.full(nk, 1.0 / nk)
    return QPoint(cfg, np.asfortranarray(e_k), np.asfortranarray(e_kq), np.asfortranarray(c_k),
                  np.asfortranarray(c_kq), e_q, np.asfortranarray(e_qmk), c_q, np.asfortranarray(c_qmk),
                  ppb, vcoud, zcousq, wk)


def redraw_high_bands(qp: QPoint, nu_t: int, seed: int) -> QPoint:
    """The same q-point with the unoccupied k+q eigenvectors above band
    no + nu_t drawn again (Haar-random in the orthogonal complement of the
    kept ones); energies, the kept bands and everything else unchanged.  The
    truncated chi0 (bands up to nu_t) is then identical, and the spread of
    results from the full band sum over such draws is what no function of
    the truncated sum can predict."""
    rng = np.random.Generator(np.random.PCG64(seed))
    no, nk = qp.cfg.no, qp.cfg.nk
    keep = no + nu_t
    c = np.array(qp.c_kq, copy=True)
    nb, nbt = c.shape[0], c.shape[1]
    if not 0 <= nu_t <= nbt - no:
        raise ValueError(f"nu_t={nu_t} outside 0..{nbt - no}")
    for k in range(nk):
        low = c[:, :keep, k]
        x = rng.standard_normal((nb, nbt - keep)) + 1j * rng.standard_normal((nb, nbt - keep))
        for _ in range(2):  # project out the kept bands (twice, for orthogonality to rounding)
            x -= low @ (low.conj().T @ x)
        q, r = np.linalg.qr(x)
        d = np.diag(r)
        c[:, keep:, k] = q * (d / np.abs(d))[None, :]
    return dataclasses.replace(qp, c_kq=np.asfortranarray(c))


# --------------------------------------------------------------------------
def omg2max_hartree(qp: QPoint) -> float:
    """omg2max of getfreq2 (S/m_freq.f90:63-64) from the band range, in Hartree."""
    allE = np.concatenate([qp.e_k.ravel(), qp.e_kq.ravel()]) / RYDBERG_EV
    return (allE.max() - allE.min()) * 0.5 + 0.2


def frhis(qp: QPoint) -> np.ndarray:
    """Histogram bin edges frhis(1:nwhis+1) in Hartree (S/m_freq.f90:95-108)."""
    cfg = qp.cfg
    aa = cfg.hist_ratio - 1.0
    bb = cfg.hist_dw / aa
    omg = omg2max_hartree(qp)
    iw = 0
    while True:
        iw += 1
        if bb * (math.exp(aa * (iw - 1)) - 1.0) > omg + 1e-6:
            break
    nwhis = iw + 2
    return np.array([bb * (math.exp(aa * (i - 1)) - 1.0) for i in range(1, nwhis + 2)])


# --------------------------------------------------------------------------
# interstitial plane-wave (IPW) part of zmel
@dataclass
class IPWCase:
    """Inputs of zmel's IPW block (S/m_zmel.f90:418-538): G-vector lists of
    the IPW product basis (ngveccR), of the plane-wave basis at q (ngvecpB1)
    and at q-k (ngvecpB2); the m_read_ppovl tables of q: ggg (Fourier
    coefficients of the interstitial step function) with its index cube
    igggi over [lo, lo+n), the G list nvgcgp2 (every Gc+G2 and Gc-G1 needed)
    with its index cube igcgp2i; PW coefficients geigq (ngpmx x ntp0, end
    states at q) and dgeigqk (ngpmx x nm, middle states at q-k); nadd; the
    phase exp(-i 2pi (R k + G).shtv) per IPW function."""
    ngveccR: np.ndarray
    ngvecpB1: np.ndarray
    ngvecpB2: np.ndarray
    nadd: np.ndarray
    ggg: np.ndarray
    igggi: np.ndarray
    ggg_lo: np.ndarray
    nvgcgp2: np.ndarray
    igcgp2i: np.ndarray
    gcp_lo: np.ndarray
    geigq: np.ndarray
    dgeigqk: np.ndarray
    phase: np.ndarray


def _cube(vecs, lo, hi):
    n = hi - lo + 1
    cube = np.zeros(tuple(n), np.int32, order="F")
    for i, v in enumerate(vecs):
        cube[tuple(v - lo)] = i + 1
    return cube


def ipw_case(seed=1, ngc=40, ngp1=30, ngp2=35, nm=7, ntp0=5, ngpmx=None, nadd=(0, 1, -1), gcut2=12,
             absent=0.1) -> IPWCase:
    """Seeded IPW inputs: integer G vectors sorted by shell, a step-function
    table with Hermitian symmetry ggg(-G) = conj(ggg(G)) inside |G|^2 <=
    gcut2 (entries beyond it, and a fraction `absent` inside, have index 0:
    no contribution), random PW coefficients."""
    rng = np.random.default_rng(seed)
    G = gvectors(max(ngc, ngp1, ngp2) + 8)
    gc = G[:ngc].copy()
    g1 = G[rng.permutation(ngp1)].copy()
    g2 = G[rng.permutation(ngp2)].copy()
    nadd = np.asarray(nadd, np.int32)
    combos = np.unique(np.concatenate([(gc[:, None, :] + g2[None, :, :]).reshape(-1, 3),
                                       (gc[:, None, :] - g1[None, :, :]).reshape(-1, 3)]), axis=0)
    lo2, hi2 = combos.min(0), combos.max(0)
    igcgp2i = _cube(combos, lo2, hi2)
    # every G1 - X - nadd and -G2 - X - nadd the two branches look up
    need = np.concatenate([(g1[None, :, :] - combos[:, None, :] - nadd).reshape(-1, 3),
                           (-g2[None, :, :] - combos[:, None, :] - nadd).reshape(-1, 3)])
    lo, hi = need.min(0) - 1, need.max(0) + 1
    r = int(np.abs(np.concatenate([lo, hi])).max())
    ax = np.arange(-r, r + 1)
    allg = np.array(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, -1).T
    inside = allg[(allg ** 2).sum(1) <= gcut2]
    half = [v for v in inside if tuple(v) >= tuple(-v)]      # one of each +-G pair
    vals = {}
    for v in half:
        if rng.random() < absent and np.any(v != 0):
            continue
        z = complex(rng.normal(), rng.normal()) / (1.0 + float((v ** 2).sum()))
        if np.all(v == 0):
            z = complex(0.7, 0.0)
        vals[tuple(v)] = z
        vals[tuple(-v)] = np.conj(z)
    keys = sorted(vals)
    ggg = np.array([vals[k] for k in keys], np.complex128)
    kv = np.array(keys, np.int32)
    cube = np.zeros(tuple(hi - lo + 1), np.int32, order="F")
    for i, v in enumerate(kv):
        if np.all(v >= lo) and np.all(v <= hi):
            cube[tuple(v - lo)] = i + 1
    ngpmx = ngpmx or max(ngp1, ngp2) + 3
    crand = lambda *sh: np.asfortranarray(rng.normal(size=sh) + 1j * rng.normal(size=sh))
    return IPWCase(np.asfortranarray(gc.T, np.int32), np.asfortranarray(g1.T, np.int32),
                   np.asfortranarray(g2.T, np.int32), nadd, ggg, cube, lo.astype(np.int32),
                   np.asfortranarray(combos.T, np.int32), igcgp2i, lo2.astype(np.int32),
                   crand(ngpmx, ntp0), crand(ngpmx, nm), np.exp(1j * rng.uniform(0, 2 * np.pi, ngc)))


# --------------------------------------------------------------------------
# k-mesh, tetrahedra and smooth bands for the tetrahedron weights
@dataclass
class TetraMesh:
    """The BZ-mesh arrays that tetwt5x_dtet4 takes (getbzdata1's output,
    S/getbzdata1.f90:247-293): qbzw the (n1+1)(n2+1)(n3+1) extended mesh
    (3 x nqbzw, 2pi/alat units), ib1bz its 1-based k index in the n1 n2 n3
    mesh, idtetf 4 x ntetf 1-based qbzw indices of the tetrahedra (6 per
    micro-cell), qlat the reciprocal vectors (columns)."""
    n: tuple
    qlat: np.ndarray
    qbzw: np.ndarray
    ib1bz: np.ndarray
    idtetf: np.ndarray

    @property
    def nqbz(self):
        return int(np.prod(self.n))

    def frac(self):
        """Fractional coordinates (3 x nqbz) of the k-points, k index i1 + n1*(i2 + n2*i3)."""
        n1, n2, n3 = self.n
        i = np.arange(self.nqbz)
        return np.stack([(i % n1) / n1, ((i // n1) % n2) / n2, (i // (n1 * n2)) / n3])


def tetra_mesh(n1: int, n2: int, n3: int, qlat=None) -> TetraMesh:
    """Each micro-cell is cut into the 6 tetrahedra that share its (0,0,0)-(1,1,1)
    diagonal (the reference picks the shortest diagonal; for a cubic cell all
    four are equal)."""
    ql = np.eye(3) if qlat is None else np.asarray(qlat, float)
    ext = [(i1, i2, i3) for i3 in range(n3 + 1) for i2 in range(n2 + 1) for i1 in range(n1 + 1)]
    widx = {e: j for j, e in enumerate(ext)}
    qbzw = np.array([ql @ np.array([e[0] / n1, e[1] / n2, e[2] / n3]) for e in ext]).T
    ib1bz = np.array([1 + (e[0] % n1) + n1 * ((e[1] % n2) + n2 * (e[2] % n3)) for e in ext], np.int32)
    paths = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for i3 in range(n3):
        for i2 in range(n2):
            for i1 in range(n1):
                for p in paths:
                    c = [0, 0, 0]
                    corners = [tuple(c)]
                    for ax in p:
                        c[ax] += 1
                        corners.append(tuple(c))
                    tets.append([widx[(i1 + a, i2 + b, i3 + d)] + 1 for a, b, d in corners])
    return TetraMesh((n1, n2, n3), ql, np.asfortranarray(qbzw), ib1bz, np.asfortranarray(np.array(tets, np.int32).T))


def tetra_bands(mesh: TetraMesh, nocc: int, nunocc: int, seed: int = 20261210, nmetal: int = 2,
                degenerate: bool = True, iq=(1, 0, 0)):
    """Smooth bands on the mesh: E_n(k) = c_n + w_n (cos 2pi f1 + cos 2pi f2 + cos 2pi f3)/3
    + a_n sin 2pi(f1 + 2 f2 - f3)/4 (eV), sorted per k as eigenvalues are.
    nmetal of the occupied bands cross E_F = 0 (Fermi-surface cuts); with
    `degenerate` the lowest band is doubled (degenerate-band symmetrisation).
    Returns ek1 (k), ek2 (k+q, q = iq/n on the mesh), both nband x nqbz in Ry,
    and E_F = 0."""
    rng = np.random.default_rng(seed)
    c = np.concatenate([np.sort(rng.uniform(-12.0, -2.0, nocc - nmetal)), rng.uniform(-0.6, 0.6, nmetal),
                        np.sort(rng.uniform(1.0, 60.0, nunocc))])
    w = np.concatenate([rng.uniform(0.5, 2.0, nocc - nmetal), rng.uniform(3.0, 5.0, nmetal),
                        rng.uniform(1.0, 4.0, nunocc)])
    a = rng.uniform(-0.3, 0.3, nocc + nunocc)
    if degenerate and nocc - nmetal >= 2:
        c[1], w[1], a[1] = c[0], w[0], a[0]

    def bands(f):
        g = (np.cos(2 * np.pi * f[0]) + np.cos(2 * np.pi * f[1]) + np.cos(2 * np.pi * f[2])) / 3.0
        h = np.sin(2 * np.pi * (f[0] + 2 * f[1] - f[2])) / 4.0
        e = c[:, None] + w[:, None] * g[None, :] + a[:, None] * h[None, :]
        return np.sort(e, axis=0) / RYDBERG_EV

    f = mesh.frac()
    fq = (f + np.array([iq[0] / mesh.n[0], iq[1] / mesh.n[1], iq[2] / mesh.n[2]])[:, None]) % 1.0
    return np.asfortranarray(bands(f)), np.asfortranarray(bands(fq)), 0.0


def tetra_frhis(ek1, ek2, oratio=1.03, dw=1e-5):
    """frhis for the band range of ek1/ek2 (getfreq2's omg2max, S/m_freq.f90:63-64)."""
    allE = np.concatenate([np.ravel(ek1), np.ravel(ek2)])
    omg = (allE.max() - allE.min()) * 0.5 + 0.2
    aa = oratio - 1.0
    bb = dw / aa
    iw = 0
    while True:
        iw += 1
        if bb * (math.exp(aa * (iw - 1)) - 1.0) > omg + 1e-6:
            break
    nwhis = iw + 2
    return np.array([bb * (math.exp(aa * (i - 1)) - 1.0) for i in range(1, nwhis + 2)])


def real_mesh(qp: QPoint, fr: np.ndarray, deltaw: float = 0.02, ffac: float = 10.0) -> np.ndarray:
    """freq_r(0:nw) in Hartree for the real-axis pole term.

    getwemax (S/m_freq.f90:197-268): the largest |e - E_F| among the external
    bands (+-2 deltaw) and the bands within ffac*esmr of E_F, plus ffac*esmr
    (Ry); getfreq (S/m_freq.f90:110-129): nw is the first iw >= 3 whose
    (frhis(iw-2)+frhis(iw-1))/2 exceeds wemax/2 (Hartree), at most nwhis-1;
    freq_r(0) = 0 and freq_r(i) the bin centres.  E_F = 0; deltaw is the
    GWinput default of the reference's samples (0.02)."""
    cfg = qp.cfg
    ef = 0.0
    eq = qp.e_q[:cfg.nq] / RYDBERG_EV
    below = eq - 2 * deltaw
    above = eq + 2 * deltaw
    omegav = min(0.0, float(below[below < ef].min())) if np.any(below < ef) else 0.0
    omegac = max(0.0, float(above[above > ef].max())) if np.any(above > ef) else 0.0
    allE = np.concatenate([qp.e_k.ravel(), qp.e_kq.ravel()]) / RYDBERG_EV
    sel = allE[allE < ef + ffac * cfg.esmr]
    emaxv = max(0.0, float(sel.max())) if sel.size else 0.0
    sel = allE[allE > ef - ffac * cfg.esmr]
    eminc = min(0.0, float(sel.min())) if sel.size else 0.0
    we = max(abs(emaxv - ef), abs(omegav - ef), abs(omegac - ef), abs(eminc - ef))
    wemax = we + ffac * cfg.esmr
    nwhis = len(fr) - 1
    nw = None
    for iw in range(3, nwhis + 1):
        if 0.5 * (fr[iw - 3] + fr[iw - 2]) > wemax / 2.0:
            nw = iw
            break
    if nw is None:
        raise ValueError("real_mesh: histogram mesh does not reach wemax")
    nw = min(nw, nwhis - 1)
    return np.concatenate([[0.0], 0.5 * (fr[:nw] + fr[1:nw + 1])])


@dataclass
class PairWeights:
    """Per-k band pairs with histogram weights, the tetwt5 output layout
    (n1b/n2b, ihw/nhw/jhw, whw; S/m_tetwt.f90:7-25)."""
    n1b: list       # per k: int array of occupied band indices (1-based)
    n2b: list       # per k: unoccupied band indices (1-based, in 1..nbt)
    ihw: list       # per k: first bin (1-based)
    nhw: list       # per k: number of bins
    jhw: list       # per k: pointer into whw (1-based)
    whw: np.ndarray # all weights


def band_windows(qp: QPoint):
    """Per-k (occupied, unoccupied) band counts: uniform (no, nu) unless the
    q-point carries uneven windows (qp.extra['no_k'], qp.extra['nu_k'])."""
    cfg = qp.cfg
    no_k = qp.extra.get("no_k", np.full(cfg.nk, cfg.no))
    nu_k = qp.extra.get("nu_k", np.full(cfg.nk, cfg.nu))
    return np.asarray(no_k, int), np.asarray(nu_k, int)


def pair_weights(qp: QPoint, fr: np.ndarray, nu_max: int = 0) -> PairWeights:
    """Pairs (occupied n <= no_k, unoccupied no_k < m <= no_k + min(nu_k, nu_max); nu_max = 0: all)."""
    cfg = qp.cfg
    nwhis = len(fr) - 1
    n1b, n2b, ihw, nhw, jhw = [], [], [], [], []
    whw_parts = []
    ptr = 1
    no_k, nu_k = band_windows(qp)
    for k in range(cfg.nk):
        nok = int(no_k[k])
        nuk = int(min(nu_k[k], nu_max) if nu_max else nu_k[k])
        occ = np.arange(1, nok + 1)
        unocc = np.arange(nok + 1, nok + nuk + 1)
        nn, mm = np.meshgrid(occ, unocc, indexing="ij")  # pair order: n outer, m inner
        nn = nn.ravel()
        mm = mm.ravel()
        d = (qp.e_kq[mm - 1, k] - qp.e_k[nn - 1, k]) / RYDBERG_EV * 0.5  # Hartree
        lo = d * (1.0 - cfg.window)
        hi = d * (1.0 + cfg.window)
        b0 = np.searchsorted(fr, lo, side="right")  # bin index (1-based) containing lo
        b1 = np.searchsorted(fr, hi, side="right")
        if b0.min() < 1 or b1.max() > nwhis:
            raise ValueError("histogram window outside the bin mesh")
        cnt = b1 - b0 + 1
        span = int(cnt.max())
        w = np.zeros((len(d), span))
        tot = math.pi * qp.wk[k]
        for s in range(span):
            b = b0 + s
            valid = b <= b1
            bl = fr[np.minimum(b, nwhis) - 1]
            br = fr[np.minimum(b, nwhis)]
            ov = np.clip(np.minimum(hi, br) - np.maximum(lo, bl), 0.0, None)
            w[:, s] = np.where(valid, tot * ov / (hi - lo), 0.0)
        n1b.append(nn.astype(np.int32))
        n2b.append(mm.astype(np.int32))
        ihw.append(b0.astype(np.int32))
        nhw.append(cnt.astype(np.int32))
        j = np.empty(len(d), np.int32)
        j[0] = ptr
        j[1:] = ptr + np.cumsum(cnt)[:-1]
        jhw.append(j)
        ptr += int(cnt.sum())
        mask = np.arange(span)[None, :] < cnt[:, None]
        whw_parts.append(w[mask])
    return PairWeights(n1b, n2b, ihw, nhw, jhw, np.concatenate(whw_parts))


def pair_weights_from_tetwt(r, jpm=0) -> PairWeights:
    """The tetrahedron weights (gettetwt's n1b/n2b/ihw/nhw/jhw/whw, per k) of
    one jpm as PairWeights for x0kf_index (jhw keeps pointing into the whole
    whw)."""
    nbnb = np.asarray(r["nbnb"])
    pick = (lambda a: a) if nbnb.ndim == 1 else (lambda a: a[..., jpm])
    nb = pick(nbnb)
    nq = len(nb)
    sl = lambda a, k: np.ascontiguousarray(pick(a)[:nb[k], k], np.int32)
    return PairWeights([sl(r["n1b"], k) for k in range(nq)], [sl(r["n2b"], k) for k in range(nq)],
                       [sl(r["ihw"], k) for k in range(nq)], [sl(r["nhw"], k) for k in range(nq)],
                       [sl(r["jhw"], k) for k in range(nq)], np.asarray(r["whw"], np.float64))


@dataclass
class X0Index:
    """The flattened index lists of x0kf_v4hz_init (S/x0kf_v4h.f90:48-142)."""
    kc: np.ndarray
    itc: np.ndarray
    itpc: np.ndarray
    jpmc: np.ndarray
    iwini: np.ndarray
    iwend: np.ndarray
    icouini: np.ndarray
    whwc: np.ndarray
    icounkmin: np.ndarray
    icounkmax: np.ndarray
    nkmin: np.ndarray
    nkmax: np.ndarray
    nkqmin: np.ndarray
    nkqmax: np.ndarray


def x0kf_index(pw: PairWeights, nctot: int = 0, npm: int = 1, nband: int = 0,
               pw_minus: PairWeights = None) -> X0Index:
    """Restates x0kf_v4hz_init job=0 then job=1 (S/x0kf_v4h.f90:96-138).

    pw holds the jpm = 1 pair tables; with npm = 2, pw_minus holds jpm = 2.
    Bands n > nband are core states (nband = 0: no core bands).  Job 0 takes
    per k the smallest n1b / n2b over all pairs and the largest over valence
    bands only; job 1 skips pairs with n1b > nkmax or n2b > nkqmax (which
    drops every core pair), maps it = nctot + n1b (valence) or n1b - nband
    (core), itp = ncc + n2b - nkqmin + 1 (valence; ncc = nctot when npm = 2)
    or n2b - nkqmin + 1 - nband (core), and takes the weights of each pair
    from whw(jhw + iw - ihw)."""
    tables = [pw] if npm == 1 else [pw, pw_minus]
    if npm == 2 and pw_minus is None:
        raise ValueError("x0kf_index: npm=2 needs the jpm=2 pair tables")
    nband = nband if nband > 0 else 1 << 30
    ncc = 0 if npm == 1 else nctot
    nk = len(pw.n1b)
    nkmin = np.full(nk, 999999, np.int32)
    nkqmin = np.full(nk, 999999, np.int32)
    nkmax = np.full(nk, -999999, np.int32)
    nkqmax = np.full(nk, -999999, np.int32)
    for t in tables:
        for k in range(nk):
            a, b = np.asarray(t.n1b[k]), np.asarray(t.n2b[k])
            if len(a) == 0:
                continue
            nkmin[k] = min(nkmin[k], a.min())
            nkqmin[k] = min(nkqmin[k], b.min())
            if (a <= nband).any():
                nkmax[k] = max(nkmax[k], a[a <= nband].max())
            if (b <= nband).any():
                nkqmax[k] = max(nkqmax[k], b[b <= nband].max())
    kc, itc, itpc, jpmc, iwini, iwend, icouini, whwc = ([] for _ in range(8))
    icounkmin = np.zeros(nk, np.int32)
    icounkmax = np.zeros(nk, np.int32)
    icoun = 0
    icount = 0
    for k in range(nk):
        icounkmin[k] = icoun + 1
        if npm == 2 and nkqmin[k] != 1:
            raise ValueError("x0kf_index: npm=2 needs nkqmin=1")
        for jpm, t in enumerate(tables, start=1):
            a, b = np.asarray(t.n1b[k], np.int64), np.asarray(t.n2b[k], np.int64)
            ih, nh, jh = (np.asarray(x, np.int64) for x in (t.ihw[k], t.nhw[k], t.jhw[k]))
            keep = (a <= nkmax[k]) & (b <= nkqmax[k])
            a, b, ih, nh, jh = a[keep], b[keep], ih[keep], nh[keep], jh[keep]
            npair = len(a)
            if npair == 0:
                continue
            kc.append(np.full(npair, k + 1))
            itc.append(np.where(a <= nband, nctot + a, a - nband))
            itpc.append(np.where(b <= nband, ncc + b - nkqmin[k] + 1, b - nkqmin[k] + 1 - nband))
            jpmc.append(np.full(npair, jpm))
            iwini.append(ih)
            iwend.append(ih + nh - 1)
            icouini.append(icount + 1 + np.concatenate([[0], np.cumsum(nh)[:-1]]))
            # bin iw of a pair has weight whw(jhw + iw - ihw): gather per pair
            off = np.repeat(jh - 1, nh) + (np.arange(int(nh.sum())) - np.repeat(np.cumsum(nh) - nh, nh))
            whwc.append(np.asarray(t.whw)[off])
            icoun += npair
            icount += int(nh.sum())
        icounkmax[k] = icoun
    cat = lambda xs, dt: np.ascontiguousarray(np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt))
    return X0Index(cat(kc, np.int32), cat(itc, np.int32), cat(itpc, np.int32), cat(jpmc, np.int32),
                   cat(iwini, np.int32), cat(iwend, np.int32), cat(icouini, np.int32),
                   cat(whwc, np.float64), icounkmin, icounkmax, nkmin, nkmax, nkqmin, nkqmax)


def make_uneven_qpoint(cfg: Config, no_range=(8, 24), nu_range=(150, 400), windows=None) -> QPoint:
    """A q-point whose k-points have their own band windows: occupied counts
    drawn from U{no_range}, unoccupied from U{nu_range} (per k, seeded), the
    uneven per-k ranges that ecalj's x0kf_v4hz_init and sxcf_scz_count
    handle (nkmin..nkqmax per k, S/x0kf_v4h.f90:91-137).  cfg.no / cfg.nu
    must be the range maxima; the k-th window uses the first no_k + nu_k
    bands, occupied energies for the first no_k of them.  windows = (no_k,
    nu_k) fixes the per-k counts instead of drawing them."""
    assert cfg.no == no_range[1] and cfg.nu == nu_range[1]
    qp = make_qpoint(cfg)
    rng = np.random.Generator(np.random.PCG64(cfg.seed + 3000))
    no_k = rng.integers(no_range[0], no_range[1] + 1, cfg.nk)
    nu_k = rng.integers(nu_range[0], nu_range[1] + 1, cfg.nk)
    if windows is not None:
        no_k, nu_k = np.asarray(windows[0], int), np.asarray(windows[1], int)
        if (len(no_k) != cfg.nk or len(nu_k) != cfg.nk or no_k.min() < no_range[0] or no_k.max() > no_range[1]
                or nu_k.min() < nu_range[0] or nu_k.max() > nu_range[1]):
            raise ValueError("make_uneven_qpoint: windows outside the ranges")
    for k in range(cfg.nk):  # energies of the k-th window: no_k occupied, then unoccupied
        for e in (qp.e_k, qp.e_kq, qp.e_qmk):
            e[:, k] = _energies(rng, int(no_k[k]), cfg.nbt - int(no_k[k]))
    qp.extra["no_k"] = no_k
    qp.extra["nu_k"] = nu_k
    return qp


def c4_qpoints(nq: int = 16, cfg: Config = None):
    """Configs of the batched uneven workload's q-points: nI from U[800, 1600]
    rounded to a multiple of 8 (the atom count), one seed per q-point."""
    cfg = cfg or CONFIGS["C4"]
    rng = np.random.Generator(np.random.PCG64(cfg.seed))
    nis = (rng.integers(800, 1601, nq) // cfg.natom) * cfg.natom
    return [cfg.scaled(ni=int(n), seed=cfg.seed + 1 + i) for i, n in enumerate(nis)]


def sigma_inputs(qp: QPoint, seed_offset: int = 1000):
    """Synthetic Sigma_x, Sigma_x(core) (Hartree) and vxc (eV) for the QP update."""
    cfg = qp.cfg
    rng = np.random.Generator(np.random.PCG64(cfg.seed + seed_offset))
    sex = rng.uniform(-0.8, -0.1, cfg.nq)
    sexcore = rng.uniform(-0.05, 0.0, cfg.nq)
    vxc = rng.uniform(-20.0, -5.0, cfg.nq)
    return sex, sexcore, vxc


def wc_phases(cfg: Config, seed_offset: int = 2000) -> np.ndarray:
    """Unit-modulus phases d_k(I): W_c at internal k is taken as
    D_k W_c(q) D_k^H (same spectrum, Hermitian, negative semidefinite)."""
    rng = np.random.Generator(np.random.PCG64(cfg.seed + seed_offset))
    ph = rng.uniform(0.0, 2.0 * math.pi, (cfg.ni, cfg.nk))
    ph[:, 0] = 0.0
    return np.asfortranarray(np.exp(1j * ph))


# --------------------------------------------------------------------------
# Gamma cell: offset directions q0i with their invariant tensors
def real_ylm(lmax: int, v: np.ndarray) -> np.ndarray:
    """Normalized real harmonics Y_L(v/|v|), L = l*l + l + m (0-based): m > 0 ~
    cos(m phi), m < 0 ~ sin(|m| phi), no Condon-Shortley sign (the convention
    of sylm/sylmnc, S/sylm.f90:1-121).  v: (..., 3) -> (..., (lmax+1)^2)."""
    v = np.asarray(v, np.float64)
    r = np.linalg.norm(v, axis=-1)
    ct = v[..., 2] / r
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = np.arctan2(v[..., 1], v[..., 0])
    out = np.zeros(v.shape[:-1] + ((lmax + 1) ** 2,))
    for m in range(lmax + 1):
        pmm = np.ones_like(ct)
        for k in range(1, m + 1):
            pmm = pmm * (2 * k - 1) * st
        p2, p1 = None, pmm
        for l in range(m, lmax + 1):
            if l == m:
                p = pmm
            elif l == m + 1:
                p = ct * (2 * m + 1) * pmm
            else:
                p = ((2 * l - 1) * ct * p1 - (l + m - 1) * p2) / (l - m)
            if l > m:
                p2, p1 = p1, p
            nrm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            c = l * l + l
            if m == 0:
                out[..., c] = nrm * p
            else:
                out[..., c + m] = math.sqrt(2.0) * nrm * p * np.cos(m * ph)
                out[..., c - m] = math.sqrt(2.0) * nrm * p * np.sin(m * ph)
    return out


def sphere_rule(n: int = 24):
    """Product rule on the unit sphere: points (N, 3) and weights (N,), exact
    for polynomials of degree < min(2n, 2n)."""
    x, w = np.polynomial.legendre.leggauss(n)
    ph = 2 * np.pi * np.arange(2 * n) / (2 * n)
    ct, p = np.meshgrid(x, ph, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    pts = np.stack([st * np.cos(p), st * np.sin(p), ct], -1).reshape(-1, 3)
    wts = (w[:, None] * np.full(2 * n, 2 * np.pi / (2 * n))[None, :]).ravel()
    return pts, wts


@dataclass
class GammaCell:
    q0i: np.ndarray        # (3, nq0i) offset directions (Cartesian, 2 pi/a)
    epinv: np.ndarray      # (3, 3, nq0i) invariant tensors
    epinvq0i: np.ndarray   # (nq0i, nq0i): <q0i_i| epinv_j |q0i_i> / |q0i_i|^2
    dmlx: np.ndarray       # (nq0i, 9): <e|epinv_i|e> = sum_L dmlx(i, L) Y_L(e)
    wklm: np.ndarray       # ((lmxax+1)^2,) Gamma-cell weights
    wbz: float             # BZ weight of q = 0
    lmxax: int


def gamma_cell(seed: int = 7, lmxax: int = 4, wbz: float = 1.0 / 512, aniso: bool = True) -> GammaCell:
    """Six offset directions and their invariant tensors without symmetry, as
    getallq0p / diele_invariant / m_q0p build them (S/m_q0p.f90:225-296,
    S/q0irre.f90:151-266): q0x = the three reciprocal axes (a slightly
    distorted cubic lattice) and three pair combinations; epinv_ = q q^T
    normalized, Gram-Schmidt orthonormalized in the Frobenius product;
    epinvq0i(i, j) = e_i . epinv_j . e_i; dmlx(i, L) = int e.epinv_i.e Y_L by
    quadrature.  wklm: positive l = 0 weight, small even l > 0 components
    (zero when aniso=False)."""
    rng = np.random.default_rng(seed)
    qlat = np.eye(3) + 0.08 * rng.standard_normal((3, 3))
    q0x = [qlat[:, i] for i in range(3)]
    for a, b in ((0, 1), (1, 2), (2, 0)):
        u, v = q0x[a], q0x[b]
        q0x.append((u - v) / 2 if np.sum((u - v) ** 2) < np.sum((u + v) ** 2) else (u + v) / 2)
    q0x = np.array(q0x).T * 0.01
    basis = []
    for i in range(6):
        e = q0x[:, i] / np.linalg.norm(q0x[:, i])
        t = np.outer(e, e)
        t = t / np.linalg.norm(t)
        for b in basis:
            t = t - np.sum(t * b) * b
        nt = np.linalg.norm(t)
        assert nt > 1e-6
        basis.append(t / nt)
    epinv = np.asfortranarray(np.stack(basis, -1))
    nq0i = 6
    epq = np.zeros((nq0i, nq0i))
    for i in range(nq0i):
        e = q0x[:, i] / np.linalg.norm(q0x[:, i])
        for j in range(nq0i):
            epq[i, j] = e @ epinv[:, :, j] @ e
    pts, wts = sphere_rule(12)
    Y = real_ylm(2, pts)
    dmlx = np.zeros((nq0i, 9))
    for i in range(nq0i):
        f = np.einsum("pa,ab,pb->p", pts, epinv[:, :, i], pts)
        dmlx[i] = (wts * f) @ Y
    nlk = (lmxax + 1) ** 2
    wklm = np.zeros(nlk)
    wklm[0] = 0.35 * wbz
    if aniso:
        for lm in range(1, nlk):
            l = int(math.isqrt(lm))
            if l % 2 == 0:
                wklm[lm] = 0.02 * wbz * rng.standard_normal() / l
    return GammaCell(np.asfortranarray(q0x), epinv, np.asfortranarray(epq), np.asfortranarray(dmlx),
                     wklm, float(wbz), lmxax)


# --------------------------------------------------------------------------
# Eigenfunction rotation: a two-atom cubic crystal with its space group
def cubic_point_group() -> np.ndarray:
    """The 48 signed permutation matrices (O_h), identity first: (3, 3, 48)."""
    import itertools

    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            g = np.zeros((3, 3))
            for i in range(3):
                g[i, perm[i]] = signs[i]
            ops.append(g)
    ops.sort(key=lambda g: (not np.allclose(g, np.eye(3)),))
    return np.stack(ops, -1)


def rotation_dlmm(g: np.ndarray, lmax: int) -> np.ndarray:
    """D_l(M, L) = int Y_M(r) Y_L(g^-1 r) dOmega in the real harmonics of
    real_ylm: the coefficients of the rotated function Y_L(g^-1 r).
    Shape (2 lmax+1, 2 lmax+1, lmax+1), index m + lmax."""
    pts, wts = sphere_rule(lmax + 2)
    gi = np.linalg.inv(g)
    Y = real_ylm(lmax, pts)
    Yr = real_ylm(lmax, pts @ gi.T)
    d = np.zeros((2 * lmax + 1, 2 * lmax + 1, lmax + 1))
    for l in range(lmax + 1):
        s = slice(l * l, l * l + 2 * l + 1)
        blk = (Y[:, s] * wts[:, None]).T @ Yr[:, s]
        d[lmax - l:lmax + l + 1, lmax - l:lmax + l + 1, l] = blk
    return d


@dataclass
class RotationCase:
    plat: np.ndarray
    qlat: np.ndarray
    pos: np.ndarray          # (3, nbas)
    nbas: int
    symops: np.ndarray       # (3, 3, ngrp)
    shtvg: np.ndarray        # (3, ngrp)
    miat: np.ndarray         # (nbas, ngrp), 1-based
    tiat: np.ndarray         # (3, nbas, ngrp)
    dlmm: np.ndarray         # (2 lmxax+1, 2 lmxax+1, lmxax+1, ngrp)
    lmxax: int
    ibas_tbl: np.ndarray     # @MNLA_CPHI tables (1-based ibas, l, k; offset = first row - 1)
    l_tbl: np.ndarray
    k_tbl: np.ndarray
    offset_tbl: np.ndarray
    ndima: int
    qtt: np.ndarray          # (3, nqtt); the first nqi are the irreducible points
    nqi: int
    iqmap: np.ndarray        # (nqtt,) 1-based: the irreducible point each q is rotated from
    igmap: np.ndarray        # (nqtt,) 1-based operation
    iqimap: np.ndarray       # (nqtt,) 1-based index among the irreducible points
    ngp: np.ndarray          # (nqtt,)
    ngpmx: int
    imx: int
    ngvecp: np.ndarray       # (3, ngpmx, nqtt)
    ngvecprev: np.ndarray    # (2 imx+1,)*3 + (nqtt,), 1-based, 9999 where absent
    alpha: dict              # Gaussian exponent of each radial function (l, k)


def rotation_case(seed: int = 3, nqi: int = 3, ntarget: int = 5, lmxax: int = 3, qcut: float = 2.3) -> RotationCase:
    """Simple cubic lattice (a = 1) with identical atoms at (0,0,0) and
    (1/2,1/2,1/2): the 48 cubic point operations with shift 0 and with shift
    (1/2,1/2,1/2) map it to itself (miat, tiat: g tau_a + s = tau_b + t_a).
    Orbitals per atom: (l,k) = (0,1) (0,2) (1,1) (2,1) (3,1), rows m = -l..l,
    numbered as sugw writes @MNLA_CPHI.  qtt: nqi irreducible points, then for
    each ntarget random operations with qtarget = g q + G0 (G0 folds into the
    first cell, plus a random lattice vector).  G lists: |q+G| < qcut."""
    rng = np.random.default_rng(seed)
    plat = np.eye(3)
    qlat = np.eye(3)
    pos = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]).T
    nbas = 2
    pg = cubic_point_group()
    ops, shifts = [], []
    for s in (np.zeros(3), np.full(3, 0.5)):
        for i in range(pg.shape[2]):
            ops.append(pg[:, :, i])
            shifts.append(s)
    ngrp = len(ops)
    symops = np.asfortranarray(np.stack(ops, -1))
    shtvg = np.asfortranarray(np.stack(shifts, -1))
    miat = np.zeros((nbas, ngrp), np.int32)
    tiat = np.zeros((3, nbas, ngrp))
    for ig in range(ngrp):
        for a in range(nbas):
            x = symops[:, :, ig] @ pos[:, a] + shtvg[:, ig]
            for b in range(nbas):
                t = x - pos[:, b]
                if np.allclose(t, np.round(t), atol=1e-12):
                    miat[a, ig] = b + 1
                    tiat[:, a, ig] = np.round(t)
                    break
            assert miat[a, ig] > 0
    dlmm = np.asfortranarray(np.stack([rotation_dlmm(symops[:, :, ig], lmxax) for ig in range(ngrp)], -1))
    orbs = [(0, 1), (0, 2), (1, 1), (2, 1), (3, 1)]
    ibas_tbl, l_tbl, k_tbl, off = [], [], [], []
    row = 0
    for a in range(nbas):
        for l, k in sorted(orbs):
            ibas_tbl.append(a + 1)
            l_tbl.append(l)
            k_tbl.append(k)
            off.append(row)
            row += 2 * l + 1
    ndima = row
    qi = rng.uniform(-0.5, 0.5, (3, nqi))
    qtt = [qi[:, i] for i in range(nqi)]
    iqmap, igmap, iqimap = list(range(1, nqi + 1)), [1] * nqi, list(range(1, nqi + 1))
    for i in range(nqi):
        for _ in range(ntarget):
            ig = int(rng.integers(1, ngrp))
            gq = symops[:, :, ig] @ qi[:, i]
            g0 = -np.round(gq) + rng.integers(-1, 2, 3)
            qtt.append(gq + qlat @ g0)
            iqmap.append(i + 1)
            igmap.append(ig + 1)
            iqimap.append(i + 1)
    qtt = np.asfortranarray(np.array(qtt).T)
    nqtt = qtt.shape[1]
    lists = []
    r = int(np.ceil(qcut)) + 3
    grid = np.array(np.meshgrid(*[np.arange(-r, r + 1)] * 3, indexing="ij")).reshape(3, -1).T
    for iq in range(nqtt):
        d2 = np.sum((qtt[:, iq][None, :] + grid @ qlat.T) ** 2, axis=1)
        keep = d2 < qcut ** 2
        gap = np.abs(d2 - qcut ** 2).min()
        assert gap > 1e-9, "cutoff sits on a shell"
        sel = grid[keep]
        order = np.lexsort((sel[:, 2], sel[:, 1], sel[:, 0], np.round(d2[keep], 12)))
        lists.append(sel[order])
    ngp = np.array([len(x) for x in lists], np.int32)
    ngpmx = int(ngp.max())
    imx = int(max(np.abs(x).max() for x in lists))
    nc = 2 * imx + 1
    ngvecp = np.zeros((3, ngpmx, nqtt), np.int32, order="F")
    ngvecprev = np.full((nc, nc, nc, nqtt), 9999, np.int32, order="F")
    for iq, x in enumerate(lists):
        ngvecp[:, :len(x), iq] = x.T
        for ig, g in enumerate(x):
            ngvecprev[g[0] + imx, g[1] + imx, g[2] + imx, iq] = ig + 1
    alpha = {(l, k): 9.0 + 3.0 * k + l for l, k in orbs}
    return RotationCase(plat, qlat, pos, nbas, symops, shtvg, miat, tiat, dlmm, lmxax,
                        np.array(ibas_tbl, np.int32), np.array(l_tbl, np.int32), np.array(k_tbl, np.int32),
                        np.array(off, np.int32), ndima, qtt, nqi, np.array(iqmap, np.int32),
                        np.array(igmap, np.int32), np.array(iqimap, np.int32), ngp, ngpmx, imx, ngvecp, ngvecprev,
                        alpha)
