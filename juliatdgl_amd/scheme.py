"""Host side of the stepper: material, grid, masks and the Stepper's coefficients.

In the reference these are Julia objects (gtdgl/src/model.jl, grid.jl, geometry.jl,
scheme.jl, field.jl:151-180, implicit.jl:126-140) that TDGLGPU reads when it builds its device
arrays.  Julia is not available to this package's tests, so the derivation is restated here, in
the reference's order of operations (IEEE doubles in Python round as Julia's Float64 do), and
`Stepper.desc()` packs it into the C ABI's ``tdgl_desc``.  A Julia caller builds the same
description from its own ``TDGL.Stepper`` instead (julia/TDGLHIP.jl).

Only what the HIP stepper runs is restated: N bands (isotropic or anisotropic masses), KWT,
Euler substeps, RKL1/RKC1 stages, the rotating frame, explicit / separable / PCG field solves,
slab and bar self-field boundaries.  Heat, drag, noise, materials per region, sculpted vacuum,
normal fluids and the open boundary are refused (``Stepper`` raises NotImplementedError).
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

INF = math.inf


# ----------------------------------------------------------------------------- model.jl
@dataclass(frozen=True)
class Band:
    """One band (model.jl:26-44): mass ratio m (stiffness 1/m), Γ, a, β, KWT γ̃, masses per axis."""
    m: float = 1.0
    Gamma: float = 1.0
    a: float = 1.0
    beta: float = 1.0
    gkwt: float = 0.0
    ma: tuple = None

    def __post_init__(self):
        ma = self.ma if self.ma is not None else (self.m, self.m, self.m)
        ma = tuple(float(x) for x in ma)
        if len(ma) != 3:
            raise ValueError("Band: m_axes is (m_x, m_y, m_z)")
        object.__setattr__(self, "ma", ma)
        if self.ma is not None and ma != (self.m, self.m, self.m):
            object.__setattr__(self, "m", ma[0])
        if not (self.m > 0 and all(x > 0 for x in ma) and self.Gamma > 0 and self.beta > 0):
            raise ValueError("Band: m (or m_axes), Γ and β must be positive")
        if self.gkwt < 0:
            raise ValueError("Band: γkwt must be >= 0")

    @staticmethod
    def make(m=1.0, g=None, Gamma=1.0, a=1.0, beta=1.0, gkwt=0.0, m_axes=None):
        mm = float(m) if g is None else 1.0 / float(g)
        ma = (mm, mm, mm) if m_axes is None else tuple(float(x) for x in m_axes)
        if m_axes is not None:
            mm = ma[0]
        return Band(mm, float(Gamma), float(a), float(beta), float(gkwt), ma)

    def stiffness(self, axis: Optional[int] = None) -> float:
        return 1.0 / self.m if axis is None else 1.0 / self.ma[axis]

    @property
    def isotropic(self) -> bool:
        return self.ma[0] == self.ma[1] == self.ma[2] == self.m


U_DIRTY = math.pi ** 4 / (14 * 1.2020569031595942)


@dataclass(frozen=True)
class Model:
    """Model{N} (model.jl:73-90): bands, symmetric Josephson matrix γ, κ, σn, bath T."""
    bands: tuple
    gamma: tuple
    kappa: float
    sigma_n: float = 1.0
    T: float = 0.0

    @staticmethod
    def make(bands: Sequence[Band], gamma, kappa, sigma_n=1.0, T=0.0):
        N = len(bands)
        g = np.asarray(gamma, dtype=float)
        if g.shape != (N, N):
            raise ValueError(f"Model: γ must be {N}x{N}")
        for mu in range(N):
            for nu in range(N):
                if g[mu, nu] != g[nu, mu]:
                    raise ValueError("Model: γ must be symmetric")
                if mu == nu and g[mu, mu] != 0:
                    raise ValueError("Model: the diagonal of γ must be zero")
        if sigma_n == "dirty":
            sigma_n = sum(b.stiffness() * b.Gamma / b.beta for b in bands) / U_DIRTY
        if not (kappa > 0 and sigma_n > 0):
            raise ValueError("Model: κ and σn must be positive")
        return Model(tuple(bands), tuple(tuple(float(x) for x in row) for row in g), float(kappa),
                     float(sigma_n), float(T))

    @property
    def N(self) -> int:
        return len(self.bands)

    def eps(self, mu: int) -> float:
        """ε_µ = a_µ − T (model.jl:95)."""
        return self.bands[mu].a - self.T


def two_band(xi_ratio, a2n=0.59, gamma12=1e-3, gamma_ratio=1.0, kappa=1.0, sigma_n=1.0, T=0.0,
             beta2=1.0, gamma_kwt=0.0) -> Model:
    """two_band (model.jl:104-111)."""
    mr = 1 / (xi_ratio * xi_ratio * (a2n / beta2))
    gk = (gamma_kwt, gamma_kwt) if np.isscalar(gamma_kwt) else tuple(gamma_kwt)
    return Model.make((Band.make(gkwt=gk[0]), Band.make(m=mr, Gamma=gamma_ratio, a=a2n, beta=beta2, gkwt=gk[1])),
                      [[0.0, gamma12], [gamma12, 0.0]], kappa=kappa, sigma_n=sigma_n, T=T)


def equilibrium(md: Model):
    """Coupled homogeneous equilibrium amplitudes (model.jl:121-141)."""
    b = md.bands
    N = md.N
    p = [math.sqrt(max(md.eps(mu), 0.0) / b[mu].beta) for mu in range(N)]
    for _ in range(100_000):
        dmax = 0.0
        for mu in range(N):
            e = md.eps(mu)
            be = b[mu].beta
            c = _jsum([0.0 if nu == mu else abs(md.gamma[mu][nu]) * p[nu] for nu in range(N)])
            x = math.sqrt(max(e, 0.0) / be) + float(np.cbrt(c / be))
            for _ in range(200):
                f = be * x ** 3 - e * x - c
                df = 3 * be * x * x - e
                if f <= 0 or df <= 0:
                    break
                dx = f / df
                x -= dx
                if dx <= 1e-15 * x:
                    break
            dmax = max(dmax, abs(x - p[mu]))
            p[mu] = x
        if dmax < 1e-14:
            break
    return tuple(p)


def _jsum(xs):
    """Julia's sum over a short collection: a left fold from the first element."""
    s = xs[0]
    for x in xs[1:]:
        s = s + x
    return s


# ----------------------------------------------------------------------------- grid.jl
@dataclass(frozen=True)
class Grid:
    """Grid (grid.jl:109-127): cells with padding, the block x1..x2+1 (1-based nodes)."""
    Nx: int
    Ny: int
    Nz: int
    x1: int
    x2: int
    y1: int
    y2: int
    z1: int
    z2: int
    dx: float
    dy: float
    dz: float

    @staticmethod
    def make(nx, ny, nz, dx, dy=None, dz=None, buf=1, bufx=None, bufy=None, bufz=None):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        bufx = buf if bufx is None else bufx
        bufy = buf if bufy is None else bufy
        bufz = buf if bufz is None else bufz
        if min(nx, ny, nz) < 2:
            raise ValueError("Grid: at least 2 cells per axis")
        if min(bufx, bufy, bufz) < 1:
            raise ValueError("Grid: at least one padding cell per face")
        return Grid(nx + 2 * bufx, ny + 2 * bufy, nz + 2 * bufz, bufx + 1, bufx + nx, bufy + 1, bufy + ny,
                    bufz + 1, bufz + nz, float(dx), float(dy), float(dz))

    @property
    def nodes(self):
        return (self.Nx + 1, self.Ny + 1, self.Nz + 1)

    @property
    def cells(self) -> int:
        """Sample cells nx·ny·nz (the unit of cell-cycles/s)."""
        return (self.x2 - self.x1 + 1) * (self.y2 - self.y1 + 1) * (self.z2 - self.z1 + 1)

    def slab_thickness(self) -> float:
        return (self.z2 - self.z1 + 1) * self.dz

    def dl(self) -> float:
        return 2.0 / (1 / (self.dx * self.dx) + 1 / (self.dy * self.dy) + 1 / (self.dz * self.dz))


@dataclass
class Masks:
    """Masks (grid.jl:135-143), node-shaped Float64 arrays, Fortran order; `vac` the vacuum of a
    sculpted sample and `mat` the material index per node (juliatdgl_amd.geometry), None without."""
    f: np.ndarray
    tx: np.ndarray
    ty: np.ndarray
    tz: np.ndarray
    tpsi: np.ndarray
    vac: Optional[np.ndarray] = None
    mat: Optional[np.ndarray] = None


def box_masks(gr: Grid) -> Masks:
    """box_masks (grid.jl:148-158)."""
    sh = gr.nodes
    f = np.ones(sh, order="F")
    tx = np.zeros(sh, order="F")
    ty = np.zeros(sh, order="F")
    tz = np.zeros(sh, order="F")
    tp = np.zeros(sh, order="F")
    x1, x2, y1, y2, z1, z2 = gr.x1 - 1, gr.x2 - 1, gr.y1 - 1, gr.y2 - 1, gr.z1 - 1, gr.z2 - 1
    tx[x1:x2 + 1, y1:y2 + 2, z1:z2 + 2] = 1.0
    ty[x1:x2 + 2, y1:y2 + 1, z1:z2 + 2] = 1.0
    tz[x1:x2 + 2, y1:y2 + 2, z1:z2 + 1] = 1.0
    tp[x1:x2 + 2, y1:y2 + 2, z1:z2 + 2] = 1.0
    return Masks(f, tx, ty, tz, tp)


def catalogue_masks(gr: Grid, mode: str = "SIM", wedge_depth: float = 0.35, wedge_half: float = 0.20,
                    wedge_depth2: float = 0.0) -> Masks:
    """build_masks(GeomSpec(GEOM_MODE=mode, wedge_depth, wedge_half, wedge_depth2)) of a catalogue
    mode or its _REF mirror (geometry.jl:268-476; juliatdgl_amd.geometry), J along x, z_mode through,
    no shifts."""
    from .geometry import GeomSpec, build_masks
    mk, _ = build_masks(GeomSpec(GEOM_MODE=mode, wedge_depth=wedge_depth, wedge_half=wedge_half,
                                 wedge_depth2=wedge_depth2), gr)
    return mk


def wedge_masks(gr: Grid, wedge_depth=0.35, wedge_half=0.20) -> Masks:
    """The catalogue WEDGE notch on face y1, through the thickness (geometry.jl:299-308,
    paint_defect! 'wedge' :121-135, build_defect_list :268-270, z_mode through)."""
    return catalogue_masks(gr, "WEDGE", wedge_depth, wedge_half)


def notch_masks(gr: Grid, xr, yr) -> Masks:
    """Condensation cut on cells x in [xa, xb), y in [ya, yb), counted from the sample's first cell
    (x1, y1), through the thickness: f = 0 on those nodes."""
    mk = box_masks(gr)
    for i in range(xr[0], xr[1]):
        for j in range(yr[0], yr[1]):
            mk.f[gr.x1 - 1 + i, gr.y1 - 1 + j, gr.z1 - 1:gr.z2] = 0.0
    return mk


def sloped_wedge_masks(gr: Grid, angle_deg=30.0, depth_frac=0.35, x_frac=0.5) -> Masks:
    """A wedge cut into face y1 whose sloped face makes `angle_deg` with the face: at the
    position x_frac along the sample the cut is depth_frac·ny deep; it shallows along +x with
    slope tan(angle) and ends in a vertical wall on its −x side."""
    mk = box_masks(gr)
    nx, ny = gr.x2 - gr.x1 + 1, gr.y2 - gr.y1 + 1
    x0 = x_frac * nx
    d0 = depth_frac * ny
    t = math.tan(math.radians(angle_deg))
    for i in range(nx + 1):
        x = float(i)
        if x < x0:
            continue
        depth = d0 - (x - x0) * t
        if depth <= 0:
            continue
        for j in range(ny + 1):
            if j < depth:
                mk.f[gr.x1 - 1 + i, gr.y1 - 1 + j, gr.z1 - 1:gr.z2] = 0.0
    return mk


def hole_masks(gr: Grid, center, radius) -> Masks:
    """A cylindrical hole through the thickness: f = 0 within `radius` cells of `center` (x, y
    in sample cells)."""
    mk = box_masks(gr)
    nx, ny = gr.x2 - gr.x1 + 1, gr.y2 - gr.y1 + 1
    for i in range(nx + 1):
        for j in range(ny + 1):
            if (i - center[0]) ** 2 + (j - center[1]) ** 2 <= radius ** 2:
                mk.f[gr.x1 - 1 + i, gr.y1 - 1 + j, gr.z1 - 1:gr.z2] = 0.0
    return mk


# ----------------------------------------------------------------------------- scheme.jl
RKC_EPS = 0.5


def _rkc_cheb(s: int, w0: float):
    """(T_s(ω₀), T_s'(ω₀)) (scheme.jl:82-89)."""
    if s == 0:
        return (1.0, 0.0)
    T0, T1, d0, d1 = 1.0, w0, 0.0, 1.0
    for _ in range(2, s + 1):
        T0, T1, d0, d1 = T1, 2 * w0 * T1 - T0, d1, 2 * T1 + 2 * w0 * d1 - d0
    return T1, d1


@dataclass
class RKCCoef:
    """RKCCoef (scheme.jl:77-100): q = ω₀/ω₁, per-stage μ, ν, c, and SK-ROCK's weights of the
    noise in the first stage sk = (ν₁, κ₁) = (sω₁/2, sω₁/ω₀)."""
    q: float
    mu: list
    nu: list
    c: list
    sk: tuple = (0.0, 0.0)

    @staticmethod
    def make(s: int, eps: float = RKC_EPS) -> "RKCCoef":
        w0 = 1 + eps / (s * s)
        Ts, dTs = _rkc_cheb(s, w0)
        w1 = Ts / dTs
        mu = [1.0] * s
        nu = [0.0] * s
        c = [0.0] * s
        for j in range(1, s + 1):
            Tj, dTj = _rkc_cheb(j, w0)
            c[j - 1] = w1 * dTj / Tj
            if j >= 2:
                mu[j - 1] = 2 * w0 * _rkc_cheb(j - 1, w0)[0] / Tj
                nu[j - 1] = -_rkc_cheb(j - 2, w0)[0] / Tj
        c[s - 1] = 1.0
        return RKCCoef(w0 / w1, mu, nu, c, (s * w1 / 2, s * w1 / w0))


def _rkc_ratio(s: int, eps: float = RKC_EPS) -> float:
    w0 = 1 + eps / (s * s)
    Ts, dTs = _rkc_cheb(s, w0)
    return (1 + w0) / (2 * Ts / dTs)


@dataclass(frozen=True)
class BandCoef:
    """BandCoef (scheme.jl:21-28)."""
    kx: float
    ky: float
    kz: float
    c: float
    eps: float
    beta: float
    cpl: tuple
    g2: float

    def divided(self, n) -> "BandCoef":
        """_divided (scheme.jl:63-64)."""
        n = float(n)
        return BandCoef(self.kx / n, self.ky / n, self.kz / n, self.c / n, self.eps, self.beta,
                        tuple(x / n for x in self.cpl), self.g2)


def _psi2_bound(md: Model):
    """|Ψ_µ|² of the coupled homogeneous equilibrium (scheme.jl:38-52)."""
    b = md.bands
    N = md.N
    p = [math.sqrt(max(md.eps(mu), 0.0) / b[mu].beta) + 1e-3 for mu in range(N)]
    done = False
    for _ in range(10_000):
        q = []
        for mu in range(N):
            c = _jsum([0.0 if nu == mu else abs(md.gamma[mu][nu]) * p[nu] for nu in range(N)]) / p[mu]
            q.append(0.5 * p[mu] + 0.5 * math.sqrt(max(md.eps(mu) + c, 0.0) / b[mu].beta))
        if max(abs(qq - pp) for qq, pp in zip(q, p)) < 1e-13:
            p = q
            done = True
            break
        p = q
    if not done:
        p = list(equilibrium(md))
    out = []
    for mu in range(N):
        if p[mu] < 2e-3 and all(nu == mu or md.gamma[mu][nu] == 0 for nu in range(N)):
            out.append(max(md.eps(mu), 0.0) / b[mu].beta)
        else:
            out.append(p[mu] * p[mu])
    return tuple(out)


def _band_bounds(md: Model, gr: Grid, l: float):
    """Stability bounds bb0 (no KWT), bbk (KWT) and the supercurrent stiffness λL (scheme.jl:107-137)."""
    bands = md.bands
    N = md.N
    ps = _psi2_bound(md)
    P2 = []
    for mu in range(N):
        p, q = ps[mu], bands[mu].a / bands[mu].beta
        P2.append(q if (q > 0 and p <= 1.01 * q) else p)
    hr = []
    for mu in range(N):
        b = bands[mu]
        if P2[mu] == b.a / b.beta and b.beta == 1:
            hr.append(2 * (b.a / b.beta))
        else:
            hr.append(max(3 * b.beta * P2[mu] - md.eps(mu), 0.0))
    dif = []
    for mu in range(N):
        b = bands[mu]
        if b.isotropic:
            dif.append(8 / (l * b.m))
        else:
            dif.append(4 * (1 / (b.ma[0] * (gr.dx * gr.dx)) + 1 / (b.ma[1] * (gr.dy * gr.dy))
                            + 1 / (b.ma[2] * (gr.dz * gr.dz))))
    bb0 = [2 * bands[mu].Gamma / (dif[mu] + hr[mu]) for mu in range(N)]
    bbk = [bb0[mu] / math.sqrt(1 + bands[mu].gkwt ** 2 * P2[mu]) if bands[mu].gkwt > 0 else bb0[mu]
           for mu in range(N)]
    ns = [bands[mu].a / bands[mu].beta if bands[mu].a > 0 else P2[mu] for mu in range(N)]
    terms = [(bands[mu].stiffness() if bands[mu].isotropic else 1 / min(bands[mu].ma)) * ns[mu] for mu in range(N)]
    lamL = _jsum(terms)
    return bb0, bbk, lamL


def _bandcoef(md: Model, mu: int, dt: float, rx: float, ry: float, rz: float, mref) -> BandCoef:
    """_bandcoef (scheme.jl:141-147)."""
    b = md.bands[mu]
    c = dt / b.Gamma
    Gh = dt / c
    return BandCoef(rx / (mref[0] * Gh), ry / (mref[1] * Gh), rz / (mref[2] * Gh), c, md.eps(mu), b.beta,
                    tuple((dt / Gh) * md.gamma[mu][nu] for nu in range(md.N)), b.gkwt ** 2)


@dataclass(frozen=True)
class FieldCoef:
    """FieldCoef (scheme.jl:149-153)."""
    d1: float
    sxy: float
    sxz: float
    syx: float
    syz: float
    szx: float
    szy: float
    a: float


def link_conductivity(gr: Grid, svac: float, sn: float):
    """σ_l/σ_n per link: 1 in the bar y1:y2+1, z1:z2+1 along x, σvac/σn elsewhere (implicit.jl:126-140)."""
    r = svac / sn
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    j = np.arange(1, Ny + 2)
    k = np.arange(1, Nz + 2)
    iny = (gr.y1 <= j) & (j <= gr.y2 + 1)
    inz = (gr.z1 <= k) & (k <= gr.z2 + 1)
    Dx = np.full((Nx, Ny + 1, Nz + 1), r, order="F")
    Dx[:, iny[:, None] & inz[None, :]] = 1.0
    Dy = np.full((Nx + 1, Ny, Nz + 1), r, order="F")
    m = (iny[:-1] & iny[1:])[:, None] & inz[None, :]
    Dy[:, m] = 1.0
    Dz = np.full((Nx + 1, Ny + 1, Nz), r, order="F")
    m = iny[:, None] & (inz[:-1] & inz[1:])[None, :]
    Dz[:, m] = 1.0
    return Dx, Dy, Dz


def bar_field(gr: Grid, nq: int = 8):
    """Field of the injected current in the bar on the outer faces, per unit HJ: the 2D
    Biot–Savart midpoint quadrature of field.jl:151-180 (nq points per cell and axis).

    The evaluation points (nodes or centres of the y and z edges) and the quadrature points share
    the step hq = dy/nq (kq = dz/nq), so with nq even every difference y − Y_a is hq (m + 1/2) and
    z − Z_b is kq (n + 1/2) for integers m, n: each field value is the sum of one kernel table
    T(m, n) over an ny x nz window, done here as two matrix products with banded indicator
    matrices (the reference sums the points one by one; the sums agree to rounding, ~1e-15
    relative, and the stepper and the oracle read the same table)."""
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    y1, y2, z1, z2, dy, dz = gr.y1, gr.y2, gr.z1, gr.z2, gr.dy, gr.dz
    ya, yb = (y1 - 1) * dy, y2 * dy
    za, zb = (z1 - 1) * dz, z2 * dz
    d = (z2 - z1 + 1) * dz
    c = (-2 / d) / (2 * math.pi)
    ny = nq * (y2 - y1 + 1)
    nz = nq * (z2 - z1 + 1)
    hq = (yb - ya) / ny
    kq = (zb - za) / nz
    if nq % 2:
        raise ValueError("bar_field: nq must be even")

    def window_sum(mj, nk, hyfield):
        # Σ_{a=1..ny} Σ_{b=1..nz} T(mj − a, nk − b) for every pair (mj, nk)
        m = np.arange(mj.min() - ny, mj.max())          # every m a window reaches
        n = np.arange(nk.min() - nz, nk.max())
        u = hq * (m + 0.5)
        v = kq * (n + 0.5)
        U, V = np.meshgrid(u, v, indexing="ij")
        r2 = U * U + V * V
        T = (-V / r2) if hyfield else (U / r2)
        Bn = ((n[:, None] <= nk[None, :] - 1) & (n[:, None] >= nk[None, :] - nz)).astype(float)
        Am = ((m[None, :] <= mj[:, None] - 1) & (m[None, :] >= mj[:, None] - ny)).astype(float)
        return (Am @ (T @ Bn)) * (c * hq * kq)

    # hy at (y nodes j = 1..Ny+1, z centres k = 1..Nz): y − Y_a = hq((j − y1) nq − a + 1/2),
    # z − Z_b = kq((k − z1) nq + nq/2 − b + 1/2)
    HY = window_sum((np.arange(1, Ny + 2) - y1) * nq, (np.arange(1, Nz + 1) - z1) * nq + nq // 2, True)
    # hz at (y centres j = 1..Ny, z nodes k = 1..Nz+1)
    HZ = window_sum((np.arange(1, Ny + 1) - y1) * nq + nq // 2, (np.arange(1, Nz + 2) - z1) * nq, False)
    F = lambda a: np.asfortranarray(a)
    hy_z = F(HY[:, [0, Nz - 1]])
    hy_y = F(HY[[0, Ny], :])
    hz_z = F(HZ[:, [0, Nz]])
    hz_y = F(HZ[[0, Ny - 1], :])
    return dict(hy_z=hy_z, hy_y=hy_y, hz_z=hz_z, hz_y=hz_y, hz_x=F(HZ.copy()), hy_x=F(HY.copy()))


def bar_field_direct(gr: Grid, nq: int = 8):
    """Field of the injected current in the bar on the outer faces, per unit HJ: the 2D
    Biot–Savart quadrature of field.jl:151-180 (midpoint rule, nq points per cell)."""
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    y1, y2, z1, z2, dy, dz = gr.y1, gr.y2, gr.z1, gr.z2, gr.dy, gr.dz
    ya, yb = (y1 - 1) * dy, y2 * dy
    za, zb = (z1 - 1) * dz, z2 * dz
    d = (z2 - z1 + 1) * dz
    c = (-2 / d) / (2 * math.pi)
    ny = nq * (y2 - y1 + 1)
    nz = nq * (z2 - z1 + 1)
    hq = (yb - ya) / ny
    kq = (zb - za) / nz
    # quadrature points in the order of the reference's loops (b outer, a inner)
    yp = np.array([ya + (a - 0.5) * hq for a in range(1, ny + 1)])
    zp = np.array([za + (b - 0.5) * kq for b in range(1, nz + 1)])

    Y, Z = np.meshgrid(yp, zp, indexing="ij")

    def h(y, z):
        # the reference sums the ny·nz quadrature points one by one; the pairwise sum here agrees
        # with it to rounding (~1e-16 relative), and the stepper and the oracle read the same table
        dyv = y - Y
        dzv = z - Z
        r2 = dyv * dyv + dzv * dzv
        hy = float(np.sum(-dzv / r2))
        hz = float(np.sum(dyv / r2))
        return (c * hy * hq * kq, c * hz * hq * kq)

    def yc(j): return (j - 0.5) * dy
    def yn(j): return (j - 1) * dy
    def zc(k): return (k - 0.5) * dz
    def zn(k): return (k - 1) * dz

    cache = {}

    def hh(y, z):
        key = (y, z)
        if key not in cache:
            cache[key] = h(y, z)
        return cache[key]

    F = lambda sh: np.zeros(sh, order="F")
    hy_z = F((Ny + 1, 2))
    for c_, k in enumerate((1, Nz)):
        for j in range(1, Ny + 2):
            hy_z[j - 1, c_] = hh(yn(j), zc(k))[0]
    hy_y = F((2, Nz))
    for c_, j in enumerate((1, Ny + 1)):
        for k in range(1, Nz + 1):
            hy_y[c_, k - 1] = hh(yn(j), zc(k))[0]
    hz_z = F((Ny, 2))
    for c_, k in enumerate((1, Nz + 1)):
        for j in range(1, Ny + 1):
            hz_z[j - 1, c_] = hh(yc(j), zn(k))[1]
    hz_y = F((2, Nz + 1))
    for c_, j in enumerate((1, Ny)):
        for k in range(1, Nz + 2):
            hz_y[c_, k - 1] = hh(yc(j), zn(k))[1]
    hz_x = F((Ny, Nz + 1))
    for j in range(1, Ny + 1):
        for k in range(1, Nz + 2):
            hz_x[j - 1, k - 1] = hh(yc(j), zn(k))[1]
    hy_x = F((Ny + 1, Nz))
    for j in range(1, Ny + 2):
        for k in range(1, Nz + 1):
            hy_x[j - 1, k - 1] = hh(yn(j), zc(k))[0]
    return dict(hy_z=hy_z, hy_y=hy_y, hz_z=hz_z, hz_y=hz_y, hz_x=hz_x, hy_x=hy_x)


FIELD_EXPLICIT, FIELD_SEPARABLE, FIELD_PCG, FIELD_EXPLICIT_SCALED = 0, 1, 2, 3


QP_NONE, QP_TWO, QP_N = 0, 1, 2


# ---------------------------------------------------------------- the open boundary (openbc.jl)
def _scaled_besseli(R: int, x: float) -> np.ndarray:
    """e^{-x} I_n(x), n = 0..R, by the trapezoid rule on (1/π)∫₀^π e^{x(cos θ − 1)} cos nθ dθ
    (_scaled_besseli!, openbc.jl:43-56)."""
    M = 32 + math.ceil(12 * math.sqrt(x))
    n = np.arange(R + 1)
    b = np.zeros(R + 1)
    for m in range(M + 1):
        th = math.pi * m / M
        w = 0.5 if (m == 0 or m == M) else 1.0
        e = w * math.exp(x * (math.cos(th) - 1))
        b += e * np.cos(n * th)
    return b / M


def lattice_green(R: int) -> np.ndarray:
    """G[0:R, 0:R, 0:R], the Green function of −Δ_h on the unit simple-cubic lattice
    (lattice_green, openbc.jl:67-93): G(n) = ∫₀^∞ Π_i e^{−2t} I_{n_i}(2t) dt by 48-point
    Gauss–Legendre on [0, 0.5, 2, 8, 32, 128, 512] and the asymptotic tail beyond."""
    x, w = np.polynomial.legendre.leggauss(48)
    edges = (0.0, 0.5, 2.0, 8.0, 32.0, 128.0, 512.0)
    T = edges[-1]
    G = np.zeros((R + 1, R + 1, R + 1))
    for p in range(len(edges) - 1):
        a, c = edges[p], edges[p + 1]
        for q in range(len(x)):
            t = (a + c) / 2 + (c - a) / 2 * x[q]
            wt = (c - a) / 2 * w[q]
            b = _scaled_besseli(R, 2 * t)
            G += ((wt * b[:, None, None]) * b[None, :, None]) * b[None, None, :]
    n = np.arange(R + 1)
    c = ((4 * n[:, None, None] ** 2 - 1) + (4 * n[None, :, None] ** 2 - 1) + (4 * n[None, None, :] ** 2 - 1)) / 16
    G += (4 * math.pi) ** -1.5 * (2 / math.sqrt(T) - c * (2 / 3) / T ** 1.5)
    return np.asfortranarray(G)


def lgf_far(i, j, k) -> float:
    """1/(4πr) with the first correction of the cubic lattice (_lgf_far, openbc.jl:96-99)."""
    r2 = float(i * i + j * j + k * k)
    r = math.sqrt(r2)
    return (1 + (5 * (i ** 4 + j ** 4 + k ** 4) / r2 ** 2 - 3) / (8 * r2)) / (4 * math.pi * r)


@dataclass
class OpenData:
    """OpenBC (openbc.jl:109-140) as the step reads it."""
    every: int
    R: int
    lgf: np.ndarray        # lattice_green(R)
    fac: float             # dx³/κ²
    kappa2: float
    Jbar: np.ndarray       # (Ny+1, Nz+1)


def open_data(gr, kappa: float, bar: dict, every: int, R: int = 8) -> OpenData:
    if not (gr.dx == gr.dy == gr.dz):
        raise ValueError(":open needs dx = dy = dz")
    Ny, Nz = gr.Ny, gr.Nz
    k2 = kappa ** 2
    Jb = np.zeros((Ny + 1, Nz + 1), order="F")
    hz_x, hy_x = bar["hz_x"], bar["hy_x"]
    for k in range(2, Nz + 1):
        for j in range(2, Ny + 1):
            Jb[j - 1, k - 1] = k2 * ((hz_x[j - 1, k - 1] - hz_x[j - 2, k - 1]) / gr.dy -
                                     (hy_x[j - 1, k - 1] - hy_x[j - 1, k - 2]) / gr.dz)
    return OpenData(int(every), int(R), lattice_green(R), gr.dx ** 3 / k2, k2, Jb)


def lap1d(n: int, h: float, dirichlet: bool):
    """The 1D second difference on n nodes with spacing h, Dirichlet (the n unknowns between two
    fixed ends) or Neumann with mirror ghosts (2/h² to the only neighbour at each end); its
    eigenbasis V (columns), W = V⁻¹ and the eigenvalues (_lap1d, exchange.jl:148-166), eigenvalues
    in increasing order."""
    L = np.zeros((n, n))
    for a in range(n):
        L[a, a] = -2 / h ** 2
        if a > 0:
            L[a, a - 1] = 1 / h ** 2
        if a < n - 1:
            L[a, a + 1] = 1 / h ** 2
    if not dirichlet and n > 1:
        L[0, 1] = 2 / h ** 2
        L[n - 1, n - 2] = 2 / h ** 2
    e, V = np.linalg.eig(L)
    e, V = np.real(e), np.real(V)
    o = np.argsort(e, kind="stable")
    e, V = e[o], V[:, o]
    return np.asfortranarray(V), np.asfortranarray(np.linalg.inv(V)), e.copy()


@dataclass
class QPData:
    """QPExchange / QPExchangeN (exchange.jl:130-191, 425-475) for a uniform, not sculpted sample:
    what qp_potential! and qp_turn! read."""
    mode: int
    K: int                  # Helmholtz problems per step
    s: tuple                # shares of σn
    G: float                # two bands: χ_σ ν (the exchange conductance)
    lam: tuple              # λ0, or the eigenvalues of S⁻¹L
    P: np.ndarray           # N: μ = P m
    Pinv: np.ndarray        # N: m = Pinv R
    turn: tuple             # two bands: m_µ dt, m = (−s_π, s_σ)
    j0: int
    nj: int
    k0: int
    nk: int
    bases: tuple            # (Vx, Wx, ex, Vy, Wy, ey, Vz, Wz, ez)
    Gm: np.ndarray = None   # N: the conductances G_µν


def qp_exchange_data(md, gr, nu, split, conductor: bool, dt: float) -> QPData:
    """QPExchange(md, gr; ν, split, conductor) for a number ν (two bands) or QPExchangeN for a
    matrix of rates (exchange.jl:170-191, 450-475)."""
    N = md.N
    b = md.bands
    if conductor:
        jr, kr = (gr.y1, gr.y2 + 1), (gr.z1, gr.z2 + 1)
    else:
        jr, kr = (1, gr.Ny + 1), (1, gr.Nz + 1)
    nj, nk = jr[1] - jr[0] + 1, kr[1] - kr[0] + 1
    bases = lap1d(gr.Nx - 1, gr.dx, True) + lap1d(nj, gr.dy, False) + lap1d(nk, gr.dz, False)
    if np.ndim(nu) == 0:
        if N != 2:
            raise ValueError(f"qp_exchange: a number is for two bands; with {N} bands give the matrix of rates")
        nu = float(nu)
        if not nu >= 0:
            raise ValueError("qp_exchange: the rate must be >= 0")
        w = tuple(b[mu].stiffness() * b[mu].Gamma / b[mu].beta for mu in range(2)) if split is None else \
            tuple(float(x) for x in split)
        if len(w) != 2 or not all(x > 0 for x in w):
            raise ValueError("qp_split: two positive shares")
        sh = (w[0] / (w[0] + w[1]), w[1] / (w[0] + w[1]))
        chis = sh[0] * md.sigma_n * b[0].Gamma / b[0].stiffness()
        G = chis * nu
        lam0 = G * (1 / sh[0] + 1 / sh[1])
        m = (-sh[1], sh[0])
        return QPData(QP_TWO, 1, sh, G, (lam0,), np.zeros((N, N)), np.zeros((N, N)),
                      tuple(m[mu] * dt for mu in range(2)), jr[0] - 1, nj, kr[0] - 1, nk, bases)
    nu = np.asarray(nu, dtype=float)
    if nu.shape != (N, N):
        raise ValueError(f"qp_exchange: an {N}x{N} matrix of rates for {N} bands")
    if not (all(nu[mu, mu] == 0 for mu in range(N)) and (nu >= 0).all()):
        raise ValueError("qp_exchange: rates >= 0 with a zero diagonal")
    w = tuple(b[mu].stiffness() * b[mu].Gamma / b[mu].beta for mu in range(N)) if split is None else \
        tuple(float(x) for x in split)
    if len(w) != N or not all(x > 0 for x in w):
        raise ValueError(f"qp_split: {N} positive shares")
    tot = 0.0
    for x in w:
        tot += x
    sh = tuple(x / tot for x in w)
    chi = [sh[mu] * md.sigma_n * b[mu].Gamma / b[mu].stiffness() for mu in range(N)]
    G = np.array([[chi[mu] * nu[mu, n] for n in range(N)] for mu in range(N)])
    if not np.allclose(G, G.T, rtol=1e-10, atol=1e-14):
        raise ValueError("qp_exchange: the rates must satisfy detailed balance, χ_µν_µν = χ_νν_νµ")
    G = (G + G.T) / 2
    L = -G.copy()
    for mu in range(N):
        acc = 0.0
        for n in range(N):
            if n != mu:
                acc += G[mu, n]
        L[mu, mu] = acc
    rs = np.sqrt(np.array(sh))
    lam, Q = np.linalg.eigh(L / np.outer(rs, rs))
    P = Q / rs[:, None]
    Pinv = Q.T * rs[None, :]
    return QPData(QP_N, N, sh, 0.0, tuple(float(max(x, 0.0)) for x in lam), P, Pinv, tuple(0.0 for _ in range(N)),
                  jr[0] - 1, nj, kr[0] - 1, nk, bases, G)


@dataclass(frozen=True)
class NoiseSpec:
    """Noise(T̃, seed; field) (noise.jl:32-41) and where the normals come from: `host` False draws
    them on the device (Philox4x32-10 keyed by the seed), True takes the caller's per step
    (tdgl_noise_draws)."""
    T: float
    seed: int
    field: bool = False
    host: bool = False


@dataclass
class Stepper:
    """The coefficients of Stepper{N} (scheme.jl:167-198) that the step reads."""
    model: Model
    grid: Grid
    masks: Masks
    dt: float
    nsub: tuple
    rkl: tuple
    bc: tuple
    bcs: tuple
    w: tuple
    wa: tuple
    fc: FieldCoef
    field_mode: int
    D: Optional[tuple]
    bar: Optional[dict]
    extrap: bool
    rotframe: bool
    hard: bool
    limit: str
    dt_scale: float
    rkl_margin: float
    rkc: Optional[tuple]
    divisor: tuple
    pcg_tol: float = 1e-8
    pcg_maxit: int = 500
    pcg_fast: bool = True
    pcg_f32: bool = True
    pcg_extrap: bool = True
    selffield_bc: str = "bar"
    qp: Optional[QPData] = None
    open: Optional[OpenData] = None
    noise: Optional[NoiseSpec] = None
    sk: tuple = ()            # _skrock(st, µ) per band: the phase kick inside the first RKC stage
    # the PCG's precision (an extension): FP64 refinement around Float32 inner solves, each to
    # pcg_inner_tol of its starting residual; pcg_rounds rounds per captured solve
    pcg_refine: bool = False
    pcg_inner_tol: float = 1e-4
    pcg_rounds: int = 3

    @property
    def N(self) -> int:
        return self.model.N

    def selffield(self, J: float) -> float:
        """HJ = J d/(2κ²) (step.jl:38)."""
        return J * self.grid.slab_thickness() / (2 * self.model.kappa ** 2)


def make_stepper(md: Model, gr: Grid, mk: Masks, *, implicit=True, substeps=None, dt_scale=0.5,
                 extrap=True, hard=False, sigma_vac=None, pcg_tol=1e-8, selffield_bc="bar",
                 rkl="rkc", rotframe=None, rkl_margin=None, rkc_damping=RKC_EPS, pcg_fast=True,
                 pcg_f32=True, pcg_extrap=True, dt=None, pcg_maxit=500, qp_exchange=None, qp_split=None,
                 pcg_refine=False, pcg_inner_tol=1e-4, pcg_rounds=3,
                 open_every=1.0, noise=None, noise_seed=1, noise_field=False, noise_draws="device",
                 **unsupported) -> Stepper:
    """Stepper(md, gr, mk; kw...) (scheme.jl:250-434) for the options the HIP stepper runs.

    `dt` (an extension): a fixed time step in place of the derived one; the substep and stage
    counts follow from it by the same rules (Euler substeps so that each stays within 0.99 of
    its band's bound, stages from r = rkl_margin·dt/(0.99·bbk)).

    `noise = T̃`, `noise_seed`, `noise_field` as the reference's (noise.jl); `noise_draws`
    "device" (Philox on the device, keyed by the seed) or "host" (the caller's draws per step).

    `pcg_refine` (an extension of the PCG solver, with pcg_f32 and the insulating padding): the FP64
    operator, residual and solution refine around Float32 inner PCG solves, each to `pcg_inner_tol`
    of its starting residual, until ||b − A x|| <= pcg_tol ||b||; `pcg_rounds` rounds a captured
    solve holds."""
    for k, v in unsupported.items():
        if v is not None and v is not False:
            raise NotImplementedError(f"Stepper option {k!r} is not run by the HIP stepper")
    if pcg_refine and not (pcg_f32 and sigma_vac is not None and float(sigma_vac) != md.sigma_n):
        raise ValueError("pcg_refine needs the PCG (sigma_vac) with its Float32 preconditioner (pcg_f32)")
    if not (0 < float(pcg_inner_tol) < 1) or not (1 <= int(pcg_rounds) <= 16):
        raise ValueError("pcg_inner_tol must be in (0, 1) and pcg_rounds in 1..16")
    if mk.vac is not None and mk.vac.any():
        raise NotImplementedError("a sculpted sample (vacuum = true in the shape file) is not run by the HIP stepper")
    if mk.mat is not None and (mk.mat != 1).any():
        raise NotImplementedError("materials per region (material = \"name\" in the shape file) are not run by the "
                                  "HIP stepper")
    N = md.N
    if noise is not None and not float(noise) >= 0:
        raise ValueError("noise: T̃ must be >= 0")
    if noise_field and noise is None:
        raise ValueError("noise_field needs noise = T̃ (the same bath for Ψ and A)")
    if noise_draws not in ("device", "host"):
        raise ValueError("noise_draws must be 'device' or 'host'")
    substeps = tuple(1 if mu == 0 else 0 for mu in range(N)) if substeps is None else tuple(substeps)
    if isinstance(rkl, bool):
        rkc = False
    else:
        s = str(rkl).lower()
        if s not in ("rkc", "rkl"):
            raise ValueError(f"rkl must be True, False, 'rkl' or 'rkc', not {rkl!r}")
        rkc = s == "rkc"
        rkl = True
    if not (0 < dt_scale <= 1):
        raise ValueError(f"dt_scale must be in (0, 1], not {dt_scale}")
    kwt = any(b.gkwt > 0 for b in md.bands)
    rotframe = kwt if rotframe is None else bool(rotframe)
    rkl_margin = float(rkl_margin if rkl_margin is not None else (1.1 if kwt else 1.0))
    if rkl_margin < 1:
        raise ValueError("rkl_margin must be >= 1")
    if rotframe and not extrap:
        raise ValueError("rotframe needs extrap = true (it turns copies of the links)")
    if any(s < 0 for s in substeps):
        raise ValueError("substeps must be >= 0")
    if implicit and not (gr.dx == gr.dy == gr.dz):
        raise ValueError("the implicit curl-curl needs dx = dy = dz (separable solve)")
    l = gr.dl()
    bands = md.bands
    bb0, bbk, lamL = _band_bounds(md, gr, l)
    bb = bb0 if rkl else bbk
    pad = sigma_vac is not None and float(sigma_vac) != md.sigma_n
    if pad and not (0 < sigma_vac <= md.sigma_n):
        raise ValueError("σvac must be in (0, σn]")
    bfield = INF if implicit else (float(sigma_vac) if pad else md.sigma_n) * l / (4 * md.kappa ** 2)
    bsc = 2 * md.sigma_n / lamL if lamL != 0 else math.copysign(INF, lamL)  # IEEE division, as Julia's
    ref = min(bfield, bsc, min(INF if substeps[mu] == 0 else bb[mu] * max(substeps[mu], 1) for mu in range(N)))
    nsub = [max(1, math.ceil(dt_scale * ref / bb[mu] - 1e-12)) if substeps[mu] == 0 else substeps[mu]
            for mu in range(N)]
    bbn = [bb[mu] if (nsub[mu] == 1 and substeps[mu] == 1) else bb[mu] * (nsub[mu] / dt_scale) for mu in range(N)]
    cands = bbn + [bfield, bsc]
    names = [f"band {mu + 1}" for mu in range(N)] + ["field (explicit curl-curl)", "explicit supercurrent"]
    dtv = 0.99 * min(cands)
    if dt_scale != 1:
        dtv *= dt_scale
    limit = names[int(np.argmin(cands))]
    if dt is not None:
        dtv = float(dt)
        limit = "fixed dt"
        nsub = [max(1, math.ceil(dtv / (0.99 * bb[mu]) - 1e-12)) if substeps[mu] == 0 else substeps[mu]
                for mu in range(N)]
    rx = dtv / (gr.dx * gr.dx)
    ry = dtv / (gr.dy * gr.dy)
    rz = dtv / (gr.dz * gr.dz)
    bc = tuple(_bandcoef(md, mu, dtv, rx, ry, rz, bands[mu].ma) for mu in range(N))
    stg = []
    for mu in range(N):
        if not rkl:
            stg.append(0)
            continue
        r = rkl_margin * dtv / (0.99 * bbk[mu])
        if r <= 1 + 1e-12:
            stg.append(0)
            continue
        s = 1
        while s * (s + 1) // 2 < r - 1e-12:
            s += 1
        if rkc:
            q = 2
            while _rkc_ratio(q, float(rkc_damping)) < r - 1e-12:
                q += 1
            # with noise a KWT band takes RKC even at the same count (SK-ROCK)
            if q < s or (noise is not None and bands[mu].gkwt > 0):
                s = -q
        stg.append(s)
    usec = [s < 0 for s in stg]
    skb = tuple(noise is not None and rkc and usec[mu] and bands[mu].gkwt > 0 for mu in range(N))
    stg = [abs(s) for s in stg]
    nsub = [1 if stg[mu] > 0 else nsub[mu] for mu in range(N)]
    if noise is not None and any(bands[mu].gkwt > 0 and not skb[mu] and stg[mu] > 0 for mu in range(N)):
        raise ValueError("noise with KWT on RKL stages: the equilibrium comes out biased (10-25 % low at "
                         "T̃ = 1e-2). Use rkl = 'rkc' (the default: the noise inside the stages, SK-ROCK) or "
                         "rkl = False (Euler substeps)")
    if noise is not None and any(bands[mu].gkwt > 0 and not skb[mu] and nsub[mu] > 1 for mu in range(N)):
        warnings.warn("noise with KWT on Euler substeps: the kick after the step over-heats the stiff phase "
                      "modes (it converges as dt resolves them); rkl = 'rkc' puts the noise inside the stages")
    rkcc = tuple(RKCCoef.make(stg[mu], float(rkc_damping)) if usec[mu] else None for mu in range(N)) if rkc else None

    def divisor(mu):
        if stg[mu] > 0:
            if rkcc is not None and rkcc[mu] is not None:
                return rkcc[mu].q
            return stg[mu] * (stg[mu] + 1) // 2
        return nsub[mu]

    bcs = tuple(bc[mu].divided(divisor(mu)) if stg[mu] > 0 else (bc[mu] if nsub[mu] == 1 else bc[mu].divided(nsub[mu]))
                for mu in range(N))
    w = tuple(bands[mu].stiffness() for mu in range(N))
    wa = tuple(tuple(bands[mu].stiffness(i) for mu in range(N)) for i in range(3))
    dts = dtv / md.sigma_n
    ks2 = md.kappa ** 2
    fc = FieldCoef(dts, (dts * ks2) * (gr.dx / gr.dy), (dts * ks2) * (gr.dx / gr.dz), (dts * ks2) * (gr.dy / gr.dx),
                   (dts * ks2) * (gr.dy / gr.dz), (dts * ks2) * (gr.dz / gr.dx), (dts * ks2) * (gr.dz / gr.dy),
                   dts * ks2 / (gr.dx * gr.dx))
    D = link_conductivity(gr, float(sigma_vac), md.sigma_n) if pad else None
    if not implicit:
        mode = FIELD_EXPLICIT_SCALED if pad else FIELD_EXPLICIT
    else:
        mode = FIELD_PCG if pad else FIELD_SEPARABLE
    sbc = str(selffield_bc).lower().lstrip(":")
    if sbc not in ("slab", "bar", "open"):
        raise ValueError("selffield_bc must be slab, bar or open")
    bar = bar_field(gr) if sbc in ("bar", "open") else None
    ob = open_data(gr, md.kappa, bar, max(1, round(open_every / dtv))) if sbc == "open" else None
    qp = None if qp_exchange is None else qp_exchange_data(md, gr, qp_exchange, qp_split, pad, dtv)
    nz = None if noise is None else NoiseSpec(float(noise), int(noise_seed) & 0xFFFFFFFFFFFFFFFF, bool(noise_field),
                                              noise_draws == "host")
    return Stepper(md, gr, mk, dtv, tuple(nsub), tuple(stg), bc, bcs, w, wa, fc, mode, D, bar, bool(extrap), rotframe,
                   bool(hard), limit, float(dt_scale), rkl_margin, rkcc, tuple(float(divisor(mu)) for mu in range(N)),
                   float(pcg_tol), int(pcg_maxit), bool(pcg_fast), bool(pcg_f32), bool(pcg_extrap), sbc, qp, ob,
                   nz, skb, bool(pcg_refine), float(pcg_inner_tol), int(pcg_rounds))
