"""ctypes mirror of include/tdglhip.h: the description, state and PCG structs, and the packing of
a host Stepper / State into them.  Shared by the HIP binding (juliatdgl_amd.hip) and by any other
implementation of the same C ABI."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from .scheme import FIELD_PCG, FIELD_EXPLICIT_SCALED, Stepper
from .state import FIELDS, State

ABI_VERSION = 5
MAXB = 4
MAXS = 64
dp = ct.POINTER(ct.c_double)


class BandCoefC(ct.Structure):
    _fields_ = [("kx", ct.c_double), ("ky", ct.c_double), ("kz", ct.c_double), ("c", ct.c_double),
                ("eps", ct.c_double), ("beta", ct.c_double), ("gamma2", ct.c_double),
                ("cpl", ct.c_double * MAXB)]


class DescC(ct.Structure):
    _fields_ = [("abi_version", ct.c_int32), ("nbands", ct.c_int32),
                ("Nx", ct.c_int32), ("Ny", ct.c_int32), ("Nz", ct.c_int32),
                ("x1", ct.c_int32), ("x2", ct.c_int32), ("y1", ct.c_int32), ("y2", ct.c_int32),
                ("z1", ct.c_int32), ("z2", ct.c_int32),
                ("dx", ct.c_double), ("dy", ct.c_double), ("dz", ct.c_double), ("dt", ct.c_double),
                ("bc", BandCoefC * MAXB), ("bcs", BandCoefC * MAXB),
                ("nsub", ct.c_int32 * MAXB), ("stages", ct.c_int32 * MAXB), ("rkc", ct.c_int32 * MAXB),
                ("divisor", ct.c_double * MAXB),
                ("rkc_mu", (ct.c_double * MAXS) * MAXB), ("rkc_nu", (ct.c_double * MAXS) * MAXB),
                ("rkc_c", (ct.c_double * MAXS) * MAXB),
                ("wa", (ct.c_double * MAXB) * 3),
                ("d1", ct.c_double), ("sxy", ct.c_double), ("sxz", ct.c_double), ("syx", ct.c_double),
                ("syz", ct.c_double), ("szx", ct.c_double), ("szy", ct.c_double), ("a", ct.c_double),
                ("field_mode", ct.c_int32), ("selffield_bc", ct.c_int32),
                ("extrap", ct.c_int32), ("rotframe", ct.c_int32), ("hard", ct.c_int32),
                ("pcg_tol", ct.c_double), ("pcg_maxit", ct.c_int32), ("pcg_fast", ct.c_int32),
                ("pcg_f32", ct.c_int32), ("pcg_extrap", ct.c_int32),
                ("f", dp), ("taux", dp), ("tauy", dp), ("tauz", dp),
                ("Dx", dp), ("Dy", dp), ("Dz", dp),
                ("bar_hy_z", dp), ("bar_hy_y", dp), ("bar_hz_z", dp), ("bar_hz_y", dp),
                ("bar_hz_x", dp), ("bar_hy_x", dp),
                ("qp_mode", ct.c_int32), ("qp_nmodes", ct.c_int32),
                ("qp_j0", ct.c_int32), ("qp_nj", ct.c_int32), ("qp_k0", ct.c_int32), ("qp_nk", ct.c_int32),
                ("qp_sigma", ct.c_double), ("qp_s", ct.c_double * MAXB), ("qp_turn", ct.c_double * MAXB),
                ("qp_lambda", ct.c_double * MAXB), ("qp_P", (ct.c_double * MAXB) * MAXB),
                ("qp_Pinv", (ct.c_double * MAXB) * MAXB), ("qp_gamma", (ct.c_double * MAXB) * MAXB),
                ("qp_Vx", dp), ("qp_Wx", dp), ("qp_ex", dp), ("qp_Vy", dp), ("qp_Wy", dp), ("qp_ey", dp),
                ("qp_Vz", dp), ("qp_Wz", dp), ("qp_ez", dp),
                ("open_every", ct.c_int32), ("open_R", ct.c_int32), ("open_fac", ct.c_double),
                ("open_kappa2", ct.c_double), ("open_lgf", dp), ("open_Jbar", dp),
                ("noise", ct.c_int32), ("noise_field", ct.c_int32), ("noise_host", ct.c_int32),
                ("noise_sk", ct.c_int32 * MAXB), ("noise_T", ct.c_double), ("noise_seed", ct.c_uint64),
                ("noise_sk_w", (ct.c_double * 2) * MAXB), ("noise_Gamma", ct.c_double * MAXB),
                ("noise_sigma_n", ct.c_double),
                ("pcg_refine", ct.c_int32), ("pcg_rounds", ct.c_int32), ("pcg_inner_tol", ct.c_double)]


class StateC(ct.Structure):
    _fields_ = [("G", dp * MAXB), ("F", dp * MAXB)] + [(k, dp) for k in FIELDS]


class PcgC(ct.Structure):
    _fields_ = [("x", dp * 3), ("xprev", dp * 3), ("warm", ct.c_int32), ("have2", ct.c_int32),
                ("iters", ct.c_int32)]


class OpenC(ct.Structure):
    _fields_ = [("count", ct.c_int32), ("hold", dp * 3), ("hnew", dp * 3)]


class QpC(ct.Structure):
    _fields_ = [("pot", dp * 4)]


class StatsC(ct.Structure):
    _fields_ = [("steps", ct.c_int64), ("calls", ct.c_int64), ("pcg_iters", ct.c_int64),
                ("last_iters", ct.c_int32), ("device_ms", ct.c_double), ("device_bytes", ct.c_int64),
                ("graph_cycles", ct.c_int64), ("rollbacks", ct.c_int64), ("pcg_budget", ct.c_int32),
                ("captures", ct.c_int64), ("folded", ct.c_int32)]


NKERNELS = 15
TUNE_MAX = 16
TUNE_KINDS = {1: "box_kc", 2: "plane_mt", 3: "zmodes_waves", 4: "fold", 5: "zcol", 6: "qparts"}


class TuneCandC(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("value", ct.c_int32), ("us", ct.c_double)]


class TuningC(ct.Structure):
    _fields_ = [("box_kc", ct.c_int32), ("plane_mt", ct.c_int32), ("zmodes_waves", ct.c_int32),
                ("ncand", ct.c_int32), ("cand", TuneCandC * TUNE_MAX), ("fold", ct.c_int32), ("zcol", ct.c_int32),
                ("qparts", ct.c_int32)]


class MeterC(ct.Structure):
    _fields_ = [("vi", ct.c_int32), ("vo", ct.c_int32), ("kappa2", ct.c_double),
                ("Min", ct.POINTER(ct.c_uint8)), ("Mou", ct.POINTER(ct.c_uint8)),
                ("line_ok", ct.POINTER(ct.c_uint8)), ("line_f1", ct.POINTER(ct.c_uint8))]


class CycleRecordC(ct.Structure):
    _fields_ = [("V", ct.c_double), ("TA", ct.c_double), ("Tf1", ct.c_double), ("T", ct.c_double * MAXB),
                ("Tabs", ct.c_double * MAXB), ("dmax", ct.c_double), ("nbig", ct.c_int32), ("pad", ct.c_int32), ("It", ct.c_double),
                ("Is", ct.c_double), ("dif", ct.c_double * MAXB), ("amax", ct.c_double), ("diffh", ct.c_double)]


class ProfileC(ct.Structure):
    _fields_ = [("launches", ct.c_int64 * NKERNELS), ("ms", ct.c_double * NKERNELS)]


def _ptr(a: np.ndarray):
    if not (a.flags.f_contiguous and a.dtype in (np.float64, np.complex128)):
        raise ValueError("arrays must be Fortran-contiguous float64/complex128")
    return a.ctypes.data_as(dp)


class PackedDesc:
    """A DescC with the numpy arrays it points to kept alive."""

    def __init__(self, st: Stepper):
        d = DescC()
        gr = st.grid
        N = st.N
        if N > MAXB:
            raise ValueError(f"at most {MAXB} bands")
        d.abi_version = ABI_VERSION
        d.nbands = N
        d.Nx, d.Ny, d.Nz = gr.Nx, gr.Ny, gr.Nz
        d.x1, d.x2, d.y1, d.y2, d.z1, d.z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
        d.dx, d.dy, d.dz = gr.dx, gr.dy, gr.dz
        d.dt = st.dt
        for mu in range(N):
            for dst, src in ((d.bc[mu], st.bc[mu]), (d.bcs[mu], st.bcs[mu])):
                dst.kx, dst.ky, dst.kz, dst.c = src.kx, src.ky, src.kz, src.c
                dst.eps, dst.beta, dst.gamma2 = src.eps, src.beta, src.g2
                for nu in range(N):
                    dst.cpl[nu] = src.cpl[nu] if nu != mu else 0.0
            d.nsub[mu] = st.nsub[mu]
            d.stages[mu] = st.rkl[mu]
            d.divisor[mu] = st.divisor[mu]
            kc = st.rkc[mu] if st.rkc is not None else None
            d.rkc[mu] = 1 if kc is not None else 0
            if kc is not None:
                if len(kc.mu) > MAXS:
                    raise ValueError(f"at most {MAXS} stages")
                for j in range(len(kc.mu)):
                    d.rkc_mu[mu][j] = kc.mu[j]
                    d.rkc_nu[mu][j] = kc.nu[j]
                    d.rkc_c[mu][j] = kc.c[j]
            elif st.rkl[mu] > MAXS:
                raise ValueError(f"at most {MAXS} stages")
            for ax in range(3):
                d.wa[ax][mu] = st.wa[ax][mu]
        fc = st.fc
        d.d1, d.sxy, d.sxz, d.syx, d.syz, d.szx, d.szy, d.a = (fc.d1, fc.sxy, fc.sxz, fc.syx, fc.syz, fc.szx,
                                                               fc.szy, fc.a)
        d.field_mode = st.field_mode
        d.selffield_bc = {"slab": 0, "bar": 1, "open": 2}[st.selffield_bc]
        d.extrap, d.rotframe, d.hard = int(st.extrap), int(st.rotframe), int(st.hard)
        d.pcg_tol, d.pcg_maxit = st.pcg_tol, st.pcg_maxit
        d.pcg_fast, d.pcg_f32, d.pcg_extrap = int(st.pcg_fast), int(st.pcg_f32), int(st.pcg_extrap)
        keep = []

        def hold(a):
            a = np.asfortranarray(a, dtype=np.float64)
            keep.append(a)
            return _ptr(a)

        mk = st.masks
        d.f, d.taux, d.tauy, d.tauz = hold(mk.f), hold(mk.tx), hold(mk.ty), hold(mk.tz)
        if st.field_mode in (FIELD_PCG, FIELD_EXPLICIT_SCALED):
            d.Dx, d.Dy, d.Dz = (hold(x) for x in st.D)
        if st.bar is not None:
            for k in ("hy_z", "hy_y", "hz_z", "hz_y", "hz_x", "hy_x"):
                setattr(d, "bar_" + k, hold(st.bar[k]))
        q = st.qp
        if q is not None:
            d.qp_mode, d.qp_nmodes = q.mode, q.K
            d.qp_j0, d.qp_nj, d.qp_k0, d.qp_nk = q.j0, q.nj, q.k0, q.nk
            d.qp_sigma = st.model.sigma_n
            for mu in range(N):
                d.qp_s[mu] = q.s[mu]
                d.qp_turn[mu] = q.turn[mu]
                for nu in range(N):
                    d.qp_P[mu][nu] = q.P[mu, nu]
                    d.qp_Pinv[mu][nu] = q.Pinv[mu, nu]
                    d.qp_gamma[mu][nu] = st.model.gamma[mu][nu]
            for k in range(q.K):
                d.qp_lambda[k] = q.lam[k]
            for name, arr in zip(("Vx", "Wx", "ex", "Vy", "Wy", "ey", "Vz", "Wz", "ez"), q.bases):
                setattr(d, "qp_" + name, hold(arr))
        ob = st.open
        if ob is not None:
            d.open_every, d.open_R, d.open_fac, d.open_kappa2 = ob.every, ob.R, ob.fac, ob.kappa2
            d.open_lgf, d.open_Jbar = hold(ob.lgf), hold(ob.Jbar)
        nz = st.noise
        if nz is not None:
            d.noise, d.noise_T, d.noise_seed = 1, nz.T, nz.seed
            d.noise_field, d.noise_host = int(nz.field), int(nz.host)
            for mu in range(N):
                d.noise_sk[mu] = int(st.sk[mu])
                kc = st.rkc[mu] if st.rkc is not None else None
                if kc is not None:
                    d.noise_sk_w[mu][0], d.noise_sk_w[mu][1] = kc.sk
                d.noise_Gamma[mu] = st.model.bands[mu].Gamma
            d.noise_sigma_n = st.model.sigma_n
        d.pcg_refine, d.pcg_rounds, d.pcg_inner_tol = int(st.pcg_refine), int(st.pcg_rounds), float(st.pcg_inner_tol)
        self.c = d
        self._keep = keep


def state_struct(s: State) -> StateC:
    sc = StateC()
    for mu in range(s.N):
        sc.G[mu] = _ptr(s.G[mu])
        sc.F[mu] = _ptr(s.F[mu])
    for k in FIELDS:
        setattr(sc, k, _ptr(getattr(s, k)))
    return sc


def pcg_ranges(gr):
    """Shapes of the PCG range arrays rx, ry, rz (implicit.jl:52)."""
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    return ((Nx - 2, Ny - 1, Nz - 1), (Nx - 1, Ny - 2, Nz - 1), (Nx - 1, Ny - 1, Nz - 2))


class PcgState:
    """Host copy of the PCG history (x, xprev, warm, have2, iters)."""

    def __init__(self, gr):
        sh = pcg_ranges(gr)
        self.x = [np.zeros(s, order="F") for s in sh]
        self.xprev = [np.zeros(s, order="F") for s in sh]
        self.warm = False
        self.have2 = False
        self.iters = 0

    def struct(self) -> PcgC:
        p = PcgC()
        for c in range(3):
            p.x[c] = _ptr(self.x[c])
            p.xprev[c] = _ptr(self.xprev[c])
        p.warm, p.have2, p.iters = int(self.warm), int(self.have2), int(self.iters)
        return p

    def load(self, p: PcgC):
        self.warm, self.have2, self.iters = bool(p.warm), bool(p.have2), int(p.iters)
