"""CPU restatement of a ladder's per-cycle measurements (gtdgl/src/protocol.jl:148-189, 416-464;
observables.jl:102-114), in the reference's loop order with sequential sums (sumlin).

TEST INFRASTRUCTURE ONLY: the checker of the device meter (tdgl_meter_cycle) and the measuring
half of an oracle backend for juliatdgl_amd.meter.  Arrays are the host State's (Fortran order,
0-based here where the reference is 1-based).
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 2 * math.pi


def _seq(v) -> float:
    """sumlin (observables.jl:8-14): a sequential sum."""
    v = np.asarray(v, dtype=float).ravel(order="F")
    return float(np.cumsum(v)[-1]) if v.size else 0.0


def _lines(gr):
    """(j, k) of the lines, 1-based, k outer and j inner (the reference's loop order)."""
    return [(j, k) for k in range(gr.z1, gr.z2 + 2) for j in range(gr.y1, gr.y2 + 2)]


def plane_sum(s, gr, plano: int, M) -> float:
    """_plane_sum (protocol.jl:148-155)."""
    tot = 0.0
    Mf = np.asarray(M).ravel(order="F")
    for q, (j, k) in enumerate(_lines(gr)):
        if Mf[q]:
            tot += s.phx[plano - 1, j - 1, k - 1]
    return tot


def theta_A(s, gr, vi: int, vo: int, line_ok) -> float:
    """_theta_A (protocol.jl:157-166): one running total over the lines and their links."""
    tot, n = 0.0, 0
    ok = np.asarray(line_ok).ravel(order="F")
    for q, (j, k) in enumerate(_lines(gr)):
        if not ok[q]:
            continue
        for ii in range(vi, vo):
            tot += s.phx[ii - 1, j - 1, k - 1]
        n += 1
    return tot / n if n > 0 else 0.0


def path_phase(G, s, gr, vi: int, vo: int) -> np.ndarray:
    """_path_phase! (protocol.jl:168-179): D[jj, kk] = Σ_{ii=vi}^{vo-1} angle((conj(G_ii) V_ii) G_ii+1)."""
    nl1, nl2 = gr.y2 - gr.y1 + 2, gr.z2 - gr.z1 + 2
    D = np.zeros((nl1, nl2), order="F")
    js, ks = slice(gr.y1 - 1, gr.y2 + 1), slice(gr.z1 - 1, gr.z2 + 1)
    a = np.conj(G[vi - 1:vo - 1, js, ks]) * s.Vx[vi - 1:vo - 1, js, ks]
    a = a * G[vi:vo, js, ks]
    ang = np.angle(a)
    D[:, :] = np.cumsum(ang, axis=0)[-1]
    return D


def unwrap(T, Dl, Dprev) -> None:
    """_unwrap! (protocol.jl:181-189), in place; round half away from zero (oround_f)."""
    d = Dl - Dprev
    r = d / TWO_PI
    d = d - TWO_PI * np.where(r >= 0, np.floor(r + 0.5), np.ceil(r - 0.5))
    T += d
    Dprev[...] = Dl


def midplane_current(s, st):
    """midplane_current (observables.jl:102-114)."""
    gr = st.grid
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    ks2 = st.model.kappa ** 2
    ii = int(math.floor(Nx / 2 + 0.5))  # oround
    It = Is = 0.0
    for kk in range(2, Nz + 1):
        for jj in range(2, Ny + 1):
            jx = ks2 * ((s.hz[ii - 1, jj - 1, kk - 1] - s.hz[ii - 1, jj - 2, kk - 1]) / gr.dy
                        - (s.hy[ii - 1, jj - 1, kk - 1] - s.hy[ii - 1, jj - 1, kk - 2]) / gr.dz)
            It += jx
            if gr.y1 <= jj <= gr.y2 + 1 and gr.z1 <= kk <= gr.z2 + 1:
                Is += jx
    return It * gr.dy * gr.dz, Is * gr.dy * gr.dz


def _jmax(a, b):
    return a + b if (a != a or b != b) else max(a, b)


class OracleMeter:
    """The per-cycle measurements of `ladder` (protocol.jl:371-372, 416-464) on host states: the
    record tdgl_meter_cycle returns."""

    def __init__(self, st, vm, s):
        self.st, self.vm = st, vm
        gr = st.grid
        sl = (slice(gr.x1 - 1, gr.x2 + 1), slice(gr.y1 - 1, gr.y2 + 1), slice(gr.z1 - 1, gr.z2 + 1))
        self._sl = sl
        self.cooper = [np.abs(g[sl]) for g in s.G]
        self.field_old = s.hz.copy()
        sh = (gr.y2 - gr.y1 + 2, gr.z2 - gr.z1 + 2)
        self.Dprev = [np.zeros(sh, order="F") for _ in s.G]
        self.T = [np.zeros(sh, order="F") for _ in s.G]
        self.dmax, self.nbig = 0.0, 0

    def cycle(self, s, first: bool) -> dict:
        st, vm, gr = self.st, self.vm, self.st.grid
        N = len(s.G)
        nin, nou = int(vm.Min.sum()), int(vm.Mou.sum())
        ok = np.asarray(vm.line_ok)
        f1 = np.asarray(vm.line_f1)
        nlin, nlin_f1 = int(ok.sum()), int(f1.sum())
        rec = {"V": plane_sum(s, gr, vm.vo, vm.Mou) / nou - plane_sum(s, gr, vm.vi, vm.Min) / nin,
               "TA": theta_A(s, gr, vm.vi, vm.vo, ok)}
        D = [path_phase(s.G[mu], s, gr, vm.vi, vm.vo) for mu in range(N)]
        if first:
            for mu in range(N):
                self.Dprev[mu][...] = D[mu]
                self.T[mu][...] = 0.0
            self.dmax, self.nbig = 0.0, 0
            Tf1, Tl, tabs = 0.0, [0.0] * N, [0.0] * N
        else:
            d = D[0] - self.Dprev[0]
            r = d / TWO_PI
            d = d - TWO_PI * np.where(r >= 0, np.floor(r + 0.5), np.ceil(r - 0.5))
            for x in np.abs(d[ok]):
                self.dmax = _jmax(self.dmax, float(x))
                self.nbig += int(x > 2.0)
            for mu in range(N):
                unwrap(self.T[mu], D[mu], self.Dprev[mu])
            okf, f1f = ok.ravel(order="F"), f1.ravel(order="F")
            Tf1 = _seq(self.T[0].ravel(order="F")[f1f]) / nlin_f1 if nlin_f1 > 0 else 0.0
            Tl = [_seq(self.T[mu].ravel(order="F")[okf]) / nlin for mu in range(N)]
            tabs = [_seq(np.abs(self.T[mu].ravel(order="F")[okf])) for mu in range(N)]
        rec.update(Tf1=Tf1, T=Tl, Tabs=tabs, dmax=self.dmax, nbig=self.nbig)
        rec["It"], rec["Is"] = midplane_current(s, st)
        dif, amax = [0.0] * N, 0.0
        for mu in range(N):
            a = np.abs(s.F[mu][self._sl])
            dif[mu] = float(np.max(np.abs(a - self.cooper[mu]))) if not np.isnan(a).any() else math.nan
            m = float(np.max(a)) if not np.isnan(a).any() else math.nan
            amax = _jmax(amax, m)
            self.cooper[mu] = a
        rec["dif"], rec["amax"] = dif, amax
        rec["diffh"] = float(np.max(np.abs(s.hz - self.field_old)))
        self.field_old = s.hz.copy()
        return rec


class OracleBackend:
    """An oracle stepper with the meter on its downloaded states: the backend of
    juliatdgl_amd.meter for the CPU side of the checks."""

    def __init__(self, oracle, st):
        self.o, self.st = oracle, st
        self.meter = None

    def step(self, n: int, H: float, HJ: float) -> None:
        self.o.step(n, H, HJ)

    def noise_draws(self, flat) -> None:
        self.o.noise_draws(flat)

    def meter_begin(self, vm) -> None:
        self.meter = OracleMeter(self.st, vm, self.o.get_state())

    def meter_cycle(self, first: bool) -> dict:
        return self.meter.cycle(self.o.get_state(), first)
