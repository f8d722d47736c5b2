"""A ladder's per-cycle measurements and verdicts on records instead of whole states.

The reference measures every ladder cycle on the host from the whole state (gtdgl/src/protocol.jl:
416-464): with its GPU backend the state goes back to the host every cycle for that.  Here the
device keeps the measurement state (the path phases of the last cycle, the unwrapped phases per
line, |Ψ| and hz of the last cycle) and returns a record of a few numbers per cycle
(tdgl_meter_cycle); this module holds what the host does with the records:

* `Voltmeter`: the voltmeter planes, masks and lines of `ladder` (protocol.jl:341-352) and of
  `crossing_times` (:271-279);
* the verdict helpers `crossings_now`, `late_crossings`, `silent_now` (:191-210) and `pendiente`
  (observables.jl:23-26);
* `crossing_times` (:259-296): the times at which the line-averaged band-1 phase across the
  voltmeter passes a further multiple of 2π;
* `Rung`: the cycle loop of one ladder point with its 'F' / 'V' closing rules (:401-489, no heat);
* `ladder` / `run_ladders`: the points of both polarities from the Meissner seed with what the
  reference records at each close (:493-545: the window slopes, the static voltage, the slips, the
  magnetization, the 'N' verdict) and the onset brackets (`bracket_Ic`, observables.jl:132-151).

Everything takes a `backend` with `step(n, H, HJ)`, `meter_begin(voltmeter)` and
`meter_cycle(first) -> record` (juliatdgl_amd.hip.HipStepper is one).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TWO_PI = 2 * math.pi


@dataclass
class Voltmeter:
    """The x-link planes vi, vo (1-based) and the masks over the lines (j, k) of the block,
    arrays of shape (y2 - y1 + 2, z2 - z1 + 2)."""
    vi: int
    vo: int
    kappa2: float
    Min: np.ndarray
    Mou: np.ndarray
    line_ok: np.ndarray
    line_f1: np.ndarray

    @classmethod
    def of_ladder(cls, st, delta_v: int) -> "Voltmeter":
        """ladder's voltmeter (protocol.jl:341-352): live links on the planes, complete lines,
        and the complete lines condensed throughout."""
        gr, mk = st.grid, st.masks
        vi, vo = gr.x1 + delta_v, gr.x2 - delta_v
        tx, f = mk.tx, mk.f
        js, ks = slice(gr.y1 - 1, gr.y2 + 1), slice(gr.z1 - 1, gr.z2 + 1)
        Min = tx[vi - 1, js, ks] > 0.5
        Mou = tx[vo - 1, js, ks] > 0.5
        if not Min.any() or not Mou.any():
            raise ValueError(f"ladder: no live links on the voltmeter planes (delta_v = {delta_v})")
        line_ok = np.all(tx[vi - 1:vo - 1, js, ks] > 0.5, axis=0)
        if not line_ok.any():
            raise ValueError("ladder: no complete line between the voltmeter planes")
        line_f1 = line_ok & np.all(f[vi - 1:vo, js, ks] >= 1 - 1e-12, axis=0)
        return cls(vi, vo, st.model.kappa ** 2, Min, Mou, line_ok, line_f1)

    @classmethod
    def of_crossings(cls, st, delta_v: int) -> "Voltmeter":
        """crossing_times' voltmeter (protocol.jl:273-279): the complete lines only."""
        gr, mk = st.grid, st.masks
        vi, vo = gr.x1 + delta_v, gr.x2 - delta_v
        if vi >= vo:
            raise ValueError(f"crossing_times: delta_v = {delta_v} leaves no voltmeter")
        js, ks = slice(gr.y1 - 1, gr.y2 + 1), slice(gr.z1 - 1, gr.z2 + 1)
        line_ok = np.all(mk.tx[vi - 1:vo - 1, js, ks] > 0.5, axis=0)
        if not line_ok.any():
            raise ValueError("crossing_times: no complete line between the voltmeter planes")
        return cls(vi, vo, st.model.kappa ** 2, line_ok.copy(), line_ok.copy(), line_ok, line_ok.copy())


def pendiente(t, y) -> float:
    """Least-squares slope of y against t (observables.jl:23-26)."""
    A = np.column_stack([np.asarray(t, dtype=float), np.ones(len(t))])
    return float(np.linalg.lstsq(A, np.asarray(y, dtype=float), rcond=None)[0][0])


def crossings_now(T1, tabs: float, m: int, dt_cyc: float) -> float:
    """_crossings_now (protocol.jl:193-199): the crossings per line a rung closing now would
    count (second-half slope × time), zero without a whole turn.  T1[1..m] is the series (index
    0 unused), `tabs` the sum over the complete lines of |T|."""
    if tabs / TWO_PI < 1:
        return 0.0
    wh = list(range(m - m // 2 + 1, m + 1))
    if len(wh) < 3:
        return 0.0
    th = [(q - wh[0]) * dt_cyc for q in wh]
    return abs(pendiente(th, [T1[q] for q in wh])) * m * dt_cyc / TWO_PI


def late_crossings(T1, m: int) -> float:
    """_late_crossings (protocol.jl:204-207): the turn since the first quarter of the point."""
    q = max(1, m // 4)
    return abs(T1[m] - T1[q]) * m / (m - q) / TWO_PI if m > q else 0.0


def silent_now(T1, tabs: float, m: int, dt_cyc: float, lim: float) -> bool:
    """_silent_now (protocol.jl:208-209)."""
    return max(crossings_now(T1, tabs, m, dt_cyc), late_crossings(T1, m)) < lim


def dl(gr) -> float:
    """grid.jl:33."""
    return 2.0 / (1 / gr.dx ** 2 + 1 / gr.dy ** 2 + 1 / gr.dz ** 2)


def dt_release10(md, gr) -> float:
    """protocol.jl:76-83: the step of release 1.0, for the tolerance per cycle."""
    l = dl(gr)
    dt1 = 5.79 * l / (4 * math.sqrt(1 + 10.0 ** 2))
    bands = [dt1 if mu == 0 else dt1 * b.m * b.Gamma for mu, b in enumerate(md.bands)]
    dt2 = 1.00 * l / (4 * md.kappa ** 2)
    dt3 = 0.03 * l / (4 * 0.50 ** 2)
    return 0.99 * min(*bands, dt2, dt2, dt3)


@dataclass
class Protocol:
    """The fields of Protocol (protocol.jl:27-68) that the cycle loop of a rung reads, with its
    defaults."""
    lMax: int = 40
    iter_min: int = 40
    mMax: int = 500_000
    t_min_GL: float = 400.0
    t_min_GL_res: float = 1584.0
    eps_conv: float = 1e-4
    eps_rate: float = 0.0
    eps_seed: float = 5e-4
    delta_v: int = 8
    V_win: int = 40
    V_win_min_res: int = 400
    use_Vconv: bool = True
    slip_band_mode: str = "band1"
    v_margin: float = 1.1
    f_strict: bool = True
    f_margin: float = 0.5
    guard_psi: float = 5.0
    n_frac: float = 0.02


def eps_cycle(pr: Protocol, st) -> float:
    """protocol.jl:85-86."""
    if pr.eps_rate > 0:
        return pr.eps_rate * (pr.lMax * st.dt)
    return pr.eps_conv * (st.dt / dt_release10(st.model, st.grid))


def crossing_times(backend, st, J: float, H: float, T: float, stop: int = 3, delta_v: int = 8, every: int = 10):
    """crossing_times (protocol.jl:259-296) with the measurements on the backend: the times at
    which the line-averaged band-1 phase across the voltmeter passes a further multiple of 2π,
    sampled every `every` steps, stopping after `stop` crossings (0: never).  Returns
    (times, t, diverged)."""
    vm = Voltmeter.of_crossings(st, delta_v)
    HJ = st.selffield(J)
    backend.meter_begin(vm)
    backend.meter_cycle(True)  # the path phase of the starting state is the reference
    times = []
    nxt = 1
    nst = round(T / st.dt)
    q = 0
    while q < nst:
        backend.step(every, H, HJ)
        q += every
        ph = backend.meter_cycle(False)["T"][0]
        if not math.isfinite(ph):
            return times, q * st.dt, True
        if abs(ph) >= nxt * TWO_PI:
            times.append(q * st.dt)
            nxt += 1
        if stop > 0 and len(times) >= stop:
            break
    return times, q * st.dt, False


@dataclass
class Rung:
    """The cycle loop of one ladder point (protocol.jl:401-489), deterministic path (no noise,
    no heat): cycles of lMax steps until the point closes 'F' (a fixed point: |Ψ| and hz settled,
    and silent across the voltmeter) or 'V' (a flow of at least v_margin × 10 crossings per line
    over the second half), or reaches mMax ('*').  Series are 1-based (index 0 unused)."""
    conv: str = "*"
    m: int = 0
    V: list = field(default_factory=lambda: [0.0])
    TA: list = field(default_factory=lambda: [0.0])
    Tf1: list = field(default_factory=lambda: [0.0])
    T: list = field(default_factory=list)
    It: list = field(default_factory=lambda: [0.0])
    Is: list = field(default_factory=lambda: [0.0])
    dmax: float = 0.0
    nbig: int = 0
    res: list = field(default_factory=list)
    slip_band: int = 0
    tabs: list = field(default_factory=list)   # Σ|T| over the complete lines per band at the close
    m_out: int = 0                             # the reference's out.m: m at the close (mMax + 1 if open)


def run_rung(backend, st, pr: Protocol, J: float, H: float, seed_point: bool = False, m_max=None) -> Rung:
    """One ladder point at current J and field H from the backend's state, its meter begun with
    the ladder's voltmeter (Voltmeter.of_ladder, backend.meter_begin) before the first point."""
    N = st.N
    dt_cyc = pr.lMax * st.dt
    iter_min = pr.iter_min
    if pr.t_min_GL > 0:
        iter_min = max(iter_min, math.ceil(pr.t_min_GL / dt_cyc))
    it_res_min = max(iter_min, math.ceil(pr.t_min_GL_res / dt_cyc))
    nwin = max(2, pr.V_win)
    eps_here = pr.eps_seed if seed_point else eps_cycle(pr, st)
    bands = list(range(N)) if pr.slip_band_mode == "any_band" else [0]
    HJ = st.selffield(J)
    r = Rung(T=[[0.0] for _ in range(N)])
    tabs = [0.0] * N
    mmax = pr.mMax if m_max is None else m_max
    m = 1
    while m <= mmax:
        backend.step(pr.lMax, H, HJ)
        rec = backend.meter_cycle(m == 1)
        r.V.append(rec["V"])
        r.TA.append(rec["TA"])
        r.Tf1.append(rec["Tf1"])
        for mu in range(N):
            r.T[mu].append(rec["T"][mu])
        tabs = rec["Tabs"]
        r.It.append(rec["It"])
        r.Is.append(rec["Is"])
        r.dmax, r.nbig = rec["dmax"], rec["nbig"]
        r.res = list(rec["dif"])
        r.tabs = list(tabs)
        amax = rec["amax"]
        if pr.guard_psi > 0 and (not math.isfinite(amax) or amax > pr.guard_psi):
            raise RuntimeError(f"the scheme diverged at J = {J}, cycle {m}: max|Ψ| = {amax}; dt = {st.dt}")
        if st.noise is not None:
            # with noise a rung is silent when it held t_min_GL_res without a phase slip on any band (:472-477)
            if m >= it_res_min and max(abs(r.T[mu][m]) for mu in range(N)) < TWO_PI:
                if not pr.f_strict or all(silent_now(r.T[mu], tabs[mu], m, dt_cyc, pr.f_margin) for mu in bands):
                    r.conv = "F"
                    break
        elif m >= iter_min and max(rec["dif"]) < eps_here and rec["diffh"] < eps_here:
            if m < 2 or max(abs(r.T[mu][m] - r.T[mu][max(1, m // 2)]) for mu in range(N)) < TWO_PI:
                if not pr.f_strict or all(silent_now(r.T[mu], tabs[mu], m, dt_cyc, pr.f_margin) for mu in bands):
                    r.conv = "F"
                    break
        if pr.use_Vconv and abs(J) > 0 and m >= max(it_res_min, 2 * nwin):
            for mu in bands:
                plateau = abs(r.T[mu][m] - r.T[mu][max(1, m // 2)]) * 2 >= pr.v_margin * 10 * TWO_PI
                if plateau and (not pr.f_strict or crossings_now(r.T[mu], tabs[mu], m, dt_cyc) >= 10):
                    r.conv = "V"
                    if pr.slip_band_mode == "any_band":
                        r.slip_band = mu + 1
                    break
            if r.conv == "V":
                break
        m += 1
    r.m = min(m, mmax)
    r.m_out = m
    return r


def meanlin(v) -> float:
    """meanlin (observables.jl:15): the sequential sum over the count."""
    v = np.asarray(v, dtype=float).ravel()
    return float(np.cumsum(v)[-1] / v.size) if v.size else math.nan


def magnetization(hz, H: float, gr) -> float:
    """magnetization (observables.jl:34-37): the mean of hz over the sample's inner plaquettes less H."""
    return meanlin(np.asarray(hz)[gr.x1 - 1:gr.x2 - 1, gr.y1 - 1:gr.y2 - 1, gr.z1 - 1:gr.z2].ravel(order="F")) - H


@dataclass
class Ladder:
    """The points of one polarity (the reference's Ladder, protocol.jl:366-371, 609-613)."""
    label: str
    J: list
    H: list
    V: np.ndarray
    Vb: np.ndarray
    Vh: np.ndarray
    Vst: np.ndarray
    m: np.ndarray
    conv: list
    res: np.ndarray
    nslip: np.ndarray
    M: np.ndarray
    psi2: np.ndarray
    dmax: np.ndarray
    nbig: np.ndarray
    Vf1: np.ndarray
    VA: np.ndarray
    t_cycle: float
    state: object = None


def ladder(backend, st, s0, Js, H, pr: Protocol, label: str = "J+") -> Ladder:
    """ladder (protocol.jl:325-621) on a backend, without output files, heat, confirmation ring or
    resume: from `s0`, point after point of the currents `Js`, each closed by `run_rung`, and what
    the reference records at the close (:493-545)."""
    gr, N = st.grid, st.N
    np_ = len(Js)
    Hs = [float(H)] * np_ if np.isscalar(H) else [float(x) for x in H]
    dt_cyc = pr.lMax * st.dt
    nwin = max(2, pr.V_win)
    vm = Voltmeter.of_ladder(st, pr.delta_v)
    nlin_f1 = int(np.sum(vm.line_f1))
    out = Ladder(label, [float(j) for j in Js], Hs, np.zeros(np_), np.full((np_, N), np.nan), np.full((np_, N), np.nan),
                 np.zeros(np_), np.zeros(np_, dtype=int), [" "] * np_, np.zeros((np_, N)), np.zeros(np_), np.zeros(np_),
                 np.zeros((np_, N)), np.zeros(np_), np.zeros(np_, dtype=int), np.full(np_, np.nan), np.full(np_, np.nan),
                 dt_cyc)
    backend.upload(s0)
    backend.meter_begin(vm)
    lastF = 0.0
    for jp in range(np_):
        J, Hc = out.J[jp], Hs[jp]
        r = run_rung(backend, st, pr, J, Hc, seed_point=(jp == 0))
        conv = r.conv
        mlast = r.m
        nav = min(nwin, mlast)
        out.Vst[jp] = meanlin(r.V[mlast - nav + 1:mlast + 1])
        if conv == "V":
            nav = min(max(pr.V_win_min_res, mlast // 2), mlast)
        wq = list(range(mlast - nav + 1, mlast + 1))
        nslip, Vc = 0.0, 0.0
        if len(wq) >= 3:
            tq = [(q - wq[0]) * dt_cyc for q in wq]
            for mu in range(N):
                out.Vb[jp, mu] = -pendiente(tq, [r.T[mu][q] for q in wq])
            Vc = out.Vb[jp, 0]
            if nlin_f1 > 0:
                out.Vf1[jp] = -pendiente(tq, [r.Tf1[q] for q in wq])
            out.VA[jp] = (r.TA[wq[-1]] - r.TA[wq[0]]) / ((wq[-1] - wq[0]) * dt_cyc)
            nslip = r.tabs[0] / TWO_PI
            if nslip < 1:
                Vc = 0.0
        wh = list(range(mlast - mlast // 2 + 1, mlast + 1))
        if len(wh) >= 3:
            th = [(q - wh[0]) * dt_cyc for q in wh]
            for mu in range(N):
                ns = r.tabs[mu] / TWO_PI
                out.Vh[jp, mu] = 0.0 if ns < 1 else -pendiente(th, [r.T[mu][q] for q in wh])
        out.V[jp], out.nslip[jp], out.m[jp], out.conv[jp] = Vc, nslip, r.m_out, conv
        out.dmax[jp], out.nbig[jp] = r.dmax, r.nbig
        out.res[jp, :] = r.res
        s = backend.download()
        out.M[jp] = magnetization(s.hz, Hc, gr)
        blk = (slice(gr.x1 - 1, gr.x2 + 1), slice(gr.y1 - 1, gr.y2 + 1), slice(gr.z1 - 1, gr.z2 + 1))
        for mu in range(N):
            out.psi2[jp, mu] = meanlin((np.abs(s.F[mu][blk]) ** 2).ravel(order="F"))
        if conv == "F" and pr.n_frac > 0:
            if abs(J) > 0 and lastF > 0 and out.psi2[jp, 0] < pr.n_frac * lastF:
                out.conv[jp] = "N"
            else:
                lastF = max(lastF, out.psi2[jp, 0])
        out.state = s
    return out


def rung_rate(L: Ladder) -> list:
    """rung_rate (protocol.jl:130-144): Vh of band 1 where measured, else V."""
    return [L.V[q] if math.isnan(L.Vh[q, 0]) else L.Vh[q, 0] for q in range(len(L.V))]


def rung_time(L: Ladder) -> list:
    return [int(m) * L.t_cycle for m in L.m]


def rung_class(crossings: float, conv: str, require_silence: bool = False) -> str:
    """_rung_class (observables.jl:148-151): 'D' resistive, 'S' silent, 'U' undecided."""
    if conv == "N" or crossings >= 10:
        return "D"
    return "S" if crossings < 1 and (not require_silence or conv in ("F", "D")) else "U"


def bracket_Ic(J, V, tint, conv=None, require_silence: bool = False) -> dict:
    """bracket_Ic (observables.jl:132-146): the highest silent current below the first resistive
    one (lo) and that one (hi)."""
    idx = [q for q in range(len(J)) if abs(J[q]) > 1e-12]
    Ja = [abs(float(J[q])) for q in idx]
    cr = [abs(V[q]) * tint[q] / TWO_PI for q in idx]
    cls = [rung_class(cr[k], " " if conv is None else conv[idx[k]], require_silence) for k in range(len(cr))]
    d = cls.index("D") if "D" in cls else None
    hi = math.inf if d is None else Ja[d]
    below = list(range(len(Ja))) if d is None else list(range(d))
    sil = [q for q in below if cls[q] == "S"]
    lo = Ja[sil[-1]] if sil else math.nan
    und = [Ja[q] for q in below if cls[q] == "U" and (math.isnan(lo) or Ja[q] > lo)]
    return dict(lo=lo, hi=hi, cls=cls, J=Ja, crossings=cr, undecided=und,
                ok=(not math.isnan(lo)) and math.isfinite(hi) and not und)


def run_ladders(backend_of, st, Js, H: float, pr: Protocol):
    """run_ladders (protocol.jl:842-861): both polarities from the Meissner seed, the seed point
    J = 0 first, and their brackets (:809-817).  `backend_of(st)` makes a fresh device stepper."""
    from .state import initial_state
    Jp = sorted(set([0.0] + [abs(float(j)) for j in Js]))
    res = {}
    for label, sign in (("J+", 1.0), ("J-", -1.0)):
        s0 = initial_state(st.model, st.grid, st.masks, H)
        res[label] = ladder(backend_of(st), st, s0, [sign * j for j in Jp], H, pr, label)
    pos, neg = res["J+"], res["J-"]
    br = {k: bracket_Ic(L.J, rung_rate(L), rung_time(L), conv=L.conv) for k, L in (("pos", pos), ("neg", neg))}
    return pos, neg, br
