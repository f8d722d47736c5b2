"""ctypes binding of the HIP stepper library (juliatdgl_amd/_lib/libtdglhip.so, C ABI of
include/tdglhip.h).  There is no fallback: if the library is missing or no HIP device is present,
the calls raise."""
from __future__ import annotations

import ctypes as ct
import os
import threading

import numpy as np

from .abi import (NKERNELS, TUNE_KINDS, CycleRecordC, MeterC, OpenC, PackedDesc, PcgState, ProfileC, QpC, StatsC, TuningC,
                  state_struct)
from .scheme import Stepper
from .state import State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_lib", "libtdglhip.so")

SYMBOLS = ("tdgl_supported", "tdgl_create", "tdgl_destroy", "tdgl_upload_state", "tdgl_download_state", "tdgl_step",
           "tdgl_get_pcg", "tdgl_set_pcg", "tdgl_get_open", "tdgl_set_open", "tdgl_get_qp", "tdgl_get_stats", "tdgl_last_error", "tdgl_device_count",
           "tdgl_set_device", "tdgl_profile_enable", "tdgl_profile_get", "tdgl_kernel_name", "tdgl_meter_begin",
           "tdgl_meter_cycle", "tdgl_noise_get", "tdgl_noise_set", "tdgl_noise_draws", "tdgl_batch_create",
           "tdgl_batch_step", "tdgl_batch_meter_cycle", "tdgl_batch_destroy", "tdgl_step_noalloc", "tdgl_autotune")

_lib = None
_lock = threading.Lock()


class TDGLError(RuntimeError):
    pass


def load():
    """The library, loaded once.  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB):
                raise TDGLError(f"the HIP stepper library is not built ({LIB}); run python -m juliatdgl_amd.build")
            L = ct.CDLL(LIB)
            vp = ct.c_void_p
            L.tdgl_supported.argtypes = [vp, ct.c_char_p, ct.c_size_t]
            L.tdgl_supported.restype = ct.c_int
            L.tdgl_create.argtypes = [vp, ct.POINTER(vp)]
            L.tdgl_destroy.argtypes = [vp]
            L.tdgl_upload_state.argtypes = [vp, vp]
            L.tdgl_download_state.argtypes = [vp, vp]
            L.tdgl_step.argtypes = [vp, ct.c_int32, ct.c_double, ct.c_double]
            L.tdgl_step_noalloc.argtypes = [vp, ct.c_int32, ct.c_double, ct.c_double]
            L.tdgl_get_pcg.argtypes = [vp, vp]
            L.tdgl_set_pcg.argtypes = [vp, vp]
            L.tdgl_get_stats.argtypes = [vp, vp]
            L.tdgl_get_qp.argtypes = [vp, vp]
            L.tdgl_autotune.argtypes = [vp, ct.c_double, ct.c_double, vp]
            L.tdgl_last_error.restype = ct.c_char_p
            L.tdgl_device_count.restype = ct.c_int
            L.tdgl_set_device.argtypes = [ct.c_int32]
            L.tdgl_profile_enable.argtypes = [vp, ct.c_int32]
            L.tdgl_profile_get.argtypes = [vp, vp]
            L.tdgl_kernel_name.argtypes = [ct.c_int32]
            L.tdgl_meter_begin.argtypes = [vp, vp]
            L.tdgl_meter_cycle.argtypes = [vp, ct.c_int32, vp]
            L.tdgl_kernel_name.restype = ct.c_char_p
            L.tdgl_noise_get.argtypes = [vp, ct.POINTER(ct.c_uint64)]
            L.tdgl_noise_set.argtypes = [vp, ct.c_uint64]
            L.tdgl_noise_draws.argtypes = [vp, ct.c_int32, vp]
            L.tdgl_batch_create.argtypes = [vp, ct.c_int32, ct.POINTER(vp)]
            L.tdgl_batch_step.argtypes = [vp, ct.c_int32, vp, vp]
            L.tdgl_batch_meter_cycle.argtypes = [vp, ct.c_int32, vp]
            L.tdgl_batch_destroy.argtypes = [vp]
            _lib = L
        return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().tdgl_last_error().decode(errors="replace")
        raise TDGLError(f"{what} failed ({rc}): {msg}")


def device_count() -> int:
    return int(load().tdgl_device_count())


def supported_desc(st: Stepper):
    """(ok, why) of the C ABI's tdgl_supported for this stepper."""
    pd = PackedDesc(st)
    buf = ct.create_string_buffer(256)
    ok = load().tdgl_supported(ct.byref(pd.c), buf, len(buf))
    return bool(ok), buf.value.decode()


def _record(r, N: int) -> dict:
    return {"V": r.V, "TA": r.TA, "Tf1": r.Tf1, "T": [r.T[k] for k in range(N)],
            "Tabs": [r.Tabs[k] for k in range(N)], "dmax": r.dmax,
            "nbig": int(r.nbig), "It": r.It, "Is": r.Is, "dif": [r.dif[k] for k in range(N)], "amax": r.amax,
            "diffh": r.diffh}


class HipStepper:
    """One device stepper: coefficients, masks, work arrays and the state of one run."""

    def __init__(self, st: Stepper):
        self.st = st
        self._desc = PackedDesc(st)
        L = load()
        if L.tdgl_device_count() < 1:
            raise TDGLError("no HIP device")
        h = ct.c_void_p()
        _check(L.tdgl_create(ct.byref(self._desc.c), ct.byref(h)), "tdgl_create")
        self.h = h
        self._shape = None

    def close(self) -> None:
        if getattr(self, "h", None):
            load().tdgl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, s: State) -> None:
        _check(load().tdgl_upload_state(self.h, ct.byref(state_struct(s))), "tdgl_upload_state")
        self._shape = s.empty_like()

    def download(self, into: State = None) -> State:
        out = into if into is not None else self._shape.empty_like()
        _check(load().tdgl_download_state(self.h, ct.byref(state_struct(out))), "tdgl_download_state")
        return out

    def step(self, n: int, H: float, HJ: float) -> None:
        _check(load().tdgl_step(self.h, int(n), float(H), float(HJ)), "tdgl_step")

    def step_noalloc(self, n: int, H: float, HJ: float) -> None:
        """n steps through tdgl_step_noalloc: the same steps with the reductions' scratch in the
        library's fixed device array, nothing allocated."""
        _check(load().tdgl_step_noalloc(self.h, int(n), float(H), float(HJ)), "tdgl_step_noalloc")

    def get_pcg(self) -> PcgState:
        p = PcgState(self.st.grid)
        c = p.struct()
        _check(load().tdgl_get_pcg(self.h, ct.byref(c)), "tdgl_get_pcg")
        p.load(c)
        return p

    def set_pcg(self, p: PcgState) -> None:
        _check(load().tdgl_set_pcg(self.h, ct.byref(p.struct())), "tdgl_set_pcg")

    def _open_shapes(self):
        gr = self.st.grid
        return ((gr.Nx + 1, gr.Ny, gr.Nz), (gr.Nx, gr.Ny + 1, gr.Nz), (gr.Nx, gr.Ny, gr.Nz + 1))

    def get_open(self):
        """(count, old, new) of the open boundary: its clock and h_δ of the last two
        recomputations (h-shaped hx, hy, hz each)."""
        sh = self._open_shapes()
        old = [np.zeros(x, order="F") for x in sh]
        new = [np.zeros(x, order="F") for x in sh]
        c = OpenC()
        for k in range(3):
            c.hold[k] = old[k].ctypes.data_as(ct.POINTER(ct.c_double))
            c.hnew[k] = new[k].ctypes.data_as(ct.POINTER(ct.c_double))
        _check(load().tdgl_get_open(self.h, ct.byref(c)), "tdgl_get_open")
        return int(c.count), old, new

    def set_open(self, count: int, old, new) -> None:
        old = [np.asfortranarray(x, dtype=np.float64) for x in old]
        new = [np.asfortranarray(x, dtype=np.float64) for x in new]
        c = OpenC()
        c.count = int(count)
        for k in range(3):
            c.hold[k] = old[k].ctypes.data_as(ct.POINTER(ct.c_double))
            c.hnew[k] = new[k].ctypes.data_as(ct.POINTER(ct.c_double))
        _check(load().tdgl_set_open(self.h, ct.byref(c)), "tdgl_set_open")

    def get_qp(self):
        """The normal-fluid potentials of the last step (tdgl_get_qp): [δ] with two fluids, [μ_1..μ_N]
        with N, node-shaped."""
        q = self.st.qp
        if q is None:
            raise TDGLError("the stepper has one normal fluid")
        gr = self.st.grid
        npot = 1 if q.mode == 1 else self.st.N   # TDGL_QP_TWO: δ alone
        out = [np.zeros((gr.Nx + 1, gr.Ny + 1, gr.Nz + 1), order="F") for _ in range(npot)]
        c = QpC()
        for k, a in enumerate(out):
            c.pot[k] = a.ctypes.data_as(ct.POINTER(ct.c_double))
        _check(load().tdgl_get_qp(self.h, ct.byref(c)), "tdgl_get_qp")
        return out

    def autotune(self, H: float, HJ: float) -> dict:
        """Times the candidate launch shapes on the handle's own state (tdgl_autotune) and keeps the
        fastest; the state is left as it was.  Returns the choice and every candidate's time."""
        c = TuningC()
        _check(load().tdgl_autotune(self.h, float(H), float(HJ), ct.byref(c)), "tdgl_autotune")
        return {"box_kc": c.box_kc, "plane_mt": c.plane_mt, "zmodes_waves": c.zmodes_waves, "fold": c.fold, "zcol": c.zcol,
                "qparts": c.qparts,
                "candidates": [(TUNE_KINDS.get(c.cand[i].kind, "?"), c.cand[i].value, c.cand[i].us)
                               for i in range(c.ncand)]}

    def meter_begin(self, vm) -> None:
        """Uploads the voltmeter `vm` (juliatdgl_amd.meter.Voltmeter) and takes |Ψ| and hz of the
        state on the device as the last cycle's (tdgl_meter_begin)."""
        self._vm = [np.ascontiguousarray(a.ravel(order="F"), dtype=np.uint8)
                    for a in (vm.Min, vm.Mou, vm.line_ok, vm.line_f1)]
        c = MeterC()
        c.vi, c.vo, c.kappa2 = int(vm.vi), int(vm.vo), float(vm.kappa2)
        c.Min, c.Mou, c.line_ok, c.line_f1 = [a.ctypes.data_as(ct.POINTER(ct.c_uint8)) for a in self._vm]
        _check(load().tdgl_meter_begin(self.h, ct.byref(c)), "tdgl_meter_begin")

    def meter_cycle(self, first: bool) -> dict:
        """The record of the cycle that just ended (tdgl_meter_cycle); `first` starts a rung."""
        r = CycleRecordC()
        _check(load().tdgl_meter_cycle(self.h, int(first), ct.byref(r)), "tdgl_meter_cycle")
        return _record(r, self.st.N)

    @property
    def noise_counter(self) -> int:
        """The noise stream's position: steps taken with noise since it began (tdgl_noise_get)."""
        c = ct.c_uint64()
        _check(load().tdgl_noise_get(self.h, ct.byref(c)), "tdgl_noise_get")
        return int(c.value)

    @noise_counter.setter
    def noise_counter(self, c: int) -> None:
        _check(load().tdgl_noise_set(self.h, int(c)), "tdgl_noise_set")

    def noise_draws(self, flat) -> None:
        """The caller's normals for the next steps (tdgl_noise_draws; juliatdgl_amd.noise.pack
        lays them out)."""
        gr, st = self.st.grid, self.st
        nn = (gr.Nx + 1) * (gr.Ny + 1) * (gr.Nz + 1)
        nl = sum(a * b * c for a, b, c in ((gr.Nx, gr.Ny + 1, gr.Nz + 1), (gr.Nx + 1, gr.Ny, gr.Nz + 1),
                                           (gr.Nx + 1, gr.Ny + 1, gr.Nz)))
        stride = st.N * 2 * nn + (nl if st.noise is not None and st.noise.field else 0)
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        if flat.size % stride:
            raise ValueError("noise draws: not a whole number of steps")
        _check(load().tdgl_noise_draws(self.h, flat.size // stride, flat.ctypes.data), "tdgl_noise_draws")

    def noise_extra(self) -> dict:
        """What a checkpoint holds of the noise stream, in place of the reference's Xoshiro state
        (noise_rng, protocol.jl:659): its seed and position."""
        if self.st.noise is None:
            return {}
        return {"noise_seed": int(self.st.noise.seed), "noise_counter": self.noise_counter}

    def noise_restore(self, w: dict) -> None:
        """Resuming the stream where a checkpoint left it (protocol.jl:382)."""
        if self.st.noise is None or "noise_counter" not in w:
            return
        if int(w.get("noise_seed", self.st.noise.seed)) != self.st.noise.seed:
            raise ValueError("resume: the checkpoint's noise stream has another seed")
        self.noise_counter = int(w["noise_counter"])

    def profile(self, on: bool) -> None:
        """Per-kernel HIP-event timing of every launch while on (clears the totals)."""
        _check(load().tdgl_profile_enable(self.h, int(on)), "tdgl_profile_enable")

    def profile_get(self) -> dict:
        c = ProfileC()
        _check(load().tdgl_profile_get(self.h, ct.byref(c)), "tdgl_profile_get")
        L = load()
        return {L.tdgl_kernel_name(k).decode(): (int(c.launches[k]), float(c.ms[k])) for k in range(NKERNELS)}

    def stats(self) -> dict:
        c = StatsC()
        _check(load().tdgl_get_stats(self.h, ct.byref(c)), "tdgl_get_stats")
        return {k: getattr(c, k) for k, _ in StatsC._fields_}


class HipBatch:
    """Many device steppers advanced together (tdgl_batch_*): one launch per kernel of a step for
    every run, the steps of a call replayed as one graph.  The steppers keep their state and
    meter; they are not stepped by themselves while the batch is open."""

    def __init__(self, devs):
        self.devs = list(devs)
        n = len(self.devs)
        arr = (ct.c_void_p * n)(*[d.h for d in self.devs])
        b = ct.c_void_p()
        _check(load().tdgl_batch_create(arr, n, ct.byref(b)), "tdgl_batch_create")
        self.b = b

    def step(self, n: int, H, HJ) -> None:
        """n steps of every run, run r at (H[r], HJ[r])."""
        H = np.ascontiguousarray(H, dtype=np.float64)
        HJ = np.ascontiguousarray(HJ, dtype=np.float64)
        if H.shape != (len(self.devs),) or HJ.shape != (len(self.devs),):
            raise ValueError("batch step: one H and one HJ per run")
        _check(load().tdgl_batch_step(self.b, int(n), H.ctypes.data, HJ.ctypes.data), "tdgl_batch_step")

    def meter_cycle(self, first: bool):
        """The records of the cycle that just ended, one per run (tdgl_batch_meter_cycle)."""
        recs = (CycleRecordC * len(self.devs))()
        _check(load().tdgl_batch_meter_cycle(self.b, int(first), recs), "tdgl_batch_meter_cycle")
        return [_record(recs[r], d.st.N) for r, d in enumerate(self.devs)]

    def close(self) -> None:
        if getattr(self, "b", None):
            load().tdgl_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
