"""ctypes binding of oracle/_build/liboracle.so (the C++ restatement of gtdgl's step!).

TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  Built by `make -C oracle`.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

from juliatdgl_amd.abi import PackedDesc, PcgC, PcgState, state_struct
from juliatdgl_amd.scheme import Stepper
from juliatdgl_amd.state import State

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        build()
        L = ct.CDLL(LIB)
        L.orc_create.restype = ct.c_void_p
        L.orc_create.argtypes = [ct.c_void_p]
        L.orc_destroy.argtypes = [ct.c_void_p]
        L.orc_set_state.argtypes = [ct.c_void_p, ct.c_void_p]
        L.orc_get_state.argtypes = [ct.c_void_p, ct.c_void_p]
        L.orc_step.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_double]
        L.orc_get_pcg.argtypes = [ct.c_void_p, ct.c_void_p]
        L.orc_set_pcg.argtypes = [ct.c_void_p, ct.c_void_p]
        L.orc_total_iters.argtypes = [ct.c_void_p]
        L.orc_total_iters.restype = ct.c_longlong
        L.orc_sep_solve.argtypes = [ct.c_void_p] + [ct.c_void_p] * 3
        L.orc_basis.argtypes = [ct.c_int] + [ct.c_void_p] * 3
        L.orc_kwt_update.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_void_p]
        L.orc_rotation.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int, ct.c_void_p]
        L.orc_field_solve.argtypes = [ct.c_void_p]
        L.orc_link_rate.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_double, ct.c_double]
        L.orc_link_rate.restype = ct.c_double
        L.orc_scis.argtypes = [ct.c_double, ct.c_void_p]
        L.orc_qp_modes.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p]
        L.orc_qp_source.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p]
        L.orc_qp_potential.argtypes = [ct.c_void_p]
        L.orc_open_state.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.orc_open_state.restype = ct.c_int
        L.orc_open_fields.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.orc_set_open.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
        L.orc_noise_get.argtypes = [ct.c_void_p]
        L.orc_noise_get.restype = ct.c_uint64
        L.orc_noise_set.argtypes = [ct.c_void_p, ct.c_uint64]
        L.orc_noise_draws.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p]
        L.orc_philox.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.orc_normal2.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_uint32, ct.c_uint32, ct.c_void_p]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's row loops (same bits with any number)."""
    import ctypes.util
    gomp = ct.CDLL(ctypes.util.find_library("gomp") or "libgomp.so.1")
    gomp.omp_set_num_threads(int(n))


class Oracle:
    """The CPU step on a host State: set_state / step / get_state, as the C ABI does on the GPU."""

    def __init__(self, st: Stepper):
        self.st = st
        self._desc = PackedDesc(st)
        self.h = lib().orc_create(ct.byref(self._desc.c))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_state(self, s: State) -> None:
        self._s = s.copy()
        lib().orc_set_state(self.h, ct.byref(state_struct(self._s)))

    def get_state(self) -> State:
        out = self._s.copy()
        lib().orc_get_state(self.h, ct.byref(state_struct(out)))
        return out

    def step(self, n: int, H: float, HJ: float) -> None:
        nz = self.st.noise
        if nz is not None and nz.host and n > 0:
            left = getattr(self, "_slots", 0) - (self.noise_counter - getattr(self, "_base", 0))
            if left < n:
                raise ValueError(f"noise: draws for {max(left, 0)} more steps, {n} asked")
        lib().orc_step(self.h, int(n), float(H), float(HJ))

    @property
    def noise_counter(self) -> int:
        """Steps taken with noise since the stream began."""
        return int(lib().orc_noise_get(self.h))

    @noise_counter.setter
    def noise_counter(self, c: int) -> None:
        lib().orc_noise_set(self.h, int(c))

    def noise_draws(self, flat) -> None:
        """The caller's normals for the next steps (juliatdgl_amd.noise.pack)."""
        from juliatdgl_amd.noise import per_step  # noqa: F401  (layout owner)
        gr, st = self.st.grid, self.st
        nn = (gr.Nx + 1) * (gr.Ny + 1) * (gr.Nz + 1)
        nl = sum(a * b * c for a, b, c in ((gr.Nx, gr.Ny + 1, gr.Nz + 1), (gr.Nx + 1, gr.Ny, gr.Nz + 1),
                                           (gr.Nx + 1, gr.Ny + 1, gr.Nz)))
        stride = st.N * 2 * nn + (nl if st.noise.field else 0)
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        if flat.size % stride:
            raise ValueError("noise draws: not a whole number of steps")
        self._draws = flat
        self._slots = flat.size // stride
        self._base = self.noise_counter
        lib().orc_noise_draws(self.h, self._slots, flat.ctypes.data)

    def get_pcg(self) -> PcgState:
        p = PcgState(self.st.grid)
        c = p.struct()
        lib().orc_get_pcg(self.h, ct.byref(c))
        p.load(c)
        return p

    def set_pcg(self, p: PcgState) -> None:
        lib().orc_set_pcg(self.h, ct.byref(p.struct()))

    def total_iters(self) -> int:
        return int(lib().orc_total_iters(self.h))

    def _qp_region(self):
        q = self.st.qp
        return (self.st.grid.Nx - 1, q.nj, q.nk)

    def qp_potential(self):
        """qp_potential! on the current state (no step); returns the solved modes (δ for two bands),
        each region-shaped ((Nx-1), nj, nk), and the sources they were solved from."""
        lib().orc_qp_potential(self.h)
        return self.qp_modes(), self.qp_sources()

    def qp_modes(self):
        out = []
        for k in range(self.st.qp.K):
            a = np.zeros(self._qp_region(), order="F")
            lib().orc_qp_modes(self.h, k, a.ctypes.data)
            out.append(a)
        return out

    def qp_sources(self):
        out = []
        for k in range(self.st.qp.K):
            a = np.zeros(self._qp_region(), order="F")
            lib().orc_qp_source(self.h, k, a.ctypes.data)
            out.append(a)
        return out

    def open_state(self):
        """(count, old, new) of the open boundary: h_δ of its last two recomputations, each the
        three h-shaped arrays."""
        gr = self.st.grid
        sh = ((gr.Nx + 1, gr.Ny, gr.Nz), (gr.Nx, gr.Ny + 1, gr.Nz), (gr.Nx, gr.Ny, gr.Nz + 1))
        n = sum(int(np.prod(x)) for x in sh)
        a, b = np.zeros(n), np.zeros(n)
        cnt = lib().orc_open_state(self.h, a.ctypes.data, b.ctypes.data)
        return cnt, _split(a, sh), _split(b, sh)

    def set_open(self, count: int, old, new) -> None:
        """The clock and ramp of the open boundary (as open_state returns them)."""
        a = np.concatenate([np.asarray(x, dtype=np.float64).ravel(order="F") for x in old])
        b = np.concatenate([np.asarray(x, dtype=np.float64).ravel(order="F") for x in new])
        lib().orc_set_open(self.h, int(count), a.ctypes.data, b.ctypes.data)

    def open_fields(self):
        """θ_δ and δJ of the last recomputation, each the three link-shaped arrays."""
        gr = self.st.grid
        sh = ((gr.Nx, gr.Ny + 1, gr.Nz + 1), (gr.Nx + 1, gr.Ny, gr.Nz + 1), (gr.Nx + 1, gr.Ny + 1, gr.Nz))
        n = sum(int(np.prod(x)) for x in sh)
        a, b = np.zeros(n), np.zeros(n)
        lib().orc_open_fields(self.h, a.ctypes.data, b.ctypes.data)
        return _split(a, sh), _split(b, sh)

    def field_solve(self) -> None:
        """The field solve of the step alone on the state as set (φ holding φₑ): implicit_curl! or
        scale_increment! (step.jl:25-29)."""
        lib().orc_field_solve(self.h)
        lib().orc_get_state(self.h, ct.byref(state_struct(self._s)))

    def sep_solve(self, BX, BY, BZ):
        """(I + aK)^{-1} on range arrays (the stepper must have the separable or PCG field)."""
        out = [np.array(b, dtype=np.float64, order="F", copy=True) for b in (BX, BY, BZ)]
        lib().orc_sep_solve(self.h, *(o.ctypes.data for o in out))
        return out


def philox(ctr, key):
    """Philox4x32-10 of one counter block (four uint32) under a key (two uint32)."""
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return [int(x) for x in out]


def normal2(seed: int, t: int, tag: int, q: int):
    """The two normals of the noise at (step t, tag, index q) under `seed`."""
    out = np.zeros(2)
    lib().orc_normal2(int(seed), int(t), int(tag), int(q), out.ctypes.data)
    return float(out[0]), float(out[1])


def _split(a, shapes):
    out, q = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(a[q:q + n].reshape(sh, order="F"))
        q += n
    return out


def basis(nn: int):
    V = np.zeros((nn, nn), order="F")
    U = np.zeros((nn - 1, nn - 1), order="F")
    s = np.zeros(nn)
    lib().orc_basis(nn, V.ctypes.data, U.ctypes.data, s.ctypes.data)
    return V, U, s


def kwt_update(g: complex, d: complex, g2: float) -> complex:
    a = np.array([g.real, g.imag])
    b = np.array([d.real, d.imag])
    o = np.zeros(2)
    lib().orc_kwt_update(a.ctypes.data, b.ctypes.data, g2, o.ctypes.data)
    return complex(o[0], o[1])


def link_rate(v: complex, u: complex, om0: float, om1: float, dt: float) -> float:
    """_link_rate! of one link: the phase of V Ū plus (ω₁ − ω₀) dt."""
    a = np.array([v.real, v.imag])
    b = np.array([u.real, u.imag])
    return float(lib().orc_link_rate(a.ctypes.data, b.ctypes.data, float(om0), float(om1), float(dt)))


def scis(a: float) -> complex:
    """_cis: the short series of e^{ia} below |a| = 0.05."""
    o = np.zeros(2)
    lib().orc_scis(float(a), o.ctypes.data)
    return complex(o[0], o[1])


def rotation(mode: int, v: complex, u: complex, c: float = 0.0, n: int = 1) -> complex:
    a = np.array([v.real, v.imag])
    b = np.array([u.real, u.imag])
    o = np.zeros(2)
    lib().orc_rotation(mode, a.ctypes.data, b.ctypes.data, float(c), int(n), o.ctypes.data)
    return complex(o[0], o[1])
