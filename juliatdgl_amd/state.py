"""State{N} of the release-2.0 step as host arrays (gtdgl/src/state.jl:51-100).

Arrays are numpy, Fortran order, in the Julia shapes (1-based index i of the reference is
index i-1 here).  Complex fields are complex128 (interleaved re, im as ComplexF64).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from .scheme import Grid, Masks, Model


def emiphi(phi: np.ndarray) -> np.ndarray:
    """Links e^{-iφ} (state.jl:62).  All +0.0 (a fresh state): 1 − 0i everywhere, as cos and sin of
    −0.0 give it, without evaluating them."""
    phi = np.asarray(phi, dtype=np.float64)
    if not phi.view(np.uint64).any():
        return np.full(phi.shape, complex(1.0, -0.0), dtype=np.complex128, order="F")
    out = np.empty(phi.shape, dtype=np.complex128, order="F")
    out.real = np.cos(-phi)
    out.imag = np.sin(-phi)
    return out


FIELDS = ("thx", "thy", "thz", "phx", "phy", "phz", "Ux", "Uy", "Uz", "Vx", "Vy", "Vz",
          "Jx", "Jy", "Jz", "hx", "hy", "hz")


@dataclass
class State:
    G: List[np.ndarray]
    F: List[np.ndarray]
    thx: np.ndarray
    thy: np.ndarray
    thz: np.ndarray
    phx: np.ndarray
    phy: np.ndarray
    phz: np.ndarray
    Ux: np.ndarray
    Uy: np.ndarray
    Uz: np.ndarray
    Vx: np.ndarray
    Vy: np.ndarray
    Vz: np.ndarray
    Jx: np.ndarray
    Jy: np.ndarray
    Jz: np.ndarray
    hx: np.ndarray
    hy: np.ndarray
    hz: np.ndarray

    @property
    def N(self) -> int:
        return len(self.G)

    @staticmethod
    def from_fields(gr: Grid, G, theta, theta_prev, h) -> "State":
        """State(gr, G; θ, θprev, h) (state.jl:69-79): F = G, V = e^{-iθ}, U = e^{-iθprev}, J = 0."""
        Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
        F = lambda a, dt=np.float64: np.array(a, dtype=dt, order="F", copy=True)
        thx, thy, thz = (F(t) for t in theta)
        phx, phy, phz = (F(t) for t in theta_prev)
        Gs = [F(g, np.complex128) for g in G]
        return State(Gs, [g.copy(order="F") for g in Gs], thx, thy, thz, phx, phy, phz,
                     emiphi(phx), emiphi(phy), emiphi(phz), emiphi(thx), emiphi(thy), emiphi(thz),
                     np.zeros((Nx, Ny + 1, Nz + 1), order="F"), np.zeros((Nx + 1, Ny, Nz + 1), order="F"),
                     np.zeros((Nx + 1, Ny + 1, Nz), order="F"), F(h[0]), F(h[1]), F(h[2]))

    def empty_like(self) -> "State":
        """A State of uninitialised arrays of the same shapes (what a download fills)."""
        e = lambda a: np.empty(a.shape, dtype=a.dtype, order="F")
        return State([e(g) for g in self.G], [e(f) for f in self.F], *(e(getattr(self, k)) for k in FIELDS))

    def copy(self) -> "State":
        return State([g.copy(order="F") for g in self.G], [f.copy(order="F") for f in self.F],
                     *(getattr(self, k).copy(order="F") for k in FIELDS))

    def arrays(self):
        """(name, array) of every field, G and F per band first."""
        out = [(f"G{mu + 1}", g) for mu, g in enumerate(self.G)] + [(f"F{mu + 1}", f) for mu, f in enumerate(self.F)]
        return out + [(k, getattr(self, k)) for k in FIELDS]


def _fresh(md: Model, gr: Grid, H: float, G) -> State:
    """The state of fresh arrays G (taken, not copied) with A = 0 and B_z = H on the z faces:
    F = G, θ = φ = 0, U = V = 1 − 0i (e^{-i·0}), J = 0 (state.jl:69-79, 86-100)."""
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    sh = ((Nx, Ny + 1, Nz + 1), (Nx + 1, Ny, Nz + 1), (Nx + 1, Ny + 1, Nz))
    z = lambda s: np.zeros(s, order="F")
    one = lambda s: np.full(s, complex(1.0, -0.0), dtype=np.complex128, order="F")
    hz = z((Nx, Ny, Nz + 1))
    hz[:, :, 0] = H
    hz[:, :, Nz] = H
    return State(G, [g.copy(order="F") for g in G], *(z(s) for s in sh), *(z(s) for s in sh),
                 *(one(s) for s in sh), *(one(s) for s in sh), *(z(s) for s in sh),
                 z((Nx + 1, Ny, Nz)), z((Nx, Ny + 1, Nz)), hz)


def initial_state(md: Model, gr: Grid, mk: Masks, H: float) -> State:
    """Meissner seed (state.jl:86-100): Ψ_µ = τψ·√(ε_µ/β_µ), A = 0, B_z = H on the z faces."""
    G = [np.asarray(mk.tpsi * np.sqrt(max(md.eps(mu), 0.0) / md.bands[mu].beta), dtype=np.complex128, order="F")
         for mu in range(md.N)]
    return _fresh(md, gr, H, [np.array(g, order="F", copy=True) for g in G])


def seeded_state(md: Model, gr: Grid, mk: Masks, H: float, seed: int, amp: float = 0.01) -> State:
    """Ψ_µ = τψ·(1 + amp·(n₁ + i n₂)) with n from a seeded normal stream (numpy PCG64), A = 0,
    B_z = H on the z faces: the seeded initial condition of the benchmark configurations."""
    rng = np.random.default_rng(seed)
    G = []
    for mu in range(md.N):
        n = rng.standard_normal(size=(2,) + mk.tpsi.shape)
        G.append(np.asarray(mk.tpsi * (1.0 + amp * (n[0] + 1j * n[1])), dtype=np.complex128, order="F"))
    return _fresh(md, gr, H, G)
