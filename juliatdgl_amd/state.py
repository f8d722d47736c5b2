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
    """Links e^{-iφ} (state.jl:62)."""
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

    def copy(self) -> "State":
        return State([g.copy(order="F") for g in self.G], [f.copy(order="F") for f in self.F],
                     *(getattr(self, k).copy(order="F") for k in FIELDS))

    def arrays(self):
        """(name, array) of every field, G and F per band first."""
        out = [(f"G{mu + 1}", g) for mu, g in enumerate(self.G)] + [(f"F{mu + 1}", f) for mu, f in enumerate(self.F)]
        return out + [(k, getattr(self, k)) for k in FIELDS]


def initial_state(md: Model, gr: Grid, mk: Masks, H: float) -> State:
    """Meissner seed (state.jl:86-100): Ψ_µ = τψ·√(ε_µ/β_µ), A = 0, B_z = H on the z faces."""
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    G = [np.asarray(mk.tpsi * np.sqrt(max(md.eps(mu), 0.0) / md.bands[mu].beta), dtype=np.complex128, order="F")
         for mu in range(md.N)]
    th = (np.zeros((Nx, Ny + 1, Nz + 1), order="F"), np.zeros((Nx + 1, Ny, Nz + 1), order="F"),
          np.zeros((Nx + 1, Ny + 1, Nz), order="F"))
    hx = np.zeros((Nx + 1, Ny, Nz), order="F")
    hy = np.zeros((Nx, Ny + 1, Nz), order="F")
    hz = np.zeros((Nx, Ny, Nz + 1), order="F")
    hz[:, :, 0] = H
    hz[:, :, Nz] = H
    return State.from_fields(gr, G, th, th, (hx, hy, hz))


def seeded_state(md: Model, gr: Grid, mk: Masks, H: float, seed: int, amp: float = 0.01) -> State:
    """Ψ_µ = τψ·(1 + amp·(n₁ + i n₂)) with n from a seeded normal stream (numpy PCG64), A = 0,
    B_z = H on the z faces: the seeded initial condition of the benchmark configurations."""
    st = initial_state(md, gr, mk, H)
    rng = np.random.default_rng(seed)
    for mu in range(md.N):
        n = rng.standard_normal(size=(2,) + st.G[mu].shape)
        g = mk.tpsi * (1.0 + amp * (n[0] + 1j * n[1]))
        st.G[mu] = np.asarray(g, dtype=np.complex128, order="F")
        st.F[mu] = st.G[mu].copy(order="F")
    return st
