"""The draws of thermal noise in the layout the stepper takes them (tdgl_noise_draws).

With `noise_draws="host"` the caller hands over the normals of each step: per band a complex node
array ξ = n₁ + i n₂ at the nodes the kick visits, then with `noise_field` a real array per link
family at the links the field kick visits.  `from_stream` lays out one step of a flat stream in
the reference's consumption order (gtdgl/src/noise.jl:47-52, 108-111, 188-196: band by band, k, j,
i outer to inner over the interior nodes that are not hard f = 0 ones, two normals per node, then
the x, y, z links), which is how the reference's GPU backend takes its Xoshiro stream
(gpu/TDGLGPU/src/TDGLGPU.jl:714-870): any stream in that order runs here number for number.
"""
from __future__ import annotations

import numpy as np

from .scheme import Stepper


def link_shapes(gr):
    return ((gr.Nx, gr.Ny + 1, gr.Nz + 1), (gr.Nx + 1, gr.Ny, gr.Nz + 1), (gr.Nx + 1, gr.Ny + 1, gr.Nz))


def kicked_nodes(st: Stepper) -> np.ndarray:
    """Mask (Nx+1, Ny+1, Nz+1) of the nodes the kick visits: the interior, without the hard f = 0
    nodes (noise.jl:47-50)."""
    gr = st.grid
    m = np.zeros((gr.Nx + 1, gr.Ny + 1, gr.Nz + 1), dtype=bool, order="F")
    m[gr.x1:gr.x2, gr.y1:gr.y2, gr.z1:gr.z2] = True
    if st.hard:
        m &= st.masks.f != 0.0
    return m


def link_ranges(gr):
    """0-based slices of the links field_kick! visits (noise.jl:188-196)."""
    return ((slice(0, gr.Nx), slice(1, gr.Ny), slice(1, gr.Nz)),
            (slice(1, gr.Nx), slice(0, gr.Ny), slice(1, gr.Nz)),
            (slice(1, gr.Nx), slice(1, gr.Ny), slice(0, gr.Nz)))


def per_step(st: Stepper) -> int:
    """Normals one step takes from a stream in the reference's order."""
    n = 2 * st.N * int(kicked_nodes(st).sum())
    if st.noise is not None and st.noise.field:
        gr = st.grid
        n += sum(int(np.prod([s.stop - s.start for s in r])) for r in link_ranges(gr))
    return n


def from_stream(st: Stepper, normals) -> tuple:
    """(ξ per band, n per link family or None) of one step from `per_step(st)` normals in the
    reference's order."""
    gr = st.grid
    z = np.asarray(normals, dtype=np.float64).ravel()
    if z.size != per_step(st):
        raise ValueError(f"one step takes {per_step(st)} normals, not {z.size}")
    m = kicked_nodes(st)
    # the nodes in loop order k, j, i (i fastest): column-major order of the mask
    idx = np.flatnonzero(m.ravel(order="F"))
    xi = []
    p = 0
    for _ in range(st.N):
        a = np.zeros(m.size, dtype=np.complex128)
        pairs = z[p:p + 2 * idx.size].reshape(-1, 2)
        a[idx] = pairs[:, 0] + 1j * pairs[:, 1]
        xi.append(a.reshape(m.shape, order="F"))
        p += 2 * idx.size
    links = None
    if st.noise is not None and st.noise.field:
        links = []
        for shp, r in zip(link_shapes(gr), link_ranges(gr)):
            a = np.zeros(shp, order="F")
            sub = a[r]
            cnt = sub.size
            a[r] = z[p:p + cnt].reshape(sub.shape, order="F")
            links.append(a)
            p += cnt
    return xi, links


def pack(st: Stepper, steps) -> np.ndarray:
    """The flat array tdgl_noise_draws takes: for each step (ξ per band, n per link family or
    None), the band arrays then the link arrays, column-major."""
    gr = st.grid
    out = []
    for xi, links in steps:
        if len(xi) != st.N:
            raise ValueError(f"{st.N} bands of draws per step")
        for a in xi:
            a = np.asarray(a, dtype=np.complex128)
            if a.shape != (gr.Nx + 1, gr.Ny + 1, gr.Nz + 1):
                raise ValueError(f"band draws are node-shaped, not {a.shape}")
            out.append(a.ravel(order="F").view(np.float64))
        if st.noise is not None and st.noise.field:
            if links is None:
                raise ValueError("noise_field: each step needs the link draws")
            for a, shp in zip(links, link_shapes(gr)):
                a = np.asarray(a, dtype=np.float64)
                if a.shape != shp:
                    raise ValueError(f"link draws of shape {shp}, not {a.shape}")
                out.append(a.ravel(order="F"))
    return np.ascontiguousarray(np.concatenate(out)) if out else np.zeros(0)
