"""The oracle's implicit field solves and temporal link rotations against what the mathematics and
the reference's own fixtures fix, independently of the oracle's code:

* the curl-curl operator is built here as a sparse matrix from its definition (the Yee circulation
  of the link phases around every plaquette whose four edges are unknowns; the unknowns are the
  links whose two end nodes are both inside the box), so that (D + aK) never passes through the
  oracle's `_apply_A!` (gtdgl/src/implicit.jl:212-236);
* the separable solve (implicit.jl:62-104) is its exact inverse at D = 1, on grids no golden state
  has;
* the PCG (implicit.jl:279-446: exact, fast and Float32-preconditioned) leaves a true residual
  ||b - (D + aK) x|| within its tolerance, equals the separable solve at D = 1, and keeps the
  zero-source rule of gtdgl/test/pcg_zero_fixture.jl (x = 0, warm, no iterations, xprev the last
  real solution, decoupled links scaled by 1/D, bit-equal continuation through a checkpoint);
* _rotfrac!, _rot! roots, _link_rate! and _cis (bands.jl:288-423) meet the 93 manufactured
  rotations of gtdgl/test/rotation_fixture.jl with the bounds of test_link_rotations.jl;
* the anisotropic masses (model.jl:36-44, scheme.jl:121-147) give a plane wave along axis i the
  Euler factor 1 - 2c - dt/(m_i Γ dx²)·λ of the discrete Neumann Laplacian's eigenvalue λ.
"""
import dataclasses
import math
import os
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from juliatdgl_amd.checkpoint import load_state, pcg_extra, restore_pcg, save_state
from juliatdgl_amd.scheme import Band, Grid, Model, box_masks, catalogue_masks, make_stepper, two_band
from juliatdgl_amd.state import initial_state, seeded_state
from oracle.oracle import Oracle, link_rate, rotation, scis


# ------------------------------------------------------------------ the operator, from its definition
def link_shapes(gr):
    return [(gr.Nx, gr.Ny + 1, gr.Nz + 1), (gr.Nx + 1, gr.Ny, gr.Nz + 1), (gr.Nx + 1, gr.Ny + 1, gr.Nz)]


def unknown_masks(gr):
    """Links whose two end nodes are both interior nodes of the box (not on its surface)."""
    nodes = np.zeros((gr.Nx + 1, gr.Ny + 1, gr.Nz + 1), dtype=bool)
    nodes[1:-1, 1:-1, 1:-1] = True
    return [nodes[:-1, :, :] & nodes[1:, :, :], nodes[:, :-1, :] & nodes[:, 1:, :], nodes[:, :, :-1] & nodes[:, :, 1:]]


def range_boxes(gr):
    """The bounding boxes of the unknowns per component (0-based slices), which the oracle's range
    vectors run over with the first index fastest."""
    out = []
    for m in unknown_masks(gr):
        idx = np.nonzero(m)
        box = tuple(slice(int(a.min()), int(a.max()) + 1) for a in idx)
        assert m[box].all() and m.sum() == m[box].size  # a box: the unknowns are a tensor range
        out.append(box)
    return out


def curl_matrix(gr):
    """C: one row per plaquette whose four edges are unknowns, the circulation e_u(r) + e_v(r+u) -
    e_u(r+v) - e_v(r) of the link phases around it; columns are the unknowns, component after
    component, each in its range box in Fortran order."""
    masks = unknown_masks(gr)
    boxes = range_boxes(gr)
    offs, ncol, colid = [], 0, []
    for c in range(3):
        ids = -np.ones(link_shapes(gr)[c], dtype=np.int64)
        sub = ids[boxes[c]]
        sub[...] = ncol + np.arange(sub.size).reshape(sub.shape, order="F")
        ids[boxes[c]] = sub
        colid.append(ids)
        ncol += sub.size
    rows, cols, vals = [], [], []
    nrow = 0
    for n in range(3):
        u, v = (n + 1) % 3, (n + 2) % 3
        # plaquettes normal to n at base node r: r_n in 1..N_n+1, r_u in 1..N_u, r_v in 1..N_v
        N = (gr.Nx, gr.Ny, gr.Nz)
        shape = [N[0] + 1, N[1] + 1, N[2] + 1]
        shape[u] -= 1
        shape[v] -= 1
        base = np.indices(shape).reshape(3, -1)
        eu = np.eye(3, dtype=np.int64)[u][:, None]
        ev = np.eye(3, dtype=np.int64)[v][:, None]
        edges = [(u, base, 1.0), (v, base + eu, 1.0), (u, base + ev, -1.0), (v, base, -1.0)]
        ok = np.ones(base.shape[1], dtype=bool)
        for ax, r, _ in edges:
            sh = link_shapes(gr)[ax]
            inb = (r[0] < sh[0]) & (r[1] < sh[1]) & (r[2] < sh[2])
            rr = np.minimum(r, np.array(sh)[:, None] - 1)
            ok &= inb & masks[ax][rr[0], rr[1], rr[2]]
        sel = np.nonzero(ok)[0]
        for ax, r, sgn in edges:
            rows.append(nrow + np.arange(sel.size))
            cols.append(colid[ax][r[0, sel], r[1, sel], r[2, sel]])
            vals.append(np.full(sel.size, sgn))
        nrow += sel.size
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nrow, ncol))


def operator(st, D=None):
    """(D + aK) with K = CᵀC (dx = dy = dz: the circulations are the curl times dx², a the
    Stepper's (dt/σn) κ²/dx²); D = 1 without a conductivity per link."""
    gr = st.grid
    C = curl_matrix(gr)
    boxes = range_boxes(gr)
    if D is None:
        d = np.ones(C.shape[1])
    else:
        d = np.concatenate([np.asarray(D[c])[boxes[c]].ravel(order="F") for c in range(3)])
    return sp.diags(d) + st.fc.a * (C.T @ C)


def flat(vs):
    return np.concatenate([np.asarray(v).ravel(order="F") for v in vs])


def split(x, gr):
    out, q = [], 0
    for b in range_boxes(gr):
        sh = tuple(s.stop - s.start for s in b)
        n = int(np.prod(sh))
        out.append(x[q:q + n].reshape(sh, order="F"))
        q += n
    return out


def test_operator_is_the_curl_curl_of_a_uniform_field_free_gauge():
    """Sanity of the construction: K annihilates pure gauges (θ = ∇χ with χ zero on the surface)
    and is symmetric positive semidefinite; the numbers of rows and columns are those of the
    plaquette and link sets."""
    gr = Grid.make(5, 4, 3, dx=0.5, buf=1)
    C = curl_matrix(gr)
    chi = np.zeros((gr.Nx + 1, gr.Ny + 1, gr.Nz + 1))
    rng = np.random.default_rng(3)
    chi[1:-1, 1:-1, 1:-1] = rng.standard_normal((gr.Nx - 1, gr.Ny - 1, gr.Nz - 1))
    grad = [np.diff(chi, axis=0), np.diff(chi, axis=1), np.diff(chi, axis=2)]
    boxes = range_boxes(gr)
    g = np.concatenate([grad[c][boxes[c]].ravel(order="F") for c in range(3)])
    assert np.abs(C @ g).max() <= 1e-13
    K = (C.T @ C).toarray()
    assert np.array_equal(K, K.T)
    assert np.linalg.eigvalsh(K).min() > -1e-10


# ------------------------------------------------------------------ the separable solve
@pytest.mark.parametrize("dims,buf", [((6, 4, 3), 1), ((7, 5, 4), 2), ((9, 3, 2), 1)])
def test_separable_solve_inverts_I_plus_aK(dims, buf):
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(*dims, dx=0.5, buf=buf)
    st = make_stepper(md, gr, box_masks(gr))
    assert st.field_mode == 1
    A = operator(st)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(A.shape[0])
    x = flat(Oracle(st).sep_solve(*split(b, gr)))
    res = np.abs(A @ x - b).max()
    assert res <= 1e-12 * np.abs(b).max(), res
    y = spla.spsolve(A.tocsc(), b)
    assert np.abs(y - x).max() <= 1e-10 * np.abs(x).max()


# ------------------------------------------------------------------ the PCG
def pcg_case(mode, tol=1e-11, dims=(8, 5, 4), buf=2):
    fast, f32, extrap = mode
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(*dims, dx=0.5, buf=buf)
    st = make_stepper(md, gr, box_masks(gr), sigma_vac=0.01, pcg_tol=tol, pcg_fast=fast, pcg_f32=f32,
                      pcg_extrap=extrap)
    assert st.field_mode == 2
    return st


def random_links(st, seed, amp_b=1e-3):
    """θ random, φ = θ + a random increment (a right-hand side on every link)."""
    gr = st.grid
    s = initial_state(st.model, gr, st.masks, 0.0)
    rng = np.random.default_rng(seed)
    for c, (t, p) in enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz"))):
        th = 0.01 * rng.standard_normal(getattr(s, t).shape)
        getattr(s, t)[...] = th
        getattr(s, p)[...] = th + amp_b * rng.standard_normal(th.shape)
    return s


MODES = [(False, False, True), (True, False, True), (True, True, True), (True, True, False)]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: "fast%d_f32%d_extrap%d" % m)
def test_pcg_true_residual_within_tolerance(mode):
    """After the oracle's PCG, the true residual of (D + aK)Δ = φₑ − θ (built here) is within
    the tolerance, cold and warm (three solves in a row with a changing source), and the
    decoupled links take Δ = (φₑ − θ)/D."""
    tol = 1e-11
    st = pcg_case(mode, tol)
    gr = st.grid
    A = operator(st, st.D)
    boxes = range_boxes(gr)
    o = Oracle(st)
    for q in range(3):
        s = random_links(st, 100 + q)
        o.set_state(s)
        o.field_solve()
        out = o.get_state()
        b = flat([(getattr(s, p) - getattr(s, t))[boxes[c]] for c, (t, p) in
                  enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz")))])
        x = flat([(getattr(out, p) - getattr(out, t))[boxes[c]] for c, (t, p) in
                  enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz")))])
        res = np.linalg.norm(b - A @ x)
        assert res <= 1.5 * tol * np.linalg.norm(b), (q, res / np.linalg.norm(b))
        # the solution, not just a small residual: against a direct solve
        xd = spla.spsolve(A.tocsc(), b)
        assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max()
        pc = o.get_pcg()
        assert pc.warm and pc.iters >= 1
        assert np.array_equal(flat(pc.x), x + 0.0) or np.abs(flat(pc.x) - x).max() <= 1e-15 * np.abs(x).max() + 1e-18
        for c, (t, p) in enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz"))):
            dec = np.ones(getattr(s, t).shape, dtype=bool)
            dec[boxes[c]] = False
            want = getattr(s, t) + (getattr(s, p) - getattr(s, t)) / st.D[c]
            got = getattr(out, p)
            assert np.abs(got[dec] - want[dec]).max() <= 1e-15
            # the coupled links are not the decoupled formula (the planted mistake of scaling all)
        assert np.abs(A @ x - b).max() < np.abs(A @ (b / np.concatenate([st.D[c][boxes[c]].ravel(order="F")
                                                                          for c in range(3)])) - b).max()


@pytest.mark.parametrize("mode", MODES, ids=lambda m: "fast%d_f32%d_extrap%d" % m)
def test_pcg_at_unit_conductivity_is_the_separable_solve(mode):
    """With D ≡ 1 the PCG solves (I + aK)Δ = b: the separable solve's answer (golden-pinned by the
    reference's prod states), to the tolerance."""
    st = pcg_case(mode, 1e-12)
    gr = st.grid
    ones = tuple(np.ones(sh, order="F") for sh in link_shapes(gr))
    st1 = dataclasses.replace(st, D=ones)
    sep = make_stepper(st.model, gr, st.masks)
    boxes = range_boxes(gr)
    s = random_links(st1, 7)
    o = Oracle(st1)
    o.set_state(s)
    o.field_solve()
    out = o.get_state()
    b = [(getattr(s, p) - getattr(s, t))[boxes[c]] for c, (t, p) in enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz")))]
    x = flat([(getattr(out, p) - getattr(out, t))[boxes[c]] for c, (t, p) in
              enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz")))])
    xs = flat(Oracle(sep).sep_solve(*b))
    assert np.abs(x - xs).max() <= 1e-11 * np.abs(xs).max()
    # the cold start is the separable solve itself: no iteration is needed
    assert o.get_pcg().iters <= 1


# ------------------------------------------------------------------ the zero-source rule (pcg_zero_fixture.jl)
def _ranges(st):
    return range_boxes(st.grid)


def _source(st, s, q):
    """source!(s, q) of pcg_zero_fixture.jl:21-30: θ = 0.01 sin(0.13 i + c) over the linear
    index, φ = θ + 0.003 on every link, and on the coupled links φ = θ + 0.001 sin(0.17 n + c + q)
    (n their running index), or φ = θ for q = 0."""
    for c, (t, p) in enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz"))):
        T, P = getattr(s, t), getattr(s, p)
        i = np.arange(1, T.size + 1)
        T[...] = (0.01 * np.sin(0.13 * i + (c + 1))).reshape(T.shape, order="F")
        P[...] = T + 0.003
        box = _ranges(st)[c]
        th = T[box]
        n = np.arange(1, th.size + 1).reshape(th.shape, order="F")
        P[box] = th + (0.0 if q == 0 else 0.001 * np.sin(0.17 * n + (c + 1) + q))


def _solve(o, s):
    o.set_state(s)
    o.field_solve()
    out = o.get_state()
    for k in ("phx", "phy", "phz"):
        getattr(s, k)[...] = getattr(out, k)


def _check_zero(st, o, s, cold=False):
    fast, extrap = st.pcg_fast or st.pcg_f32, st.pcg_extrap and st.pcg_fast
    previous = [x.copy() for x in o.get_pcg().x]
    expected = [getattr(s, t) + (getattr(s, p) - getattr(s, t)) / st.D[c]
                for c, (t, p) in enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz")))]
    _solve(o, s)
    pc = o.get_pcg()
    assert all(not x.any() for x in pc.x)
    assert pc.warm and pc.iters == 0
    for c, (t, p) in enumerate(zip(("thx", "thy", "thz"), ("phx", "phy", "phz"))):
        box = _ranges(st)[c]
        assert np.array_equal(getattr(s, p)[box], getattr(s, t)[box])
        expected[c][box] = getattr(s, t)[box]
        assert np.array_equal(getattr(s, p), expected[c])
    if fast:
        assert pc.have2 == ((not cold) and extrap)
        if not cold and extrap:
            assert all(np.array_equal(a, b) for a, b in zip(pc.xprev, previous))


def _snap(o):
    p = o.get_pcg()
    return (p.warm, [x.copy() for x in p.x], p.have2, [x.copy() for x in p.xprev] if p.have2 else None)


def _snap_equal(a, b):
    return (a[0] == b[0] and all(np.array_equal(x, y) for x, y in zip(a[1], b[1])) and a[2] == b[2]
            and (a[3] is None) == (b[3] is None) and (a[3] is None or all(np.array_equal(x, y) for x, y in zip(a[3], b[3]))))


@pytest.mark.parametrize("mode", MODES, ids=lambda m: "fast%d_f32%d_extrap%d" % m)
def test_pcg_zero_source_and_checkpoint(mode):
    """pcg_zero_rhs_checks (gtdgl/test/pcg_zero_fixture.jl:5-82) on the oracle: two_band(ξ ratio 4,
    κ 2) on a 4x4x4 box with buf 2, SIM masks, σvac = 0.01."""
    fast, f32, extrap = mode
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(4, 4, 4, dx=0.5, buf=2)
    mk = catalogue_masks(gr, "SIM")

    def newstepper():
        return make_stepper(md, gr, mk, sigma_vac=0.01, pcg_fast=fast, pcg_f32=f32, pcg_extrap=extrap)

    st = newstepper()
    o = Oracle(st)
    s = initial_state(md, gr, mk, 0.0)
    _source(st, s, 0)
    _check_zero(st, o, s, cold=True)
    for q in (1, 2, 3):
        _source(st, s, q)
        _solve(o, s)
        assert any(x.any() for x in o.get_pcg().x)
    _source(st, s, 0)
    _check_zero(st, o, s)
    # the ladder checkpoint's keys (pcg_x*, and pcg_p* with a second solution)
    extra = pcg_extra(o)
    assert ("pcg_p1" in extra) == (bool(fast) and o.get_pcg().have2)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "zero.npz")
        save_state(path, s, **extra)
        resumed, raw = load_state(path, gr)
    o2 = Oracle(newstepper())
    restore_pcg(o2, raw)
    for q in (0, 4, 5):
        _source(st, s, q)
        _source(st, resumed, q)
        _solve(o, s)
        _solve(o2, resumed)
        for k in ("phx", "phy", "phz"):
            assert np.array_equal(getattr(s, k), getattr(resumed, k))
        assert _snap_equal(_snap(o), _snap(o2))
    _source(st, s, 0)
    _check_zero(st, o, s)
    _source(st, s, 0)
    _check_zero(st, o, s)  # the previous solution is now zero too
    p = o.get_pcg()  # _pcg_reset!: a cold start
    p.warm, p.have2 = False, False
    o.set_pcg(p)
    _source(st, s, 0)
    _check_zero(st, o, s, cold=True)


# ------------------------------------------------------------------ temporal rotations (rotation_fixture.jl)
def rotation_inputs():
    angles = [0.0, math.pi, -math.pi, 0.02, -0.02, 0.4, -0.4]
    for t in (0.05, math.asin(0.05), math.pi - 0.05, math.pi - math.asin(0.05)):
        for s in (-1, 1):
            for d in (-1e-9, 0.0, 1e-9):
                angles.append(s * t + d)
    angles = sorted(set(angles))
    ph = [a for _ in range(3) for a in angles]
    scale = [r for r in (1.0, 1 - 2.6e-7, 1 + 2.6e-7) for _ in angles]
    n = len(ph)
    ph.append(ph[-1])
    scale.append(scale[-1])
    V = np.ones((n + 1, 2, 2), dtype=complex)
    U = np.empty((n + 1, 2, 2), dtype=complex)
    for i in range(n + 1):
        U[i] = scale[i] * complex(math.cos(-ph[i]), math.sin(-ph[i]))
    omega = np.array([[[(i + 1) / 64 + (j + 1) / 32 - (k + 1) / 16 for k in range(2)] for j in range(2)]
                      for i in range(n + 1)])
    phase = np.angle(V * np.conj(U))
    return V, U, omega, phase, n


def test_rotation_fixture_has_the_reference_count():
    V, U, omega, phase, n = rotation_inputs()
    assert n == 93


@pytest.mark.parametrize("c", [0.25, 0.5, 0.9, 1.0])
def test_rotfrac_against_cis(c):
    V, U, _, phase, _ = rotation_inputs()
    L = np.array([rotation(0, v, u, c) for v, u in zip(V.ravel(), U.ravel())])
    ref = V.ravel() * np.exp(1j * c * phase.ravel())
    large = np.abs(phase.ravel()) > 0.1
    assert np.abs(L - ref).max() < 4e-7
    assert np.abs(L[large] - ref[large]).max() < 2e-14
    assert np.abs(np.abs(L) - 1).max() < 3e-7


@pytest.mark.parametrize("n", [2, 5, 9])
def test_rot_roots_against_cis(n):
    V, U, _, phase, _ = rotation_inputs()
    R = np.array([rotation(1, v, u, 0.0, n) for v, u in zip(V.ravel(), U.ravel())])
    ref = np.exp(1j * phase.ravel() / n)
    large = np.abs(phase.ravel()) > 0.1
    assert np.abs(R - ref).max() < 4e-8
    assert np.abs(R[large] - ref[large]).max() < 2e-14


def test_link_rate_and_cis_series():
    V, U, omega, phase, n = rotation_inputs()
    dt = 1 / 32
    for axis, shape in enumerate([(n, 2, 2), (n + 1, 1, 2), (n + 1, 2, 1)]):
        off = [0, 0, 0]
        off[axis] = 1
        errs, big = [], []
        for I in np.ndindex(*shape):
            J = tuple(a + b for a, b in zip(I, off))
            got = link_rate(V[I], U[I], omega[I], omega[J], dt)
            want = phase[I] + (omega[J] - omega[I]) * dt
            errs.append(abs(got - want))
            if abs(phase[I]) > 0.1:
                big.append(abs(got - want))
        assert max(errs) < 4e-8
        assert max(big) < 2e-14
    assert abs(scis(0.05) - complex(math.cos(0.05), math.sin(0.05))) < 2.605e-7


# ------------------------------------------------------------------ anisotropic masses
def plane_wave_factor(md, gr, axis, m, dt, delta=1e-7):
    """The Euler factor of one step on Ψ = 1 + δ·cos(π m (p − ½)/n) along `axis` (p = 1..n over the
    block's interior nodes, the ghosts copying their neighbour: the discrete Neumann mode)."""
    st = make_stepper(md, gr, box_masks(gr), rkl=False, dt=dt)
    assert st.nsub == (1,) and st.rkl == (0,)
    s = initial_state(md, gr, st.masks, 0.0)
    lo = (gr.x1, gr.y1, gr.z1)[axis]   # first interior node, 0-based = (x1 + 1) - 1
    hi = (gr.x2, gr.y2, gr.z2)[axis]
    nint = hi - lo
    prof = np.zeros(s.G[0].shape[axis])
    p = np.arange(1, nint + 1)
    prof[lo:hi] = np.cos(math.pi * m * (p - 0.5) / nint)
    prof[lo - 1] = prof[lo]
    prof[hi] = prof[hi - 1]
    shp = [1, 1, 1]
    shp[axis] = -1
    mask = st.masks.tpsi
    G = mask * (1.0 + delta * prof.reshape(shp))
    s.G[0][...] = G
    s.F[0][...] = G
    o = Oracle(st)
    o.set_state(s)
    o.step(1, 0.0, 0.0)
    out = o.get_state().G[0]
    sl = tuple(slice(a, b + 1) for a, b in ((gr.x1, gr.x2 - 1), (gr.y1, gr.y2 - 1), (gr.z1, gr.z2 - 1)))
    d0 = (G[sl] - 1).real
    d1 = (out[sl] - 1).real
    big = np.abs(d0) > 0.3 * delta
    g = d1[big] / d0[big]
    return float(g.mean()), float(g.std()), 2 - 2 * math.cos(math.pi * m / nint), st


@pytest.mark.parametrize("m", [1, 2])
def test_anisotropic_masses_plane_wave_rates(m):
    """One band, m_axes = (0.8, 0.8, 3.2): the excess of the Euler decay over the uniform 2c is
    dt/(m_i Γ dx²)·λ along each axis, so z relaxes 4x slower than x (m_z/m_xy = 4)."""
    ma = (0.8, 0.8, 3.2)
    md = Model.make((Band.make(a=1.0, m_axes=ma),), [[0.0]], kappa=2.0)
    gr = Grid.make(8, 8, 8, dx=0.5, buf=1)
    dt = 0.002
    rate = {}
    for axis in (0, 1, 2):
        g, sd, lam, st = plane_wave_factor(md, gr, axis, m, dt)
        c = dt / md.bands[0].Gamma
        assert sd <= 1e-5 * (1 - g)          # one mode: the same factor on every node
        excess = 1 - g - 2 * c
        want = dt / (ma[axis] * md.bands[0].Gamma * gr.dx ** 2) * lam
        assert abs(excess - want) <= 1e-5 * want, (axis, excess, want)
        rate[axis] = excess
    assert abs(rate[0] / rate[2] - 4.0) <= 1e-4
    assert abs(rate[0] / rate[1] - 1.0) <= 1e-6
