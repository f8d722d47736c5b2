"""The oracle's thermal noise (oracle/oracle.cpp: noise_fill .. field_kick) against what fixes it:
the Philox4x32-10 known-answer vectors (Random123's kat_vectors), the exact equilibrium of
decoupled nodes and its relaxation rate (gtdgl/test/test_physics.jl:475-497), the same with KWT
(:499-535), and the Stepper's rules for noise (scheme.jl:336-360, 424; test_physics.jl:737-747).
The layout of the caller's draws (juliatdgl_amd.noise) against the reference's loop order."""
import math

import numpy as np
import pytest

from juliatdgl_amd.noise import from_stream, kicked_nodes, pack, per_step
from juliatdgl_amd.scheme import Band, Grid, Model, box_masks, make_stepper, two_band
from juliatdgl_amd.state import initial_state


def test_philox_known_answers():
    from oracle.oracle import philox
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert philox([f, f, f, f], [f, f]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normals_are_standard_and_independent():
    from oracle.oracle import normal2
    z = np.array([normal2(7, t, tag, q) for t in range(40) for tag in (0, 1, 16) for q in range(250)])
    a, b = z[:, 0], z[:, 1]
    for x in (a, b):
        assert abs(x.mean()) < 0.02 and abs(x.var() - 1) < 0.03
        assert abs(np.mean(x ** 4) - 3) < 0.2
    assert abs(np.mean(a * b)) < 0.02
    # another seed, another stream
    assert normal2(8, 0, 0, 0) != normal2(7, 0, 0, 0) and normal2(7, 1, 0, 0) != normal2(7, 0, 0, 0)


def test_stream_layout_follows_the_reference_loop_order():
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(6, 4, 4, dx=0.5)
    mk = box_masks(gr)
    mk.f[gr.x1 + 1, gr.y1 + 1, gr.z1 + 1] = 0.0       # one hard node: no draws there
    st = make_stepper(md, gr, mk, hard=True, noise=1e-3, noise_field=True, noise_draws="host")
    m = kicked_nodes(st)
    assert not m[gr.x1 + 1, gr.y1 + 1, gr.z1 + 1] and m.sum() == (gr.x2 - gr.x1) * (gr.y2 - gr.y1) * (gr.z2 - gr.z1) - 1
    n = per_step(st)
    z = np.arange(n, dtype=float)
    xi, links = from_stream(st, z)
    # band 1, first interior node (x1+1, y1+1, z1+1) 1-based, takes the first two numbers
    assert xi[0][gr.x1, gr.y1, gr.z1] == 0 + 1j
    assert xi[0][gr.x1 + 1, gr.y1, gr.z1] == 2 + 3j                # i fastest
    assert xi[1][gr.x1, gr.y1, gr.z1] == 2 * m.sum() + (2 * m.sum() + 1) * 1j
    nb = 4 * m.sum()
    assert links[0][0, 1, 1] == nb and links[0][1, 1, 1] == nb + 1   # x links: i = 1.. first
    assert links[2][gr.Nx - 1, gr.Ny - 1, gr.Nz - 1] == n - 1          # the last z link is the last
    flat = pack(st, [(xi, links)])
    assert flat.size == st.N * 2 * m.size + sum(a.size for a in links)


def test_host_draws_equal_the_generator_draws():
    """The caller's draws taken as the device's generator would make them give the same run, for
    the SK-ROCK path (KWT on RKC stages) with the field kick."""
    from oracle.oracle import Oracle, normal2
    md = two_band(xi_ratio=4.0, kappa=2.0, gamma_kwt=10.0)
    gr = Grid.make(8, 6, 4, dx=0.5, buf=2)
    mk = box_masks(gr)
    kw = dict(sigma_vac=0.01, noise=1e-3, noise_seed=5, noise_field=True)
    st_d = make_stepper(md, gr, mk, **kw)
    st_h = make_stepper(md, gr, mk, noise_draws="host", **kw)
    assert any(st_d.sk) and st_d.rkc is not None
    s0 = initial_state(md, gr, mk, 0.03)
    a = Oracle(st_d)
    a.set_state(s0)
    a.step(3, 0.03, st_d.selffield(0.1))
    b = Oracle(st_h)
    b.set_state(s0)
    steps = []
    for t in range(3):
        xi = []
        for mu in range(st_d.N):
            x = np.zeros((gr.Nx + 1, gr.Ny + 1, gr.Nz + 1), dtype=complex, order="F")
            for (i, j, k) in zip(*np.nonzero(kicked_nodes(st_d))):
                q = int(i + (gr.Nx + 1) * (j + (gr.Ny + 1) * k))
                r, im = normal2(5, t, mu, q)
                x[i, j, k] = r + 1j * im
            xi.append(x)
        links = []
        for c, shp in enumerate(((gr.Nx, gr.Ny + 1, gr.Nz + 1), (gr.Nx + 1, gr.Ny, gr.Nz + 1), (gr.Nx + 1, gr.Ny + 1, gr.Nz))):
            y = np.zeros(shp, order="F")
            for idx in np.ndindex(*shp):
                q = int(idx[0] + shp[0] * (idx[1] + shp[1] * idx[2]))
                y[idx] = normal2(5, t, 16 + c, q)[0]
            links.append(y)
        steps.append((xi, links))
    b.noise_draws(pack(st_h, steps))
    b.step(3, 0.03, st_h.selffield(0.1))
    with pytest.raises(ValueError, match="draws for 0 more steps"):
        b.step(1, 0.03, st_h.selffield(0.1))
    sa, sb = a.get_state(), b.get_state()
    for mu in range(st_d.N):
        assert np.array_equal(sa.G[mu], sb.G[mu])
    assert np.array_equal(sa.hz, sb.hz) and np.array_equal(sa.thx, sb.thx)
    assert a.noise_counter == b.noise_counter == 3
    # and the noise did something
    a0 = Oracle(make_stepper(md, gr, mk, sigma_vac=0.01))
    a0.set_state(s0)
    a0.step(3, 0.03, st_d.selffield(0.1))
    assert np.max(np.abs(a0.get_state().G[0] - sa.G[0])) > 1e-4


def decoupled_equilibrium(backend_of, steps_of):
    """test_physics.jl:478-497 on a backend: |Ψ|² of decoupled nodes (m = 10⁸) is Gaussian with
    mean ε/β = 1 and variance T̃/(dV β) (the Euler–Maruyama bias 1/(1 − ε dt/Γ)); its correlation
    at lag Γ/(4ε) is e^{−1/2}."""
    G, T = 2.0, 0.01
    md = Model.make((Band.make(m=1e8, Gamma=G),), np.zeros((1, 1)), kappa=2.0)
    gr = Grid.make(10, 10, 2, dx=0.5)
    mk = box_masks(gr)
    st = make_stepper(md, gr, mk, dt_scale=0.02, noise=T, noise_seed=11)
    b = backend_of(st, initial_state(md, gr, mk, 0.0))
    run, get = steps_of(b)
    dV = gr.dx * gr.dy * gr.dz
    lag = round(G / 4 / st.dt)
    run(round(20 / st.dt))
    sl = (slice(gr.x1, gr.x2), slice(gr.y1, gr.y2), slice(gr.z1, gr.z2))
    n = lambda: (np.abs(get()[sl]) ** 2).ravel()
    sm = s2 = c = 0.0
    N = 0
    prev = n()
    for _ in range(round(1500 / (lag * st.dt))):
        run(lag)
        x = n()
        sm += x.sum()
        s2 += (x * x).sum()
        c += ((x - 1) * (prev - 1)).sum()
        N += x.size
        prev = x
    mean = sm / N
    var = s2 / N - mean ** 2
    rho = (c / N) / var
    assert mean == pytest.approx(1.0, abs=0.01)
    assert var == pytest.approx(T / dV / (1 - st.dt / G), rel=0.04)
    assert rho == pytest.approx(math.exp(-0.5), abs=0.03)


def test_stepper_rules_for_noise():
    """test_physics.jl:737-747 and scheme.jl:424: KWT on RKL stages with noise is refused, RKC is
    taken by a KWT band with noise (SK-ROCK), the Euler substeps run; noise_field needs noise."""
    md = two_band(xi_ratio=4.0, kappa=2.0, gamma_kwt=10.0)
    gr = Grid.make(12, 8, 4, dx=0.5)
    mk = box_masks(gr)
    with pytest.raises(ValueError, match="RKL stages"):
        make_stepper(md, gr, mk, noise=1e-3, rkl=True)
    st = make_stepper(md, gr, mk, noise=1e-3)
    assert st.rkc is not None and all(st.sk[mu] == (st.rkl[mu] > 0) for mu in range(st.N))
    with pytest.warns(UserWarning, match="Euler substeps"):
        st = make_stepper(md, gr, mk, noise=1e-3, rkl=False)
    assert all(s == 0 for s in st.rkl)
    assert make_stepper(md, gr, mk, rkl=True).rkc is None
    with pytest.raises(ValueError, match="noise_field needs noise"):
        make_stepper(two_band(xi_ratio=4.0, kappa=2.0), gr, mk, noise_field=True)
    with pytest.raises(ValueError, match=">= 0"):
        make_stepper(two_band(xi_ratio=4.0, kappa=2.0), gr, mk, noise=-1.0)


def kwt_equilibrium_and_rates(backend_of, steps_of):
    """test_physics.jl:503-535 on a backend: decoupled nodes with γ̃ = 3 (s = √10) keep the mean
    and variance of |Ψ|² of exp(−F/T̃); the phase diffuses s times faster (⟨cos Δφ⟩ at its lag is
    e^{−1/2}) and the amplitude relaxes s times slower (the correlation of |Ψ|² at its lag)."""
    G, T, gk = 2.0, 0.01, 3.0
    sk = math.sqrt(1 + gk * gk)
    md = Model.make((Band.make(m=1e8, Gamma=G, gkwt=gk),), np.zeros((1, 1)), kappa=2.0)
    gr = Grid.make(10, 10, 2, dx=0.5)
    mk = box_masks(gr)
    st = make_stepper(md, gr, mk, dt_scale=0.02, noise=T, noise_seed=5)
    b = backend_of(st, initial_state(md, gr, mk, 0.0))
    dV = gr.dx * gr.dy * gr.dz
    run, get = steps_of(b)
    run(round(40 / st.dt))
    sl = (slice(gr.x1, gr.x2), slice(gr.y1, gr.y2), slice(gr.z1, gr.z2))
    lagn = round(G * sk / 4 / st.dt)
    lagp = round(2 * G * dV / (T * sk) * 0.5 / st.dt)
    sm = s2 = cp = 0.0
    N = Np = 0
    prev = get()[sl].copy()
    since = 0
    total = round(3000 / st.dt)
    q = 0
    while q < total:
        # the next event: a sample of |Ψ|² every lagn steps, of the phase every lagp steps
        nxt = min(lagn - since % lagn, lagp - since, total - q)
        run(nxt)
        q += nxt
        since += nxt
        g = None
        if since % lagn == 0:
            g = get()[sl]
            x = np.abs(g) ** 2
            sm += x.sum()
            s2 += (x * x).sum()
            N += x.size
        if since == lagp:
            g = get()[sl] if g is None else g
            cp += np.sum(np.real(g * np.conj(prev)) / (np.abs(g) * np.abs(prev)))
            Np += g.size
            prev = g.copy()
            since = 0
    mean = sm / N
    var = s2 / N - mean ** 2
    assert mean == pytest.approx(1.0, abs=0.01)
    assert var == pytest.approx(T / dV, rel=0.06)
    assert cp / Np == pytest.approx(math.exp(-0.5), abs=0.03)
    xs, ys = [], []
    for _ in range(400):
        a = np.abs(get()[sl]) ** 2
        run(lagn)
        c = np.abs(get()[sl]) ** 2
        xs.append(a.ravel() - 1)
        ys.append(c.ravel() - 1)
    xs, ys = np.concatenate(xs), np.concatenate(ys)
    assert np.sum(xs * ys) / np.sum(xs * xs) == pytest.approx(math.exp(-0.5), abs=0.04)


def curl_matrix(gr, L, P):
    """curl_h! (gtdgl/src/field.jl:124-137) as a matrix from the link phases L to the plaquettes P."""
    li = {l: c for c, l in enumerate(L)}
    Cm = np.zeros((len(P), len(L)))
    dx, dy, dz = gr.dx, gr.dy, gr.dz
    for r, (f, i, j, k) in enumerate(P):
        if f == 1:
            terms, den = (((3, i, j + 1, k), 1), ((3, i, j, k), -1), ((2, i, j, k + 1), -1), ((2, i, j, k), 1)), dy * dz
        elif f == 2:
            terms, den = (((1, i, j, k + 1), 1), ((1, i, j, k), -1), ((3, i + 1, j, k), -1), ((3, i, j, k), 1)), dz * dx
        else:
            terms, den = (((2, i + 1, j, k), 1), ((2, i, j, k), -1), ((1, i, j + 1, k), -1), ((1, i, j, k), 1)), dx * dy
        for l, sgn in terms:
            if l in li:
                Cm[r, li[l]] += sgn / den
    return Cm


def field_equipartition(backend_of, steps_of, kappa, svac, buf, teq, tav):
    """test_physics.jl:701-733 on a backend: in a normal box at H = 0 the field energy
    F_A = κ²dV Σ h² of the free plaquettes averages Σ_m (T̃/2)/(1 + λ_m dt/2) over the transverse
    modes of D^{-1/2}CᵀC D^{-1/2} (the implicit Euler of the field).  Returns the ratio and max|Ψ|."""
    md = Model.make((Band.make(m=1.0, Gamma=1.0),), np.zeros((1, 1)), kappa=kappa)
    gr = Grid.make(6, 4, 4, dx=0.5, buf=buf)
    mk = box_masks(gr)
    mk.f[...] = 0.0
    T = 0.02
    st = make_stepper(md, gr, mk, hard=True, noise=T, noise_seed=3, noise_field=True, sigma_vac=svac)
    s = initial_state(md, gr, mk, 0.0)
    s.G[0][...] = 0
    s.F[0][...] = 0
    Nx, Ny, Nz, dx = gr.Nx, gr.Ny, gr.Nz, gr.dx
    L = ([(1, i, j, k) for k in range(2, Nz + 1) for j in range(2, Ny + 1) for i in range(1, Nx + 1)] +
         [(2, i, j, k) for k in range(2, Nz + 1) for j in range(1, Ny + 1) for i in range(2, Nx + 1)] +
         [(3, i, j, k) for k in range(1, Nz + 1) for j in range(2, Ny + 1) for i in range(2, Nx + 1)])
    P = ([(1, i, j, k) for k in range(2, Nz) for j in range(2, Ny) for i in range(2, Nx + 1)] +
         [(2, i, j, k) for k in range(2, Nz) for j in range(2, Ny + 1) for i in range(2, Nx)] +
         [(3, i, j, k) for k in range(2, Nz + 1) for j in range(2, Ny) for i in range(2, Nx)])
    Cm = curl_matrix(gr, L, P)
    Dl = np.ones(len(L)) if st.D is None else np.array([st.D[f - 1][i - 1, j - 1, k - 1] for f, i, j, k in L])
    Ds = 1 / np.sqrt(Dl)
    ev = np.linalg.eigvalsh((Ds[:, None] * (Cm.T @ Cm)) * Ds[None, :])
    ev = ev[ev > 1e-9 * ev.max()]
    expect = np.sum((T / 2) / (1 + (md.kappa ** 2 * dx ** 2 / md.sigma_n) * ev * (st.dt / 2)))
    pi = [np.array([p[1] - 1 for p in P if p[0] == f]) for f in (1, 2, 3)]
    pj = [np.array([p[2] - 1 for p in P if p[0] == f]) for f in (1, 2, 3)]
    pk = [np.array([p[3] - 1 for p in P if p[0] == f]) for f in (1, 2, 3)]
    b = backend_of(st, s)
    run, get_h = steps_of(b, field=True)

    def FA():
        hs = get_h()
        return md.kappa ** 2 * dx ** 3 * sum(np.sum(hs[f][pi[f], pj[f], pk[f]] ** 2) for f in range(3))

    run(round(teq / st.dt))
    acc, n = 0.0, 0
    # FA after every 10th step of the averaging time (test_physics.jl:725-728)
    nav = round(tav / st.dt)
    for _ in range(nav // 10):
        run(10)
        acc += FA()
        n += 1
    run(nav % 10)
    return acc / n / expect


def _oracle_backend(st, s):
    from oracle.oracle import Oracle
    o = Oracle(st)
    o.set_state(s)
    return o


def _oracle_steps(o, field=False):
    run = lambda n: o.step(n, 0.0, 0.0)
    if field:
        def get_h():
            s = o.get_state()
            return (s.hx, s.hy, s.hz)
        return run, get_h
    return run, lambda: o.get_state().G[0]


def test_equilibrium_of_decoupled_nodes():
    decoupled_equilibrium(_oracle_backend, _oracle_steps)


def test_kwt_noise_equilibrium_and_rates():
    kwt_equilibrium_and_rates(_oracle_backend, _oracle_steps)


def test_field_noise_equipartition_stiff_modes():
    """The stiff variant of test_physics.jl:730 (κ = 1, λdt up to 0.87), on the oracle."""
    r = field_equipartition(_oracle_backend, _oracle_steps, 1.0, None, 1, 100.0, 1000.0)
    assert r == pytest.approx(1.0, abs=0.01)
