"""The oracle's normal fluids (qp_exchange, gtdgl/src/exchange.jl) against what pins them: the
reference's own checks of test_physics.jl:338-432 (the δ solve satisfies its discrete equation,
a fast exchange is one fluid, two bands written as a matrix are the two-band model, N potentials
keep the gauge Σ s_µ μ_µ = 0 and solve their coupled equations), and the conversion C_µ against
its closed form on a manufactured state (a phase that varies along every axis, unit links)."""
import numpy as np
import pytest

from juliatdgl_amd.scheme import Band, Grid, Model, box_masks, make_stepper, wedge_masks
from juliatdgl_amd.state import initial_state


def _model(kappa=25.0, gamma12=1e-3, gkwt=0.0):
    """two_band(xi_ratio=4, κ) of test_physics.jl:341 (a2n = 0.59)."""
    R, a = 4.0, 0.59
    m = 1 / (R * R * a)
    return Model.make((Band.make(gkwt=gkwt), Band.make(m=m, a=a, gkwt=gkwt)), [[0, gamma12], [gamma12, 0]],
                      kappa=kappa)


def _three(kappa=25.0):
    """The frustrated three-band model of test_physics.jl:399-400."""
    return Model.make((Band.make(), Band.make(m=2.0, a=0.8), Band.make(m=0.5, Gamma=1.5, a=0.6)),
                      [[0, -0.05, -0.05], [-0.05, 0, -0.05], [-0.05, -0.05, 0]], kappa=kappa)


def _warm(md, gr, mk, nsteps=300, H=0.03, J=0.5, **kw):
    """initial_state and `nsteps` one-fluid oracle steps (test_physics.jl:343-345)."""
    from oracle.oracle import Oracle
    st0 = make_stepper(md, gr, mk, **kw)
    HJ = st0.selffield(J)
    o = Oracle(st0)
    o.set_state(initial_state(md, gr, mk, H))
    o.step(nsteps, H, HJ)
    return o.get_state(), st0, H, HJ


def _node(a, gr, j0, k0):
    """A region array (Nx-1, nj, nk) on the node grid (zero on the x faces and off the region)."""
    full = np.zeros((gr.Nx + 1, gr.Ny + 1, gr.Nz + 1))
    n1, nj, nk = a.shape
    full[1:gr.Nx, j0:j0 + nj, k0:k0 + nk] = a
    return full


def _residual(P, S, gr, sig, lam_term, j0, nj, k0, nk):
    """max |σn ∇²P − λ(P) − S| on the unknowns, mirror ghosts on the y and z faces of the region,
    P = 0 on the x faces (the check of test_physics.jl:349-360)."""
    dx = gr.dx
    res = 0.0
    for k in range(k0, k0 + nk):
        for j in range(j0, j0 + nj):
            jm = j0 + 1 if j == j0 else j - 1
            jp = j0 + nj - 2 if j == j0 + nj - 1 else j + 1
            km = k0 + 1 if k == k0 else k - 1
            kp = k0 + nk - 2 if k == k0 + nk - 1 else k + 1
            i = np.arange(1, gr.Nx)
            lap = (P[i + 1, j, k] + P[i - 1, j, k] + P[i, jp, k] + P[i, jm, k] + P[i, j, kp] + P[i, j, km] -
                   6 * P[i, j, k]) / dx ** 2
            res = max(res, float(np.max(np.abs(sig * lap - lam_term(i, j, k) - S[i, j, k]))))
    return res


@pytest.mark.parametrize("pad", [False, True])
def test_delta_solves_its_discrete_equation(pad):
    from oracle.oracle import Oracle
    md = _model()
    gr = Grid.make(24, 12, 4, dx=0.5, buf=2 if pad else 1)
    mk = wedge_masks(gr)
    kw = dict(sigma_vac=0.01) if pad else {}
    s0, st0, H, HJ = _warm(md, gr, mk, 200, **kw)
    st1 = make_stepper(md, gr, mk, qp_exchange=1.0, **kw)
    q = st1.qp
    # the shares and the conductance of test_physics.jl:347
    assert q.s[0] == pytest.approx(1 / (1 + 16 * 0.59)) and q.G == pytest.approx(q.s[0] * 1.0) and st1.dt == st0.dt
    o = Oracle(st1)
    o.set_state(s0)
    (d,), (S,) = o.qp_potential()
    delta, Sn = _node(d, gr, q.j0, q.k0), _node(S, gr, q.j0, q.k0)
    res = _residual(delta, Sn, gr, md.sigma_n, lambda i, j, k: q.lam[0] * delta[i, j, k], q.j0, q.nj, q.k0, q.nk)
    assert np.max(np.abs(S)) > 0.1 and res < 1e-11, res
    if pad:  # the region is the bar of the sample and its contacts
        assert (q.j0, q.nj) == (gr.y1 - 1, gr.y2 - gr.y1 + 2)


def test_conversion_closed_form():
    """C_µ on a manufactured state, Ψ_µ = r_µ e^{iφ_µ}, unit links: the τ-weighted differences of
    sin(φ[+e] − φ) along each axis plus Σ_ν γ_µν r_µ r_ν sin(φ_ν − φ_µ) (exchange.jl:316-337)."""
    from oracle.oracle import Oracle
    md = _model(gamma12=-0.07)
    gr = Grid.make(10, 8, 6, dx=0.5)
    mk = box_masks(gr)
    st = make_stepper(md, gr, mk, qp_exchange=2.0)
    s = initial_state(md, gr, mk, 0.0)
    I, J, K = np.meshgrid(np.arange(1, gr.Nx + 2), np.arange(1, gr.Ny + 2), np.arange(1, gr.Nz + 2), indexing="ij")
    r = (0.9, 0.6)
    phi = (0.013 * I ** 2 + 0.02 * J ** 2 - 0.05 * K, -0.01 * I ** 2 + 0.03 * J + 0.04 * K ** 2)
    for mu in range(2):
        s.G[mu] = np.asfortranarray(r[mu] * np.exp(1j * phi[mu]))
        s.F[mu] = s.G[mu].copy(order="F")
    o = Oracle(st)
    o.set_state(s)
    _, (S,) = o.qp_potential()
    q = st.qp
    rx2 = 1 / gr.dx ** 2
    C = []
    for mu in range(2):
        p = phi[mu]
        g = [md.bands[mu].stiffness(a) for a in range(3)]
        c = np.zeros_like(p)
        i, j, k = np.ogrid[gr.x1 - 1:gr.x2 + 1, gr.y1 - 1:gr.y2 + 1, gr.z1 - 1:gr.z2 + 1]
        tx, ty, tz = mk.tx, mk.ty, mk.tz
        cx = tx[i, j, k] * np.sin(p[i + 1, j, k] - p[i, j, k]) + tx[i - 1, j, k] * np.sin(p[i - 1, j, k] - p[i, j, k])
        cy = ty[i, j, k] * np.sin(p[i, j + 1, k] - p[i, j, k]) + ty[i, j - 1, k] * np.sin(p[i, j - 1, k] - p[i, j, k])
        cz = tz[i, j, k] * np.sin(p[i, j, k + 1] - p[i, j, k]) + tz[i, j, k - 1] * np.sin(p[i, j, k - 1] - p[i, j, k])
        nu = 1 - mu
        c[i, j, k] = r[mu] ** 2 * (g[0] * cx + g[1] * cy + g[2] * cz) * rx2 + \
            md.gamma[mu][nu] * r[mu] * r[nu] * np.sin(phi[nu][i, j, k] - p[i, j, k])
        C.append(c)
    want = np.zeros_like(C[0])
    want[gr.x1 - 1:gr.x2 + 1, gr.y1 - 1:gr.y2 + 1, gr.z1 - 1:gr.z2 + 1] = \
        (C[1] / q.s[1] - C[0] / q.s[0])[gr.x1 - 1:gr.x2 + 1, gr.y1 - 1:gr.y2 + 1, gr.z1 - 1:gr.z2 + 1]
    got = _node(S, gr, q.j0, q.k0)
    assert np.max(np.abs(want)) > 1.0
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))


def test_fast_exchange_is_one_fluid():
    """ν = 1e6: the step of today to O(1/ν) after 300 steps (test_physics.jl:363-366)."""
    from oracle.oracle import Oracle
    md = _model()
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    H = 0.03
    st0 = make_stepper(md, gr, mk)
    stb = make_stepper(md, gr, mk, qp_exchange=1e6)
    HJ = st0.selffield(0.5)
    a, b = Oracle(st0), Oracle(stb)
    s = initial_state(md, gr, mk, H)
    a.set_state(s)
    b.set_state(s)
    a.step(300, H, HJ)
    b.step(300, H, HJ)
    sa, sb = a.get_state(), b.get_state()
    for mu in range(2):
        assert np.max(np.abs(sa.G[mu] - sb.G[mu])) < 1e-4
    # and a slow exchange is not (the potential does act)
    c = Oracle(make_stepper(md, gr, mk, qp_exchange=0.1))
    c.set_state(s)
    c.step(300, H, HJ)
    assert np.max(np.abs(c.get_state().G[0] - sa.G[0])) > 1e-6


def _chi(md, s):
    return [s[mu] * md.sigma_n * md.bands[mu].Gamma / md.bands[mu].stiffness() for mu in range(md.N)]


def test_two_bands_as_a_matrix_are_the_two_band_model():
    """QPExchangeN with two bands gives δ = μ_π − μ_σ and the same steps (test_physics.jl:380-396)."""
    from oracle.oracle import Oracle
    md = _model()
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    s0, st0, H, HJ = _warm(md, gr, mk, 200)
    sa = make_stepper(md, gr, mk, qp_exchange=1.0)
    chi = _chi(md, sa.qp.s)
    sb = make_stepper(md, gr, mk, qp_exchange=[[0.0, 1.0], [chi[0] / chi[1], 0.0]])
    assert sb.qp.mode == 2 and np.allclose(sb.qp.s, sa.qp.s)
    oa, ob = Oracle(sa), Oracle(sb)
    oa.set_state(s0)
    ob.set_state(s0)
    (d,), _ = oa.qp_potential()
    M, _ = ob.qp_potential()
    P = sb.qp.P
    mu_ = [sum(P[m, k] * M[k] for k in range(2)) for m in range(2)]
    assert np.max(np.abs(d)) > 1e-3
    assert np.max(np.abs((mu_[1] - mu_[0]) - d)) < 1e-12 * max(1, np.max(np.abs(d)))
    oa.step(200, H, HJ)
    ob.step(200, H, HJ)
    a, b = oa.get_state(), ob.get_state()
    for m in range(2):
        assert np.max(np.abs(a.G[m] - b.G[m])) < 1e-10


def test_three_fluids_gauge_and_equations():
    """N = 3: Σ s_µ μ_µ = 0, and each μ_µ solves σn∇²μ − S⁻¹Lμ = S⁻¹C − ΣC with a different
    conductance per pair (test_physics.jl:397-426)."""
    from oracle.oracle import Oracle
    md = _three()
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    s3, st3, H, HJ = _warm(md, gr, mk, 200, J=0.3)
    w = [md.bands[m].stiffness() * md.bands[m].Gamma / md.bands[m].beta for m in range(3)]
    sh = [x / sum(w) for x in w]
    chi = _chi(md, sh)
    rates = [[0.0 if m == n else 0.5 / chi[m] for n in range(3)] for m in range(3)]
    stn = make_stepper(md, gr, mk, qp_exchange=rates)
    o = Oracle(stn)
    o.set_state(s3)
    M, _ = o.qp_potential()
    P = stn.qp.P
    pot = [sum(P[m, k] * M[k] for k in range(3)) for m in range(3)]
    assert np.max(np.abs(pot[2])) > 1e-3
    assert np.max(np.abs(sum(stn.qp.s[m] * pot[m] for m in range(3)))) < 1e-12
    Gp = np.array([[0.0, 0.5, 0.1], [0.5, 0.0, 2.0], [0.1, 2.0, 0.0]])
    stg = make_stepper(md, gr, mk, qp_exchange=[[0.0 if m == n else Gp[m, n] / chi[m] for n in range(3)]
                                                 for m in range(3)])
    og = Oracle(stg)
    og.set_state(s3)
    Mg, _ = og.qp_potential()
    q = stg.qp
    Pg = q.P
    pot = [_node(sum(Pg[m, k] * Mg[k] for k in range(3)), gr, q.j0, q.k0) for m in range(3)]
    # the right sides R_µ = C_µ/s_µ − ΣC back from the sources of the modes, m_k = Σ_µ Pinv[k][µ] R_µ
    # (Pinv = Qᵀ S^{1/2} is invertible)
    Src = og.qp_sources()
    Rm = np.linalg.inv(q.Pinv)
    R = [_node(sum(Rm[m, k] * Src[k] for k in range(3)), gr, q.j0, q.k0) for m in range(3)]
    res = 0.0
    for m in range(3):
        def lam(i, j, k, m=m):
            return sum(Gp[m, n] * (pot[m][i, j, k] - pot[n][i, j, k]) for n in range(3)) / q.s[m]
        res = max(res, _residual(pot[m], R[m], gr, md.sigma_n, lam, q.j0, q.nj, q.k0, q.nk))
    assert max(np.max(np.abs(r)) for r in R) > 0.1 and res < 1e-10, res


def test_rates_without_detailed_balance_are_refused():
    md = _three()
    gr = Grid.make(12, 8, 4, dx=0.5)
    mk = wedge_masks(gr)
    with pytest.raises(ValueError):
        make_stepper(md, gr, mk, qp_exchange=[[0, 1, 2], [2, 0, 3], [3, 6, 0]])
    with pytest.raises(ValueError):
        make_stepper(md, gr, mk, qp_exchange=0.5)   # a number is two bands only
