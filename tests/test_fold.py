"""The algebra of the parity-folded transforms that the Float32 preconditioner runs on the device
(k_fold, the quarter x-y transforms with half-size bases, k_zmodes_q, k_unfold), restated in numpy
and checked against the oracle's dense separable solve (gtdgl/src/implicit.jl:62-83): the mirror
symmetry of the DCT-II and DST-I columns, the folds and unfolds of even and odd lengths, the
placement of X, Y and Z of one mode in the quarters, and the whole solve to rounding."""
import numpy as np
import pytest

from juliatdgl_amd.scheme import Grid, box_masks, make_stepper, two_band
from oracle.oracle import Oracle, basis


def half(M, p):
    """Columns 2cc + p of the basis on its first ceil(L/2) rows, zero-padded to a square."""
    L = M.shape[0]
    Lh = (L + 1) // 2
    H = np.zeros((Lh, Lh))
    for cc in range(Lh):
        if 2 * cc + p < L:
            H[:, cc] = M[:Lh, 2 * cc + p]
    return H


def fold1(a, axis):
    """(even, odd) folds along `axis`: a[i] ± a[L-1-i]; an odd length's centre goes to the even one alone."""
    L = a.shape[axis]
    Lh, h = (L + 1) // 2, L // 2
    a = np.moveaxis(a, axis, 0)
    F0 = np.zeros((Lh,) + a.shape[1:])
    F1 = np.zeros_like(F0)
    F0[:h] = a[:h] + a[::-1][:h]
    F1[:h] = a[:h] - a[::-1][:h]
    if L % 2:
        F0[h] = a[h]
    return np.moveaxis(F0, 0, axis), np.moveaxis(F1, 0, axis)


def unfold1(R0, R1, L, axis):
    h = L // 2
    R0, R1 = np.moveaxis(R0, axis, 0), np.moveaxis(R1, axis, 0)
    out = np.zeros((L,) + R0.shape[1:])
    out[:h] = R0[:h] + R1[:h]
    out[L - 1 - np.arange(h)] = R0[:h] - R1[:h]
    if L % 2:
        out[h] = R0[h] + R1[h]
    return np.moveaxis(out, 0, axis)


@pytest.mark.parametrize("nn", [5, 6, 19, 20, 47])
def test_basis_columns_are_even_or_odd(nn):
    V, U, _ = basis(nn)
    for M in (V, U):
        L = M.shape[0]
        for p in range(L):
            assert np.abs(M[::-1, p] - (-1) ** p * M[:, p]).max() <= 1e-14


@pytest.mark.parametrize("nn", [6, 7, 19])
def test_folded_products_are_the_full_ones(nn):
    """Mᵀa by the folded halves (the forward transform) and M t by the unfolded ones (the inverse)."""
    rng = np.random.default_rng(nn)
    for M in basis(nn)[:2]:
        L = M.shape[0]
        a = rng.standard_normal(L)
        F0, F1 = fold1(a, 0)
        T = np.zeros(L)
        for p, F in ((0, F0), (1, F1)):
            T[p::2] = (half(M, p).T @ F)[: len(T[p::2])]
        assert np.abs(T - M.T @ a).max() <= 1e-13
        t = rng.standard_normal(L)
        R = [half(M, p) @ np.pad(t[p::2], (0, (L + 1) // 2 - len(t[p::2]))) for p in (0, 1)]
        assert np.abs(unfold1(R[0], R[1], L, 0) - M @ t).max() <= 1e-13


def folded_sep(B, gr, a):
    """The device's folded preconditioner in numpy: fold x and y, quarter transforms, the z pass with
    the projector reading X, Y, Z of each mode from their quarters, inverse quarter transforms, unfold."""
    nx, ny, nz = gr.Nx - 1, gr.Ny - 1, gr.Nz - 1
    Vx, Ux, sx = basis(nx)
    Vy, Uy, sy = basis(ny)
    Vz, Uz, sz = basis(nz)
    Mx, My, Mz = [Ux, Vx, Vx], [Vy, Uy, Vy], [Vz, Vz, Uz]
    Q = []
    for c in range(3):
        Fx = fold1(B[c], 0)
        q = {}
        for px in (0, 1):
            Fy = fold1(Fx[px], 1)
            for py in (0, 1):
                t = np.einsum("im,ijz->mjz", half(Mx[c], px), Fy[py])
                q[px, py] = np.einsum("jn,mjz->mnz", half(My[c], py), t)
        Q.append(q)
    for c in range(3):
        for k in Q[c]:
            Q[c][k] = np.einsum("kz,ijk->ijz", Mz[c], Q[c][k])
    for mz in range(nz):
        for my in range(ny):
            for mx in range(nx):
                k1, k2, k3 = sx[mx], sy[my], sz[mz]
                k2s = k1 * k1 + k2 * k2 + k3 * k3
                if k2s == 0:
                    continue
                px, py = mx & 1, my & 1
                # Z at (mm, qq) of quarter (px, py); Y at (mm, qq - 1 + py) of (px, 1 - py); X at (mm - 1 + px, qq) of (1 - px, py)
                mm, qq = mx >> 1, my >> 1
                zq, yq, xq = Q[2][px, py], Q[1][px, 1 - py], Q[0][1 - px, py]
                x1 = xq[mm - 1 + px, qq, mz] if mx > 0 else 0.0
                y1 = yq[mm, qq - 1 + py, mz] if my > 0 else 0.0
                z1 = zq[mm, qq, mz - 1] if mz > 0 else 0.0
                d = (k1 * x1 + k2 * y1 + k3 * z1) / k2s
                f = 1 / (1 + a * k2s)
                if mx > 0:
                    xq[mm - 1 + px, qq, mz] = k1 * d + f * (x1 - k1 * d)
                if my > 0:
                    yq[mm, qq - 1 + py, mz] = k2 * d + f * (y1 - k2 * d)
                if mz > 0:
                    zq[mm, qq, mz - 1] = k3 * d + f * (z1 - k3 * d)
    out = []
    for c in range(3):
        R = {}
        for (px, py), t in Q[c].items():
            t = np.einsum("zk,ijk->ijz", Mz[c], t)
            t = np.einsum("im,mjz->ijz", half(Mx[c], px), t)
            R[px, py] = np.einsum("jn,inz->ijz", half(My[c], py), t)
        Lx, Ly = B[c].shape[0], B[c].shape[1]
        rows = [unfold1(R[px, 0], R[px, 1], Ly, 1) for px in (0, 1)]
        out.append(unfold1(rows[0], rows[1], Lx, 0))
    return out


@pytest.mark.parametrize("dims", [(6, 4, 3), (7, 5, 4), (8, 6, 2)])
def test_folded_solve_is_the_dense_solve(dims):
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(*dims, dx=0.5, buf=1)
    st = make_stepper(md, gr, box_masks(gr))
    rng = np.random.default_rng(1)
    sh = [(gr.Nx - 2, gr.Ny - 1, gr.Nz - 1), (gr.Nx - 1, gr.Ny - 2, gr.Nz - 1), (gr.Nx - 1, gr.Ny - 1, gr.Nz - 2)]
    B = [rng.standard_normal(s) for s in sh]
    ref = Oracle(st).sep_solve(*[b.copy(order="F") for b in B])
    got = folded_sep(B, gr, st.fc.a)
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() <= 1e-13 * max(1.0, np.abs(r).max())


ZQ_SLACK = 128


def zq_bases(nz):
    """The quarter z pass's padded bases, as the library builds them: V(n, mz) and U(n, mz - 1) (zero
    for mz = 0 or n = nz - 1) by rows, then by columns, each nz x nz followed by ZQ_SLACK zeros."""
    V, U, _ = basis(nz)
    zb = [np.zeros(nz * nz + ZQ_SLACK) for _ in range(4)]
    for n in range(nz):
        for mz in range(nz):
            u = U[n, mz - 1] if (mz >= 1 and n < nz - 1) else 0.0
            zb[0][n * nz + mz] = V[n, mz]
            zb[1][n * nz + mz] = u
            zb[2][mz * nz + n] = V[n, mz]
            zb[3][mz * nz + n] = u
    return zb


def zpass_blocks(buf, h, off, nx, ny, nz, s, a, G):
    """k_zmodes_q block by block: lanes over Z's quarter flattened along mm, G waves a block, wave g
    the modes (positions) g mzc .. g mzc + MZ - 1 read without guards from the padded bases; every
    index checked against its buffer.  `buf` holds the quarters [z][j][i] at off[c][2 px + py]."""
    zb = zq_bases(nz)
    mzc = (nz + G - 1) // G
    MZ = next(k for k in (2, 4, 6, 8, 12, 16) if k >= mzc)
    hzx, hzy = h[2]
    nblk = -(-hzx * hzy // 64)
    out = buf.copy()
    lane = np.arange(64)

    def at(c, quarter, i, j, z):
        idx = off[c][quarter] + i + h[c][0] * j + h[c][0] * h[c][1] * z
        assert (idx >= 0).all() and (idx < buf.size).all()
        return idx

    for pq in range(4):
        px, py = pq >> 1, pq & 1
        for b in range(nblk):
            f = b * 64 + lane
            mm, qq = f % hzx, f // hzx
            mx, my = 2 * mm + px, 2 * qq + py
            live = (qq < hzy) & (mx < nx) & (my < ny)
            hx, hy = live & (mx > 0), live & (my > 0)
            xi = np.minimum(np.maximum(mm - 1 + px, 0), h[0][0] - 1); xj = np.minimum(qq, h[0][1] - 1)
            yi = np.minimum(mm, h[1][0] - 1); yj = np.minimum(np.maximum(qq - 1 + py, 0), h[1][1] - 1)
            zi = np.minimum(mm, h[2][0] - 1); zj = np.minimum(qq, h[2][1] - 1)
            C = np.zeros((3, nz, 64))
            for z in range(nz):
                C[0, z] = np.where(hx, buf[at(0, 2 * (1 - px) + py, xi, xj, z)], 0)
                C[1, z] = np.where(hy, buf[at(1, 2 * px + (1 - py), yi, yj, z)], 0)
                C[2, z] = np.where(live & (z < nz - 1), buf[at(2, pq, zi, zj, min(z, nz - 2))], 0)
            acc = np.zeros((G, 3, MZ, 64))
            for g in range(G):
                m0 = g * mzc
                for n in range(nz):
                    for r in range(MZ):
                        k = n * nz + m0 + r
                        assert k < zb[0].size
                        acc[g, 0, r] += C[0, n] * zb[0][k]
                        acc[g, 1, r] += C[1, n] * zb[0][k]
                        acc[g, 2, r] += C[2, n] * zb[1][k]
            P = np.zeros((3, nz, 64))
            for g in range(G):
                for r in range(mzc):
                    mz = g * mzc + r
                    if mz >= nz:
                        continue
                    k1 = np.where(live, s[0][np.minimum(mx, nx - 1)], 0)
                    k2 = np.where(live, s[1][np.minimum(my, ny - 1)], 0)
                    k3 = s[2][mz]
                    x1 = np.where(hx, acc[g, 0, r], 0)
                    y1 = np.where(hy, acc[g, 1, r], 0)
                    z1 = np.where(live & (mz > 0), acc[g, 2, r], 0)
                    k2s = k1 * k1 + k2 * k2 + k3 * k3
                    ok = live & (k2s != 0)
                    d = np.where(ok, (k1 * x1 + k2 * y1 + k3 * z1) / np.where(ok, k2s, 1), 0)
                    fa = 1 / (1 + a * k2s)
                    P[0, mz] = np.where(ok & hx, k1 * d + fa * (x1 - k1 * d), x1)
                    P[1, mz] = np.where(ok & hy, k2 * d + fa * (y1 - k2 * d), y1)
                    P[2, mz] = np.where(ok & (mz > 0), k3 * d + fa * (z1 - k3 * d), z1)
            for g in range(G):
                m0 = g * mzc
                o = np.zeros((3, MZ, 64))
                for qm in range(nz):
                    for r in range(MZ):
                        k = qm * nz + m0 + r
                        assert k < zb[2].size
                        o[0, r] += P[0, qm] * zb[2][k]
                        o[1, r] += P[1, qm] * zb[2][k]
                        o[2, r] += P[2, qm] * zb[3][k]
                for r in range(mzc):
                    n = m0 + r
                    if n >= nz:
                        continue
                    out[at(0, 2 * (1 - px) + py, xi, xj, n)[hx]] = o[0, r][hx]
                    out[at(1, 2 * px + (1 - py), yi, yj, n)[hy]] = o[1, r][hy]
                    if n < nz - 1:
                        out[at(2, pq, zi, zj, n)[live]] = o[2, r][live]
    return out


@pytest.mark.parametrize("dims,G", [((6, 4, 3), 4), ((7, 5, 4), 8), ((9, 5, 10), 4), ((9, 5, 10), 8)])
def test_grouped_z_pass_is_the_dense_one(dims, G):
    """The device's grouped z pass (zpass_blocks) against the dense z pass and projector of
    folded_sep, on random quarters: the padded bases and the unguarded group reads add nothing."""
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(*dims, dx=0.5, buf=1)
    st = make_stepper(md, gr, box_masks(gr))
    nx, ny, nz = gr.Nx - 1, gr.Ny - 1, gr.Nz - 1
    ext = [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1)]
    h = [((e[0] + 1) // 2, (e[1] + 1) // 2) for e in ext]
    off, o = [], 0
    for c in range(3):
        off.append([o + k * h[c][0] * h[c][1] * ext[c][2] for k in range(4)])
        o += 4 * h[c][0] * h[c][1] * ext[c][2]
    rng = np.random.default_rng(dims[2] + G)
    buf = rng.standard_normal(o)
    s = [basis(k)[2] for k in (nx, ny, nz)]
    a = st.fc.a
    got = zpass_blocks(buf, h, off, nx, ny, nz, s, a, G)
    # the dense pass on the same quarters
    Vz, Uz, _ = basis(nz)
    Mz = [Vz, Vz, Uz]
    Q = [{(k >> 1, k & 1): buf[off[c][k]: off[c][k] + h[c][0] * h[c][1] * ext[c][2]].reshape(
        (ext[c][2], h[c][1], h[c][0])).transpose(2, 1, 0).copy() for k in range(4)} for c in range(3)]
    for c in range(3):
        for k in Q[c]:
            Q[c][k] = np.einsum("kz,ijk->ijz", Mz[c], Q[c][k])
    sx, sy, sz = s
    for mz in range(nz):
        for my in range(ny):
            for mx in range(nx):
                k1, k2, k3 = sx[mx], sy[my], sz[mz]
                k2s = k1 * k1 + k2 * k2 + k3 * k3
                if k2s == 0:
                    continue
                px, py, mm, qq = mx & 1, my & 1, mx >> 1, my >> 1
                zq, yq, xq = Q[2][px, py], Q[1][px, 1 - py], Q[0][1 - px, py]
                x1 = xq[mm - 1 + px, qq, mz] if mx > 0 else 0.0
                y1 = yq[mm, qq - 1 + py, mz] if my > 0 else 0.0
                z1 = zq[mm, qq, mz - 1] if mz > 0 else 0.0
                d = (k1 * x1 + k2 * y1 + k3 * z1) / k2s
                f = 1 / (1 + a * k2s)
                if mx > 0:
                    xq[mm - 1 + px, qq, mz] = k1 * d + f * (x1 - k1 * d)
                if my > 0:
                    yq[mm, qq - 1 + py, mz] = k2 * d + f * (y1 - k2 * d)
                if mz > 0:
                    zq[mm, qq, mz - 1] = k3 * d + f * (z1 - k3 * d)
    want = buf.copy()
    for c in range(3):
        for k in range(4):
            t = np.einsum("zk,ijk->ijz", Mz[c], Q[c][k >> 1, k & 1])
            want[off[c][k]: off[c][k] + t.size] = t.transpose(2, 1, 0).ravel()
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
