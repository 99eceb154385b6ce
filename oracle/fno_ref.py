"""CPU reference of the Fourier neural operator (felis_amd/csrc/felis_fno.hip).

TEST INFRASTRUCTURE ONLY: tests compare the GPU network's outputs and
gradients with this numpy restatement. It forms the same products as the
kernels (transform matrices, channel maps, per-mode mixing), rounding both
operands of every product to BF16 as the kernels do (rounding=True), and
accumulates in FP64; with rounding=False it is the exact function, whose
derivative the gradients restate and whose spectral part equals the
truncated two-dimensional FFT (tests/test_fno_ref.py).

Parameters are one flat vector in the kernels' order (layout()); maps are
channel-last [map][y][x][channel] on a 64 x 64 grid.
"""
from __future__ import annotations

import numpy as np

from oracle.mlp_ref import bf16, gelu, gelu_d

N = 64
P = N * N


def layout(cin, cout, w=32, m=8, L=4, ph=64):
    """Offsets and shapes of the parameter vector: name -> (offset, shape)."""
    out, o = {}, 0

    def put(name, shape):
        nonlocal o
        out[name] = (o, shape)
        o += int(np.prod(shape))

    put("lw", (cin, w))
    put("lb", (w,))
    for l in range(L):
        put(f"W{l}", (w, w))
        put(f"b{l}", (w,))
        put(f"Wm{l}", (2 * m, m, 2 * w, 2 * w))
    put("q1", (w, ph))
    put("c1", (ph,))
    put("q2", (ph, cout))
    put("c2", (cout,))
    out["_size"] = o
    return out


def entries(lay):
    """(name, offset, shape) of every parameter block."""
    return [(k, v[0], v[1]) for k, v in lay.items() if k != "_size"]


def unpack(Pv, lay):
    Pv = np.asarray(Pv, np.float64)
    return {k: Pv[o:o + int(np.prod(s))].reshape(s) for k, o, s in entries(lay)}


def tables(m):
    """Fx [2m][N], Fxi [N][2m], Gy [4m][2N], Giy [2N][4m] as the kernels build them."""
    x = np.arange(N)
    kx = np.arange(m)
    ph = 2 * np.pi * np.outer(kx, x) / N
    Fx = np.concatenate([np.cos(ph), -np.sin(ph)])
    al = np.where(kx == 0, 1.0, 2.0) / (N * N)
    Fxi = np.concatenate([al[:, None] * np.cos(ph), -al[:, None] * np.sin(ph)]).T
    ky = np.concatenate([np.arange(m), np.arange(-m, 0)])  # 0..m-1, then -m..-1
    y = np.arange(N)
    phy = 2 * np.pi * np.outer(ky, y) / N
    c, s = np.cos(phy), np.sin(phy)
    Gy = np.zeros((4 * m, 2 * N))
    Gy[0::2, 0::2] = c
    Gy[0::2, 1::2] = s
    Gy[1::2, 0::2] = -s
    Gy[1::2, 1::2] = c
    Giy = np.zeros((2 * N, 4 * m))
    Giy[0::2, 0::2] = c.T
    Giy[0::2, 1::2] = -s.T
    Giy[1::2, 0::2] = s.T
    Giy[1::2, 1::2] = c.T
    return Fx.astype(np.float32), Fxi.astype(np.float32), Gy.astype(np.float32), Giy.astype(np.float32)


def _r(rounding):
    return (lambda a: bf16(a).astype(np.float64)) if rounding else (lambda a: np.asarray(a, np.float64))


def spectral(hv, Wm, m, rounding=True):
    """The spectral part of a Fourier layer on hv [B][N][N][w] with mode
    weights Wm [2m][m][2w][2w], by the kernels' products: [B][N][N][w], and
    the kept modes Y [B][2m][m][2w] (the mixing's input)."""
    r = _r(rounding)
    Fx, Fxi, Gy, Giy = tables(m)
    B, w = hv.shape[0], hv.shape[-1]
    Xf = np.einsum("kx,byxc->bykc", r(Fx), r(hv))            # [B][y][(r,kx)][c]
    Xf2 = Xf.reshape(B, 2 * N, m * w)                          # rows (y, r)
    Yv = np.einsum("ij,bjn->bin", r(Gy), r(Xf2))               # [B][(ky,r')][(kx,c)]
    Y = Yv.reshape(B, 2 * m, 2, m, w).transpose(0, 1, 3, 2, 4).reshape(B, 2 * m, m, 2 * w)
    Z = np.einsum("bkjr,kjrs->bkjs", r(Y), r(Wm))
    Zv = Z.reshape(B, 2 * m, m, 2, w).transpose(0, 1, 3, 2, 4).reshape(B, 4 * m, m * w)
    Xb = np.einsum("ij,bjn->bin", r(Giy), r(Zv)).reshape(B, N, 2 * m, w)
    return np.einsum("xk,bykc->byxc", r(Fxi), r(Xb)), Y


def forward(Pv, X, cin, cout, w=32, m=8, L=4, ph=64, mu=None, isd=None, rounding=True):
    """Output [B][P][cout] and the cache for backward. X [B][P][cin]."""
    lay = layout(cin, cout, w, m, L, ph)
    p = unpack(Pv, lay)
    r = _r(rounding)
    B = X.shape[0]
    x = np.asarray(X, np.float32)
    if mu is not None:
        x = ((x - mu) * isd).astype(np.float32)
    x = x.astype(np.float64)
    h = r(x) @ r(p["lw"]) + p["lb"]
    cache = dict(x=x, h=[h], a=[], Y=[])
    for l in range(L):
        S, Y = spectral(h.reshape(B, N, N, w), p[f"Wm{l}"], m, rounding)
        a = (r(h) @ r(p[f"W{l}"]) + p[f"b{l}"]).reshape(B, N, N, w) + S
        a = a.reshape(B, P, w)
        cache["a"].append(a)
        cache["Y"].append(Y)
        h = gelu(a) if l < L - 1 else a
        if l < L - 1:
            cache["h"].append(h)
    zq = r(h) @ r(p["q1"]) + p["c1"]
    q = gelu(zq)
    out = r(q) @ r(p["q2"]) + p["c2"]
    cache.update(zq=zq, q=q, hl=h)
    return out, cache


def loss_and_dout(out, E, R, wb):
    """sum_b wb sum_{p,h} |o E - R|^2 / (B P) and its derivative by out (the
    (re, im) pairs of the cout channels are complex)."""
    B = out.shape[0]
    o = out[..., 0::2] + 1j * out[..., 1::2]
    e = E[..., 0::2] + 1j * E[..., 1::2]
    t = R[..., 0::2] + 1j * R[..., 1::2]
    d = o * e - t
    wv = np.asarray(wb, np.float64)[:, None, None]
    loss = float(np.sum(wv * np.abs(d) ** 2) / (B * P))
    g = 2.0 * wv * d * np.conj(e) / (B * P)
    dout = np.empty_like(out, dtype=np.float64)
    dout[..., 0::2] = g.real
    dout[..., 1::2] = g.imag
    return loss, dout


def gradients(Pv, X, E, R, wb, cin, cout, w=32, m=8, L=4, ph=64, mu=None, isd=None, rounding=True):
    """Loss and the gradient of every parameter, as the kernels form them."""
    lay = layout(cin, cout, w, m, L, ph)
    p = unpack(Pv, lay)
    r = _r(rounding)
    Fx, Fxi, Gy, Giy = tables(m)
    out, c = forward(Pv, X, cin, cout, w, m, L, ph, mu, isd, rounding)
    B = out.shape[0]
    loss, dout = loss_and_dout(out, E, R, wb)
    G = {k: np.zeros(s) for k, o, s in entries(lay)}
    flat = lambda a: a.reshape(-1, a.shape[-1])  # noqa: E731
    G["q2"] = flat(r(c["q"])).T @ flat(r(dout))
    G["c2"] = flat(dout).sum(0)
    dq = r(dout) @ r(p["q2"]).T
    dz = dq * gelu_d(c["zq"])
    G["q1"] = flat(r(c["hl"])).T @ flat(r(dz))
    G["c1"] = flat(dz).sum(0)
    dh = r(dz) @ r(p["q1"]).T
    for l in range(L - 1, -1, -1):
        h = c["h"][l]
        da = dh * gelu_d(c["a"][l]) if l < L - 1 else dh
        G[f"W{l}"] = flat(r(h)).T @ flat(r(da))
        G[f"b{l}"] = flat(da).sum(0)
        dh = r(da) @ r(p[f"W{l}"]).T
        dav = da.reshape(B, N, N, w)
        dXb = np.einsum("xk,byxc->bykc", r(Fxi), r(dav))              # [B][y][(r,kx)][c]
        dXb2 = dXb.reshape(B, 2 * N, m * w)
        dZv = np.einsum("ij,bin->bjn", r(Giy), r(dXb2))                # [B][(ky,r')][(kx,c)]
        dZ = dZv.reshape(B, 2 * m, 2, m, w).transpose(0, 1, 3, 2, 4).reshape(B, 2 * m, m, 2 * w)
        G[f"Wm{l}"] = np.einsum("bkjr,bkjs->kjrs", r(c["Y"][l]), r(dZ))
        dY = np.einsum("bkjs,kjrs->bkjr", r(dZ), r(p[f"Wm{l}"]))
        dYv = dY.reshape(B, 2 * m, m, 2, w).transpose(0, 1, 3, 2, 4).reshape(B, 4 * m, m * w)
        dXf = np.einsum("ij,bin->bjn", r(Gy), r(dYv)).reshape(B, N, 2 * m, w)
        dh = dh + np.einsum("kx,bykc->byxc", r(Fx), r(dXf)).reshape(B, P, w)
    G["lw"] = flat(r(c["x"])).T @ flat(r(dh))
    G["lb"] = flat(dh).sum(0)
    Gv = np.zeros(lay["_size"])
    for k, o, s in entries(lay):
        Gv[o:o + int(np.prod(s))] = G[k].reshape(-1)
    return loss, Gv


def spectral_fft(h, m):
    """The spectral part of a Fourier layer whose mode mixing is the identity,
    from numpy's FFT: the kept modes (kx 0..m-1, ky 0..m-1 and -m..-1) of the
    2D transform of h [B][N][N][w], weight 2 on kx > 0, real part of the
    inverse transform."""
    H = np.fft.fft2(h, axes=(1, 2))  # [B][ky][kx][c]
    keep = np.zeros((N, N))
    ky = np.r_[0:m, N - m:N]
    keep[np.ix_(ky, np.arange(m))] = 1.0
    keep[:, 1:m] *= 2.0
    return np.real(np.fft.ifft2(H * keep[None, :, :, None], axes=(1, 2)))
