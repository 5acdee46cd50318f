"""CPU oracle of the warm-start CNN (test infrastructure only).

float64 restatement of the network of nemo3d_amd/csrc/nh_cnn.hip: conv
3x3x3 with zero padding + ReLU layers, 1x1x1 head, weights rounded to
BF16 (round to nearest even) as the device uses them.  The backward pass
is written out by hand and pinned by central finite differences of the
forward pass (tests/test_learned.py).
"""
import numpy as np


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(c_in, width, layers, c_out):
    offs = []
    off = 0
    ci = c_in
    for _ in range(layers):
        w = (off, off + width * ci * 27)
        off = w[1]
        b = (off, off + width)
        off = b[1]
        offs.append((w, b, width, ci))
        ci = width
    w = (off, off + c_out * ci)
    off = w[1]
    b = (off, off + c_out)
    off = b[1]
    offs.append((w, b, c_out, ci))
    return offs, off


def _shift(x, dx, dy, dz):
    """y[p] = x[p + d] with zero padding; x: (gx, gy, gz, C)."""
    y = np.zeros_like(x)
    gx, gy, gz = x.shape[:3]
    sx = slice(max(0, -dx), gx - max(0, dx)), slice(max(0, dx), gx - max(0, -dx))
    sy = slice(max(0, -dy), gy - max(0, dy)), slice(max(0, dy), gy - max(0, -dy))
    sz = slice(max(0, -dz), gz - max(0, dz)), slice(max(0, dz), gz - max(0, -dz))
    y[sx[0], sy[0], sz[0]] = x[sx[1], sy[1], sz[1]]
    return y


OFFS = [(k // 9 - 1, (k // 3) % 3 - 1, k % 3 - 1) for k in range(27)]


def forward(params, feats, dims, c_in, width, layers, c_out, keep=False,
            round_bf16=True):
    offs, n = layout(c_in, width, layers, c_out)
    p = bf16(params) if round_bf16 else np.asarray(params, dtype=np.float64)
    x = np.asarray(feats, dtype=np.float64).reshape(*dims, c_in)
    cache = []
    for (w, b, co, ci) in offs[:-1]:
        W = p[w[0]:w[1]].reshape(co, ci, 27)
        pre = np.broadcast_to(p[b[0]:b[1]], x.shape[:3] + (co,)).copy()
        for k, d in enumerate(OFFS):
            pre += _shift(x, *d) @ W[:, :, k].T
        cache.append((x, pre))
        x = np.maximum(pre, 0.0)
    (w, b, co, ci) = offs[-1]
    W = p[w[0]:w[1]].reshape(co, ci)
    out = x @ W.T + p[b[0]:b[1]]
    if keep:
        return out.reshape(-1, c_out), cache, x
    return out.reshape(-1, c_out)


def backward(params, feats, dims, c_in, width, layers, c_out, dout,
             round_bf16=True):
    """Gradient of sum(out * dout) with respect to params (float64)."""
    offs, n = layout(c_in, width, layers, c_out)
    p = bf16(params) if round_bf16 else np.asarray(params, dtype=np.float64)
    out, cache, a = forward(params, feats, dims, c_in, width, layers, c_out,
                            keep=True, round_bf16=round_bf16)
    g = np.zeros(n)
    d = np.asarray(dout, dtype=np.float64).reshape(*dims, c_out)
    (w, b, co, ci) = offs[-1]
    W = p[w[0]:w[1]].reshape(co, ci)
    g[w[0]:w[1]] = np.einsum("xyzo,xyzi->oi", d, a).ravel()
    g[b[0]:b[1]] = d.sum(axis=(0, 1, 2))
    da = d @ W
    for l in range(layers - 1, -1, -1):
        (w, b, co, ci) = offs[l]
        x, pre = cache[l]
        dpre = np.where(pre > 0, da, 0.0)
        W = p[w[0]:w[1]].reshape(co, ci, 27)
        gW = np.zeros((co, ci, 27))
        dx = np.zeros_like(x)
        for k, dd in enumerate(OFFS):
            xs = _shift(x, *dd)
            gW[:, :, k] = np.einsum("xyzo,xyzi->oi", dpre, xs)
            dx += _shift(dpre @ W[:, :, k], -dd[0], -dd[1], -dd[2])
        g[w[0]:w[1]] = gW.ravel()
        g[b[0]:b[1]] = dpre.sum(axis=(0, 1, 2))
        da = dx
    return g


def map_out(out, point, slot, nb, ne):
    """Network output (P, c_out) -> start block (ne, A*nb) complex: atom a,
    orbital o of state e reads channels 2*((e*2 + slot[a])*nb + o) + {0, 1}
    of grid point point[a]."""
    out = np.asarray(out, dtype=np.float64)
    A = point.size
    Y = np.zeros((ne, A * nb), dtype=np.complex128)
    o = np.arange(nb)
    for e in range(ne):
        ch = 2 * ((e * 2 + slot[:, None]) * nb + o[None, :])
        Y[e] = (out[point[:, None], ch] + 1j * out[point[:, None], ch + 1]).reshape(-1)
    return Y


def map_grad(G, point, slot, nb, ne, n_points, c_out):
    """Adjoint of map_out for real/imaginary channel gradients."""
    dout = np.zeros((n_points, c_out))
    o = np.arange(nb)
    for e in range(ne):
        g = G[e].reshape(-1, nb)
        ch = 2 * ((e * 2 + slot[:, None]) * nb + o[None, :])
        dout[point[:, None], ch] = g.real
        dout[point[:, None], ch + 1] = g.imag
    return dout


def subspace_loss(Y, V, cluster, out_cluster):
    """L = mean_e (1 - sum_{f in cluster of e} |<v_f, y_e>|^2 / |y_e|^2) and
    G_e = dL/dRe(y_e) + i dL/dIm(y_e)."""
    ne = Y.shape[0]
    C = np.conj(V) @ Y.T  # C[f, e] = <v_f, y_e>
    L = 0.0
    G = np.zeros_like(Y)
    for e in range(ne):
        sel = np.asarray(cluster) == out_cluster[e]
        b = np.vdot(Y[e], Y[e]).real
        a = np.sum(np.abs(C[sel, e]) ** 2)
        L += (1 - a / b) / ne
        Py = V[sel].T @ C[sel, e]
        G[e] = 2 * (-Py / b + a * Y[e] / b ** 2) / ne
    return L, G
