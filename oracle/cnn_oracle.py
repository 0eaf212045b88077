"""cnn_oracle.py -- numpy restatement of the warm-start network (TEST
INFRASTRUCTURE ONLY: the checker of libfpm_amd's fpm_cnn_* kernels, never part
of the product).

The network has no counterpart in the reference.  This file restates, in
float64, exactly what fused-fourier-ptycho_amd/csrc/cnn_api.hip computes:

  encoder level l = 0..L-1:  [avgpool 2x2 if l > 0] -> conv3x3+ReLU -> conv3x3+ReLU  (e_l)
  decoder level l = L-2..0:  upsample 2x (nearest) -> conv3x3+ReLU (d_a);
                             s = d_a + e_l; conv3x3+ReLU (d_b)
  head: conv3x3 (no ReLU) -> depth-to-space by u -> (amplitude, phase) planes
  pupil head (K = pupil_terms > 0): mean of e_{L-1} over its pixels (Cb
      values) -> W1 (64 x Cb) + b1, ReLU -> W2 (K x 64) + b2: K coefficients
  loss: mean squared error over all output values, plus pupil_weight x the
      mean squared error of the K coefficients

Parameters use the library's vector layout: per conv W (Cout x 9 x Cin, tap
= 3 (dy+1) + (dx+1), Cin of the first conv rounded up to 32) then b (Cout);
then, with a pupil head, W1, b1, W2, b2 (FP32 in the library, no BF16).

parallax_shift restates the warm start's pupil measurement (global
Lucas-Kanade shifts of the brightfield images, fpm_parallax_shifts).

With ``bf16=True`` values are rounded to BF16 (round to nearest even) where
the library stores BF16: the input features, every activation, the weights
used in the products, the gradients entering each conv's products.  With
``bf16=False`` everything is float64, which is what the finite-difference
checks of the backward pass use.
"""
from __future__ import annotations

import numpy as np


def round_bf16(x):
    """Round float64/float32 values to the nearest BF16 (ties to even)."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_shapes(levels, base, in_ch, up):
    """[(cin, cout, relu)] of every conv in parameter order."""
    cin_pad = (in_ch + 31) // 32 * 32
    convs = []
    prev = cin_pad
    for l in range(levels):
        c = base << l
        convs += [(prev, c, 1), (c, c, 1)]
        prev = c
    for l in range(levels - 2, -1, -1):
        c = base << l
        convs += [(prev, c, 1), (c, c, 1)]
        prev = c
    convs.append((base, 2 * up * up, 0))
    return convs


def split_params(p, convs):
    out, k = [], 0
    for cin, cout, _ in convs:
        w = p[k:k + cout * 9 * cin].reshape(cout, 9, cin)
        k += cout * 9 * cin
        b = p[k:k + cout]
        k += cout
        out.append((w, b))
    assert k == p.size
    return out


def _shift(x, dy, dx):
    """x(b, y + dy, x + dx, c) with zero padding."""
    B, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv3x3(x, w, b):
    out = np.zeros(x.shape[:3] + (w.shape[0],))
    for tap in range(9):
        out += _shift(x, tap // 3 - 1, tap % 3 - 1) @ w[:, tap, :].T
    return out + (b if b is not None else 0.0)


def conv3x3_input_grad(gz, w):
    """adjoint of conv3x3 in x: gx(p, ci) = sum_tap,co gz(p - off(tap), co) w(co, tap, ci)"""
    gx = np.zeros(gz.shape[:3] + (w.shape[2],))
    for tap in range(9):
        gx += _shift(gz, -(tap // 3 - 1), -(tap % 3 - 1)) @ w[:, tap, :]
    return gx


def conv3x3_weight_grad(gz, x):
    cout, cin = gz.shape[3], x.shape[3]
    gw = np.zeros((cout, 9, cin))
    gzf = gz.reshape(-1, cout)
    for tap in range(9):
        gw[:, tap, :] = gzf.T @ _shift(x, tap // 3 - 1, tap % 3 - 1).reshape(-1, cin)
    return gw


def avgpool2(x):
    return 0.25 * ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))


def avgpool2_adj(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def upsample2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample2_adj(g):
    return (g[:, 0::2, 0::2] + g[:, 0::2, 1::2]) + (g[:, 1::2, 0::2] + g[:, 1::2, 1::2])


def depth_to_space(yf, u):
    """(B, H, W, 2u^2) -> (B, 2, uH, uW), channel c u^2 + sy u + sx -> plane c, (u y + sy, u x + sx)."""
    B, H, W, _ = yf.shape
    t = yf.reshape(B, H, W, 2, u, u).transpose(0, 3, 1, 4, 2, 5)
    return t.reshape(B, 2, u * H, u * W)


def space_to_depth(o, u):
    B, _, OH, OW = o.shape
    H, W = OH // u, OW // u
    return o.reshape(B, 2, H, u, W, u).transpose(0, 2, 4, 1, 3, 5).reshape(B, H, W, 2 * u * u)


PUPIL_HIDDEN = 64


def pupil_head_forward(e, W1, b1, W2, b2):
    """e (B, h, w, Cb) -> (pooled (B, Cb), hidden (B, 64), coefficients (B, K))"""
    pooled = e.mean(axis=(1, 2))
    hid = np.maximum(pooled @ W1.T + b1, 0.0)
    return pooled, hid, hid @ W2.T + b2


def pupil_head_backward(g_coef, pooled, hid, W1, W2, hw):
    """gradients (gW1, gb1, gW2, gb2) and the gradient of every pixel of e
    (B, Cb): each of the hw pixels receives g_pooled / hw"""
    gW2, gb2 = g_coef.T @ hid, g_coef.sum(axis=0)
    g_hid = (g_coef @ W2) * (hid > 0)
    gW1, gb1 = g_hid.T @ pooled, g_hid.sum(axis=0)
    return gW1, gb1, gW2, gb2, (g_hid @ W1) / hw


class Net:
    """Forward / backward of the network for a parameter vector."""

    def __init__(self, levels, base, in_ch, up, bf16=True, pupil_terms=0, pupil_weight=1.0):
        self.L, self.C, self.in_ch, self.u, self.bf16 = levels, base, in_ch, up, bf16
        self.cin_pad = (in_ch + 31) // 32 * 32
        self.convs = layer_shapes(levels, base, in_ch, up)
        self.K, self.lam = pupil_terms, pupil_weight
        self.cb = base << (levels - 1)
        self.nconv = sum(co * 9 * ci + co for ci, co, _ in self.convs)
        self.npupil = (PUPIL_HIDDEN * self.cb + PUPIL_HIDDEN + self.K * PUPIL_HIDDEN + self.K) if self.K else 0
        self.nparam = self.nconv + self.npupil

    def pupil_params(self, params):
        """(W1, b1, W2, b2) views of the pupil head's parameters"""
        p = np.asarray(params, dtype=np.float64)[self.nconv:]
        H, cb, K = PUPIL_HIDDEN, self.cb, self.K
        W1 = p[:H * cb].reshape(H, cb)
        b1 = p[H * cb:H * cb + H]
        W2 = p[H * cb + H:H * cb + H + K * H].reshape(K, H)
        b2 = p[H * cb + H + K * H:]
        return W1, b1, W2, b2

    def _r(self, x):
        return round_bf16(x) if self.bf16 else np.asarray(x, dtype=np.float64)

    def features(self, feats):
        """planar (B, Cf, n, n) -> NHWC padded to cin_pad channels"""
        B, Cf, n, _ = feats.shape
        x = np.zeros((B, n, n, self.cin_pad))
        x[..., :Cf] = np.transpose(feats, (0, 2, 3, 1))
        return self._r(x)

    def forward(self, params, feats):
        """(output planes (B, 2, uN, uN), cache); cache["coef"] holds the
        pupil coefficients (B, K) when the network has a pupil head"""
        P = split_params(np.asarray(params, dtype=np.float64)[:self.nconv], self.convs)
        Wb = [(self._r(w), b) for w, b in P]
        cache = {"W": Wb}
        r = self._r
        x = self.features(feats)
        cache["in"] = x
        k = 0
        enc = []
        for l in range(self.L):
            if l > 0:
                x = r(avgpool2(enc[-1]["eb"]))
            ea = r(np.maximum(conv3x3(x, *Wb[k]), 0.0))
            eb = r(np.maximum(conv3x3(ea, *Wb[k + 1]), 0.0))
            enc.append({"x": x, "ea": ea, "eb": eb, "k": k})
            k += 2
            x = eb
        dec = {}
        prev = enc[-1]["eb"]
        for l in range(self.L - 2, -1, -1):
            u_ = upsample2(prev)
            da = r(np.maximum(conv3x3(u_, *Wb[k]), 0.0))
            s = r(da + enc[l]["eb"])
            db = r(np.maximum(conv3x3(s, *Wb[k + 1]), 0.0))
            dec[l] = {"up": u_, "da": da, "s": s, "db": db, "k": k}
            k += 2
            prev = db
        yf = conv3x3(prev, *Wb[k])
        cache.update(enc=enc, dec=dec, last=prev, khead=k, yf=yf)
        if self.K:
            cache["pupil"] = pupil_head_forward(enc[-1]["eb"], *self.pupil_params(params))
            cache["coef"] = cache["pupil"][2]
        return depth_to_space(yf, self.u), cache

    def loss_and_grad(self, params, feats, targets, pupil_targets=None):
        out, c = self.forward(params, feats)
        d = out - targets
        count = d.size
        loss = float(np.sum(d * d) / count)
        g_pupil = None
        if self.K:
            dp = c["coef"] - pupil_targets
            loss += self.lam * float(np.sum(dp * dp)) / dp.size
            pooled, hid, _ = c["pupil"]
            W1, _, W2, _ = self.pupil_params(params)
            e = c["enc"][-1]["eb"]
            g_pupil = pupil_head_backward(2.0 * self.lam * dp / dp.size, pooled, hid, W1, W2,
                                          e.shape[1] * e.shape[2])
        gy_head = space_to_depth(2.0 * d / count, self.u)
        grads = [None] * len(self.convs)
        r, Wb = self._r, c["W"]

        def conv_bwd(k, x, y, gy, need_gx=True):
            relu = self.convs[k][2]
            gz = gy * (y > 0) if relu else gy
            gz = r(gz)
            grads[k] = (conv3x3_weight_grad(gz, x), gz.sum(axis=(0, 1, 2)))
            return conv3x3_input_grad(gz, Wb[k][0]) if need_gx else None

        g_last = conv_bwd(c["khead"], c["last"], None, gy_head)
        enc, dec = c["enc"], c["dec"]
        g = {}
        gprev = g_last
        for l in range(0, self.L - 1):
            dl = dec[l]
            gs = conv_bwd(dl["k"] + 1, dl["s"], dl["db"], gprev)
            g[("eb", l)] = gs.copy()
            gu = conv_bwd(dl["k"], dl["up"], dl["da"], gs)
            gprev = upsample2_adj(gu)
        g[("eb", self.L - 1)] = gprev
        if g_pupil is not None:
            g[("eb", self.L - 1)] = gprev + g_pupil[4][:, None, None, :]
        for l in range(self.L - 1, -1, -1):
            e = enc[l]
            gea = conv_bwd(e["k"] + 1, e["ea"], e["eb"], g[("eb", l)])
            gx = conv_bwd(e["k"], e["x"], e["ea"], gea, need_gx=l > 0)
            if l > 0:
                g[("eb", l - 1)] = g.get(("eb", l - 1), 0.0) + avgpool2_adj(gx)
        flat = []
        for gw, gb in grads:
            flat += [gw.ravel(), gb.ravel()]
        if g_pupil is not None:
            flat += [t.ravel() for t in g_pupil[:4]]
        return loss, np.concatenate(flat), out


# ------------------------------------------------------- parallax shifts
def _bilinear_clamped(q, y, x):
    """bilinear samples of plane q at (y, x) (arrays), coordinates clamped to
    the plane, cell corner (min(floor y, H-2), min(floor x, W-2))"""
    H, W = q.shape
    y = np.clip(y, 0.0, H - 1.0)
    x = np.clip(x, 0.0, W - 1.0)
    y0 = np.minimum(np.floor(y).astype(np.int64), H - 2)
    x0 = np.minimum(np.floor(x).astype(np.int64), W - 2)
    fy, fx = y - y0, x - x0
    a = q[y0, x0] + fx * (q[y0, x0 + 1] - q[y0, x0])
    b = q[y0 + 1, x0] + fx * (q[y0 + 1, x0 + 1] - q[y0 + 1, x0])
    return a + fy * (b - a)


def parallax_shift(ref, img, iters):
    """(dy, dx) with img(y, x) ~ ref(y - dy, x - dx): Lucas-Kanade on the
    mean/std-normalised planes, w(y, x) = img(y + dy, x + dx) bilinear with
    clamped coordinates, over the interior pixels g = central differences of
    (ref + w) / 2, e = w - ref, [sum g g^T] delta = -[sum g e], d += delta
    (what libfpm_amd's fpm_parallax_shifts computes, in float64)."""
    def normalised(a):
        a = np.asarray(a, dtype=np.float64)
        m, v = a.mean(), a.var()
        return (a - m) / np.sqrt(v) if v > 0 else a * 0.0
    r, q = normalised(ref), normalised(img)
    H, W = r.shape
    yy, xx = np.meshgrid(np.arange(1, H - 1, dtype=np.float64), np.arange(1, W - 1, dtype=np.float64), indexing="ij")
    rc = r[1:-1, 1:-1]
    d = np.zeros(2)
    for _ in range(iters):
        w = lambda oy, ox: _bilinear_clamped(q, yy + d[0] + oy, xx + d[1] + ox)  # noqa: E731
        wc = w(0, 0)
        g0 = 0.25 * ((r[2:, 1:-1] - r[:-2, 1:-1]) + (w(1, 0) - w(-1, 0)))
        g1 = 0.25 * ((r[1:-1, 2:] - r[1:-1, :-2]) + (w(0, 1) - w(0, -1)))
        e = wc - rc
        s00, s01, s11 = np.sum(g0 * g0), np.sum(g0 * g1), np.sum(g1 * g1)
        b0, b1 = np.sum(g0 * e), np.sum(g1 * e)
        det = s00 * s11 - s01 * s01
        if not det > 0:
            break
        d = d - np.array([s11 * b0 - s01 * b1, s00 * b1 - s01 * b0]) / det
    return d
