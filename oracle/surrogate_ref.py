"""CPU restatement of the learned indicator surrogate (test infrastructure only: the
checker of the device network's forward pass and of its hand-written backward pass;
nothing in the product imports it).

The network (no counterpart in the reference; it stands in for the dual solve of
estimate_error, src/em_fwd.cc:1102-1180):

    h0 = relu(F W0 + b0)
    h_l = relu([h_{l-1}, mean_nb h_{l-1}] W_l + b_l),   l = 1..3
    y = h_3 wo + bo

with GEMM operands rounded to bf16 (round to nearest even) and products accumulated
in float64 here (fp32 on the device).  `loss_grad` is the analytic gradient of the
mean squared error; `fd_grad` is a central finite difference of the same loss
without bf16 rounding, which pins the analytic formulas independently.
"""
import numpy as np

FEAT, HID, LAYERS = 16, 64, 3


def layout():
    off, o = {}, 0
    off["w0"] = o; o += FEAT * HID
    off["b0"] = o; o += HID
    for l in range(LAYERS):
        off["w%d" % (l + 1)] = o; o += 2 * HID * HID
        off["b%d" % (l + 1)] = o; o += HID
    off["wo"] = o; o += HID
    off["bo"] = o; o += 1
    off["total"] = o
    return off


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    L = layout()
    p = np.asarray(p, dtype=np.float64)
    W = {"w0": p[L["w0"]:L["w0"] + FEAT * HID].reshape(FEAT, HID), "b0": p[L["b0"]:L["b0"] + HID]}
    for l in range(1, LAYERS + 1):
        W["w%d" % l] = p[L["w%d" % l]:L["w%d" % l] + 2 * HID * HID].reshape(2 * HID, HID)
        W["b%d" % l] = p[L["b%d" % l]:L["b%d" % l] + HID]
    W["wo"] = p[L["wo"]:L["wo"] + HID]
    W["bo"] = p[L["bo"]]
    return W


def nb_mean(h, nb):
    valid = nb >= 0
    deg = valid.sum(1)
    s = np.zeros_like(h)
    for f in range(4):
        idx = nb[:, f]
        s[valid[:, f]] += h[idx[valid[:, f]]]
    return np.where(deg[:, None] > 0, s / np.maximum(deg, 1)[:, None], 0.0)


def nb_mean_T(d, nb):
    """Transpose of nb_mean applied to d (scatter of d[n] / deg(n) to n's neighbours)."""
    valid = nb >= 0
    deg = valid.sum(1)
    out = np.zeros_like(d)
    for f in range(4):
        rows = np.nonzero(valid[:, f])[0]
        np.add.at(out, nb[rows, f], d[rows] / deg[rows][:, None])
    return out


def forward(p, F, nb, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W = unpack(p)
    cache = []
    z = r(F) @ r(W["w0"]) + W["b0"]
    h = np.maximum(z, 0.0)
    cache.append((np.asarray(F, dtype=np.float64), z))
    for l in range(1, LAYERS + 1):
        cat = np.concatenate([h, nb_mean(h, nb)], axis=1)
        z = r(cat) @ r(W["w%d" % l]) + W["b%d" % l]
        cache.append((cat, z))
        h = np.maximum(z, 0.0)
    y = r(h) @ r(W["wo"]) + W["bo"]
    return y, h, cache


def loss_grad(p, F, nb, target, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W = unpack(p)
    T = F.shape[0]
    y, h, cache = forward(p, F, nb, round_bf16)
    e = y - np.asarray(target, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / T
    L = layout()
    g = np.zeros(L["total"])
    g[L["wo"]:L["wo"] + HID] = r(h).T @ r(dy)
    g[L["bo"]] = dy.sum()
    dh = np.outer(dy, W["wo"])  # an fp32 product on the device (no matrix core)
    for l in range(LAYERS, 0, -1):
        cat, z = cache[l]
        dz = dh * (z > 0)
        g[L["w%d" % l]:L["w%d" % l] + 2 * HID * HID] = (r(cat).T @ r(dz)).ravel()
        g[L["b%d" % l]:L["b%d" % l] + HID] = dz.sum(0)
        dcat = r(dz) @ r(W["w%d" % l]).T
        dh = dcat[:, :HID] + nb_mean_T(dcat[:, HID:], nb)
    F0, z0 = cache[0]
    dz = dh * (z0 > 0)
    g[L["w0"]:L["w0"] + FEAT * HID] = (r(F0).T @ r(dz)).ravel()
    g[L["b0"]:L["b0"] + HID] = dz.sum(0)
    return y, loss, g


def fd_grad(p, F, nb, target, idx, eps=1e-6):
    """Central differences of the (unrounded) loss for parameter indices idx."""
    p = np.asarray(p, dtype=np.float64)
    out = []
    for i in idx:
        q = p.copy(); q[i] += eps
        lp = np.mean((forward(q, F, nb, False)[0] - target) ** 2)
        q[i] -= 2 * eps
        lm = np.mean((forward(q, F, nb, False)[0] - target) ** 2)
        out.append((lp - lm) / (2 * eps))
    return np.array(out)
