"""CPU oracle of the value network (test infrastructure only: tests/ use it
to check the GPU kernels; nothing in the product imports it).

A numpy restatement, in float64, of the network pca_amd/value.py runs with
the kernels of include/pcavalue.h: per-test features -> 512 -> 256 (ReLU),
mean over valid tests, depth embedding, -> 256 (ReLU) -> 10 outputs (a
reach logit and logits over 0..8 remaining steps), binary cross-entropy on
the reach logit plus 0.5 x the mean cross-entropy of the remaining steps
over the positive states (the loss of python/train.py:44-60), and the
backward pass.  With bf16=True it rounds at the same
points the kernels store BF16 (weights, activations, the gradients fed to
the next GEMM), so the GPU must agree up to FP32 accumulation order; with
bf16=False it is the exact float64 function whose gradients
test_value_oracle_gradients checks against central finite differences.

The reference's model (python/model.py:16-84) has another trunk; its two
heads and their loss are what is kept, so what is followed here is that
loss and the mathematics.
"""
import numpy as np

PER_TEST, H1, H2, EMB, H3, MAX_T = 180, 512, 256, 16, 256, 32
NDIST = 9              # remaining steps 0..8
NH = 1 + NDIST         # reach logit + remaining-steps logits
DIST_WEIGHT = 0.5


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def features(rec):
    """[B, 32, 16] uint16 records -> [B*32, 180] 0/1 features."""
    rec = np.asarray(rec, dtype=np.uint32)
    words = rec[..., :11]                                   # in0, in1, target, r0..r7
    bits = (words[..., None] >> np.arange(16)) & 1          # [B, 32, 11, 16]
    flags = (rec[..., 11:12] >> np.arange(4)) & 1           # [B, 32, 4]
    f = np.concatenate([bits.reshape(*rec.shape[:2], 176), flags], axis=-1)
    return f.reshape(-1, PER_TEST).astype(np.float64)


def unpack(p):
    """Flat parameter vector (pca_amd.value.ValueNet.LAYOUT order) -> dict."""
    shapes = [("W1", (PER_TEST, H1)), ("b1", (H1,)), ("W2", (H1, H2)), ("b2", (H2,)), ("E", (9, EMB)),
              ("W3", (H2 + EMB, H3)), ("b3", (H3,)), ("W4", (H3, NH)), ("b4", (NH,))]
    out, off = {}, 0
    for name, shape in shapes:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[off:off + n], dtype=np.float64).reshape(shape)
        off += n
    return out


def n_params():
    return (PER_TEST * H1 + H1 + H1 * H2 + H2 + 9 * EMB + (H2 + EMB) * H3 + H3 + H3 * NH + NH)


def forward_backward(p, X, nvalid, depth, label, remaining=None, dist_weight=DIST_WEIGHT, bf16_mode=True):
    """Returns (head outputs [B, NH], loss, flat gradient like p).  The
    distance term is taken over the states with label > 0.5, against
    remaining (clamped to 0..8); remaining=None leaves the reach term
    alone."""
    P = unpack(p)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, W3, W4 = r(P["W1"]), r(P["W2"]), r(P["W3"]), r(P["W4"])
    B = len(nvalid)
    h1 = r(np.maximum(X @ W1 + P["b1"], 0))
    h2 = r(np.maximum(h1 @ W2 + P["b2"], 0))
    h2b = h2.reshape(B, MAX_T, H2)
    mask = (np.arange(MAX_T)[None, :] < np.asarray(nvalid)[:, None]).astype(np.float64)
    pooled = (h2b * mask[..., None]).sum(1) / np.maximum(np.asarray(nvalid), 1)[:, None]
    z = r(np.concatenate([pooled, P["E"][np.asarray(depth)]], axis=1))
    h3 = r(np.maximum(z @ W3 + P["b3"], 0))
    y = h3 @ W4 + P["b4"]                                   # [B, NH]
    t = np.asarray(label, dtype=np.float64)
    x = y[:, 0]
    loss = np.mean(np.maximum(x, 0) - t * x + np.log1p(np.exp(-np.abs(x))))
    dy = np.zeros_like(y)
    dy[:, 0] = (1 / (1 + np.exp(-x)) - t) / B
    pos = t > 0.5
    n_pos = int(pos.sum())
    if remaining is not None and dist_weight and n_pos:
        rem = np.clip(np.asarray(remaining), 0, NDIST - 1)
        d = y[pos, 1:]
        m = d.max(1, keepdims=True)
        lse = (m + np.log(np.exp(d - m).sum(1, keepdims=True)))[:, 0]
        loss += dist_weight * np.mean(lse - d[np.arange(n_pos), rem[pos]])
        soft = np.exp(d - lse[:, None])
        soft[np.arange(n_pos), rem[pos]] -= 1
        dy[pos, 1:] = dist_weight * soft / n_pos
    dy16 = r(dy)
    g = {}
    g["W4"] = h3.T @ dy16
    g["b4"] = dy.sum(0)
    dh3 = (dy16 @ W4.T) * (h3 > 0)
    dh3_16 = r(dh3)
    g["W3"] = z.T @ dh3_16
    g["b3"] = dh3.sum(0)
    dz = dh3_16 @ W3.T
    dpool = dz[:, :H2] / np.maximum(np.asarray(nvalid), 1)[:, None]
    g["E"] = np.zeros_like(P["E"])
    np.add.at(g["E"], np.asarray(depth), dz[:, H2:])
    dh2 = (np.repeat(dpool[:, None, :], MAX_T, axis=1) * mask[..., None]).reshape(-1, H2) * (h2 > 0)
    dh2_16 = r(dh2)
    g["W2"] = h1.T @ dh2_16
    g["b2"] = dh2.sum(0)
    dh1 = (dh2_16 @ W2.T) * (h1 > 0)
    dh1_16 = r(dh1)
    g["W1"] = X.T @ dh1_16
    g["b1"] = dh1.sum(0)
    flat = np.concatenate([g[n].reshape(-1) for n in ("W1", "b1", "W2", "b2", "E", "W3", "b3", "W4", "b4")])
    return y, loss, flat
