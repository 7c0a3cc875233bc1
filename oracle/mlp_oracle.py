"""CPU restatement of the warm-start network (TEST INFRASTRUCTURE ONLY).

Loaded by tests/ and nothing else; the product runs the network on the GPU
(plasticity_amd/csrc/cp_mlp.h).  There is no reference counterpart: the
reference's local loop starts from zero slip (calculatePlasticity.cc:311).

Network: 18 -> 64 -> 64 -> 64 -> 12, SiLU, input (x - mean) * scale, output
y * out_scale.  Parameter blob: per layer W^T[N][K] (padded K, N = (32,64),
(64,64), (64,64), (64,16)) then b[N]; then mean[18], scale[18], out_scale[12].

quant=True mirrors the GPU's arithmetic: the operands of every matrix product
(weights, activations, output gradients) are rounded to bfloat16 (round to
nearest even), products are summed in double here (FP32 on the GPU), biases and
bias gradients stay unrounded.  quant=False is plain float64, which the
finite-difference check of the gradient uses.
"""
import numpy as np

NIN, NOUT, NH = 18, 12, 64
LK = (32, 64, 64, 64)
LN = (64, 64, 64, 16)
OFF_W = (0, 2112, 6272, 10432)
OFF_B = (2048, 6208, 10368, 11456)
NPARAM = 11472
OFF_MEAN, OFF_SCALE, OFF_OSCALE = NPARAM, NPARAM + NIN, NPARAM + 2 * NIN
NBLOB = NPARAM + 2 * NIN + NOUT


def bf16(x):
    """float32 -> bfloat16 -> float32, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def layers(blob):
    """[(W^T (N x K), b (N))] of the four layers, float64."""
    blob = np.asarray(blob)
    out = []
    for l in range(4):
        WT = blob[OFF_W[l]:OFF_W[l] + LN[l] * LK[l]].astype(np.float64).reshape(LN[l], LK[l])
        b = blob[OFF_B[l]:OFF_B[l] + LN[l]].astype(np.float64)
        out.append((WT, b))
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def _q(x, quant):
    """bf16-rounded (via float32) when quant, else the float64 value itself."""
    return bf16(np.asarray(x, np.float32)).astype(np.float64) if quant else np.asarray(x, np.float64)


def forward(blob, X, quant=True):
    """Returns (y normalised (n x 16), cache).  Columns 12..15 are padding."""
    blob = np.asarray(blob)
    X = np.asarray(X)
    n = X.shape[0]
    if quant:
        mean = blob[OFF_MEAN:OFF_MEAN + NIN].astype(np.float32)
        scale = blob[OFF_SCALE:OFF_SCALE + NIN].astype(np.float32)
        xn = (X.astype(np.float32) - mean) * scale
    else:
        xn = (X.astype(np.float64) - blob[OFF_MEAN:OFF_MEAN + NIN]) * blob[OFF_SCALE:OFF_SCALE + NIN]
    a = np.zeros((n, LK[0]))
    a[:, :NIN] = _q(xn, quant)
    acts, zs = [a], []
    L = layers(blob)
    for l in range(4):
        WT, b = L[l]
        z = a @ _q(WT, quant).T + b
        if l < 3:
            zs.append(z)
            a = _q(_silu(z), quant)
            acts.append(a)
        else:
            y = z
    return y, {"acts": acts, "zs": zs, "L": L}


def predict(blob, X, quant=True):
    y, _ = forward(blob, X, quant)
    return y[:, :NOUT] * np.asarray(blob[OFF_OSCALE:OFF_OSCALE + NOUT], np.float64)


def loss_grad(blob, X, Y, quant=True):
    """Loss mean_{samples, outputs} (y - Y/out_scale)^2 and its gradient w.r.t. the NPARAM weights
    and biases (zero on padding)."""
    blob = np.asarray(blob)
    y, c = forward(blob, X, quant)
    n = X.shape[0]
    osc = np.asarray(blob[OFF_OSCALE:OFF_OSCALE + NOUT], np.float64)
    if quant:
        t = (np.asarray(Y, np.float32) / blob[OFF_OSCALE:OFF_OSCALE + NOUT].astype(np.float32)).astype(np.float64)
    else:
        t = np.asarray(Y, np.float64) / osc
    d = np.zeros((n, LN[3]))
    d[:, :NOUT] = y[:, :NOUT] - t
    inv = 1.0 / (n * NOUT)
    loss = float((d ** 2).sum() * inv)
    G = 2.0 * d * inv
    grad = np.zeros(NPARAM)
    acts, zs, L = c["acts"], c["zs"], c["L"]
    for l in (3, 2, 1, 0):
        Gq = _q(G, quant)
        dWT = Gq.T @ acts[l]                     # (N x K)
        grad[OFF_W[l]:OFF_W[l] + LN[l] * LK[l]] = dWT.ravel()
        grad[OFF_B[l]:OFF_B[l] + LN[l]] = G.sum(axis=0)
        if l > 0:
            WT = _q(L[l][0], quant)
            dA = Gq @ WT                         # (n x K)
            G = dA * _dsilu(zs[l - 1])
    return loss, grad
