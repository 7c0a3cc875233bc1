"""Learned warm start of the local return map (host side of include/plasticity_amd.h's pa_mlp_*).

The reference runs the local active-set loop of calculatePlasticity (calculatePlasticity.cc:311-441)
from zero slip.  Run to convergence (Max Solver Iterations > 1, stopping test
||b_PA||_inf <= Stress Tolerance, :322) it takes 3-9 local solves per plastic quadrature point.  A
small network predicts the converged slip increments from the trial state; the loop starts from
the prediction and its own stopping test still decides convergence, so the plain return map is
both the fallback and the judge of the prediction.

Everything here drives the GPU through the C ABI: training pairs come from the GPU qp kernels on
states of a GPU run, the network is trained by the library's own BF16 matrix-core kernels.
"""
import ctypes

import numpy as np

from . import native
from .native import check, f64, ptr

NIN, NOUT, NPARAM, NBLOB = 18, 12, 11472, 11520
LK = (32, 64, 64, 64)
LN = (64, 64, 64, 16)
OFF_W = (0, 2112, 6272, 10432)
OFF_B = (2048, 6208, 10368, 11456)
OFF_MEAN, OFF_SCALE, OFF_OSCALE = NPARAM, NPARAM + NIN, NPARAM + 2 * NIN


def init_blob(X, Y, seed=0):
    """Glorot-uniform weights (seeded), zero biases, input mean/scale and output scale from data."""
    rng = np.random.default_rng(seed)
    blob = np.zeros(NBLOB, np.float32)
    for l in range(4):
        fan_in = NIN if l == 0 else 64
        nout = NOUT if l == 3 else 64
        lim = np.sqrt(6.0 / (fan_in + nout))
        WT = np.zeros((LN[l], LK[l]), np.float32)
        WT[:nout, :fan_in] = rng.uniform(-lim, lim, size=(nout, fan_in))
        blob[OFF_W[l]:OFF_W[l] + LN[l] * LK[l]] = WT.ravel()
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    blob[OFF_MEAN:OFF_MEAN + NIN] = X.mean(axis=0)
    blob[OFF_SCALE:OFF_SCALE + NIN] = 1.0 / np.maximum(X.std(axis=0), 1e-12)
    rms = float(np.sqrt(np.mean(Y ** 2))) if Y.size else 1.0
    blob[OFF_OSCALE:OFF_OSCALE + NOUT] = rms if rms > 0 else 1.0
    return blob


class WarmStartNet:
    """pa_mlp handle: the 18-64-64-64-12 SiLU network on the GPU."""

    def __init__(self, blob=None):
        h = ctypes.c_void_p()
        check(native.lib().pa_mlp_create(ctypes.byref(h)))
        self._h = h
        if blob is not None:
            self.set_params(blob)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            native.lib().pa_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, blob):
        blob = np.ascontiguousarray(blob, np.float32)
        assert blob.size == NBLOB
        check(native.lib().pa_mlp_set_params(self._h, ptr(blob)))

    def get_params(self):
        blob = np.zeros(NBLOB, np.float32)
        check(native.lib().pa_mlp_get_params(self._h, ptr(blob)))
        return blob

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NIN)
        Y = np.zeros((len(X), NOUT), np.float32)
        check(native.lib().pa_mlp_predict(self._h, ctypes.c_longlong(len(X)), ptr(X), ptr(Y)))
        return Y

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NIN)
        Y = np.ascontiguousarray(Y, np.float32).reshape(-1, NOUT)
        g = np.zeros(NPARAM, np.float32)
        loss = ctypes.c_double()
        check(native.lib().pa_mlp_grad(self._h, ctypes.c_longlong(len(X)), ptr(X), ptr(Y), ptr(g), ctypes.byref(loss)))
        return loss.value, g

    def train(self, X, Y, epochs=20, batch=1024, lr=2e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NIN)
        Y = np.ascontiguousarray(Y, np.float32).reshape(-1, NOUT)
        losses = np.zeros(epochs)
        check(native.lib().pa_mlp_train(self._h, ctypes.c_longlong(len(X)), ptr(X), ptr(Y), ctypes.c_int(epochs),
                                        ctypes.c_int(batch), ctypes.c_float(lr), ctypes.c_ulonglong(seed), ptr(losses)))
        return losses


def qp_update_warm(ctx, F, Fp, Fe_conv, s, grain, net=None, flag=0):
    """pa_qp_update_warm over a batch: the local loop from the network's prediction (net=None: from
    zero, the reference's start).  Returns the pa_qp_update outputs plus iters (local solves),
    feat (network inputs) and x0 (predictions, net only)."""
    F = f64(F).reshape(-1, 9)
    n = len(F)
    ns = ctx.n_slip
    Fp = f64(Fp).reshape(n, 9)
    Fe_conv = f64(Fe_conv).reshape(n, 9)
    s = f64(s).reshape(n, ns)
    grain = np.ascontiguousarray(grain, dtype=np.uint32).reshape(n)
    out = {k: np.zeros((n, 9)) for k in ("P", "T", "Fe", "Fp")}
    out["dPdF"] = np.zeros((n, 81)) if flag else None
    out["s"] = np.zeros((n, ns))
    out["slip_inc"] = np.zeros((n, ns))
    out["n_active"] = np.zeros(n, np.int32)
    out["iters"] = np.zeros(n, np.int32)
    out["feat"] = np.zeros((n, NIN), np.float32)
    out["x0"] = np.zeros((n, NOUT), np.float32) if net is not None else None
    check(native.lib().pa_qp_update_warm(ctx._h, net.handle if net is not None else None, ctypes.c_longlong(n), ptr(F), ptr(Fp),
                                         ptr(Fe_conv), ptr(s), ptr(grain), ctypes.c_int(flag), ptr(out["P"]), ptr(out["T"]),
                                         ptr(out["dPdF"]), ptr(out["Fe"]), ptr(out["Fp"]), ptr(out["s"]), ptr(out["slip_inc"]),
                                         ptr(out["n_active"]), ptr(out["iters"]), ptr(out["feat"]), ptr(out["x0"])))
    return out


def qp_deformation_gradients(u, nel, span, elems=None):
    """F = I + grad u at the 2x2x2 Gauss points of every Q1 element of the structured box
    (getElementalValues.cc:129-138), qp order Q = 8e + qx + 2qy + 4qz.  u: 3 per node.  elems
    (optional): only these elements, in the given order."""
    nx, ny, nz = nel
    h = np.asarray(span, float) / np.array([nx, ny, nz])
    U = np.asarray(u, float).reshape(nz + 1, ny + 1, nx + 1, 3)
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    if elems is None:
        # corner displacements of every element, vertex v = vx + 2vy + 4vz
        corners = np.stack([U[vz:vz + nz, vy:vy + ny, vx:vx + nx] for vz in (0, 1) for vy in (0, 1) for vx in (0, 1)], axis=3)
        corners = corners.reshape(nx * ny * nz, 8, 3)  # element e = ex + nx (ey + ny ez)
    else:
        e = np.asarray(elems, np.int64)
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        corners = np.stack([U[ez + vz, ey + vy, ex + vx] for vz in (0, 1) for vy in (0, 1) for vx in (0, 1)], axis=1)
    F = np.zeros((len(corners), 8, 3, 3))
    for q in range(8):
        xi = (g[q & 1], g[(q >> 1) & 1], g[(q >> 2) & 1])
        dN = np.zeros((8, 3))
        for v in range(8):
            vv = (v & 1, (v >> 1) & 1, (v >> 2) & 1)
            val = [xi[d] if vv[d] else 1.0 - xi[d] for d in range(3)]
            der = [1.0 if vv[d] else -1.0 for d in range(3)]
            dN[v] = [der[0] * val[1] * val[2] / h[0], val[0] * der[1] * val[2] / h[1], val[0] * val[1] * der[2] / h[2]]
        F[:, q] = np.eye(3) + np.einsum("evi,vj->eij", corners, dN)
    return F.reshape(-1, 9)


def collect_trial_states(ctx, nel, span, elem_grain, increments, max_newton):
    """Run `increments` increments of ctx and record, before every Newton iteration after the
    first, the trial state each quadrature point's next assembly sees: F from the current u, with
    the converged history (Fp_conv, Fe_conv, s_conv) of the increment.  Returns a dict of arrays."""
    nq = 8 * int(np.prod(nel))
    grain = np.repeat(np.asarray(elem_grain, np.uint32), 8)
    out = {"F": [], "Fp": [], "Fe": [], "s": [], "grain": []}
    for inc in range(increments):
        Fp = ctx.get_state("Fp").reshape(nq, 9)
        Fe = ctx.get_state("Fe").reshape(nq, 9)
        s = ctx.get_state("s").reshape(nq, -1)
        for it in range(max_newton):
            ctx.newton_iteration(inc, it)
            u = ctx.get_state("u")
            out["F"].append(qp_deformation_gradients(u, nel, span))
            out["Fp"].append(Fp)
            out["Fe"].append(Fe)
            out["s"].append(s)
            out["grain"].append(grain)
        ctx.commit_increment()
    return {k: np.concatenate(v) for k, v in out.items()}
