"""CPU reference of the closure surrogate (test infrastructure).

The surrogate is a 4-layer 3x3 convolutional network over the cell grid:
4 input channels -> 32 -> 32 -> 32 -> 2, ReLU after the first three layers.
Grids are laid out [batch, channel, ix, iy] (cell index iy + Ny*ix); the x
axis pads periodically when both x walls are periodic, else with zeros, and
likewise y.  Flat parameter order, per layer: W[co][tap][ci] with
tap = 3*(dx+1) + (dy+1), then b[co].

`forward`/`backward` mirror the GPU kernels' rounding: inputs, weights,
stored activations and the gradients fed to the products are rounded to
bfloat16 at the same points, products are summed in float64 here (the GPU
sums in float32), so the two agree to float32 accumulation error.  With
`bf16=False` they are the plain float64 network, checked against torch
autograd and finite differences in tests/test_surrogate_ref.py.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

CHANNELS = (4, 32, 32, 32, 2)
BETA1, BETA2, EPS = 0.9, 0.999, 1e-8


def layer_shapes():
    return [(CHANNELS[i + 1], 9, CHANNELS[i]) for i in range(4)]


def n_params() -> int:
    return sum(co * 9 * ci + co for co, _, ci in layer_shapes())


def unflatten(theta):
    """flat vector -> [(W[co,tap,ci], b[co])] as float64 tensors"""
    t = torch.as_tensor(np.asarray(theta, np.float64))
    out, o = [], 0
    for co, _, ci in layer_shapes():
        w = t[o:o + co * 9 * ci].reshape(co, 9, ci)
        o += co * 9 * ci
        b = t[o:o + co]
        o += co
        out.append((w, b))
    return out


def flatten(layers) -> np.ndarray:
    parts = []
    for w, b in layers:
        parts += [w.reshape(-1), b.reshape(-1)]
    return torch.cat(parts).detach().numpy().astype(np.float64)


def to_kernel(w):
    """W[co,tap,ci] -> torch conv weight [co,ci,3,3] (tap = 3*i + j)"""
    co, _, ci = w.shape
    return w.reshape(co, 3, 3, ci).permute(0, 3, 1, 2)


def from_kernel(k):
    co, ci = k.shape[:2]
    return k.permute(0, 2, 3, 1).reshape(co, 9, ci)


def bf16(t, on=True):
    return t.to(torch.bfloat16).to(t.dtype) if on else t


def pad(x, perx, pery):
    """[B,C,X,Y] -> [B,C,X+2,Y+2]"""
    x = F.pad(x, (0, 0, 1, 1), mode="circular") if perx else F.pad(x, (0, 0, 1, 1))
    x = F.pad(x, (1, 1, 0, 0), mode="circular") if pery else F.pad(x, (1, 1, 0, 0))
    return x


def unpad_grad(g, perx, pery):
    """adjoint of pad: fold the wrapped halo back (periodic) or drop it"""
    if pery:
        g = g.clone()
        g[:, :, :, 1] += g[:, :, :, -1]
        g[:, :, :, -2] += g[:, :, :, 0]
    g = g[:, :, :, 1:-1]
    if perx:
        g = g.clone()
        g[:, :, 1, :] += g[:, :, -1, :]
        g[:, :, -2, :] += g[:, :, 0, :]
    return g[:, :, 1:-1, :]


def conv(x, w, b, perx, pery):
    return F.conv2d(pad(x, perx, pery), to_kernel(w)) + b.view(1, -1, 1, 1)


def forward(theta, feat, perx, pery, bf=True):
    """feat [B,4,X,Y] -> (out [B,2,X,Y], cache)"""
    layers = unflatten(theta)
    x = bf16(torch.as_tensor(np.asarray(feat, np.float64)), bf)
    acts = [x]
    for li, (w, b) in enumerate(layers):
        z = conv(acts[-1], bf16(w, bf), b, perx, pery)
        if li < 3:
            acts.append(bf16(torch.relu(z), bf))
        else:
            out = z
    return out, acts


def loss_and_grad(theta, feat, target, perx, pery, bf=True):
    """mean squared error over all outputs and its gradient (flat)"""
    layers = unflatten(theta)
    out, acts = forward(theta, feat, perx, pery, bf)
    t = torch.as_tensor(np.asarray(target, np.float64))
    diff = out - t
    loss = float((diff * diff).mean())
    g = bf16(2.0 * diff / diff.numel(), bf)
    grads = [None] * 4
    for li in range(3, -1, -1):
        w, _ = layers[li]
        xin = pad(acts[li], perx, pery)
        dk = torch.nn.grad.conv2d_weight(xin, to_kernel(w).shape, g)
        grads[li] = (from_kernel(dk), g.sum(dim=(0, 2, 3)))
        if li == 0:
            break
        dx = unpad_grad(F.conv_transpose2d(g, to_kernel(bf16(w, bf))), perx, pery)
        g = bf16(dx * (acts[li] > 0), bf)
    return loss, flatten(grads), out


def adam(theta, m, v, grad, step, lr):
    """one Adam update (float64); step counts from 1"""
    m = BETA1 * m + (1.0 - BETA1) * grad
    v = BETA2 * v + (1.0 - BETA2) * grad * grad
    mh = m / (1.0 - BETA1 ** step)
    vh = v / (1.0 - BETA2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + EPS), m, v


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (the GPU's initializer draws its own)"""
    rng = np.random.default_rng(seed)
    parts = []
    for co, _, ci in layer_shapes():
        bound = np.sqrt(6.0 / (9 * ci))
        parts += [rng.uniform(-bound, bound, co * 9 * ci), np.zeros(co)]
    return np.concatenate(parts)
