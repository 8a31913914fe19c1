"""Float32 PyTorch reference of the learned denoiser (CPU): the same features, layer shapes and bf16 rounding
points as the HIP kernels, for the parity tests of the denoiser."""
import numpy as np
import torch

LAYERS = ((8, 32), (32, 32), (32, 32), (32, 32), (32, 3))


def bf16_round(x):
    """float32 -> nearest-even bfloat16, returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def features(accum, mom, zmin, spp):
    """(h, w, 8) input features (bf16-rounded) and (h, w, 4) mean with the variance of the mean in w; f32
    arithmetic in the kernel's order."""
    f = np.float32
    n = np.maximum(accum[..., 3], f(1.0))
    mr, mg, mb = accum[..., 0] / n, accum[..., 1] / n, accum[..., 2] / n
    vr = np.maximum(mom[..., 0] / n - mr * mr, f(0.0))
    vg = np.maximum(mom[..., 1] / n - mg * mg, f(0.0))
    vb = np.maximum(mom[..., 2] / n - mb * mb, f(0.0))
    var = (vr + vg + vb) / f(3.0)
    x = np.zeros(accum.shape[:2] + (8,), np.float32)
    x[..., 0], x[..., 1], x[..., 2] = mr, mg, mb
    x[..., 3] = np.sqrt(var)
    x[..., 4] = f(8.0) * np.sqrt(var / n)
    hit = zmin < f(3.0e38)
    x[..., 5] = np.where(hit, f(1.0) / (f(1.0) + np.where(hit, zmin, f(0.0))), f(0.0))
    x[..., 6] = mom[..., 3] / (n * f(spp))
    mean = np.stack([mr, mg, mb, (vr + vg + vb) / n], -1).astype(np.float32)
    return bf16_round(x), mean


def split(params):
    out, o = [], 0
    for cin, cout in LAYERS:
        w = params[o:o + cout * 9 * cin].reshape(cout, 3, 3, cin)
        o += cout * 9 * cin
        out.append((w, params[o:o + cout]))
        o += cout
    return out


class _RoundST(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return torch.from_numpy(bf16_round(x.detach().numpy()))

    @staticmethod
    def backward(ctx, g):
        return g


def forward(x8, mean, params, grad=False):
    """x8 (b, h, w, 8), mean (b, h, w, 4), params (n,) -> output rgb (b, h, w, 3) as a torch tensor.  Weights are
    bf16-rounded; hidden activations are rounded to bf16 after the ReLU (straight-through for gradients)."""
    x = torch.from_numpy(np.ascontiguousarray(x8, np.float32)).permute(0, 3, 1, 2)
    ws = []
    for w, b in split(np.asarray(params, np.float32)):
        wt = torch.from_numpy(bf16_round(w)).permute(0, 3, 1, 2).contiguous().requires_grad_(grad)
        bt = torch.from_numpy(np.ascontiguousarray(b, np.float32)).requires_grad_(grad)
        ws.append((wt, bt))
    h = x
    for li, (wt, bt) in enumerate(ws):
        h = torch.nn.functional.conv2d(h, wt, bt, padding=1)
        if li < len(ws) - 1:
            h = _RoundST.apply(torch.relu(h))
    m = torch.from_numpy(np.ascontiguousarray(mean[..., :3], np.float32)).permute(0, 3, 1, 2)
    return (m + h).permute(0, 2, 3, 1), ws


def pair_weight(ensembles):
    """Training weight of a pair whose input averages `ensembles` ensembles."""
    return float(np.sqrt(np.float32(ensembles / 4.0)))


def loss_grad(x8, mean, target, params, weights=None):
    out, ws = forward(x8, mean, params, grad=True)
    t = torch.from_numpy(np.ascontiguousarray(target[..., :3], np.float32))
    wt = torch.ones(out.shape[0]) if weights is None else torch.tensor(np.asarray(weights, np.float32))
    loss = (((out - t) ** 2).mean(dim=(1, 2, 3)) * wt).mean()
    loss.backward()
    g = []
    for wt, bt in ws:
        g.append(wt.grad.permute(0, 2, 3, 1).reshape(-1).numpy())
        g.append(bt.grad.numpy())
    return float(loss.detach()), np.concatenate(g)
