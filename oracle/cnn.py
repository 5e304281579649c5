"""Oracle: the 3x3x3 circular-padding CNN closure in torch (CPU, float32).
TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

The HIP path keeps weights, activations, pre-activations and data gradients
in BF16 and accumulates in FP32; with ``emulate=True`` this oracle rounds at
the same storage points, so the two agree to FP32 accumulation-order noise
amplified by the occasional BF16 rounding flip.  With ``emulate=False`` it
is the plain FP32 network, whose explicit backward is pinned against torch
autograd (tests/test_cnn_oracle.py).  The 1x1x1 special case (only the centre
tap non-zero) is the reference's pointwise layer: pointwise_apply
(src/nets.jl:63-81) followed by the activation.

Parameter layout (as sc_cnn_get_params): per layer W[t][cout_p][cin_p] with
tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) (output voxel v reads input v+(dz,dy,dx)),
then three hidden biases.  Tensors: x (B, n^3, cin_p) channels-last, voxel
index x + n (y + n z).
"""
import numpy as np
import torch
import torch.nn.functional as F

C_GELU = 0.7978845608028654


def bf(t, on=True):
    return t.to(torch.bfloat16).to(t.dtype) if on else t


def gelu(z):
    return 0.5 * z * (1 + torch.tanh(C_GELU * (z + 0.044715 * z**3)))


def gelu_grad(z):
    u = C_GELU * (z + 0.044715 * z**3)
    t = torch.tanh(u)
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * z * z)


def torch_weight(W):
    """(27, cout, cin) -> conv3d weight (cout, cin, 3, 3, 3)"""
    W = torch.as_tensor(np.asarray(W))
    if W.dtype not in (torch.float32, torch.float64):
        W = W.to(torch.float32)
    co, ci = W.shape[1], W.shape[2]
    return W.reshape(3, 3, 3, co, ci).permute(3, 4, 0, 1, 2).contiguous()


def conv(x, Wt):
    return F.conv3d(F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular"), Wt)


def conv_data_grad(d, Wt):
    """transpose of conv: correlation with the mirrored, channel-swapped kernel"""
    return conv(d, Wt.flip(2, 3, 4).transpose(0, 1).contiguous())


def conv_weight_grad(x, d, shape):
    return torch.nn.grad.conv3d_weight(F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular"), shape, d)


def to_ncdhw(a, n):
    B, _, C = a.shape
    return a.reshape(B, n, n, n, C).permute(0, 4, 1, 2, 3).contiguous()


def to_nvc(a):
    B, C = a.shape[:2]
    return a.permute(0, 2, 3, 4, 1).reshape(B, -1, C).contiguous()


def split_params(h, shapes, width):
    Ws, off = [], 0
    for co, ci in shapes:
        Ws.append(h[off: off + 27 * co * ci].reshape(27, co, ci))
        off += 27 * co * ci
    bs = [h[off + i * width: off + (i + 1) * width] for i in range(3)]
    return Ws, bs


def tracefree(y, dim=1):
    """trace removal of the packed stress channels (xx, yy, zz, xy, yz, zx):
    the first three lose their mean (the `symm` layer, src/nets.jl:275-293).
    The map is an orthogonal projection, so it is also its own transpose."""
    y = y.clone()
    d = y.narrow(dim, 0, 3)
    d -= d.sum(dim=dim, keepdim=True) / 3
    return y


def forward(Ws, bs, x, n, emulate=True, dtype=torch.float32, tracefree_out=False):
    """x: (B, n^3, cin_p) float tensor.  Returns y (B, n^3, cout_p) and the
    stored tensors needed by backward.  ``dtype=torch.float64`` (with
    emulate=False) gives the smooth FP64 network for finite differences;
    ``tracefree_out`` applies the trace removal to the output."""
    Wt = [bf(torch_weight(W).to(dtype), emulate) for W in Ws]
    a = to_ncdhw(bf(torch.as_tensor(x, dtype=dtype), emulate), n)
    acts, zs = [a], []
    for l in range(3):
        z = conv(a, Wt[l]) + torch.as_tensor(np.asarray(bs[l]), dtype=dtype).view(1, -1, 1, 1, 1)
        zs.append(bf(z, emulate))
        a = bf(gelu(z), emulate)
        acts.append(a)
    y = conv(a, Wt[3])
    if tracefree_out:
        y = tracefree(y)
    return to_nvc(y), dict(Wt=Wt, acts=acts, zs=zs, tracefree=tracefree_out)


def backward(store, dy, n, emulate=True):
    """dy: (B, n^3, cout_p).  Returns (weight grads [(27,cout,cin)], bias
    grads, dx (B, n^3, cin_p))."""
    Wt, acts, zs = store["Wt"], store["acts"], store["zs"]
    d = to_ncdhw(bf(torch.as_tensor(dy, dtype=Wt[0].dtype), emulate), n)
    if store.get("tracefree"):
        d = tracefree(d)
    gW, gb = [None] * 4, [None] * 3
    gW[3] = conv_weight_grad(acts[3], d, Wt[3].shape)
    d = bf(conv_data_grad(d, Wt[3]) * gelu_grad(zs[2]), emulate)
    for l in (2, 1, 0):
        gW[l] = conv_weight_grad(acts[l], d, Wt[l].shape)
        gb[l] = d.sum(dim=(0, 2, 3, 4))
        if l > 0:
            d = bf(conv_data_grad(d, Wt[l]) * gelu_grad(zs[l - 1]), emulate)
        else:
            d = conv_data_grad(d, Wt[0])
    flat = [g.permute(2, 3, 4, 0, 1).reshape(27, g.shape[0], g.shape[1]) for g in gW]
    return flat, gb, to_nvc(d)


def mse(y, tgt, nout):
    """loss and dL/dy for the MSE over the first nout channels"""
    y = torch.as_tensor(y, dtype=torch.float32)
    tgt = torch.as_tensor(tgt, dtype=torch.float32)
    e = y[..., :nout] - tgt
    N = e.numel()
    dy = torch.zeros_like(y)
    dy[..., :nout] = 2 * e / N
    return float((e.double() ** 2).sum() / N), dy


def rel_loss(y, tgt, nout):
    """loss and dL/dy of the reference's relative loss over the batch,
    sum|y - t|^2 / (sum|t|^2 + eps(Float32)) (create_loss,
    src/training.jl:3-9), on the first nout channels"""
    y = torch.as_tensor(y, dtype=torch.float32)
    tgt = torch.as_tensor(tgt, dtype=torch.float32)
    e = (y[..., :nout] - tgt).double()
    den = float((tgt.double() ** 2).sum()) + float(np.finfo(np.float32).eps)
    dy = torch.zeros_like(y)
    dy[..., :nout] = (2 * e / den).float()
    return float((e**2).sum() / den), dy


def pointwise_mlp(Ws, bs, x):
    """the 1x1x1 case: pointwise_apply (X W + b) then GELU per hidden layer,
    no activation on the last layer (src/nets.jl:63-81, 247-298)"""
    a = np.asarray(x, np.float64)
    for l in range(3):
        W = np.asarray(Ws[l][13], np.float64)  # centre tap, (cout, cin)
        z = a @ W.T + np.asarray(bs[l], np.float64)
        a = 0.5 * z * (1 + np.tanh(C_GELU * (z + 0.044715 * z**3)))
    return a @ np.asarray(Ws[3][13], np.float64).T
