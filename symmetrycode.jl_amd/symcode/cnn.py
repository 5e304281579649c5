"""Host-side wrapper of the learned closure (sc_cnn_* in include/symcode.h):
a 3x3x3 circular-padding CNN on BF16 MFMA with FP32 master weights and Adam.

torch supplies device buffers only; every operation runs in
libsymcode_hip.so.  Tensors: input x bf16 (batch, n^3, cin_p), output y fp32
(batch, n^3, cout_p), target fp32 (batch, n^3, cout), dy bf16 like x.
"""
import ctypes

import numpy as np
import torch

from ._lib import check_cnn, lib


def _p(t):
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class ConvClosureNet:
    def __init__(self, n, batch_max=1, cin=4, width=64, cout=6, seed=0, device=0, tracefree=False):
        """tracefree: the output's xx, yy, zz channels lose their mean (the
        reference nets' last layer, src/nets.jl:275-293; needs cout 6)"""
        net = ctypes.c_void_p()
        rc = lib().sc_cnn_create(n, batch_max, cin, width, cout, seed, ctypes.byref(net))
        if rc != 0:
            if rc == 1:
                raise ValueError("sc_cnn_create: n must be a multiple of 64 and <= 128, width 64, cin <= 8, cout <= 16")
            raise RuntimeError("sc_cnn_create failed: " + lib().sc_cnn_last_error(None).decode())
        self.net = net
        info = (ctypes.c_int64 * 8)()
        check_cnn(lib().sc_cnn_info(net, info), net)
        self.n, self.batch_max, self.cin, self.width, self.cout, self.cin_p, self.cout_p, self.nparams = list(info)
        self.device = torch.device("cuda", device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check_cnn(lib().sc_cnn_set_stream(net, ctypes.c_void_p(stream)), net)
        self.nvox = self.n**3
        self.tracefree = False
        if tracefree:
            self.set_tracefree(True)

    def set_tracefree(self, on=True):
        if on and self.cout != 6:
            raise ValueError("trace removal needs cout = 6 (xx, yy, zz, xy, yz, zx)")
        check_cnn(lib().sc_cnn_set_tracefree(self.net, 1 if on else 0), self.net)
        self.tracefree = bool(on)

    def __del__(self):
        try:
            if getattr(self, "net", None):
                lib().sc_cnn_destroy(self.net)
                self.net = None
        except Exception:
            pass

    # --------------------------------------------------------- parameters
    def layer_shapes(self):
        """[(cout_p, cin_p)] per conv layer, in the flat parameter order."""
        return [(self.width, self.cin_p), (self.width, self.width), (self.width, self.width),
                (self.cout_p, self.width)]

    def get_params(self):
        h = np.empty(self.nparams, np.float32)
        check_cnn(lib().sc_cnn_get_params(self.net, h.ctypes.data_as(ctypes.c_void_p)), self.net)
        return h

    def set_params(self, h):
        h = np.ascontiguousarray(h, np.float32)
        check_cnn(lib().sc_cnn_set_params(self.net, h.ctypes.data_as(ctypes.c_void_p)), self.net)

    def get_grads(self):
        h = np.empty(self.nparams, np.float32)
        check_cnn(lib().sc_cnn_get_grads(self.net, h.ctypes.data_as(ctypes.c_void_p)), self.net)
        return h

    def zero_grad(self):
        check_cnn(lib().sc_cnn_zero_grad(self.net), self.net)

    def unflatten(self, h):
        """flat parameters -> ([W_l (27, cout_p, cin_p)], [b_l])"""
        out, off = [], 0
        for co, ci in self.layer_shapes():
            out.append(h[off:off + 27 * co * ci].reshape(27, co, ci))
            off += 27 * co * ci
        biases = [h[off + i * self.width: off + (i + 1) * self.width] for i in range(3)]
        return out, biases

    # ------------------------------------------------------------ tensors
    def new_input(self, batch):
        return torch.zeros((batch, self.nvox, self.cin_p), dtype=torch.bfloat16, device=self.device)

    def new_output(self, batch):
        return torch.zeros((batch, self.nvox, self.cout_p), dtype=torch.float32, device=self.device)

    def pack_input(self, fields, scale=1.0, extra=None, x=None):
        """fields: torch (batch, ncomp, n, n, n) or (ncomp, n, n, n) physical
        (float32/float64) -> bf16 input; extra = constant extra channel."""
        if fields.dim() == 4:
            fields = fields.unsqueeze(0)
        fields = fields.contiguous()
        batch, ncomp = fields.shape[0], fields.shape[1]
        x = x if x is not None else self.new_input(batch)
        check_cnn(lib().sc_cnn_pack_input(self.net, _p(fields), 1 if fields.dtype == torch.float64 else 0, ncomp,
                                          self.nvox, ncomp * self.nvox, batch, scale,
                                          1 if extra is not None else 0, 0.0 if extra is None else float(extra),
                                          _p(x)), self.net)
        return x

    def pack_target(self, fields, scale=1.0, out=None):
        """fields: (batch, cout, n, n, n) or (cout, n, n, n) physical -> fp32
        target (batch, n^3, cout) times scale."""
        if fields.dim() == 4:
            fields = fields.unsqueeze(0)
        fields = fields.contiguous()
        batch = fields.shape[0]
        if fields.shape[1] != self.cout:
            raise ValueError(f"expected {self.cout} components")
        out = out if out is not None else torch.empty((batch, self.nvox, self.cout), dtype=torch.float32,
                                                      device=self.device)
        check_cnn(lib().sc_cnn_pack_target(self.net, _p(fields), 1 if fields.dtype == torch.float64 else 0,
                                           self.nvox, self.cout * self.nvox, batch, scale, _p(out)), self.net)
        return out

    def forward(self, x, y=None):
        batch = x.shape[0]
        y = y if y is not None else self.new_output(batch)
        check_cnn(lib().sc_cnn_forward(self.net, _p(x), _p(y), batch), self.net)
        return y

    LOSSES = {"mse": 0, "relative": 1}

    def loss(self, y, target, kind="mse", dy=None):
        """(loss, dy): "mse" the mean squared error, "relative" the reference's
        sum|y - t|^2 / (sum|t|^2 + eps) over the batch (create_loss,
        src/training.jl:3-9)"""
        if kind not in self.LOSSES:
            raise ValueError(f"loss kind must be one of {sorted(self.LOSSES)}")
        batch = y.shape[0]
        dy = dy if dy is not None else torch.zeros((batch, self.nvox, self.cout_p), dtype=torch.bfloat16,
                                                   device=self.device)
        loss = ctypes.c_double()
        check_cnn(lib().sc_cnn_loss(self.net, _p(y), _p(target.contiguous()), _p(dy), batch, self.LOSSES[kind],
                                    ctypes.byref(loss)), self.net)
        return loss.value, dy

    def mse(self, y, target, dy=None):
        return self.loss(y, target, "mse", dy)

    def backward(self, x, dy, want_dx=False):
        batch = x.shape[0]
        dx = torch.zeros((batch, self.nvox, self.cin_p), dtype=torch.float32, device=self.device) if want_dx else None
        check_cnn(lib().sc_cnn_backward(self.net, _p(x), _p(dy), None if dx is None else _p(dx), batch), self.net)
        return dx

    def profile(self, enable=True):
        """start (clearing) or stop per-launch HIP-event timing"""
        check_cnn(lib().sc_cnn_profile(self.net, 1 if enable else 0), self.net)

    def profile_report(self):
        """{name: (launches, total ms, executed flop)} since profile(True)"""
        buf = ctypes.create_string_buffer(1 << 14)
        check_cnn(lib().sc_cnn_profile_report(self.net, buf, len(buf)), self.net)
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms, flop = line.split()
            out[name] = (int(cnt), float(ms), float(flop))
        return out

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check_cnn(lib().sc_cnn_adam(self.net, lr, beta1, beta2, eps), self.net)

    def adamw(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, clip=0.0):
        """ClipNorm(clip) per parameter array (clip <= 0: none), then AdamW
        (src/training.jl:50)"""
        check_cnn(lib().sc_cnn_adamw(self.net, lr, beta1, beta2, eps, weight_decay, clip), self.net)

    def train_step(self, x, target, lr=1e-3, loss="mse", clip=0.0, weight_decay=0.0):
        """one optimiser step on the loss; returns the loss before the step"""
        self.zero_grad()
        y = self.forward(x)
        value, dy = self.loss(y, target, loss)
        self.backward(x, dy)
        self.adamw(lr, weight_decay=weight_decay, clip=clip)
        return value
