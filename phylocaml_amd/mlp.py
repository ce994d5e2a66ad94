"""The SPR-scoring MLP (phy_mlp_* in include/phylo_hip.h): 32 -> 128 -> 128 -> 1,
GELU, BF16-input MFMA GEMMs with FP32 master weights and Adam, all in
libphylo_hip.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as D
from ._lib import check, lib


class MLP:
    def __init__(self, ctx: D.Context, n_in: int = 32, n_hidden: int = 128, seed: int = 0):
        torch = __import__("torch")
        self.ctx, self.n_in, self.H = ctx, int(n_in), int(n_hidden)
        h = C.c_void_p()
        ctx.check(lib.phy_mlp_create(ctx.handle, self.n_in, self.H, seed, C.byref(h)), "mlp_create")
        self.handle = h
        self.n_params = int(lib.phy_mlp_n_params(h))
        dev = f"cuda:{ctx.device}"
        self._loss = torch.zeros((1,), dtype=torch.float64, device=dev)
        self._grads = torch.zeros((self.n_params,), dtype=torch.float32, device=dev)

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        check(lib.phy_mlp_get_params(self.ctx.handle, self.handle,
                                     p.ctypes.data_as(C.POINTER(C.c_float))), self.ctx.handle)
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(lib.phy_mlp_set_params(self.ctx.handle, self.handle,
                                     p.ctypes.data_as(C.POINTER(C.c_float))), self.ctx.handle)

    def forward(self, x):
        """x: device float32 [B, n_in] -> device float32 [B]."""
        torch = __import__("torch")
        x = x.contiguous()
        y = torch.empty((x.shape[0],), dtype=torch.float32, device=x.device)
        self.ctx.check(lib.phy_mlp_forward(self.ctx.handle, self.handle, D.dptr(x), x.shape[0],
                                           D.dptr(y)), "mlp_forward")
        return y

    def gradients(self, x, t):
        """(host gradient vector, loss) without changing the parameters."""
        x, t = x.contiguous(), t.contiguous()
        self.ctx.check(lib.phy_mlp_backward(self.ctx.handle, self.handle, D.dptr(x), D.dptr(t),
                                            x.shape[0], D.dptr(self._grads), D.dptr(self._loss)),
                       "mlp_backward")
        return self._grads.cpu().numpy().astype(np.float64), float(self._loss.cpu()[0])

    def step(self, x, t, lr: float = 1e-3, want_loss: bool = True):
        x, t = x.contiguous(), t.contiguous()
        self.ctx.check(lib.phy_mlp_step(self.ctx.handle, self.handle, D.dptr(x), D.dptr(t),
                                        x.shape[0], float(lr),
                                        D.dptr(self._loss) if want_loss else None), "mlp_step")
        return float(self._loss.cpu()[0]) if want_loss else None

    def close(self):
        if getattr(self, "handle", None):
            lib.phy_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
