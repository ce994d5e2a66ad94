"""Device plumbing: the phy_ctx context on torch's current HIP stream, compiled
topologies, and nibble-block packing.

PyTorch only provides device memory and the stream here; all compute goes
through libphylo_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


def words(n_sites: int) -> int:
    """uint64 words of a nibble block of n_sites sites."""
    return (int(n_sites) + 15) // 16


def pack_nibbles(sets: np.ndarray) -> np.ndarray:
    """[..., S] 4-bit sets (uint8) -> [..., words(S)] uint64, site s in word
    s//16, bits 4*(s%16).. (the layout of include/phylo_hip.h)."""
    sets = np.asarray(sets, dtype=np.uint8)
    S = sets.shape[-1]
    W = words(S)
    pad = np.zeros(sets.shape[:-1] + (W * 16,), dtype=np.uint8)
    pad[..., :S] = sets & 15
    b = (pad[..., 0::2] | (pad[..., 1::2] << 4)).astype(np.uint8)
    return np.ascontiguousarray(b).view(np.uint64).reshape(sets.shape[:-1] + (W,))


def unpack_nibbles(words_: np.ndarray, n_sites: int) -> np.ndarray:
    """Inverse of pack_nibbles."""
    w = np.ascontiguousarray(words_, dtype=np.uint64)
    b = w.view(np.uint8).reshape(w.shape[:-1] + (w.shape[-1] * 8,))
    out = np.empty(w.shape[:-1] + (w.shape[-1] * 16,), dtype=np.uint8)
    out[..., 0::2] = b & 15
    out[..., 1::2] = b >> 4
    return out[..., :n_sites]


def _torch():
    import torch  # device memory and streams only

    return torch


def dptr(t) -> int | None:
    """Raw device address of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return int(t.data_ptr())


class Context:
    """phy_ctx bound to a device and a HIP stream (default: torch's current
    stream on that device, so torch.cuda.Event timing sees the kernels)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device: the phylo_hip kernels need an MI355X (gfx950)")
        torch.cuda.set_device(device)
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = C.c_void_p()
        check(lib.phy_ctx_create(device, C.c_void_p(stream), C.byref(h)), None, "ctx_create")
        self.handle = h
        self.device = device
        self.stream = stream

    def sync(self) -> None:
        check(lib.phy_ctx_sync(self.handle), self.handle, "ctx_sync")

    def check(self, rc: int, what: str) -> None:
        check(rc, self.handle, what)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.phy_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Topology:
    """A rooted binary tree compiled into device programs (phy_topo).

    children: 2(N-1) ints, children of internal node N+k at [2k], [2k+1].
    ctx=None compiles on the host only (programs inspectable, no device)."""

    def __init__(self, n_leaves: int, children, root: int, ctx: Context | None = None):
        self.n_leaves = int(n_leaves)
        self.children = np.ascontiguousarray(children, dtype=np.int32)
        self.root = int(root)
        self.ctx = ctx
        h = C.c_void_p()
        rc = lib.phy_topo_create(ctx.handle if ctx else None, self.n_leaves,
                                 ptr(self.children, C.c_int32), self.root, C.byref(h))
        check(rc, ctx.handle if ctx else None, "topo_create")
        self.handle = h

    @property
    def n_nodes(self) -> int:
        return 2 * self.n_leaves - 1

    def info(self):
        dd, ud = C.c_int32(), C.c_int32()
        parent = np.zeros(self.n_nodes, dtype=np.int32)
        check(lib.phy_topo_info(self.handle, C.byref(dd), C.byref(ud), ptr(parent, C.c_int32)))
        return dd.value, ud.value, parent

    def programs(self):
        N = self.n_leaves
        dp = np.zeros(4 * (N - 1), np.int32)
        df = np.zeros(N, np.int32)
        up = np.zeros(4 * (N - 1), np.int32)
        uf = np.zeros(2 * N - 1, np.int32)
        check(lib.phy_topo_programs(self.handle, ptr(dp, C.c_int32), ptr(df, C.c_int32),
                                    ptr(up, C.c_int32), ptr(uf, C.c_int32)))
        return dp.reshape(-1, 4), df, up.reshape(-1, 4), uf

    def split(self):
        """(cut nodes, {("down"|"up", k): (entries [n, 4], operand ids)}) of the
        split Fitch programs; no cut nodes when the tree runs whole."""
        n = C.c_int32()
        cuts = np.zeros(32, np.int32)
        check(lib.phy_topo_split(self.handle, C.byref(n), ptr(cuts, C.c_int32)))
        cuts = cuts[: n.value]
        progs = {}
        for up, name in ((0, "down"), (1, "up")):
            for k in range(len(cuts) + 1 if len(cuts) else 0):
                no, nf = C.c_int32(), C.c_int32()
                check(lib.phy_topo_split_program(self.handle, up, k, None, C.byref(no), None,
                                                 C.byref(nf)))
                pg = np.zeros(4 * no.value, np.int32)
                ff = np.zeros(max(nf.value, 1), np.int32)
                check(lib.phy_topo_split_program(self.handle, up, k, ptr(pg, C.c_int32),
                                                 C.byref(no), ptr(ff, C.c_int32), C.byref(nf)))
                progs[(name, k)] = (pg.reshape(-1, 4), ff[: nf.value])
        return cuts, progs

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.phy_topo_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_device(a: np.ndarray, device: int = 0):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{device}")


def empty(shape, dtype, device: int = 0):
    torch = _torch()
    return torch.empty(shape, dtype=dtype, device=f"cuda:{device}")


def zeros(shape, dtype, device: int = 0):
    torch = _torch()
    return torch.zeros(shape, dtype=dtype, device=f"cuda:{device}")


__all__ = ["Context", "Topology", "words", "pack_nibbles", "unpack_nibbles", "dptr",
           "to_device", "empty", "zeros", "_lib"]
