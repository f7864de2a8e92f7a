"""Many independent reconstruction problems as one block-diagonal operator on the GPU.

  ProjectorBatch([op_0, ..., op_{n-1}])

Each op is a projector of elsa_amd.projectors with its own volume, trajectory and detector
(all of one dtype). The batch's volume is the problems' volumes flattened and packed back to
back (problem j at vol_offsets[j]), its sinogram likewise; pack() / unpack() convert between
the packed vector and per-problem arrays. apply / applyAdjoint give, for every problem, what
the problem's own operator gives, bit for bit; for FP32 Joseph operators, FP32 Siddon
operators or FP32 voxel projectors (one of these kinds per batch) with the default kernels
every pass is one kernel launch over all problems per kernel. The solvers of elsa_amd.solvers
accept a batch in place of an operator and then solve every problem independently (its own
step lengths and convergence), with per-problem outputs as rows.

No counterpart in elsa: BlockLinearOperator (elsa/operators/BlockLinearOperator.h:31-38)
stacks operators over a shared domain or range; a batch is the block-diagonal arrangement.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .projectors import LinearOperator


class _FlatDescriptor:
    def __init__(self, n):
        self.shape = (int(n),)

    def getNumberOfCoefficients(self):
        return self.shape[0]


class ProjectorBatch(LinearOperator):
    def __init__(self, ops):
        ops = list(ops)
        if not ops:
            raise ValueError("ProjectorBatch needs at least one operator")
        self.ops = ops  # the operators must outlive the batch
        self._L = _lib.load()
        self._dtype = ops[0]._dtype
        handles = (ctypes.c_void_p * len(ops))(*[op.handle.value for op in ops])
        self._b = ctypes.c_void_p()
        _lib.check(self._L.ehp_batch_create(handles, len(ops), ctypes.byref(self._b)), "ProjectorBatch")
        n = len(ops)
        self.vol_offsets = np.zeros(n + 1, np.int64)
        self.sino_offsets = np.zeros(n + 1, np.int64)
        fused = ctypes.c_int()
        i64p = ctypes.POINTER(ctypes.c_int64)
        _lib.check(self._L.ehp_batch_layout(self._b, self.vol_offsets.ctypes.data_as(i64p),
                                            self.sino_offsets.ctypes.data_as(i64p), ctypes.byref(fused)))
        self.fused = bool(fused.value)
        self._n_vol, self._n_sino = int(self.vol_offsets[-1]), int(self.sino_offsets[-1])
        super().__init__(_FlatDescriptor(self._n_vol), _FlatDescriptor(self._n_sino))

    def __len__(self):
        return len(self.ops)

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            self._L.ehp_batch_destroy(b)
            self._b = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._b

    def _tdt(self):
        import torch

        return torch.float32 if self._dtype == np.float32 else torch.float64

    def pack(self, arrays, adjoint_space=False):
        """Per-problem volumes (or sinograms with adjoint_space=True) -> one packed GPU tensor."""
        import torch

        off = self.sino_offsets if adjoint_space else self.vol_offsets
        if len(arrays) != len(self.ops):
            raise ValueError(f"expected {len(self.ops)} arrays, got {len(arrays)}")
        out = torch.empty(int(off[-1]), dtype=self._tdt(), device="cuda")
        for j, a in enumerate(arrays):
            t = torch.as_tensor(np.ascontiguousarray(a, dtype=self._dtype)) if isinstance(a, np.ndarray) else a
            if t.numel() != off[j + 1] - off[j]:
                raise ValueError(f"array {j} has {t.numel()} elements, expected {off[j + 1] - off[j]}")
            out[off[j]:off[j + 1]].copy_(t.reshape(-1))
        return out

    def unpack(self, packed, adjoint_space=False):
        """Packed tensor -> list of per-problem views shaped like each operator's volume (or
        sinogram with adjoint_space=True)."""
        off = self.sino_offsets if adjoint_space else self.vol_offsets
        res = []
        for j, op in enumerate(self.ops):
            d = op.getRangeDescriptor() if adjoint_space else op.getDomainDescriptor()
            res.append(packed[off[j]:off[j + 1]].reshape(d.shape))
        return res

    def _run(self, x, out, adjoint):
        import torch

        n_in, n_out = (self._n_sino, self._n_vol) if adjoint else (self._n_vol, self._n_sino)
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == self._tdt()):
            raise ValueError("batch input must be a GPU tensor of the operators' dtype")
        xt = x.contiguous()
        if xt.numel() != n_in:
            raise ValueError(f"input has {xt.numel()} elements, expected {n_in}")
        if out is None:
            out = torch.empty(n_out, dtype=xt.dtype, device=xt.device)
        elif not (out.is_cuda and out.dtype == xt.dtype and out.is_contiguous() and out.numel() == n_out):
            raise ValueError("out must be a contiguous GPU tensor of the right size and dtype")
        if out.data_ptr() == xt.data_ptr():
            raise ValueError("input and output must not alias")
        stream = ctypes.c_void_p(torch.cuda.current_stream(xt.device).cuda_stream)
        fn = self._L.ehp_batch_adjoint if adjoint else self._L.ehp_batch_forward
        _lib.check(fn(self._b, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(out.data_ptr()), stream),
                   "applyAdjoint" if adjoint else "apply")
        return out
