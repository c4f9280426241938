"""Training step: energy + force loss and its gradient with respect to the MLP weights.

Restates the loss of FeNNol's train_step (training/utils.py:140-525) for the
ANI path with ``mse`` terms on ``total_energy`` and ``forces``:

    L = w_E * sum_b (E_b - E_ref,b)^2 / n_E + w_F * sum_ia (F_ia - F_ref,ia)^2 / n_F

with n_E the number of systems and n_F = 3 x the number of real atoms
(training/utils.py:329-343, weights :496-501).  The reference differentiates
``jax.grad`` of the energy with respect to the coordinates once more with
respect to the weights (:508).  Here the gradient is assembled by hand:

* energy part: dE_b/dW by the reverse pass of the species MLP;
* force part:  sum_i r_i . dF_i/dW = - d/dW sum_j g_j(W) . dA_j, where
  r = F - F_ref, g_j = dE_j/dA_j is the MLP input gradient and
  dA = dAEV/dx . r is the AEV tangent along r (``fnl_aev_tangent``, the AEV
  terms recomputed on the GPU).  dA does not depend on W, so only the MLP is
  differentiated twice: forward primal and tangent chains, then the reverse of
  both (CELU'' enters through the tangent chain).

The MLP part is ``fnl_train_mlp_grad``: FP32 MFMA kernels that run both chains
on shared weight operands, fuse the CELU parts and the start of the reverse
pass into their epilogues, and reduce the weight gradients over atoms in a
fixed order (the same bits every run); the AEV, its tangent, the forces and the
energies come from the other HIP kernels.  No library GEMM runs.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from .potential import ANIPotential, SystemBuffers


class TrainingStep:
    """Loss and dL/dW for the species networks of an FP32 :class:`ANIPotential`."""

    def __init__(self, potential: ANIPotential, params: dict, w_energy: float = 1.0, w_forces: float = 10.0,
                 celu_alpha: float = 0.1):
        if potential.precision != "fp32":
            raise ValueError("the training step uses the FP32 kernels")
        self.pot = potential
        self.w_e = float(w_energy)
        self.w_f = float(w_forces)
        self.alpha = celu_alpha
        dev = potential.device
        self.species = list(params.keys())
        if self.species != list(potential.nets.species):
            raise ValueError("params must list the potential's networks in its order")
        self.weights = {}
        self.cnets = (_abi.TrainNet * len(self.species))()
        for n, (s, layers) in enumerate(params.items()):
            self.weights[s] = [(torch.tensor(np.asarray(W), dtype=torch.float32, device=dev).contiguous(),
                                torch.tensor(np.asarray(b), dtype=torch.float32, device=dev).contiguous())
                               for W, b in layers]
            (W1, b1), (W2, b2), (W3, b3), (W4, b4) = self.weights[s]
            c = self.cnets[n]
            c.w1, c.b1, c.w2, c.b2 = W1.data_ptr(), b1.data_ptr(), W2.data_ptr(), b2.data_ptr()
            c.w3, c.b3, c.w4, c.b4 = W3.data_ptr(), b3.data_ptr(), W4.data_ptr(), b4.data_ptr()
        self.k_in = potential.config.aev_length
        self.gstride = int(potential.lib.fnl_train_grad_stride(self.k_in))
        self._ws = None

    def tangent(self, sb: SystemBuffers, x32: torch.Tensor, disp: torch.Tensor) -> torch.Tensor:
        """AEV tangent rows [N, k_pad] along disp [N, 3] (float32)."""
        out = torch.empty(sb.g_aev.shape, dtype=torch.float32, device=sb.g_aev.device)
        P = _abi.ptr
        _abi.check(self.pot.lib.fnl_aev_tangent(
            P(x32), P(disp), P(sb.species_idx), P(sb.batch_index), sb.n_atoms, sb.pbc, P(sb.cells), P(sb.rcells),
            P(sb.nbr), P(sb.n_nbr), sb.max_nbr, P(sb.nbr_shift), P(sb.nbr_ang), P(sb.n_nbr_ang), sb.max_nbr_ang,
            ctypes.byref(self.pot.cparams), P(out), out.shape[1], _abi.stream_handle()),
            "fnl_aev_tangent")
        return out

    def __call__(self, sb: SystemBuffers, x32: torch.Tensor, e_ref: torch.Tensor, f_ref: torch.Tensor):
        """Returns (loss, {species: [(dW, db) per layer]}, energies [B], forces [N,3])."""
        pot = self.pot
        pot.compute(sb, x32, forces=True, rebuild=True, check=True)
        N = sb.n_atoms
        dev = pot.device
        E = sb.energy[: sb.n_sys]
        n_e = float(sb.n_sys)
        n_f = 3.0 * float(np.count_nonzero(np.asarray(sb.species[:N]) > 0))
        F = torch.empty((N, 3), dtype=torch.float32, device=dev)
        rF = torch.empty((N, 3), dtype=torch.float32, device=dev)
        cE_atom = torch.empty(N, dtype=torch.float32, device=dev)
        loss = torch.empty(1, dtype=torch.float64, device=dev)
        part = torch.empty(max(1, (N + 255) // 256), dtype=torch.float64, device=dev)
        P = _abi.ptr
        e_ref64 = e_ref.to(torch.float64).contiguous()
        f_ref32 = f_ref.to(torch.float32).contiguous()
        # forces, residuals, energy adjoints and the loss in two launches; the loss stays on the
        # device until every launch of the step is queued (one wait)
        _abi.check(pot.lib.fnl_train_residuals(
            P(E), P(e_ref64), sb.n_sys, P(sb.grad_self), P(sb.grad_nbr), 1.0, P(f_ref32), P(sb.species_idx),
            P(sb.batch_index), N, self.w_e, self.w_f, n_e, n_f, P(F), P(rF), P(cE_atom), P(loss), P(part),
            _abi.stream_handle()), "fnl_train_residuals")
        dA = self.tangent(sb, x32, rF)
        cD = -2.0 * self.w_f / n_f
        grads = self.mlp_grad(sb, dA, cE_atom, cD)
        return float(loss), grads, E, F

    def mlp_grad(self, sb: SystemBuffers, dA: torch.Tensor, c_energy: torch.Tensor, c_force: float) -> dict:
        """dL/dW of every network from the AEV rows of ``sb``, the tangent rows dA [N, k_pad],
        per-atom energy adjoints (atom order) and the force adjoint (fnl_train_mlp_grad)."""
        pot = self.pot
        # atoms grouped by network in network order: the potential's tile permutation
        counts = [int(np.sum(sb.species[: int(np.sum(sb.natoms))] == _z(s))) for s in self.species]
        net_rows = (ctypes.c_int32 * (len(counts) + 1))(*np.concatenate([[0], np.cumsum(counts)]).astype(int))
        n_rows = int(net_rows[len(counts)])
        need = int(pot.lib.fnl_train_workspace(n_rows, len(counts), self.k_in))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=pot.device)
        grad = torch.empty(len(counts) * self.gstride, dtype=torch.float32, device=pot.device)
        k_live = pot.config.live_column_blocks(sb.extra.get("aev_mask", 0xFFFFFFFF))
        P = _abi.ptr
        _abi.check(pot.lib.fnl_train_mlp_grad(
            P(sb.aev), sb.aev.shape[1], P(dA), dA.shape[1], self.k_in, self.cnets, len(counts), net_rows,
            P(sb.perm), P(c_energy), float(c_force), float(self.alpha), k_live, P(self._ws), self._ws.numel(),
            P(grad), _abi.stream_handle()), "fnl_train_mlp_grad")
        K, H1, H2, H3 = self.k_in, 256, 192, 160
        shapes = [(K, H1), (H1,), (H1, H2), (H2,), (H2, H3), (H3,), (H3, 1), (1,)]
        out = {}
        for n, s in enumerate(self.species):
            g = grad[n * self.gstride: (n + 1) * self.gstride]
            parts, o = [], 0
            for shp in shapes:
                size = int(np.prod(shp))
                parts.append(g[o: o + size].view(*shp))
                o += size
            out[s] = [(parts[0], parts[1]), (parts[2], parts[3]), (parts[4], parts[5]), (parts[6], parts[7])]
        return out


def _z(symbol: str) -> int:
    from .units import SYMBOL_TO_Z

    return SYMBOL_TO_Z[symbol]
