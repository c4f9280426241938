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

The MLP part runs as plain FP32 library GEMMs (torch.mm on rocBLAS/hipBLASLt);
the AEV, its tangent, the forces and the energies come from the HIP kernels.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from .potential import ANIPotential, SystemBuffers


def _celu_parts(z: torch.Tensor, alpha: float):
    e = torch.exp(torch.clamp(z, max=0.0) / alpha)
    pos = z > 0
    h = torch.where(pos, z, alpha * (e - 1.0))
    d1 = torch.where(pos, torch.ones_like(z), e)
    d2 = torch.where(pos, torch.zeros_like(z), e / alpha)
    return h, d1, d2


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
        self.weights = {}
        for s, layers in params.items():
            self.weights[s] = [(torch.tensor(np.asarray(W), dtype=torch.float32, device=dev),
                                torch.tensor(np.asarray(b), dtype=torch.float32, device=dev)) for W, b in layers]
        self.k_in = potential.config.aev_length

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
        real = torch.as_tensor(sb.species > 0, device=pot.device)
        E = sb.energy[: sb.n_sys]
        F = pot.forces(sb)
        n_e = float(sb.n_sys)
        n_f = 3.0 * float(real.sum().item())
        rE = E - e_ref.to(torch.float64)
        rF = (F - f_ref) * real[:, None]
        loss = self.w_e * float((rE ** 2).sum()) / n_e + self.w_f * float((rF.double() ** 2).sum()) / n_f
        dA = self.tangent(sb, x32, rF.contiguous())
        cE_atom = ((2.0 * self.w_e / n_e) * rE)[sb.batch_index[:N].long()].float()
        cD = -2.0 * self.w_f / n_f
        grads = {}
        species_z = torch.as_tensor(sb.species, device=pot.device)
        aev = sb.aev[:, : self.k_in].float()
        for n, s in enumerate(self.species):
            idx = torch.nonzero(species_z == _z(s)).flatten()
            grads[s] = self._species_grad(s, aev[idx], dA[idx, : self.k_in], cE_atom[idx], cD)
        return loss, grads, E, F

    def _species_grad(self, s: str, A: torch.Tensor, dA: torch.Tensor, cE: torch.Tensor, cD: float):
        (W1, b1), (W2, b2), (W3, b3), (W4, b4) = self.weights[s]
        a = self.alpha
        if A.shape[0] == 0:
            return [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in self.weights[s]]
        # forward primal and tangent chains
        z1 = A @ W1 + b1
        h1, s1, t1 = _celu_parts(z1, a)
        dz1 = dA @ W1
        dh1 = s1 * dz1
        z2 = h1 @ W2 + b2
        h2, s2, t2 = _celu_parts(z2, a)
        dz2 = dh1 @ W2
        dh2 = s2 * dz2
        z3 = h2 @ W3 + b3
        h3, s3, t3 = _celu_parts(z3, a)
        dz3 = dh2 @ W3
        dh3 = s3 * dz3
        w4 = W4[:, 0]
        # reverse through E = h3 w4 + b4 (adjoint cE) and D = dh3 w4 (adjoint cD)
        gW4 = (h3 * cE[:, None]).sum(0) + cD * dh3.sum(0)
        gb4 = cE.sum().reshape(1)
        bh3 = cE[:, None] * w4[None, :]
        bdh3 = torch.full_like(dh3, cD) * w4[None, :]
        bdz3 = s3 * bdh3
        bz3 = t3 * dz3 * bdh3 + s3 * bh3
        gW3 = dh2.T @ bdz3 + h2.T @ bz3
        gb3 = bz3.sum(0)
        bdh2 = bdz3 @ W3.T
        bh2 = bz3 @ W3.T
        bdz2 = s2 * bdh2
        bz2 = t2 * dz2 * bdh2 + s2 * bh2
        gW2 = dh1.T @ bdz2 + h1.T @ bz2
        gb2 = bz2.sum(0)
        bdh1 = bdz2 @ W2.T
        bh1 = bz2 @ W2.T
        bdz1 = s1 * bdh1
        bz1 = t1 * dz1 * bdh1 + s1 * bh1
        gW1 = dA.T @ bdz1 + A.T @ bz1
        gb1 = bz1.sum(0)
        return [(gW1, gb1), (gW2, gb2), (gW3, gb3), (gW4[:, None], gb4)]


def _z(symbol: str) -> int:
    from .units import SYMBOL_TO_Z

    return SYMBOL_TO_Z[symbol]
