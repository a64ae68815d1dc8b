"""Interferometer fitting: a Clements mesh trained by Adam on the negative
log-likelihood of boson-sampling samples, with permanent gradients from the GPU backward
kernel (no reference counterpart: theboss samples but does not fit interferometers).

Mesh: m(m-1)/2 Mach-Zehnder units in the rectangular (Clements et al., Optica 3, 1460
(2016)) layout; unit (theta, phi) on modes (k, k+1) acts as
    [[e^{i phi} cos theta, -sin theta], [e^{i phi} sin theta, cos theta]].
Output phases do not change |perm|, so they are not parametrised.
Likelihood of output occupation t for collision-free input modes S:
    p(t) = |perm(U[rows(t), S])|^2 / prod_j t_j!
"""
import math
from typing import List, Sequence, Tuple

import numpy as np

from . import _device as D
from .autograd import permanent


def clements_layout(m: int) -> List[Tuple[int, int]]:
    """(column, first mode) of every unit, m(m-1)/2 of them."""
    units = []
    for col in range(m):
        for k in range(col % 2, m - 1, 2):
            units.append((col, k))
    return units


def clements_unitary(theta, phi, m: int):
    """Differentiable m x m unitary from the mesh parameters (torch tensors of length
    m(m-1)/2, real): U = T_last ... T_1, unit u acting on rows (k, k+1). The units of one
    column act on the disjoint row pairs (p, p+1), (p+2, p+3), ... (p = column parity), so
    a column is one batched update of strided row views (about a dozen tensor operations
    per column instead of per unit, and no index gathers whose backward must sort)."""
    T = D.torch()
    U = T.eye(m, dtype=T.complex128, device=theta.device)
    u0 = 0
    for col in range(m):
        p = col % 2
        L = len(range(p, m - 1, 2))  # units in this column
        if L == 0:
            continue
        th, ph = theta[u0:u0 + L], phi[u0:u0 + L]
        u0 += L
        c = T.cos(th).to(T.complex128)[:, None]
        s = T.sin(th).to(T.complex128)[:, None]
        e = T.exp(1j * ph.to(T.complex128))[:, None]
        a, b = U[p:p + 2 * L:2], U[p + 1:p + 2 * L:2]
        mid = T.stack((e * c * a - s * b, e * s * a + c * b), dim=1).reshape(2 * L, m)
        U = T.cat((U[:p], mid, U[p + 2 * L:]), dim=0)
    return U


def sample_index_tensors(samples: np.ndarray, input_modes: Sequence[int]):
    """Row index matrix (S x n) of the effective matrices and log prod t_j! per sample."""
    samples = np.asarray(samples, dtype=np.int64)
    rows = np.stack([np.repeat(np.arange(samples.shape[1]), s) for s in samples])
    logfact = np.array([sum(math.lgamma(c + 1) for c in s) for s in samples])
    return rows, logfact


def negative_log_likelihood(U, rows, cols, logfact):
    """Mean NLL of the samples under U. rows: S x n long tensor, cols: n long tensor."""
    T = D.torch()
    A = U[rows][:, :, cols]  # S x n x n effective matrices (differentiable gather)
    p = permanent(A)
    logp = T.log(p.real ** 2 + p.imag ** 2) - logfact
    return -logp.mean()


def fit_clements(samples: np.ndarray, input_modes: Sequence[int], m: int, steps: int = 200,
                 lr: float = 0.05, seed: int = 0, init=None, lr_end: float = None,
                 restarts: int = 1):
    """Adam on the NLL; returns (theta, phi, history of NLL values). With lr_end the learning
    rate falls from lr to lr_end along a half cosine over the steps (large early steps cross
    the rugged start of the landscape, small late ones settle into the basin). The
    landscape has poor local minima that some starting points cannot leave: restarts > 1
    runs the fit from the random starts seed, seed + 1, ... and returns the run with the
    lowest final NLL (with init, every restart starts there and restarts is pointless)."""
    T = D.require_gpu()
    dev = D.device()
    if restarts < 1:
        raise ValueError(f"restarts={restarts} < 1")
    rows, logfact = sample_index_tensors(samples, input_modes)
    rows_t = T.from_numpy(rows).to(dev)
    cols_t = T.tensor(list(input_modes), dtype=T.long, device=dev)
    lf_t = T.from_numpy(logfact).to(dev)
    nunits = m * (m - 1) // 2
    best = None
    for r in range(restarts):
        g = T.Generator(device="cpu").manual_seed(seed + r)
        if init is None:
            theta = (T.rand(nunits, generator=g, dtype=T.float64) * math.pi / 2).to(dev)
            phi = (T.rand(nunits, generator=g, dtype=T.float64) * 2 * math.pi).to(dev)
        else:
            theta, phi = (x.detach().clone().to(dev) for x in init)
        theta.requires_grad_(True)
        phi.requires_grad_(True)
        opt = T.optim.Adam([theta, phi], lr=lr)
        hist = []
        for step in range(steps):
            if lr_end is not None:
                frac = step / max(steps - 1, 1)
                for grp in opt.param_groups:
                    grp["lr"] = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * frac))
            opt.zero_grad()
            U = clements_unitary(theta, phi, m)
            loss = negative_log_likelihood(U, rows_t, cols_t, lf_t)
            loss.backward()
            opt.step()
            hist.append(float(loss.detach().cpu()))
        if best is None or hist[-1] < best[2][-1]:
            best = (theta.detach(), phi.detach(), hist)
    return best
