"""Tensor-level wrappers of the C ABI. Inputs are torch tensors on a ROCm device;
complex values are torch.complex128, index/multiplicity arrays torch.int32.
Every function launches on the current stream and returns device tensors.
"""
import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _device as D
from . import _lib as L


def _c128(t, name):
    T = D.torch()
    if t.dtype != T.complex128 or not t.is_cuda:
        raise TypeError(f"{name} must be a complex128 GPU tensor, got {t.dtype} on {t.device}")
    return t.contiguous()


def _i32(t, name):
    T = D.torch()
    if t.dtype != T.int32 or not t.is_cuda:
        raise TypeError(f"{name} must be an int32 GPU tensor, got {t.dtype} on {t.device}")
    return t.contiguous()


def glynn_batch(A, precision: int = L.TB_PREC_FP64):
    """Permanents of a batch of square matrices A[b] (B x n x n) by Glynn/Gray."""
    T = D.require_gpu()
    if A.dim() != 3 or A.shape[1] != A.shape[2]:
        raise ValueError(f"glynn_batch expects B x n x n, got {tuple(A.shape)}")
    A = _c128(A, "A")
    B, n = A.shape[0], A.shape[1]
    out = T.empty(B, dtype=T.complex128, device=A.device)
    lib = L.load()
    need = lib.tb_glynn_workspace_size(n, B, precision)
    ws = D.workspace(need, A.device) if need else None
    L.check(lib.tb_glynn_batch(D.ptr(A), n, B, n * n, precision, D.ptr(out),
                               D.ptr(ws) if ws is not None else None, need,
                               D.stream_ptr(A.device)))
    return out


def guan_perm_batch(U, out_modes, out_mult, in_modes, in_mult, formula: int = L.TB_FORMULA_RYSER):
    """Multiplicity-aware permanents (Ryser or Chin-Huh over Guan codes). Item b has
    output modes out_modes[b] with multiplicities out_mult[b] and input modes
    in_modes[b] with multiplicities in_mult[b] on the shared matrix U."""
    T = D.require_gpu()
    U = _c128(U, "U")
    out_modes, out_mult = _i32(out_modes, "out_modes"), _i32(out_mult, "out_mult")
    in_modes, in_mult = _i32(in_modes, "in_modes"), _i32(in_mult, "in_mult")
    B = out_modes.shape[0]
    out = T.empty(B, dtype=T.complex128, device=U.device)
    L.check(L.load().tb_guan_perm_batch(
        D.ptr(U), U.shape[1], D.ptr(out_modes), D.ptr(out_mult), out_modes.shape[1],
        D.ptr(in_modes), D.ptr(in_mult), in_modes.shape[1], B, formula, D.ptr(out),
        D.stream_ptr(U.device)))
    return out


def guan_perm_outer(U, out_modes, out_mult, in_modes, in_mult, formula: int = L.TB_FORMULA_RYSER):
    """All pairs of output rows j and input rows i in one launch: an (n_out, n_in) tensor
    holding perm(U[t_j, s_i]); each pair enumerates the row with fewer Guan codes."""
    T = D.require_gpu()
    U = _c128(U, "U")
    out_modes, out_mult = _i32(out_modes, "out_modes"), _i32(out_mult, "out_mult")
    in_modes, in_mult = _i32(in_modes, "in_modes"), _i32(in_mult, "in_mult")
    n_out, n_in = out_modes.shape[0], in_modes.shape[0]
    out = T.empty((n_out, n_in), dtype=T.complex128, device=U.device)
    L.check(L.load().tb_guan_perm_outer(
        D.ptr(U), U.shape[1], D.ptr(out_modes), D.ptr(out_mult), out_modes.shape[1], n_out,
        D.ptr(in_modes), D.ptr(in_mult), in_modes.shape[1], n_in, formula, D.ptr(out),
        D.stream_ptr(U.device)))
    return out


def cc_submatrices_batch(U, out_modes, out_mult, in_modes, in_mult, k: int,
                         collision_free: bool):
    """All input-deleted minors per item: result[b, d] = perm(U[t_b, s_b - e_d])."""
    T = D.require_gpu()
    U = _c128(U, "U")
    out_modes, out_mult = _i32(out_modes, "out_modes"), _i32(out_mult, "out_mult")
    in_modes, in_mult = _i32(in_modes, "in_modes"), _i32(in_mult, "in_mult")
    B, in_max = in_modes.shape
    out_max = out_modes.shape[1]
    minors = T.empty((B, in_max), dtype=T.complex128, device=U.device)
    lib = L.load()
    need = lib.tb_cc_workspace_size(k, out_max, in_max, B) if collision_free else 0
    ws = D.workspace(need, U.device) if need else None
    L.check(lib.tb_cc_submatrices_batch(
        D.ptr(U), U.shape[1], D.ptr(out_modes), D.ptr(out_mult), out_max, D.ptr(in_modes),
        D.ptr(in_mult), in_max, B, int(k), 1 if collision_free else 0, D.ptr(minors),
        D.ptr(ws) if ws is not None else None, need, D.stream_ptr(U.device)))
    return minors


def sweep_minors_batch(Vt, k: int):
    """Binary C&C sweep on gathered columns Vt (B x k x ldv, ldv >= k-1)."""
    T = D.require_gpu()
    Vt = _c128(Vt, "Vt")
    B, kk, ldv = Vt.shape
    if kk < k:
        raise ValueError("Vt has fewer than k particle rows")
    minors = T.empty((B, k), dtype=T.complex128, device=Vt.device)
    lib = L.load()
    need = lib.tb_sweep_workspace_size(k, B)
    ws = D.workspace(need, Vt.device)
    L.check(lib.tb_sweep_minors_batch(D.ptr(Vt), k, ldv, kk * ldv, B, D.ptr(minors), D.ptr(ws),
                                      need, D.stream_ptr(Vt.device)))
    return minors


def perm_grad_batch(A) -> Tuple["object", "object"]:
    """perm(A[b]) and d perm / d A[b][i][j] = perm(A[b] without row i, column j)."""
    T = D.require_gpu()
    if A.dim() != 3 or A.shape[1] != A.shape[2]:
        raise ValueError(f"perm_grad_batch expects B x n x n, got {tuple(A.shape)}")
    A = _c128(A, "A")
    B, n = A.shape[0], A.shape[1]
    perm = T.empty(B, dtype=T.complex128, device=A.device)
    grad = T.empty((B, n, n), dtype=T.complex128, device=A.device)
    lib = L.load()
    need = lib.tb_perm_grad_workspace_size(n, B)
    ws = D.workspace(need, A.device) if need else None
    L.check(lib.tb_perm_grad_batch(D.ptr(A), n, B, n * n, D.ptr(perm), D.ptr(grad),
                                   D.ptr(ws) if ws is not None else None, need,
                                   D.stream_ptr(A.device)))
    return perm, grad


def gccb_config(m: int, n: int, samples: int, chunk_bits: Optional[Sequence[int]] = None):
    cfg = L.GccbConfig()
    cfg.m, cfg.n, cfg.samples = int(m), int(n), int(samples)
    for k in range(L.TB_SWEEP_MAX_K + 1):
        cfg.chunk_bits[k] = 0
    if chunk_bits is not None:
        for k, c in enumerate(chunk_bits):
            if k <= L.TB_SWEEP_MAX_K:
                cfg.chunk_bits[k] = int(c)
    return cfg


def gccb_sample(U, particles, samples: int, seed: int, transmissivity: float = 1.0,
                forced_order=None, forced_uniforms=None, forced_steps=None,
                chunk_bits: Optional[Sequence[int]] = None, trace: bool = False, out=None):
    """Batched GCC-B sampling on the device. Returns (samples x m int32, trace or None).

    U is m x m (all samples) or samples x m x m (one interferometer per sample); in the
    second form particles is samples x n, row s holding sample s's forced_steps[s]
    particles first (the rest is padding with valid mode indices), and forced_steps is
    required."""
    T = D.require_gpu()
    U = _c128(U, "U")
    particles = _i32(particles, "particles")
    per_sample = U.dim() == 3
    if per_sample:
        if U.shape[0] != samples or U.shape[1] != U.shape[2]:
            raise ValueError(f"per-sample U must be {samples} x m x m, got {tuple(U.shape)}")
        if particles.dim() != 2 or particles.shape[0] != samples:
            raise ValueError("per-sample particles must be samples x n")
        if forced_steps is None:
            raise ValueError("per-sample particle lists need forced_steps")
        m, n = U.shape[1], particles.shape[1]
    else:
        if U.dim() != 2 or U.shape[0] != U.shape[1] or particles.dim() != 1:
            raise ValueError("U must be m x m and particles a vector of mode indices")
        m, n = U.shape[0], particles.shape[0]
    cfg = gccb_config(m, n, samples, chunk_bits)
    cfg.per_sample = 1 if per_sample else 0
    cfg.u_stride = m * m if per_sample else 0
    lib = L.load()
    need = lib.tb_gccb_workspace_size(ctypes.byref(cfg))
    ws = D.workspace(need, U.device)
    if out is None:
        out = T.empty((samples, m), dtype=T.int32, device=U.device)
    elif (out.dtype != T.int32 or not out.is_cuda or tuple(out.shape) != (samples, m)
          or not out.is_contiguous()):
        raise ValueError(f"out must be a contiguous int32 GPU tensor of shape ({samples}, {m})")
    tr = T.zeros((samples, n, n), dtype=T.complex128, device=U.device) if trace else None
    if forced_order is not None:
        fo_host = forced_order.detach().cpu().numpy()
        if fo_host.shape != (samples, n) or not all(
                np.array_equal(np.sort(row), np.arange(n)) for row in fo_host):
            raise ValueError("forced_order must hold one permutation of range(n) per sample")
        if per_sample:
            fs_chk = forced_steps.detach().cpu().numpy()
            if not all(np.array_equal(np.sort(row[:l]), np.arange(l))
                       for row, l in zip(fo_host, fs_chk)):
                raise ValueError("with per-sample particles, forced_order row s must start "
                                 "with a permutation of range(forced_steps[s])")
    if forced_uniforms is not None:
        fu_host = forced_uniforms.detach().cpu().numpy()
        if fu_host.shape != (samples, n) or np.any(fu_host < 0) or np.any(fu_host >= 1):
            raise ValueError("forced_uniforms must be samples x n values in [0, 1)")
    if forced_steps is not None:
        fs_host = forced_steps.detach().cpu().numpy()
        if fs_host.shape != (samples,) or np.any(fs_host < 0) or np.any(fs_host > n):
            raise ValueError("forced_steps must be samples values in [0, n]")
    p_host = particles.detach().cpu().numpy()
    if np.any(p_host < 0) or np.any(p_host >= m):
        raise ValueError("particles must be mode indices in [0, m)")
    fo = _i32(forced_order, "forced_order") if forced_order is not None else None
    fu = forced_uniforms.contiguous() if forced_uniforms is not None else None
    if fu is not None and (fu.dtype != T.float64 or not fu.is_cuda):
        raise TypeError("forced_uniforms must be a float64 GPU tensor")
    fs = _i32(forced_steps, "forced_steps") if forced_steps is not None else None
    L.check(lib.tb_gccb_sample(
        ctypes.byref(cfg), D.ptr(U), U.shape[-1], D.ptr(particles), ctypes.c_uint64(seed & (2**64 - 1)),
        float(transmissivity),
        D.ptr(fo) if fo is not None else None, D.ptr(fu) if fu is not None else None,
        D.ptr(fs) if fs is not None else None, D.ptr(out), D.ptr(tr) if tr is not None else None,
        D.ptr(ws), need, D.stream_ptr(U.device)))
    return out, tr


def outcome_count(particles_number: int, modes_number: int, losses: bool = False) -> int:
    """Length of generate_possible_states(particles_number, modes_number, losses)."""
    c = L.load().tb_outcome_count(int(particles_number), int(modes_number), 1 if losses else 0)
    L.check(min(c, 0))
    return int(c)


def outcome_histogram(samples, particles_number: int, losses: bool = False):
    """Counts of the samples (S x m int32 GPU tensor) over the outcome list
    generate_possible_states(particles_number, m, losses), in that order, on the device.
    Returns (counts: int64 tensor, outside: number of samples not in the list)."""
    T = D.require_gpu()
    samples = _i32(samples, "samples")
    if samples.dim() != 2:
        raise ValueError(f"samples must be S x m, got {tuple(samples.shape)}")
    S, m = samples.shape
    count = outcome_count(particles_number, m, losses)
    counts = T.zeros(count, dtype=T.int32, device=samples.device)  # uint32 counters
    outside = T.zeros(1, dtype=T.int32, device=samples.device)
    L.check(L.load().tb_outcome_histogram(D.ptr(samples), m, S, m, int(particles_number),
                                          1 if losses else 0, D.ptr(counts), D.ptr(outside),
                                          D.stream_ptr(samples.device)))
    return counts.to(T.int64), int(outside.item())
