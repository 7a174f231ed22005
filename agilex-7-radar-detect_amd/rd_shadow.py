"""FP32 differentiable shadow of PFB -> FFT -> beamformer -> power -> soft CA-CFAR.

The forward and backward passes of the linear stages and the power run in the
HIP kernels of include/rd_shadow.h (``rd_sh_*``), on torch's current stream;
the soft CFAR is a few torch operations whose gradient torch derives.
:class:`FrontEnd` is the ``torch.autograd.Function``:

    p = FrontEnd.apply(x, W, gain, h)

* ``x``    complex64 [Q, C, F*N]  channel samples (int16 I/Q / 2^15), Q CPIs
* ``W``    complex64 [B, C]       beam weights (Q1.15 / 2^15), no conjugate
                                  (model/cpp/beamformer/beamformer_model.hpp:162-179)
* ``gain`` complex64 [C]          per-channel complex gain/phase correction
* ``h``    complex64 [N, T]       PFB taps, phase-major (pfb_model.hpp:108-111)
* ``p``    float32  [Q, B, F*N]   |y|^2, y = sum_c W[b][c] gain[c] X[c] with
                                  X = DFT(PFB(x)) / N per frame

The channel gain multiplies each channel's samples; PFB and FFT are linear per
channel, so it is applied as effective weights W'[b][c] = W[b][c] gain[c]
(exactly the same map). The FFT output X is the one activation kept for the
backward pass ("checkpointed"); y is recomputed from X and W' by the backward
kernel. Gradients for x flow back through the FFT and PFB adjoint kernels.

The soft CA-CFAR (:func:`soft_cfar_logits`) is
``sigmoid((p - alpha*mean) / tau)`` with p and the reference-cell mean both in
units of that mean, i.e. ``z = (p/mean - alpha) / tau``; the hard decision of
the bit-exact CFAR (cfar_model.hpp:296-301, C*N*2^8 > A*S) is ``z > 0``.

:func:`reference` is the same model in float64 written with plain torch ops
(torch.fft, matmul) for the gradient checks; FP64 is used on the GPU only
there.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

import rd_amd

SH_EXPORTS = ["rd_sh_pfb", "rd_sh_pfb_bwd", "rd_sh_fft", "rd_sh_bf_power", "rd_sh_bf_bwd_work_bytes",
              "rd_sh_bf_power_bwd"]

_bound = None


def _lib():
    global _bound
    if _bound is None:
        L = rd_amd.load().lib
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.rd_sh_pfb.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp]
        L.rd_sh_pfb_bwd.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp]
        L.rd_sh_fft.argtypes = [vp, vp, vp, u32, u32, i32, C.c_float, vp]
        L.rd_sh_bf_power.argtypes = [vp, vp, vp, u32, u32, u64, u32, vp, vp]
        L.rd_sh_bf_bwd_work_bytes.argtypes = [u32, u32]
        L.rd_sh_bf_bwd_work_bytes.restype = C.c_size_t
        L.rd_sh_bf_power_bwd.argtypes = [vp, vp, vp, vp, u32, u32, u64, u32, vp, vp, vp, C.c_size_t]
        for n in SH_EXPORTS:
            if n != "rd_sh_bf_bwd_work_bytes":
                getattr(L, n).restype = i32
        _bound = L
    return _bound


def _check(name, rc):
    if rc != 0:
        raise rd_amd.RdError(name, rc, rd_amd.load().lib.rd_status_str(rc).decode())


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _p(t):
    assert t.is_cuda and t.is_contiguous(), "device tensors, contiguous"
    return t.data_ptr()


_TW = {}


def twiddles(n: int, device) -> torch.Tensor:
    """exp(-2 pi i j / n), j < n/2, evaluated in float64 then rounded to complex64."""
    key = (n, str(device))
    if key not in _TW:
        j = np.arange(n // 2)
        _TW[key] = torch.from_numpy(np.exp(-2j * np.pi * j / n).astype(np.complex64)).to(device)
    return _TW[key]


# ----------------------------------------------------------------------------- kernel wrappers
def pfb(x: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    """x complex64 [R, F*N], h [N, T] -> u [R, F, N]."""
    R, L = x.shape
    N, T = h.shape
    F = L // N
    u = torch.empty((R, F, N), dtype=torch.complex64, device=x.device)
    _check("rd_sh_pfb", _lib().rd_sh_pfb(_stream(), _p(x), _p(h), R, N, T, F, _p(u)))
    return u


def pfb_adjoint(g_u: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    """grad u [R, F, N] -> grad x [R, F*N]."""
    R, F, N = g_u.shape
    T = h.shape[1]
    g_x = torch.empty((R, F * N), dtype=torch.complex64, device=g_u.device)
    _check("rd_sh_pfb_bwd", _lib().rd_sh_pfb_bwd(_stream(), _p(g_u), _p(h), R, N, T, F, _p(g_x)))
    return g_x


def fft(x: torch.Tensor, inverse=False, scale=1.0) -> torch.Tensor:
    """Transforms along the last axis (natural order)."""
    n = x.shape[-1]
    out = torch.empty_like(x)
    _check("rd_sh_fft", _lib().rd_sh_fft(_stream(), _p(x), _p(twiddles(n, x.device)), n, x.numel() // n,
                                          1 if inverse else 0, float(scale), _p(out)))
    return out


def bf_power(X: torch.Tensor, W: torch.Tensor, want_y=False):
    """X [Q, C, M], W [B, C] -> p [Q, B, M] (and y)."""
    Q, Cn, M = X.shape
    B = W.shape[0]
    p = torch.empty((Q, B, M), dtype=torch.float32, device=X.device)
    y = torch.empty((Q, B, M), dtype=torch.complex64, device=X.device) if want_y else None
    _check("rd_sh_bf_power", _lib().rd_sh_bf_power(_stream(), _p(X), _p(W), Cn, B, M, Q, _p(y) if want_y else None,
                                                    _p(p)))
    return (p, y) if want_y else p


def bf_power_bwd(X: torch.Tensor, W: torch.Tensor, g_p: torch.Tensor, want_gX=True):
    Q, Cn, M = X.shape
    B = W.shape[0]
    L = _lib()
    wb = int(L.rd_sh_bf_bwd_work_bytes(Cn, B))
    work = torch.empty(wb // 4, dtype=torch.float32, device=X.device)
    g_W = torch.empty((B, Cn), dtype=torch.complex64, device=X.device)
    g_X = torch.empty_like(X) if want_gX else None
    _check("rd_sh_bf_power_bwd", L.rd_sh_bf_power_bwd(_stream(), _p(X), _p(W), _p(g_p), Cn, B, M, Q,
                                                       _p(g_X) if want_gX else None, _p(g_W), _p(work), wb))
    return g_W, g_X


# ----------------------------------------------------------------------------- autograd
class FrontEnd(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, W, gain, h):
        Q, Cn, L = x.shape
        N, T = h.shape
        F = L // N
        x = x.contiguous()
        u = pfb(x.reshape(Q * Cn, L), h.contiguous())
        X = fft(u.reshape(-1, N), scale=1.0 / N).reshape(Q, Cn, F * N)
        Wp = (W * gain[None, :]).contiguous()
        p = bf_power(X, Wp)
        ctx.save_for_backward(X, W, gain, h)
        ctx.dims = (Q, Cn, F, N)
        return p

    @staticmethod
    def backward(ctx, g_p):
        X, W, gain, h = ctx.saved_tensors
        Q, Cn, F, N = ctx.dims
        need_x = ctx.needs_input_grad[0]
        Wp = (W * gain[None, :]).contiguous()
        g_Wp, g_X = bf_power_bwd(X, Wp, g_p.contiguous(), want_gX=need_x)
        g_W = g_Wp * gain.conj()[None, :] if ctx.needs_input_grad[1] else None
        g_gain = (g_Wp * W.conj()).sum(0) if ctx.needs_input_grad[2] else None
        g_x = None
        if need_x:  # adjoint of DFT/N is the conjugate transform / N, then the PFB adjoint
            g_u = fft(g_X.reshape(-1, N), inverse=True, scale=1.0 / N)
            g_x = pfb_adjoint(g_u.reshape(Q * Cn, F, N), h.contiguous()).reshape(Q, Cn, F * N)
        return g_x, g_W, g_gain, None


# ----------------------------------------------------------------------------- soft CFAR and loss
def soft_cfar_logits(p: torch.Tensor, n_fft: int, guard: int, ref: int, alpha: float, tau: float) -> torch.Tensor:
    """p [Q, B, F*N] -> logits z [Q, B, F, N - 2(guard+ref)] for the bins whose
    whole window fits (the others are suppressed, cfar_model.hpp:268-278).
    Symmetric windows: lead cells i+guard+1..i+guard+ref, lag cells
    i-guard-ref..i-guard-1."""
    Q, B, L = p.shape
    P = p.reshape(Q, B, L // n_fft, n_fft)
    lo = guard + ref
    hi = n_fft - lo
    S = None
    for d in range(guard + 1, guard + ref + 1):
        t = P[..., lo + d: hi + d] + P[..., lo - d: hi - d]
        S = t if S is None else S + t
    mean = S / (2 * ref)
    return (P[..., lo:hi] / mean - alpha) / tau


def loss_fn(z: torch.Tensor, target: torch.Tensor, weight: torch.Tensor | None = None) -> torch.Tensor:
    """Binary cross-entropy of sigmoid(z) against the true target cells;
    `weight` (optional, same shape) zeroes cells that count neither way."""
    return torch.nn.functional.binary_cross_entropy_with_logits(z, target, weight=weight)


def reference(x, W, gain, h):
    """The front end in plain torch ops (any dtype; float64 for the checks)."""
    Q, Cn, L = x.shape
    N, T = h.shape
    F = L // N
    xr = x.reshape(Q, Cn, F, N)
    u = h[:, 0] * xr
    for k in range(1, min(T, F)):  # taps reaching before the first beat see the zero history
        shifted = torch.cat([torch.zeros_like(xr[:, :, :k]), xr[:, :, : F - k]], dim=2)
        u = u + h[:, k] * shifted
    X = torch.fft.fft(u, dim=-1) / N
    Wp = W * gain[None, :]
    y = torch.matmul(Wp, X.reshape(Q, Cn, F * N))
    return y.real ** 2 + y.imag ** 2


# ----------------------------------------------------------------------------- quantise back
def q15_pack(z: np.ndarray) -> np.ndarray:
    """complex -> packed Q1.15 words {im, re} (round half to even, saturate)."""
    re = np.clip(np.rint(np.real(z) * 32768.0), -32768, 32767).astype(np.int64) & 0xFFFF
    im = np.clip(np.rint(np.imag(z) * 32768.0), -32768, 32767).astype(np.int64) & 0xFFFF
    return ((im << 16) | re).astype(np.uint32)


def unpack_cx(words: np.ndarray) -> np.ndarray:
    w = np.asarray(words, np.uint32)
    re = (w & 0xFFFF).astype(np.int64)
    im = (w >> 16).astype(np.int64)
    re = np.where(re >= 0x8000, re - 0x10000, re)
    im = np.where(im >= 0x8000, im - 0x10000, im)
    return (re + 1j * im) / 32768.0


def quantise_weights(W: torch.Tensor, gain: torch.Tensor) -> np.ndarray:
    """The trained beam weights with the channel gains folded in
    (W'[b][c] = W[b][c] gain[c]), as packed Q1.15 beam-major words for the
    bit-exact chain. A common scale keeps every component inside Q1.15 (the
    CA-CFAR decision is invariant to it up to rounding)."""
    Wp = (W * gain[None, :]).detach().cpu().numpy().astype(np.complex128)
    m = max(np.abs(Wp.real).max(), np.abs(Wp.imag).max())
    if m > 32767 / 32768:
        Wp = Wp * (32767 / 32768) / m
    return q15_pack(Wp).reshape(-1)
