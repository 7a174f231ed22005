"""Seeded synthetic CPI scenes and coefficient sets for the radar chain.

The reference defines no stimulus for a composed chain (its only scene generator,
model/python/gen_target_iq.py, targets the FPGA pipeline), so the workloads here
are defined from the published configuration text:

* complex int16 I/Q per channel, channel-major ``iq[ch][frame*N + t]``, packed
  ``{im, re}`` in one uint32 with the real part in the low half
  (model/cpp/fxp/fxp.hpp:288-297);
* complex Gaussian noise of standard deviation ``sigma`` LSB per component plus
  plane-wave tones on a half-wavelength uniform linear array: channel ``a`` sees
  phase ``pi * a * sin(theta)``; amplitude ``sigma * 10**(snr_db/20)``;
* every real value quantised round-to-nearest-even then saturated to 16 bits
  (NUMERICS.md section 4 rule applied to reals, as model/python/gen_fft_twiddles.py:97-115);
* PFB taps: a Kaiser-windowed sinc prototype of length P*T with cutoff 1/P,
  quantised to Q1.15 and stored phase-major, tap-minor (pfb_model.hpp:108-111):
  ``coeff[p*T + k] = h[k*P + p]``;
* beam weights ``W[b][a] = q15(conj(s_b[a]) / CH)`` beam-major
  (beamformer_model.hpp:137-139);
* CFAR configuration words ``[enable, mode, out_mode, guard_lead, guard_lag,
  ref_lead, ref_lag, alpha]`` with alpha in UQ8.8 (cfar_model.hpp:58-67).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SCHED_ALL = 0xFFFFFFFF
MAX_GUARD = 31  # 5-bit guard count (cfar_model.hpp:61)
MAX_REF = 63    # 6-bit reference count (cfar_model.hpp:62)


def q16(v: np.ndarray) -> np.ndarray:
    """Real -> int16 by round-to-nearest-even then saturate."""
    r = np.rint(v)  # numpy rint rounds half to even
    return np.clip(r, -32768, 32767).astype(np.int64)


def pack(re: np.ndarray, im: np.ndarray) -> np.ndarray:
    re = np.asarray(re, np.int64) & 0xFFFF
    im = np.asarray(im, np.int64) & 0xFFFF
    return ((im << 16) | re).astype(np.uint32)


def to_complex(p: np.ndarray) -> np.ndarray:
    p = np.asarray(p, np.uint32)
    re = (p & 0xFFFF).astype(np.int64)
    im = (p >> 16).astype(np.int64)
    re = np.where(re >= 0x8000, re - 0x10000, re)
    im = np.where(im >= 0x8000, im - 0x10000, im)
    return re + 1j * im


def q15_cx(z: np.ndarray) -> np.ndarray:
    return pack(q16(np.real(z) * 32768.0), q16(np.imag(z) * 32768.0))


def kaiser_taps(P: int, T: int, beta: float = 8.0, peak: float = 0.9) -> np.ndarray:
    """Kaiser-windowed sinc channelizer prototype, cutoff 1/P, Q1.15, phase-major."""
    L = P * T
    n = np.arange(L) - (L - 1) / 2.0
    h = np.sinc(n / P) * np.kaiser(L, beta)
    h = h / np.max(np.abs(h)) * peak
    hq = q16(h * 32768.0)
    coeff = np.zeros(P * T, np.uint32)
    for p in range(P):
        for k in range(T):
            coeff[p * T + k] = pack(hq[k * P + p], 0)
    return coeff


def steering_weights(CH: int, angles_deg) -> np.ndarray:
    a = np.arange(CH)
    rows = []
    for th in np.deg2rad(np.asarray(angles_deg, float)):
        s = np.exp(1j * np.pi * a * np.sin(th))
        rows.append(np.conj(s) / CH)
    return q15_cx(np.array(rows)).reshape(-1)


@dataclass
class Scene:
    name: str
    CH: int
    N: int
    T: int
    F: int
    B: int
    sigma: float
    guard: int
    ref: int
    alpha: int = 0x0800
    spc: int = 2
    sched: int = SCHED_ALL
    targets: list = field(default_factory=list)  # (bin, angle_deg, snr_db)
    beam_angles: list = field(default_factory=list)
    seed: int = 1

    @property
    def samples(self) -> int:
        return self.CH * self.N * self.F

    def cfar_cfg(self):
        return np.array([1, 0, 0, self.guard, self.guard, self.ref, self.ref, self.alpha], np.uint32)

    def taps(self) -> np.ndarray:
        return kaiser_taps(self.N, self.T)

    def weights(self) -> np.ndarray:
        return steering_weights(self.CH, self.beam_angles)

    def iq(self, seed: int | None = None) -> np.ndarray:
        """Noise plus the targets' plane waves; the targets are separable in
        (channel, time), so their sum is one (CH x K) @ (K x L) product."""
        rng = np.random.default_rng(self.seed if seed is None else seed)
        CH, L = self.CH, self.N * self.F
        re = rng.standard_normal((CH, L)) * self.sigma
        im = rng.standard_normal((CH, L)) * self.sigma
        if self.targets:
            a = np.arange(CH)
            t = np.arange(L)
            steer = np.empty((CH, len(self.targets)), np.complex128)
            tones = np.empty((len(self.targets), L), np.complex128)
            for i, (k, ang, snr) in enumerate(self.targets):
                amp = self.sigma * 10 ** (snr / 20.0)
                ph0 = rng.uniform(0, 2 * np.pi)
                steer[:, i] = amp * np.exp(1j * (np.pi * a * np.sin(np.deg2rad(ang)) + ph0))
                tones[i] = np.exp(1j * (2 * np.pi * (k / self.N) * t))
            sig = steer @ tones
            re += sig.real
            im += sig.imag
        return pack(q16(re), q16(im)).reshape(-1)


def _random_targets(rng, n, N, snr_lo=6.0, snr_hi=25.0, ang=60.0):
    return [(int(rng.integers(8, N - 8)), float(rng.uniform(-ang, ang)), float(rng.uniform(snr_lo, snr_hi)))
            for _ in range(n)]


def config(name: str, seed: int = 1) -> Scene:
    """The published workloads: tiny (CPU-runnable), medium (plain GPU), large (split-int8 MFMA)."""
    rng = np.random.default_rng(1000 + seed)
    if name == "tiny":
        return Scene("tiny", CH=4, N=64, T=8, F=16, B=4, sigma=512.0, guard=2, ref=8,
                     targets=[(5, -20.0, 20.0), (17, 0.0, 20.0), (40, 25.0, 20.0)],
                     beam_angles=[-30.0, -10.0, 10.0, 30.0], seed=seed)
    if name == "medium":
        return Scene("medium", CH=16, N=1024, T=16, F=256, B=16, sigma=256.0, guard=4, ref=16,
                     targets=_random_targets(rng, 8, 1024),
                     beam_angles=list(np.linspace(-60.0, 60.0, 16)), seed=seed)
    if name == "large":
        return Scene("large", CH=64, N=4096, T=16, F=512, B=64, sigma=128.0, guard=4, ref=32,
                     targets=_random_targets(rng, 32, 4096),
                     beam_angles=list(np.rad2deg(np.arcsin(np.linspace(-0.95, 0.95, 64)))), seed=seed)
    raise ValueError(name)


def uneven_batch(n_cpi: int = 256, seed: int = 4):
    """Batched uneven CPIs: channels in {8,16,32,64}, FFT in {256,1024,4096},
    frames uniform in 32..512, beams = channels, 0-16 targets, sigma 256."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_cpi):
        CH = int(rng.choice([8, 16, 32, 64]))
        N = int(rng.choice([256, 1024, 4096]))
        F = int(rng.integers(32, 513))
        nt = int(rng.integers(0, 17))
        out.append(Scene(f"cpi{i}", CH=CH, N=N, T=16, F=F, B=CH, sigma=256.0, guard=4, ref=16,
                         targets=_random_targets(rng, nt, N),
                         beam_angles=list(np.rad2deg(np.arcsin(np.linspace(-0.9, 0.9, CH)))), seed=seed * 7919 + i))
    return out
