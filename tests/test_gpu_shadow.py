"""The FP32 differentiable shadow (rd_shadow, include/rd_shadow.h).

* forward kernels (PFB, FFT, beamformer + power) against the same model in
  float64 torch ops;
* backward kernels against float64 autograd of that model (vector-Jacobian
  products for x, W and the channel gains);
* the full loss (soft CA-CFAR, binary cross-entropy): FP32 gradients against
  float64 central finite differences, relative error < 1e-4;
* the shadow against the bit-exact chain, as the reference's dot_float
  cross-check does (beamformer_model.hpp:186-223): float beamformer within
  1/2 LSB plus float error of every non-saturating bit-exact output, float FFT
  within the rounding bound of the fixed-point stages;
* training steps lower the loss, and the quantised weights (gains folded in)
  run through the unchanged bit-exact chain, equal to the oracle.
"""
import numpy as np
import pytest

import rd_amd
import rd_scenario

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sh():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import rd_shadow

    return rd_shadow


def _rand_cx(g, shape, scale=1.0, dtype=torch.complex64):
    re = torch.randn(shape, generator=g, dtype=torch.float64)
    im = torch.randn(shape, generator=g, dtype=torch.float64)
    return (torch.complex(re, im) * scale).to(dtype).cuda()


def _rel(a, b):
    a, b = a.detach().to(torch.complex128).cpu(), b.detach().to(torch.complex128).cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


@pytest.mark.parametrize("Q,C,B,N,T,F", [(2, 4, 3, 64, 4, 8), (1, 16, 16, 1024, 16, 4), (3, 7, 5, 8, 3, 5),
                                         (1, 33, 64, 4096, 16, 2), (2, 64, 20, 256, 8, 3)])
def test_forward_kernels_match_fp64(sh, Q, C, B, N, T, F):
    g = torch.Generator().manual_seed(Q * 100 + C + N)
    x = _rand_cx(g, (Q, C, F * N), 0.1)
    h = _rand_cx(g, (N, T), 0.2)
    W = _rand_cx(g, (B, C), 0.1)
    gain = _rand_cx(g, (C,), 0.5) + 1
    u = sh.pfb(x.reshape(Q * C, F * N), h)
    X = sh.fft(u.reshape(-1, N), scale=1.0 / N).reshape(Q, C, F * N)
    p, y = sh.bf_power(X, (W * gain[None, :]).contiguous(), want_y=True)
    # float64 model
    x64, h64, W64, g64 = (t.to(torch.complex128) for t in (x, h, W, gain))
    xr = x64.reshape(Q, C, F, N)
    u64 = sum(h64[:, k] * torch.cat([torch.zeros_like(xr[:, :, :k]), xr[:, :, :F - k]], 2) for k in range(min(T, F)))
    X64 = torch.fft.fft(u64, dim=-1) / N
    y64 = torch.matmul(W64 * g64[None, :], X64.reshape(Q, C, F * N))
    assert _rel(u.reshape(Q, C, F, N), u64) < 1e-5
    assert _rel(X.reshape(Q, C, F, N), X64) < 1e-5
    assert _rel(y, y64) < 1e-5
    assert _rel(p, y64.abs() ** 2) < 1e-5
    p_ref = sh.reference(x64, W64, g64, h64)
    assert _rel(p, p_ref) < 1e-5


@pytest.mark.parametrize("Q,C,B,N,T,F", [(2, 4, 3, 64, 4, 8), (1, 32, 32, 1024, 16, 3), (2, 64, 64, 128, 5, 4)])
def test_backward_kernels_match_fp64_autograd(sh, Q, C, B, N, T, F):
    g = torch.Generator().manual_seed(7 + C + N)
    x = _rand_cx(g, (Q, C, F * N), 0.1).requires_grad_()
    h = _rand_cx(g, (N, T), 0.2)
    W = _rand_cx(g, (B, C), 0.1).requires_grad_()
    gain = (_rand_cx(g, (C,), 0.3) + 1).requires_grad_()
    gp = torch.randn((Q, B, F * N), generator=g, dtype=torch.float64).float().cuda()
    p = sh.FrontEnd.apply(x, W, gain, h)
    gx, gW, gg = torch.autograd.grad(p, (x, W, gain), grad_outputs=gp)
    x64, W64, g64 = (t.detach().to(torch.complex128).requires_grad_() for t in (x, W, gain))
    p64 = sh.reference(x64, W64, g64, h.to(torch.complex128))
    rx, rW, rg = torch.autograd.grad(p64, (x64, W64, g64), grad_outputs=gp.double())
    assert _rel(gx, rx) < 1e-4
    assert _rel(gW, rW) < 1e-4
    assert _rel(gg, rg) < 1e-4


def _scene_tensors(sh, Q, C, B, N, T, F, seed, targets_per=2, guard=2, ref=4):
    """Q seeded scenes (int16 I/Q with channel gain/phase errors), as the shadow
    sees them, plus the target mask over the soft-CFAR cells."""
    rng = np.random.default_rng(seed)
    err = 10 ** (rng.uniform(-1, 1, C) / 20) * np.exp(1j * np.deg2rad(rng.uniform(-10, 10, C)))
    angles = np.rad2deg(np.arcsin(np.linspace(-0.8, 0.8, B)))
    xs, mask = [], np.zeros((Q, B, F, N), np.float32)
    for q in range(Q):
        tg = [(int(rng.integers(guard + ref + 1, N - guard - ref - 1)), float(rng.uniform(-50, 50)),
               float(rng.uniform(15, 25))) for _ in range(targets_per)]
        sc = rd_scenario.Scene(f"s{q}", CH=C, N=N, T=T, F=F, B=B, sigma=256.0, guard=guard, ref=ref, targets=tg,
                               beam_angles=list(angles), seed=seed * 31 + q)
        z = rd_scenario.to_complex(sc.iq()).reshape(C, -1) * err[:, None]
        iq = rd_scenario.pack(rd_scenario.q16(z.real), rd_scenario.q16(z.imag))
        xs.append(iq)
        for k, ang, _ in tg:
            b = int(np.argmin(np.abs(np.sin(np.deg2rad(angles)) - np.sin(np.deg2rad(ang)))))
            mask[q, b, :, k] = 1.0
    lo = guard + ref
    taps, W0 = sc.taps(), sc.weights()
    return dict(iq=xs, taps=taps, W0=W0, mask=torch.from_numpy(mask[..., lo:N - lo]).cuda(), guard=guard, ref=ref,
                sc=sc)


def _to_cx(words, shape):
    return torch.from_numpy(rd_scenario.to_complex(np.asarray(words)).reshape(shape) / 32768.0)


def test_loss_gradients_vs_fp64_finite_differences(sh):
    Q, C, B, N, T, F = 2, 4, 4, 64, 4, 6
    d = _scene_tensors(sh, Q, C, B, N, T, F, seed=3)
    x = torch.stack([_to_cx(w, (C, F * N)) for w in d["iq"]]).to(torch.complex64).cuda()
    h = _to_cx(d["taps"], (N, T)).to(torch.complex64).cuda()
    W = _to_cx(d["W0"], (B, C)).to(torch.complex64).cuda().requires_grad_()
    gain = torch.ones(C, dtype=torch.complex64, device="cuda").requires_grad_()
    alpha, tau = 8.0, 0.1

    def loss32(W_, g_):
        p = sh.FrontEnd.apply(x, W_, g_, h)
        return sh.loss_fn(sh.soft_cfar_logits(p, N, d["guard"], d["ref"], alpha, tau), d["mask"])

    L = loss32(W, gain)
    gW, gg = torch.autograd.grad(L, (W, gain))
    x64, h64, m64 = x.to(torch.complex128), h.to(torch.complex128), d["mask"].double()

    def loss64(W_, g_):
        p = sh.reference(x64, W_, g_, h64)
        return sh.loss_fn(sh.soft_cfar_logits(p, N, d["guard"], d["ref"], alpha, tau), m64)

    W64, g64 = W.detach().to(torch.complex128), gain.detach().to(torch.complex128)
    eps = 1e-6
    fd, an = [], []
    for which, base, grad in (("W", W64, gW), ("g", g64, gg)):
        flat = base.reshape(-1)
        for i in range(flat.numel()):
            for part, unit in (("re", 1.0), ("im", 1j)):
                dv = torch.zeros_like(flat)
                dv[i] = unit * eps
                dv = dv.reshape(base.shape)
                if which == "W":
                    lp, lm = loss64(W64 + dv, g64), loss64(W64 - dv, g64)
                else:
                    lp, lm = loss64(W64, g64 + dv), loss64(W64, g64 - dv)
                fd.append(float((lp - lm) / (2 * eps)))
                gi = grad.reshape(-1)[i]
                an.append(float(gi.real if part == "re" else gi.imag))
    fd, an = np.array(fd), np.array(an)
    rel = np.linalg.norm(an - fd) / np.linalg.norm(fd)
    assert rel < 1e-4, rel
    assert np.max(np.abs(an - fd)) < 1e-4 * np.max(np.abs(fd))


def test_shadow_tracks_the_bit_exact_chain(gpu, sh):
    """Float beamformer of the bit-exact aligned FFT output within 1/2 LSB (plus
    float error) of every non-saturating bit-exact beam sample; float FFT of the
    bit-exact PFB output within the fixed-point rounding bound."""
    sc = rd_scenario.config("medium")
    sc.F = 8
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    ch.load(sc.iq(), sc.taps(), sc.weights())
    ch.run()
    got = ch.fetch(("u", "X", "y", "yflags"))
    u = _to_cx(got["u"], (sc.CH * sc.F, sc.N)).to(torch.complex64).cuda()
    Xf = sh.fft(u, scale=1.0 / sc.N).cpu().numpy().reshape(-1)
    Xb = rd_scenario.to_complex(got["X"]) / 32768.0
    # each of log2(N) scaled stages rounds by <= 1/2 LSB per component; twiddle
    # products add <= 1/2 LSB each
    bound = (np.log2(sc.N) + 1) / 32768.0
    err = np.abs(Xf - Xb)
    assert err.max() < bound, err.max() * 32768
    X = _to_cx(got["X"], (1, sc.CH, sc.F * sc.N)).to(torch.complex64).cuda()
    W = _to_cx(sc.weights(), (sc.B, sc.CH)).to(torch.complex64).cuda()
    _, y = sh.bf_power(X, W, want_y=True)
    yf = y.cpu().numpy().reshape(-1) * 32768.0
    yb = rd_scenario.to_complex(got["y"])
    ok = got["yflags"] == 0
    assert np.abs(yf.real - yb.real)[ok].max() <= 0.5 + 1e-2
    assert np.abs(yf.imag - yb.imag)[ok].max() <= 0.5 + 1e-2


def test_training_then_quantise_into_the_bit_exact_chain(gpu, orc, sh):
    Q, C, B, N, T, F = 4, 8, 8, 256, 16, 8
    d = _scene_tensors(sh, Q, C, B, N, T, F, seed=11, targets_per=3)
    x = torch.stack([_to_cx(w, (C, F * N)) for w in d["iq"]]).to(torch.complex64).cuda()
    h = _to_cx(d["taps"], (N, T)).to(torch.complex64).cuda()
    W = _to_cx(d["W0"], (B, C)).to(torch.complex64).cuda().requires_grad_()
    gain = torch.ones(C, dtype=torch.complex64, device="cuda").requires_grad_()
    opt = torch.optim.Adam([W, gain], lr=2e-3)
    losses = []
    for _ in range(15):
        opt.zero_grad()
        p = sh.FrontEnd.apply(x, W, gain, h)
        L = sh.loss_fn(sh.soft_cfar_logits(p, N, d["guard"], d["ref"], 8.0, 0.1), d["mask"])
        L.backward()
        opt.step()
        losses.append(float(L))
    assert losses[-1] < losses[0]
    Wq = sh.quantise_weights(W, gain)
    sc = d["sc"]
    cfg = np.array([1, 0, 0, d["guard"], d["guard"], d["ref"], d["ref"], 0x800], np.uint32)
    ch = rd_amd.Chain(gpu, C, N, T, F, B, cfg)
    ch.load(d["iq"][0], d["taps"], Wq)
    n = ch.run()
    got = ch.fetch(("power",), n_events=n)
    o = orc.chain(d["iq"][0], d["taps"], Wq, C, N, T, F, B, cfg, 31, 63, keep=("power", "events"))
    assert n == o["n_events"]
    assert np.array_equal(got["power"], o["power"])
    assert np.array_equal(got["events"], o["events"])
    del sc


def test_config5_gradients_at_full_size(sh):
    """The stated configuration: 64 CPIs x 32 ch x 1024-pt x 128 frames, 32x32
    complex weights and 32 channel gains. The FP32 gradients of the soft-CFAR
    loss (hand-written HIP backward kernels) against float64 autograd of the
    plain-torch model and float64 central finite differences: relative error
    below 1e-4."""
    import rd_workloads

    pb = rd_workloads.ShadowProblem(torch, n_cpi=64, n_train=48)
    assert pb.x.shape == (64, 32, 1024 * 128) and pb.W0.shape == (32, 32)
    rel, fd = pb.gradient_check(pb.W0, torch.ones(32, dtype=torch.complex64, device="cuda"), fd_dirs=2)
    assert rel < 1e-4, rel
    assert max(fd) < 1e-4, fd
    del pb
    torch.cuda.empty_cache()
