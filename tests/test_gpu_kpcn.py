"""The learned denoiser's HIP kernels (MFMA convolutions, weight gradients,
pooling, radiance pyramid, multi-scale filter reconstruction and its
backward, loss, Adam) and the whole network's gradients against the CPU
oracle (oracle/kpcn_oracle.py, pinned by finite differences).

Bounds: BF16-stored outputs within one BF16 rounding of the oracle evaluated
on the same BF16 operands (|d| <= 2^-7 |ref| + 1e-6); FP32 weight gradients
within rel-L2 1e-5 (exact BF16 products, FP32 sums in another order); the
whole network's gradients, which pass through 17 layers of BF16 activations
and gradients (a rounding that lands on the other side of a tie in one layer
moves every later layer's gradient), within 1.5x of the distance BF16
rounding itself puts between the BF16-emulating oracle and its FP64 form
(2e-3 to 9e-2 per layer here), at cosine >= 0.995 per layer and 0.999 over
all parameters.  The 1.5x: FP32 instead of FP64 accumulation alone, run
through the oracle on these inputs, moves the logits by 6e-4 .. 2.6e-3 and
flips up to a quarter of the coarsest level's BF16 activations, as large as
the GPU's own distance from the oracle."""
import ctypes
import math

import numpy as np
import pytest

import kpcn_check
import kpcn_oracle as ko
from gpu_util import require_gpu
from stratusgfx_amd import _abi, kpcn

pytestmark = pytest.mark.gpu

H, W = 37, 45  # partial 16x16 tiles on both axes; levels 37x45 .. 3x3


def _bf(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda().to(torch.bfloat16)


def _np(t):
    return t.float().cpu().numpy().astype(np.float64)


def _close_bf16(got, want, what):
    tol = 2.0 ** -7 * np.abs(want) + 1e-6
    bad = np.abs(got - want) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} off, worst {np.abs(got - want).max():.3e}"


@pytest.fixture(scope="module")
def lib():
    require_gpu()
    return _abi.load()


def _rand_layer(seed):
    rng = np.random.default_rng(seed)
    Wt = ko.to_bf16(rng.normal(0, 0.1, (9, 32, 32)))
    b = rng.normal(0, 0.1, 32).astype(np.float32)
    x = ko.to_bf16(rng.normal(0, 1, (H, W, 32)))
    return rng, Wt, b, x


def test_bf16_encode_is_nearest_even_on_every_pattern(lib):
    """sgi_kpcn_bf16_encode (the conversion every BF16 store of the network makes)
    on all 2^32 FP32 bit patterns: each non-NaN pattern u gives the integer
    round-to-nearest-even (u + 0x7fff + ((u >> 16) & 1)) >> 16 (denormals kept,
    overflow to infinity, signed zeros), each NaN a NaN."""
    torch = require_gpu()
    n = 1 << 28
    bad = nan_bad = 0
    out = torch.empty(n, dtype=torch.int16, device="cuda")
    for c in range(16):
        u = torch.arange(c * n, (c + 1) * n, dtype=torch.int64, device="cuda")
        f = u.to(torch.int32).view(torch.float32)  # the low 32 bits, as FP32
        _abi.call(lib, "sgi_kpcn_bf16_encode", f.data_ptr(), n, out.data_ptr(), 0)
        got = out.to(torch.int64) & 0xFFFF
        want = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
        nan = (((u >> 23) & 0xFF) == 0xFF) & ((u & 0x7FFFFF) != 0)
        bad += int(((got != want) & ~nan).sum())
        nan_bad += int((nan & ~((((got >> 7) & 0xFF) == 0xFF) & ((got & 0x7F) != 0))).sum())
        del u, f
    torch.cuda.synchronize()
    print(f"[bf16] 2^32 patterns: {bad} non-NaN mismatches, {nan_bad} NaNs not kept NaN", flush=True)
    assert bad == 0 and nan_bad == 0


def test_conv3x3_forward(lib):
    torch = require_gpu()
    _, Wt, b, x = _rand_layer(0)
    # every device tensor stays referenced until the kernels have run
    xd, wd, bd = _bf(torch, x), _bf(torch, Wt), torch.from_numpy(b).cuda()
    for relu, f32 in ((1, 0), (0, 1)):
        out = torch.empty((H, W, 32), dtype=torch.float32 if f32 else torch.bfloat16, device="cuda")
        _abi.call(lib, "sgi_kpcn_conv3x3", xd.data_ptr(), 0, wd.data_ptr(), bd.data_ptr(), relu, H, W,
                  out.data_ptr(), f32, 0)
        torch.cuda.synchronize()
        want = ko.conv3x3(x, Wt, b.astype(np.float64))
        if relu:
            want = np.maximum(want, 0)
        got = _np(out)
        if f32:
            assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
        else:
            _close_bf16(got, ko.to_bf16(want), "conv bf16")


def test_conv3x3_input_gradient(lib):
    """dgrad = conv with the mirrored, transposed weights and the ReLU gate."""
    torch = require_gpu()
    rng, Wt, _, dy = _rand_layer(1)
    y = ko.to_bf16(np.maximum(rng.normal(0, 1, (H, W, 32)), 0))
    params = np.zeros(ko.PARAMS, np.float32)
    params[:ko.LAYER_WEIGHTS] = Wt.reshape(-1)
    wts = kpcn.Weights(params)
    out = torch.empty((H, W, 32), dtype=torch.bfloat16, device="cuda")
    dyd, yd = _bf(torch, dy), _bf(torch, y)
    _abi.call(lib, "sgi_kpcn_conv3x3", dyd.data_ptr(), yd.data_ptr(), wts.layer_wt(0).data_ptr(), 0, 0, H, W,
              out.data_ptr(), 0, 0)
    torch.cuda.synchronize()
    _, _, dx = ko.conv3x3_backward(np.zeros((H, W, 32)), dy * (y > 0), Wt)
    _close_bf16(_np(out), ko.to_bf16(dx), "dgrad")


def test_conv3x3_weight_gradient(lib):
    torch = require_gpu()
    rng, _, _, x = _rand_layer(2)
    dy = ko.to_bf16(rng.normal(0, 1, (H, W, 32)))
    y = ko.to_bf16(np.maximum(rng.normal(0, 1, (H, W, 32)), 0))
    parts = torch.empty(int(lib.sgi_kpcn_wgrad_partials(H, W)), dtype=torch.float32, device="cuda")
    grad = torch.empty(ko.LAYER_PARAMS, dtype=torch.float32, device="cuda")
    xd, dyd = _bf(torch, x), _bf(torch, dy)
    for mask in (None, y):
        md = _bf(torch, mask) if mask is not None else None
        _abi.call(lib, "sgi_kpcn_conv3x3_wgrad", xd.data_ptr(), dyd.data_ptr(), md.data_ptr() if md is not None else 0,
                  H, W, parts.data_ptr(), grad.data_ptr(), 0)
        torch.cuda.synchronize()
        dz = dy if mask is None else dy * (mask > 0)
        dW, db, _ = ko.conv3x3_backward(x, dz, np.zeros((9, 32, 32)))
        g = grad.cpu().numpy().astype(np.float64)
        want = np.concatenate([dW.reshape(-1), db])
        e = np.linalg.norm(g - want) / np.linalg.norm(want)
        print(f"wgrad mask={mask is not None} rel-L2 {e:.2e}")
        assert e <= 1e-5


def test_pool_and_upsample(lib):
    torch = require_gpu()
    rng = np.random.default_rng(3)
    x = ko.to_bf16(rng.normal(0, 1, (H, W, 32)))
    h, w = (H + 1) // 2, (W + 1) // 2
    low = ko.to_bf16(rng.normal(0, 1, (h, w, 32)))
    prev = ko.to_bf16(rng.normal(0, 1, (H, W, 32)))
    xd, lowd = _bf(torch, x), _bf(torch, low)
    out = torch.empty((h, w, 32), dtype=torch.bfloat16, device="cuda")
    _abi.call(lib, "sgi_kpcn_pool", xd.data_ptr(), H, W, out.data_ptr(), 0)
    up = torch.empty((H, W, 32), dtype=torch.bfloat16, device="cuda")
    _abi.call(lib, "sgi_kpcn_upsample_add", xd.data_ptr(), lowd.data_ptr(), H, W, up.data_ptr(), 0)
    dlow = torch.empty((h, w, 32), dtype=torch.bfloat16, device="cuda")
    _abi.call(lib, "sgi_kpcn_upsample_backward", xd.data_ptr(), H, W, dlow.data_ptr(), 0, 0)
    dlow_acc = lowd.clone()
    _abi.call(lib, "sgi_kpcn_upsample_backward", xd.data_ptr(), H, W, dlow_acc.data_ptr(), 1, 0)
    din = _bf(torch, prev)
    _abi.call(lib, "sgi_kpcn_pool_backward", lowd.data_ptr(), H, W, din.data_ptr(), 1, 0)
    din0 = torch.empty((H, W, 32), dtype=torch.bfloat16, device="cuda")
    _abi.call(lib, "sgi_kpcn_pool_backward", lowd.data_ptr(), H, W, din0.data_ptr(), 0, 0)
    torch.cuda.synchronize()
    _close_bf16(_np(out), ko.to_bf16(ko.pool(x)), "pool")
    _close_bf16(_np(up), ko.to_bf16(x + ko.upsample(low, H, W)), "upsample_add")
    _close_bf16(_np(dlow), ko.to_bf16(ko.upsample_backward(x)), "upsample_backward")
    _close_bf16(_np(dlow_acc), ko.to_bf16(low + ko.upsample_backward(x)), "upsample_backward accumulate")
    _close_bf16(_np(din), ko.to_bf16(prev + ko.pool_backward(low, H, W)), "pool_backward accumulate")
    _close_bf16(_np(din0), ko.to_bf16(ko.pool_backward(low, H, W)), "pool_backward")


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def test_radiance_pyramid(lib):
    """Levels 1..4 of the 2x2-mean pyramid (partial tiles, odd level sizes)."""
    torch = require_gpu()
    rng = np.random.default_rng(10)
    for hh, ww in ((H, W), (16, 16), (70, 33)):
        rad = rng.uniform(0, 3, (hh, ww, 3)).astype(np.float32)
        n = int(lib.sgi_kpcn_pyramid_floats(hh, ww))
        want = ko.pyramid(rad.astype(np.float64))[1:]
        assert n == sum(p.size for p in want)
        rd = _dev(torch, rad)
        pyr = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        _abi.call(lib, "sgi_kpcn_radiance_pyramid", rd.data_ptr(), hh, ww, pyr.data_ptr(), 0)
        torch.cuda.synchronize()
        got = pyr.cpu().numpy().astype(np.float64)
        off = 0
        for l, p in enumerate(want, 1):
            g = got[off:off + p.size].reshape(p.shape)
            off += p.size
            assert np.abs(g - p).max() <= 1e-6 * np.abs(p).max(), (hh, ww, l)


def test_filter_apply_levels_and_backward(lib):
    """One reconstruction level (filter alone, filter blended with the upsampled
    coarser output), its backward and the bilinear transpose against the oracle."""
    torch = require_gpu()
    rng = np.random.default_rng(4)
    h, w = (H + 1) // 2, (W + 1) // 2
    logits = rng.normal(0, 2, (H, W, 32)).astype(np.float32)
    rad = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    coarse = rng.uniform(0, 2, (h, w, 3)).astype(np.float32)
    dout = rng.normal(0, 1, (H, W, 3)).astype(np.float32)
    derr = rng.normal(0, 1, (H, W)).astype(np.float32)
    ld, rd, cd, dd, ed = (_dev(torch, a) for a in (logits, rad, coarse, dout, derr))
    f64 = lambda a: a.astype(np.float64)
    for with_coarse in (False, True):
        out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        err = torch.empty((H, W), dtype=torch.float32, device="cuda")
        c_ptr = cd.data_ptr() if with_coarse else 0
        _abi.call(lib, "sgi_kpcn_apply", ld.data_ptr(), rd.data_ptr(), c_ptr, H, W, out.data_ptr(), err.data_ptr(), 0)
        dl = torch.empty((H, W, 32), dtype=torch.bfloat16, device="cuda")
        dup = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        _abi.call(lib, "sgi_kpcn_apply_backward", ld.data_ptr(), rd.data_ptr(), c_ptr, dd.data_ptr(),
                  ed.data_ptr() if with_coarse else 0, H, W, dl.data_ptr(), dup.data_ptr() if with_coarse else 0, 0)
        dc = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
        if with_coarse:
            _abi.call(lib, "sgi_kpcn_upsample_bilinear_backward", dup.data_ptr(), H, W, dc.data_ptr(), 0)
        torch.cuda.synchronize()
        c64 = f64(coarse) if with_coarse else None
        o, e = ko.apply(f64(logits), f64(rad), c64)
        assert np.abs(out.cpu().numpy() - o).max() <= 1e-5 * np.abs(o).max(), with_coarse
        assert np.array_equal(err.cpu().numpy(), logits[..., ko.ERR])
        dlo, dupo = ko.apply_backward(f64(logits), f64(rad), c64, f64(dout), f64(derr) if with_coarse else None)
        _close_bf16(_np(dl), ko.to_bf16(dlo), f"dlogits coarse={with_coarse}")
        if with_coarse:
            assert np.abs(dup.cpu().numpy() - dupo).max() <= 1e-6 * np.abs(dupo).max()
            dco = ko.upsample_bilinear_backward(dup.cpu().numpy().astype(np.float64))
            assert np.abs(dc.cpu().numpy() - dco).max() <= 1e-5 * np.abs(dco).max()


@pytest.mark.parametrize("ratio", [False, True])
def test_loss(lib, ratio):
    """sgi_kpcn_loss, and sgi_kpcn_loss_ratio (the error channel against the log ratio of the output's
    relative squared error to an alternative output's)."""
    torch = require_gpu()
    rng = np.random.default_rng(11)
    out = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    logits = rng.normal(0, 2, (H, W, 32)).astype(np.float32)
    target = rng.uniform(0.1, 1.5, (H, W, 3)).astype(np.float32)
    alt = (target + rng.normal(0, 0.3, (H, W, 3))).astype(np.float32)
    od, ld, td, ad = (_dev(torch, a) for a in (out, logits, target, alt))
    dout = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    derr = torch.empty((H, W), dtype=torch.float32, device="cuda")
    sums = torch.zeros(2, dtype=torch.float32, device="cuda")
    if ratio:
        _abi.call(lib, "sgi_kpcn_loss_ratio", od.data_ptr(), ld.data_ptr(), td.data_ptr(), ad.data_ptr(), H, W,
                  ctypes.c_float(0.25), ctypes.c_float(0.5), dout.data_ptr(), derr.data_ptr(), sums.data_ptr(), 0)
    else:
        _abi.call(lib, "sgi_kpcn_loss", od.data_ptr(), ld.data_ptr(), td.data_ptr(), H, W, ctypes.c_float(0.25),
                  ctypes.c_float(0.5), dout.data_ptr(), derr.data_ptr(), sums.data_ptr(), 0)
    torch.cuda.synchronize()
    L, do, de = ko.loss(out.astype(np.float64), logits[..., ko.ERR].astype(np.float64), target.astype(np.float64),
                        0.25, alt.astype(np.float64) if ratio else None)
    s = sums.cpu().numpy().astype(np.float64)
    P = H * W
    assert abs(s[0] / (3 * P) + 0.25 * s[1] / P - L) <= 1e-4 * abs(L)
    assert np.abs(dout.cpu().numpy() - 0.5 * do).max() <= 1e-5 * np.abs(0.5 * do).max()
    assert np.abs(derr.cpu().numpy() - 0.5 * de).max() <= 1e-4 * np.abs(0.5 * de).max()


def test_adam_step(lib):
    torch = require_gpu()
    rng = np.random.default_rng(5)
    p0 = rng.normal(0, 0.1, ko.PARAMS).astype(np.float32)
    g = rng.normal(0, 1e-3, ko.PARAMS).astype(np.float32)
    wts = kpcn.Weights(p0)
    tr_m = torch.zeros(ko.PARAMS, dtype=torch.float32, device="cuda")
    tr_v = torch.zeros(ko.PARAMS, dtype=torch.float32, device="cuda")
    gd = torch.from_numpy(g).cuda()
    for step in (1, 2):
        _abi.call(lib, "sgi_kpcn_adam", wts.params.data_ptr(), gd.data_ptr(), tr_m.data_ptr(), tr_v.data_ptr(),
                  ctypes.c_float(1e-3), ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8), step,
                  wts.w.data_ptr(), wts.wt.data_ptr(), wts.biases.data_ptr(), 0)
    torch.cuda.synchronize()
    m = np.zeros_like(p0, np.float64)
    v = np.zeros_like(p0, np.float64)
    p = p0.astype(np.float64)
    for step in (1, 2):
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        p -= 1e-3 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    got = wts.params.cpu().numpy()
    assert np.abs(got - p).max() <= 1e-6
    # the BF16 copies follow the new parameters, the transposed copy mirrors the taps
    Wl = got[:ko.LAYER_WEIGHTS].reshape(9, 32, 32)
    assert np.array_equal(_np(wts.layer_w(0)).reshape(9, 32, 32), ko.to_bf16(Wl))
    assert np.array_equal(_np(wts.layer_wt(0)).reshape(9, 32, 32), ko.to_bf16(Wl[::-1].transpose(0, 2, 1)))
    assert np.array_equal(wts.biases[:32].cpu().numpy(), got[ko.LAYER_WEIGHTS:ko.LAYER_PARAMS])


def test_network_gradients_match_oracle(lib):
    """One training forward / backward of the whole network on the GPU against the BF16-emulating oracle."""
    torch = require_gpu()
    rng = np.random.default_rng(6)
    params = kpcn.init_params(7)
    for l in range(ko.LAYERS):
        params[l * ko.LAYER_PARAMS + ko.LAYER_WEIGHTS:(l + 1) * ko.LAYER_PARAMS] = rng.normal(0, 0.05, 32)
    feat = np.zeros((H, W, 32))
    feat[..., :10] = rng.normal(0, 1, (H, W, 10))
    feat = ko.to_bf16(feat)
    rad = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    target = rng.uniform(0.1, 1.5, (H, W, 3)).astype(np.float32)
    wts = kpcn.Weights(params)
    tr = kpcn.Trainer(wts, H, W)
    tr.zero_grad()
    rd = torch.from_numpy(np.ascontiguousarray(rad)).cuda()
    td = torch.from_numpy(np.ascontiguousarray(target)).cuda()
    fd = _bf(torch, feat)
    tr.forward(fd, rd)
    tr.loss_backward(td, err_weight=0.1)
    torch.cuda.synchronize()
    kpcn_check.compare(tr, params, feat, rad, target, 0.1)


def test_denoise_learned_threshold_semantics(lib):
    """threshold <= 0: every tile falls back, so the outputs are sgi_denoise's bit for bit;
    threshold = inf: no tile falls back and gi is the network's reconstruction, bit for bit the one
    the training path's separate kernels compute; a threshold that flags some tiles: the flagged
    tiles' outputs are sgi_denoise's, the others the network's, bit for bit."""
    from gpu_util import Case
    from stratusgfx_amd import scene as sc
    from stratusgfx_amd.gi import GiPass

    torch = require_gpu()
    case = Case(sc.CORNELL_640x360)
    w = case.w
    Hh, Ww = w.height, w.width
    gi = GiPass(Ww, Hh, w.tile_size, w.max_vpls_per_tile)
    gi.set_frame(case.dev.frame)
    gi.set_csm(case.dev.csm_struct())
    n = w.num_vpls
    g = case.dev.gbuffer()
    upd = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    vis = torch.zeros((1 + n,), dtype=torch.int32, device="cuda")
    gi.cull_global(case.dev.vpls, n, upd, vis)
    ent = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    gi.visible_atlas_entries(case.dev.entries, vis, n, ent)
    f32 = dict(dtype=torch.float32, device="cuda")
    col, res = torch.zeros((Hh, Ww, 3), **f32), torch.zeros((Hh, Ww, 4), **f32)
    gi.gather_ex(g, upd, vis, ent, case.dev.cubes(), None, col, res, mode=_abi.GATHER_SAMPLES, samples=8)
    lit = torch.zeros((Hh, Ww, 3), **f32)
    gi.direct_light(g, lit)
    wts = kpcn.Weights(kpcn.init_params(0))
    ws = gi.kpcn_workspace()
    n_tiles = ((Ww + 15) // 16) * ((Hh + 15) // 16)

    def run_learned(thr):
        outs = [torch.full((Hh, Ww, c), -1.0, **f32) for c in (3, 3, 4)]
        hist = torch.full((Hh, Ww), -1.0, **f32)
        fb = torch.full((1 + n_tiles,), -1, dtype=torch.int32, device="cuda")
        gi.denoise_learned(g, None, vis, lit, col, res, None, None, wts.net(), thr, ws, outs[0], outs[1], outs[2],
                           hist, fb)
        return outs, hist, fb

    outs, hist, fb = run_learned(-1.0)
    ref = [torch.zeros((Hh, Ww, c), **f32) for c in (3, 3, 4)]
    rh = torch.zeros((Hh, Ww), **f32)
    gi.denoise(g, None, vis, lit, col, res, None, None, ref[0], ref[1], ref[2], rh)
    torch.cuda.synchronize()
    assert int(fb[0]) == n_tiles and bool((fb[1:] == 1).all())
    for a, b in zip(outs + [hist], ref + [rh]):
        assert torch.equal(a, b)

    outs, hist, fb = run_learned(float("inf"))
    feat = torch.empty((Hh, Ww, 32), dtype=torch.bfloat16, device="cuda")
    rad = torch.empty((Hh, Ww, 3), **f32)
    gi.kpcn_features(g, res, feat, rad)
    tr = kpcn.Trainer(wts, Hh, Ww)
    out = tr.forward(feat, rad)  # separate kernels: convolutions, pyramid, one apply per level
    torch.cuda.synchronize()
    assert int(fb[0]) == 0
    assert torch.equal(outs[1], out)
    assert torch.all(hist == 1.0)
    assert torch.equal(outs[0], lit + col * out)

    # a threshold between: each flagged tile takes sgi_denoise's outputs bit for bit (the a-trous
    # passes run over the flagged tiles and a margin wider than their reach), every other tile the
    # network's
    net_outs, net_hist = outs, hist
    lo, hi = -30.0, 30.0  # log thresholds: every tile falls back at lo, none at hi; bisect
    for _ in range(60):
        thr = math.exp(0.5 * (lo + hi))
        outs, hist, fb = run_learned(thr)
        torch.cuda.synchronize()
        c = int(fb[0])
        if 0 < c < n_tiles:
            break
        lo, hi = (math.log(thr), hi) if c == n_tiles else (lo, math.log(thr))
    count = int(fb[0])
    assert 0 < count < n_tiles, "no threshold splits the tiles"
    flags = fb[1:].view((Hh + 15) // 16, (Ww + 15) // 16).bool()
    assert int(flags.sum()) == count
    mask = flags.repeat_interleave(16, 0).repeat_interleave(16, 1)[:Hh, :Ww]
    print(f"threshold {thr:.3g}: {count} of {n_tiles} tiles fall back", flush=True)
    for a, at, nt in zip(outs + [hist], ref + [rh], net_outs + [net_hist]):
        m = mask if a.dim() == 2 else mask[..., None]
        assert torch.equal(a, torch.where(m, at, nt))
