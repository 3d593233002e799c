"""One example of the learned denoiser's training set at its full size, and
one training step on it:

- the example (stratusgfx_amd.train_kpcn's sample path): a seed-generated
  2000-box city block with 4096 VPLs (6x256x256 FP16 cubes), the block's
  default pose at 512x512; input = the noisy gather of 8 VPLs drawn per
  pixel plus the G-buffer features (BF16); target = the converged render over
  every visible VPL with FP64 sums.  The visible list equals the oracle's; on
  the 8-row band with the most GI-receiving pixels in each half of the image
  the noisy gather and the converged gather are within 1e-4 rel-L2 of the
  oracle's; the radiance equals reservoir.rgb / max(reservoir.a, 1e-5) and
  every feature channel is the BF16 rounding of its definition (log(1 + 64
  radiance), albedo, decoded normal, log(1 + view distance) / 8, zeros);
- one training step on the whole 512x512 example with the trainer's
  settings (error-head weight 1e-4, gradient scale 1000 / batch 4): level
  logits, reconstruction, loss and every layer's gradient against the
  BF16-emulating oracle (bounds in kpcn_check), from a fresh network and from
  the shipped weights; the Adam step's parameters equal the update rule on
  the GPU's gradient; the loss after the step, from the updated BF16 weight
  copies, against the oracle on the updated parameters."""
import ctypes
import dataclasses

import numpy as np
import pytest

import kpcn_check
import kpcn_oracle as ko
import oracle as orc
from gpu_util import require_gpu
from stratusgfx_amd import _abi, kpcn
from stratusgfx_amd import scene as sc
from stratusgfx_amd.train_kpcn import _sample

pytestmark = pytest.mark.gpu

SIZE, VPLS, SPP = 512, 4096, 8
ERR_WEIGHT, SCALE, LR = 1e-4, 1000.0 / 4, 2e-3
BAND = 8
REL_L2 = 1e-4


@pytest.fixture(scope="module")
def example():
    torch = require_gpu()
    print(f"[case] c3 example: city seed 1000, {VPLS} VPLs, {SIZE}x{SIZE}, {SPP} draws per pixel", flush=True)
    w = sc.Workload("city-train-1000", _abi.SCENE_CITY, 1000, SIZE, SIZE, VPLS, 256, _abi.TEXEL_F16, 1024, 16, 64)
    scene = sc.Scene(w)
    v = _abi.View()  # the block's default pose: the first of each training block's views
    _abi.call(scene.lib, "sgi_scene_default_view", scene.handle, SIZE, SIZE, ctypes.byref(v))
    scene.set_view(v)
    inp = sc.generate_device(scene)
    s = _sample(inp, SIZE, SIZE, VPLS, SPP, -1, keep_frame=True)
    torch.cuda.synchronize()
    print(f"[c3] example: eye ({v.eye[0]:.1f}, {v.eye[1]:.1f}, "
          f"{v.eye[2]:.1f}) yaw {v.yaw_deg:.1f} pitch {v.pitch_deg:.1f}; "
          f"{int(s.extra['vis'][0].item())} VPLs visible", flush=True)
    return dict(inp=inp, s=s)


def test_example_against_the_oracle(example):
    inp, s = example["inp"], example["s"]
    h = sc.to_host(inp)
    hg, hcsm = h.gbuffer(), h.csm_struct()
    upd, vis = orc.cull_global(h.frame, hcsm, h.vpls)
    c = int(vis[0])
    assert np.array_equal(s.extra["vis"].cpu().numpy()[:1 + c], vis)
    ent = h.entries[vis[1:]]
    res = s.extra["res"].cpu().numpy()
    target = s.target.cpu().numpy()
    # in each half of the image the 8-row band with the most pixels that receive GI (sky and
    # pixels out of every VPL's reach receive none)
    lit_rows = (target.sum(-1) > 0).reshape(SIZE // BAND, BAND * SIZE).sum(-1)
    half = SIZE // BAND // 2
    bands = [int(np.argmax(lit_rows[:half])) * BAND, (half + int(np.argmax(lit_rows[half:]))) * BAND]
    print(f"[c3] {float((target.sum(-1) > 0).mean()):.3f} of the pixels receive GI; bands {bands}: "
          f"{[int(lit_rows[b // BAND]) for b in bands]} of {BAND * SIZE} pixels", flush=True)
    for y0, y1 in ((b, b + BAND) for b in bands):
        _, o_noisy = orc.gather(h.frame, hg, upd, c, ent, h.cubes(), mode=orc.GATHER_SAMPLES, samples=SPP,
                                rejection_limit=-1, y0=y0, y1=y1)
        _, o_conv = orc.gather(h.frame, hg, upd, c, ent, h.cubes(), mode=orc.GATHER_CONVERGED, y0=y0, y1=y1)
        o_target = o_conv[..., :3] / np.maximum(o_conv[..., 3:4], 1e-5)
        e_n = orc.rel_l2(res[y0:y1], o_noisy)
        e_t = orc.rel_l2(target[y0:y1], o_target)
        relmse = float(np.mean((res[y0:y1, :, :3] / np.maximum(res[y0:y1, :, 3:4], 1e-5) - o_target) ** 2
                               / (o_target ** 2 + 1e-2)))
        print(f"[c3] rows {y0}..{y1}: noisy gather rel-L2 {e_n:.2e}, converged target rel-L2 {e_t:.2e}; "
              f"noisy relMSE vs target {relmse:.3f}", flush=True)
        assert e_n <= REL_L2 and e_t <= REL_L2
        assert np.abs(o_target).sum() > 0


def _features(h, res):
    """The network input restated in FP64 from the G-buffer and the noisy gather."""
    H, W = h.height, h.width
    rad = res[..., :3].astype(np.float64) / np.maximum(res[..., 3:4].astype(np.float64), 1e-5)
    f = np.zeros((H, W, 32))
    f[..., 0:3] = np.log1p(np.maximum(rad, 0.0) * 64.0)
    f[..., 3:6] = h.albedo[..., :3]
    n = h.normal.astype(np.float64) * 2.0 - 1.0
    f[..., 6:9] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    M = np.array(h.frame.invProjectionView, np.float64).reshape(4, 4).T  # column-major
    u = (np.arange(W) + 0.5) / W * 2.0 - 1.0
    v = (np.arange(H) + 0.5) / H * 2.0 - 1.0
    U, V = np.meshgrid(u, v)
    clip = np.stack([U, V, h.depth.astype(np.float64) * 2.0 - 1.0, np.ones_like(U)], -1)
    P = clip @ M.T
    P = P[..., :3] / P[..., 3:4]
    d = np.linalg.norm(np.array(h.frame.viewPosition, np.float64) - P, axis=-1)
    f[..., 9] = np.log1p(d) / 8.0
    return rad, f


def test_example_features(example):
    inp, s = example["inp"], example["s"]
    h = sc.to_host(dataclasses.replace(inp, atlases=[]))  # the G-buffer only
    res = s.extra["res"].cpu().numpy()
    rad, f = _features(h, res)
    got_rad = s.rad.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(got_rad - rad) <= 4e-7 * np.abs(rad) + 1e-30)
    got = s.feat.float().cpu().numpy().astype(np.float64)
    tol = 2.0 ** -8 * np.abs(f) * 1.01 + 1e-6
    bad = np.abs(got - f) > tol
    print(f"[c3] features: {int(bad.sum())} of {bad.size} outside one BF16 rounding; "
          f"radiance feature mean {f[..., 0:3].mean():.3f}, distance feature mean {f[..., 9].mean():.3f}", flush=True)
    assert not bad.any()
    assert np.all(got[..., 10:] == 0)


@pytest.mark.parametrize("weights", ["fresh", "shipped"])
def test_training_step_against_the_oracle(example, weights):
    torch = require_gpu()
    s = example["s"]
    params = kpcn.init_params(3) if weights == "fresh" else kpcn.load_params()
    feat = s.feat.float().cpu().numpy().astype(np.float64)
    rad = s.rad.cpu().numpy()
    target = s.target.cpu().numpy()
    wts = kpcn.Weights(params)
    tr = kpcn.Trainer(wts, SIZE, SIZE)
    tr.zero_grad()
    tr.forward(s.feat, s.rad)
    tr.loss_backward(s.target, err_weight=ERR_WEIGHT, scale=SCALE)
    torch.cuda.synchronize()
    L0, _ = kpcn_check.compare(tr, params, feat, rad, target, ERR_WEIGHT, SCALE, tag=f"[c3 {weights}] ")
    # the Adam step (first step: m = (1 - b1) g, v = (1 - b2) g^2, bias-corrected)
    g = tr.grad.cpu().numpy().astype(np.float64)
    tr.step(LR)
    torch.cuda.synchronize()
    p1 = wts.params.cpu().numpy()
    want_p1 = params.astype(np.float64) - LR * g / (np.abs(g) + 1e-8)
    assert np.abs(p1 - want_p1).max() <= 1e-6
    Wl = p1[:ko.LAYER_WEIGHTS].reshape(9, 32, 32)
    assert np.array_equal(wts.layer_w(0).float().cpu().numpy().reshape(9, 32, 32), ko.to_bf16(Wl).astype(np.float32))
    # the loss after the step, from the BF16 copies the step wrote
    tr.zero_grad()
    tr.forward(s.feat, s.rad)
    tr.loss_backward(s.target, err_weight=ERR_WEIGHT, scale=SCALE)
    torch.cuda.synchronize()
    H = W = SIZE
    L1 = kpcn_check.loss_of_sums(tr.loss_sums.cpu().numpy(), H, W, ERR_WEIGHT)
    _, _, want_L1, _ = kpcn_check.oracle_step(p1, feat, rad.astype(np.float64), target.astype(np.float64),
                                              ERR_WEIGHT, SCALE)
    _, _, want_L1_64, _ = kpcn_check.oracle_step(p1, feat, rad.astype(np.float64), target.astype(np.float64),
                                                 ERR_WEIGHT, SCALE, bf16=False)
    print(f"[c3 {weights}] loss before the step {L0:.6e}, after {L1:.6e} (oracle {want_L1:.6e}, "
          f"FP64 oracle {want_L1_64:.6e})", flush=True)
    assert abs(L1 - want_L1) <= max(2.0 * abs(want_L1 - want_L1_64), 1e-3 * abs(want_L1))
