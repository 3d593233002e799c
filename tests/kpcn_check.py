"""One training forward / backward of the learned denoiser on the GPU
(stratusgfx_amd.kpcn.Trainer) against the BF16-emulating oracle
(oracle/kpcn_oracle.py) on the same features, radiance and target.

Bounds: level logits within rel-L2 1e-2, the reconstruction within 1e-3; each
layer's gradient within 1.5x of the distance BF16 rounding itself puts between
the BF16-emulating oracle and its FP64 form (at least 2e-3), at cosine >=
0.995 per layer and 0.999 over all parameters."""
import numpy as np

import kpcn_oracle as ko


def oracle_step(params, feat, rad64, t64, err_weight, scale, bf16=True):
    """(logits, reconstructions, loss, grads [layers][layer params]) of the oracle; the output
    gradients are multiplied by scale before the backward pass, as sgi_kpcn_loss does."""
    net = ko.Net(np.asarray(params, np.float64), bf16=bf16)
    logits = net.forward(feat)
    o, err = ko.reconstruct(logits, rad64)
    L, dout, derr = ko.loss(o[0], err, t64, err_weight)
    grads = net.backward(ko.reconstruct_backward(logits, rad64, o, scale * dout, scale * derr))
    return logits, o, L, np.asarray(grads, np.float64).reshape(ko.LAYERS, ko.LAYER_PARAMS)


def loss_of_sums(sums, H, W, err_weight):
    """The loss from sgi_kpcn_loss's two accumulated sums (before the means)."""
    s = np.asarray(sums, np.float64)
    return s[0] / (3.0 * H * W) + err_weight * s[1] / (H * W)


def compare(tr, params, feat, rad, target, err_weight, scale=1.0, tag=""):
    """Checks the trainer's last forward / loss_backward (run by the caller with these arguments)
    against the oracle; returns (GPU loss, oracle loss)."""
    rad64, t64 = np.asarray(rad, np.float64), np.asarray(target, np.float64)
    H, W = rad64.shape[:2]
    want_logits, want_o, want_L, want = oracle_step(params, feat, rad64, t64, err_weight, scale)
    for l in range(ko.LEVELS):
        got_l = tr.logits[l].cpu().numpy().astype(np.float64)
        e_log = np.linalg.norm(got_l - want_logits[l]) / np.linalg.norm(want_logits[l])
        print(f"{tag}level {l} logits rel-L2 {e_log:.2e}", flush=True)
        assert e_log <= 1e-2, l
    e_out = np.linalg.norm(tr.o[0].cpu().numpy() - want_o[0]) / np.linalg.norm(want_o[0])
    print(f"{tag}reconstruction rel-L2 {e_out:.2e}", flush=True)
    assert e_out <= 1e-3
    # the size of BF16 rounding itself: the same oracle in FP64
    _, _, L64, want64 = oracle_step(params, feat, rad64, t64, err_weight, scale, bf16=False)
    got_L = loss_of_sums(tr.loss_sums.cpu().numpy(), H, W, err_weight)
    print(f"{tag}loss GPU {got_L:.6e} oracle {want_L:.6e} (FP64 oracle {L64:.6e})", flush=True)
    assert abs(got_L - want_L) <= max(2.0 * abs(want_L - L64), 1e-3 * abs(want_L))
    got = tr.grad.cpu().numpy().astype(np.float64).reshape(ko.LAYERS, ko.LAYER_PARAMS)
    bad = []
    for l in range(ko.LAYERS):
        a, b = got[l], want[l]
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        noise = np.linalg.norm(b - want64[l]) / np.linalg.norm(want64[l])
        cos = (a @ b) / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"{tag}layer {l:2d} grad rel-L2 {e:.2e} (BF16 vs FP64 oracle {noise:.2e}) cos {cos:.6f}", flush=True)
        if not (e <= 1.5 * max(noise, 2e-3) and cos >= 0.995):
            bad.append(l)
    assert not bad, f"layers {bad}"
    cos_all = (got.ravel() @ want.ravel()) / (np.linalg.norm(got) * np.linalg.norm(want))
    print(f"{tag}all layers cos {cos_all:.6f}", flush=True)
    assert cos_all >= 0.999
    return got_L, want_L
