"""GPU checks of the proposal network's hand-written training and inference
kernels against the FP64 numpy restatement (oracle/cnn_ref.py, itself pinned
by finite differences), and of what training gains.

Tolerances: FP32 forward within 1e-5 of the logit scale; FP32 gradients within
1e-4 of each parameter block's largest gradient; BF16-weight inference equal
to the FP64 forward of the BF16-rounded weights within 1e-3 in probability."""
import math
import sys

import numpy as np
import pytest
import torch

from bsoft_amd import search, synth
from bsoft_amd.proposal import (EmbeddingNet, OrientationClasses, ProposalNet, class_references,
                                learned_search, pair_set, training_set)
from bsoft_amd.recon import map_transform
from oracle import cnn_ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16_round(p: np.ndarray) -> np.ndarray:
    u = p.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("S,ch,ncls,B", [(16, (8, 8, 16, 8), 37, 5), (32, (32, 64, 128, 128), 300, 16)])
def test_forward_and_gradients_match_oracle(S, ch, ncls, B):
    say(f"cnn S={S} ch={ch} classes={ncls} batch={B}")
    net = ProposalNet(S, ncls, ch, max_batch=64, seed=3)
    rng = np.random.default_rng(S)
    x = rng.normal(size=(B, S, S)).astype(np.float32)
    y = rng.integers(0, ncls, B).astype(np.int32)
    xd = torch.as_tensor(x).cuda()
    logits = net.forward(xd).cpu().numpy()
    g, loss = net.gradients(xd, torch.as_tensor(y).cuda())
    p = net.params().cpu().numpy().astype(np.float64)
    ol, _ = cnn_ref.forward(p, x.astype(np.float64), S, ch, ncls)
    assert np.abs(logits - ol).max() <= 1e-5 * max(1.0, np.abs(ol).max())
    L, og, olosses = cnn_ref.loss_and_grad(p, x.astype(np.float64), y, S, ch, ncls)
    np.testing.assert_allclose(loss.cpu().numpy(), olosses, rtol=1e-5, atol=1e-5)
    g = g.cpu().numpy()
    off, n, feat, _ = cnn_ref.layout(S, ch, ncls)
    for o, ob, c in off:
        for a, b in ((o, ob), (ob, ob + c)):
            ref = og[a:b]
            err = np.abs(g[a:b] - ref).max()
            say(f"  block [{a}:{b}] max|g| {np.abs(ref).max():.3e} err {err:.3e}")
            assert err <= 1e-4 * np.abs(ref).max() + 1e-7


@pytest.mark.parametrize("S,ch,d,B", [(16, (8, 8, 16, 8), 12, 6), (32, (16, 32, 64, 64), 64, 24)])
def test_pair_gradients_match_oracle(S, ch, d, B):
    """bsx_cnn_pair_gradients (contrastive loss of B particle / reference
    pairs) against the FP64 restatement, and the embedding's unit rows."""
    say(f"pairs S={S} ch={ch} dim={d} batch={B}")
    net = EmbeddingNet(S, d, ch, max_batch=64, seed=4)
    rng = np.random.default_rng(S + 1)
    x = rng.normal(size=(2 * B, S, S)).astype(np.float32)
    xd = torch.as_tensor(x).cuda()
    tau = 0.2
    g, loss = net.pair_gradients(xd[:B], xd[B:], tau)
    p = net.params().cpu().numpy().astype(np.float64)
    L, og, olosses = cnn_ref.pair_loss_and_grad(p, x.astype(np.float64), B, tau, S, ch, d)
    np.testing.assert_allclose(loss.cpu().numpy(), olosses, rtol=1e-5, atol=1e-5)
    g = g.cpu().numpy()
    off, n, feat, _ = cnn_ref.layout(S, ch, d)
    for o, ob, c in off:
        for a, b in ((o, ob), (ob, ob + c)):
            ref = og[a:b]
            err = np.abs(g[a:b] - ref).max()
            say(f"  block [{a}:{b}] max|g| {np.abs(ref).max():.3e} err {err:.3e}")
            assert err <= 1e-4 * np.abs(ref).max() + 1e-7
    # embeddings: unit rows of the (BF16-weight) outputs
    e = net.embed(xd).cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(e, axis=1), 1.0, atol=1e-5)
    z, _ = cnn_ref.forward(bf16_round(p), x.astype(np.float64), S, ch, d)
    u = z / np.linalg.norm(z, axis=1, keepdims=True)
    assert np.abs(e - u).max() <= 2e-2
    # match top-k: the best references by cosine, probabilities a softmax over all
    cls, prob = net.match_topk(torch.as_tensor(e[:B], dtype=torch.float32).cuda(),
                               torch.as_tensor(e[B:], dtype=torch.float32).cuda(), 3, tau)
    cos = e[:B] @ e[B:].T
    for i in range(B):
        best = np.argsort(-cos[i], kind="stable")[:3]
        assert np.allclose(cos[i][cls[i].cpu().numpy()], cos[i][best], atol=1e-5)
        sm = np.exp(cos[i] / tau - (cos[i] / tau).max())
        sm /= sm.sum()
        np.testing.assert_allclose(prob[i].cpu().numpy(), sm[cls[i].cpu().numpy()], rtol=1e-4)


@pytest.mark.parametrize("ch", [(8, 8, 16, 8), (32, 64, 32, 64)])
def test_training_reduces_loss_and_bf16_inference(ch):
    """Adam memorises a fixed batch; BF16 inference (VALU path for small
    channel counts, MFMA implicit GEMM when the channels are multiples of
    32) equals the FP64 forward of the BF16-rounded weights with BF16
    activations."""
    S, ncls, B = 16, 20, 64
    net = ProposalNet(S, ncls, ch, max_batch=B, seed=5)
    rng = np.random.default_rng(1)
    x = torch.as_tensor(rng.normal(size=(B, S, S)).astype(np.float32)).cuda()
    y = torch.as_tensor(rng.integers(0, ncls, B).astype(np.int32)).cuda()
    first = net.train_step(x, y, 3e-3).mean().item()
    for _ in range(150):
        last = net.train_step(x, y, 3e-3).mean().item()
    say(f"  loss {first:.3f} -> {last:.3f}")
    assert last < 0.2 * first
    cls, prob = net.topk(x, 5)
    p = bf16_round(net.params().cpu().numpy())
    mfma = all(c % 32 == 0 for c in ch[:3])
    ol, _ = cnn_ref.forward(p, x.cpu().numpy().astype(np.float64), S, ch, ncls,
                            round_act=bf16_round if mfma else None)
    e = np.exp(ol - ol.max(axis=1, keepdims=True))
    pr = e / e.sum(axis=1, keepdims=True)
    order = np.argsort(-pr, axis=1, kind="stable")[:, :5]
    cls, prob = cls.cpu().numpy(), prob.cpu().numpy()
    np.testing.assert_allclose(prob, np.take_along_axis(pr, cls, 1), atol=1e-3)
    # the chosen classes are the most probable ones (up to near-ties)
    assert np.all(np.take_along_axis(pr, cls, 1)[:, -1] >= np.take_along_axis(pr, order, 1)[:, -1] - 1e-3)
    assert np.mean(cls[:, 0] == y.cpu().numpy()) > 0.9


def test_classes_assign_nearest():
    cl = OrientationClasses(20.0)
    rng = np.random.default_rng(2)
    m = synth.view_to_matrix(synth.uniform_views(300, rng))
    a = cl.assign(m)
    tr = np.einsum("cij,nij->nc", cl.mats, m)
    assert np.mean(a == tr.argmax(axis=1)) > 0.99


def test_trained_proposal_beats_untrained():
    """A short training run on extracted projections (64^3 blob map, 32x32
    network input) puts the true class into the top-16 proposals far more
    often than the untrained network, and the learned pipeline's final
    orientations agree with the exhaustive search's."""
    n, R, S = 64, 16, 32
    rng = np.random.default_rng(7)
    bm = synth.make_blob_map(n, 40, rng)
    ft = map_transform(bm.volume().cuda())
    classes = OrientationClasses(15.0)
    s = search.Search(n, R, 1.5, topk=8)
    say(f"classes {classes.n}; generating training set")
    x, y, _ = training_set(ft, n, 40000, classes, s, S, rng, snr=0.5)
    net = ProposalNet(S, classes.n, (16, 32, 64, 64), max_batch=512, seed=11)
    ps = synth.make_particles(bm, 400, n, rng, snr=0.5, seed=7)
    truth = synth.view_to_matrix(ps.parts["view"])
    tcls = torch.as_tensor(classes.assign(truth)).cuda()
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    xin = s.cnn_input(S)
    c0, _ = net.topk(xin, 16)
    hit0 = (c0.long() == tcls[:, None]).any(dim=1).float().mean().item()
    hist = net.fit(x, y, epochs=6, batch=256, lr=2e-3, log=say)
    c1, p1 = net.topk(xin, 16)
    hit1 = (c1.long() == tcls[:, None]).any(dim=1).float().mean().item()
    say(f"  top-16 hit rate untrained {hit0:.3f} trained {hit1:.3f}; loss {hist[0]:.3f} -> {hist[-1]:.3f}")
    assert hit1 > 0.8 and hit1 > 5 * hit0
    # learned vs exhaustive on the same particles
    views = search.views_c1(4.0, angle_step_deg=4.0)
    lm = search.view_matrices(views)
    s.set_reference(ft, lm)
    s.set_classes(classes.members(lm))
    ex = search.exhaustive_search(s)
    le, st, _, _ = learned_search(s, net, k=16, min_prob=0.0)
    d = search.rotation_angle_deg(le.mats.cpu().numpy(), ex.mats.cpu().numpy())
    et = search.rotation_angle_deg(ex.mats.cpu().numpy(), truth)
    say(f"  learned vs exhaustive median {np.median(d):.2f} deg (exhaustive vs truth "
        f"{np.median(et):.2f}); {st.candidates} candidates/particle")
    assert np.median(d) <= 1.0



def test_embedding_generalises_to_unseen_map():
    """The contrastive embedding network trained on particle / clean
    projection pairs of eight blob maps proposes orientations for a ninth
    map it has never seen: the true 15-degree class is among the top-16
    cosine matches of the map's own class projections far more often than
    with the untrained network, and the learned pipeline (proposals -> BF16
    GEMM over their member views -> FP32 FRC -> refinement) ends where the
    exhaustive search ends."""
    n, R, S, snr, tau = 64, 16, 32, 0.2, 0.1
    sp = search.Search(n, R, 1.5, topk=8)
    s_ref = search.Search(n, R, 1.5, topk=8, ctf=False)
    xps, xrs = [], []
    for m in range(8):
        rng = np.random.default_rng(100 + m)
        bm = synth.make_blob_map(n, 30, rng)
        ft = map_transform(bm.volume().cuda())
        xp, xr, _ = pair_set(ft, n, 3000, sp, s_ref, S, rng, snr=snr)
        xps.append(xp)
        xrs.append(xr)
    xp, xr = torch.cat(xps), torch.cat(xrs)
    net = EmbeddingNet(S, 64, (16, 32, 64, 64), max_batch=512, seed=5)
    rng = np.random.default_rng(7)
    bm = synth.make_blob_map(n, 40, rng)                 # the unseen map
    ft = map_transform(bm.volume().cuda())
    classes = OrientationClasses(15.0)
    s = search.Search(n, R, 1.5, topk=8)
    ps = synth.make_particles(bm, 400, n, rng, snr=snr, seed=7)
    truth = synth.view_to_matrix(ps.parts["view"])
    tcls = torch.as_tensor(classes.assign(truth)).cuda()
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    xin = s.cnn_input(S)

    def hit_rate():
        refs = class_references(net, ft, n, classes, s_ref, S)
        c, _ = net.match_topk(net.embed(xin), refs, 16, tau)
        return (c.long() == tcls[:, None]).any(dim=1).float().mean().item(), refs

    h0, _ = hit_rate()
    hist = net.fit_pairs(xp, xr, epochs=6, batch=128, lr=2e-3, tau=tau, log=say)
    h1, refs = hit_rate()
    say(f"  {len(xp)} pairs from 8 maps; {classes.n} classes; top-16 hit rate on the unseen map "
        f"untrained {h0:.3f} trained {h1:.3f}; pair loss {hist[0]:.3f} -> {hist[-1]:.3f}")
    assert hist[-1] < 0.5 * hist[0]
    assert h1 >= 0.9 and h1 > h0
    lm = search.view_matrices(search.views_c1(4.0, angle_step_deg=4.0))
    s.set_reference(ft, lm)
    s.set_classes(classes.members(lm))
    ex = search.exhaustive_search(s)
    le, st, _, _ = learned_search(s, net, k=16, min_prob=0.0, refs=refs, tau=tau)
    d = search.rotation_angle_deg(le.mats.cpu().numpy(), ex.mats.cpu().numpy())
    say(f"  learned vs exhaustive median {np.median(d):.2f} deg, {np.mean(d < 1):.3f} within 1 deg")
    assert np.median(d) <= 1.0
