"""Learned light-cluster proposal on the GPU, through the C ABI.

The network kernels round inputs, weights, hidden activations and
back-propagated deltas to bfloat16 (8 significant bits) and accumulate in
FP32; the checker is the binary64 network of oracle/vr_mlp.c evaluated on the
same bf16-rounded inputs and weights.  Tolerances (stated per test) follow
from the bf16 unit roundoff u = 2^-9 compounded over the 3 layers:
  forward   total variation distance of q per sample <= 0.02, row sums 1 +- 1e-4
  gradient  per parameter block ||g - g_ref|| <= 0.03 ||g_ref||, cosine >= 0.9995
  loss      relative error <= 5e-3
  Adam      FP32 update vs binary64 update from the same gradient: 1e-6 rel + 1e-7 abs
Sampling (k_proposal) is checked exactly against the network's own
probabilities: slot mass = (1-eps) q + eps P, stratified counts per cluster
within one of 32 m, and the cluster alias tables reproduce the in-cluster
power shares.  Everything
downstream of the proposal (candidate RIS, visibility, temporal and spatial
reuse, shading) is compared with the oracle bit for bit, the oracle taking the
GPU's candidate slots as its candidate lists.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from _parity import params_for_oracle, run_config
from smallproject_amd import Renderer, _lib, configs, default_config
from smallproject_amd._lib import LIGHT_DTYPE

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (0, 576), "b1": (576, 640), "W2": (640, 4736), "b2": (4736, 4800), "W3": (4800, 21184),
          "b3": (21184, 21440)}


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest even) and back."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def random_params(rng, out_scale=1.0) -> np.ndarray:
    """He-uniform weights (rounded to bf16, as the kernels use them) and
    small FP32 biases."""
    P = np.zeros(_lib.MLP_PARAMS, np.float32)
    for name, fan, scale in (("W1", 9, 1.0), ("W2", 64, 1.0), ("W3", 64, out_scale)):
        a, b = BLOCKS[name]
        P[a:b] = bf16(scale * np.sqrt(6.0 / fan) * rng.uniform(-1, 1, b - a))
    for name in ("b1", "b2", "b3"):
        a, b = BLOCKS[name]
        P[a:b] = rng.uniform(-0.1, 0.1, b - a)
    return P


def random_inputs(rng, B: int) -> np.ndarray:
    """Positions in [0,1)^3 (rounded to bf16) + one-hot face; 1 in 16 rows
    is the all-zero input of a pixel without a hit."""
    x = np.zeros((B, 9), np.float32)
    x[:, :3] = bf16(rng.uniform(0, 1, (B, 3)).astype(np.float32))
    x[np.arange(B), 3 + rng.integers(0, 6, B)] = 1.0
    x[::16] = 0.0
    return x


def context(nlights=512, nclusters=256, cluster_of=None):
    r = Renderer(default_config(16, 16))
    L = np.zeros(nlights, LIGHT_DTYPE)
    L["position"] = np.arange(nlights) * 7919 % (1 << 30)
    L["voxel"] = 0x010F0FFF
    L["size"] = 1
    r.set_lights(L)
    c = (np.arange(nlights) % nclusters) if cluster_of is None else cluster_of
    r.set_light_clusters(c.astype(np.uint32), 256 if cluster_of is None else nclusters)
    return r


def test_probs_match_oracle(gpu):
    rng = np.random.default_rng(21)
    r = context()
    for out_scale in (0.125, 1.0, 3.0):
        P = random_params(rng, out_scale)
        r.mlp_set_params(P)
        assert np.array_equal(r.mlp_get_params(), P)
        for B in (1, 63, 1000):
            x = random_inputs(rng, B)
            q = r.mlp_probs(x)
            _, qo = O.mlp_forward(P.astype(np.float64), x.astype(np.float64))
            tv = 0.5 * np.abs(q - qo).sum(axis=1)
            assert np.all(np.abs(q.sum(axis=1) - 1.0) <= 1e-4), "probabilities do not sum to 1"
            assert tv.max() <= 0.02, f"out_scale {out_scale} B {B}: max TV distance {tv.max():.4f}"


def test_probs_mask_empty_clusters(gpu):
    """Clusters without lights get probability 0; the rest is the softmax
    restricted to the non-empty clusters."""
    rng = np.random.default_rng(22)
    nl, k = 600, 200
    r = context(nl, k, cluster_of=np.arange(nl) % k)
    P = random_params(rng, 1.0)
    r.mlp_set_params(P)
    x = random_inputs(rng, 300)
    q = r.mlp_probs(x)
    assert np.all(q[:, k:] == 0.0)
    z, _ = O.mlp_forward(P.astype(np.float64), x.astype(np.float64))
    z = z[:, :k]
    qo = np.exp(z - z.max(axis=1, keepdims=True))
    qo /= qo.sum(axis=1, keepdims=True)
    assert 0.5 * np.abs(q[:, :k] - qo).sum(axis=1).max() <= 0.02


def test_grad_matches_oracle(gpu):
    rng = np.random.default_rng(23)
    r = context()
    for B in (1, 1000, 4096):
        P = random_params(rng, 1.0)
        r.mlp_set_params(P)
        x = random_inputs(rng, B)
        t = rng.integers(0, 256, B).astype(np.uint32)
        w = rng.choice([0.0, 1.0, 0.37, 1.9], B).astype(np.float32)
        w[0] = 1.0
        loss, g = r.mlp_grad(x, t, w)
        lo, go = O.mlp_loss_grad(P.astype(np.float64), x.astype(np.float64), t, w.astype(np.float64))
        assert abs(loss - lo) <= 5e-3 * abs(lo), f"B {B}: loss {loss} vs {lo}"
        for name, (a, b) in BLOCKS.items():
            ga, gr = g[a:b].astype(np.float64), go[a:b]
            nr = np.linalg.norm(gr)
            if nr == 0.0:
                assert np.all(ga == 0.0), f"B {B} {name}: gradient should vanish"
                continue
            rel = np.linalg.norm(ga - gr) / nr
            cos = float(ga @ gr) / (np.linalg.norm(ga) * nr)
            assert rel <= 0.03 and cos >= 0.9995, f"B {B} {name}: rel {rel:.4f} cos {cos:.6f}"


def test_adam_step_matches_oracle(gpu):
    """vr_mlp_step = the batch gradient (as vr_mlp_grad computes it) + Adam
    (Kingma & Ba, Algorithm 1, beta 0.9/0.999, eps 1e-8) on FP32 masters."""
    rng = np.random.default_rng(24)
    r = context()
    P = random_params(rng, 1.0)
    r.mlp_set_params(P)
    x = random_inputs(rng, 2000)
    t = rng.integers(0, 256, 2000).astype(np.uint32)
    w = np.ones(2000, np.float32)
    Pr = P.astype(np.float64)
    m = np.zeros_like(Pr)
    v = np.zeros_like(Pr)
    for step in (1, 2, 3):
        loss_g, g = r.mlp_grad(x, t, w)
        loss_s = r.mlp_step(x, t, w, lr=1e-3)
        assert loss_s == loss_g
        O.adam_step(Pr, g.astype(np.float64), m, v, step, lr=1e-3)
        got = r.mlp_get_params().astype(np.float64)
        err = np.abs(got - Pr)
        assert np.all(err <= 1e-6 * np.abs(Pr) + 1e-7), f"step {step}: max err {err.max():.3g}"
        Pr = got.copy()  # continue from the device's FP32 state
    assert loss_s < O.mlp_loss_grad(P.astype(np.float64), x.astype(np.float64), t)[0]


def _learned_frame(seed=25, W=96, H=64, nclusters=16, out_scale=2.0, mix=0.1, lights=None,
                   draw=_lib.CLUSTER_DRAW_POWER):
    sc = dataclasses.replace(configs.CONFIGS["c1_point_128"], source=_lib.SOURCE_LEARNED)
    if lights is not None:
        sc = dataclasses.replace(sc, lights=lights)
    w = configs.build_world(sc, W, H)
    w.force_sync()
    r = w.renderer
    clusters = w.cluster_lights(nclusters, 8, 1)
    r.set_cluster_draw(draw)
    r.set_light_clusters(clusters, nclusters)
    r.set_proposal_mix(mix)
    r.mlp_set_params(random_params(np.random.default_rng(seed), out_scale))
    pos, tgt = sc.camera(0)
    w.look_at(pos, tgt)
    p = w.prepare_frame()
    w.finish_frame()
    r.render(p)
    return w, r, p, clusters


def _pixel_inputs(r, p, W, H):
    """The proposal's inputs: hit position / 1024 (FP32, in the kernel's
    operation order) and the one-hot face; zero where the ray missed."""
    hw, d, _ = r.hits()
    D = O.camera_rays(params_for_oracle(p), W, H)
    E = np.array(p.E, np.float32)
    x = np.zeros((W * H, 9), np.float32)
    hit = (hw & 0xFFFFFF) != 0
    pos = (D * d[:, None] + E[None, :]) * np.float32(1.0 / 1024.0)
    x[hit, :3] = pos[hit]
    face = ((hw >> 24) & 3) + 3 * ((hw >> 26) & 1)
    x[np.nonzero(hit)[0], 3 + face[hit]] = 1.0
    return x, hit


@pytest.mark.parametrize("lights", [None, 12], ids=["64_lights", "12_lights"])
def test_proposal_sampling(gpu, lights):
    """Slots hold (cluster, mixture mass m_c = (1-eps) q_c + eps P_c) with q
    the network's own probabilities at the pixel's input; with S = min(lights,
    candidates) slots read by the candidate pass, every cluster gets floor or
    ceil of S m_c of the S stratified slots (+-1 for the FP32 walk): with 12
    lights the strata cover the mixture with 12 slots, not 12 of 32."""
    W, H, K, eps = 96, 64, 8 if lights else 16, 0.25
    w, r, p, clusters = _learned_frame(W=W, H=H, nclusters=K, mix=eps, lights=lights)
    S = min(len(w.lights()), 32)
    idx, mass = r.candidates()
    idx, mass = idx[:S], mass[:S]
    n = W * H
    assert np.all(idx < K), "every slot holds a non-empty cluster"
    t = r.cluster_tables(len(w.lights()))
    x, hit = _pixel_inputs(r, p, W, H)
    assert hit.mean() > 0.3 and (~hit).sum() > 0
    q = r.mlp_probs(x)                                   # [n, 256]
    keep, mix = np.float32(1.0) - np.float32(eps), np.float32(eps)
    m = keep * q + mix * t["power"][None, :]             # float32, the kernel's operation order
    px = np.broadcast_to(np.arange(n), idx.shape)
    assert np.allclose(mass, m[px, idx], rtol=1e-6, atol=0), "slot mass != (1-eps) q + eps P"
    counts = np.zeros((n, 256))
    np.add.at(counts, (px.ravel(), idx.ravel()), 1.0)
    bad = (counts < np.floor(S * m) - 1) | (counts > np.ceil(S * m) + 1)
    assert not bad.any(), f"{bad.any(axis=1).sum()} pixels with counts outside the strata"


@pytest.mark.parametrize("draw", [_lib.CLUSTER_DRAW_POWER, _lib.CLUSTER_DRAW_UNIFORM], ids=["power", "uniform"])
def test_cluster_tables(gpu, draw):
    """P_c sums to 1 and is the cluster's share of the power-source pmf;
    each cluster's alias table reproduces its in-cluster shares, which are
    the lights' power shares, or 1/n_c for the uniform in-cluster draw."""
    w, r, p, clusters = _learned_frame(nclusters=16, draw=draw)
    L = w.lights()
    t = r.cluster_tables(len(L))
    assert abs(float(t["power"].astype(np.float64).sum()) - 1.0) < 1e-5
    start = t["start"]
    for c in range(16):
        c0, c1 = int(start[c]), int(start[c + 1])
        members = t["lights"][c0:c1]
        assert np.all(clusters[members] == c)
        sh = t["share"][members].astype(np.float64)
        assert abs(sh.sum() - 1.0) < 1e-5
        cnt = c1 - c0
        if draw == _lib.CLUSTER_DRAW_UNIFORM:
            assert np.all(t["share"][members] == np.float32(1.0 / cnt))
            assert np.all(t["prob"][c0:c1] == 1.0)
        implied = np.zeros(cnt)
        for k in range(cnt):
            pk = float(t["prob"][c0 + k])
            implied[k] += pk / cnt
            implied[int(t["alias"][c0 + k])] += (1.0 - pk) / cnt
        assert np.allclose(implied, sh, atol=1e-6), f"cluster {c}: alias table != shares"


@pytest.mark.parametrize("draw", [_lib.CLUSTER_DRAW_POWER, _lib.CLUSTER_DRAW_UNIFORM], ids=["power", "uniform"])
def test_learned_frames_match_oracle(gpu, draw):
    """Candidate RIS with the learned slots, shadow rays, temporal and
    spatial reuse and shading: bit-exact against the oracle fed with the
    same slots, over 4 frames of the point-light scene and 2 of the
    area-light scene, with either in-cluster draw."""
    def setup(w, r):
        r.set_cluster_draw(draw)
        r.set_light_clusters(w.cluster_lights(16, 8, 1), 16)
        r.mlp_set_params(random_params(np.random.default_rng(26), 2.0))

    for name, frames, W, H in (("c1_point_128", 4, 160, 96), ("c2_area_1024", 2, 192, 108)):
        sc = configs.CONFIGS[name]
        res = run_config(sc, frames, W, H, source=_lib.SOURCE_LEARNED, setup=setup)
        for cmp in res:
            assert cmp.ok, f"{name}\n{cmp.report()}"


@pytest.mark.parametrize("lights", [64, 12])
def test_learned_source_unbiased(gpu, lights):
    """Without reuse the per-frame estimate is unbiased for any source pdf
    that covers the lights: the image mean over 24 independent frames agrees
    between the learned proposal and the power source within 4 standard
    errors, also with fewer lights than candidates (12: the candidate pass
    then draws 12 slots, which the proposal's strata must cover)."""
    means = {}
    for source in (_lib.SOURCE_POWER, _lib.SOURCE_LEARNED):
        sc = dataclasses.replace(configs.CONFIGS["c1_point_128"], source=source, lights=lights)
        w = configs.build_world(sc, 128, 72)
        w.force_sync()
        r = w.renderer
        if source == _lib.SOURCE_LEARNED:
            k = min(16, lights)
            r.set_light_clusters(w.cluster_lights(k, 8, 1), k)
            r.mlp_set_params(random_params(np.random.default_rng(27), 2.0))
        pos, tgt = sc.camera(2)
        vals = []
        for f in range(24):
            w.look_at(pos, tgt)
            p = w.prepare_frame()
            w.finish_frame()
            p.temporal, p.spatial = 0, 0
            r.render(p)
            vals.append(float(r.frame()[:, :3].astype(np.float64).mean()))
        means[source] = np.array(vals)
    a, b = means[_lib.SOURCE_POWER], means[_lib.SOURCE_LEARNED]
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    assert abs(a.mean() - b.mean()) <= 4 * se + 1e-3 * a.mean(), f"power {a.mean()} learned {b.mean()} se {se}"


def test_dynamic_run_falls_back_by_its_rule(gpu):
    """The dynamic configuration's learned-vs-plain run: each check compares
    MSE x frame time and picks the next source by that rule; the learned
    context's first period always uses the proposal."""
    from smallproject_amd.evaluate import dynamic
    r = dynamic("c5_dynamic_1024", 160, 96, frames=24, finetune_every=4, check_every=8, gt_spp=32,
                nsamples=1 << 12, train_steps=2, pretrain_frames=2)
    assert len(r["checks"]) == 3
    assert r["checks"][0]["source_until_now"] == "learned"
    for a, b in zip(r["checks"], r["checks"][1:]):
        assert b["source_until_now"] == a["next_source"]
    for c in r["checks"]:
        better = c["mse_learned_context"] * c["ms_learned_context"] <= c["mse_plain"] * c["ms_plain"]
        assert c["next_source"] == ("learned" if better else "plain")
        assert c["mse_plain"] > 0 and c["mse_learned_context"] > 0


def test_learned_back_to_back_frames_equal_synchronised_frames(gpu):
    """A lone context queues a learned frame's proposal behind its primary
    pass on its second stream, beside the previous frame's last stage, once
    the last queued training step is done.  Frames queued back to back, with
    training steps left queued between some of them, must come out bit for
    bit as the same frames each waited for (slots, hits, final reservoirs,
    radiance)."""
    outs = []
    for wait in (False, True):
        sc = dataclasses.replace(configs.CONFIGS["c1_point_128"], source=_lib.SOURCE_LEARNED)
        w = configs.build_world(sc, 160, 96)
        w.force_sync()
        r = w.renderer
        r.set_light_clusters(w.cluster_lights(16, 8, 1), 16)
        r.mlp_set_params(random_params(np.random.default_rng(31), 2.0))
        for f in range(6):
            pos, tgt = sc.camera(f)
            w.look_at(pos, tgt)
            w.render()
            if f in (1, 3):
                r.mlp_train(1 << 12, 1e-3, 2, want_loss=False)
            if wait:
                r.sync()
        idx, mass = r.candidates()
        hw, d, sd = r.hits()
        o = [idx, mass.view(np.uint32), hw, d.view(np.uint32), sd, r.frame().view(np.uint32)]
        a, b, m = r.reservoirs(0)
        o += [a.view(np.uint32), b.view(np.uint32), m]
        outs.append(o)
        w.close()
    for k, (g, e) in enumerate(zip(*outs)):
        assert np.array_equal(g, e), f"output {k}: {np.mean(g != e):.6f} differ"


def test_trained_proposal_beats_the_plain_source(gpu):
    """What the trained network gains over the method without it: on the
    many-light configuration (65,536 emitters in 256 clusters, at 480x270),
    the proposal trained online from the plain method's frames lowers the
    training loss, and with it the renderer's MSE against a converged
    reference (512 spp) is at most 0.7x the plain power source's on the
    first frame (measured 0.51x) and lower on average over the first 4 and
    over all 8 frames (MIS spatial merge, which has no bias floor).  The
    weights stay frozen during the comparison, so the result measures the
    pre-trained proposal alone.  The full-resolution equal-time figures,
    with training kept on during the comparison, are bench.py's quality
    report."""
    from smallproject_amd import evaluate
    r = evaluate.equal_time("c3_clusters_1024", width=480, height=270, train_frames=16, eval_frames=8, gt_spp=512,
                            train_steps=8, flags=_lib.FLAG_MIS_SPATIAL, online=False)
    eP, eL = r["plain"]["mse_per_frame"], r["learned"]["mse_per_frame"]
    print("loss", r["train_loss_first"], "->", r["train_loss_last"])
    print("plain  ", [round(e, 4) for e in eP])
    print("learned", [round(e, 4) for e in eL])
    assert r["train_loss_last"] < r["train_loss_first"]
    assert eL[0] <= 0.7 * eP[0]
    assert np.mean(eL[:4]) < np.mean(eP[:4])
    assert np.mean(eL) < np.mean(eP)

