"""CPU: pins of the oracle's ReSTIR DI restatement.

The reservoir arithmetic (wavefront_restir.glsl:51-180) is checked against
what resampled importance sampling fixes, not against a retyped formula:
selection frequencies proportional to the candidate weights, the unbiased
contribution weight W = w_sum / (M * target) carried unchanged through a
temporal merge at equal targets and through the history limit, the stored
12-byte form round-tripping, and the half-float packing against numpy's IEEE
binary16 conversion.  The estimator as a whole is pinned by its expectation:
with one candidate and no history, ReSTIR at a primary hit is plain one-sample
area-light NEE, so its image mean must converge to the plain estimator's.
"""

import numpy as np
import pytest

from ordinarylight_amd import cameras, packing, workloads
from oracle import native as O

NONE = 0xFFFFFFFF


def test_half_conversion_matches_ieee_binary16():
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.uniform(-70000, 70000, 4000),
        rng.uniform(-1, 1, 6000) * 2.0 ** rng.integers(-30, 17, 6000),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -14,
         1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf]]).astype(np.float32)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([O.float_to_half(float(v)) for v in vals], np.uint16)
    assert np.array_equal(got, ref)
    bits = np.arange(0, 65536, 7, dtype=np.uint16)
    back = np.array([O.half_to_float(int(b)) for b in bits], np.float32)
    want = bits.view(np.float16).astype(np.float32)
    assert np.all((back == want) | (np.isnan(back) & np.isnan(want)))


def test_selection_frequencies_follow_weights():
    """RIS: candidate i is kept with probability w_i / sum(w)."""
    w = np.array([1.0, 2.0, 0.0, 3.0, 4.0], np.float32)
    rng = np.random.default_rng(11)
    trials = 20000
    hits = np.zeros(len(w))
    for _ in range(trials):
        r, sel = O.reservoir_updates(np.arange(len(w)), w, w, rng.random(len(w)).astype(np.float32))
        hits[r.light] += 1
    p = w / w.sum()
    sigma = np.sqrt(trials * p * (1 - p))
    assert hits[2] == 0
    assert np.all(np.abs(hits - trials * p) <= 4 * sigma + 1e-9), hits
    assert r.w_sum == pytest.approx(w.sum()) and r.count == len(w)


def test_first_positive_candidate_is_always_taken_and_zero_weights_never():
    r, sel = O.reservoir_updates([7, 8, 9], [0.0, 2.0, 5.0], [0.0, 2.0, 0.0], [0.5, 0.999, 0.0])
    assert list(sel) == [False, True, False] and r.light == 8 and r.target == 2.0
    empty, sel = O.reservoir_updates([1, 2], [0.0, 0.0], [0.0, 0.0], [0.1, 0.2])
    assert empty.light == NONE and not sel.any() and empty.normalization() == 0.0


def test_temporal_merge_and_limit_keep_the_contribution_weight():
    rng = np.random.default_rng(5)
    t = rng.uniform(0.5, 4.0, 12).astype(np.float32)
    src, _ = O.reservoir_updates(np.arange(12) + 100, t, t, rng.random(12).astype(np.float32))
    W = src.normalization()
    # Merging into an empty reservoir at the same target transfers W and M.
    dst = O.Reservoir().merge(src, src.target, 0.3)
    assert dst.light == src.light and dst.count == src.count
    assert dst.normalization() == pytest.approx(W, rel=2e-3)
    # The history limit scales w_sum with M: W is unchanged.
    lim = src.limited(5.0)
    assert lim.count == 5.0 and lim.w_sum == pytest.approx(src.w_sum * 5.0 / 12.0, rel=1e-6)
    assert lim.normalization() == pytest.approx(W, rel=2e-3)
    assert src.limited(20.0).words.tolist() == src.words.tolist()
    # Merge weight of the temporal form: target_here * w_sum / target_src, M added.
    base, _ = O.reservoir_updates([1], [1.0], [1.0], [0.0])
    merged = O.Reservoir(base.words.copy()).merge(src, 2.0, 0.999)
    assert merged.count == 1 + src.count
    assert merged.w_sum == pytest.approx(1.0 + 2.0 * src.w_sum / src.target, rel=2e-3)


def test_pairwise_and_canonical_spatial_weights():
    src, _ = O.reservoir_updates([3, 4], [2.0, 2.0], [2.0, 2.0], [0.0, 0.9])
    assert src.count == 2.0
    # Canonical: target_here * w_sum / (target_src * M_src), one represented sample.
    c = O.Reservoir().merge(src, 1.5, 0.5, mode="canonical")
    assert c.count == 1.0 and c.w_sum == pytest.approx(1.5 * 4.0 / (2.0 * 2.0), rel=1e-3)
    # Pairwise: times the balance 2 t_there / (t_here + t_there); equal targets give 1.
    p = O.Reservoir().merge(src, 2.0, 0.5, mode="pairwise", target_there=2.0)
    assert p.w_sum == pytest.approx(2.0 * 4.0 / 4.0, rel=1e-3)
    q = O.Reservoir().merge(src, 1.0, 0.5, mode="pairwise", target_there=3.0)
    assert q.w_sum == pytest.approx(1.0 * 4.0 / 4.0 * (6.0 / 4.0), rel=1e-3)
    z = O.Reservoir().merge(src, 1.0, 0.5, mode="pairwise", target_there=0.0)
    assert z.light == NONE and z.w_sum == 0.0


def test_stored_form_round_trip_and_limits():
    r, _ = O.reservoir_updates([123456], [3.25], [3.25], [0.0], b1=[0.125], b2=[0.75])
    s = r.stored()
    assert s[0] == (123456 | (1 << 25)) and s[3] == 0
    back = O.Reservoir.loaded(s)
    assert back.light == 123456 and back.count == 1.0 and back.target == 3.25
    assert back.barycentrics == (0.125, 0.75) and back.w_sum == 3.25
    # Empty reservoirs store light 2^25 - 1 and load empty.
    e = O.Reservoir().stored()
    assert e[0] == 33554431 and O.Reservoir.loaded(e).light == NONE
    # Counts clamp at 127 and weights at the largest half.
    big, _ = O.reservoir_updates([5], [1.0], [1e6], [0.0], represented=[300.0])
    sb = big.stored()
    assert sb[0] >> 25 == 127 and O.half_to_float(int(sb[2]) & 0xFFFF) == 65504.0


def test_single_candidate_without_history_is_plain_area_nee():
    """Expectation pin of the whole ReSTIR estimator (FP64 oracle)."""
    scene, cam = workloads.instanced_spheres(1, meshes=3, subdivisions=2, instances=60, box=12.0, dielectric=0.2,
                                             emitters=40, emitter_size=0.3)
    orc = O.OracleScene(packing.pack_scene(scene), "f64")
    b, p = cameras.camera_basis(cam), cameras.camera_projection(cam)
    frames = 800
    r = np.array([orc.render_restir(b, p, 32, 24, 1, f, candidates=1, max_bounces=2,
                                    environment_samples=0)[..., :3].mean() for f in range(frames)])
    q = np.array([orc.render(b, p, 32, 24, 1, f, max_bounces=2, unified=False,
                             environment_samples=0)[..., :3].mean() for f in range(frames)])
    orc.close()
    se = np.sqrt(r.var() / frames + q.var() / frames)
    assert abs(r.mean() - q.mean()) <= 4 * se, (r.mean(), q.mean(), se)
    assert abs(r.mean() / q.mean() - 1.0) < 0.02
