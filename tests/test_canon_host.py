"""The device arithmetic header (canon.h) compiled for the host agrees bit for
bit with the oracle -- the same FP32 sequences, written twice.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import rng

TWO_PI_F = np.float32(2.0) * np.float32(math.pi)


def test_canon_matches_oracle(H, oracle_mod):
    O = oracle_mod
    r = rng(90)
    ang = TWO_PI_F / np.float32(30)
    for _ in range(3000):
        p1, p2 = r.uniform(-100, 100, (2, 3)).astype(np.float32)
        n1, n2 = r.normal(size=(2, 3))
        n1 = (n1 / np.linalg.norm(n1)).astype(np.float32)
        n2 = (n2 / np.linalg.norm(n2)).astype(np.float32)
        for tail in (0, 1):
            assert H.ch_ppf_key(p1, n1, p2, n2, ang, np.float32(5.3), tail) == \
                O.ppf_key(p1, n1, p2, n2, ang, np.float32(5.3), tail)[0]
        rt = np.zeros(12, np.float32)
        H.ch_frame(p1, n1, rt)
        assert np.array_equal(rt, O.frame(p1, n1))
        a = H.ch_alpha(rt, p2)
        assert np.float32(a).view(np.uint32) == np.float32(O.alpha(rt, p2)).view(np.uint32)
        am, as_ = r.uniform(-math.pi, math.pi, 2).astype(np.float32)
        for tail in (0, 1):
            assert H.ch_bin(am, as_, 30, tail) == O.lib().orc_angle_bin(float(am), float(as_), 30, tail)
        y, x = r.normal(size=2).astype(np.float32)
        assert np.float32(H.ch_atan2f(y, x)) == np.float32(O.lib().orc_atan2f(float(y), float(x)))
    # poses: compose_pose with the oracle's cos/sin of bin j
    for j in range(30):
        p1, p2 = r.uniform(-100, 100, (2, 3)).astype(np.float32)
        n1, n2 = r.normal(size=(2, 3))
        Ts = O.frame(p1, (n1 / np.linalg.norm(n1)).astype(np.float32))
        Tm = O.frame(p2, (n2 / np.linalg.norm(n2)).astype(np.float32))
        a0 = TWO_PI_F * np.float32(j) / np.float32(29)
        a = np.float32(np.float64(a0) - math.pi)
        out = np.zeros(12, np.float32)
        H.ch_pose(Ts, Tm, np.float32(math.cos(float(a))), np.float32(math.sin(float(a))), out)
        assert np.array_equal(out, O.pose(Ts, Tm, j, 30))
        R = out.reshape(3, 4)[:, :3].copy()
        q1, q2 = np.zeros(4, np.float32), np.zeros(4, np.float32)
        H.ch_quat(np.ascontiguousarray(R.ravel()), q1)
        O.lib().orc_quat_from_rot(np.ascontiguousarray(R.ravel()), q2)
        assert np.array_equal(q1, q2)


def test_branch_free_bin_equals_both_rules(H, oracle_mod):
    """The voting kernels' branch-free bin (one wrap chosen from the
    unwrapped angle) equals the reference's body and tail rules, each with
    its own wrap order, on random angles and on the edges where a wrap
    result lands on 0 or 2pi."""
    O = oracle_mod
    r = rng(91)
    pi = np.float32(math.pi)
    am = np.concatenate([r.uniform(-math.pi, math.pi, 20000).astype(np.float32),
                         np.array([pi, -pi, 0.0, -0.0, np.nextafter(pi, 0), np.nextafter(-pi, 0)],
                                  np.float32)])
    as_ = np.concatenate([r.uniform(-math.pi, math.pi, 20000).astype(np.float32),
                          np.array([-pi, pi, pi, -pi, 0.0, np.float32(1e-7)], np.float32)])
    # pairs that make alphaModel - offset hit the wrap boundaries exactly
    edge_m = np.array([pi, -pi, pi, -pi, 0.0], np.float32)
    edge_s = np.array([pi, -pi, -pi, pi, -pi], np.float32)
    am = np.concatenate([am, edge_m, as_[:500]])
    as_ = np.concatenate([as_, edge_s, am[:500]])
    for n in (30, 29, 64):
        for tail in (0, 1):
            for x, y in zip(am, as_):
                ref = O.lib().orc_angle_bin(float(x), float(y), n, tail)
                assert H.ch_bin_wrapped(x, y, n, tail) == ref, (x, y, n, tail)
                # the voting loop's form (added wrap constant, float clamp)
                assert H.ch_bin_wrapped_fast(x, y, n, tail) == ref, (x, y, n, tail)


def test_libm_atan2f_divergence(H, oracle_mod):
    """How often glibc's atan2f (the reference's scalar-tail path,
    util.cpp:698,712-730, and computeAlpha's scalar form, :477-484) would
    change a key, an alpha or an angle bin against the canonical atan2f,
    over every ordered pair of the config-1 model's sampled points (the
    pairs trainModel hashes).  Reported, with a sanity bound: the two
    functions differ by at most an ulp or two, so a key changes only when a
    feature sits within an ulp of a quantisation step."""
    O = oracle_mod
    from surface_match_amd import synth
    mp, mn = synth.workload(1)["model"]
    m = O.train(mp, mn, 0.04)
    sp, sn = m.sampled()
    D = float(np.linalg.norm(np.asarray(m.boxMax) - np.asarray(m.boxMin)))
    step_a = np.float32(2.0) * np.float32(math.pi) / np.float32(30)
    step_d = np.float32(D) * np.float32(0.04)
    for tail in (0, 1):
        out = np.zeros(4, np.int64)
        H.ch_libm_divergence(np.ascontiguousarray(sp), np.ascontiguousarray(sn), len(sp), step_a,
                             step_d, tail, out)
        pairs, keys, alphas, bins = (int(x) for x in out)
        print(f"tail={tail}: {pairs} pairs; keys differ {keys} ({keys / pairs:.2e}), alphas "
              f"{alphas} ({alphas / pairs:.2e}), bins {bins} ({bins / pairs:.2e})")
        assert keys / pairs < 5e-3 and bins / pairs < 5e-3
    O.lib().orc_free_model(C.byref(m))
