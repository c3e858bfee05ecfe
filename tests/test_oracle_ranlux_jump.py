"""The skip-ahead restatement of the ranlxd streams (oracle/ranlux_jump.py)
pinned against the known-answer vectors of devel/nompi/random/check1.c, the
oracle's own ranlxd (itself pinned against those vectors) and the recursion
stepped one word at a time.  The device's parallel generation jumps its
streams with the same algebra."""
import json
import os
import random

import numpy as np

import pyoracle as po
import ranlux_jump as rj

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_known_answers_by_jump():
    k = json.load(open(os.path.join(GOLD, "ranlux_kat.json")))["ranlxd"]
    n0 = (k["calls"] - 1) * k["per_call"] + k["offset"]
    got = [round(v * 2.0 ** 48) for v in rj.uniforms(k["seed"], k["level"], n0, n0 + 48)]
    assert got == [int(v) for v in k["values_times_2p48"]]


def test_jump_equals_oracle_stream():
    for level, seed in [(1, 1), (2, 12345), (1, 2 ** 30 + 7)]:
        n = 48 * 7 + 13
        ref = np.empty(n)
        po.lib().oc_kat_ranlxd(level, seed, po.ptr(ref), n)
        got = np.array(rj.uniforms(seed, level, 0, n))
        assert np.array_equal(got, ref), (level, seed)


def test_jump_equals_stepping():
    rnd = random.Random(11)
    lanes, carries = rj.seed_lanes(4711)
    states = [(lanes[l], carries[l]) for l in range(4)]
    states += [([rnd.randrange(rj.B) for _ in range(12)], rnd.randrange(2)) for _ in range(6)]
    for w, c in states:
        for n in (1, 2, 5, 11, 12, 13, 202, 397, 3 * 397 + 5):
            assert rj.jump(w, c, n) == rj.step(w, c, n), n


def test_residue_round_trip():
    rnd = random.Random(5)
    for _ in range(20):
        z = rnd.randrange(rj.M)
        w, c = rj.from_z(z)
        assert all(0 <= x < rj.B for x in w)
        assert rj.to_z(w, c) == z
    # a b = 1 mod m: one step back of a jump of one step
    assert rj.A * rj.B % rj.M == 1
