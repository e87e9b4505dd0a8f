"""Workload generators reproduce the committed fixtures (CPU only)."""
import json
import os

from conftest import ROOT
from highmap_amd import workloads as W

GOLD = os.path.join(ROOT, "tests", "golden")


def test_splitmix64_known_answer():
    # first output of splitmix64 seeded with 0 (the generator's published
    # reference stream begins 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4)
    g = W.SplitMix64(0)
    assert g.next() == 0xE220A8397B1DCDAF
    assert g.next() == 0x6E789E6AA1B965F4


def test_batch_manifest_fixture():
    with open(os.path.join(GOLD, "batch_manifest.json")) as f:
        assert W.batch_manifest() == json.load(f)
    m = W.batch_manifest()
    assert len(m) == 96 and [x["seed"] for x in m] == list(range(100, 196))
    assert all(100 <= x["iterations"] <= 400 for x in m)
    assert all(x["ny"] in W.SIDES and x["nx"] in (x["ny"], x["ny"] // 2) for x in m)


def test_serve_requests_fixture_and_statistics():
    with open(os.path.join(GOLD, "serve_requests_head.json")) as f:
        assert W.serve_requests()[:50] == json.load(f)
    r = W.serve_requests()
    assert len(r) == 2000
    rate = len(r) / r[-1]["arrival_s"]
    assert 180 < rate < 220
    assert all(256 <= x["side"] <= 4096 and x["side"] % 64 == 0 for x in r)
    assert all(50 <= x["iterations"] <= 500 for x in r)
    frac = sum(x["kind"] == "vpipes+thermal" for x in r) / len(r)
    assert 0.66 < frac < 0.74
