"""The committed 64-chunk batch manifest is what the generator draws (seed 0)."""
import json
import os

from unnatural_planets_amd.spec import BATCH_SHAPES, BATCH_SIZES, batch_manifest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_manifest_fixture_matches_generator():
    fx = json.load(open(os.path.join(HERE, "golden", "batch_manifest.json")))
    assert fx["seed"] == 0
    assert fx["chunks"] == batch_manifest(0, 64)


def test_manifest_ranges():
    m = batch_manifest(0, 64)
    assert len(m) == 64
    assert {c["n"] for c in m} <= set(BATCH_SIZES) and {c["shape"] for c in m} <= set(BATCH_SHAPES)
    assert all(3 <= c["octaves"] <= 6 for c in m)
    assert len({c["seed"] for c in m}) == 64
