"""The stack engine's comparator networks (ndfiltrs_amd/csrc/ndf_stack_networks.inc).

Each network is proven by the 0-1 principle in gen_stack_networks.verify():
column sorts and merges on every 0-1 input of their class, whole networks
exhaustively where that is small and otherwise through every two-list merge
they are built from, plus random keys with ties and all-ones sentinels. The
committed header must be exactly what the generator writes, and a planted
mistake in a network must be caught.
"""

from __future__ import annotations

import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "ndfiltrs_amd", "csrc", "gen_stack_networks.py")


def load_generator():
    spec = importlib.util.spec_from_file_location("gen_stack_networks", SCRIPT)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


GEN = load_generator()


@pytest.fixture(scope="module")
def networks() -> dict:
    return {cfg: GEN.build(*cfg) for cfg in GEN.CONFIGS}


@pytest.mark.parametrize("cfg", GEN.CONFIGS, ids=["x".join(map(str, c)) for c in GEN.CONFIGS])
def test_network_is_correct(networks, cfg) -> None:
    GEN.verify(networks[cfg])


def test_header_matches_generator() -> None:
    with open(GEN.HEADER) as f:
        committed = f.read()
    assert committed == GEN.render(), "ndf_stack_networks.inc is stale: run gen_stack_networks.py"


def test_merge_matches_sorted_union_for_many_lengths() -> None:
    rng = np.random.default_rng(3)
    for m in range(0, 12):
        for n in range(0, 12):
            comps, out = GEN.merge(list(range(m)), list(range(m, m + n)))
            keys = np.sort(rng.integers(0, 4, size=(200, m)), axis=1)
            more = np.sort(rng.integers(0, 4, size=(200, n)), axis=1)
            both = np.concatenate([keys, more], axis=1)
            got = GEN.apply(comps, both)[:, out] if m + n else both
            assert np.array_equal(got, np.sort(both, axis=1)), (m, n)


def test_planted_mistakes_are_caught(networks) -> None:
    """Dropping one comparator, or swapping the direction of one, from any
    network of a configuration makes verify() fail."""
    rng = np.random.default_rng(8)
    for cfg in ((5, 1, 5), (3, 3, 3)):
        net = networks[cfg]
        for part in ("slice", "half", "full"):
            comps = net[part]
            for index in rng.choice(len(comps), size=min(6, len(comps)), replace=False):
                for mistake in ("drop", "flip"):
                    broken = dict(net)
                    changed = list(comps)
                    if mistake == "drop":
                        del changed[index]
                    else:
                        a, b = changed[index]
                        changed[index] = (b, a)
                    broken[part] = changed
                    with pytest.raises(AssertionError):
                        GEN.verify(broken, random_trials=500)
