"""Generate the comparator networks of the stack engine (ndf_stack_networks.inc).

The stack engine computes a median / rank over a box window of K0 slices
along the slowest window axis, each slice an SH x SW box (SH = 1 for a 2-D
window, whose slices are row segments). Per output it merges K0 sorted
slices, so the networks here are:

  col    sorts one column of SH keys (wires 0 .. SH-1);
  slice  merges SW sorted columns (column c on wires c*SH .. c*SH+SH-1) into
         one sorted slice; `slice_out[i]` is the wire holding rank i;
  half   merges K0 sorted slices (slice s, in rank order, on wires
         s*L .. s*L+L-1 with L = SH*SW) and keeps only the comparators that
         ranks 0 .. W/2 depend on (W = K0*L); `half_out[r]` is rank r's wire.
         Omitted samples carry the all-ones key and sort last, so a median's
         two middle ranks never exceed W/2;
  full   the same merge with every rank (rank and percentile statistics).

Merges are Batcher's odd-even merge (K. E. Batcher, "Sorting networks and
their applications", 1968): merge the odd- and the even-indexed elements of
both lists recursively, then compare neighbours once; it works for lists of
any lengths. A merge tree is searched over its shapes for the fewest
comparators. `verify()` proves every network by the 0-1 principle: a
comparator network sorts (or selects a rank of) every input of its class iff
it does so for every 0-1 input of that class, and the class here is "lists
already sorted", whose 0-1 inputs are 0^a 1^(n-a) per list. Where the 0-1
inputs of a whole network are too many, every two-list merge it is built
from is proven instead (a tree of correct merges is correct, and dropping
comparators outside the cone of the kept ranks leaves those ranks unchanged)
and the whole network is checked on random inputs.

Run `python gen_stack_networks.py` in this directory to rewrite the header.
"""

from __future__ import annotations

import itertools
import os
import sys

import numpy as np

# (K0, SH, SW): slices along the slowest window axis, slice rows, slice columns.
CONFIGS = [
    (3, 1, 3),
    (5, 1, 5),
    (7, 1, 7),
    (9, 1, 9),
    (3, 3, 3),
    (3, 5, 5),
    (3, 7, 7),
]

# Live wire counts of the network engine (one window of up to LIVE taps,
# padded with all-ones keys, sorted per output).
SORT_LIVES = [9, 16, 25, 32, 49, 64, 72, 96, 128, 160]

# (core, extra) wire classes of the network engine's two-row step: a
# footprint whose windows at rows y and y + 1 share `core` taps and differ
# by `extra` taps each.
PAIR_CLASSES = [(16, 8), (32, 12), (48, 20), (64, 24), (96, 32), (128, 32)]

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "ndf_stack_networks.inc")


def merge(xs: list[int], ys: list[int]) -> tuple[list[tuple[int, int]], list[int]]:
    """Batcher's odd-even merge of the sorted wire lists xs and ys (any
    lengths). Returns the comparators (min goes to the first wire) and the
    wires of the merged list in rank order."""
    m, n = len(xs), len(ys)
    if m == 0:
        return [], list(ys)
    if n == 0:
        return [], list(xs)
    if m == 1 and n == 1:
        return [(xs[0], ys[0])], [xs[0], ys[0]]
    c_odd, v = merge(xs[0::2], ys[0::2])
    c_even, w = merge(xs[1::2], ys[1::2])
    comps = c_odd + c_even
    out = [v[0]]
    i = 0
    while i < len(w) or i + 1 < len(v):
        if i < len(w) and i + 1 < len(v):
            comps.append((w[i], v[i + 1]))
            out += [w[i], v[i + 1]]
        elif i < len(w):
            out.append(w[i])
        else:
            out.append(v[i + 1])
        i += 1
    return comps, out


def merge_tree(lists: list[list[int]], shape) -> tuple[list[tuple[int, int]], list[int]]:
    """Merge sorted wire lists along a tree: `shape` is an index or a pair of
    shapes."""
    if isinstance(shape, int):
        return [], lists[shape]
    ca, a = merge_tree(lists, shape[0])
    cb, b = merge_tree(lists, shape[1])
    cm, out = merge(a, b)
    return ca + cb + cm, out


def tree_shapes(items: tuple[int, ...]):
    """Every binary tree over the items in this order (Catalan many; the
    configurations keep this small)."""
    if len(items) == 1:
        yield items[0]
        return
    for split in range(1, len(items)):
        for left in tree_shapes(items[:split]):
            for right in tree_shapes(items[split:]):
                yield (left, right)


def cone(comps: list[tuple[int, int]], outputs: list[int]) -> list[tuple[int, int]]:
    """The comparators that the given output wires depend on, in order."""
    need = set(outputs)
    kept = []
    for a, b in reversed(comps):
        if a in need or b in need:
            kept.append((a, b))
            need.add(a)
            need.add(b)
    kept.reverse()
    return kept


def sort_network(n: int) -> tuple[list[tuple[int, int]], list[int]]:
    """Sorting network on wires 0..n-1 as a merge tree of single wires (the
    cheapest tree shape)."""
    best = None
    for shape in tree_shapes(tuple(range(n))) if n <= 9 else [balanced(tuple(range(n)))]:
        comps, out = merge_tree([[i] for i in range(n)], shape)
        if best is None or len(comps) < len(best[0]):
            best = (comps, out)
    return best


def balanced(items: tuple[int, ...]):
    if len(items) == 1:
        return items[0]
    h = (len(items) + 1) // 2
    return (balanced(items[:h]), balanced(items[h:]))


def build(k0: int, sh: int, sw: int) -> dict:
    length = sh * sw
    width = k0 * length
    col, col_out = sort_network(sh) if sh > 1 else ([], [0])
    # Slice: merge SW sorted columns; columns are symmetric, so the balanced
    # tree and the left-deep tree are the candidates.
    columns = [list(range(c * sh, (c + 1) * sh)) for c in range(sw)]
    best = None
    for shape in tree_shapes(tuple(range(sw))) if sw <= 9 else [balanced(tuple(range(sw)))]:
        comps, out = merge_tree(columns, shape)
        if best is None or len(comps) < len(best[0]):
            best = (comps, out)
    slice_comps, slice_out = best
    # Output merges over the K0 slices: the fewest comparators for the lower
    # half (the median's ranks), and for the full order.
    slices = [list(range(s * length, (s + 1) * length)) for s in range(k0)]
    best_half = best_full = None
    for shape in tree_shapes(tuple(range(k0))):
        comps, out = merge_tree(slices, shape)
        half = cone(comps, out[: width // 2 + 1])
        if best_half is None or len(half) < len(best_half[0]):
            best_half = (half, out[: width // 2 + 1])
        if best_full is None or len(comps) < len(best_full[0]):
            best_full = (comps, out)
    net = dict(k0=k0, sh=sh, sw=sw, length=length, width=width, col=col, col_out=col_out,
               slice=slice_comps, slice_out=slice_out, half=best_half[0], half_out=best_half[1],
               full=best_full[0], full_out=best_full[1])
    if sh == 1:
        net.update(build_pair(k0, length))
    return net


def build_pair(k0: int, length: int) -> dict:
    """Networks of a 2-D window's two-output step: outputs p and p + 1 share
    the k0 - 1 row slices p + 1 .. p + k0 - 1 (the core). `core` merges them
    (slice s on wires s*L ..), keeping what the final merges read; `fin_half`
    / `fin_full` merge the sorted core (ranks on wires 0 .. C-1) with one
    more sorted slice (wires C .. W-1) into the lower half / the full order.
    `core_out[r]` is the wire of the core's rank r."""
    width = k0 * length
    core_width = (k0 - 1) * length
    fin, fin_out = merge(list(range(core_width)), list(range(core_width, width)))
    fin_half = cone(fin, fin_out[: width // 2 + 1])
    used_half = {w for pair in fin_half for w in pair} | set(fin_out[: width // 2 + 1])
    slices = [list(range(s * length, (s + 1) * length)) for s in range(k0 - 1)]
    best = None
    for shape in tree_shapes(tuple(range(k0 - 1))):
        comps, out = merge_tree(slices, shape)
        if best is None or len(comps) < len(best[0]):
            best = (comps, out)
    core, core_out = best
    # The median's core keeps only the ranks its final merge reads.
    need = sorted(r for r in range(core_width) if r in used_half)
    core_half = cone(core, [core_out[r] for r in need])
    return dict(core_width=core_width, core=core, core_half=core_half, core_out=core_out,
                fin_half=fin_half, fin_half_out=fin_out[: width // 2 + 1], fin_full=fin, fin_full_out=fin_out)


def build_sort(live: int) -> dict:
    """Sorting network of the network engine: a balanced merge tree of single
    wires; `half` keeps the comparators ranks 0 .. live/2 depend on."""
    comps, out = merge_tree([[i] for i in range(live)], balanced(tuple(range(live))))
    half_ranks = out[: live // 2 + 1]
    return dict(live=live, full=comps, full_out=out, half=cone(comps, half_ranks), half_out=half_ranks)


def build_pair_sort(core: int, extra: int) -> dict:
    """Networks of the network engine's two-row step. The core (the taps
    both windows share, padded with all-ones keys to `core` wires) and each
    window's extra taps (`extra` wires) are sorted by merge-tree sorts; `fin`
    merges the sorted core (wires 0 .. core-1) with one sorted extra (wires
    core ..). The median's networks keep only what ranks 0 .. W/2 of the
    final merge depend on (W = core + extra)."""
    width = core + extra
    fin, fin_out = merge(list(range(core)), list(range(core, width)))
    fin_half = cone(fin, fin_out[: width // 2 + 1])
    used = {w for pair in fin_half for w in pair} | set(fin_out[: width // 2 + 1])
    core_net = build_sort(core)
    extra_net = build_sort(extra)
    core_need = [core_net["full_out"][r] for r in range(core) if r in used]
    extra_need = [extra_net["full_out"][r] for r in range(extra) if core + r in used]
    return dict(core=core, extra=extra, width=width,
                core_full=core_net["full"], core_half=cone(core_net["full"], core_need), core_out=core_net["full_out"],
                extra_full=extra_net["full"], extra_half=cone(extra_net["full"], extra_need),
                extra_out=extra_net["full_out"],
                fin_full=fin, fin_full_out=fin_out, fin_half=fin_half, fin_half_out=fin_out[: width // 2 + 1])


def verify_pair_sort(net: dict, random_trials: int = 3000, seed: int = 5) -> None:
    """The two-row networks: every merge they are built from on every 0-1
    input, and the pipeline (sort core, sort extra, merge) on random keys
    with ties and sentinels against np.sort."""
    core, extra, width = net["core"], net["extra"], net["width"]
    verify_merge(core, extra)
    for n in (core, extra):
        for m, k in tree_merge_sizes(tuple(range(n))):
            verify_merge(m, k)
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 6, size=(random_trials, width)).astype(np.uint32)
    keys[rng.random(keys.shape) < 0.2] = 0xFFFFFFFF
    expect = np.sort(keys, axis=1)
    for half in (False, True):
        tag = "half" if half else "full"
        c = apply(net["core_" + tag], keys[:, :core])[:, net["core_out"]]
        e = apply(net["extra_" + tag], keys[:, core:])[:, net["extra_out"]]
        got = apply(net["fin_" + tag], np.concatenate([c, e], axis=1))[:, net["fin_" + tag + "_out"]]
        want = expect[:, : width // 2 + 1] if half else expect
        assert np.array_equal(got, want), f"two-row {tag} networks ({core}, {extra}) do not sort"


def tree_merge_sizes(items: tuple[int, ...]) -> set[tuple[int, int]]:
    """The (m, n) list lengths a balanced merge tree of single wires merges."""
    if len(items) == 1:
        return set()
    h = (len(items) + 1) // 2
    return {(h, len(items) - h)} | tree_merge_sizes(items[:h]) | tree_merge_sizes(items[h:])


def verify_sort(net: dict, random_trials: int = 3000, seed: int = 1) -> None:
    """Every two-list merge of the tree on every 0-1 input, then the whole
    network (full and half) on random keys with ties and sentinels, and on
    every 0-1 input when there are few enough."""
    live = net["live"]
    for m, n in tree_merge_sizes(tuple(range(live))):
        verify_merge(m, n)
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 6, size=(random_trials, live)).astype(np.uint32)
    keys[rng.random(keys.shape) < 0.2] = 0xFFFFFFFF
    if 2**live <= EXHAUSTIVE_LIMIT:
        keys = np.concatenate([keys, np.array(list(itertools.product((0, 1), repeat=live)), dtype=np.uint32)])
    expect = np.sort(keys, axis=1)
    assert np.array_equal(apply(net["full"], keys)[:, net["full_out"]], expect), "sort network does not sort"
    assert np.array_equal(apply(net["half"], keys)[:, net["half_out"]], expect[:, : live // 2 + 1]), \
        "half sort network does not give the lower half"


def apply(comps: list[tuple[int, int]], values: np.ndarray) -> np.ndarray:
    """Apply comparators to rows of `values` (one input per row)."""
    v = values.copy()
    for a, b in comps:
        lo = np.minimum(v[:, a], v[:, b])
        hi = np.maximum(v[:, a], v[:, b])
        v[:, a] = lo
        v[:, b] = hi
    return v


def sorted_01(lists: int, length: int) -> np.ndarray:
    """Every 0-1 input made of `lists` sorted lists of `length` (rows)."""
    rows = []
    for zeros in itertools.product(range(length + 1), repeat=lists):
        row = []
        for z in zeros:
            row += [0] * z + [1] * (length - z)
        rows.append(row)
    return np.array(rows, dtype=np.uint8)


EXHAUSTIVE_LIMIT = 1 << 18


def verify_merge(m: int, n: int) -> None:
    """Batcher's merge of sorted lists of m and n keys, on every 0-1 input."""
    comps, out = merge(list(range(m)), list(range(m, m + n)))
    x = np.array([[0] * a + [1] * (m - a) + [0] * b + [1] * (n - b) for a in range(m + 1) for b in range(n + 1)],
                 dtype=np.uint8)
    assert np.array_equal(apply(comps, x)[:, out], np.sort(x, axis=1)), f"merge({m}, {n}) is wrong"


def merge_sizes(lists: int, length: int) -> set[tuple[int, int]]:
    """Every (m, n) a merge tree over `lists` lists of `length` may merge."""
    return {(a * length, b * length) for a in range(1, lists) for b in range(1, lists - a + 1)}


def verify(net: dict, random_trials: int = 2000, seed: int = 0) -> None:
    """Raise AssertionError unless every network of `net` is correct."""
    sh, sw, k0, length, width = net["sh"], net["sw"], net["k0"], net["length"], net["width"]
    for m, n in merge_sizes(sw, sh) | merge_sizes(k0, length):
        verify_merge(m, n)
    # Column sort: all 2^SH 0-1 inputs.
    if sh > 1:
        x = np.array(list(itertools.product((0, 1), repeat=sh)), dtype=np.uint8)
        y = apply(net["col"], x)[:, net["col_out"]]
        assert np.array_equal(y, np.sort(x, axis=1)), "column network does not sort"
    # Slice merge: every 0-1 input of SW sorted columns (when few enough).
    if (sh + 1) ** sw <= EXHAUSTIVE_LIMIT:
        x = sorted_01(sw, sh)
        y = apply(net["slice"], x)[:, net["slice_out"]]
        assert np.array_equal(y, np.sort(x, axis=1)), "slice network does not merge"
    # Output merges: every 0-1 input of K0 sorted slices (when few enough).
    if (length + 1) ** k0 <= EXHAUSTIVE_LIMIT:
        x = sorted_01(k0, length)
        expect = np.sort(x, axis=1)
        y = apply(net["full"], x)[:, net["full_out"]]
        assert np.array_equal(y, expect), "full network does not merge"
        y = apply(net["half"], x)[:, net["half_out"]]
        assert np.array_equal(y, expect[:, : width // 2 + 1]), "half network does not give the lower half"
    if sh == 1:
        verify_pair(net, seed)
    # The whole pipeline on random keys with ties and sentinels.
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 6, size=(random_trials, width)).astype(np.uint32)
    keys[rng.random(keys.shape) < 0.2] = 0xFFFFFFFF
    cols = keys.reshape(random_trials, k0, sw, sh)
    if sh > 1:
        flat = cols.reshape(-1, sh)
        cols = apply(net["col"], flat)[:, net["col_out"]].reshape(random_trials, k0, sw, sh)
    sl = cols.reshape(-1, length)
    sl = apply(net["slice"], sl)[:, net["slice_out"]].reshape(random_trials, width)
    expect = np.sort(keys, axis=1)
    assert np.array_equal(apply(net["full"], sl)[:, net["full_out"]], expect)
    assert np.array_equal(apply(net["half"], sl)[:, net["half_out"]], expect[:, : width // 2 + 1])


def verify_pair(net: dict, seed: int = 0, random_trials: int = 3000) -> None:
    """The two-output networks: every merge they are built from on every 0-1
    input, the core + final pipeline exhaustively over sorted 0-1 slices when
    few enough, and on random keys."""
    k0, length, width = net["k0"], net["length"], net["width"]
    core_width = net["core_width"]
    verify_merge(core_width, length)
    for m, n in merge_sizes(k0 - 1, length):
        verify_merge(m, n)

    def run(slices_keys: np.ndarray, extra: np.ndarray, half: bool) -> np.ndarray:
        core = apply(net["core_half" if half else "core"], slices_keys)[:, net["core_out"]]
        both = np.concatenate([core, extra], axis=1)
        part = "fin_half" if half else "fin_full"
        return apply(net[part], both)[:, net[part + "_out"]]

    rng = np.random.default_rng(seed + 7)
    keys = rng.integers(0, 5, size=(random_trials, width)).astype(np.uint32)
    keys[rng.random(keys.shape) < 0.2] = 0xFFFFFFFF
    keys = np.sort(keys.reshape(random_trials, k0, length), axis=2).reshape(random_trials, width)
    if (length + 1) ** k0 <= EXHAUSTIVE_LIMIT:
        keys = np.concatenate([keys, sorted_01(k0, length).astype(np.uint32)])
    expect = np.sort(keys, axis=1)
    got = run(keys[:, :core_width], keys[:, core_width:], False)
    assert np.array_equal(got, expect), "two-output full networks do not merge"
    got = run(keys[:, :core_width], keys[:, core_width:], True)
    assert np.array_equal(got, expect[:, : width // 2 + 1]), "two-output half networks miss the lower half"


def _pairs(name: str, comps: list[tuple[int, int]]) -> str:
    body = ", ".join(f"{{{a}, {b}}}" for a, b in comps) if comps else "{0, 0}"
    return f"    static constexpr uint8_t {name}[][2] = {{{body}}};\n    static constexpr int {name}_count = {len(comps)};\n"


def _list(name: str, values: list[int]) -> str:
    return f"    static constexpr uint8_t {name}[] = {{{', '.join(str(v) for v in values)}}};\n"


def render() -> str:
    out = [
        "// Comparator networks of the stack engine. Generated by gen_stack_networks.py",
        "// (same directory); edit that script, not this file.",
        "// Net<K0, SH, SW>: K0 slices of SH x SW keys; W = K0 * SH * SW taps.",
        "#pragma once",
        "",
        "#include <stdint.h>",
        "",
        "namespace ndf {",
        "namespace stack_net {",
        "",
        "template <int K0, int SH, int SW>",
        "struct Net;",
        "",
        "// Sort<LIVE>: sorting network of the network engine on LIVE wires (a",
        "// window's taps padded with all-ones keys).",
        "template <int LIVE>",
        "struct Sort;",
        "",
        "// PairSort<C, E>: the network engine's two-row step (C shared taps, E",
        "// taps per window beyond them).",
        "template <int C, int E>",
        "struct PairSort;",
        "",
    ]
    for live in SORT_LIVES:
        net = build_sort(live)
        out.append(f"template <>\nstruct Sort<{live}> {{")
        out.append(f"    static constexpr int W = {live}, H = {live // 2 + 1};")
        text = _pairs("half", net["half"]) + _list("half_out", net["half_out"])
        text += _pairs("full", net["full"]) + _list("full_out", net["full_out"])
        out.append(text.rstrip("\n"))
        out.append("};\n")
    for core, extra in PAIR_CLASSES:
        net = build_pair_sort(core, extra)
        out.append(f"template <>\nstruct PairSort<{core}, {extra}> {{")
        out.append(f"    static constexpr int C = {core}, E = {extra}, W = {core + extra}, H = {(core + extra) // 2 + 1};")
        text = ""
        for part in ("core", "extra", "fin"):
            text += _pairs(part + "_full", net[part + "_full"]) + _pairs(part + "_half", net[part + "_half"])
        text += _list("core_out", net["core_out"]) + _list("extra_out", net["extra_out"])
        text += _list("fin_full_out", net["fin_full_out"]) + _list("fin_half_out", net["fin_half_out"])
        out.append(text.rstrip("\n"))
        out.append("};\n")
    for k0, sh, sw in CONFIGS:
        net = build(k0, sh, sw)
        out.append(f"template <>\nstruct Net<{k0}, {sh}, {sw}> {{")
        out.append(f"    static constexpr int L = {net['length']}, W = {net['width']}, H = {net['width'] // 2 + 1};")
        text = _pairs("col", net["col"]) + _list("col_out", net["col_out"])
        text += _pairs("slice", net["slice"]) + _list("slice_out", net["slice_out"])
        text += _pairs("half", net["half"]) + _list("half_out", net["half_out"])
        text += _pairs("full", net["full"]) + _list("full_out", net["full_out"])
        if sh == 1:
            text += f"    static constexpr int C = {net['core_width']};\n"
            text += _pairs("core", net["core"]) + _pairs("core_half", net["core_half"])
            text += _list("core_out", net["core_out"])
            text += _pairs("fin_half", net["fin_half"]) + _list("fin_half_out", net["fin_half_out"])
            text += _pairs("fin_full", net["fin_full"]) + _list("fin_full_out", net["fin_full_out"])
        out.append(text.rstrip("\n"))
        out.append("};\n")
    out += ["}  // namespace stack_net", "}  // namespace ndf", ""]
    return "\n".join(out)


def main() -> None:
    for live in SORT_LIVES:
        net = build_sort(live)
        verify_sort(net)
        print(f"Sort<{live}>: half={len(net['half'])} full={len(net['full'])}", file=sys.stderr)
    for core, extra in PAIR_CLASSES:
        net = build_pair_sort(core, extra)
        verify_pair_sort(net)
        print(f"PairSort<{core}, {extra}>: core half={len(net['core_half'])} extra half={len(net['extra_half'])} "
              f"fin half={len(net['fin_half'])} per output {(len(net['core_half']) + 2 * len(net['extra_half']) + 2 * len(net['fin_half'])) / 2}",
              file=sys.stderr)
    for cfg in CONFIGS:
        net = build(*cfg)
        verify(net)
        extra = ""
        if cfg[1] == 1:
            extra = (f" pair: core={len(net['core'])} core_half={len(net['core_half'])} "
                     f"fin_half={len(net['fin_half'])} fin_full={len(net['fin_full'])}")
        print(f"Net{cfg}: W={net['width']} col={len(net['col'])} slice={len(net['slice'])} "
              f"half={len(net['half'])} full={len(net['full'])}{extra}", file=sys.stderr)
    with open(HEADER, "w") as f:
        f.write(render())


if __name__ == "__main__":
    main()
