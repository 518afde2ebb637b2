"""Write ``tests/golden/kat.json``: known-answer cases of the median/rank path.

Every case restates a literal input and its expected output from one of the
reference's own tests (cited per case). Nothing here imports or runs the
reference; expected values are the literals those tests assert, or the
closed-form expression the test itself states (cited in ``source``).

Values are stored as hex bit patterns so signed zeros, infinities and NaN
payloads survive the round trip. Run from the repository root:

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "kat.json")

F32_NAN = 0x7FC00000
F64_NAN = 0x7FF8000000000000


def bits(values, dtype) -> list[str]:
    array = np.asarray(values, dtype=dtype)
    u = np.uint32 if array.dtype == np.float32 else np.uint64
    return [f"{int(b):x}" for b in array.ravel().view(u)]


def case(name, source, function, values, dtype, kwargs, expected, *, shape=None, rank=None):
    """expected: {"all": [...]} full output, or {"at": {flat_index: value}}."""
    array = np.asarray(values, dtype=dtype)
    if shape is not None:
        array = array.reshape(shape)
    entry = {
        "name": name,
        "source": source,
        "function": function,
        "dtype": np.dtype(dtype).name,
        "shape": list(array.shape),
        "input": bits(array, dtype),
        "kwargs": kwargs,
    }
    if rank is not None:
        entry["rank"] = rank
    if "all" in expected:
        entry["expected_all"] = bits(np.asarray(expected["all"], dtype=dtype), dtype)
    elif "at" in expected:
        entry["expected_at"] = {str(k): bits([v], dtype)[0] for k, v in expected["at"].items()}
    else:
        entry["expected_at"] = {str(k): f"{v:x}" for k, v in expected["at_bits"].items()}
    return entry


def cases() -> list[dict]:
    out = []
    f32, f64 = np.float32, np.float64

    # tests/test_reference.py:8-20 (nearest 1-D fixture).
    out.append(case("reference_nearest_1d", "tests/test_reference.py:8-20", "median_filter",
                    [1.0, 9.0, 2.0], f64, {"size": 3, "mode": "nearest"}, {"all": [1.0, 2.0, 2.0]}))
    # tests/test_reference.py:23-34 (constant cval 10 on 2x2).
    for dt in (f32,):
        out.append(case("reference_constant_2d", "tests/test_reference.py:23-34", "median_filter",
                        [[1.0, 2.0], [3.0, 4.0]], dt, {"size": 3, "mode": "constant", "cval": 10.0},
                        {"all": np.full((2, 2), 10.0)}))
    # tests/test_reference.py:37-47 (unfiltered axes in 3-D).
    out.append(case("reference_3d_unfiltered_axes", "tests/test_reference.py:37-47", "median_filter",
                    [[[1.0]], [[9.0]], [[2.0]]], f64,
                    {"size": [3, 1, 1], "mode": ["nearest", "reflect", "wrap"]},
                    {"all": [[[1.0]], [[2.0]], [[2.0]]]}))
    # src/window.rs:311-325: 3x3 window, modes (constant, nearest), cval -1.
    # Samples at center 0 are [-1,-1,-1, 0,0,1, 3,3,4]; rank r gives the r-th
    # sorted sample, at centre 4 the samples are 0..8.
    grid = np.arange(9.0).reshape(3, 3)
    at0 = sorted([-1.0, -1.0, -1.0, 0.0, 0.0, 1.0, 3.0, 3.0, 4.0])
    for r in range(9):
        out.append(case(f"window_rect_rank{r}", "src/window.rs:311-325", "rank_filter", grid, f64,
                        {"size": [3, 3], "mode": ["constant", "nearest"], "cval": -1.0},
                        {"at": {0: at0[r], 4: float(r)}}, rank=r))
    # src/window.rs:327-342: footprint [1,0,1,1], origin -1, reflect:
    # samples [1,3,4] at centre 2 and [4,5,4] at centre 5.
    line = np.arange(6.0)
    for r, (a, b) in enumerate(zip([1.0, 3.0, 4.0], [4.0, 4.0, 5.0])):
        out.append(case(f"window_footprint_origin_rank{r}", "src/window.rs:327-342", "rank_filter", line, f64,
                        {"footprint": [1, 0, 1, 1], "origin": -1, "mode": "reflect"},
                        {"at": {2: a, 5: b}}, rank=r))
    out.append(case("window_footprint_origin_median", "src/window.rs:327-342", "median_filter", line, f64,
                    {"footprint": [1, 0, 1, 1], "origin": -1, "mode": "reflect"}, {"at": {2: 3.0, 5: 4.0}}))
    # src/reduction.rs:137-143: rank 1 of [3, -0, 0, 1] is +0.0.
    out.append(case("reduction_rank_signed_zero", "src/reduction.rs:137-143", "rank_filter",
                    [3.0, -0.0, 0.0, 1.0], f64, {"size": 4, "mode": "nearest"}, {"at_bits": {2: 0}}, rank=1))
    # src/reduction.rs:185-194: propagate gives canonical NaN for a payload NaN.
    payload = np.array([0x7FF0000000000001], dtype=np.uint64).view(f64)[0]
    out.append(case("reduction_propagate_canonical_nan", "src/reduction.rs:185-194", "median_filter",
                    [1.0, payload, 3.0], f64, {"size": 3, "mode": "nearest"}, {"at_bits": {1: F64_NAN}}))
    # src/reduction.rs:196-219: omit {ignore_inf, min_count=2}.
    out.append(case("reduction_omit_compacts", "src/reduction.rs:196-209", "nanmedian_filter",
                    [1.0, np.nan, np.inf, 3.0], f64,
                    {"size": 4, "mode": "nearest", "ignore_inf": True, "min_count": 2}, {"at": {2: 2.0}}))
    out.append(case("reduction_omit_min_count", "src/reduction.rs:210-219", "nanmedian_filter",
                    [1.0, np.nan, np.inf], f64,
                    {"size": 3, "mode": "constant", "cval": np.nan, "ignore_inf": True, "min_count": 2},
                    {"at_bits": {1: F64_NAN}}))
    # src/reduction.rs:11-14: ignore_inf drops both +inf and -inf (the 3-D
    # case of tests/test_semantics.py:28-50 mixes -inf with ignore_inf); with
    # ignore_inf off, -inf is an ordinary lowest sample.
    out.append(case("reduction_ignore_inf_drops_negative", "src/reduction.rs:11-14", "nanmedian_filter",
                    [1.0, -np.inf, 3.0], f64, {"size": 3, "mode": "nearest", "ignore_inf": True},
                    {"at": {1: 2.0}}))
    out.append(case("reduction_negative_inf_retained", "src/reduction.rs:11-14", "nanmedian_filter",
                    [1.0, -np.inf, 3.0], f64, {"size": 3, "mode": "nearest"}, {"at": {1: 1.0}}))
    # src/reduction.rs:222-246: signed zeros, extremes, mixed infinities.
    out.append(case("reduction_signed_zero_median", "src/reduction.rs:223-230", "median_filter",
                    [0.0, -0.0, -0.0], f64, {"size": 3, "mode": "nearest"},
                    {"at_bits": {1: 0x8000000000000000}}))
    out.append(case("reduction_mixed_infinities", "src/reduction.rs:239-245", "nanmedian_filter",
                    [-np.inf, np.inf, np.nan], f64, {"size": 3, "mode": "constant", "cval": np.nan},
                    {"at_bits": {1: F64_NAN}}))
    # src/reduction.rs:248-259: f32 [1, nextafter(1)] -> f64 midpoint cast once
    # ((1 + (1 + 2^-23)) / 2 = 1 + 2^-24, a tie that rounds to even: 1.0).
    upper = np.nextafter(np.float32(1.0), np.float32(2.0))
    out.append(case("reduction_f32_cast_once", "src/reduction.rs:248-259", "nanmedian_filter",
                    [1.0, float(upper), np.nan], f32, {"size": 3, "mode": "constant", "cval": np.nan},
                    {"at_bits": {1: 0x3F800000}}))
    # tests/test_reducers_contract.py:9-20: overflow-safe midpoint of MAX and
    # its predecessor equals previous + (maximum - previous) / 2 in the dtype.
    for dt in (f32, f64):
        maximum = np.finfo(dt).max
        previous = np.nextafter(maximum, dt(0), dtype=dt)
        expected = dt(previous + (maximum - previous) / dt(2))
        out.append(case(f"contract_max_midpoint_{np.dtype(dt).name}", "tests/test_reducers_contract.py:9-20",
                        "nanmedian_filter", [maximum, previous, np.nan], dt,
                        {"size": 3, "mode": "constant", "cval": np.nan}, {"at": {1: expected}}))
        minimum = -np.finfo(dt).max
        nxt = np.nextafter(minimum, dt(0), dtype=dt)
        expected = dt(minimum + (nxt - minimum) / dt(2))
        out.append(case(f"contract_negative_extremes_{np.dtype(dt).name}",
                        "tests/test_reducers_contract.py:38-48", "nanmedian_filter", [minimum, nxt, np.nan], dt,
                        {"size": 3, "mode": "constant", "cval": np.nan}, {"at": {1: expected}}))
        out.append(case(f"contract_repeated_{np.dtype(dt).name}", "tests/test_reducers_contract.py:38-48",
                        "nanmedian_filter", [2.0, 2.0, 4.0, 4.0, np.nan], dt,
                        {"size": 5, "mode": "constant", "cval": np.nan}, {"at": {2: 3.0}}))
        for inf in (-np.inf, np.inf):
            out.append(case(f"contract_equal_infinities_{np.dtype(dt).name}_{'neg' if inf < 0 else 'pos'}",
                            "tests/test_reducers_contract.py:31-35", "nanmedian_filter", [inf, inf, np.nan], dt,
                            {"size": 3, "mode": "constant", "cval": np.nan}, {"at": {1: inf}}))
    # tests/test_reducers_contract.py:23-28: mixed infinities -> NaN, +-0 -> +0.
    out.append(case("contract_infinity_pairings", "tests/test_reducers_contract.py:23-26", "nanmedian_filter",
                    [-np.inf, np.inf, np.nan], f64, {"size": 3, "mode": "constant", "cval": np.nan},
                    {"at_bits": {1: F64_NAN}}))
    out.append(case("contract_signed_zero_midpoint", "tests/test_reducers_contract.py:26-28", "nanmedian_filter",
                    [-0.0, 0.0, np.nan], f64, {"size": 3, "mode": "constant", "cval": np.nan},
                    {"at_bits": {1: 0}}))
    # tests/test_semantics.py:54-61: generated NaN and signed zero.
    out.append(case("semantics_generated_nan", "tests/test_semantics.py:54-57", "nanmedian_filter", [np.nan],
                    f64, {"size": 1, "min_count": 1}, {"at_bits": {0: F64_NAN}}))
    out.append(case("semantics_signed_zero", "tests/test_semantics.py:59-60", "median_filter",
                    [-0.0, -0.0, 0.0], f64, {"size": 3, "mode": "nearest"}, {"at_bits": {1: 0x8000000000000000}}))
    # tests/test_semantics.py:63-76: cval participates in the policy.
    out.append(case("semantics_cval_retained", "tests/test_semantics.py:64-66", "nanmedian_filter", [np.nan],
                    f64, {"size": 3, "mode": "constant", "cval": 2.0, "min_count": 2}, {"at": {0: 2.0}}))
    out.append(case("semantics_cval_omitted", "tests/test_semantics.py:67-75", "nanmedian_filter", [np.nan],
                    f64, {"size": 3, "mode": "constant", "cval": np.inf, "ignore_inf": True, "min_count": 1},
                    {"at_bits": {0: F64_NAN}}))
    # tests/test_semantics.py:103-123: canonical NaN for both dtypes.
    for dt, nan in ((f32, F32_NAN), (f64, F64_NAN)):
        name = np.dtype(dt).name
        out.append(case(f"semantics_canonical_propagated_{name}", "tests/test_semantics.py:111-122",
                        "median_filter", [np.nan], dt, {"size": 1}, {"at_bits": {0: nan}}))
        out.append(case(f"semantics_canonical_below_min_count_{name}", "tests/test_semantics.py:112-123",
                        "nanmedian_filter", [1.0, np.nan, np.nan], dt,
                        {"size": 3, "mode": "nearest", "min_count": 2}, {"at_bits": {1: nan}}))
    # tests/test_api.py:108-115: integer-index size/min_count.
    out.append(case("api_min_count_two", "tests/test_api.py:108-115", "nanmedian_filter", [1.0, np.nan, 3.0],
                    f64, {"size": 3, "min_count": 2}, {"all": [1.0, 2.0, 3.0]}))
    # src/plan.rs:206-222 (doc example): skip NaN over a 1x3 window.
    out.append(case("plan_doc_example", "src/plan.rs:206-222", "nanmedian_filter",
                    [1.0, 2.0, np.nan, 4.0, 5.0, 6.0], f64, {"size": [1, 3]},
                    {"all": [[1.0, 1.5, 2.0], [4.0, 5.0, 6.0]]}, shape=(2, 3)))
    # tests/test_footprint.py:279-292: even footprint [1, 1] gives the midpoint
    # of (x[i-1], x[i]); rank 1 gives the upper one.
    values = np.array([0, 4, 1, 7, 2, 9, 3], dtype=f64)
    previous = np.concatenate([values[:1], values[:-1]])
    for dt in (f32, f64):
        v = values.astype(dt)
        p = previous.astype(dt)
        out.append(case(f"footprint_even_midpoint_{np.dtype(dt).name}", "tests/test_footprint.py:279-292",
                        "median_filter", v, dt, {"footprint": [1, 1], "mode": "nearest"}, {"all": (p + v) / 2}))
        out.append(case(f"footprint_even_upper_rank_{np.dtype(dt).name}", "tests/test_footprint.py:289-292",
                        "rank_filter", v, dt, {"footprint": [1, 1], "mode": "nearest"},
                        {"all": np.maximum(p, v)}, rank=1))
    # src/boundary.rs:77-117: literal map tables via rank 0 of a size-1 window
    # shifted past the edges is not expressible; the tables are stored raw.
    return out


def fixpix_case(name, source, values, mask, dtype, kwargs, expected):
    array = np.asarray(values, dtype=dtype)
    return {
        "name": name,
        "source": source,
        "dtype": np.dtype(dtype).name,
        "shape": list(array.shape),
        "input": bits(array, dtype),
        "mask": [int(v) for v in np.asarray(mask, dtype=bool).ravel()],
        "kwargs": kwargs,
        "expected_all": bits(np.asarray(expected, dtype=dtype), dtype),
    }


def fixpix_cases() -> list[dict]:
    """Literal answers of fixpix(method="median") in the reference's doc example and tests."""
    out = []
    nan, inf = np.nan, np.inf
    for dt in (np.float32, np.float64):
        tag = np.dtype(dt).name
        out.append(fixpix_case(f"doc_example_{tag}", "src/repair.rs:99-104", [1.0, 9.0, 99.0, 3.0, 5.0],
                               [0, 0, 1, 0, 0], dt, {"size": 3, "cval": 0.0}, [1.0, 9.0, 6.0, 3.0, 5.0]))
        out.append(fixpix_case(f"never_uses_repaired_{tag}", "tests/test_fixpix_methods.py:146-168",
                               [0.0, 100.0, 200.0, 10.0], [0, 1, 1, 0], dt, {"size": 3, "mode": "reflect"},
                               [0.0, 0.0, 10.0, 10.0]))
        out.append(fixpix_case(f"edge_cval_{tag}", "tests/test_fixpix_methods.py:171-196", [99.0, 5.0], [1, 0], dt,
                               {"size": 3, "mode": "constant", "cval": 11.0}, [8.0, 5.0]))
        out.append(fixpix_case(f"inf_retained_{tag}", "tests/test_fixpix_methods.py:198-223", [inf, 99.0, 2.0],
                               [0, 1, 0], dt, {"size": 3, "mode": "reflect"}, [inf, inf, 2.0]))
        out.append(fixpix_case(f"inf_omitted_{tag}", "tests/test_fixpix_methods.py:198-225", [inf, 99.0, 2.0],
                               [0, 1, 0], dt, {"size": 3, "mode": "reflect", "ignore_inf": True}, [inf, 2.0, 2.0]))
        out.append(fixpix_case(f"min_count_enough_{tag}", "tests/test_fixpix_methods.py:228-252", [99.0, 7.0],
                               [1, 0], dt, {"size": 5, "mode": "nearest", "min_count": 2}, [7.0, 7.0]))
        out.append(fixpix_case(f"min_count_short_{tag}", "tests/test_fixpix_methods.py:228-252", [99.0, 7.0],
                               [1, 0], dt, {"size": 5, "mode": "nearest", "min_count": 3}, [nan, 7.0]))
        for size in (1, 3, 5):
            out.append(fixpix_case(f"fully_masked_size{size}_{tag}", "tests/test_fixpix_methods.py:433-439",
                                   np.arange(6.0).reshape(2, 3), np.ones((2, 3)), dt, {"size": size},
                                   np.full((2, 3), nan)))
    return out


BOUNDARY_TABLES = {
    "source": "src/boundary.rs:77-187",
    "length4": {
        "coordinates": [-5, -4, -3, -2, -1, 0, 1, 2, 3, 4, 5, 6, 7, 8],
        "reflect": [3, 3, 2, 1, 0, 0, 1, 2, 3, 3, 2, 1, 0, 0],
        "mirror": [1, 2, 3, 2, 1, 0, 1, 2, 3, 2, 1, 0, 1, 2],
        "nearest": [0, 0, 0, 0, 0, 0, 1, 2, 3, 3, 3, 3, 3, 3],
        "wrap": [3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0],
    },
    "constant_length2": {"coordinates": [-2, -1, 0, 1, 2, 3], "mapped": [None, None, 0, 1, None, None]},
    "length2": {
        "coordinates": [-5, -4, -3, -2, -1, 0, 1, 2, 3, 4, 5, 6],
        "reflect": [0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1],
        "mirror": [1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0],
        "nearest": [0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1],
        "wrap": [1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0],
    },
    "many_periods": [
        {"index": 1000000003, "length": 4, "mode": "reflect", "mapped": 3},
        {"index": -1000000003, "length": 4, "mode": "mirror", "mapped": 1},
        {"index": -1000000003, "length": 4, "mode": "wrap", "mapped": 1},
    ],
}


def main() -> None:
    data = {
        "about": "Known answers restated from the ndfiltrs reference tests (see each 'source').",
        "boundary": BOUNDARY_TABLES,
        "cases": cases(),
        "fixpix_median": fixpix_cases(),
    }
    with open(OUT, "w") as handle:
        json.dump(data, handle, indent=1)
        handle.write("\n")
    print(f"wrote {len(data['cases'])} cases to {OUT}")


if __name__ == "__main__":
    main()
