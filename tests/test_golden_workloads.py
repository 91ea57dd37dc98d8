"""Golden fixture of the benchmark inputs (tests/golden/workload_fields.json, written by
tests/golden/make_workload_fields.py): the seeded design fields must regenerate bit for bit, and the
boundary-condition regions must select the node sets of the benchmark specification (C0: the 33 x 33
clamped nodes of x = 0 and the 33 load nodes of x = 64, z = 16; C1: 129 load nodes; C2: the 65-node roller
edge x = 128, y = 0 and the 65 load nodes x = 0, y = 64; C4: 193 x 193 clamped and 193 load nodes)."""
import importlib.util
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _maker():
    spec = importlib.util.spec_from_file_location("make_workload_fields",
                                                  os.path.join(HERE, "golden", "make_workload_fields.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "workload_fields.json")) as fh:
        return json.load(fh)


def test_fields_regenerate_bit_for_bit(golden):
    mk = _maker()
    for name, (ne, x) in mk.fields().items():
        ref = golden["fields"][name]
        assert list(ne) == ref["grid"] and x.size == ref["size"], name
        assert mk.digest(x) == ref["sha256"], name
        assert 1e-3 <= x.min() and x.max() <= 1.0, name


def test_smoothed_fields_meet_the_recipe(golden):
    # mean 0.3 after the clamp-and-shift of the smoothed recipe; noise field within 0.3 +- 0.05
    f = golden["fields"]
    for name in ("c1_smoothed_256x128x128_seed1", "c3_smoothed_160x80x80_seed3"):
        assert abs(f[name]["mean"] - 0.3) < 1e-6
    c0 = f["c0_noise_64x32x32_seed0"]
    assert 0.25 <= c0["min"] and c0["max"] <= 0.35


def test_bc_regions_select_the_specified_nodes(golden):
    mk = _maker()
    expected = {"c0": [1089, 33], "c1": [129 * 129, 129], "c2": [65, 65], "c4": [193 * 193, 193]}
    got = mk.region_counts()
    for cfg, counts in expected.items():
        assert [r["nodes"] for r in got[cfg]] == counts, cfg
        assert [r["nodes"] for r in golden["regions"][cfg]] == counts, cfg
