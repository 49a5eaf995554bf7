"""Host-side logic of the GEMM tile autotuner (no GPU): the built-in rule,
batch buckets, and the JSON table round trip with validation."""
import json

import pytest

from khepri_amd import autotune


def test_default_rule_and_buckets():
    assert autotune.default_tile(162, 162) == 7 and autotune.default_tile(81, 162) == 5
    assert autotune.default_tile(96, 1000) == 5 and autotune.default_tile(97, 97) == 7
    assert [autotune._bucket(b) for b in (1, 15, 16, 255, 256, 65536)] == [0, 0, 1, 1, 2, 2]
    assert set(autotune.TILES) == set(range(8))


def test_table_load_validates(tmp_path):
    good = {"162x162x162:NN:b2": {"shape": [162, 162, 162, 0, 0, 1024], "tile": 7,
                                  "ms": {"mfma32x64": 1.0}, "timed_batch": 1024}}
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"meta": {}, "table": good}))
    assert autotune.load(p, install_now=False) == good
    bad = {"k": {"shape": [1, 2, 3], "tile": 0}}
    p.write_text(json.dumps({"meta": {}, "table": bad}))
    with pytest.raises(Exception):
        autotune.load(p, install_now=False)
    bad = {"k": {"shape": [1, 2, 3, 0, 0, 1], "tile": 9}}
    p.write_text(json.dumps({"meta": {}, "table": bad}))
    with pytest.raises(Exception):
        autotune.load(p, install_now=False)


def test_install_goes_through_the_library():
    """install() validates through khx_gemm_tune_set (host-only call)."""
    table = {"k": {"shape": [10, 10, 10, 0, 0, 4], "tile": 3}}
    autotune.install(table)
    autotune.clear(table)
    with pytest.raises(ValueError):
        autotune.install({"k": {"shape": [10, 10, 10, 5, 0, 4], "tile": 3}})
