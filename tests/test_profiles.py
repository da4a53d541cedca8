"""The rocprofv3 summaries under profiles/ carry, in the form bench.py reads
them (`key: value` lines), every per-node figure its roofline object quotes,
and each figure lies in its physical range.  CPU only."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402

INSTRUCTION_KEYS = ("valu_wave_instructions_per_node", "valu_busy", "active_lane_fraction",
                    "lds_bank_conflict_rate")
FRACTIONS = ("valu_busy", "active_lane_fraction", "lds_bank_conflict_rate")


@pytest.mark.parametrize("config,kind", [("c2", "manhattan"), ("c3", "manhattan"), ("c4", "pdb")])
def test_instruction_summaries(config, kind):
    for key in INSTRUCTION_KEYS:
        v = bench.profile_figure(config, kind, "instructions", key)
        assert v is not None, (config, key)
        if key in FRACTIONS:
            assert 0.0 < v < 1.0, (config, key, v)
        else:
            assert 0.5 < v < 50.0, (config, key, v)


def test_pdb_traffic_summary():
    tb = bench.profile_figure("c4", "pdb", "traffic", "traffic_bytes_per_node")
    miss = bench.profile_figure("c4", "pdb", "traffic", "l2_misses_per_node")
    hits = bench.profile_figure("c4", "pdb", "traffic", "l2_hits_per_node")
    assert tb is not None and miss is not None and hits is not None
    # a generated node makes at least one table read; misses are 64 B lines
    assert 1.0 <= miss + hits < 4.0
    assert miss * 64 <= tb < 256


def test_each_summary_names_its_command():
    for name in sorted(os.listdir(os.path.join(REPO, "profiles"))):
        text = open(os.path.join(REPO, "profiles", name)).read()
        assert "rocprofv3" in text, name
        assert "bench.py" in text, name
