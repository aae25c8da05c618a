"""The committed tuning table (golaypair_amd/tuning.tsv, written by
python -m golaypair_amd.autotune) is well formed: "key value" lines with the
keys the library reads (golaypair_amd/csrc/tuning.hip)."""
import os
import re

import golaypair_amd as g

TABLE = os.path.join(os.path.dirname(g.__file__), "tuning.tsv")
KEYS = [r"unc_waves/\d+", r"unc_leaf_budget/\d+", r"unc_counter_rounds/\d+", r"unc_counter_streams/\d+", r"gen_ranks_per_lane/\d+/\d+"]


def test_tuning_table_format():
    assert os.path.exists(TABLE)
    entries = 0
    for line in open(TABLE):
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        key, value = line.split()
        assert any(re.fullmatch(p, key) for p in KEYS), key
        assert int(value) >= 0 if key.startswith("unc_counter_rounds/") else int(value) > 0
        entries += 1
    head = open(TABLE).read()
    assert "golaypair_amd.autotune" in head
