"""The autotuner's table (host only): bucket selection by family, skew and
floor(log2 W), the bitmap-length rule, and loading from $WTC_TUNING_FILE or
the committed table beside the library (a malformed file leaves the
built-in rows in force)."""
import json
import os
import subprocess
import sys

import wtc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wedge-parallel-triangle-counting_amd")

TABLE = """# family skew log2w spread a_div a_cap kernel waves ways split_waves
alone any 0 2 0 0 walk 4 4 8
alone any 10 5 0 0 staged 2 8 8
alone skew 10 7 64 4096 split 8 2 16   # skewed graphs from 2^10 wedges
alone any 20 9 0 0 walk 1 1 4
batch any 0 3 0 0 walk 2 4 8
batch flat 12 6 0 0 walk 4 4 8
"""


def _probe(env_file, queries):
    """Runs the autotuner in a fresh process (the table is read once per
    process) and returns its answers."""
    code = (
        "import json, sys; sys.path.insert(0, %r); import wtc\n"
        "out = [wtc.tuning_source()]\n"
        "for n, m, W, b in %r:\n"
        "    out.append(wtc.autotune_shape(n, m, W, b).as_dict())\n"
        "print(json.dumps(out))\n" % (PKG, queries))
    env = dict(os.environ)
    env["WTC_TUNING_FILE"] = env_file
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_table_rows_are_selected_by_bucket(tmp_path):
    f = tmp_path / "t.tsv"
    f.write_text(TABLE)
    q = [(100, 50, 500, 0),          # 2^8 wedges: first row
         (5000, 1000, 5000, 0),      # 2^12, W/m = 5: flat -> staged row
         (5000, 100, 4096, 0),       # 2^12, W/m = 40: skewed -> split row, a = n/64 -> 96
         (10 ** 6, 10 ** 4, 2 ** 20, 0),  # 2^20: the `any` row above the skewed one
         (300, 1000, 100, 1),        # batch, 2^6
         (300, 1000, 5000, 1),       # batch, 2^12, flat
         (300, 10, 5000, 1),         # batch, 2^12, skewed: no skewed row -> the 2^0 row
         (10 ** 7, 100, 2 ** 15, 0)]  # skewed, a capped at 4096
    src, *ans = _probe(str(f), q)
    assert src == str(f)
    assert ans[0]["spread"] == 2 and ans[0]["kernel"] == wtc.KERNEL_WALK
    assert ans[1]["spread"] == 5 and ans[1]["kernel"] == wtc.KERNEL_STAGED
    assert ans[1]["waves_per_block"] == 2 and ans[1]["probe_ways"] == 8
    assert ans[2]["spread"] == 7 and ans[2]["kernel"] == wtc.KERNEL_SPLIT and ans[2]["split_waves"] == 16
    assert ans[2]["adjacency_matrix_len"] == 96  # ceil(5000/64) = 79 -> 96
    assert ans[3]["spread"] == 9 and ans[3]["adjacency_matrix_len"] == 0
    assert ans[4]["spread"] == 3
    assert ans[5]["spread"] == 6
    assert ans[6]["spread"] == 3
    assert ans[7]["adjacency_matrix_len"] == 4096


def test_malformed_table_keeps_builtin_rows(tmp_path):
    f = tmp_path / "bad.tsv"
    f.write_text("alone any 0 2 0 0 walk 3 4 8\n")  # 3 waves per block is not a block shape
    src, ans = _probe(str(f), [(16775427, 260385033, 41737375097, 0)])
    assert src == "built-in"
    assert ans["spread"] >= 1


def test_committed_table_loads_and_matches_autotune():
    """The library reads the committed table beside it (when present), and
    autotune() is the (spread, a) of autotune_shape()."""
    committed = os.path.join(PKG, "tuning", "mi355x.tsv")
    src = wtc.tuning_source()
    if os.path.exists(committed) and "WTC_TUNING_FILE" not in os.environ:
        assert os.path.realpath(src) == os.path.realpath(committed)
    for n, m, W in ((1000, 5000, 30000), (1048465, 15701962, 1226291937), (16775427, 260385033, 41737375097)):
        t = wtc.autotune_shape(n, m, W)
        assert wtc.autotune(n, m, W) == (t.spread, t.adjacency_matrix_len)
        assert t.adjacency_matrix_len % 32 == 0 and t.adjacency_matrix_len <= ((n + 31) // 32) * 32
