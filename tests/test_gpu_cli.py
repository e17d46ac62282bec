"""The tc / tc_big_graphs command lines end to end on the GPU: same flags and
results table as the reference's main (tc.cu:816-918)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wedge-parallel-triangle-counting_amd")
HEADER = ["graph", "n", "m", "s", "a", "triangles", "prepro (s)", "GPU copy (s)", "GPU exec (s)",
          "GPU total (s)", "CPU+GPU (s)"]


@pytest.mark.parametrize("binary", ["tc", "tc_big_graphs"])
def test_cli_table(golden, binary):
    for fx in golden[:6]:
        flag = "-m" if fx["format"] == "mm" else "-e"
        args = [os.path.join(PKG, binary), flag, fx["path"], "-s", "5", "-a", "64", "-l", "3"]
        if fx["zero_indexed"]:
            args.append("-z")
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = r.stdout.strip().splitlines()
        assert len(lines) == 1 + 3  # header + loop count rows (warm-up row dropped)
        assert re.split(r"\s{2,}", lines[0].strip()) == HEADER
        for row in lines[1:]:
            cols = row.split()
            assert cols[0] == fx["path"]
            n, m, s, a, tri = (int(x) for x in cols[1:6])
            assert (n, m, s, a, tri) == (fx["n"], fx["m"], 5, 64, fx["triangles"])
            copy, exe, total = (float(x) for x in cols[7:10])
            assert abs(copy + exe - total) < 2e-6
            assert float(cols[10]) >= total - 1e-6


def test_cli_csr_cache(golden, tmp_path):
    """WTC_CSR_CACHE: the first run writes the oriented CSR file, the second
    loads it; a changed input (new mtime) rebuilds it.  Same table either way."""
    fx = next(f for f in golden if f["format"] != "mm")
    src = tmp_path / "g.txt"
    src.write_bytes(open(fx["path"], "rb").read())
    cache = tmp_path / "g.csr"
    env = dict(os.environ, WTC_CSR_CACHE=str(cache))
    args = [os.path.join(PKG, "tc"), "-e", str(src), "-s", "3", "-a", "32", "-l", "1"]
    if fx["zero_indexed"]:
        args.append("-z")

    def run():
        r = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode == 0, r.stderr
        return [int(x) for x in r.stdout.strip().splitlines()[1].split()[1:6]]

    first = run()
    assert cache.exists()
    stamp = cache.stat().st_mtime_ns
    assert run() == first == [fx["n"], fx["m"], 3, 32, fx["triangles"]]
    assert cache.stat().st_mtime_ns == stamp  # loaded, not rewritten
    os.utime(src, ns=(stamp + 10 ** 9, stamp + 10 ** 9))
    assert run() == first
    assert cache.stat().st_mtime_ns != stamp  # key changed: rebuilt and rewritten


STYLED_HEADER = ["graph", "n", "m", "a", "wedge style", "triangles", "prepro (s)", "GPU copy (s)",
                 "GPU exec (s)", "GPU total (s)", "CPU+GPU (s)"]


@pytest.mark.parametrize("binary", ["tc_arrow", "tc_vertex", "tc_edge"])
def test_experiment_cli_tables(golden, binary):
    """The experiments' command lines: tc_arrow takes tc's flags and table;
    tc_vertex / tc_edge take -w <style> instead of -s and print the style."""
    styled = binary != "tc_arrow"
    for fx in golden[:5]:
        flag = "-m" if fx["format"] == "mm" else "-e"
        for style in ((0, 1, 2) if styled else (None,)):
            args = [os.path.join(PKG, binary), flag, fx["path"], "-a", "64", "-l", "2"]
            args += ["-w", str(style)] if styled else ["-s", "5"]
            if fx["zero_indexed"]:
                args.append("-z")
            r = subprocess.run(args, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            lines = r.stdout.strip().splitlines()
            assert len(lines) == 1 + 2
            assert re.split(r"\s{2,}", lines[0].strip()) == (STYLED_HEADER if styled else HEADER)
            for row in lines[1:]:
                cols = row.split()
                if styled:
                    assert [int(x) for x in cols[1:4]] == [fx["n"], fx["m"], 64]
                    assert cols[4] == ("outgoing", "arrow", "mixed")[style]
                    assert int(cols[5]) == fx["triangles"], (binary, fx["file"], style)
                else:
                    assert [int(x) for x in cols[1:6]] == [fx["n"], fx["m"], 5, 64, fx["triangles"]]
