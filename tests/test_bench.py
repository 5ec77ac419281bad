"""bench.py on the CPU: it measures one GPU and refuses more, and the table
traffic it reports follows the access counts it is given."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    assert "one GPU" in out.stdout + out.stderr
    assert not out.stdout.strip().startswith("{")  # no result line


def test_table_traffic_model():
    b = load_bench()
    assert b.tt_access_bytes({}, {})["algorithmic_gb"] == 0
    nodes = {2: 10**9, 3: 10**8, 5: 10**6}
    none = {r: (0, 0, 0, 0) for r in nodes}

    def with_(r, c):
        d = dict(none)
        d[r] = c
        return d

    base = b.tt_access_bytes(nodes, none)
    # every board of a level with a duplicate set is inserted once
    assert base["set_inserts"] == sum(nodes.values())
    # a table hit is not inserted; a probe of a deep entry costs twice a shallow one
    hit = b.tt_access_bytes(nodes, with_(3, (10**6, 10**6, 0, 0)))
    assert hit["set_inserts"] == base["set_inserts"] - 10**6
    shallow = b.tt_access_bytes(nodes, with_(3, (10**6, 0, 0, 0)))["algorithmic_gb"] - base["algorithmic_gb"]
    deep = b.tt_access_bytes(nodes, with_(5, (10**6, 0, 0, 0)))["algorithmic_gb"] - base["algorithmic_gb"]
    assert abs(deep - 2 * shallow) < 1e-12 and shallow > 0
    # rounding a random access up to a whole line never shrinks it
    full = b.tt_access_bytes(nodes, {2: (0, 0, 0, 5 * 10**8), 3: (10**7, 0, 5 * 10**6, 10**7), 5: (10**5, 10, 10**5, 0)})
    assert full["line_rounded_gb"] > full["algorithmic_gb"] > 0
