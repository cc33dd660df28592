"""CPU checks of bench.py's contract: a request for more than one GPU ends
with a JSON error line before anything touches a GPU, and every measured
traffic figure the benchmark can report (keyed to the kernel source it was
profiled with) cites a committed profile that names the same source."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench_module():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_json_line():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1 and "one GPU" in json.loads(lines[0])["error"]


def test_traffic_figures_cite_profiles_of_their_source():
    b = _bench_module()
    spmm_id, tiles_id = b.kernel_source_id(), b.kernel_source_id(b.TILES_SOURCE)
    current = 0
    for table, sid in ((b.MEASURED_TRAFFIC, spmm_id), (b.CONFIG_TRAFFIC, spmm_id), (b.TILE_TRAFFIC, tiles_id)):
        for key, (nbytes, profile, source) in table.items():
            text = open(os.path.join(ROOT, profile)).read()
            assert source in text, (key, profile, source)
            assert nbytes > 0
            current += source == sid
    # the headline's figure is live for the kernel source in the tree
    key = ("rmat", 18, 256, 1, "auto", "locality")
    assert b.MEASURED_TRAFFIC[key][2] == spmm_id
    assert current >= 4


def test_profile_commands_use_existing_bench_flags():
    """The commands in the committed profile summaries run committed files:
    every long option in their code blocks is one bench.py defines or one
    of rocprofv3's."""
    import glob
    import re
    flags = set(re.findall(r'add_argument\("(--[a-z0-9-]+)"', open(os.path.join(ROOT, "bench.py")).read()))
    rocprof = {"--kernel-trace", "--stats", "--output-format", "--pmc"}
    files = sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md")))
    assert files
    for f in files:
        text = open(f).read()
        used = set()
        for block in re.findall(r"```(.*?)```", text, flags=re.S):
            assert "bench.py" in block, f
            used |= set(re.findall(r"(?<![\w-])(--[a-z0-9-]+)", block))
        assert not used - flags - rocprof, (f, sorted(used - flags - rocprof))
