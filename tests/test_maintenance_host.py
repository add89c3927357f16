"""The host index's maintenance logic (cluster splits and merges at both levels, index.hpp:1004-1685)
built with AddressSanitizer / UBSan against CPU stand-ins of the C-ABI calls it makes
(tests/maint/maint_check.cpp), checking the index's structural invariants after every operation:
each live row in exactly one slot and the row map pointing at it, deleted rows unmapped, counts and
capacities consistent, every L1 cluster in exactly one meso cluster (its mesocluster_id) holding
its centroid. No GPU is involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_maintenance_invariants_under_asan(tmp_path):
    exe = str(tmp_path / "maint_check")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "pdx_amd", "csrc", "host"), os.path.join(ROOT, "tests", "maint", "maint_check.cpp"),
           "-o", exe, "-pthread"]
    subprocess.run(cmd, check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "maintenance invariants hold" in r.stdout
    # every scenario split and merged L1 clusters; the last one split its meso cluster
    lines = [l for l in r.stdout.splitlines() if l.startswith("clusters")]
    assert len(lines) == 5
