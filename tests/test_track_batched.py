"""The batched delta / ratio trackers (several lookups in flight per thread,
internship_amd/csrc/mrp_track.h, reading their draws by index from the
stream's ring of generated words) against the one-lookup-at-a-time loops that
state the reference semantics (`DeterminateNextVertex`, `Tr`,
core/render.cuh:93-172) on the plain stream: a host build of the same
templates on random rays through a field with negative regions, batch widths
1-6, step caps 1..10000, fresh and resumed streams.  Result bits, lookup
counts, stream positions and the draws that follow must be identical.  The same
program checks that block_order (the first-vertex stage's 8 x 8 work order)
visits every pixel exactly once."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_batched_trackers_match_sequential(tmp_path):
    exe = str(tmp_path / "track_batched")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "internship_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "track_batched.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout
