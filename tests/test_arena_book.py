"""The scratch-arena rules (fri-cc_amd/csrc/arena.h) on the host: slots grow
on demand, growth waits for the device before freeing, and a pinned arena
(held by a recorded graph) never frees or releases memory -- growth is an
error instead.  Built with g++ against a stand-in allocator (no GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_arena_book(tmp_path):
    exe = tmp_path / "arena_book_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "fri-cc_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "arena_book_test.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "arena book ok" in r.stdout
