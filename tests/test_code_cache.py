"""The opt-in on-disk cache of hipRTC code objects (KTT_HIP_CACHE_DIR): an entry per (source,
kernel, options), reused on the next compilation, and a new entry when an included header
changes. Runs hipRTC on the CPU (no device needed), in a child process so the environment
variable takes effect."""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]

SCRIPT = """
import sys, time
sys.path.insert(0, {repo!r})
from ktt_amd import capi
start = time.perf_counter()
capi.compile_check({source!r}, "directCoulombSum", {config!r})
print(time.perf_counter() - start)
"""


def compile_once(cache: Path, source: Path, config: str | None) -> float:
    code = SCRIPT.format(repo=str(REPO), source=str(source), config=config)
    result = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                            env={**os.environ, "KTT_HIP_CACHE_DIR": str(cache)})
    assert result.returncode == 0, result.stderr
    return float(result.stdout.strip().splitlines()[-1])


def entries(cache: Path) -> set[str]:
    return {p.name for p in cache.glob("*.co")}


def test_code_objects_are_cached_and_invalidated(tmp_path):
    kernels = tmp_path / "kernels"
    shutil.copytree(REPO / "ktt_amd" / "kernels", kernels)
    cache = tmp_path / "cache"
    source = kernels / "CoulombSum3d.hip"

    first = compile_once(cache, source, None)
    assert len(entries(cache)) == 1
    again = compile_once(cache, source, None)
    assert len(entries(cache)) == 1
    assert again < first, (first, again)

    compile_once(cache, source, "POINTS_X=4,ROWS_PER_THREAD=8")
    assert len(entries(cache)) == 2

    # Editing an included header changes the key.
    header = kernels / "coulomb_core.h"
    header.write_text(header.read_text() + "\n// edited\n")
    os.utime(header, (time.time() + 5, time.time() + 5))
    compile_once(cache, source, None)
    assert len(entries(cache)) == 3
