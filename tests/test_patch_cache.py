"""The C++ drivers' binary patch cache (include/bbrt/drivers.hpp), on the host: a directory of reference-format
patch files is parsed once and written to the cache; the second load comes from the cache with every parsed
value bit-identical; changing a file makes the cache stale, so the next load parses the text again."""
import os
import subprocess

import pytest

from billardballraytracing3d_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "drivers_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cpp/drivers_check is built by __graft_entry__.build()")
def test_patch_cache_round_trip_and_staleness(tmp_path):
    sc, _, _, _ = S.two_spheres_plane_source()
    sub = S.subset_scene(sc, list(range(0, sc.n_patches, 7)))
    pdir = tmp_path / "patches"
    pdir.mkdir()
    S.write_patch_dir(sub, str(pdir))
    cdir = tmp_path / "cache"
    cdir.mkdir()
    r = subprocess.run([EXE, "cache", str(pdir), str(cdir)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"patches={sub.n_patches}" in r.stdout and "equal=1" in r.stdout
    assert len(list(cdir.iterdir())) == 1        # one cache file, rewritten in place after the change
