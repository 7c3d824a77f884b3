"""CPU check of the closed-form brick cut of a ray (dda_math.h SegEnum)
against the step-by-step integer walk, including rational crossing ties,
negative coordinates and origins on brick faces (tests/cpp/segments_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path, include_dir):
    exe = str(tmp_path / "segments_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + include_dir,
                    os.path.join(ROOT, "tests", "cpp", "segments_check.cpp"), "-o", exe],
                   check=True, timeout=300)
    return exe


def test_segments_match_walk(tmp_path):
    exe = _build(tmp_path, os.path.join(ROOT, "vdb_mapping_amd", "csrc"))
    for seed in (1, 2):
        r = subprocess.run([exe, "60000", str(seed)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        assert "0 differ" in r.stdout


def test_segment_check_catches_a_dropped_tie_rule(tmp_path):
    """The same check fails when the tie-break of count_before is removed."""
    src = open(os.path.join(ROOT, "vdb_mapping_amd", "csrc", "dda_math.h")).read()
    planted = src.replace("if (r == 0 && (q & 1)) {", "if (false) {")
    assert planted != src
    (tmp_path / "inc").mkdir()
    (tmp_path / "inc" / "dda_math.h").write_text(planted)
    exe = _build(tmp_path, str(tmp_path / "inc"))
    r = subprocess.run([exe, "20000", "1"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
