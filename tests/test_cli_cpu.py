"""The drop-in executables' argument and input handling, without a GPU:
usage errors exit 2 (wih_twos.cpp:103-135), input errors print
"fatal: <what>" and exit 1 (:1860-1863) before any device work, and the
--serve protocol rejects malformed headers with 2 (study.cpp:365-370)."""

import subprocess

import numpy as np
import pytest

from cli_files import BIN, write_ply, write_points

pytestmark = pytest.mark.skipif(not (BIN / "wih_twos").exists(), reason="drop-in executables not built")


def _run(tool, args, stdin=None):
    return subprocess.run([str(BIN / tool), *args], capture_output=True, text=True, input=stdin, timeout=60)


@pytest.fixture()
def inputs(tmp_path):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    f = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    write_ply(tmp_path / "m.ply", v, f)
    write_points(tmp_path / "p.csv", [[0.1, 0.1, 0.1], [0.2, 0.1, 0.1]])
    base = ["--mesh", str(tmp_path / "m.ply"), "--points", str(tmp_path / "p.csv"),
            "--output", str(tmp_path / "f.csv"), "--summary", str(tmp_path / "s.json")]
    return tmp_path, base


@pytest.mark.parametrize("extra,needle", [
    (["--walks", "1"], "invalid numeric option"),
    (["--epsilon", "0"], "invalid numeric option"),
    (["--mode", "voxel"], "not provided by the MI355X backend"),
    (["--mode", "diffwos"], "Differential WoS requires --cavities"),
    (["--bogus", "1"], "unknown option --bogus"),
    (["--walks"], "missing value for --walks"),
    (["--walks", "x"], "invalid value for --walks"),
    (["--cavities", "c.csv"], "invalid numeric option"),  # cavities need a radius
    (["--streams", "other"], "--streams must be"),
    (["--walks-max", "64"], "need --streams philox"),
])
def test_usage_errors_exit_2(inputs, extra, needle):
    _, base = inputs
    r = _run("wih_twos", base + extra)
    assert r.returncode == 2 and needle in r.stderr


def test_required_flags(inputs):
    r = _run("wih_twos", [])
    assert r.returncode == 2 and "are required" in r.stderr


@pytest.mark.parametrize("mesh_text,needle", [
    ("ply\nformat binary_little_endian 1.0\nelement vertex 3\nelement face 1\nend_header\n",
     "expected a nonempty ASCII triangle PLY mesh"),
    ("ply\nformat ascii 1.0\nelement vertex 3\nelement face 1\nend_header\n0 0 0\n1 0 0\n",
     "truncated PLY vertex section"),
    ("ply\nformat ascii 1.0\nelement vertex 3\nelement face 1\nend_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 7\n",
     "expected valid triangular PLY faces"),
    ("ply\nformat ascii 1.0\nelement vertex 4\nelement face 1\nend_header\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n4 0 1 2 3\n",
     "expected valid triangular PLY faces"),
])
def test_mesh_errors_are_fatal(inputs, mesh_text, needle):
    tmp, base = inputs
    (tmp / "m.ply").write_text(mesh_text)
    r = _run("wih_twos", base)
    assert r.returncode == 1 and r.stderr.strip() == f"fatal: {needle}"


def test_points_errors_are_fatal(inputs):
    tmp, base = inputs
    (tmp / "p.csv").write_text("# only a comment\n\n")
    r = _run("wih_twos", base)
    assert r.returncode == 1 and "points file contains no points" in r.stderr
    (tmp / "p.csv").unlink()
    r = _run("wih_twos", base)
    assert r.returncode == 1 and "cannot open points" in r.stderr


def test_valid_inputs_reach_the_device(inputs):
    """Inputs parse; without a GPU the device context is what fails."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    _, base = inputs
    r = _run("wih_twos", base)
    assert r.returncode == 1 and r.stderr.startswith("fatal: device")


@pytest.mark.parametrize("stdin", ["x\n", "0\n", "201\n", "3\nwih_study\n--mesh\n"])
def test_serve_rejects_malformed_requests(stdin):
    r = _run("wih_study", ["--serve"], stdin=stdin)
    assert r.returncode == 2 and "DONE" not in r.stdout


def test_serve_reports_a_failed_request(inputs, tmp_path):
    """A request with a missing config fails with status 1 (reported as
    DONE 1) before any device work, and the session ends with that status."""
    _, base = inputs
    args = ["wih_study", *base, "--config", str(tmp_path / "missing.txt")]
    r = _run("wih_study", ["--serve"], stdin=f"{len(args)}\n" + "\n".join(args) + "\n")
    assert r.returncode == 1 and r.stdout.strip() == "DONE 1" and "bad sensor config" in r.stderr


def test_study_rejects_other_solvers(inputs, tmp_path):
    _, base = inputs
    cfg = tmp_path / "c.txt"
    cfg.write_text("0 0 0 0.1\n0.5 0 0 0.1 3\n")
    for extra, needle in ((["--mode", "mg"], "mode mg is not provided"), (["--sensor-side", "2"], "sensor side")):
        r = _run("wih_study", [*base, "--config", str(cfg), *extra])
        assert r.returncode == 1 and needle in r.stderr, (extra, r.stderr)
