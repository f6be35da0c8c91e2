"""The drop-in executables end to end on an MI355X: files in the reference's
formats in, the reference's outputs out.

- wih_twos on the Stanford Bunny reproduces benchmarks/compiled_bunny/output/
  field-twos-0016.csv (per-site u, score, score_se, steps, limit_hits) and
  walk-traces.csv (the --trace-output rows), with the binary32 query radius
  those files were written with;
- wih_study reproduces retained submission-study requests (paired and
  mean-only), alone and as a --serve session;
- wih_twos --streams philox equals the library's estimate of the same call.
"""

import json
import subprocess

import numpy as np
import pytest

import study_fixtures as SF
from cli_files import BIN, read_csv, write_ply, write_points

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


def _run(tool, args, stdin=None, timeout=240):
    print("cli", tool, " ".join(args[:2]), flush=True)
    return subprocess.run([str(BIN / tool), *args], capture_output=True, text=True, input=stdin, timeout=timeout)


@pytest.fixture(scope="module")
def bunny_files(golden, tmp_path_factory):
    d = tmp_path_factory.mktemp("bunny")
    m = np.load(golden / "bunny_69674.npz")
    write_ply(d / "bunny.ply", m["vertices"], m["faces"])
    sites = np.load(golden / "bunny_sites.npy")
    write_points(d / "points.csv", sites)
    meta = json.loads((golden / "bunny_meta.json").read_text())
    return d, sites, meta


def test_wih_twos_reproduces_the_reference_field(bunny_files, golden):
    d, sites, meta = bunny_files
    f = meta["fields"]["twos_16"]
    ref = np.load(golden / "field_twos_16.npz")
    r = _run("wih_twos", ["--mesh", str(d / "bunny.ply"), "--points", str(d / "points.csv"),
                          "--output", str(d / "field.csv"), "--summary", str(d / "summary.json"),
                          "--walks", str(f["walks"]), "--seed", str(f["seed"]), "--repeats", "2",
                          "--epsilon", str(f["epsilon"]), "--max-steps", str(f["max_steps"]),
                          "--binary32-radius", "1"])
    assert r.returncode == 0, r.stderr
    out = read_csv(d / "field.csv")
    np.testing.assert_array_equal(out["site"], np.arange(len(sites)))
    np.testing.assert_array_equal(np.c_[out["x"], out["y"], out["z"]], sites)
    assert np.sum(out["steps"] != ref["steps"]) <= 1
    same = out["steps"] == ref["steps"]
    np.testing.assert_allclose(out["u"][same], ref["u"][same], rtol=1e-11)
    np.testing.assert_allclose(out["score"][same], ref["score"][same], rtol=1e-10)
    np.testing.assert_allclose(out["score_se"][same], ref["score_se"][same], rtol=1e-7)
    np.testing.assert_array_equal(out["limit_hits"], ref["limit_hits"])
    s = json.loads((d / "summary.json").read_text())
    assert s["sites"] == len(sites) and s["walks_per_site"] == f["walks"] and len(s["timings_seconds"]) == 2
    assert s["faces"] == 69674 and s["max_step_limit_hits"] == 0
    assert abs(s["mean_steps"] - ref["steps"].sum() / (len(sites) * f["walks"])) < 1e-3


def test_wih_twos_trace_output(bunny_files, golden):
    d, sites, meta = bunny_files
    traces = np.load(golden / "bunny_walk_traces.npy")
    t = meta["traces"]
    r = _run("wih_twos", ["--mesh", str(d / "bunny.ply"), "--points", str(d / "points.csv"),
                          "--output", str(d / "f2.csv"), "--summary", str(d / "s2.json"), "--walks", "4",
                          "--repeats", "1", "--seed", str(t["seed"]), "--trace-output", str(d / "trace.csv"),
                          "--trace-sites", str(t["sites"]), "--binary32-radius", "1"])
    assert r.returncode == 0, r.stderr
    got = read_csv(d / "trace.csv")
    rows = np.c_[got["path"], got["step"], got["x"], got["y"], got["z"], got["radius"]]
    assert rows.shape == traces.shape
    np.testing.assert_array_equal(rows[:, :2], traces[:, :2])
    assert np.abs(rows[:, 2:] - traces[:, 2:]).max() <= 1e-13


def _study_args(d, req, name):
    # The study driver passes an empty --cavities value when there are no
    # holes (benchmarks/submission_study/**/raw/*/request.json).
    return ["--mesh", str(d / "branch.ply"), "--config", str(d / f"{name}.cfg"), "--points",
            str(d / f"{name}.pts"), "--cavities", str(d / f"{name}.holes") if len(req.holes) else "",
            "--cavity-radius", repr(req.cavity_radius),
            "--output", str(d / f"{name}.csv"), "--summary", str(d / f"{name}.json"), "--mode", "twos",
            "--source-mode", "housing", "--walks", str(req.walks), "--repeats", "1", "--seed", str(req.seed),
            "--threads", "4", "--max-steps", str(req.max_steps), "--epsilon", repr(req.epsilon),
            "--hot-only", "1" if req.hot else "0", "--fd", "0", "--diff", "0", "--grid", "65,47,27",
            "--mg-cycles", "120", "--mg-tolerance", "1e-7", "--diff-offset", "0.004"]


@pytest.fixture(scope="module")
def study_files(golden, tmp_path_factory):
    d = tmp_path_factory.mktemp("study")
    v, f, requests = SF.load(golden)
    write_ply(d / "branch.ply", v, f)
    picks = {"paired": next(r for r in requests if not r.hot and r.walks == 1024),
             "hot": next(r for r in requests if r.hot)}
    for name, req in picks.items():
        (d / f"{name}.cfg").write_text(req.config)
        write_points(d / f"{name}.pts", req.points, sep=" ")
        write_points(d / f"{name}.holes", req.holes, sep=" ")
    return d, picks


def _check_study(d, name, req):
    out = read_csv(d / f"{name}.csv")
    SF.compare(req, out["u"], out["p"], out["score"], out["same_index"], out["steps"].astype(np.int64),
               out["caps"].astype(np.int64), rtol=1e-6)
    s = json.loads((d / f"{name}.json").read_text())
    assert s["steps"] == int(req.counts[:, 0].sum()) and s["caps"] == 0


def test_wih_study_requests(study_files):
    d, picks = study_files
    for name, req in picks.items():
        r = _run("wih_study", _study_args(d, req, name))
        assert r.returncode == 0, r.stderr
        _check_study(d, name, req)


def test_wih_study_serve_session(study_files):
    d, picks = study_files
    lines = []
    for name, req in picks.items():
        args = ["wih_study", *_study_args(d, req, name)]
        lines.append(f"{len(args)}\n" + "\n".join(args) + "\n")
        (d / f"{name}.csv").unlink(missing_ok=True)
    r = _run("wih_study", ["--serve"], stdin="".join(lines))
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["DONE", "0", "DONE", "0"]
    for name, req in picks.items():
        _check_study(d, name, req)


def test_wih_twos_philox_equals_the_library(bunny_files, golden):
    """The production path through the executable (Philox streams, binary32
    walks, adaptive budgets, distance grid) gives the library's results for
    the same call, to the last bit."""
    d, sites, _ = bunny_files
    args = ["--mesh", str(d / "bunny.ply"), "--points", str(d / "points.csv"), "--output", str(d / "ph.csv"),
            "--summary", str(d / "ph.json"), "--streams", "philox", "--walks", "64", "--walks-max", "1024",
            "--ci-rel", "0.05", "--seed", "77", "--repeats", "1", "--epsilon", "5e-4", "--max-steps", "2048",
            "--distance-grid", "1"]
    r = _run("wih_twos", args)
    assert r.returncode == 0, r.stderr
    out = read_csv(d / "ph.csv")
    m = np.load(golden / "bunny_69674.npz")
    ctx = twos.Context()
    try:
        g = ctx.geometry(twos.MESH, mesh=ctx.mesh(m["vertices"], m["faces"]))
        lib = ctx.estimate(g, sites, [twos.reference_affine_side()], combine=twos.SQUARE, precision=32, seed=77,
                           walks=64, walks_max=1024, ci_rel=0.05, eps=5e-4, max_steps=2048, distance_grid=True,
                           use_graph=True)
    finally:
        ctx.close()
    np.testing.assert_array_equal(out["steps"], lib.steps)
    np.testing.assert_array_equal(out["u"], lib.u)
    np.testing.assert_array_equal(out["score"], lib.score)
    np.testing.assert_array_equal(out["score_se"], lib.score_se)
    s = json.loads((d / "ph.json").read_text())
    assert s["streams"] == "philox" and s["walks_per_site"] == pytest.approx(np.mean(lib.walks))
