"""CSV snapshots (simp_output_*): the writer reproduces, byte for byte, the
files the reference's own output routine (in_out/output.f90) writes for
the same arrays -- golden files made by tests/golden/make_output_golden.py
with the compiled reference -- and the reader reads the last snapshot
back.  Host code only (no GPU)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from simp_amd import _lib, output

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "output_csv")
FILES = ("particle_momentum.csv", "particle_state.csv", "particle_other.csv")


def load_case(name):
    g = np.load(os.path.join(GOLD, name, "input.npz"))
    snaps = []
    for i in range(int(g["snapshots"])):
        snaps.append({k: g[f"s{i}_{k}"] for k in ("time", "x", "vx", "mass", "p", "hsml", "A",
                                                  "h", "eps", "sigma")})
    return int(g["n"]), int(g["nb"]), snaps


def write(folder, nb, snaps):
    output.create(folder)
    for s in snaps:
        output.append(folder, s["x"], s["vx"], s["mass"], s["p"], s["hsml"], s["A"], s["h"],
                      s["eps"], s["sigma"], nb, float(s["time"]))


@pytest.mark.parametrize("case", ["edge", "box500"])
def test_writer_matches_the_reference_output_routine_bytewise(tmp_path, case):
    n, nb, snaps = load_case(case)
    write(str(tmp_path), nb, snaps)
    for fn in FILES:
        with open(os.path.join(GOLD, case, fn), "rb") as f:
            want = f.read()
        with open(tmp_path / fn, "rb") as f:
            got = f.read()
        assert got == want, (case, fn)


def test_reader_returns_the_last_snapshot(tmp_path):
    n, nb, snaps = load_case("box500")
    s0 = snaps[0]
    # two snapshots: the reader must return the second
    s1 = {k: (v * 1.5 if k in ("x", "vx", "p", "eps", "sigma") else v) for k, v in s0.items()}
    s1["time"] = float(s0["time"]) + 3600.0
    write(str(tmp_path), nb, [s0, s1])
    r = output.read_last(str(tmp_path))
    assert r["n"] == n and r["nb"] == nb
    # every field equals the printed value (8 significant digits), with the
    # reference reader's unit conversions (initialize.f90:282, 300-301)
    rnd = lambda a: np.array([float("%.7e" % v) for v in np.ravel(a)]).reshape(np.shape(a))  # noqa: E731
    assert np.array_equal(r["x"], rnd(s1["x"] / 1000) * 1000)
    assert np.array_equal(r["vx"], rnd(s1["vx"] * 100) / 100)
    for k in ("mass", "p", "hsml", "A", "h", "eps", "sigma"):
        assert np.array_equal(r[k], rnd(s1[k])), k
    assert r["current_time"] == rnd(s1["time"] / 3600) * 3600
    # and writing what was read reproduces the last snapshot's lines
    os.makedirs(tmp_path / "again")
    write(str(tmp_path / "again"), r["nb"], [dict(r, time=r["current_time"])])
    for fn in FILES[:2]:
        last = open(tmp_path / fn).read().splitlines()[-n:]
        again = open(tmp_path / "again" / fn).read().splitlines()
        # positions and velocities went through a unit conversion and back:
        # compare the values, not the text, for the momentum file
        if fn == FILES[0]:
            a = np.array([[float(v) for v in ln.split(",")] for ln in last])
            b = np.array([[float(v) for v in ln.split(",")] for ln in again])
            assert np.allclose(a, b, rtol=2e-8, atol=0), fn
        else:
            assert last == again, fn


def test_reader_reports_capacity(tmp_path):
    n, nb, snaps = load_case("edge")
    write(str(tmp_path), nb, snaps[:1])
    with pytest.raises(_lib.SimpError) as ei:
        output.read_last(str(tmp_path), cap=n - 1)
    assert ei.value.status == _lib.SIMP_E_CAPACITY


def test_missing_folder_is_an_argument_error(tmp_path):
    with pytest.raises(_lib.SimpError) as ei:
        output.read_last(str(tmp_path / "nowhere"))
    assert ei.value.status == _lib.SIMP_E_ARG


@pytest.mark.gpu
def test_device_run_snapshot_prints_as_the_reference_run(tmp_path):
    """20 FP64 predictor-corrector steps of the box2500 golden scenario on
    the device, with diagnostics: the snapshot of its first 500 particles is
    the file the reference's output routine wrote from the reference's own
    arrays after the same steps (golden box500)."""
    from golden_io import load, scenario
    from simp_amd.api import Simp
    print("[gpu] device snapshot box2500, 20 steps", flush=True)
    g = load("box2500_pc20.npz")
    sc = scenario(g)
    K = int(g["out_itimestep"])
    dev = Simp.for_scenario(sc)
    dev.set_diagnostics(True)
    dev.time_integration(sc, K)
    st = dev.download()
    dg = dev.download_diagnostics()
    n, nb, snaps = load_case("box500")
    output.create(str(tmp_path))
    output.append(str(tmp_path), st["x"][:n], st["vx"][:n], g["in_mass"][:n], dg["p"][:n],
                  st["hsml"][:n], st["A"][:n], st["h"][:n], dg["eps"][:n], dg["sigma"][:n], nb,
                  float(snaps[0]["time"]))
    for fn in FILES:
        with open(os.path.join(GOLD, "box500", fn), "rb") as f:
            want = f.read()
        with open(tmp_path / fn, "rb") as f:
            got = f.read()
        assert got == want, fn
