"""Reader/writer of the reference's printed vectors and points, and the IB
checkpoint file."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from afmibins_amd import formats as FM

HERE = os.path.dirname(os.path.abspath(__file__))

SAMPLE = np.array([0.0, -0.0, 1.0, -3.0, 0.5, 1.0 / 3.0, 2.0 / 3.0, 1e-5, -1.234567e-12, 123456.0, 1234567.0,
                   9.9999995e5, 6.02214076e23, 5e-324, 1.7976931348623157e308, math.inf, -math.inf, math.nan])


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_writer_matches_a_default_ostream(tmp_path):
    exe = str(tmp_path / "ostream_print")
    subprocess.check_call(["g++", "-O1", os.path.join(HERE, "native", "ostream_print.cpp"), "-o", exe])
    pt = (17, 0.25, -1.0 / 7.0, 1e-7)
    stdin = f"{SAMPLE.size} " + " ".join(float(x).hex() for x in SAMPLE) + f" {pt[0]} " + \
        " ".join(float(x).hex() for x in pt[1:]) + "\n"
    out = subprocess.run([exe], input=stdin, capture_output=True, text=True, check=True).stdout.splitlines()
    assert FM.format_vector(SAMPLE) == out[0]
    assert FM.format_point(*pt) == out[1]
    # and the reader takes the ostream text back (six significant digits)
    got = FM.parse_vector(out[0])
    fin = np.isfinite(SAMPLE)
    assert np.array_equal(np.isnan(got), np.isnan(SAMPLE))
    assert np.all(got[np.isinf(SAMPLE)] == SAMPLE[np.isinf(SAMPLE)])
    assert np.allclose(got[fin], SAMPLE[fin], rtol=5e-6, atol=0)
    idx, xyz = FM.parse_points(out[1])
    assert idx.tolist() == [17] and np.allclose(xyz[0], pt[1:], rtol=5e-6)


def test_reader_several_vectors_points_and_empty():
    text = "solution: [1, 2.5, -3e-05]'\nresidual [ ]'\nnext [4,5]' and 0 (0.5, 0.25, 0.125)\n12 (1e-05, -2, inf)"
    vs = FM.parse_vectors(text)
    assert [v.tolist() for v in vs] == [[1.0, 2.5, -3e-05], [], [4.0, 5.0]]
    idx, xyz = FM.parse_points(text)
    assert idx.tolist() == [0, 12]
    assert xyz.tolist() == [[0.5, 0.25, 0.125], [1e-05, -2.0, math.inf]]
    with pytest.raises(ValueError):
        FM.parse_vector("no vector here")


def test_writer_round_trips_at_full_precision():
    rng = np.random.default_rng(4)
    x = rng.standard_normal(1000) * 10.0 ** rng.integers(-30, 30, 1000)
    assert np.array_equal(FM.parse_vector(FM.format_vector(x, precision=17)), x)


def test_checkpoint_round_trip_and_checks(tmp_path):
    rng = np.random.default_rng(5)
    n, m = 16, 7
    arrays = dict(u=rng.standard_normal((n, n)), v=rng.standard_normal((n, n)), X=rng.uniform(0, 1, (m, 2)),
                  R=rng.uniform(0, 1, (m, 2)), k=rng.uniform(0, 2, m), ds=np.full(m, 0.01))
    p = str(tmp_path / "ck.npz")
    FM.save_checkpoint(p, **arrays, n=n, mu=1e-3, rho=1.0, dt=2e-5, tol=1e-10, tol_mode=1, precision=1, step=42)
    meta, got = FM.load_checkpoint(p)
    assert meta["n"] == n and meta["step"] == 42 and meta["precision"] == 1 and meta["markers"] == m
    assert meta["mu"] == 1e-3 and meta["dt"] == 2e-5
    for name, a in arrays.items():
        assert np.array_equal(got[name], a), name
    assert "start0" not in got and "start1" not in got
    # the CG forms' start potentials: either, both or none
    s1 = rng.standard_normal((n, n))
    FM.save_checkpoint(p, **arrays, n=n, mu=1e-3, rho=1.0, dt=2e-5, tol=1e-10, precision=1, starts=(None, s1))
    meta, got = FM.load_checkpoint(p)
    assert meta["starts"] == [False, True, False, False] and "start0" not in got
    assert np.array_equal(got["start1"], s1)
    s3 = rng.standard_normal((n, n))
    FM.save_checkpoint(p, **arrays, n=n, mu=1e-3, rho=1.0, dt=2e-5, tol=1e-10, precision=1,
                       starts=(s1, s1, None, s3))
    meta, got = FM.load_checkpoint(p)
    assert meta["starts"] == [True, True, False, True] and np.array_equal(got["start3"], s3)
    with pytest.raises(ValueError):
        FM.save_checkpoint(p, **arrays, n=n, mu=1e-3, rho=1.0, dt=2e-5, tol=1e-10, starts=(s1[:3], None))
    with pytest.raises(ValueError):
        FM.save_checkpoint(p, **dict(arrays, u=arrays["u"][:3]), n=n, mu=1e-3, rho=1.0, dt=2e-5, tol=1e-10)
    bad = str(tmp_path / "other.npz")
    np.savez(bad, meta=np.array(['{"format": "something else"}']))
    with pytest.raises(ValueError):
        FM.load_checkpoint(bad)
