"""Movie files (reference movie_output.py:120-287): the header records this build
writes against the files the reference's own writer produced
(tests/golden/files16_*_mov.bin, written by tests/golden/make_golden.py), and
the movie-string parsing; the frames are compared on the GPU
(tests/test_gpu_facade.py)."""

import io
import json
import os
import struct

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MOVIES = ["Br CMB", "Vr CMB", "Vp CMB", "T CMB", "Br EQ", "Vr EQ", "T EQ", "Bt EQ"]


def records(buf: bytes):
    """Split a big-endian Fortran sequential file into its record payloads."""
    out, i = [], 0
    while i < len(buf):
        n = struct.unpack(">i", buf[i:i + 4])[0]
        out.append(buf[i + 4:i + 4 + n])
        assert struct.unpack(">i", buf[i + 4 + n:i + 8 + n])[0] == n
        i += 8 + n
    return out


def _params():
    from magic_amd.params import Params
    meta = json.load(open(os.path.join(GOLD, "files16.json")))
    return Params(dict(meta["cfg"]))


@pytest.mark.parametrize("ms", MOVIES)
def test_movie_header_matches_reference(ms):
    from magic_amd import grid
    from magic_amd.movie_output import parse_movie_string, write_movie_header
    p = _params()
    rad = grid.radial(p)
    ftype, nsurf, const, stem = parse_movie_string(ms, p)
    ref = records(open(os.path.join(GOLD, f"files16_{stem}.bin"), "rb").read())
    f = io.BytesIO()
    write_movie_header(f, p, rad, nsurf, const, [ftype])
    got = records(f.getvalue())
    assert len(ref) == 10 + 2 * 2 and len(got) == 10
    for k in (0, 1, 3, 4, 5):                      # version, sizes, types, run id, grid sizes
        assert got[k] == ref[k], k
    assert struct.unpack(">d", got[2])[0] == pytest.approx(struct.unpack(">d", ref[2])[0],
                                                           rel=1e-14)
    for k in (6, 7, 8, 9):                         # parameters, radii, colatitudes, longitudes
        a, b = np.frombuffer(got[k], ">f4"), np.frombuffer(ref[k], ">f4")
        assert a.shape == b.shape
        np.testing.assert_allclose(a, b, rtol=2e-7, atol=1e-7)
    # frame records: time, then one (n_theta, n_phi) or (n_r, n_phi) float32 field
    n = (p.n_theta if nsurf == 1 else p.n_r_max) * p.n_phi
    assert len(ref[10]) == 4 and len(ref[11]) == 4 * n


def test_movie_strings_and_equator_row():
    from magic_amd.movie_output import equator_row, parse_movie_string
    p = _params()
    assert parse_movie_string("br cmb", p)[:2] == (1, 1)
    assert parse_movie_string("S EQUAT", p)[3] == "T_EQU_mov"
    for bad in ("Br", "Xx CMB", "Br MID"):
        with pytest.raises(ValueError):
            parse_movie_string(bad, p)
    # the row the reference reads: sorted index of grid row n_theta/2 - 1, used as a row
    T = 48
    q = T // 2 - 1                                  # odd: a southern row, node (q-1)/2
    assert equator_row(T) == T - 1 - (q - 1) // 2
    assert equator_row(8) == 6                      # q = 3: southern node 1 -> sorted index 6
