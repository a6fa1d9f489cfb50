"""On-disk formats (kerrsplat_amd/io.py): ehtim-style FITS Stokes images and uvfits observations, the reference's
fits.jl and uvfits.jl, on the reference's own fixtures (test/data, copied to tests/golden/ehtim) at the bars of its
gate (test/test_uvfits.jl:23-80).

* read_uvfits on ehtim's synthetic EHT 2017 observation (noisy and noiseless, 386 rows each) against ehtim's own
  parse of the same files row by row: header quantities exact, times to 1e-6 h, u and v to 2e-7 relative (float32
  storage), visibilities to 1e-7 Jy, noise to 1e-9 Jy, stations and integration times exact.
* read_stokes_fits on the source image: 32 x 32 over 200 microarcseconds, 1.4 Jy in total; its direct transform
  (the oracle's DFT) reproduces the noiseless uvfits visibilities to 5e-7 Jy, and a mirrored sky fails by > 1e-2
  (the sky conventions of the reader); the noisy file gives a reduced chi2 of one against it.
* write_stokes_fits then read_stokes_fits / read_stokes_movie round trips: values, pixel size, frequency, MJD,
  the movie's frame times and camera.
"""
import csv
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from kerrsplat_amd import io

EHT = os.path.join(GOLDEN, "ehtim")
JY = 1e-23


def ehtim_rows(name):
    rows = list(csv.DictReader(open(os.path.join(EHT, f"{name}_ehtim.csv"))))
    f = lambda k: np.array([float(r[k]) for r in rows])
    vis = np.stack([f(s + "re") + 1j * f(s + "im") for s in "IQUV"], axis=1)
    sig = np.stack([f("sig" + s) for s in "IQUV"], axis=1)
    return rows, f("time_h"), f("u"), f("v"), vis, sig, f("tint")


@pytest.mark.parametrize("name", ["synth_eht2017", "synth_eht2017_noiseless"])
def test_uvfits_against_ehtim(name):
    obs = io.read_uvfits(os.path.join(EHT, f"{name}.uvfits"))
    rows, t, u, v, vis, sig, tint = ehtim_rows(name)
    assert len(obs) == len(rows) == 386
    assert obs.mjd == 57854 and obs.freq == 230e9 and obs.bandwidth == 1.856e9 and obs.source == "SYNTH"
    assert abs(obs.ra - 12.513728717168174) < 1e-9 and abs(obs.dec - 12.39112323919932) < 1e-9
    assert sorted(obs.stations) == ["AA", "AP", "AZ", "JC", "LM", "PV", "SM", "SR"]
    assert np.abs(obs.time - t).max() < 1e-6
    assert all(obs.stations[obs.s1[k]] == rows[k]["t1"].strip() and obs.stations[obs.s2[k]] == rows[k]["t2"].strip() for k in range(386))
    assert np.max(np.abs(obs.u - u) / (np.abs(u) + 1)) < 2e-7 and np.max(np.abs(obs.v - v) / (np.abs(v) + 1)) < 2e-7
    assert np.abs(obs.vis - vis).max() < 1e-7
    assert np.abs(obs.sigma - sig).max() < 1e-9
    assert np.array_equal(obs.tint, tint)
    # the circular products are the Stokes combinations: RR = I + V, LL = I - V, RL = Q + iU, LR = Q - iU
    assert np.allclose(obs.coh[:, 0], obs.vis[:, 0] + obs.vis[:, 3], rtol=0, atol=1e-12)
    assert np.allclose(obs.coh[:, 2], obs.vis[:, 1] + 1j * obs.vis[:, 2], rtol=0, atol=1e-12)


def screen_frames(S):
    na, nb, _ = S.shape
    return S.transpose(1, 0, 2).reshape(1, na * nb, 4)


def test_source_image_against_the_observation(oracle):
    img, hdr = io.read_stokes_fits(os.path.join(EHT, "synth_eht2017_image.fits"))
    psize = hdr["psize_deg"] * math.pi / 180
    assert img.shape == (32, 32, 4) and abs(psize - 200e-6 / 206264.806247 / 32) < 1e-13
    assert abs(img[..., 0].sum() - 1.4) < 1e-9 and hdr["freq"] == 230e9 and hdr["mjd"] == 57854.0
    stokes = img * (JY / psize ** 2)
    clean = io.read_uvfits(os.path.join(EHT, "synth_eht2017_noiseless.uvfits"))
    model = oracle.visibilities(screen_frames(stokes), 32, 32, psize, clean.u, clean.v)[0]
    assert np.abs(model - clean.vis).max() < 5e-7
    mirrored = oracle.visibilities(screen_frames(stokes[::-1]), 32, 32, psize, clean.u, clean.v)[0]
    assert np.abs(mirrored - clean.vis).max() > 1e-2
    noisy = io.read_uvfits(os.path.join(EHT, "synth_eht2017.uvfits"))
    r = (model - noisy.vis) / noisy.sigma
    chi2 = float(np.sum(r.real ** 2 + r.imag ** 2))
    assert 0.85 < chi2 / (8 * len(noisy)) < 1.15


def test_fits_round_trip_and_movie(tmp_path):
    rng = np.random.default_rng(2)
    M, D = 6.5e9, 16.8e6
    psize_M = 0.25
    frames = []
    for k, mjd in enumerate((60000.0, 60000.5)):
        im = rng.uniform(0.0, 1.0, size=(12, 10, 4)) * 1e-3
        p = tmp_path / f"frame{k}.fits"
        io.write_stokes_fits(str(p), im, psize_M, M_solar=M, D_pc=D, freq=230e9, mjd=mjd)
        frames.append((str(p), im))
    back, hdr = io.read_stokes_fits(frames[0][0])
    from kerrsplat_amd.api import gravitational_radius
    L = gravitational_radius(M)
    psize_rad = psize_M * L / (D * 3.085678e18)
    assert abs(hdr["psize_deg"] - psize_rad * 180 / math.pi) < 1e-18 and hdr["freq"] == 230e9 and hdr["mjd"] == 60000.0
    assert np.allclose(back * JY / psize_rad ** 2, frames[0][1], rtol=1e-14, atol=0)
    movie, cam, L2 = io.read_stokes_movie([f[0] for f in frames[::-1]], M_solar=M, D_pc=D, sigma=[1e-3] * 4)
    assert movie.data.shape == (12, 10, 2, 1, 4) and L2 == L
    assert np.allclose(movie.data[:, :, 1, 0, :], frames[1][1], rtol=1e-13, atol=0)
    assert movie.times[0] == 0.0 and abs(movie.times[1] - 0.5 * 86400 / io.time_unit(M)) < 1e-6
    assert cam.size == (12, 10) and abs(cam.alphas[1] - cam.alphas[0] - psize_M) < 1e-12 and abs(cam.alphas[0] + 5.5 * psize_M) < 1e-12
