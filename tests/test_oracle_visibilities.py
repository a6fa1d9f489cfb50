"""The oracle's visibilities, visibility seeds and χ² seeds, pinned without the reference running.

* Sign and sky axes: the reference pins `visibilities` against ehtim's observation of a known
  source (test/test_uvfits.jl:41-54): the direct transform of test/data/synth_eht2017_image.fits
  (32 x 32 Stokes, Jy/pixel) must reproduce ehtim's noiseless visibilities of it
  (test/data/synth_eht2017_noiseless_ehtim.csv), and the mirrored image must not. Both files are
  committed under tests/golden/ehtim. The bar is the reference's, 5e-7 Jy on a 1.4 Jy source (the
  float32 storage of ehtim's file); the oracle lands at 9e-8 Jy, the mirrored image at 0.9 Jy.
* The seed is the adjoint of the transform: sum_pixels seed . dI = Re sum_k w_k conj(V_k(dI)) for
  any w and dI (device_visibilities.jl:76-88), checked on random inputs.
* The χ² seeds (frame_chi2_seed!, device_visibilities.jl:97-111; _frame_seed!, Fit.jl:537-577):
  the seed of the weighted residuals is the gradient of χ², checked against central differences.
"""
import os

import numpy as np
import pytest

JY = 1e-23
HERE = os.path.join(os.path.dirname(__file__), "golden", "ehtim")


def read_stokes_fits(path):
    """A minimal reader for the fixture: one image HDU per Stokes parameter (STOKES card), BITPIX -64.
    img[i, j] follows the FITS axes (i along NAXIS1), as FITS.jl reads them (fits.jl:18-40)."""
    raw = open(path, "rb").read()
    off, imgs, psize_deg = 0, {}, None
    while off < len(raw):
        hdr = {}
        done = False
        while not done:
            block = raw[off:off + 2880]
            off += 2880
            for c in range(36):
                card = block[80 * c:80 * c + 80].decode("ascii")
                if card.startswith("END"):
                    done = True
                    break
                if card[8:10] == "= ":
                    hdr[card[:8].strip()] = card[10:].split("/")[0].strip().strip("'").strip()
        n1, n2, bp = int(hdr["NAXIS1"]), int(hdr["NAXIS2"]), int(hdr["BITPIX"])
        assert bp == -64
        nbytes = n1 * n2 * 8
        data = np.frombuffer(raw[off:off + nbytes], dtype=">f8").reshape(n2, n1)
        off += (nbytes + 2879) // 2880 * 2880
        imgs[hdr.get("STOKES", "I")[0]] = data.T.copy()   # [i, j] with i along NAXIS1
        if psize_deg is None:
            psize_deg = abs(float(hdr["CDELT2"]))
    n = imgs["I"].shape
    S = np.stack([imgs.get(s, np.zeros(n)) for s in "IQUV"], axis=-1)
    return S, psize_deg


def ehtim_csv(path):
    rows = np.genfromtxt(path, delimiter=",", names=True, dtype=None, encoding="ascii")
    u, v = rows["u"].astype(float), rows["v"].astype(float)
    vis = np.stack([rows[s + "re"] + 1j * rows[s + "im"] for s in "IQUV"], axis=-1)
    return u, v, vis


def screen_frames(S):
    """(na, nb, 4) image -> (1, na * nb, 4) in screen order (index i + na j)."""
    na, nb, _ = S.shape
    return S.transpose(1, 0, 2).reshape(1, na * nb, 4)


def test_visibility_sign_against_ehtim(oracle):
    S, psize_deg = read_stokes_fits(os.path.join(HERE, "synth_eht2017_image.fits"))
    psize = psize_deg * np.pi / 180
    assert S.shape == (32, 32, 4) and abs(psize - 200e-6 / 206264.806247 / 32) < 1e-13
    assert abs(S[..., 0].sum() - 1.4) < 1e-9
    u, v, vis = ehtim_csv(os.path.join(HERE, "synth_eht2017_noiseless_ehtim.csv"))
    assert len(u) == 386
    stokes = S * (JY / psize ** 2)
    model = oracle.visibilities(screen_frames(stokes), 32, 32, psize, u, v)[0]
    err = np.abs(model - vis).max()
    assert err < 5e-7, err
    assert abs(model[np.argmin(u * u + v * v), 0].real - 1.4) < 1e-3
    mirrored = oracle.visibilities(screen_frames(stokes[::-1]), 32, 32, psize, u, v)[0]
    assert np.abs(mirrored - vis).max() > 1e-2


def test_visibility_seed_is_the_adjoint(oracle):
    rng = np.random.default_rng(3)
    na, nb, nuv, nf = 9, 7, 23, 2
    psize = 2e-11
    u, v = rng.normal(scale=5e9, size=nuv), rng.normal(scale=5e9, size=nuv)
    tap = rng.uniform(0.5, 1.0, size=nuv)
    dI = rng.normal(size=(nf, na * nb, 4))
    W = rng.normal(size=(nf, nuv, 4)) + 1j * rng.normal(size=(nf, nuv, 4))
    V = oracle.visibilities(dI, na, nb, psize, u, v, taper=tap)
    seed = oracle.vis_seed(W, na, nb, psize, u, v, taper=tap)
    lhs = np.sum(seed * dI, axis=(1, 2))
    rhs = np.real(np.sum(W * np.conj(V), axis=(1, 2)))
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=0)
    # a planted transposition of the seed's pixel axes breaks the identity
    bad = seed.reshape(nf, nb, na, 4).transpose(0, 2, 1, 3).reshape(nf, na * nb, 4)
    assert not np.allclose(np.sum(bad * dI, axis=(1, 2)), rhs, rtol=1e-6)


def vis_chi2(oracle, I, na, nb, psize, u, v, tap, g, d, sig):
    V = oracle.visibilities(I[None], na, nb, psize, u, v, taper=tap)[0]
    r = (g[:, None] * V - d) / sig
    return float(np.sum(np.abs(r) ** 2)), r


def test_visibility_chi2_seed_is_the_gradient(oracle):
    """w = 2 conj(g) r / sigma through the adjoint transform is d chi2 / d I (frame_chi2_seed!)."""
    rng = np.random.default_rng(4)
    na, nb, nuv = 6, 5, 17
    psize = 3e-11
    scale = psize ** 2 / JY
    u, v = rng.normal(scale=4e9, size=nuv), rng.normal(scale=4e9, size=nuv)
    tap = rng.uniform(0.5, 1.0, size=nuv)
    g = np.exp(rng.normal(scale=0.1, size=nuv) + 1j * rng.normal(scale=0.3, size=nuv))
    I = rng.uniform(0.5, 1.0, size=(na * nb, 4)) / scale
    d = rng.normal(size=(nuv, 4)) + 1j * rng.normal(size=(nuv, 4))
    sig = rng.uniform(0.1, 0.3, size=(nuv, 4))
    chi, r = vis_chi2(oracle, I, na, nb, psize, u, v, tap, g, d, sig)
    w = 2 * np.conj(g)[:, None] * r / sig
    seed = oracle.vis_seed(w[None], na, nb, psize, u, v, taper=tap)[0]
    for p, q in [(0, 0), (7, 1), (13, 2), (29, 3)]:
        h = 1e-6 / scale
        Ip, Im = I.copy(), I.copy()
        Ip[p, q] += h
        Im[p, q] -= h
        fd = (vis_chi2(oracle, Ip, na, nb, psize, u, v, tap, g, d, sig)[0] - vis_chi2(oracle, Im, na, nb, psize, u, v, tap, g, d, sig)[0]) / (2 * h)
        assert abs(fd - seed[p, q]) <= 1e-6 * np.abs(seed).max(), (p, q, fd, seed[p, q])
