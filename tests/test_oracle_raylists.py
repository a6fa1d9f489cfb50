"""The oracle's per-ray splat lists (ray_lists, dual_gradient.jl:283-359): for every ray the splats whose
support sphere one of its step-taking samples enters at the observation time.

Pinned by:
* brute force on a small camera: the oracle's stored samples of every ray, their quasi-Cartesian positions
  and the bounding-sphere test of polarized.jl:110-120 evaluated in numpy, with the horizon guard; pairs
  closer than 1e-9 to the sphere are left out of the comparison (rounding may put them on either side);
* the reference's defining property: the transport of a ray over its listed splats alone equals the
  transport over all splats, bit for bit (also with the half-orbit cutoff), since a splat outside its
  support adds exactly zero; a list that dropped a splat the ray sees fails it;
* the half-orbit cutoff only removes splats (lists with nmax = 0 are subsets of the uncut lists), a splat far
  from the hole is listed for no ray, and each frame's lists equal those of a one-frame call at its time.
"""
import math

import numpy as np
import pytest

from kerrsplat_amd import synthetic as syn

A, TH, N = 0.5, math.radians(60.0), 96
SUPPORT2 = -2.0 * math.log(1e-6)


@pytest.fixture(scope="module")
def camera():
    n, fov = 10, 12.0
    al, be = np.meshgrid(np.linspace(-fov, fov, n), np.linspace(-fov, fov, n), indexing="ij")
    return al.ravel(order="F"), be.ravel(order="F")


def splats(n=12):
    P = syn.thermal_splats(4, A, n)
    P[3:6] -= 0.4                 # narrower splats: each ray sees a few of them
    P[20] = np.linspace(-0.05, 0.05, n)
    return P


def brute_force(oracle, camera, P, t_obs):
    """(hits, margin): margin[i, j] = min over the step-taking samples of ray j of (d^2 - R^2) / R^2."""
    s = oracle.samples(A, TH, camera[0], camera[1], N)
    hor = (1.0 + math.sqrt(1.0 - A * A)) * (1.0 + 1e-3)
    take = ((s["flags"] & 1) != 0) & (s["r"] > hor)
    x, y, z = syn.quasi_cartesian_ks(A, s["r"], s["theta"], s["phi"])
    t = t_obs - s["t"]
    ns = P.shape[1]
    margin = np.full((ns, len(camera[0])), np.inf)
    for i in range(ns):
        ph = P[20, i] * (t - P[10, i])
        cx = P[0, i] * np.cos(ph) - P[1, i] * np.sin(ph)
        cy = P[0, i] * np.sin(ph) + P[1, i] * np.cos(ph)
        R2 = SUPPORT2 * math.exp(2.0 * max(P[3, i], P[4, i], P[5, i]))
        d2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - P[2, i]) ** 2
        m = np.where(take, (d2 - R2) / R2, np.inf)
        margin[i] = m.min(axis=1)
    return margin <= 0, margin


def test_hits_match_brute_force(oracle, camera):
    P = splats()
    hits = oracle.ray_hits(A, TH, camera[0], camera[1], N, P, [7.0])[:, :, 0]
    ref, margin = brute_force(oracle, camera, P, 7.0)
    clear = np.abs(margin) > 1e-9
    assert np.array_equal(hits.astype(bool)[clear], ref[clear])
    assert clear.sum() >= hits.size - 2
    per_ray = hits.sum(axis=0)
    assert per_ray.max() > 0 and per_ray.min() < P.shape[1]          # the lists are proper subsets
    assert (per_ray > 0).sum() > 10


@pytest.mark.parametrize("nmax,slab", [(-1, 0.0), (0, 0.5)])
def test_transport_over_the_lists_is_the_full_transport(oracle, camera, nmax, slab):
    P = splats()
    L = syn.length_unit()
    hits = oracle.ray_hits(A, TH, camera[0], camera[1], N, P, [0.0], nmax=nmax, slab=slab)[:, :, 0]
    full = oracle.polarized_image(A, TH, camera[0], camera[1], N, P, 0.0, 230e9, L, nmax=nmax, slab=slab)[0]
    assert np.abs(full[:, 0]).max() > 0
    seen = 0
    for j in range(len(camera[0])):
        ids = np.flatnonzero(hits[:, j])
        one = ([camera[0][j]], [camera[1][j]])
        if len(ids) == 0:
            assert np.all(full[j] == 0.0)
            continue
        sub = oracle.polarized_image(A, TH, one[0], one[1], N, P[:, ids], 0.0, 230e9, L, nmax=nmax, slab=slab)[0]
        assert np.array_equal(sub[0], full[j]), j
        seen += full[j, 0] != 0.0
    assert seen > 10
    # a list with one seen splat dropped changes the ray's image
    j = int(np.argmax(np.abs(full[:, 0])))
    ids = np.flatnonzero(hits[:, j])
    contrib = [i for i in ids if oracle.polarized_image(A, TH, [camera[0][j]], [camera[1][j]], N, P[:, [i]], 0.0, 230e9, L,
                                                        nmax=nmax, slab=slab)[0][0, 0] != 0.0]
    drop = np.setdiff1d(ids, contrib[:1])
    sub = oracle.polarized_image(A, TH, [camera[0][j]], [camera[1][j]], N, P[:, drop], 0.0, 230e9, L, nmax=nmax, slab=slab)[0]
    assert not np.array_equal(sub[0], full[j])


def test_cutoff_frames_and_far_splats(oracle, camera):
    P = splats()
    far = P[:, :1].copy()
    far[0:3, 0] = [2e3, -2e3, 5e2]
    Q = np.hstack([P, far])
    h = oracle.ray_hits(A, TH, camera[0], camera[1], N, Q, [0.0, 40.0])
    assert not h[-1].any()
    cut = oracle.ray_hits(A, TH, camera[0], camera[1], N, Q, [0.0, 40.0], nmax=0, slab=0.5)
    assert np.all(cut <= h) and cut.sum() < h.sum()
    for f, t in enumerate([0.0, 40.0]):
        assert np.array_equal(h[:, :, f], oracle.ray_hits(A, TH, camera[0], camera[1], N, Q, [t])[:, :, 0])
    assert not np.array_equal(h[:, :, 0], h[:, :, 1])              # the pattern rotation moves the splats
