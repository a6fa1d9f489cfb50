"""Host side of the batched screens: the uneven tiles of the batched configuration's main screen (every
pixel in exactly one tile, square tiles of the allowed sizes), its extra screens, the tiles' cameras and
index lists against the full screen's pixels, and the concatenated camera's part table."""
import numpy as np
import pytest

from kerrsplat_amd import synthetic as syn
from kerrsplat_amd.api import Camera, concatenate_screens, screen_tiles


def test_config_screens_sizes():
    main, splits, extras = syn.config_screens(2)
    assert main.size == (256, 256)
    cover = np.zeros((256, 256), dtype=int)
    for a0, a1, b0, b1 in splits:
        assert a1 - a0 == b1 - b0 and a1 - a0 in (32, 64, 128)
        cover[a0:a1, b0:b1] += 1
    assert (cover == 1).all()
    assert len({a1 - a0 for a0, a1, _, _ in splits}) > 1          # uneven
    assert len(extras) == 32
    assert all(48 <= c.size[0] <= 160 and c.size[0] == c.size[1] for c in extras)
    # the same draw every time
    again = syn.config_screens(2)
    assert again[1] == splits and [c.size for c in again[2]] == [c.size for c in extras]


def test_screen_tiles_and_concatenation():
    main = Camera((-10.0, 10.0), (-8.0, 8.0), (12, 10))
    splits = [(0, 6, 0, 5), (6, 12, 0, 5), (0, 12, 5, 10)]
    cams, idx = screen_tiles(main, splits)
    for c, ix, (a0, a1, b0, b1) in zip(cams, idx, splits):
        assert c.size == (a1 - a0, b1 - b0)
        assert np.array_equal(c.alphas, main.alphas[ix]) and np.array_equal(c.betas, main.betas[ix])
    assert np.array_equal(np.sort(np.concatenate(idx)), np.arange(main.npix))
    extra = Camera((-10.0, 10.0), (-10.0, 10.0), 7)
    cat = concatenate_screens(*cams, extra)
    assert cat.parts == [(6, 5), (6, 5), (12, 5), (7, 7)]
    assert cat.npix == main.npix + 49 and cat.size == (cat.npix, 1)
    assert np.array_equal(cat.alphas[-49:], extra.alphas)
    with pytest.raises(ValueError):
        screen_tiles(main, splits[:2])
