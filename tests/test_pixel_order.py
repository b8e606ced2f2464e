"""The order in which a job's path slots cover its pixels (photon-v2_amd/csrc/pixel_order.h,
compiled for the host by tests/hostemu): a bijection for every image size, with full 8 x 8
tiles in the interior, and tile_slot its inverse."""
import numpy as np
import pytest

from hostemu import emu


@pytest.mark.parametrize("W,H", [(1, 1), (8, 8), (16, 8), (7, 5), (13, 21), (64, 3), (3, 64), (100, 37), (257, 129),
                                 (1024, 1024), (1920, 1080), (640, 360)])
def test_slot_order_is_a_bijection_with_its_inverse(W, H):
    pix, slot = emu.pixel_order(W, H)
    n = W * H
    assert np.array_equal(np.sort(pix), np.arange(n))        # every pixel exactly once
    assert np.array_equal(slot[pix], np.arange(n))            # tile_slot inverts tile_pixel
    assert np.array_equal(pix[slot], np.arange(n))


def test_full_tiles_are_8x8_blocks():
    W, H = 40, 24
    pix, _ = emu.pixel_order(W, H)
    for k in range((W // 8) * (H // 8)):
        x, y = pix[64 * k:64 * k + 64] % W, pix[64 * k:64 * k + 64] // W
        assert x.max() - x.min() == 7 and y.max() - y.min() == 7
        # rows within the tile, left to right
        assert np.array_equal(x - x.min(), np.tile(np.arange(8), 8))
        assert np.array_equal(y - y.min(), np.repeat(np.arange(8), 8))
