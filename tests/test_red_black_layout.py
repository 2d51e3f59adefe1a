"""The premises of the red-black colour-split layout, checked on the CPU for
box shapes of every parity: natural rows 2m and 2m + 1 of a box always
differ in colour (parity of x + y + z), so m = v >> 1 numbers each colour's
rows in order; the positions rb_pos(c, m) = 128 (m >> 6) + 64 c + (m & 63)
are distinct and stay inside the scene's whole tiles of 512 rows; and every
position decodes back to its row the way the device kernels decode it
(vx_grid.hip: rb_pos, rb_row). The device path itself is covered by
test_gpu_red_black.py."""
import itertools

import numpy as np
import pytest

TILE = 512


def colours(dims):
    nx, ny, _ = dims
    v = np.arange(dims[0] * dims[1] * dims[2])
    return ((v % nx) + (v // nx) % ny + v // (nx * ny)) & 1


def rb_pos(c, m):
    return ((m >> 6) << 7) | (c << 6) | (m & 63)


SHAPES = [(a, b, c) for a in (1, 2, 3, 5, 8) for b in (1, 2, 3, 4, 7) for c in (1, 2, 3, 5)] + [
    (64, 64, 64), (12, 12, 30), (511, 1, 1), (512, 1, 1), (513, 1, 1), (33, 31, 29), (20, 18, 16), (127, 3, 3)]


@pytest.mark.parametrize("dims", SHAPES)
def test_pairs_differ_in_colour_and_positions_fit(dims):
    n = dims[0] * dims[1] * dims[2]
    col = colours(dims)
    assert np.all(col[0:n - 1:2] != col[1:n:2])
    v = np.arange(n)
    pos = rb_pos(col, v >> 1)
    assert len(np.unique(pos)) == n
    assert pos.max() < -(-n // TILE) * TILE
    # decode as the kernels do: colour from the position, the pair's member of that colour
    s = np.arange(-(-n // TILE) * TILE)
    c = (s >> 6) & 1
    m = ((s >> 7) << 6) | (s & 63)
    v0 = 2 * m
    c0 = np.where(v0 < n, colours_at(dims, v0), -1)
    vv = np.where((v0 < n) & (c0 == c), v0, v0 + 1)
    ok = vv < n
    assert ok.sum() == n
    assert np.array_equal(pos[vv[ok]], s[ok])


def colours_at(dims, v):
    nx, ny, _ = dims
    return ((v % nx) + (v // nx) % ny + v // (nx * ny)) & 1


def test_every_bond_joins_two_colours():
    for dims in itertools.product((2, 3, 4), repeat=3):
        col = colours(dims).reshape(dims[2], dims[1], dims[0])
        assert np.all(col[:, :, 1:] != col[:, :, :-1])
        assert np.all(col[:, 1:, :] != col[:, :-1, :])
        assert np.all(col[1:, :, :] != col[:-1, :, :])
