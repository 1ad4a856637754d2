"""The render path's cell storage (shaders/slang/common/cellStateUtils.slang):
conversions between the packed search layout and the r8ui image / Morton-tiled
buffer forms, against a numpy model written from the layout's definition.
Host-side functions of the C-ABI library; no GPU needed."""
import numpy as np
import pytest

import cave_amd as cv
import np_model as NP


def morton3(x, y, z):
    # bit b of x / y / z goes to position 3b / 3b+1 / 3b+2 (8x8x8 block)
    m = 0
    for b in range(3):
        m |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return m


def model(dense, layout):
    Z, Y, X = dense.shape
    iw = (X + 1) // 2
    pad = np.zeros((Z, Y, 2 * iw), np.uint32)
    pad[:, :, :X] = dense
    texel = pad[:, :, 0::2] | (pad[:, :, 1::2] << 4)  # low nibble = even x
    if layout == cv.RENDER_IMAGE:
        return texel.astype(np.uint8).reshape(-1)
    nbx, nby, nbz = (iw + 7) // 8, (Y + 7) // 8, (Z + 7) // 8
    buf = np.zeros(nbx * nby * nbz * 512, np.uint32)
    for z in range(Z):
        for y in range(Y):
            for tx in range(iw):
                blk = ((z // 8) * nby + y // 8) * nbx + tx // 8
                buf[blk * 512 + morton3(tx % 8, y % 8, z % 8)] = texel[z, y, tx]
    return buf


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (3, 5, 2), (17, 9, 8), (31, 16, 17), (40, 3, 9)])
@pytest.mark.parametrize("layout", [0, 1], ids=["image", "tiled"])
def test_render_layout_round_trip(dims, layout):
    X, Y, Z = dims
    dense = np.random.default_rng(X * 7 + Y + Z).integers(0, 16, size=(Z, Y, X)).astype(np.uint8)
    packed = NP.pack(dense)
    r = cv.packed_to_render(layout, packed, X, Y, Z)
    assert np.array_equal(r, model(dense, layout))
    assert np.array_equal(cv.render_to_packed(layout, r, X, Y, Z), packed)


def test_tiled_known_answer():
    # cells (0,0,0)=3, (1,0,0)=5 share texel 0 -> byte 0x53 at element 0; texel
    # (1,0,0) is Morton position 1, texel (0,1,0) position 2, (0,0,1) position 4,
    # texel (8,0,0) starts the second 8x8x8 block (element 512)
    dense = np.zeros((2, 2, 18), np.uint8)
    dense[0, 0, 0], dense[0, 0, 1] = 3, 5
    dense[0, 0, 2] = 7
    dense[0, 1, 0] = 9
    dense[1, 0, 1] = 11
    dense[0, 0, 16] = 2
    r = cv.packed_to_render(cv.RENDER_TILED, NP.pack(dense), 18, 2, 2)
    assert r[0] == 0x53 and r[1] == 0x07 and r[2] == 0x09 and r[4] == 0xB0 and r[512] == 0x02
    assert int(np.count_nonzero(r)) == 5
