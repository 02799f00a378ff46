"""The runtime reader (dynlight_hip.mirror.TriangleView) against the reference's own reader tests
(src/bake/mirror.rs:138-199): words written by the builder (serialize_triangles, restated in the
oracle) read back through the shader's wrapping u32 math."""
import pytest

import oracle
from dynlight_hip.mirror import TriangleView
from test_oracle_kats import pack, serialize

U32_MAX = 0xFFFFFFFF


@pytest.fixture(scope="module")
def ol():
    return oracle.load()


def test_reads_back_what_the_builder_wrote(ol):
    # padded 7x7 grid of the crop (1, 0)-(3, 2); the bit of lightmap texel (2, 1) is local (3, 3)
    bits = [0] * 49
    bits[3 + 3 * 7] = 1
    words = serialize(ol, [(1, 0, 3, 2)], [[(9, pack(ol, bits, 1), [0x04030201, 0x08070605])]], True)
    view = TriangleView(words, 0, True)
    assert view.light_count == 1
    assert view.bounds == (U32_MAX, U32_MAX - 1, 7)
    entry = view.light(0)
    assert entry.light_index == 9
    assert view.shadow_bit(entry, 2, 1) is True
    assert view.shadow_bit(entry, 1, 1) is False
    # the padding starts two texels before the crop: lightmap (-1, -2) is local (0, 0)
    assert view.shadow_bit(entry, U32_MAX, U32_MAX - 1) is False
    # 8 bits per pixel: pixel index 5 is byte 1 of the second word
    x, y = (5 + view.bounds[0]) & U32_MAX, view.bounds[1]
    assert view.bounce_texel(entry, x, y, 8) == 6.0 / 255.0


def test_soft_shadows_read_back_as_fractions(ol):
    # padded 6x6 grid of 4-bit levels of the crop (0, 0)-(1, 1); lightmap (0, 0) is local (2, 2)
    levels = [0] * 36
    levels[2 + 2 * 6] = 15
    levels[3 + 2 * 6] = 5
    words = serialize(ol, [(0, 0, 1, 1)], [[(0, pack(ol, levels, 4), None)]], False)
    view = TriangleView(words, 0, False)
    entry = view.light(0)
    assert view.shadow_value(entry, 0, 0, 4) == 1.0
    assert view.shadow_value(entry, 1, 0, 4) == 5.0 / 15.0
    assert view.shadow_value(entry, 0, 1, 4) == 0.0


def test_colored_bounce_texels_cross_word_boundaries(ol):
    # 6 bits per channel, 18 per pixel: pixel 1 spans bits 18-35, across the first word
    r, g, b = 17, 42, 63
    pixel = r | g << 6 | b << 12
    stream = pixel << 18
    bounce = [stream & U32_MAX, stream >> 32]
    words = serialize(ol, [(0, 0, 0, 0)], [[(3, [0], bounce)]], True)
    view = TriangleView(words, 0, True)
    e = view.light(0)
    x, y = (1 + view.bounds[0]) & U32_MAX, view.bounds[1]
    assert view.bounce_texel_rgb(e, x, y, 6) == (17 / 63, 42 / 63, 1.0)
