"""Reader of a mesh's triangle words as the runtime shader reads them (src/bake/mirror.rs:12-130,
src/baked_lighting/shaders/triangles.wesl:1-60): triangle headers, light entries, shadow values
and bounce texels, plus distance-cube values."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MAX_TRIANGLE_LIGHTS = 256  # mirror.rs:10
DISTANCE_CUBE_WORDS = 64 * 64 * 3


@dataclass
class LightEntry:
    light_index: int
    shadow_offset: int
    bounce_offset: int


class TriangleView:
    """TriangleView::load (mirror.rs:38-50)."""

    def __init__(self, words: np.ndarray, triangle_index: int, bounce_layout: bool):
        self.words = words
        off = triangle_index * 4
        light_data = int(words[off])
        self.bounds = (int(words[off + 1]), int(words[off + 2]), int(words[off + 3]))
        self.light_count = min(int(words[light_data]), MAX_TRIANGLE_LIGHTS)
        self.light_data_offset = light_data + 1
        self.bounce_layout = bounce_layout

    def light(self, k: int) -> LightEntry:
        stride = 3 if self.bounce_layout else 2
        o = self.light_data_offset + k * stride
        return LightEntry(int(self.words[o]), int(self.words[o + 1]),
                          int(self.words[o + 2]) if self.bounce_layout else 0)

    def entries(self):
        return [self.light(k) for k in range(self.light_count)]

    def grid_index(self, x: int, y: int) -> int:
        lx = (x - self.bounds[0]) & 0xFFFFFFFF
        ly = (y - self.bounds[1]) & 0xFFFFFFFF
        return (ly * self.bounds[2] + lx) & 0xFFFFFFFF

    def shadow_value(self, e: LightEntry, x: int, y: int, bits: int):
        if e.shadow_offset == 0:
            return None
        i = self.grid_index(x, y)
        per = 32 // bits
        mask = (1 << bits) - 1
        word = int(self.words[e.shadow_offset + i // per])
        return ((word >> ((i % per) * bits)) & mask) / mask

    def shadow_bit(self, e: LightEntry, x: int, y: int):
        v = self.shadow_value(e, x, y, 1)
        return None if v is None else v > 0.5

    def bounce_texel(self, e: LightEntry, x: int, y: int, bits: int) -> float:
        """bounce_texel (mirror.rs:96-104): the stored (square-rooted) white bounce value."""
        i = self.grid_index(x, y)
        per = 32 // bits
        mask = (1 << bits) - 1
        word = int(self.words[e.bounce_offset + i // per])
        return ((word >> ((i % per) * bits)) & mask) / mask

    def bounce_texel_rgb(self, e: LightEntry, x: int, y: int, bits: int):
        i = self.grid_index(x, y)
        pixel_bits = 3 * bits
        mask = (1 << bits) - 1
        off, shift = i * pixel_bits // 32, i * pixel_bits % 32
        at = e.bounce_offset + off
        word = int(self.words[at]) >> shift
        if shift + pixel_bits > 32:
            word |= (int(self.words[at + 1]) << (32 - shift)) & 0xFFFFFFFF
        return tuple(((word >> (c * bits)) & mask) / mask for c in range(3))


def triangle_count(words: np.ndarray) -> int:
    """BakedMesh::triangle_count (format/types.rs:200-205)."""
    return int(words[0]) // 4 if len(words) else 0


def distance_cube_value(cubes: np.ndarray, light: int, index: int) -> float:
    """distance_cube_value (mirror.rs:127-130)."""
    packed = int(cubes[light * DISTANCE_CUBE_WORDS + (index >> 1)])
    return float(np.array([(packed >> ((index & 1) * 16)) & 0xFFFF], np.uint16).view(np.float16)[0])
