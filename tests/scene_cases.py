"""Small bake inputs built from the reference's test helpers (src/bake/tests/mod.rs:23-137),
covering every light shape and the edge cases of the texel and seam logic."""
import math

import numpy as np

from dynlight_hip import abi, scenes
from dynlight_hip.abi import BakeInput, LightData, MeshData


def opts(density=16.0, **kw):
    o = abi.default_options()
    o.pixel_density = density
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def box_on_floor(extra_lights=(), density=16.0, **kw):
    meshes = [scenes.floor(1, (0, 0, 0), 4.0), scenes.cuboid(2, (-0.5, 0.0, -0.5), (0.5, 1.0, 0.5))]
    return BakeInput(meshes, list(extra_lights)), opts(density, **kw)


def point_case():
    return box_on_floor([scenes.point_light(10, (0.3, 3.0, 0.2), 8.0), scenes.point_light(11, (-2.5, 1.5, 2.0), 5.0)])


def panel_case():
    panel = scenes.area_light(20, (0.2, 3.0, -0.1), 9.0, scenes.quat_from_rotation_x(-math.pi / 2), (1.5, 0.8))
    tube = scenes.area_light(21, (-1.5, 2.0, 1.0), 7.0, (0.0, 0.0, 0.0, 1.0), (2.0, 0.0))
    return box_on_floor([panel, tube], area_light_samples=16, area_light_min_samples=4)


def sun_case():
    hard = scenes.sun_light(30, (0.4, 1.0, 0.3), 0.0)
    soft = scenes.sun_light(31, (-0.5, 0.8, 0.2), 4.0)
    return box_on_floor([hard, soft], area_light_samples=16, area_light_min_samples=16)


def sky_case():
    return box_on_floor([scenes.sky_light(40), scenes.point_light(41, (1.0, 2.0, 1.0), 6.0)], sky_samples=32)


def mixed_case():
    meshes = [scenes.cuboid(1, (-5.0, 0.0, -5.0), (5.0, 4.0, 5.0), inward=True),
              scenes.cuboid(2, (-1.0, 0.0, -1.0), (0.5, 2.0, 0.5)),
              scenes.cuboid(3, (0.0, 0.0, 0.0), (1.2, 1.0, 1.5)),  # overlaps box 2: buried texels
              scenes.cuboid(4, (2.0, 0.0, -3.0), (2.05, 3.5, -2.95))]  # a thin pillar: touched texels
    lights = [scenes.point_light(100, (0.0, 3.5, 2.0), 9.0),
              scenes.area_light(101, (-3.0, 3.9, -3.0), 8.0, scenes.quat_from_rotation_x(-math.pi / 2), (1.0, 1.0)),
              scenes.point_light(102, (3.0, 1.0, 3.0), 4.0),
              scenes.sky_light(103)]
    return BakeInput(meshes, lights), opts(12.0, area_light_samples=8, area_light_min_samples=2, sky_samples=16)


def odd_geometry_case():
    """Flat-normal mesh, a degenerate triangle, non-finite UVs, a non-occluding mesh, UVs
    outside the unit square and a light exactly on a texel."""
    pos = np.array([[0, 0, 0], [2, 0, 0], [2, 0, 2], [0, 0, 2], [1, 0, 1]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [np.nan, 0.5]], np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2, 0, 4, 2], np.uint32)  # last triangle is degenerate (collinear)
    flat = MeshData(1, pos, idx, uv)
    glass = scenes.floor(2, (1.0, 1.0, 1.0), 0.5)
    glass.occludes = False
    outside = scenes.floor(3, (4.0, 0.0, 0.0), 1.0)
    outside.uv1 = outside.uv1 * np.float32(1.7) - np.float32(0.3)
    lights = [scenes.point_light(50, (1.0, 0.0, 1.0), 3.0),
              scenes.point_light(51, (1.0, 2.0, 1.0), 6.0)]
    return BakeInput([flat, glass, outside], lights), opts(8.0)


def _albedo(m, rgb):
    m.albedo = rgb
    return m


def bounce_point_case():
    meshes = [_albedo(scenes.cuboid(1, (-3.0, 0.0, -3.0), (3.0, 2.5, 3.0), inward=True), (0.8, 0.3, 0.2)),
              _albedo(scenes.cuboid(2, (-0.6, 0.0, -0.4), (0.4, 1.2, 0.6)), (0.1, 0.7, 0.4))]
    a = scenes.point_light(60, (1.0, 2.0, 1.2), 6.0, bounce=True)
    b = scenes.point_light(61, (-1.5, 1.0, -1.0), 3.5, bounce=True)
    b.bounce_samples = 12
    b.bounce_intensity = 1.7
    c = scenes.point_light(62, (2.0, 1.0, -2.0), 4.0)
    return BakeInput(meshes, [a, b, c]), opts(10.0, colored_bounce=1, photon_cube_resolution=64)


def bounce_sun_case():
    sun = scenes.sun_light(70, (0.3, 1.0, 0.4), 0.0)
    sun.illumination = abi.ILLUMINATION_SINGLE_BOUNCE
    sun.bounce_samples = 8
    return box_on_floor([sun, scenes.point_light(71, (0.5, 2.0, 0.5), 5.0, bounce=True)], density=10.0,
                        photon_cube_resolution=48)


def bounce_dither_case():
    inp, o = bounce_point_case()
    lights = inp.lights
    lights[0].bounce_compression = abi.BOUNCE_BITS4
    lights[1].bounce_compression = abi.BOUNCE_BITS6
    o.bounce_compression = abi.BOUNCE_BITS5
    lights[2].illumination = abi.ILLUMINATION_SINGLE_BOUNCE
    return BakeInput(inp.meshes, lights), o


def bounce_white_dither_case():
    inp, o = bounce_dither_case()
    o.colored_bounce = 0
    return inp, o


def maps_case():
    """Sun maps with an odd texel count (the last word's upper half is padding), two suns in slot
    order, and sky probes at a spacing that leaves probes inside the boxes."""
    meshes = [scenes.floor(1, (0, 0, 0), 4.0), scenes.cuboid(2, (-0.5, 0.0, -0.5), (0.5, 1.0, 0.5)),
              scenes.cuboid(3, (1.0, 0.0, 1.0), (2.5, 2.0, 1.6))]
    lights = [scenes.sun_light(51, (0.3, 1.0, 0.2), 0.53), scenes.sun_light(50, (-0.7, 0.5, 0.1), 0.0),
              scenes.sky_light(52)]
    return BakeInput(meshes, lights), opts(4.0, sun_map_resolution=33, sky_samples=20, sky_probe_spacing=0.35,
                                           area_light_samples=4, area_light_min_samples=4)


def half_transparent(mesh_id, center, half, rgba=(255, 255, 255, 0, 255, 255, 255, 255), width=2, height=1):
    """A plane whose UV0 is its UV1 and whose alpha texture is transparent on one side
    (src/bake/tests/occluder.rs:87-97)."""
    plane = scenes.floor(mesh_id, center, half)
    plane.uv0 = plane.uv1.copy()
    plane.alpha = (width, height, np.array(rgba, np.uint8), 0.5)
    return plane


def alpha_case():
    """A half-transparent screen over a floor: alpha-tested point, area, sun and sky lights pass
    the transparent half; a light without transparency is blocked by both halves.  A tiled 3x5
    texture with repeat wrap and UV0 outside the unit square exercises the wrap."""
    screen = half_transparent(2, (0.0, 1.0, 0.0), 1.0)
    tiles = half_transparent(3, (2.5, 0.8, 2.5), 0.6, width=3, height=5,
                             rgba=[v for i in range(15) for v in (255, 255, 255, 255 if (i * 7) % 3 else 40)])
    tiles.uv0 = tiles.uv1 * np.float32(2.7) - np.float32(0.35)
    meshes = [scenes.floor(1, (0, 0, 0), 4.0), screen, tiles]
    pa = scenes.point_light(60, (0.1, 2.5, 0.2), 6.0)
    pa.transparency = abi.TRANSPARENCY_ENABLED
    pb = scenes.point_light(61, (-0.2, 2.2, -0.1), 6.0)
    area = scenes.area_light(62, (2.4, 2.0, 2.6), 6.0, scenes.quat_from_rotation_x(-math.pi / 2), (0.6, 0.4))
    area.transparency = abi.TRANSPARENCY_ENABLED
    sun = scenes.sun_light(63, (0.2, 1.0, 0.1), 1.0)
    sun.transparency = abi.TRANSPARENCY_ENABLED_MAX
    sky = scenes.sky_light(64)
    sky.transparency = abi.TRANSPARENCY_ENABLED
    return BakeInput(meshes, [pa, pb, area, sun, sky]), opts(10.0, area_light_samples=8, area_light_min_samples=4,
                                                            sky_samples=16, sun_map_resolution=48, sky_probe_spacing=0.5)


CASES = {"point": point_case, "panel": panel_case, "sun": sun_case, "sky": sky_case, "mixed": mixed_case,
         "odd_geometry": odd_geometry_case, "bounce_point": bounce_point_case, "bounce_sun": bounce_sun_case,
         "bounce_dither": bounce_dither_case, "bounce_white_dither": bounce_white_dither_case, "maps": maps_case, "alpha": alpha_case}
