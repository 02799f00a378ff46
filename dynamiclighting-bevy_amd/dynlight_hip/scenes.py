"""Synthetic bake inputs: the reference's test-scene helpers and the seeded benchmark scenes.

Helpers follow src/bake/tests/mod.rs:23-137 (`floor`, `cuboid` with its 3x2 UV1 atlas at 0.8
fill, `point_light`, `area_light`, `sun_light`, `sky_light`).  The procedural scenes are:

* ``box_room``      -- closed 10x4x10 m room (12 inward triangles) plus 2 seeded boxes, 4 point
                        lights of range 6-12 m, 16 texels/m, 1-bit shadows, no bounce.
* ``interior``      -- 8x8 grid of 6x3x6 m rooms with 1x2 m door openings, seeded boxes and
                        pillars up to ~200k triangles, albedo U(0.1,0.9)^3, 256 point/spot
                        lights of range 4-12 m with one colored bounce, 32 texels/m.
* ``outdoor``       -- 1024x512 fBm heightfield (~1M triangles) plus 2000 box buildings, a
                        soft sun with one bounce, 64 rectangle/tube area lights, a baked sky,
                        4 texels/m (terrain capped at 2048^2).
* ``batch``         -- 512 independent small interiors (grids of 6x3x6 m rooms with boxes):
                        log-uniform 500-5e4 triangles, 1-64 point/spot/area lights, 1e3-2e6
                        texels, bounce on odd jobs.

Randomness is PCG32 seeded through SplitMix64, restated here so scenes are identical on every
machine.
"""
from __future__ import annotations

import math

import numpy as np

from .abi import (BakeInput, LightData, MeshData, ILLUMINATION_DIRECT,
                  ILLUMINATION_SINGLE_BOUNCE, SHAPE_POINT, SHAPE_RECTANGLE, SHAPE_SKY,
                  SHAPE_SUN, SHAPE_TUBE, default_options, DYNAMIC_ANGULAR)

MASK64 = (1 << 64) - 1


def splitmix64(v: int) -> int:
    v = (v + 0x9E3779B97F4A7C15) & MASK64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK64
    return v ^ (v >> 31)


class Pcg32:
    """PCG32 XSH-RR (same recurrence as src/effects/rng.rs:48-71)."""

    def __init__(self, seed: int, stream: int = 54):
        self.state = 0
        self.inc = ((stream << 1) | 1) & MASK64
        self.next_u32()
        self.state = (self.state + seed) & MASK64
        self.next_u32()

    @classmethod
    def seeded(cls, seed: int) -> "Pcg32":
        s = splitmix64(seed)
        return cls(s, splitmix64(s))

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & MASK64
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((x >> rot) | (x << ((32 - rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u32() >> 8) / 16777216.0)

    def below(self, n: int) -> int:
        return self.next_u32() % n

    def u32s(self, count: int) -> np.ndarray:
        """The next `count` outputs of next_u32, computed at once: the state after k steps is an
        affine map of the current one, whose coefficients come from a prefix composition."""
        if count <= 0:
            return np.zeros(0, np.uint32)
        with np.errstate(over="ignore"):
            a = np.full(count, 6364136223846793005, np.uint64)
            c = np.full(count, self.inc, np.uint64)
            shift = 1
            while shift < count:  # (a[k], c[k]) becomes the map over steps 0..k
                na, nc = a.copy(), c.copy()
                na[shift:] = a[shift:] * a[:-shift]
                nc[shift:] = a[shift:] * c[:-shift] + c[shift:]
                a, c, shift = na, nc, shift * 2
            s0 = np.uint64(self.state)
            old = np.empty(count, np.uint64)
            old[0] = s0
            old[1:] = a[:-1] * s0 + c[:-1]
            self.state = int(a[-1] * s0 + c[-1])
        x = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = old >> np.uint64(59)
        out = (x >> rot) | (x << ((np.uint64(32) - rot) & np.uint64(31)))
        return (out & np.uint64(0xFFFFFFFF)).astype(np.uint32)

    def uniforms(self, count: int, lo=0.0, hi=1.0) -> np.ndarray:
        """`count` successive uniform(lo, hi) draws."""
        return lo + (hi - lo) * ((self.u32s(count) >> 8).astype(np.float64) / 16777216.0)


# ---- reference test helpers (src/bake/tests/mod.rs:23-137) -----------------------------------
def floor(mesh_id: int, center, half: float) -> MeshData:
    corners = [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]
    c = np.asarray(center, dtype=np.float32)
    pos = np.array([[c[0] + x * half, c[1], c[2] + z * half] for x, z in corners], dtype=np.float32)
    uv1 = np.array([[(x + 1) * 0.5, (z + 1) * 0.5] for x, z in corners], dtype=np.float32)
    return MeshData(mesh_id, pos, np.array([0, 2, 1, 0, 3, 2], dtype=np.uint32), uv1,
                    normals=np.tile(np.array([0, 1, 0], np.float32), (4, 1)))


_FACES = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)


def _cuboid_template():
    """Unit-cube corners (local in [-1,1]), normals, indices and UV1 of `cuboid`."""
    local, normals, uv1, idx = [], [], [], []
    for face, n in enumerate(_FACES):
        u = np.array([1, 0, 0], np.float32) if abs(n[1]) > 0.5 else np.array([0, 1, 0], np.float32)
        v = np.cross(n, u).astype(np.float32)
        base = len(local)
        cell = np.array([face % 3, face // 3], np.float32)
        for a, b in [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]:
            local.append(n + u * np.float32(a) + v * np.float32(b))
            normals.append(n)
            inner = np.array([a + 1.0, b + 1.0], np.float32) * np.float32(0.5) * np.float32(0.8) + np.float32(0.1)
            uv1.append((cell + inner) / np.array([3.0, 2.0], np.float32))
        idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    return (np.array(local, np.float32), np.array(normals, np.float32),
            np.array(idx, np.uint32), np.array(uv1, np.float32))


_TEMPLATE = _cuboid_template()


def cuboid(mesh_id: int, mn, mx, inward: bool = False, albedo=(1.0, 1.0, 1.0)) -> MeshData:
    """An axis-aligned box with outward faces (inward for room shells)."""
    local, normals, idx, uv1 = _TEMPLATE
    mn = np.asarray(mn, np.float32)
    mx = np.asarray(mx, np.float32)
    pos = ((mn + mx) * np.float32(0.5) + local * (mx - mn) * np.float32(0.5)).astype(np.float32)
    if inward:
        idx = idx.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1).copy()
        normals = -normals
    return MeshData(mesh_id, pos, idx.copy(), uv1.copy(), normals=normals.copy(), albedo=tuple(albedo),
                    content_hash=splitmix64(mesh_id))


def point_light(light_id: int, position, radius: float, bounce: bool = False) -> LightData:
    return LightData(light_id, tuple(float(x) for x in position), float(radius),
                     illumination=ILLUMINATION_SINGLE_BOUNCE if bounce else ILLUMINATION_DIRECT)


def quat_from_rotation_x(angle: float):
    s, c = math.sin(angle * 0.5), math.cos(angle * 0.5)
    return (s, 0.0, 0.0, c)


def quat_from_rotation_arc(frm, to):
    """Quat::from_rotation_arc for unit vectors (used by sun_light)."""
    frm = np.asarray(frm, np.float64)
    to = np.asarray(to, np.float64)
    to = to / np.linalg.norm(to)
    d = float(np.dot(frm, to))
    if d < -1.0 + 1e-6:
        axis = np.cross(frm, [1.0, 0, 0])
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(frm, [0, 1.0, 0])
        axis /= np.linalg.norm(axis)
        return (float(axis[0]), float(axis[1]), float(axis[2]), 0.0)
    c = np.cross(frm, to)
    q = np.array([c[0], c[1], c[2], 1.0 + d])
    q /= np.linalg.norm(q)
    return tuple(float(x) for x in q)


def area_light(light_id: int, position, radius: float, rotation, size) -> LightData:
    shape = SHAPE_RECTANGLE if size[1] > 0 else SHAPE_TUBE
    return LightData(light_id, tuple(position), float(radius), shape=shape,
                     rotation=tuple(rotation), size=tuple(size))


def sun_light(light_id: int, toward, degrees: float) -> LightData:
    return LightData(light_id, (500.0, -300.0, 70.0), 0.5, shape=SHAPE_SUN,
                     rotation=quat_from_rotation_arc([0, 0, 1.0], toward), angular_diameter=degrees)


def sky_light(light_id: int) -> LightData:
    return LightData(light_id, (-40.0, 900.0, 3.0), 0.5, shape=SHAPE_SKY)


# ---- box room: closed room, 2 boxes, 4 point lights ------------------------------------------
def box_room(seed: int = 1):
    """Returns (BakeInput, options) of the closed-room scene."""
    rng = Pcg32.seeded(seed)
    meshes = [cuboid(1, (-5.0, 0.0, -5.0), (5.0, 4.0, 5.0), inward=True)]
    for k in range(2):
        ext = [rng.uniform(0.2, 3.0) for _ in range(3)]
        cx = rng.uniform(-5.0 + ext[0] / 2, 5.0 - ext[0] / 2)
        cz = rng.uniform(-5.0 + ext[2] / 2, 5.0 - ext[2] / 2)
        meshes.append(cuboid(2 + k, (cx - ext[0] / 2, 0.0, cz - ext[2] / 2),
                             (cx + ext[0] / 2, ext[1], cz + ext[2] / 2)))
    lights = []
    for k in range(4):
        p = (rng.uniform(-4.9, 4.9), rng.uniform(0.1, 3.9), rng.uniform(-4.9, 4.9))
        lights.append(point_light(100 + k, p, rng.uniform(6.0, 12.0)))
    opts = default_options()
    opts.pixel_density = 16.0
    opts.dynamic_geometry = DYNAMIC_ANGULAR
    return BakeInput(meshes, lights), opts


# ---- interior: 8x8 rooms with doors, boxes/pillars, 256 bounce lights -------------------------
def _room_shell(ox: float, oz: float, w: float, h: float, d: float, doors: dict):
    """Inward-facing room shell with 1x2 m door openings; one 3x2 atlas cell per face."""
    pos, nrm, uv, idx = [], [], [], []

    def quad(p0, p1, p2, p3, n, cell, uvq):
        base = len(pos)
        for p, q in zip((p0, p1, p2, p3), uvq):
            pos.append(p)
            nrm.append(n)
            uv.append(((cell[0] + 0.1 + 0.8 * q[0]) / 3.0, (cell[1] + 0.1 + 0.8 * q[1]) / 2.0))
        idx.extend([base, base + 1, base + 2, base, base + 2, base + 3])

    x0, x1, z0, z1 = ox, ox + w, oz, oz + d
    # floor (normal +y) and ceiling (normal -y): windings chosen so cross(b-a, c-a) = normal
    quad((x0, 0, z0), (x0, 0, z1), (x1, 0, z1), (x1, 0, z0), (0, 1, 0), (0, 0),
         [(0, 0), (0, 1), (1, 1), (1, 0)])
    quad((x0, h, z0), (x1, h, z0), (x1, h, z1), (x0, h, z1), (0, -1, 0), (1, 0),
         [(0, 0), (1, 0), (1, 1), (0, 1)])
    # walls: (name, origin, along, normal, length, cell)
    walls = [("-x", (x0, 0, z0), (0, 0, 1), (1, 0, 0), d, (2, 0)),
             ("+x", (x1, 0, z1), (0, 0, -1), (-1, 0, 0), d, (0, 1)),
             ("-z", (x1, 0, z0), (-1, 0, 0), (0, 0, 1), w, (1, 1)),
             ("+z", (x0, 0, z1), (1, 0, 0), (0, 0, -1), w, (2, 1))]
    for name, o, a, n, length, cell in walls:
        o = np.array(o, np.float64)
        a = np.array(a, np.float64)
        up = np.array([0.0, 1.0, 0.0])

        def wall_quad(s0, s1, t0, t1):
            p = lambda s, t: tuple(o + a * s + up * t)  # noqa: E731
            # cross(along, up) points into the room for these orientations
            quad(p(s0, t0), p(s1, t0), p(s1, t1), p(s0, t1), n, cell,
                 [(s0 / length, t0 / h), (s1 / length, t0 / h), (s1 / length, t1 / h), (s0 / length, t1 / h)])

        if doors.get(name):
            dl, dr = length / 2 - 0.5, length / 2 + 0.5
            wall_quad(0.0, dl, 0.0, h)
            wall_quad(dr, length, 0.0, h)
            wall_quad(dl, dr, 2.0, h)
        else:
            wall_quad(0.0, length, 0.0, h)
    pos = np.array(pos, np.float32)
    nrm = np.array(nrm, np.float32)
    idx = np.array(idx, np.uint32)
    # fix windings so every face normal points along the stored normal
    tri = idx.reshape(-1, 3)
    a, b, c = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    fn = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", fn, nrm[tri[:, 0]]) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return pos, nrm, tri.reshape(-1).astype(np.uint32), np.array(uv, np.float32)


def interior(seed: int = 2, target_triangles: int = 200_000, lights: int = 256,
             pillar_share: float = 0.3, grid: int = 8, density: float = 32.0):
    """Returns (BakeInput, options) of the procedural interior scene."""
    rng = Pcg32.seeded(seed)
    W, H, D = 6.0, 3.0, 6.0
    positions, normals, indices, uvs, ids, albedo = [], [], [], [], [], []
    vcount = [0]
    icount = [0]

    def add(pos, nrm, idx, uv, mesh_id, alb):
        positions.append(pos)
        normals.append(nrm)
        indices.append(idx)
        uvs.append(uv)
        ids.append(mesh_id)
        albedo.append(alb)
        vcount.append(vcount[-1] + len(pos))
        icount.append(icount[-1] + len(idx))

    mesh_id = 1
    for gz in range(grid):
        for gx in range(grid):
            doors = {"-x": gx > 0, "+x": gx < grid - 1, "-z": gz > 0, "+z": gz < grid - 1}
            p, n, i, u = _room_shell(gx * W, gz * D, W, H, D, doors)
            add(p, n, i, u, mesh_id, (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9)))
            mesh_id += 1
    shell_tris = icount[-1] // 3
    boxes = max(0, (target_triangles - shell_tris) // 12)
    local, tnrm, tidx, tuv = _TEMPLATE
    for _ in range(boxes):
        room = rng.below(grid * grid)
        rx, rz = (room % grid) * W, (room // grid) * D
        if rng.uniform() < pillar_share:
            ext = [rng.uniform(0.2, 1.0), H, 0.0]
            ext[2] = rng.uniform(0.2, 1.0)
        else:
            ext = [rng.uniform(0.2, 3.0), rng.uniform(0.2, 3.0), rng.uniform(0.2, 3.0)]
        cx = rx + rng.uniform(ext[0] / 2, W - ext[0] / 2)
        cz = rz + rng.uniform(ext[2] / 2, D - ext[2] / 2)
        mn = np.array([cx - ext[0] / 2, 0.0, cz - ext[2] / 2], np.float32)
        mx = np.array([cx + ext[0] / 2, ext[1], cz + ext[2] / 2], np.float32)
        pos = ((mn + mx) * np.float32(0.5) + local * (mx - mn) * np.float32(0.5)).astype(np.float32)
        add(pos, tnrm, tidx, tuv, mesh_id, (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9)))
        mesh_id += 1
    light_list = []
    for k in range(lights):
        room = rng.below(grid * grid)
        rx, rz = (room % grid) * W, (room // grid) * D
        p = (rx + rng.uniform(0.1, W - 0.1), rng.uniform(0.1, H - 0.1), rz + rng.uniform(0.1, D - 0.1))
        light_list.append(point_light(10_000 + k, p, rng.uniform(4.0, 12.0), bounce=True))
    inp = BakeInput.from_arrays(
        ids=np.array(ids, np.uint64), positions=np.concatenate(positions), indices=np.concatenate(indices),
        uv1=np.concatenate(uvs), normals=np.concatenate(normals),
        vertex_offsets=np.array(vcount, np.int64), index_offsets=np.array(icount, np.int64),
        albedo=np.array(albedo, np.float32), content_hash=np.array([splitmix64(i) for i in ids], np.uint64),
        lights=light_list)
    opts = default_options()
    opts.pixel_density = density
    opts.colored_bounce = 1
    opts.dynamic_geometry = DYNAMIC_ANGULAR
    return inp, opts


def _fbm(rng: Pcg32, nx: int, nz: int, octaves: int) -> np.ndarray:
    """Value-noise fBm on an nx x nz vertex grid, scaled to [0, 1]: octave o has a lattice of
    2^(o+2) cells across x, smoothstep-interpolated, amplitude 1/2^o."""
    xs = np.arange(nx, dtype=np.float64) / max(1, nx - 1)
    zs = np.arange(nz, dtype=np.float64) / max(1, nx - 1)
    out = np.zeros((nz, nx))
    for o in range(octaves):
        cells = 2 ** (o + 2)
        cz = int(math.ceil(zs[-1] * cells)) + 1 if nz > 1 else 1
        lattice = np.array([rng.uniform() for _ in range((cells + 2) * (cz + 2))]).reshape(cz + 2, cells + 2)
        fx, fz = xs * cells, zs * cells
        ix, iz = np.floor(fx).astype(int), np.floor(fz).astype(int)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        a = lattice[iz][:, ix] * (1 - sx) + lattice[iz][:, ix + 1] * sx
        b = lattice[iz + 1][:, ix] * (1 - sx) + lattice[iz + 1][:, ix + 1] * sx
        out += (a * (1 - sz)[:, None] + b * sz[:, None]) / 2 ** o
    lo, hi = out.min(), out.max()
    return (out - lo) / (hi - lo) if hi > lo else out * 0.0


def outdoor(seed: int = 3, nx: int = 1024, nz: int = 512, buildings: int = 2000, area_lights: int = 64,
            density: float = 4.0, octaves: int = 6, height: float = 20.0):
    """Returns (BakeInput, options) of the outdoor scene: an nx x nz heightfield at 1 m spacing
    (fBm heights 0-20 m, planar UV1, one mesh), seeded box buildings (footprint 4-20 m, height
    3-30 m), a soft sun (0.53 degrees, toward (0.45, 0.75, 0.5), one bounce of 32 samples),
    rectangle and tube area lights (0.5-4 m, range 8-20 m, 16 samples) and a baked sky (64 rays)."""
    rng = Pcg32.seeded(seed)
    h = (_fbm(rng, nx, nz, octaves) * height).astype(np.float32)
    gx, gz = np.meshgrid(np.arange(nx, dtype=np.float32), np.arange(nz, dtype=np.float32))
    terrain_pos = np.stack([gx, h, gz], axis=-1).reshape(-1, 3).astype(np.float32)
    du = np.gradient(h.astype(np.float64), axis=1)
    dv = np.gradient(h.astype(np.float64), axis=0)
    nrm = np.stack([-du, np.ones_like(du), -dv], axis=-1)
    nrm = (nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)).reshape(-1, 3).astype(np.float32)
    uv = np.stack([gx / max(1, nx - 1), gz / max(1, nz - 1)], axis=-1).reshape(-1, 2).astype(np.float32)
    q = (np.arange(nz - 1)[:, None] * nx + np.arange(nx - 1)[None, :]).reshape(-1)
    a, b, c, d = q, q + 1, q + nx + 1, q + nx
    idx = np.stack([a, c, b, a, d, c], axis=-1).reshape(-1).astype(np.uint32)  # counter-clockwise from +Y
    positions, normals, indices, uvs, ids, albedo = [terrain_pos], [nrm], [idx], [uv], [1], [(0.35, 0.45, 0.3)]
    vcount, icount = [0, len(terrain_pos)], [0, len(idx)]
    local, tnrm, tidx, tuv = _TEMPLATE
    for k in range(buildings):
        ext = [rng.uniform(4.0, 20.0), rng.uniform(3.0, 30.0), rng.uniform(4.0, 20.0)]
        cx = rng.uniform(ext[0] / 2, nx - 1 - ext[0] / 2)
        cz = rng.uniform(ext[2] / 2, nz - 1 - ext[2] / 2)
        base = float(h[min(nz - 1, int(cz)), min(nx - 1, int(cx))]) - 1.0
        mn = np.array([cx - ext[0] / 2, base, cz - ext[2] / 2], np.float32)
        mx = np.array([cx + ext[0] / 2, base + ext[1], cz + ext[2] / 2], np.float32)
        pos = ((mn + mx) * np.float32(0.5) + local * (mx - mn) * np.float32(0.5)).astype(np.float32)
        positions.append(pos)
        normals.append(tnrm)
        indices.append(tidx)
        uvs.append(tuv)
        ids.append(2 + k)
        albedo.append((rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8)))
        vcount.append(vcount[-1] + len(pos))
        icount.append(icount[-1] + len(tidx))
    sun = sun_light(100_000, (0.45, 0.75, 0.5), 0.53)
    sun.illumination = ILLUMINATION_SINGLE_BOUNCE
    sun.bounce_samples = 32
    light_list = [sun, sky_light(100_001)]
    for k in range(area_lights):
        x, z = rng.uniform(0.0, nx - 1), rng.uniform(0.0, nz - 1)
        y = float(h[min(nz - 1, int(z)), min(nx - 1, int(x))]) + rng.uniform(1.0, 6.0)
        length = rng.uniform(0.5, 4.0)
        radius = rng.uniform(8.0, 20.0)
        if k % 2 == 0:
            light_list.append(area_light(200_000 + k, (x, y, z), radius, quat_from_rotation_x(-math.pi / 2),
                                         (length, rng.uniform(0.5, 4.0))))
        else:
            yaw = rng.uniform(0.0, math.pi)
            light_list.append(area_light(200_000 + k, (x, y, z), radius, (0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2)),
                                         (length, 0.0)))
    inp = BakeInput.from_arrays(
        ids=np.array(ids, np.uint64), positions=np.concatenate(positions), indices=np.concatenate(indices),
        uv1=np.concatenate(uvs), normals=np.concatenate(normals),
        vertex_offsets=np.array(vcount, np.int64), index_offsets=np.array(icount, np.int64),
        albedo=np.array(albedo, np.float32), content_hash=np.array([splitmix64(i) for i in ids], np.uint64),
        lights=light_list)
    opts = default_options()
    opts.pixel_density = density
    opts.area_light_samples = 16
    opts.area_light_min_samples = 16
    opts.sky_samples = 64
    opts.sun_map_resolution = 1024
    opts.sky_probe_spacing = 1.0
    opts.dynamic_geometry = DYNAMIC_ANGULAR
    return inp, opts


# ---- batch: many independent small scenes of uneven size -------------------------------------
def batch_job(j: int, seed: int = 4):
    """Returns (BakeInput, options) of job j of the batch workload (job seed splitmix64(seed ^ j)).

    A job is a small interior in the style of ``interior``: a row-major grid of R = max(1,
    round(T / 3000)) rooms of 6x3x6 m with 1x2 m doors between neighbours, filled with boxes and
    pillars (30%, full height) to a log-uniform triangle count T in [500, 5e4].  Lights: a
    log-uniform count in [1, 64] placed in random rooms, 40% point, 30% spot (baked as points),
    30% rectangle/tube area lights (0.5-2 m), range U(4, 12), single bounce for odd j.  The
    density is solved from a log-uniform texel target in [1e3, 2e6] with a 0.64 atlas fill:
    density = sqrt(target / (0.64 * total area))."""
    rng = Pcg32.seeded(splitmix64((seed ^ j) & MASK64))
    tris = int(round(math.exp(rng.uniform(math.log(500.0), math.log(5e4)))))
    nlights = min(64, int(math.exp(rng.uniform(0.0, math.log(65.0)))))
    texels = math.exp(rng.uniform(math.log(1e3), math.log(2e6)))
    W, H, D = 6.0, 3.0, 6.0
    rooms = max(1, int(round(tris / 3000.0)))
    gx = int(math.ceil(math.sqrt(rooms)))
    origin = [((r % gx) * W, (r // gx) * D) for r in range(rooms)]
    occupied = set((r % gx, r // gx) for r in range(rooms))
    positions, normals, indices, uvs = [], [], [], []
    area = 0.0
    for r in range(rooms):
        cx, cz = r % gx, r // gx
        doors = {"-x": (cx - 1, cz) in occupied, "+x": (cx + 1, cz) in occupied,
                 "-z": (cx, cz - 1) in occupied, "+z": (cx, cz + 1) in occupied}
        p, n, idx, uv = _room_shell(origin[r][0], origin[r][1], W, H, D, doors)
        positions.append(p)
        normals.append(n)
        indices.append(idx)
        uvs.append(uv)
        t = p[idx.reshape(-1, 3)].astype(np.float64)
        area += float(np.sum(np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)) / 2)
    shell_tris = sum(len(x) for x in indices) // 3
    boxes = max(1, (tris - shell_tris) // 12)
    u = rng.uniforms(boxes * 9).reshape(boxes, 9)
    room = np.minimum((u[:, 0] * rooms).astype(np.int64), rooms - 1)
    pillar = u[:, 1] < 0.3
    ext = np.where(pillar[:, None], np.stack([0.2 + 0.8 * u[:, 2], np.full(boxes, H), 0.2 + 0.8 * u[:, 4]], 1),
                   0.2 + 2.8 * u[:, 2:5])
    ox = np.array([o[0] for o in origin])[room]
    oz = np.array([o[1] for o in origin])[room]
    cxs = ox + ext[:, 0] / 2 + u[:, 5] * (W - ext[:, 0])
    czs = oz + ext[:, 2] / 2 + u[:, 6] * (D - ext[:, 2])
    mn = np.stack([cxs - ext[:, 0] / 2, np.zeros(boxes), czs - ext[:, 2] / 2], 1).astype(np.float32)
    mx = np.stack([cxs + ext[:, 0] / 2, ext[:, 1], czs + ext[:, 2] / 2], 1).astype(np.float32)
    albedo = np.concatenate([np.full((rooms, 3), 0.8, np.float32),
                             (0.1 + 0.8 * rng.uniforms(boxes * 3).reshape(boxes, 3)).astype(np.float32)])
    local, tnrm, tidx, tuv = _TEMPLATE
    box_pos = ((mn + mx)[:, None, :] * np.float32(0.5) + local[None] * (mx - mn)[:, None, :] * np.float32(0.5))
    area += float(np.sum(2 * (ext[:, 0] * ext[:, 1] + ext[:, 1] * ext[:, 2] + ext[:, 0] * ext[:, 2])))
    vcount = np.concatenate([[0], np.cumsum([len(x) for x in positions]),
                             sum(len(x) for x in positions) + len(local) * np.arange(1, boxes + 1)])
    icount = np.concatenate([[0], np.cumsum([len(x) for x in indices]),
                             sum(len(x) for x in indices) + len(tidx) * np.arange(1, boxes + 1)])
    positions.append(box_pos.reshape(-1, 3).astype(np.float32))
    normals.append(np.tile(tnrm, (boxes, 1)))
    indices.append(np.tile(tidx, boxes))
    uvs.append(np.tile(tuv, (boxes, 1)))
    ids = np.arange(1, rooms + boxes + 1, dtype=np.uint64)
    light_list = []
    for k in range(nlights):
        kind = rng.uniform()
        r = rng.below(rooms)
        p = (origin[r][0] + rng.uniform(0.1, W - 0.1), rng.uniform(0.1, H - 0.1), origin[r][1] + rng.uniform(0.1, D - 0.1))
        radius = rng.uniform(4.0, 12.0)
        if kind < 0.7:
            light = point_light(1000 + k, p, radius)
        else:
            length = rng.uniform(0.5, 2.0)
            if rng.uniform() < 0.5:
                light = area_light(1000 + k, p, radius, quat_from_rotation_x(-math.pi / 2), (length, rng.uniform(0.5, 2.0)))
            else:
                yaw = rng.uniform(0.0, math.pi)
                light = area_light(1000 + k, p, radius, (0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2)), (length, 0.0))
        if j % 2 == 1:
            light.illumination = ILLUMINATION_SINGLE_BOUNCE
        light_list.append(light)
    inp = BakeInput.from_arrays(
        ids=ids, positions=np.concatenate(positions), indices=np.concatenate(indices), uv1=np.concatenate(uvs),
        normals=np.concatenate(normals), vertex_offsets=vcount.astype(np.int64), index_offsets=icount.astype(np.int64),
        albedo=albedo, content_hash=np.array([splitmix64(int(i)) for i in ids], np.uint64), lights=light_list)
    opts = default_options()
    opts.pixel_density = float(math.sqrt(texels / (0.64 * area)))
    opts.dynamic_geometry = DYNAMIC_ANGULAR
    return inp, opts


def batch(count: int = 512, seed: int = 4):
    """The batch workload: `count` independent jobs (lists of BakeInput and options)."""
    jobs = [batch_job(j, seed) for j in range(count)]
    return [j[0] for j in jobs], [j[1] for j in jobs]


def service_room(k: int, seed: int = 6, origin=None) -> BakeInput:
    """Room k of the serve workload's room requests: one room of the batch workload's generator
    (``batch_job``), a closed 6x3x6 m shell filled with boxes and pillars to a log-uniform
    triangle count in [500, 3000] (a batch room's share), a log-uniform light count in [1, 64]
    (40% point, 30% spot baked as point, 30% rectangle/tube), range U(4, 12), single bounce for
    odd k.  It stands 100 + 80k m along -x from the interior, so that no light or burial ray of
    one reaches the other; mesh ids from 10^7 + 10^5 k, light ids from 2 * 10^7 + 10^4 k (above
    the interior's and every earlier room's).  `origin` (x, z) places the room elsewhere."""
    rng = Pcg32.seeded(splitmix64((seed ^ (k * 0x9E3779B1)) & MASK64))
    tris = int(round(math.exp(rng.uniform(math.log(500.0), math.log(3000.0)))))
    nlights = min(64, int(math.exp(rng.uniform(0.0, math.log(65.0)))))
    W, H, D = 6.0, 3.0, 6.0
    ox, oz = (-100.0 - 80.0 * k, 0.0) if origin is None else (float(origin[0]), float(origin[1]))
    p, n, idx, uv = _room_shell(ox, oz, W, H, D, {})
    positions, normals, indices, uvs = [p], [n], [idx], [uv]
    boxes = max(1, (tris - len(idx) // 3) // 12)
    u = rng.uniforms(boxes * 9).reshape(boxes, 9)
    pillar = u[:, 1] < 0.3
    ext = np.where(pillar[:, None], np.stack([0.2 + 0.8 * u[:, 2], np.full(boxes, H), 0.2 + 0.8 * u[:, 4]], 1),
                   0.2 + 2.8 * u[:, 2:5])
    cxs = ox + ext[:, 0] / 2 + u[:, 5] * (W - ext[:, 0])
    czs = oz + ext[:, 2] / 2 + u[:, 6] * (D - ext[:, 2])
    mn = np.stack([cxs - ext[:, 0] / 2, np.zeros(boxes), czs - ext[:, 2] / 2], 1).astype(np.float32)
    mx = np.stack([cxs + ext[:, 0] / 2, ext[:, 1], czs + ext[:, 2] / 2], 1).astype(np.float32)
    albedo = np.concatenate([np.full((1, 3), 0.8, np.float32),
                             (0.1 + 0.8 * rng.uniforms(boxes * 3).reshape(boxes, 3)).astype(np.float32)])
    local, tnrm, tidx, tuv = _TEMPLATE
    box_pos = ((mn + mx)[:, None, :] * np.float32(0.5) + local[None] * (mx - mn)[:, None, :] * np.float32(0.5))
    vcount = np.concatenate([[0, len(p)], len(p) + len(local) * np.arange(1, boxes + 1)])
    icount = np.concatenate([[0, len(idx)], len(idx) + len(tidx) * np.arange(1, boxes + 1)])
    positions.append(box_pos.reshape(-1, 3).astype(np.float32))
    normals.append(np.tile(tnrm, (boxes, 1)))
    indices.append(np.tile(tidx, boxes))
    uvs.append(np.tile(tuv, (boxes, 1)))
    ids = np.arange(10_000_000 + 100_000 * k, 10_000_000 + 100_000 * k + 1 + boxes, dtype=np.uint64)
    light_list = []
    for q in range(nlights):
        kind = rng.uniform()
        lp = (ox + rng.uniform(0.1, W - 0.1), rng.uniform(0.1, H - 0.1), oz + rng.uniform(0.1, D - 0.1))
        radius = rng.uniform(4.0, 12.0)
        lid = 20_000_000 + 10_000 * k + q
        if kind < 0.7:
            light = point_light(lid, lp, radius)
        else:
            length = rng.uniform(0.5, 2.0)
            if rng.uniform() < 0.5:
                light = area_light(lid, lp, radius, quat_from_rotation_x(-math.pi / 2), (length, rng.uniform(0.5, 2.0)))
            else:
                yaw = rng.uniform(0.0, math.pi)
                light = area_light(lid, lp, radius, (0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2)), (length, 0.0))
        if k % 2 == 1:
            light.illumination = ILLUMINATION_SINGLE_BOUNCE
        light_list.append(light)
    return BakeInput.from_arrays(
        ids=ids, positions=np.concatenate(positions), indices=np.concatenate(indices), uv1=np.concatenate(uvs),
        normals=np.concatenate(normals), vertex_offsets=vcount.astype(np.int64), index_offsets=icount.astype(np.int64),
        albedo=albedo, content_hash=np.array([splitmix64(int(i)) for i in ids], np.uint64), lights=light_list)


# ---- serve: rebake requests against the resident interior -------------------------------------
def serve_requests(inp: BakeInput, count: int = 2000, seed: int = 5, rate: float = 20.0, grid: int = 8,
                   room=(6.0, 3.0, 6.0), room_share: float = 0.2):
    """Requests arriving as a Poisson process of `rate` per second against the interior scene.
    A share 1 - room_share are light requests: with equal odds, move 1-4 existing lights (same ids,
    displaced U(0.5, 3) m horizontally) or add 1-4 new lights from the interior's distribution
    (a random room, range U(4, 12), one bounce).  The rest add a room of geometry; they are
    returned with kind "room" and no lights.  Returns a list of (kind, lights, arrival_s)."""
    rng = Pcg32.seeded(seed)
    W, H, D = room
    t = 0.0
    out = []
    next_id = 1_000_000
    for r in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        if rng.uniform() < room_share:
            out.append(("room", [], t))
            continue
        k = 1 + rng.below(4)
        lights = []
        if rng.uniform() < 0.5:
            picked = []
            while len(picked) < k:
                i = rng.below(len(inp.lights))
                if i not in picked:
                    picked.append(i)
            for i in sorted(picked):
                base = inp.lights[i]
                d, a = rng.uniform(0.5, 3.0), rng.uniform(0.0, 2.0 * math.pi)
                x = min(grid * W - 0.1, max(0.1, base.position[0] + d * math.cos(a)))
                z = min(grid * D - 0.1, max(0.1, base.position[2] + d * math.sin(a)))
                lights.append(point_light(base.id, (x, base.position[1], z), base.radius, bounce=True))
            out.append(("move", lights, t))
        else:
            for _ in range(k):
                room_i = rng.below(grid * grid)
                rx, rz = (room_i % grid) * W, (room_i // grid) * D
                p = (rx + rng.uniform(0.1, W - 0.1), rng.uniform(0.1, H - 0.1), rz + rng.uniform(0.1, D - 0.1))
                lights.append(point_light(next_id, p, rng.uniform(4.0, 12.0), bounce=True))
                next_id += 1
            out.append(("add", lights, t))
    return out


def mesh_sample_mask(ids: np.ndarray, modulus: int) -> np.ndarray:
    """Deterministic receiver subset: meshes with splitmix64(id) % modulus == 0."""
    return np.array([splitmix64(int(i)) % modulus == 0 for i in ids], dtype=bool)
