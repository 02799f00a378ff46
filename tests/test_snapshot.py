"""The ECS caller's conversions behind the C ABI against the reference's own tests:
mesh_content_hash (src/shared/mesh_hash.rs:113-144, including its pinned value), BakeId::from_path
(src/shared/bake_id.rs:56-90, FNV-1a checked against its published vectors), assign_ids
(src/bake/identity.rs:268-289) and MeshData::to_world (src/bake/snapshot.rs:299-318)."""
import numpy as np

import dynlight_hip
import oracle
from dynlight_hip import abi, scenes
from helpers import assert_scenes_equal

POSITIONS = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
UVS = [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]


# ---- src/shared/mesh_hash.rs:113-144 -------------------------------------------------------------
def test_hash_is_stable_and_pinned():
    a = dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 2], UVS)
    assert a == dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 2], UVS)
    assert a == 11_458_312_689_423_647_656, "the reference's pinned value"


def test_every_input_changes_the_hash():
    base = dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 2], UVS)
    moved = [list(p) for p in POSITIONS]
    moved[1][0] = 1.0001
    assert base != dynlight_hip.mesh_content_hash(moved, [0, 1, 2], UVS)
    assert base != dynlight_hip.mesh_content_hash(POSITIONS, [0, 2, 1], UVS)
    unwrapped = [list(u) for u in UVS]
    unwrapped[2][1] = 0.5
    assert base != dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 2], unwrapped)


def test_only_the_triangles_count():
    base = dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 2], UVS)
    assert base == dynlight_hip.mesh_content_hash(POSITIONS + [[5.0, 5.0, 5.0]], [0, 1, 2], UVS + [[0.5, 0.5]])
    order = [0, 1, 2, 2, 1, 0]
    shared = dynlight_hip.mesh_content_hash(POSITIONS, order, UVS)
    assert shared == dynlight_hip.mesh_content_hash([POSITIONS[i] for i in order], range(6), [UVS[i] for i in order])


def test_missing_vertices_are_hashed_apart():
    assert dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 3], UVS) != dynlight_hip.mesh_content_hash(POSITIONS, [0, 1, 2], UVS)
    assert dynlight_hip.mesh_content_hash([], [], []) != dynlight_hip.mesh_content_hash([], [0], [])


# ---- src/shared/bake_id.rs:56-90 -----------------------------------------------------------------
def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv_matches_published_test_vectors_and_paths_are_length_prefixed():
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8
    for path in (["Level", "Hall", "Lamp"], ["a"], ["ab", "c"], ["", "x"]):
        want = fnv1a64(b"".join(len(s).to_bytes(8, "little") + s.encode() for s in path))
        assert dynlight_hip.bake_id_from_path(path) == want


def test_paths_are_stable_and_unambiguous():
    lamp = dynlight_hip.bake_id_from_path(["Level", "Hall", "Lamp"])
    assert lamp == dynlight_hip.bake_id_from_path(["Level", "Hall", "Lamp"])
    assert lamp != dynlight_hip.bake_id_from_path(["Level", "Hall", "Lamp2"])
    assert dynlight_hip.bake_id_from_path(["ab", "c"]) != dynlight_hip.bake_id_from_path(["a", "bc"])
    assert dynlight_hip.bake_id_from_path([]) == 0xCBF29CE484222325


# ---- src/bake/identity.rs:268-289 ----------------------------------------------------------------
def test_derived_ids_are_disambiguated_in_tiebreak_order():
    ids = dynlight_hip.assign_ids([(None, ["Lamp"], 0, (2.0, 0, 0)), (7, ["Lamp"], 0, (0.0, 0, 0)),
                                   (None, ["Lamp"], 0, (1.0, 0, 0)), (None, ["Desk"], 0, (0.0, 0, 0))])
    lamp = dynlight_hip.bake_id_from_path(["Lamp"])
    assert ids[1] == 7
    assert ids[2] == lamp, "the first in tiebreak order keeps the plain id"
    assert ids[0] == dynlight_hip.bake_id_from_path(["Lamp", "#1"])
    assert ids[3] == dynlight_hip.bake_id_from_path(["Desk"])


def test_mesh_ties_break_on_content_hash_before_position():
    ids = dynlight_hip.assign_ids([(None, ["Wall"], 9, (0.0, 0, 0)), (None, ["Wall"], 3, (5.0, 0, 0)),
                                   (None, ["Wall"], 3, (-1.0, 0, 0))])
    # hash 3 before 9; between the two hash-3 walls, position bits: 5.0 (0x40a00000) < -1.0 (0xbf800000)
    assert ids[1] == dynlight_hip.bake_id_from_path(["Wall"])
    assert ids[2] == dynlight_hip.bake_id_from_path(["Wall", "#1"])
    assert ids[0] == dynlight_hip.bake_id_from_path(["Wall", "#2"])


# ---- src/bake/snapshot.rs:299-318 ----------------------------------------------------------------
def test_mirrored_transforms_keep_front_faces():
    local = [[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]
    normals = [[0.0, 1.0, 0.0]] * 3
    uv = [[0.0, 0.0], [0.0, 1.0], [1.0, 0.0]]
    mirror = [1, 0, 0, 0, -2, 0, 0, 0, 1, 0, 0, 0]  # Transform::from_scale((1, -2, 1))
    mesh = dynlight_hip.mesh_to_world(1, mirror, local, [0, 1, 2], uv, normals)
    assert list(mesh.indices) == [0, 2, 1]
    assert np.allclose(mesh.normals[0], [0, -1, 0], atol=1e-6)
    a, b, c = (mesh.positions[i] for i in mesh.indices)
    face = np.cross(b - a, c - a)
    assert np.allclose(face / np.linalg.norm(face), [0, -1, 0], atol=1e-6)
    identity = dynlight_hip.mesh_to_world(1, [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], local, [0, 1, 2], uv, normals)
    assert mesh.content_hash == identity.content_hash, "hashes local data"


def test_singular_transforms_drop_normals_and_normals_use_the_inverse_transpose():
    local = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]
    n = [[1.0, 1.0, 0.0]] * 3
    flat = dynlight_hip.mesh_to_world(1, [1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0], local, [0, 1, 2], UVS, n)
    assert flat.normals is None, "a singular transform bakes flat normals"
    # a shear (x += y): the normal (1, 1, 0) maps to the inverse transpose (1, 0, 0) direction
    shear = [1, 0, 0, 1, 1, 0, 0, 0, 1, 3, 4, 5]
    m = dynlight_hip.mesh_to_world(1, shear, local, [0, 1, 2], UVS, n)
    assert np.allclose(m.normals[0], [1, 0, 0], atol=1e-6)
    assert np.allclose(m.positions[1], [4, 4, 5]), "points take the translation"


def test_a_transformed_scene_bakes_like_the_same_scene_built_in_place():
    """snapshot -> BakeInput -> bake: a box moved by a translation bakes like the box built there."""
    box = scenes.cuboid(2, (-0.5, 0.0, -0.5), (0.5, 1.0, 0.5))
    moved = dynlight_hip.mesh_to_world(2, [1, 0, 0, 0, 1, 0, 0, 0, 1, 1.5, 0.0, -0.5], box.positions, box.indices, box.uv1,
                                       box.normals)
    built = scenes.cuboid(2, (1.0, 0.0, -1.0), (2.0, 1.0, 0.0))
    assert np.array_equal(moved.positions, built.positions) and np.array_equal(moved.indices, built.indices)
    light = scenes.point_light(5, (1.0, 3.0, 0.0), 8.0)
    floor = scenes.floor(1, (0, 0, 0), 4.0)
    o = abi.default_options()
    o.pixel_density = 12.0
    moved.content_hash = built.content_hash
    assert_scenes_equal(oracle.bake(abi.BakeInput([floor, moved], [light]), o),
                        oracle.bake(abi.BakeInput([floor, built], [light]), o))
