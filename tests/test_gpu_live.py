"""The live scene of a rebake service (dl_service_apply): the service starts from a bake of its
scene and applies light edits (moved, added, changed, removed lights) by tracing only the changed
lights; after every edit the live scene must equal dl_bake of the resident meshes with the edited
light set (and, on the small scene, the oracle's bake): words, lights, light BVH, distance cubes
and stats."""
import dataclasses

import numpy as np
import pytest

import oracle
from dynlight_hip import BakeError, Service, abi, scenes
from helpers import assert_scenes_equal

pytestmark = pytest.mark.gpu


def moved(light, dx, dy=0.0, dz=0.0, **kw):
    x, y, z = light.position
    return dataclasses.replace(light, position=(x + dx, y + dy, z + dz), **kw)


def small():
    return scenes.interior(seed=9, target_triangles=3000, lights=12, grid=2, density=12.0)


def test_the_live_scene_starts_as_the_bake(ctx):
    inp, o = small()
    svc = Service(ctx, inp, o)
    try:
        want = ctx.bake(inp, o)
        assert_scenes_equal(svc.initial, want)
        assert_scenes_equal(svc.scene(), want)
        ptr, n = svc.words()
        assert ptr != 0 and n == sum(len(m.triangles) for m in want.meshes)
    finally:
        svc.close()


def test_edits_equal_a_bake_of_the_edited_light_set(ctx):
    inp, o = small()
    lights = list(inp.lights)
    svc = Service(ctx, inp, o)
    try:
        a, b, c = lights[0], lights[3], lights[7]
        new = scenes.point_light(90_001, (4.0, 1.5, 2.0), 6.0, bounce=True)
        edits = [
            ("move one light", [moved(a, 1.0, 0.0, 0.5)], []),
            ("add a light", [new], []),
            ("move two, one across rooms", [moved(b, 0.5, 0.2, -0.3), moved(c, -4.0, 0.0, 0.0)], []),
            ("change a radius and a bounce setting", [dataclasses.replace(lights[5], radius=3.0, bounce_samples=8)], []),
            ("remove a light", [], [lights[2].id]),
            ("move the added light and remove another", [moved(new, -1.5, 0.0, 1.0)], [lights[9].id]),
        ]
        current = {l.id: l for l in lights}
        for step, (name, changed, removed) in enumerate(edits):
            for l in changed:
                current[l.id] = l
            for i in removed:
                del current[i]
            info = svc.apply(changed, removed)
            print(f"edit {step} ({name}): {info['dirty_meshes']} dirty meshes, {info['traced_pairs']} traced pairs, "
                  f"{info['total_ms']:.1f} ms", flush=True)
            assert info["version"] == step + 1
            got = svc.scene()
            now = inp.with_lights(sorted(current.values(), key=lambda l: l.id))
            assert_scenes_equal(got, ctx.bake(now, o))
            if step in (0, len(edits) - 1):
                assert_scenes_equal(got, oracle.bake(now, o, threads=8))
    finally:
        svc.close()


def test_edit_errors(ctx):
    inp, o = small()
    svc = Service(ctx, inp, o)
    try:
        with pytest.raises(BakeError) as e:
            svc.apply([scenes.sun_light(5, (0.0, 1.0, 0.0), 0.5)])
        assert e.value.code == abi.DL_E_ARGUMENT
        with pytest.raises(BakeError) as e:
            svc.apply([], [123456789])
        assert e.value.code == abi.DL_E_ARGUMENT
        with pytest.raises(BakeError) as e:
            svc.apply([inp.lights[0], inp.lights[0]])
        assert e.value.code == abi.DL_E_DUPLICATE_ID
        bad = dataclasses.replace(inp.lights[1], radius=-1.0)
        with pytest.raises(BakeError) as e:
            svc.apply([bad])
        assert e.value.code == abi.DL_E_INVALID_LIGHT
        # a failed edit leaves the live scene as it was
        assert_scenes_equal(svc.scene(), ctx.bake(inp, o))
    finally:
        svc.close()


def test_an_edit_of_the_full_interior_equals_its_bake(ctx):
    inp, o = scenes.interior(2)
    svc = Service(ctx, inp, o)
    try:
        lights = list(inp.lights)
        rng = np.random.default_rng(5)
        changed = [moved(lights[int(k)], *rng.uniform(-2.0, 2.0, 3) * np.array([1, 0.2, 1])) for k in rng.choice(256, 4, replace=False)]
        info = svc.apply(changed)
        print(f"interior edit: {info['dirty_meshes']} dirty meshes, {info['traced_pairs']} traced pairs, "
              f"trace {info['trace_ms']:.1f} ms, assemble {info['assemble_ms']:.1f} ms", flush=True)
        current = {l.id: l for l in lights}
        for l in changed:
            current[l.id] = l
        assert_scenes_equal(svc.scene(), ctx.bake(inp.with_lights(sorted(current.values(), key=lambda l: l.id)), o))
    finally:
        svc.close()


def _edit_stream(lights, n, seed):
    """n edits of a light set: moves of one or two lights, a bounce change, removals and additions,
    each (changed lights, removed ids); returns the edits and the light set after each."""
    rng = np.random.default_rng(seed)
    current = {l.id: l for l in lights}
    edits, after = [], []
    for k in range(n):
        ids = sorted(current)
        kind = k % 5
        if kind == 3 and len(ids) > 4:
            gone = int(ids[int(rng.integers(len(ids)))])
            if current[gone].shape in (abi.SHAPE_SUN, abi.SHAPE_SKY):
                kind = 0
            else:
                del current[gone]
                edits.append(([], [gone]))
                after.append(dict(current))
                continue
        if kind == 4:
            new = scenes.point_light(80_000 + k, tuple(rng.uniform([0.5, 0.5, 0.5], [7.5, 2.5, 7.5])), 5.0, bounce=bool(k % 2))
            current[new.id] = new
            edits.append(([new], []))
            after.append(dict(current))
            continue
        local = [i for i in ids if current[i].shape not in (abi.SHAPE_SUN, abi.SHAPE_SKY)]
        pick = rng.choice(local, 1 + (kind == 1), replace=False)
        changed = []
        for i in pick:
            l = moved(current[int(i)], *rng.uniform(-1.0, 1.0, 3) * np.array([1, 0.2, 1]))
            if kind == 2:
                l = dataclasses.replace(l, bounce_samples=int(rng.integers(4, 16)))
            current[l.id] = l
            changed.append(l)
        edits.append((changed, []))
        after.append(dict(current))
    return edits, after


def test_replayed_edits_overlap_and_match_edits_one_at_a_time(ctx):
    """Continuous batching of edits (dl_service_replay_edits): edits arriving together trace on
    several streams at once and join the live scene together, yet wherever the mix forms the live
    scene it equals the scene after the same edits applied one at a time (edits far apart, one
    stream, a budget that admits one edit at a time), and the last one equals a bake."""
    inp, o = small()
    edits, after = _edit_stream(list(inp.lights), 10, seed=3)
    one = Service(ctx, inp, o)
    try:
        spaced = [(c, r, 0.15 * i) for i, (c, r) in enumerate(edits)]
        lat1, up1, sum1 = one.replay_edits(spaced, workers=1, budget_bytes=1, checksums=True)
    finally:
        one.close()
    assert all(sum1), "edits far apart each form the live scene"
    assert [u["reserved"] for u in up1] == [1] * len(edits)
    many = Service(ctx, inp, o)
    try:
        together = [(c, r, 0.0) for c, r in edits]
        lat4, up4, sum4 = many.replay_edits(together, workers=4, checksums=True)
        print("latency ms, one at a time:", [round(x, 1) for x in lat1], flush=True)
        print("latency ms, all at once on four streams:", [round(x, 1) for x in lat4], flush=True)
        print("joint assemblies:", [u["reserved"] for u in up4], flush=True)
        formed = [k for k, x in enumerate(sum4) if x]
        assert formed and formed[-1] == len(edits) - 1
        assert all(sum4[k] == sum1[k] for k in formed)
        assert [u["version"] for u in up4] == list(range(1, len(edits) + 1))
        assert all(x > 0 for x in lat4)
        final = inp.with_lights(sorted(after[-1].values(), key=lambda l: l.id))
        assert_scenes_equal(many.scene(), ctx.bake(final, o))
    finally:
        many.close()
    # the first edit applied alone gives the fingerprint recorded after it
    alone = Service(ctx, inp, o)
    try:
        c, r = edits[0]
        _, _, s = alone.replay_edits([(c, r, 0.0)], workers=1, checksums=True)
        assert s[0] == sum1[0]
        alone.apply(*edits[1])
        _, _, s = alone.replay_edits([(edits[2][0], edits[2][1], 0.0)], workers=2, checksums=True)
        assert s[0] == sum1[2]
    finally:
        alone.close()


def test_replay_errors_keep_the_edits_before_them(ctx):
    inp, o = small()
    lights = list(inp.lights)
    svc = Service(ctx, inp, o)
    try:
        with pytest.raises(BakeError) as e:
            svc.replay_edits([([moved(lights[0], 0.5)], [], 1.0), ([moved(lights[1], 0.5)], [], 0.5)])
        assert e.value.code == abi.DL_E_ARGUMENT
        with pytest.raises(BakeError) as e:
            svc.replay_edits([([moved(lights[0], 0.5)], [], 0.0)], workers=0)
        assert e.value.code == abi.DL_E_ARGUMENT
        first = moved(lights[0], 0.5)
        with pytest.raises(BakeError) as e:
            svc.replay_edits([([first], [], 0.0), ([], [987654321], 0.0), ([moved(lights[1], 0.5)], [], 0.0)], workers=3)
        assert e.value.code == abi.DL_E_ARGUMENT
        current = {l.id: l for l in lights}
        current[first.id] = first
        assert_scenes_equal(svc.scene(), ctx.bake(inp.with_lights(sorted(current.values(), key=lambda l: l.id)), o))
    finally:
        svc.close()


def test_added_rooms_join_the_live_scene_as_a_bake_of_the_whole(ctx):
    """Room requests: a room (a scene of its own, isolated from the live scene) joins it; the live
    scene then equals the bake of the whole scene (the oracle's too, for the first room), and
    light edits that follow, in the room and in the first scene, keep it so."""
    inp, o = small()
    svc = Service(ctx, inp, o)
    try:
        room0, room1 = scenes.service_room(0), scenes.service_room(1)
        info = svc.add_room(room0)
        print(f"room 0: {room0.mesh_count} meshes, {len(room0.lights)} lights, {info['total_ms']:.1f} ms", flush=True)
        whole = inp.joined(room0)
        assert_scenes_equal(svc.scene(), ctx.bake(whole, o))
        assert_scenes_equal(svc.scene(), oracle.bake(whole, o, threads=8))
        info = svc.add_room(room1)
        whole = whole.joined(room1)
        assert_scenes_equal(svc.scene(), ctx.bake(whole, o))
        # edits after the rooms: a room light moved, an interior light moved, a light added in a room
        lights = {l.id: l for l in whole.lights}
        a = moved(room0.lights[0], 0.5, 0.0, 0.3)
        b = moved(inp.lights[2], -0.7, 0.0, 0.4)
        c = scenes.point_light(20_000_900, (-100.0 - 80.0 + 3.0, 1.5, 3.0), 5.0, bounce=True)
        svc.apply([a, b, c])
        for l in (a, b, c):
            lights[l.id] = l
        assert_scenes_equal(svc.scene(), ctx.bake(whole.with_lights(sorted(lights.values(), key=lambda l: l.id)), o))
    finally:
        svc.close()


def test_rooms_that_are_not_isolated_are_refused(ctx):
    inp, o = small()
    svc = Service(ctx, inp, o)
    try:
        before = svc.scene()
        # a room inside the live scene's bounds, and one whose light reaches the live scene
        for origin in ((3.0, 3.0), (14.0, 0.0)):  # inside the live bounds; its lights reach them
            with pytest.raises(BakeError) as e:
                svc.add_room(scenes.service_room(0, origin=origin))
            assert e.value.code == abi.DL_E_ARGUMENT
        near = scenes.service_room(0)
        # mesh ids not above the live ones
        low = scenes.box_room(3)[0]
        with pytest.raises(BakeError) as e:
            svc.add_room(low)
        assert e.value.code == abi.DL_E_ARGUMENT
        assert_scenes_equal(svc.scene(), before)
        # a light edit that reaches two parts is refused once a room is there
        svc.add_room(near)
        far = scenes.point_light(70_001, (-60.0, 1.5, 3.0), 80.0, bounce=False)
        with pytest.raises(BakeError) as e:
            svc.apply([far])
        assert e.value.code == abi.DL_E_ARGUMENT
    finally:
        svc.close()
