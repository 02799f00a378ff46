"""Benchmark: texel-light visibility samples/s of the MI355X bake on the procedural interior.

One step is one full bake (dl_bake: BVH build, association, texels, burial, shadow rays,
post-passes, serialized words, distance cubes) of the interior scene: an 8x8 grid of 6x3x6 m
rooms with door openings, seeded boxes and pillars to 200k triangles, 256 point/spot lights of
range 4-12 m, 32 texels/m, 1-bit shadows and one colored bounce from every light.  A sample is one (texel, light) pair for which the
reference's texel_shadow returns Some (src/bake/mesh_bake.rs:241-247).

Prints one JSON line.  `roofline` prices the dominant stage, the bounce gather's sample setup (per
chunk of gather jobs: the fast pass k_gather_split, the list of the samples it leaves open, the
exact pass k_gather_open), at its algorithmic bytes: 48 B per job (job record, texel position and
coordinates, ray offset, traced count and bits), 40 B per open sample, 16 B per photon read, 44 B
per traced sample (ray and contribution written) and 64 B per photon tile of each bounce light,
with the counts from an instrumented (untimed) bake, divided by the span's average time from HIP
events on the bake stream.  `traffic` is the rocprofv3-measured HBM bytes of the two kernels per
launch recorded in profiles/interior_bake_kernels.md; `valu_roofline` their VALU instructions per
launch (rocprofv3 SQ_INSTS_VALU, also recorded there) over the same span against the issue peak,
the bound these kernels actually meet.  `other_kernels` gives the any-hit traversal of the
gather rays (k_any_stream: per ray, 112 B per 4-wide BVH node read plus 48 B per triangle tested
plus its 32-byte ray record and 4-byte list entry) and the direct shadow-ray kernel (k_trace) as
work rates counted per use of each record (mostly L2-served, so not HBM figures), beside the HBM
bytes rocprofv3 measured for them.  `cpu_baseline` times the CPU oracle on a 1/64 receiver sample
of the same scene (every mesh still occludes) and extrapolates to the whole bake, and records the
host CPU.  `batch` runs the 512-job batch workload in child processes at 4 (HIP's default) and 16
hardware queues, and `serve` the rebake service at 20 requests/s (see --workload).
"""
import argparse
import json
import os
import sys
import time

# HIP spreads a process's streams over GPU_MAX_HW_QUEUES hardware queues (4 unless the
# environment sets it before HIP starts).  The batch and serve workloads report the count they ran
# with; the default interior run measures the batch at 4 and at 16 queues in child processes.
HW_QUEUES = os.environ.get("GPU_MAX_HW_QUEUES", "4 (HIP default)")

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "dynamiclighting-bevy_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
# VALU issue peak: one wave64 VALU instruction per cycle per CU (4 SIMDs of 16 lanes), 256 CUs at
# 2.4 GHz (the same rate gives MI355X's 78.6 TFLOP/s of FP64 FMA)
VALU_PEAK_GIPS = 256 * 2.4
TILE_BYTES = 64.0    # one photon tile: 16 B of lit bits (9x9 texels) and a 48-B box
NODE_BYTES = 112.0  # a 4-wide node as the kernels read it: six float4 of child bounds, one int4 of child slots
TRI_BYTES = 48.0    # v0 | occluder index, e1, e2 as three float4
RAY_BYTES = 32.0    # origin | length, direction | flags
JOB_BYTES = 48.0    # gather job: (texel, entry) pair, texel position and coordinates, ray offset, traced count and bits
PHOTON_BYTES = 16.0  # one photon record, read by the samples the photon tiles leave open
TRACED_BYTES = 44.0  # ray record and RGB contribution written per traced sample
OPEN_BYTES = 40.0    # per sample left to the exact pass: its draws written and read (32 B), its list entry (8 B)


PROFILE = os.path.join(ROOT, "profiles", "interior_bake_kernels.md")


def profile_figure(tag, kernel, path=PROFILE):
    """A per-dispatch figure recorded in the profile summary as a line `TAG kernel value`: TRAFFIC is
    the HBM bytes per dispatch from rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by a plain-copy
    calibration), VALU the VALU instructions per dispatch (SQ_INSTS_VALU); None when absent."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == tag and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel):
    return profile_figure("TRAFFIC", kernel)


def host_cpu():
    """The CPU the baseline ran on: model, sockets, cores and threads (lscpu), and the threads this
    process may use."""
    import subprocess
    info = {"threads_used": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and keys[k.strip()] not in info:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def other_kernel(name, alg_bytes, launch_s, launches, extra):
    """A traversal kernel's algorithmic bytes per launch over its HIP-event launch time, next to the
    HBM bytes per launch rocprofv3 measured for it."""
    traffic = measured_traffic(name)
    alg = alg_bytes / launch_s / 1e9 if launch_s > 0 else 0.0
    hbm = traffic / launch_s / 1e9 if (traffic and launch_s > 0) else None
    return {"kernel": name, "work_rate_GBps": alg, "work_rate_note": "bytes counted per use; mostly served from L2",
            "alg_bytes_per_launch": alg_bytes, "launch_ms": launch_s * 1000.0, "launches_per_step": launches,
            "hbm_bytes_per_launch": traffic, "hbm_GBps": hbm, "hbm_frac": hbm / HBM_PEAK_GBS if hbm else None, **extra}


def child_run(argv, env_extra, timeout):
    """Runs bench.py with `argv` in a child process (its progress goes to our stderr) and returns the
    JSON line it prints, or None when it fails."""
    import subprocess
    env = dict(os.environ)
    env.update(env_extra)
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__)] + argv, env=env, stdout=subprocess.PIPE, text=True,
                           timeout=timeout)
    except subprocess.TimeoutExpired:
        log(f"child {' '.join(argv)} timed out")
        return None
    if r.returncode != 0:
        log(f"child {' '.join(argv)} exited with {r.returncode}")
        return None
    for line in reversed(r.stdout.strip().splitlines()):
        if line.startswith("{"):
            return json.loads(line)
    return None


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(inp, opts, full_samples, modulus=64, cube_lights=2):
    """The CPU oracle on a bounded sample, extrapolated to the whole bake.

    Meshes: receivers with splitmix64(id) % modulus == 0 baked against every occluder and light
    (photon cubes at 64^2 there: the gather's cost does not depend on the cube resolution).
    Photon cubes: `cube_lights` lights at the full resolution.  Whole-bake CPU time =
    mesh time x (full samples / sample samples) + cube time x (bounce lights / cube_lights).
    """
    import numpy as np

    import oracle
    from dynlight_hip import abi, scenes

    threads = len(os.sched_getaffinity(0))
    mask = scenes.mesh_sample_mask(inp.mesh_array["id"], modulus)
    o = abi.dl_bake_options.from_buffer_copy(opts)
    o.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
    o.photon_cube_resolution = 64
    s = oracle.bake_receivers(inp, o, mask, threads=threads)
    meshes_s = oracle.last_timings()["meshes_ms"] / 1000.0
    # the cubes of the `cube_lights` lights nearest the first room shell, at full resolution (the
    # oracle builds the cubes of lights that reach a sampled mesh: the shell is the sample)
    shell = inp._positions[inp.vertex_offsets[0]:inp.vertex_offsets[1]]
    centre = (shell.min(axis=0) + shell.max(axis=0)) * 0.5
    near = sorted(inp.lights, key=lambda l: float(np.sum((np.asarray(l.position) - centre) ** 2)))[:cube_lights]
    few = abi.BakeInput.from_arrays(
        ids=inp.mesh_array["id"], positions=inp._positions, indices=inp._indices, uv1=inp._uv1, normals=inp._normals,
        vertex_offsets=inp.vertex_offsets, index_offsets=inp.index_offsets, albedo=inp.mesh_array["albedo"],
        lights=near)
    o2 = abi.dl_bake_options.from_buffer_copy(opts)
    o2.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
    first = np.zeros(inp.mesh_count, bool)
    first[0] = True
    oracle.bake_receivers(few, o2, first, threads=threads)
    cubes_s = oracle.last_timings()["photon_cubes_ms"] / 1000.0
    bounce_lights = sum(1 for l in inp.lights if l.illumination == abi.ILLUMINATION_SINGLE_BOUNCE)
    est = meshes_s * full_samples / max(1, s.stats["samples"]) + cubes_s * bounce_lights / cube_lights
    return {"value": full_samples / est, "unit": "samples/s", "cores": threads, "kind": "port", "cpu": host_cpu(),
            "sample": f"CPU oracle on {threads} threads: receivers splitmix64(id) % {modulus} == 0 "
                      f"({int(np.sum(mask))} of {len(mask)} meshes, {s.stats['samples']} samples, "
                      f"{s.stats['bounce_rays']} bounce rays) against all occluders and lights in "
                      f"{meshes_s:.1f} s, plus {cube_lights} photon cubes at full resolution in {cubes_s:.1f} s; "
                      f"extrapolated whole bake {est:.0f} s"}


def batch_cpu_baseline(ins, ops, stride=65, budget_s=25.0):
    """The CPU oracle on every `stride`-th batch job (alternating bounce and no bounce) until
    `budget_s` seconds of bake time have passed; samples/s over the jobs baked."""
    import oracle

    threads = len(os.sched_getaffinity(0))
    done, samples, spent = [], 0, 0.0
    for j in range(0, len(ins), stride):
        t0 = time.perf_counter()
        s = oracle.bake(ins[j], ops[j], threads=threads)
        spent += time.perf_counter() - t0
        samples += s.stats["samples"]
        done.append(j)
        if spent > budget_s:
            break
    return {"value": samples / spent, "unit": "samples/s", "cores": threads, "kind": "port", "cpu": host_cpu(),
            "sample": f"CPU oracle on {threads} threads, whole bakes of jobs {done} ({samples} samples in {spent:.1f} s)"}


def batch_main(args):
    """Batch workload: 512 independent rooms (log-uniform 500-5e4 triangles, 1-64 point/spot/area
    lights, 1e3-2e6 texels, bounce on odd jobs).  `plain` bakes the jobs one after another, each
    with its own launches on one stream (dl_bake_batch with one stream is exactly a loop of
    dl_bake); `batched` hands them to worker streams that overlap on the device."""
    import dynlight_hip
    from dynlight_hip import scenes

    ins, ops = scenes.batch(args.jobs)
    ctx = dynlight_hip.Context(0)
    log(f"batch: {len(ins)} jobs, {sum(i.triangle_count for i in ins)} triangles, "
        f"{sum(len(i.lights) for i in ins)} lights")
    streams = [int(x) for x in args.streams.split(",")]
    ctx.bake_batch(ins[:16], ops[:16], streams=max(streams), read=False)
    log("warmup done")
    # launch shapes from the autotuner (timed on a 32-job sample; not part of the timed passes)
    tuned, tune_s = None, 0.0
    if not args.no_tune:
        import tempfile
        t0 = time.perf_counter()
        tuned = ctx.autotune_batch(ins, ops, max_streams=16,
                                   cache_path=os.path.join(tempfile.gettempdir(), "dynlight_tune.txt"))
        tune_s = time.perf_counter() - t0
        log(f"autotune ({tune_s:.1f} s): {tuned}")
        if tuned["streams"] not in streams:
            streams.append(tuned["streams"])
    runs = {}
    ref = None
    for s in [1] + streams:
        best = None
        for step in range(args.steps):
            t0 = time.perf_counter()
            st = ctx.bake_batch(ins, ops, streams=s, read="stats")
            dt = time.perf_counter() - t0
            log(f"{s} stream(s), pass {step + 1}/{args.steps}: {dt:.2f} s")
            best = dt if best is None else min(best, dt)
            ref = ref or st
            assert st == ref, "batched stats differ from per-job stats"
        runs[s] = best
    samples = sum(x["samples"] for x in ref)
    best_s = min(streams, key=lambda k: runs[k])
    out = {
        "metric": "texel-light visibility samples/s (1 GPU, batch of independent bakes)",
        "value": samples / runs[best_s],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": 1000.0 * runs[best_s],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch: {len(ins)} independent rooms, log-uniform 500-5e4 triangles, 1-64 "
                               "point/spot/area lights, 1e3-2e6 texels, bounce on odd jobs",
                   "jobs": len(ins), "samples_per_step": samples,
                   "bounce_rays_per_step": sum(x["bounce_rays"] for x in ref),
                   "texels": sum(x["texels"] for x in ref), "streams": best_s, "hw_queues": HW_QUEUES},
        "plain_value": samples / runs[1],
        "plain_ms": 1000.0 * runs[1],
        "batched_ms_by_streams": {str(k): 1000.0 * runs[k] for k in streams},
        "speedup_vs_plain": runs[1] / runs[best_s],
        "autotune": None if tuned is None else {
            **{k: tuned[k] for k in ("streams", "trace_block", "stream_block", "packet", "burial_block", "burial_lanes",
                                     "trace_stream", "trials", "from_cache")},
            "sample_ms": tuned["ms"], "tune_s": tune_s, "tuned_streams_ms": 1000.0 * runs[tuned["streams"]],
            "tuned_speedup_vs_plain": runs[1] / runs[tuned["streams"]]},
    }
    if not args.no_cpu:
        log("cpu baseline (oracle on a job sample)")
        out["cpu_baseline"] = batch_cpu_baseline(ins, ops)
        out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def serve_main(args):
    """Serving: the interior stays resident as a live scene (dl_service) while requests arrive as a
    Poisson process: 80% light edits (move or add 1-4 point lights with one bounce), 20% rooms (a
    room of the batch workload's generator joining the live scene as a part of its own).  Each is
    traced on the first free worker stream, admitted in arrival order within a device-memory
    budget (continuous batching: requests that arrive while others trace overlap on the device),
    and joins the live scene in arrival order, which then equals a full bake of the whole edited
    scene.  Latency runs from arrival to the live scene holding the request.  The first requests
    are also served one at a time on an idle service (arrivals 0.25 s apart): their latency is the
    service time alone, and their live-scene checksums must equal the mix's."""
    import numpy as np

    import dynlight_hip
    from dynlight_hip import scenes

    inp, opts = scenes.interior(2)
    ctx = dynlight_hip.Context(0)
    reqs = scenes.serve_requests(inp, count=args.requests, rate=args.rate)
    edits, is_room = [], []
    for kind, l, t in reqs:
        if kind == "room":
            edits.append((scenes.service_room(sum(is_room)), None, t))
        else:
            edits.append((l, [], t))
        is_room.append(kind == "room")
    is_room = np.array(is_room)
    rooms = int(is_room.sum())
    n_alone = min(40, len(edits))

    t0 = time.perf_counter()
    svc = dynlight_hip.Service(ctx, inp, opts)
    create_s = time.perf_counter() - t0
    log(f"service created in {create_s:.2f} s (initial bake of {len(inp.lights)} lights)")
    alone_edits = [(l, r, 0.25 * i) for i, (l, r, _) in enumerate(edits[:n_alone])]
    lat_a, up_a, sum_a = svc.replay_edits(alone_edits, workers=1, checksums=True)
    svc.close()
    lat_a = np.array(lat_a)
    log(f"{n_alone} edits alone: p50 {np.percentile(lat_a, 50):.1f} ms, p99 {np.percentile(lat_a, 99):.1f} ms")

    svc = dynlight_hip.Service(ctx, inp, opts)
    t0 = time.perf_counter()
    lat, ups, sums = svc.replay_edits(edits, workers=args.workers, checksums=True)
    wall = time.perf_counter() - t0
    svc.close()
    lat = np.array(lat)
    # the mix forms the live scene after the last edit of each joint assembly (checksum nonzero)
    formed = [k for k in range(n_alone) if sums[k]]
    same = bool(formed) and all(sums[k] == sum_a[k] for k in formed)
    samples = sum(u["traced_samples"] for u in ups)
    light_lat, room_lat = lat[~is_room], lat[is_room]
    log(f"{len(edits)} requests ({rooms} rooms) in {wall:.1f} s; p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms; light edits p50 {np.percentile(light_lat, 50):.1f} ms, p99 "
        f"{np.percentile(light_lat, 99):.1f} ms; live scenes formed among the first {n_alone}: {len(formed)}, equal to "
        f"alone: {same}")
    mean = lambda key: float(np.mean([u[key] for u in ups]))
    out = {
        "metric": f"p50/p99 rebake latency (1 GPU, resident interior, requests at {len(edits) / max(1e-9, edits[-1][2]):.1f}/s)",
        "value": float(np.percentile(lat, 99)),
        "unit": "ms (p99)",
        "n_gpus": 1,
        "steps": len(edits),
        "warmup": 0,
        "ms_per_step": float(np.mean(lat)),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serve: interior live scene (200k triangles, 256 bounce lights, 32 texels/m); Poisson "
                               f"arrivals {args.rate:g}/s; 80% edits moving or adding 1-4 point lights with one bounce, "
                               "20% rooms joining the live scene",
                   "requests": len(reqs), "light_edits": int((~is_room).sum()), "room_requests": rooms,
                   "workers": args.workers},
        "p50_ms": float(np.percentile(lat, 50)),
        "p90_ms": float(np.percentile(lat, 90)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()),
        "light_edit_p50_ms": float(np.percentile(light_lat, 50)),
        "light_edit_p99_ms": float(np.percentile(light_lat, 99)),
        "room_p50_ms": float(np.percentile(room_lat, 50)) if rooms else None,
        "room_p99_ms": float(np.percentile(room_lat, 99)) if rooms else None,
        "alone_p50_ms": float(np.percentile(lat_a, 50)),
        "alone_p99_ms": float(np.percentile(lat_a, 99)),
        "alone_equals_mix": bool(same),
        "alone_compared_scenes": len(formed),
        "mean_joint_edits": float(np.mean([u["reserved"] for u in ups])),
        "mean_trace_ms": mean("trace_ms"),
        "mean_assemble_ms": mean("assemble_ms"),
        "mean_queue_ms": mean("queue_ms"),
        "mean_dirty_meshes": mean("dirty_meshes"),
        "arrival_rate_per_s": len(edits) / max(1e-9, edits[-1][2]),
        "requests_per_s": len(edits) / wall,
        "samples_per_s": samples / wall,
        "service_create_s": create_s,
    }
    print(json.dumps(out), flush=True)


def outdoor_main(args):
    """Outdoor: a 1024x512 fBm heightfield (1,045,506 triangles) with 2000 box buildings, a soft sun
    (0.53 degrees, 16 samples, one bounce of 32 samples), 64 rectangle/tube area lights (16 samples),
    a baked sky (64 rays), 4 texels/m, sun map 1024^2 and sky probes at 1 m."""
    import dynlight_hip
    from dynlight_hip import scenes

    inp, opts = scenes.outdoor(3)
    ctx = dynlight_hip.Context(0)
    log(f"outdoor scene: {inp.mesh_count} meshes, {inp.triangle_count} triangles, {len(inp.lights)} lights")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        ctx.bake(inp, opts)
        log(f"warmup bake {w + 1}/{args.warmup}: {time.perf_counter() - t0:.2f} s")
    t0 = time.perf_counter()
    for step in range(args.steps):
        s = ctx.bake(inp, opts)
        tm = ctx.timings()
        log(f"timed bake {step + 1}/{args.steps}: {tm['total_ms'] / 1000.0:.2f} s")
    wall = time.perf_counter() - t0
    out = {
        "metric": "texel-light visibility samples/s (1 GPU, outdoor)",
        "value": s.stats["samples"] * args.steps / wall,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "outdoor: 1024x512 heightfield + 2000 buildings, soft sun (16 samples, bounce 32), "
                               "64 area lights (16 samples), sky (64 rays), 4 texels/m",
                   "meshes": inp.mesh_count, "triangles": inp.triangle_count, **s.stats},
        "rays_per_s": (s.stats["rays"] + s.stats["bounce_rays"]) * args.steps / wall,
        "stages_ms": {k: tm[k] for k in ("upload_ms", "texels_ms", "trace_ms", "trace_kernel_ms", "bounce_ms",
                                         "bounce_kernel_ms", "post_ms", "maps_ms", "download_ms", "total_ms")},
        "sun_maps": len(s.sun_maps), "sky_probes": s.sky_probes is not None,
    }
    print(json.dumps(out), flush=True)


def box_main(args):
    """The CPU-runnable case: a closed 10x4x10 m room (12 triangles) with 2 seeded boxes, 4 point
    lights of range 6-12 m, 16 texels/m, 1-bit shadows, no bounce.  The oracle runs the whole bake
    too, on every host thread and on one (best of 3 each)."""
    import oracle
    import dynlight_hip
    from dynlight_hip import scenes

    inp, opts = scenes.box_room(1)
    ctx = dynlight_hip.Context(0)
    for _ in range(max(1, args.warmup)):
        ctx.bake(inp, opts)
    steps = max(args.steps, 20)
    t0 = time.perf_counter()
    for _ in range(steps):
        s = ctx.bake(inp, opts)
    wall = time.perf_counter() - t0
    samples = s.stats["samples"]
    threads = len(os.sched_getaffinity(0))
    cpu = {}
    for n in (threads, 1):
        best = None
        for _ in range(3):
            t1 = time.perf_counter()
            oracle.bake(inp, opts, threads=n)
            dt = time.perf_counter() - t1
            best = dt if best is None else min(best, dt)
        cpu[n] = best
    out = {
        "metric": "texel-light visibility samples/s (1 GPU, box room)",
        "value": samples * steps / wall,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": 1000.0 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "box room: 10x4x10 m, 2 boxes, 4 point lights r 6-12 m, 16 texels/m, 1-bit, no bounce",
                   **s.stats},
        "cpu_baseline": {"value": samples / cpu[threads], "unit": "samples/s", "cores": threads, "kind": "port",
                         "cpu": host_cpu(),
                         "sample": f"the whole bake on the CPU oracle, best of 3: {cpu[threads] * 1000:.1f} ms on {threads} "
                                   f"threads, {cpu[1] * 1000:.1f} ms on 1 thread"},
        "cpu_one_thread_samples_per_s": samples / cpu[1],
    }
    out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="interior run: skip the batch workload's child runs")
    ap.add_argument("--no-serve", action="store_true", help="interior run: skip the serve workload's child run")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--workload", choices=("interior", "batch", "serve", "outdoor", "box"), default="interior",
                    help="interior: the 200k-triangle interior bake; batch: 512 independent small bakes; "
                         "serve: light rebakes of the resident interior; outdoor: the 1M-triangle outdoor bake; "
                         "box: the CPU-runnable box room")
    ap.add_argument("--requests", type=int, default=400, help="serve workload: requests (80%% are light requests)")
    ap.add_argument("--workers", type=int, default=2, help="serve workload: worker streams")
    ap.add_argument("--rate", type=float, default=20.0, help="serve workload: request arrivals per second")
    ap.add_argument("--jobs", type=int, default=512, help="batch workload: number of jobs")
    ap.add_argument("--streams", default="2,4,8,16", help="batch workload: worker stream counts to time (the fastest is reported)")
    ap.add_argument("--no-tune", action="store_true", help="batch workload: keep the default launch shapes (no autotuner)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4, "dispatches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(a sharded bake, receiver meshes split across GPUs, is designed but not built)", flush=True)
        sys.exit(2)

    if args.workload == "batch":
        batch_main(args)
        return
    if args.workload == "serve":
        serve_main(args)
        return
    if args.workload == "outdoor":
        outdoor_main(args)
        return
    if args.workload == "box":
        box_main(args)
        return

    import dynlight_hip
    from dynlight_hip import abi, scenes

    inp, opts = scenes.interior(2)
    ctx = dynlight_hip.Context(0)
    log(f"interior scene: {inp.mesh_count} meshes, {inp.triangle_count} triangles, {len(inp.lights)} lights")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        ctx.bake(inp, opts)
        log(f"warmup bake {w + 1}/{args.warmup}: {time.perf_counter() - t0:.2f} s")
    # instrumented bake: node and triangle counts of the shadow-ray kernel (untimed)
    ctx.set_counting(True)
    scene = ctx.bake(inp, opts)
    counts = ctx.timings()
    ctx.set_counting(False)
    log("instrumented bake done")
    samples = scene.stats["samples"]
    t_total, trace_ms = 0.0, 0.0
    t0 = time.perf_counter()
    for step in range(args.steps):
        s = ctx.bake(inp, opts)
        tm = ctx.timings()
        log(f"timed bake {step + 1}/{args.steps}: {tm['total_ms'] / 1000.0:.2f} s")
        trace_ms += tm["trace_kernel_ms"]
        assert s.stats["samples"] == samples
    t_total = time.perf_counter() - t0
    ms_per_step = 1000.0 * t_total / args.steps
    value = samples * args.steps / t_total
    nodes, tris = counts["trace_node_visits"], counts["trace_tri_tests"]
    b_alg = NODE_BYTES * nodes + TRI_BYTES * tris + samples / 8.0
    kernel_s = trace_ms / args.steps / 1000.0
    achieved = b_alg / kernel_s / 1e9
    # The any-hit kernel of the bounce gather: per launch, the BVH node records and triangle records
    # its visibility rays fetch plus each traced ray's 32-byte record and 4-byte list entry; counts
    # come from the instrumented bake.
    bnodes, btris = counts["bounce_node_visits"], counts["bounce_tri_tests"]
    launches = max(1, tm["gather_trace_launches"])
    g_alg = (NODE_BYTES * bnodes + TRI_BYTES * btris + (RAY_BYTES + 4.0) * counts["gather_traced"]) / launches
    g_launch_s = tm["gather_trace_ms"] / launches / 1000.0
    g_achieved = g_alg / g_launch_s / 1e9 if g_launch_s > 0 else 0.0
    # The sample setup of the gather (the largest stage, one span per chunk of jobs: the fast pass
    # k_gather_split, the list of open samples and the exact pass k_gather_open): per job, per open
    # sample, per photon read and per traced sample, plus each bounce light's photon tiles once.
    r_launches = max(1, tm["gather_rays_launches"])
    res = opts.photon_cube_resolution
    tiles = 6 * ((res + 7) // 8) ** 2 * sum(1 for l in inp.lights if l.illumination == abi.ILLUMINATION_SINGLE_BOUNCE)
    r_alg = (JOB_BYTES * counts["gather_jobs"] + OPEN_BYTES * counts["gather_open"]
             + PHOTON_BYTES * counts["gather_photon_reads"] + TRACED_BYTES * counts["gather_traced"]
             + TILE_BYTES * tiles) / r_launches
    r_launch_s = tm["gather_rays_ms"] / r_launches / 1000.0
    r_achieved = r_alg / r_launch_s / 1e9 if r_launch_s > 0 else 0.0
    vs, vo = profile_figure("VALU", "k_gather_split"), profile_figure("VALU", "k_gather_open")
    r_valu = vs + vo if vs and vo else None
    valu = None
    if r_valu and r_launch_s > 0:
        gips = r_valu / r_launch_s / 1e9
        valu = {"bound": "valu", "achieved": gips, "peak": VALU_PEAK_GIPS, "unit": "G wave-instructions/s",
                "frac": gips / VALU_PEAK_GIPS, "valu_per_launch": r_valu,
                "note": "SQ_INSTS_VALU per dispatch from rocprofv3 over the HIP-event launch time; the kernel's limit"}
    out = {
        "metric": "texel-light visibility samples/s (1 GPU)",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "interior: 8x8 rooms 6x3x6 m with doors, 200k triangles, 256 point/spot "
                               "lights r 4-12 m, 32 texels/m, 1-bit shadows + one colored bounce (32 samples, "
                               "512^2 photon cubes)",
                   "meshes": inp.mesh_count, "triangles": inp.triangle_count,
                   "texels": scene.stats["texels"], "samples_per_step": samples,
                   "rays_per_step": scene.stats["rays"], "bounce_rays_per_step": scene.stats["bounce_rays"]},
        "trace_kernel_ms": trace_ms / args.steps,
        "trace_samples_per_s": samples / kernel_s,
        "roofline": {"kernel": "gather sample setup (k_gather_split + k_gather_open)", "bound": "hbm", "achieved": r_achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r_achieved / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("k_gather_split") + measured_traffic("k_gather_open"))
                     if measured_traffic("k_gather_split") and measured_traffic("k_gather_open") else None,
                     "alg_bytes_per_launch": r_alg, "launch_ms": r_launch_s * 1000.0, "launches_per_step": r_launches,
                     "jobs_per_step": counts["gather_jobs"], "samples_per_step": counts["gather_samples"],
                     "traced_per_step": counts["gather_traced"], "photon_reads_per_step": counts["gather_photon_reads"],
                     "open_per_step": counts["gather_open"],
                     "bytes": f"{JOB_BYTES:.0f} B per job, {OPEN_BYTES:.0f} B per open sample, {PHOTON_BYTES:.0f} B per "
                              f"photon read, {TRACED_BYTES:.0f} B per traced sample (ray and contribution written), "
                              f"{TILE_BYTES:.0f} B per photon tile of each bounce light",
                     "valu_roofline": valu},
        # The other two traversal kernels, priced per use of each node and triangle record: most of
        # those reads are served from L2, so this is a work rate, not an HBM bandwidth; the HBM
        # bytes they move come from rocprofv3 (profiles/interior_bake_kernels.md).
        "other_kernels": [other_kernel("k_any_stream", g_alg, g_launch_s, launches,
                                       {"node_fetches_per_step": bnodes, "triangle_tests_per_step": btris,
                                        "bytes": f"{NODE_BYTES:.0f} B per node record, {TRI_BYTES:.0f} B per triangle "
                                                 f"record, {RAY_BYTES:.0f} B per ray record, 4 B per list entry"}),
                          other_kernel("k_trace", b_alg / max(1, tm["trace_launches"]), kernel_s / max(1, tm["trace_launches"]),
                                       tm["trace_launches"],
                                       {"node_fetches_per_step": nodes, "triangle_tests_per_step": tris,
                                        "bytes": f"{NODE_BYTES:.0f} B per node record, {TRI_BYTES:.0f} B per triangle "
                                                 "record, the level bits of each sample"})],
        "stages_ms": {k: tm[k] for k in ("upload_ms", "texels_ms", "trace_ms", "bounce_ms", "post_ms", "maps_ms",
                                         "download_ms", "total_ms", "bounce_kernel_ms", "gather_trace_ms",
                                         "gather_rays_ms")},
    }
    if not args.no_cpu:
        log("cpu baseline (oracle on a receiver sample)")
        out["cpu_baseline"] = cpu_baseline(inp, opts, samples)
        out["vs_cpu"] = value / out["cpu_baseline"]["value"]
    del scene, s
    ctx.close()
    if not args.no_batch:
        # the batched form against a loop of plain bakes (512 jobs), at HIP's default of 4 hardware
        # queues and at 16
        out["batch"] = {}
        for q in ("4", "16"):
            log(f"batch workload with GPU_MAX_HW_QUEUES={q} (child process)")
            b = child_run(["--workload", "batch", "--steps", "1", "--no-cpu", "--streams", "4,8,16"],
                          {"GPU_MAX_HW_QUEUES": q}, 900)
            if b:
                out["batch"][f"hw_queues_{q}"] = {k: b[k] for k in ("value", "plain_ms", "ms_per_step", "speedup_vs_plain",
                                                                   "batched_ms_by_streams", "autotune")}
                out["batch"][f"hw_queues_{q}"]["streams"] = b["config"]["streams"]
    if not args.no_serve:
        log("serve workload (child process)")
        v = child_run(["--workload", "serve"], {}, 900)
        if v:
            out["serve"] = {k: v[k] for k in v if k.endswith("_ms") or k.endswith("_per_s") or k in (
                "steps", "alone_equals_mix", "alone_compared_scenes", "mean_dirty_meshes", "mean_joint_edits")}
            out["serve"]["config"] = v["config"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
