#!/usr/bin/env python3
"""Benchmark of the field-evaluation + meshing hot path on one MI355X.

Workload (one "step"): torus R=1.0 r=0.4 with 6-octave ridged gradient noise and a
Voronoi F1 layer (jitter 0.9) blended by a low-frequency gradient mask, seed 42,
evaluated in FP32 on a 512^3 grid over [-1.6, 1.6]^3, then marching cubes at iso 0
(vertices deduplicated, with normals) in one stream-ordered upf_field_mesh_async call:
the field kernel also stores the cells' sign masks, which classification reads instead
of the field; the scan and the emission follow with the counts kept on the device (mesh
buffers sized once before timing).  `step_two_calls_ms` times the same step as
upf_field_mesh_count (16-byte count read-back) + upf_mc_emit; `mc_ms` the stand-alone
marching cubes on the stored field.  Inputs are generated on the device; nothing
crosses PCIe inside the timed region.  One untimed pass over every phase precedes the
timed steps, so that all figures are taken at the clocks the GPU reaches under load.

Prints one JSON line.  `value` = grid samples per second of the whole step (field +
meshing) in Gsamples/s.  Also reported: the field kernel alone, marching cubes in
cells/s, the fused forward+adjoint kernel, the 64-chunk batch, the 1024^2 texel bake,
the narrow-band planets, the full material stack, rooflines, and the CPU oracle timed
on the same grid (the reference binary cannot be built: its cage engine is absent).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector, vendor figure (packed FP32)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
# HBM bytes per launch of k_eval_grid_row on the 512^3 workload as timed for `roofline`
# (upf_field_eval_grid_ex, no sign words), from rocprofv3 --pmc: FETCH_SIZE 302 KiB (x 2.0
# by the 1 GiB copy calibration: the kernel reads only its arguments) and WRITE_SIZE
# 524288 KiB (x 1.0: each node written once, 512 MiB).  Launches that also store the
# sign words read 576 MiB of writes: 16 MiB of 8-byte words counted as 64 MiB, each word
# a partial line written by another workgroup (profiles/field_torus_ridged_voronoi_512.txt).
FIELD_TRAFFIC_512 = 2 * 302 * 1024 + 524288 * 1024


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, grid, seconds: float):
    """The CPU oracle (FP32, -O3, OpenMP over z-slabs) on the same grid: the whole grid,
    repeated to about `seconds` of work when one pass is shorter, else a centred block
    of z-slabs of about `seconds`; plus a short single-thread run for the per-core rate."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    from unnatural_planets_amd import spec as S

    o = Oracle()
    threads = int(o.L.upfo_num_threads())
    nx, ny, nz = S.grid_dims(grid)

    def run(k0, nslab, reps=1):
        sub = S.Grid()
        for a in range(3):
            sub.origin[a], sub.spacing[a], sub.dims[a] = grid.origin[a], grid.spacing[a], grid.dims[a]
        # slabs k0..k0+nslab of the grid (the slab origin is rounded once more than the
        # grid's own node coordinates; the slab is a timing sample, not a parity case)
        sub.origin[2] = np.float32(grid.origin[2]) + np.float32(k0) * np.float32(grid.spacing[2])
        sub.dims[2] = nslab
        t = time.perf_counter()
        for _ in range(reps):
            o.eval_grid(spec, sub)
        return time.perf_counter() - t

    nprobe = min(nz, 4 * threads)
    per_slab = run(nz // 2 - nprobe // 2, nprobe) / nprobe
    if per_slab * nz <= seconds:
        reps = max(1, int(seconds / (per_slab * nz)))
        t = run(0, nz, reps)
        nslab = nz
        what = f"the whole {nx}x{ny}x{nz} grid x{reps}"
    else:
        reps = 1
        nslab = int(max(nprobe, min(nz, seconds / per_slab)))
        t = run(max(0, nz // 2 - nslab // 2), nslab)
        what = f"{nslab} central z-slabs of the {nx}x{ny}x{nz} grid"
    samples = nx * ny * nslab * reps
    # single thread, two slabs
    o.L.upfo_set_num_threads(1)
    t1 = run(nz // 2, 2)
    o.L.upfo_set_num_threads(threads)
    return {
        "value": samples / t / 1e9,
        "unit": "Gsamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{what} ({samples} samples, {t:.1f} s) on {threads} threads of '{cpu_model()}'",
        "single_thread_value": 2 * nx * ny / t1 / 1e9,
    }


def bench_bake(upf, S, torch, reps: int):
    """128^3 sphere + 4-octave gradient fBm (seed 1234) meshed, then baked into a 1024^2
    atlas: atlas + rasterize + material, timed with HIP events on the launch stream."""
    L = upf.lib()
    s = S.spec_sphere_fbm(1234)
    g = S.grid_sphere_fbm(128)
    mesh = upf.mesh_generate_generic(s, g)
    w = h = 1024
    b = S.bake_spec(s)
    T, V = mesh.n_triangles, mesh.n_vertices
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    uv = torch.empty((T, 3, 2), dtype=torch.float32, device="cuda")
    tri = torch.empty((h, w), dtype=torch.int32, device="cuda")
    albedo = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    bump = torch.empty((h, w), dtype=torch.float32, device="cuda")
    rough = torch.empty((h, w), dtype=torch.float32, device="cuda")
    rbytes = L.upf_bake_rasterize_workspace_bytes(T, w, h)
    rws = torch.empty(rbytes, dtype=torch.uint8, device="cuda")
    P = [ctypes.c_void_p(t.data_ptr()) for t in (mesh.positions, mesh.normals, mesh.indices, uv, tri, albedo, bump,
                                                 rough, rws)]

    def atlas():
        upf.check(L.upf_atlas_per_triangle(T, w, h, P[3], sp), "atlas")

    def raster():
        upf.check(L.upf_bake_rasterize_ws(P[3], T, w, h, P[4], P[8], rbytes, sp), "rasterize")

    def raster_atomic():
        upf.check(L.upf_bake_rasterize(P[3], T, w, h, P[4], sp), "rasterize")

    def material():
        upf.check(L.upf_bake(ctypes.byref(b), P[0], P[1], V, P[2], T, P[3], P[4], w, h, P[5], P[6], P[7], sp), "bake")

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_atlas, t_raster_atomic, t_raster, t_mat = timed(atlas), timed(raster_atomic), timed(raster), timed(material)
    covered = int((tri >= 0).sum())
    # algorithmic bytes of k_bake: triangle id in + 20 B out per texel, the triangle's UV
    # and indices, and the position and normal of its vertices
    mbytes = w * h * (4 + 20) + T * (24 + 12) + V * 24
    gbs = mbytes / t_mat / 1e9
    return {
        "workload": "128^3 sphere r=1 + 4-oct gradient fBm (seed 1234) -> marching cubes -> 1024^2 atlas",
        "triangles": T, "vertices": V, "texels": w * h, "covered_texels": covered,
        "atlas_us": round(t_atlas * 1e6, 2), "rasterize_us": round(t_raster * 1e6, 2),
        "rasterize_atomic_us": round(t_raster_atomic * 1e6, 2),
        "material_us": round(t_mat * 1e6, 2),
        "gtexels_per_s": round(w * h / (t_atlas + t_raster + t_mat) / 1e9, 3),
        "roofline": {"bound": "hbm", "kernel": "k_bake", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": mbytes},
    }


def bench_post(upf, S, torch, reps: int):
    """Mesh post-processing on the 512^3 torus mesh (meshRemoveDisconnected, the water
    mask of sources/meshGeneration.cpp:41,62-95) and texture post-processing of 4096^2
    images (dilation by 7 and U8, sources/textures.cpp:73-81; the glTF conversions of
    sources/generator.cpp:136-139).  The mesh calls return their counts to the host, so
    each is timed on the host clock around a synchronized call; the image calls are timed
    with HIP events on the launch stream."""
    import time
    s = S.spec_torus_ridged_voronoi(42)
    g = S.grid_torus(512)
    mesh = upf.mesh_generate_generic(s, g, remove_disconnected=False)
    torch.cuda.synchronize()

    def host_timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    t_rd, kept = host_timed(lambda: upf.mesh_remove_disconnected(mesh))
    t_wm, water = host_timed(lambda: upf.mesh_water_mask(s, mesh, 0.0))
    out = {"mesh": {"workload": "512^3 torus + ridged + Voronoi mesh", "vertices": mesh.n_vertices,
                    "triangles": mesh.n_triangles, "remove_disconnected_ms": round(t_rd * 1e3, 3),
                    "kept_triangles": kept.n_triangles, "water_mask_ms": round(t_wm * 1e3, 3),
                    "water_triangles": water.n_triangles}}
    del mesh, kept, water

    n = 4096
    gen = torch.Generator(device="cuda").manual_seed(5)
    holes = torch.rand((n // 8, n // 8), device="cuda", generator=gen) < 0.3
    holes = holes.repeat_interleave(8, 0).repeat_interleave(8, 1)
    imgs = []
    for c in (3, 2, 1):
        im = torch.rand((n, n, c), device="cuda", generator=gen)
        im[holes] = float("nan")
        imgs.append(im.squeeze(2) if c == 1 else im)
    stream = torch.cuda.current_stream()

    def ev_timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    work = [im.clone() for im in imgs]

    def dilate_all():
        for w, im in zip(work, imgs):
            w.copy_(im)
            upf.image_dilate(w, 7)

    t_copy = ev_timed(lambda: [w.copy_(im) for w, im in zip(work, imgs)])
    t_dil = ev_timed(dilate_all) - t_copy
    u8 = [upf.image_to_u8(w) for w in work]
    t_u8 = ev_timed(lambda: [upf.image_to_u8(w) for w in work])
    t_nrm = ev_timed(lambda: upf.image_height_to_normal(u8[2], 1.0))
    t_pbr = ev_timed(lambda: upf.image_special_to_gltf_pbr(u8[1]))
    texels = n * n
    # dilation: the operation reads and writes the 6 channels (24 B) of every texel once
    # (the rounds run in LDS, so this is also its minimal HBM traffic)
    dil_bytes = texels * 6 * 4 * 2
    out["texture"] = {"workload": "4096^2 albedo (3 ch), special (2 ch), height (1 ch), 30% uncovered in 8x8 blocks",
                      "dilate7_ms": round(t_dil * 1e3, 3), "dilate_gbs": round(dil_bytes / t_dil / 1e9, 1),
                      "to_u8_ms": round(t_u8 * 1e3, 3), "height_to_normal_ms": round(t_nrm * 1e3, 3),
                      "special_to_gltf_ms": round(t_pbr * 1e3, 3)}
    return out


def bench_material(upf, S, torch, reps: int):
    """The full material stack (coloringDefault, poles and flowers on) on a reference
    planet: sphere r=1100 m + the lakes elevation mode (seed 3) meshed on the 128^3
    reference box, its per-texel bake into a 1024^2 atlas, and terrainTile at 1M points on
    the planet's shell.  Both kernels are VALU-bound; rates are reported."""
    L = upf.lib()
    f = S.spec_elevation_mode("lakes", 3)
    mesh = upf.mesh_generate_generic(f, S.grid_reference(128))
    m = S.material_spec(f, "default", "land", seed=7)
    w = h = 1024
    T, V = mesh.n_triangles, mesh.n_vertices
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    uv = upf.atlas_per_triangle(T, w, h)
    tri = upf.bake_rasterize(uv, w, h)
    albedo = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    special = torch.empty((h, w, 2), dtype=torch.float32, device="cuda")
    height = torch.empty((h, w), dtype=torch.float32, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    n = 1 << 20
    g = torch.Generator(device="cuda").manual_seed(5)
    d = torch.randn((n, 3), device="cuda", generator=g)
    d = d / d.norm(dim=1, keepdim=True)
    r = 1080.0 + 60.0 * torch.rand((n, 1), device="cuda", generator=g)
    pos = (d * r).contiguous()
    nrm = d.contiguous()
    tiles = torch.empty((n, 16), dtype=torch.float32, device="cuda")

    def bake():
        upf.check(L.upf_bake_material(ctypes.byref(m), vp(mesh.positions), vp(mesh.normals), V, vp(mesh.indices), T,
                                      vp(uv), vp(tri), w, h, vp(albedo), vp(special), vp(height), sp), "bake_material")

    def points():
        upf.check(L.upf_material_tiles(ctypes.byref(m), vp(pos), vp(nrm), n, vp(tiles), sp), "material_tiles")

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_bake, t_pts = timed(bake), timed(points)
    covered = int((tri >= 0).sum())
    return {
        "workload": "sphere r=1100 m + lakes elevation (seed 3), 128^3 reference box -> marching cubes -> 1024^2 "
                    "atlas, coloringDefault with poles and flowers (seed 7); terrainTile at 2^20 shell points",
        "triangles": T, "covered_texels": covered,
        "bake_ms": round(t_bake * 1e3, 3), "gtexels_per_s": round(w * h / t_bake / 1e9, 3),
        "tiles_ms": round(t_pts * 1e3, 3), "mtiles_per_s": round(n / t_pts / 1e6, 1),
    }


def bench_adjoint_256(upf, S, torch, timed, reps: int):
    """The differentiation workload: torus R=1, r=0.4 + 4-octave gradient fBm + cellular
    F1 (seed 7) on a 256^3 grid over [-1.6, 1.6]^3; the fused forward + adjoint grid
    kernel (f, df/dx per node and the 76 dL/dtheta sums for dL/df = 1) against the plain
    field on the same grid."""
    L = upf.lib()
    s, g = S.spec_torus_fbm_voronoi(7), S.uniform_grid(-1.6, 1.6, 256)
    n = 256 ** 3
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    f = torch.empty((256, 256, 256), dtype=torch.float32, device="cuda")
    dLdf = torch.ones_like(f)
    dfdx = torch.empty((256, 256, 256, 3), dtype=torch.float32, device="cuda")
    dth = torch.zeros((S.THETA_COUNT,), dtype=torch.float32, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def adj():
        upf.check(L.upf_field_adjoint_grid(ctypes.byref(s), ctypes.byref(g), vp(dLdf), vp(f), vp(dfdx), vp(dth), sp),
                  "adjoint 256")

    def fwd():
        upf.check(L.upf_field_eval_grid(ctypes.byref(s), ctypes.byref(g), vp(f), sp), "eval 256")
    adj()
    fwd()
    t_adj, t_fwd = timed(adj, reps), timed(fwd, reps)
    aflops = float(L.upf_adjoint_flops_per_sample(ctypes.byref(s)))
    flops = float(L.upf_flops_per_sample(ctypes.byref(s)))
    return {
        "workload": "torus R=1 r=0.4 + 4-oct gradient fBm + cellular F1 (seed 7), 256^3 over [-1.6,1.6]^3, "
                    "f + df/dx + 76 dL/dtheta",
        "gsamples_per_s": round(n / t_adj / 1e9, 4), "ms": round(t_adj * 1e3, 4), "flops_per_sample": aflops,
        "tflops": round(aflops * n / t_adj / 1e12, 3),
        "frac_fp32_peak": round(aflops * n / t_adj / 1e12 / PEAK_FP32_TFLOPS, 4),
        "plain_ms": round(t_fwd * 1e3, 4), "plain_frac_fp32_peak": round(flops * n / t_fwd / 1e12 / PEAK_FP32_TFLOPS, 4),
        "vs_plain_time_ratio": round(t_adj / t_fwd, 3),
    }


def bench_batch(upf, S, torch):
    """The uneven 64-chunk batch (committed manifest, seed 0): per-chunk launches (one
    stream-ordered upf_field_mesh_async per chunk, the fastest single-grid path: fused
    field + classification, no host read-back) against one autotuned, graph-captured
    batched run.  Buffers are allocated before either timed region."""
    import json

    from unnatural_planets_amd.batch import Batch
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "batch_manifest.json")))["chunks"]
    chunks = S.batch_chunks(man)
    samples = sum(m["n"] ** 3 for m in man)
    L = upf.lib()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        sp = ctypes.c_void_p(stream.cuda_stream)
        # per-chunk path: a workspace per grid size, output buffers sized for the largest chunk
        b = Batch(chunks)
        nv_all, nt_all = b.size_outputs()
        fields = [torch.empty(m["n"] ** 3, dtype=torch.float32, device="cuda") for m in man]
        ws = {m["n"]: upf.McWorkspace(chunks[i].grid) for i, m in enumerate(man)}
        pos = torch.empty((nv_all, 3), dtype=torch.float32, device="cuda")
        nrm = torch.empty_like(pos)
        idx = torch.empty((nt_all, 3), dtype=torch.int32, device="cuda")
        counts = torch.zeros((len(man), 2), dtype=torch.int64, device="cuda")
        mb = S.MeshBuffers()
        mb.positions, mb.normals, mb.indices = pos.data_ptr(), nrm.data_ptr(), idx.data_ptr()
        mb.cap_vertices, mb.cap_triangles = nv_all, nt_all

        def per_chunk():
            for i, m in enumerate(man):
                g = chunks[i].grid
                w = ws[m["n"]]
                upf.check(L.upf_field_mesh_async(ctypes.byref(chunks[i].spec), ctypes.byref(g), 0.0,
                                                 ctypes.c_void_p(fields[i].data_ptr()), None, ctypes.byref(mb),
                                                 ctypes.c_void_p(counts[i].data_ptr()),
                                                 ctypes.c_void_p(w.buf.data_ptr()), w.bytes, sp), "field_mesh_async")

        def timed(fn, reps):
            fn()
            stream.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                fn()
            stream.synchronize()
            return (time.perf_counter() - t) / reps

        t_seq = timed(per_chunk, 3)
        best = b.autotune(stream)
        b.capture(stream)
        t_graph = timed(lambda: b.replay(stream), 5)
        t_eager = timed(lambda: b.run(stream), 5)
    return {
        "chunks": len(man), "samples": samples, "vertices": nv_all, "triangles": nt_all,
        "per_chunk_ms": round(t_seq * 1e3, 3), "batched_graph_ms": round(t_graph * 1e3, 3),
        "batched_eager_ms": round(t_eager * 1e3, 3), "speedup_vs_per_chunk": round(t_seq / t_graph, 3),
        "batched_gsamples_per_s": round(samples / t_graph / 1e9, 3),
        "autotuned": {"field_variant": best.field_variant, "blocks_per_cu": best.blocks_per_cu,
                      "classify_z": best.classify_z},
    }


def bench_nb(upf, S, torch, reps: int, n: int = 1025, tex: int = 4096):
    """Narrow band: per planet a coarse pass over the 8-cell block corners of an n^3 fine
    grid over [-1.6, 1.6]^3, fine evaluation and meshing of the active blocks (analytic
    normals), then a tex^2 per-triangle atlas bake.  Buffers are sized by a first pass;
    the timed passes then run all planets back to back."""
    import numpy as np
    L = upf.lib()
    g = S.uniform_grid(-1.6, 1.6, n)
    band = S.nb_band(g)  # |grad f| < 2; a block's points lie within diag/2 of a corner
    planets = S.nb_planets()
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    # sizing pass
    max_act = max_v = max_t = 0
    ws = None
    for _, s in planets:
        m, _, st = upf.nb_extract(s, g, band, workspace=ws)
        max_act, max_v, max_t = max(max_act, st.active_blocks), max(max_v, st.n_vertices), max(max_t, st.n_triangles)
        ws = upf.NbWorkspace(g, max_act)
        del m
    pos = torch.empty((max_v, 3), dtype=torch.float32, device="cuda")
    nrm = torch.empty_like(pos)
    idx = torch.empty((max_t, 3), dtype=torch.int32, device="cuda")
    uv = torch.empty((max_t, 3, 2), dtype=torch.float32, device="cuda")
    tri = torch.empty((tex, tex), dtype=torch.int32, device="cuda")
    rbytes = L.upf_bake_rasterize_workspace_bytes(max_t, tex, tex)
    rws = torch.empty(rbytes, dtype=torch.uint8, device="cuda")
    albedo = torch.empty((tex, tex, 3), dtype=torch.float32, device="cuda")
    bump = torch.empty((tex, tex), dtype=torch.float32, device="cuda")
    rough = torch.empty((tex, tex), dtype=torch.float32, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    bakes = [S.bake_spec(s) for _, s in planets]
    stats = [S.NbStats() for _ in planets]

    def mesh_one(i):
        mb = S.MeshBuffers()
        mb.positions, mb.normals, mb.indices = pos.data_ptr(), nrm.data_ptr(), idx.data_ptr()
        mb.cap_vertices, mb.cap_triangles = max_v, max_t
        upf.check(L.upf_nb_extract(ctypes.byref(planets[i][1]), ctypes.byref(g), 0.0, band, ctypes.byref(mb), None,
                                   vp(ws.buf), ws.bytes, ctypes.byref(stats[i]), sp), "nb_extract")

    def bake_one(i):
        T, V = int(stats[i].n_triangles), int(stats[i].n_vertices)
        upf.check(L.upf_atlas_per_triangle(T, tex, tex, vp(uv), sp), "atlas")
        upf.check(L.upf_bake_rasterize_ws(vp(uv), T, tex, tex, vp(tri), vp(rws), rbytes, sp), "rasterize")
        upf.check(L.upf_bake(ctypes.byref(bakes[i]), vp(pos), vp(nrm), V, vp(idx), T, vp(uv), vp(tri), tex, tex,
                             vp(albedo), vp(bump), vp(rough), sp), "bake")

    for i in range(len(planets)):  # warm-up
        mesh_one(i)
        bake_one(i)
    torch.cuda.synchronize()
    t_mesh = t_bake = 0.0
    for _ in range(reps):
        for i in range(len(planets)):
            t0 = time.perf_counter()
            mesh_one(i)  # synchronises twice inside (count read-backs)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            bake_one(i)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            t_mesh += t1 - t0
            t_bake += t2 - t1
    npl = len(planets) * reps
    t_mesh, t_bake = t_mesh / npl, t_bake / npl
    act = [int(st.active_blocks) for st in stats]
    blocks = int(stats[0].blocks)
    # roofline of the mesh pass, per planet on average (host clock around each synchronised
    # call, so both fractions are lower bounds for the kernels):
    #   VALU: the coarse block corners and the 9^3 brick nodes of each active block, each
    #         at the spec's flops per sample (the vertex normals' gradient work is not
    #         counted);
    #   HBM:  the coarse corner field (written and read), the dense block -> rank map (4 B
    #         per block, written and read), each active block's 9^3 brick (4 B per node)
    #         and node table (2 B per node), each written once and read once, and the mesh
    #         (24 B per vertex, 12 B per triangle) written.
    nb_per_axis = (n - 1) // S.NB_BLOCK
    corners = (nb_per_axis + 1) ** 3
    mean_flops = float(np.mean([float(L.upf_flops_per_sample(ctypes.byref(sp_))) * (corners + a * 729)
                                for (_, sp_), a in zip(planets, act)]))
    mean_bytes = float(np.mean([2 * 4 * corners + 2 * 4 * blocks + a * 729 * (4 + 2) * 2 + 24 * int(st.n_vertices) +
                                12 * int(st.n_triangles) for a, st in zip(act, stats)]))
    mesh_tflops = mean_flops / t_mesh / 1e12
    mesh_gbs = mean_bytes / t_mesh / 1e9
    # the bake: triangle id in + 20 B out per texel, UV and indices per triangle, position
    # and normal per vertex (the 1024^2 bake's count), over atlas + rasterize + material
    bake_bytes = float(np.mean([tex * tex * 24 + 36 * int(st.n_triangles) + 24 * int(st.n_vertices) for st in stats]))
    bake_gbs = bake_bytes / t_bake / 1e9
    del pos, nrm, idx, uv, tri, albedo, bump, rough, ws, rws
    torch.cuda.empty_cache()

    # the dense sweep of the same fine grid for planet 0, for comparison
    s0 = planets[0][1]
    field = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    dws = upf.McWorkspace(g)
    nv, nt = ctypes.c_int64(), ctypes.c_int64()

    def dense():
        upf.check(L.upf_field_eval_grid(ctypes.byref(s0), ctypes.byref(g), vp(field), sp), "eval")
        upf.check(L.upf_mc_count(vp(field), ctypes.byref(g), 0.0, vp(dws.buf), dws.bytes, ctypes.byref(nv),
                                 ctypes.byref(nt), sp), "count")
    dense()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dense()
    torch.cuda.synchronize()
    t_dense = (time.perf_counter() - t0) / reps
    del field, dws
    torch.cuda.empty_cache()
    return {
        "workload": f"8 planets (sphere, rounded cube, torus, capsule x2; 5-oct gradient fBm, seeds 1000-1007), "
                    f"{n}^3 fine grid over [-1.6,1.6]^3 in 8^3-cell blocks, band {band:.4f}, {tex}^2 bake",
        "fine_nodes": n ** 3, "blocks": blocks, "active_blocks_mean": round(float(np.mean(act)), 1),
        "active_fraction_mean": round(float(np.mean(act)) / blocks, 5),
        "triangles_mean": round(float(np.mean([st.n_triangles for st in stats])), 1),
        "mesh_ms_per_planet": round(t_mesh * 1e3, 3), "bake_ms_per_planet": round(t_bake * 1e3, 3),
        "latency_ms_per_planet": round((t_mesh + t_bake) * 1e3, 3),
        "effective_gsamples_per_s": round(n ** 3 / t_mesh / 1e9, 2),
        "mesh_roofline": {"valu_tflops": round(mesh_tflops, 2), "valu_frac": round(mesh_tflops / PEAK_FP32_TFLOPS, 4),
                          "hbm_gbs": round(mesh_gbs, 1), "hbm_frac": round(mesh_gbs / PEAK_HBM_GBS, 4),
                          "flops_per_planet": round(mean_flops), "bytes_per_planet": round(mean_bytes)},
        "bake_roofline": {"hbm_gbs": round(bake_gbs, 1), "hbm_frac": round(bake_gbs / PEAK_HBM_GBS, 4),
                          "bytes_per_planet": round(bake_bytes)},
        "dense_field_count_ms_planet0": round(t_dense * 1e3, 3),
        "speedup_mesh_vs_dense_sweep": round(t_dense / t_mesh, 2),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512, help="grid nodes per axis")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-bake", action="store_true")
    ap.add_argument("--no-nb", action="store_true", help="skip the narrow-band 8-planet workload")
    ap.add_argument("--no-material", action="store_true", help="skip the full material stack")
    ap.add_argument("--no-post", action="store_true", help="skip the mesh and texture post-processing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 3 times (the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = 256 * 1024 * 1024  # 1 GiB of float32
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_launch": n * 4, "launches": 3}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the field and "
              "meshing shard into independent z-slabs). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import unnatural_planets_amd as upf
    from unnatural_planets_amd import spec as S

    L = upf.lib()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    spec = S.spec_torus_ridged_voronoi(42)
    grid = S.grid_torus(args.n)
    nx, ny, nz = S.grid_dims(grid)
    nodes = nx * ny * nz
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    field = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    ws = upf.McWorkspace(grid, dev)
    spec_p, grid_p = ctypes.byref(spec), ctypes.byref(grid)

    def eval_field(variant=upf.EVAL_AUTO):
        opts = upf.EvalOpts()
        opts.variant = variant
        upf.check(L.upf_field_eval_grid_ex(spec_p, grid_p, ctypes.c_void_p(field.data_ptr()), ctypes.byref(opts), sp),
                  "eval")

    nv, nt = ctypes.c_int64(), ctypes.c_int64()

    def mc_count():
        upf.check(L.upf_mc_count(ctypes.c_void_p(field.data_ptr()), grid_p, 0.0, ctypes.c_void_p(ws.buf.data_ptr()),
                                 ws.bytes, ctypes.byref(nv), ctypes.byref(nt), sp), "count")

    def field_count():
        # the field kernel also stores the sign masks; classification reads only those
        upf.check(L.upf_field_mesh_count(spec_p, grid_p, 0.0, ctypes.c_void_p(field.data_ptr()), None,
                                         ctypes.c_void_p(ws.buf.data_ptr()), ws.bytes, ctypes.byref(nv),
                                         ctypes.byref(nt), sp), "field_mesh_count")

    # size the mesh buffers once (outside the timed region)
    eval_field()
    mc_count()
    cap_v, cap_t = int(nv.value * 1.05) + 1024, int(nt.value * 1.05) + 1024
    pos = torch.empty((cap_v, 3), dtype=torch.float32, device=dev)
    nrm = torch.empty((cap_v, 3), dtype=torch.float32, device=dev)
    idx = torch.empty((cap_t, 3), dtype=torch.int32, device=dev)
    mb = S.MeshBuffers()
    mb.positions, mb.normals, mb.indices = pos.data_ptr(), nrm.data_ptr(), idx.data_ptr()
    mb.cap_vertices, mb.cap_triangles = cap_v, cap_t

    def mc_emit():
        mb.n_vertices, mb.n_triangles = nv.value, nt.value
        upf.check(L.upf_mc_emit(ctypes.c_void_p(field.data_ptr()), grid_p, 0.0, ctypes.byref(mb),
                                ctypes.c_void_p(ws.buf.data_ptr()), ws.bytes, sp), "emit")

    counts = torch.full((2,), -1, dtype=torch.int64, device=dev)

    def step():
        # field + classification + scan + emission in stream order: the counts stay on the
        # device (the buffers above hold the mesh with 5% to spare) and nothing synchronises
        upf.check(L.upf_field_mesh_async(spec_p, grid_p, 0.0, ctypes.c_void_p(field.data_ptr()), None,
                                         ctypes.byref(mb), ctypes.c_void_p(counts.data_ptr()),
                                         ctypes.c_void_p(ws.buf.data_ptr()), ws.bytes, sp), "field_mesh_async")

    def step_two_calls():  # upf_field_mesh_count (count read-back) + upf_mc_emit
        field_count()
        mc_emit()

    def run_steps(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    # per-phase timing with HIP events on the stream the library launches on
    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    reps = max(3, args.steps)
    # the GPU's clocks climb over its first ~50 ms of work: one untimed pass over every
    # phase first, so the step and the phases are all timed at the same clocks
    for fn in (eval_field, lambda: eval_field(upf.EVAL_PER_LANE), lambda: (mc_count(), mc_emit()), field_count):
        timed(fn, reps)

    t_step = run_steps(step)
    if counts.tolist() != [nv.value, nt.value]:
        raise RuntimeError(f"mesh counts {counts.tolist()} != {[nv.value, nt.value]}")
    t_step_two = run_steps(step_two_calls)
    V, T = int(nv.value), int(nt.value)
    t_field = timed(eval_field, reps)
    t_field_lane = timed(lambda: eval_field(upf.EVAL_PER_LANE), reps)
    mc_counts = torch.full((2,), -1, dtype=torch.int64, device=dev)

    def mc_async():  # stored field -> mesh in stream order, no count read-back
        upf.check(L.upf_mc_extract_async(ctypes.c_void_p(field.data_ptr()), grid_p, 0.0, ctypes.byref(mb),
                                         ctypes.c_void_p(mc_counts.data_ptr()), ctypes.c_void_p(ws.buf.data_ptr()),
                                         ws.bytes, sp), "mc_extract_async")

    t_mc_sync = timed(lambda: (mc_count(), mc_emit()), reps)
    t_mc = timed(mc_async, reps)
    if mc_counts.tolist() != [nv.value, nt.value]:
        raise RuntimeError(f"async mesh counts {mc_counts.tolist()} != {[nv.value, nt.value]}")
    t_fused = timed(field_count, reps)

    flops = float(L.upf_flops_per_sample(spec_p))
    field_tflops = flops * nodes / t_field / 1e12
    mc_bytes = 4 * nodes + 24 * V + 12 * T
    mc_gbs = mc_bytes / t_mc / 1e9

    out = {
        "metric": "SDF+noise Gsamples/s & MC cells/s at 512^3; fwd+adjoint vs plain, % FP32 peak",
        "value": round(nodes / t_step / 1e9, 4),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural field on a uniform grid, seed 42)",
        "config": {"workload": f"torus R=1 r=0.4 + 6-oct ridged gradient + Voronoi F1 (jitter 0.9), "
                               f"mask-blended, seed 42, {nx}^3 FP32 grid over [-1.6,1.6]^3, marching cubes iso 0",
                   "grid": [nx, ny, nz], "vertices": V, "triangles": T, "parallelism": "single GPU"},
        "field_gsamples_per_s": round(nodes / t_field / 1e9, 4),
        "field_ms": round(t_field * 1e3, 4),
        "field_per_lane_ms": round(t_field_lane * 1e3, 4),
        "flops_per_sample": flops,
        "mc_cells_per_s": round(cells / t_mc, 1),
        "mc_ms": round(t_mc * 1e3, 4),
        "mc_count_emit_ms": round(t_mc_sync * 1e3, 4),
        "field_plus_classify_ms": round(t_fused * 1e3, 4),
        "step_two_calls_ms": round(t_step_two * 1e3, 4),
        "roofline": {"bound": "valu", "kernel": "k_eval_grid_row", "achieved": round(field_tflops, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(field_tflops / PEAK_FP32_TFLOPS, 4),
                     "traffic": FIELD_TRAFFIC_512 if nx == ny == nz == 512 else None},
        "roofline_mc": {"bound": "hbm", "kernel": "marching cubes (signs+classify+scan+emit, upf_mc_extract_async)", "achieved": round(mc_gbs, 1),
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(mc_gbs / PEAK_HBM_GBS, 4),
                        "algorithmic_bytes": mc_bytes},
    }

    if not args.no_adjoint:
        dLdf = torch.ones((nz, ny, nx), dtype=torch.float32, device=dev)
        dfdx = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device=dev)
        dth = torch.zeros((S.THETA_COUNT,), dtype=torch.float32, device=dev)

        def adj():
            upf.check(L.upf_field_adjoint_grid(spec_p, grid_p, ctypes.c_void_p(dLdf.data_ptr()),
                                               ctypes.c_void_p(field.data_ptr()), ctypes.c_void_p(dfdx.data_ptr()),
                                               ctypes.c_void_p(dth.data_ptr()), sp), "adjoint")
        adj()
        t_adj = timed(adj, max(2, args.steps // 2))
        aflops = float(L.upf_adjoint_flops_per_sample(spec_p))
        out["adjoint"] = {
            "gsamples_per_s": round(nodes / t_adj / 1e9, 4),
            "ms": round(t_adj * 1e3, 4),
            "flops_per_sample": aflops,
            "tflops": round(aflops * nodes / t_adj / 1e12, 3),
            "frac_fp32_peak": round(aflops * nodes / t_adj / 1e12 / PEAK_FP32_TFLOPS, 4),
            "vs_plain_time_ratio": round(t_adj / t_field, 3),
        }
        del dLdf, dfdx

    if not args.no_adjoint:
        out["adjoint_256"] = bench_adjoint_256(upf, S, torch, timed, max(3, args.steps // 2))

    if not args.no_batch:
        out["batch"] = bench_batch(upf, S, torch)

    if not args.no_bake:
        out["bake"] = bench_bake(upf, S, torch, max(5, args.steps))

    if not args.no_nb:
        out["narrow_band"] = bench_nb(upf, S, torch, 2)

    if not args.no_material:
        out["material"] = bench_material(upf, S, torch, max(3, args.steps // 2))

    if not args.no_post:
        out["post"] = bench_post(upf, S, torch, max(3, args.steps // 2))

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, grid, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
