"""ctypes wrapper of oracle/liboracle.so, the CPU restatement used as the checker."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from unnatural_planets_amd.spec import THETA_COUNT, BakeSpec, FieldSpec, Grid, grid_dims

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_P = ctypes.c_void_p
_SP = ctypes.POINTER(FieldSpec)
_GP = ctypes.POINTER(Grid)


def _f32p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    def __init__(self, path: str = PATH):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        sig = {
            "upfo_hash3": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
            "upfo_hash1": (ctypes.c_uint32, [ctypes.c_uint32]),
            "upfo_noise_f64": (ctypes.c_double, [ctypes.c_int, ctypes.c_uint32, _P, ctypes.c_double]),
            "upfo_noise_f32": (ctypes.c_float, [ctypes.c_int, ctypes.c_uint32, _P, ctypes.c_float]),
            "upfo_eval_points_f32": (None, [_SP, _P, ctypes.c_int64, _P]),
            "upfo_eval_points_f64": (None, [_SP, _P, ctypes.c_int64, _P]),
            "upfo_eval_grid_f32": (None, [_SP, _GP, _P]),
            "upfo_eval_grid_f64": (None, [_SP, _GP, _P]),
            "upfo_grad_points": (None, [_SP, _P, ctypes.c_int64, _P, _P]),
            "upfo_theta_grad": (None, [_SP, _P, ctypes.c_int64, _P, _P]),
            "upfo_flops_at": (ctypes.c_longlong, [_SP, _P]),
            "upfo_layer_expand": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, _P, _P]),
            "upfo_mc_table": (None, [_P, _P]),
            "upfo_case_codes": (None, [_P, _GP, ctypes.c_float, _P]),
            "upfo_mc_run": (ctypes.c_void_p, [_P, _GP, ctypes.c_float, ctypes.c_int]),
            "upfo_mc_counts": (None, [_P, _P, _P]),
            "upfo_mc_copy": (None, [_P, _P, _P, _P]),
            "upfo_mc_free": (None, [_P]),
            "upfo_atlas_side": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int]),
            "upfo_atlas": (None, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _P]),
            "upfo_rasterize": (None, [_P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _P]),
            "upfo_bake": (None, [ctypes.POINTER(BakeSpec), _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P, _P,
                                 _P]),
            "upfo_image_dilate": (None, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "upfo_image_to_u8": (None, [_P, ctypes.c_int64, _P]),
            "upfo_bake_material": (None, [ctypes.c_void_p, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P]),
            "upfo_material_tiles": (None, [ctypes.c_void_p, _P, _P, ctypes.c_int64, _P]),
            "upfo_material_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P, _P, ctypes.c_int64, _P]),
            "upfo_voronoi_plane": (None, [ctypes.c_uint32, ctypes.c_int, ctypes.c_float, _P, _P, ctypes.c_int64, _P]),
            "upfo_material_noise": (ctypes.c_float, [ctypes.c_int, ctypes.c_uint32, ctypes.c_float, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                                     ctypes.c_int, _P]),
            "upfo_rgb_to_hsv": (None, [_P, _P]),
            "upfo_hsv_to_rgb": (None, [_P, _P]),
            "upfo_voronoi_cell_points": (None, [ctypes.c_uint32, ctypes.c_int, ctypes.c_float, _P, _P]),
            "upfo_voronoi_nearest": (None, [ctypes.c_uint32, ctypes.c_int, ctypes.c_float, _P, ctypes.c_int64, _P]),
            "upfo_image_height_to_normal": (None, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, _P]),
            "upfo_image_special_to_pbr": (None, [_P, ctypes.c_int64, _P]),
            "upfo_mesh_remove_disconnected": (None, [ctypes.c_int64, ctypes.c_int64, _P, _P, _P]),
            "upfo_mesh_water_keep": (None, [ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P]),
            "upfo_num_threads": (ctypes.c_int, []),
            "upfo_set_num_threads": (None, [ctypes.c_int]),
            "upfo_powf_n": (None, [_P, _P, _P, ctypes.c_int64]),
            "upfo_atanf_n": (None, [_P, _P, ctypes.c_int64]),
            "upfo_rescale": (ctypes.c_double, [ctypes.c_double] * 5),
            "upfo_sharp_edge": (ctypes.c_double, [ctypes.c_double] * 2),
            "upfo_terrace": (ctypes.c_double, [ctypes.c_double] * 2),
            "upfo_smooth_min": (ctypes.c_double, [ctypes.c_double] * 3),
            "upfo_smooth_max": (ctypes.c_double, [ctypes.c_double] * 3),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        self.L = L

    # ---- noise / hash
    def hash3(self, seed, x, y, z):
        return self.L.upfo_hash3(seed, x, y, z)

    def noise(self, noise: int, seed: int, q, jitter: float = 1.0, f32: bool = False):
        if f32:
            a = np.ascontiguousarray(q, dtype=np.float32)
            return self.L.upfo_noise_f32(noise, seed, _f32p(a), jitter)
        a = np.ascontiguousarray(q, dtype=np.float64)
        return self.L.upfo_noise_f64(noise, seed, _f32p(a), jitter)

    # ---- field
    def eval_points(self, s: FieldSpec, xyz, f64: bool = False):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        n = xyz.shape[0]
        out = np.empty(n, dtype=np.float64 if f64 else np.float32)
        fn = self.L.upfo_eval_points_f64 if f64 else self.L.upfo_eval_points_f32
        fn(ctypes.byref(s), _f32p(xyz), n, _f32p(out))
        return out

    def eval_grid(self, s: FieldSpec, g: Grid, f64: bool = False):
        nx, ny, nz = grid_dims(g)
        out = np.empty((nz, ny, nx), dtype=np.float64 if f64 else np.float32)
        fn = self.L.upfo_eval_grid_f64 if f64 else self.L.upfo_eval_grid_f32
        fn(ctypes.byref(s), ctypes.byref(g), _f32p(out))
        return out

    def grad_points(self, s: FieldSpec, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        n = xyz.shape[0]
        f = np.empty(n, dtype=np.float64)
        g = np.empty((n, 3), dtype=np.float64)
        self.L.upfo_grad_points(ctypes.byref(s), _f32p(xyz), n, _f32p(f), _f32p(g))
        return f, g

    def theta_grad(self, s: FieldSpec, xyz, dLdf):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        dLdf = np.ascontiguousarray(dLdf, dtype=np.float32)
        out = np.empty(THETA_COUNT, dtype=np.float64)
        self.L.upfo_theta_grad(ctypes.byref(s), _f32p(xyz), xyz.shape[0], _f32p(dLdf), _f32p(out))
        return out

    def flops_at(self, s: FieldSpec, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        return int(self.L.upfo_flops_at(ctypes.byref(s), _f32p(a)))

    def layer_expand(self, fractal, octaves, frequency, lacunarity, gain):
        fr = np.zeros(8, dtype=np.float32)
        am = np.zeros(8, dtype=np.float32)
        self.L.upfo_layer_expand(fractal, octaves, frequency, lacunarity, gain, _f32p(fr), _f32p(am))
        return fr, am

    # ---- marching cubes
    def mc_table(self):
        nt = np.zeros(256, dtype=np.int32)
        tr = np.zeros((256, 15), dtype=np.int32)
        self.L.upfo_mc_table(_f32p(nt), _f32p(tr))
        return nt, tr

    def case_codes(self, field, g: Grid, iso=0.0):
        nx, ny, nz = grid_dims(g)
        field = np.ascontiguousarray(field, dtype=np.float32)
        out = np.empty((nz - 1, ny - 1, nx - 1), dtype=np.uint8)
        self.L.upfo_case_codes(_f32p(field), ctypes.byref(g), iso, _f32p(out))
        return out

    def mc(self, field, g: Grid, iso=0.0, normals=True):
        field = np.ascontiguousarray(field, dtype=np.float32)
        h = self.L.upfo_mc_run(_f32p(field), ctypes.byref(g), iso, 1 if normals else 0)
        nv, nt = ctypes.c_int64(), ctypes.c_int64()
        self.L.upfo_mc_counts(h, ctypes.byref(nv), ctypes.byref(nt))
        pos = np.empty((nv.value, 3), dtype=np.float32)
        nrm = np.empty((nv.value, 3), dtype=np.float32)
        idx = np.empty((nt.value, 3), dtype=np.uint32)
        self.L.upfo_mc_copy(h, _f32p(pos), _f32p(nrm) if normals else None, _f32p(idx))
        self.L.upfo_mc_free(h)
        return pos, (nrm if normals else None), idx

    # ---- texel bake
    def atlas_side(self, n, w, h):
        return int(self.L.upfo_atlas_side(n, w, h))

    def atlas(self, n, w, h):
        uv = np.zeros((n, 3, 2), dtype=np.float32)
        self.L.upfo_atlas(n, w, h, _f32p(uv))
        return uv

    def rasterize(self, uv, w, h):
        uv = np.ascontiguousarray(uv, dtype=np.float32)
        tri = np.empty((h, w), dtype=np.int32)
        self.L.upfo_rasterize(_f32p(uv), uv.shape[0], w, h, _f32p(tri))
        return tri

    def bake(self, b: BakeSpec, pos, nrm, idx, uv, tri):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        uv = np.ascontiguousarray(uv, dtype=np.float32)
        tri = np.ascontiguousarray(tri, dtype=np.int32)
        h, w = tri.shape
        albedo = np.empty((h, w, 3), dtype=np.float32)
        bump = np.empty((h, w), dtype=np.float32)
        rough = np.empty((h, w), dtype=np.float32)
        self.L.upfo_bake(ctypes.byref(b), _f32p(pos), _f32p(nrm), _f32p(idx), _f32p(uv), _f32p(tri), w, h,
                         _f32p(albedo), _f32p(bump), _f32p(rough))
        return albedo, bump, rough

    def image_dilate(self, img, rounds):
        img = np.array(img, dtype=np.float32, order="C", copy=True)
        h, w = img.shape[:2]
        c = 1 if img.ndim == 2 else img.shape[2]
        self.L.upfo_image_dilate(_f32p(img), w, h, c, rounds)
        return img

    def image_to_u8(self, img):
        img = np.ascontiguousarray(img, dtype=np.float32)
        out = np.empty(img.shape, dtype=np.uint8)
        self.L.upfo_image_to_u8(_f32p(img), img.size, _f32p(out))
        return out

    def voronoi_nearest(self, seed, ppc, cell_size, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pos)
        self.L.upfo_voronoi_nearest(seed, ppc, cell_size, _f32p(pos), len(pos), _f32p(out))
        return out

    def material_tiles(self, m, pos, nrm):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((len(pos), 16), dtype=np.float32)
        self.L.upfo_material_tiles(ctypes.byref(m), _f32p(pos), _f32p(nrm), len(pos), _f32p(out))
        return out

    def bake_material(self, m, pos, nrm, idx, uv, tri):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        uv = np.ascontiguousarray(uv, dtype=np.float32)
        tri = np.ascontiguousarray(tri, dtype=np.int32)
        h, w = tri.shape
        ac = 4 if m.purpose == 2 else 3
        albedo = np.empty((h, w, ac), dtype=np.float32)
        special = np.empty((h, w, 2), dtype=np.float32)
        height = np.empty((h, w), dtype=np.float32)
        self.L.upfo_bake_material(ctypes.byref(m), _f32p(pos), _f32p(nrm), _f32p(idx), _f32p(uv), _f32p(tri), w, h,
                                  _f32p(albedo), _f32p(special), _f32p(height))
        return albedo, special, height

    def material_step(self, m, step, pos, nrm, tiles):
        """One generate* step on prepared tiles (a [n, 16] float32 array in upf_tile
        layout, biome and type as int32 bits); returns the updated copy."""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 3)
        io = np.array(tiles, dtype=np.float32, order="C", copy=True).reshape(-1, 16)
        rc = self.L.upfo_material_step(ctypes.byref(m), step, _f32p(pos), _f32p(nrm), len(pos), _f32p(io))
        assert rc == 0, rc
        return io

    def voronoi_plane(self, seed, ppc, cell_size, pos, nrm):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pos)
        self.L.upfo_voronoi_plane(seed, ppc, cell_size, _f32p(pos), _f32p(nrm), len(pos), _f32p(out))
        return out

    def material_noise(self, kind, seed, freq, p, fractal=0, octaves=3, lac=2.0, gain=0.5, dist=0, ret=0):
        a = np.ascontiguousarray(p, dtype=np.float32)
        return self.L.upfo_material_noise(kind, seed, freq, fractal, octaves, lac, gain, dist, ret, _f32p(a))

    def rgb_to_hsv(self, c):
        a = np.ascontiguousarray(c, dtype=np.float32)
        o = np.empty(3, dtype=np.float32)
        self.L.upfo_rgb_to_hsv(_f32p(a), _f32p(o))
        return o

    def hsv_to_rgb(self, c):
        a = np.ascontiguousarray(c, dtype=np.float32)
        o = np.empty(3, dtype=np.float32)
        self.L.upfo_hsv_to_rgb(_f32p(a), _f32p(o))
        return o

    def voronoi_cell_points(self, seed, ppc, cell_size, cell):
        c = np.ascontiguousarray(cell, dtype=np.int32)
        out = np.empty((ppc, 3), dtype=np.float32)
        self.L.upfo_voronoi_cell_points(seed, ppc, cell_size, _f32p(c), _f32p(out))
        return out

    def image_height_to_normal(self, height_u8, strength=1.0):
        height_u8 = np.ascontiguousarray(height_u8, dtype=np.uint8)
        h, w = height_u8.shape
        out = np.empty((h, w, 3), dtype=np.uint8)
        self.L.upfo_image_height_to_normal(_f32p(height_u8), w, h, strength, _f32p(out))
        return out

    def image_special_to_pbr(self, special_u8):
        special_u8 = np.ascontiguousarray(special_u8, dtype=np.uint8)
        n = special_u8.size // 2
        out = np.empty(special_u8.shape[:-1] + (3,), dtype=np.uint8)
        self.L.upfo_image_special_to_pbr(_f32p(special_u8), n, _f32p(out))
        return out

    # ---- the project's pow and atan (upf_math.hpp, shared by the oracle and the device)
    def powf(self, x, e):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(e, dtype=np.float32), x.shape))
        out = np.empty_like(x)
        self.L.upfo_powf_n(_f32p(x), _f32p(e), _f32p(out), len(x))
        return out

    def atanf(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        out = np.empty_like(x)
        self.L.upfo_atanf_n(_f32p(x), _f32p(out), len(x))
        return out

    # ---- mesh post-processing: keep flags and the order-preserving compaction
    def remove_disconnected_flags(self, nv, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        kv = np.zeros(nv, dtype=np.uint8)
        kt = np.zeros(len(idx), dtype=np.uint8)
        self.L.upfo_mesh_remove_disconnected(nv, len(idx), _f32p(idx), _f32p(kv), _f32p(kt))
        return kv.astype(bool), kt.astype(bool)

    def water_keep_flags(self, nv, idx, valid):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        valid = np.ascontiguousarray(valid, dtype=np.uint8)
        kv = np.zeros(nv, dtype=np.uint8)
        kt = np.zeros(len(idx), dtype=np.uint8)
        self.L.upfo_mesh_water_keep(nv, len(idx), _f32p(idx), _f32p(valid), _f32p(kv), _f32p(kt))
        return kv.astype(bool), kt.astype(bool)


def compact_mesh(pos, nrm, idx, kv, kt):
    """Order-preserving compaction by keep flags (the layout both libupf and the oracle use)."""
    newid = np.cumsum(kv) - 1
    out_idx = newid[np.asarray(idx, dtype=np.int64)[kt]].astype(np.uint32)
    return pos[kv], (nrm[kv] if nrm is not None else None), out_idx


def mesh_edge_stats(idx: np.ndarray):
    """(n_unique_edges, n_boundary_edges, n_bad_edges, oriented_ok) for a triangle list.
    An interior edge must appear exactly twice, once in each direction."""
    idx = np.asarray(idx, dtype=np.int64)
    e = np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]], axis=0)
    key = np.minimum(e[:, 0], e[:, 1]) * (1 << 32) + np.maximum(e[:, 0], e[:, 1])
    direction = (e[:, 0] < e[:, 1]).astype(np.int64)
    order = np.argsort(key, kind="stable")
    key, direction = key[order], direction[order]
    uniq, start, counts = np.unique(key, return_index=True, return_counts=True)
    dsum = np.add.reduceat(direction, start)
    boundary = int(np.sum(counts == 1))
    bad = int(np.sum(counts > 2))
    oriented = bool(np.all(dsum[counts == 2] == 1))
    return len(uniq), boundary, bad, oriented
