"""ctypes mirror of the POD structs and enums in include/upf.h, plus spec builders.

The builders mirror how the reference configures its fields:
  * `layer(...)` ~ NoiseFunctionCreateConfig + newNoiseFunction
    (reference sources/elevationModes.cpp:16-27, 83-87);
  * `FieldSpec.compose` ~ terrainSdfLand/Water/Navigation/Elevation/ElevationRaw
    (reference sources/modesConfiguration.cpp:230-279);
  * `shape=` ~ the base-shape functions of sources/sdf.cpp.
"""
from __future__ import annotations

import ctypes
import math

MAX_LAYERS = 4
MAX_OCTAVES = 8
MAX_SHAPE_PARAMS = 8
THETA_COUNT = 76

# status
OK, ERR_INVALID_VALUE, ERR_BAD_SPEC, ERR_CAPACITY, ERR_WORKSPACE, ERR_BAD_GRID, ERR_HIP, \
    ERR_NULL, ERR_EMPTY_MESH, ERR_BAND = range(10)
NB_BLOCK = 8

SHAPE = {"none": 0, "sphere": 1, "box": 2, "torus": 3, "capsule": 4, "cylinder": 5, "plane": 6, "octahedron": 7,
         "double_torus": 8, "torus_cross": 9, "h2o": 10, "bowl": 11, "inside_cube": 12, "wormhole": 13,
         "pipe": 14, "hemispheres": 15, "h3o": 16, "h4o": 17, "gear": 18, "mandelbulb": 19, "mobius": 20,
         "fibers": 21, "asteroid": 22, "twisted_plane": 23, "belt": 24, "triangular_prism": 25,
         "hexagonal_prism": 26, "twisted_hex_prism": 27, "tetrahedron": 28, "sine_net": 29, "knot": 30}
SINE_NET_WEIGHTS = 625
# the reference's parameters (sources/sdf.cpp, metres) for the parametric extended shapes
SHAPE_REFERENCE_PARAMS = {"cylinder": (10000.0, 800.0), "octahedron": (800.0, 100.0),
                          "double_torus": (900.0, 400.0, 200.0), "torus_cross": (750.0, 400.0, 200.0),
                          "plane": (0.0, 0.0, -1.0, 0.0),
                          # the reference draws these at run time: a fixed draw in its range
                          "fibers": (37.0, -12.5, 71.0), "asteroid": (0.7, 1234.0)}
NOISE = {"gradient": 1, "value": 2, "cellular": 3, "simplex": 4}
# elevation modes: the reference's fixed elevation functions (sources/elevationModes.cpp)
ELEV = {"layers": 0, "none": 1, "simple": 2, "legacy": 3, "lakes": 4, "islands": 5, "craters": 6}
ELEV_LAYERS = {0: None, 1: 0, 2: 1, 3: 2, 4: 4, 5: 4, 6: 2}
FRACTAL = {"none": 0, "fbm": 1, "ridged": 2}
BLEND = {"add": 0, "mask": 1, "lerp": 2}
COMPOSE = {"land": 0, "water": 1, "navigation": 2, "elevation": 3, "elevation_raw": 4}


class Layer(ctypes.Structure):
    _fields_ = [
        ("noise", ctypes.c_int32), ("fractal", ctypes.c_int32), ("octaves", ctypes.c_int32),
        ("blend", ctypes.c_int32), ("seed", ctypes.c_uint32), ("weight", ctypes.c_float),
        ("jitter", ctypes.c_float), ("reserved", ctypes.c_float),
        ("freq", ctypes.c_float * MAX_OCTAVES), ("amp", ctypes.c_float * MAX_OCTAVES),
    ]


class FieldSpec(ctypes.Structure):
    _fields_ = [
        ("shape", ctypes.c_int32), ("compose", ctypes.c_int32), ("n_layers", ctypes.c_int32),
        ("elevation_mode", ctypes.c_int32), ("shape_params", ctypes.c_float * MAX_SHAPE_PARAMS),
        ("elevation_scale", ctypes.c_float), ("reserved", ctypes.c_int32),
        ("shape_weights", ctypes.c_void_p),
        ("layers", Layer * MAX_LAYERS),
    ]


class Grid(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("spacing", ctypes.c_float * 3),
                ("dims", ctypes.c_int32 * 3), ("reserved", ctypes.c_int32)]


class MeshBuffers(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("normals", ctypes.c_void_p),
                ("indices", ctypes.c_void_p), ("cap_vertices", ctypes.c_int64),
                ("cap_triangles", ctypes.c_int64), ("n_vertices", ctypes.c_int64),
                ("n_triangles", ctypes.c_int64)]


assert ctypes.sizeof(Layer) == 96
assert ctypes.sizeof(FieldSpec) == 16 + 32 + 16 + 4 * 96
assert ctypes.sizeof(Grid) == 40


def f32(x: float) -> float:
    """Round a Python float to the nearest float32."""
    return ctypes.c_float(x).value


def expand_octaves(fractal: int, octaves: int, frequency: float, lacunarity: float, gain: float):
    """Per-octave (freq, amp) in float32, in the order documented at upf_layer."""
    oct_ = 1 if fractal == FRACTAL["none"] else octaves
    g = f32(abs(f32(gain)))
    s, a = f32(1.0), g
    for _ in range(1, oct_):
        s = f32(s + a)
        a = f32(a * g)
    w = f32(1.0) if fractal == FRACTAL["none"] else f32(f32(1.0) / s)
    f = f32(frequency)
    freq, amp = [0.0] * MAX_OCTAVES, [0.0] * MAX_OCTAVES
    for k in range(oct_):
        freq[k], amp[k] = f, w
        f = f32(f * f32(lacunarity))
        w = f32(w * f32(gain))
    return freq, amp


def layer(noise="gradient", fractal="fbm", octaves=4, frequency=1.0, lacunarity=2.0, gain=0.5,
          weight=1.0, seed=0, jitter=1.0, blend="add") -> Layer:
    L = Layer()
    L.noise = NOISE[noise] if isinstance(noise, str) else int(noise)
    L.fractal = FRACTAL[fractal] if isinstance(fractal, str) else int(fractal)
    L.octaves = 1 if L.fractal == 0 else int(octaves)
    L.blend = BLEND[blend] if isinstance(blend, str) else int(blend)
    L.seed = int(seed) & 0xFFFFFFFF
    L.weight = weight
    L.jitter = jitter
    freq, amp = expand_octaves(L.fractal, L.octaves, frequency, lacunarity, gain)
    for k in range(MAX_OCTAVES):
        L.freq[k], L.amp[k] = freq[k], amp[k]
    return L


def field_spec(shape="sphere", shape_params=(1.0,), layers=(), compose="land",
               elevation_scale=1.0) -> FieldSpec:
    s = FieldSpec()
    s.shape = SHAPE[shape] if isinstance(shape, str) else int(shape)
    s.compose = COMPOSE[compose] if isinstance(compose, str) else int(compose)
    for i, v in enumerate(shape_params):
        s.shape_params[i] = v
    s.elevation_scale = elevation_scale
    s.n_layers = len(layers)
    for i, L in enumerate(layers):
        s.layers[i] = L
    return s


def elevation_mode_layers(mode, seed: int = 0):
    """The noise layers of an elevation mode, as upf_spec_set_elevation_mode fills them
    (sources/elevationModes.cpp:14-198; the k-th noise seeded seed + k; cage's unset
    lacunarity and gain taken as 2 and 0.5)."""
    m = ELEV[mode] if isinstance(mode, str) else int(mode)
    s0, s1, s2, s3 = (int(seed) + k for k in range(4))
    if m == ELEV["simple"]:
        return [layer("simplex", "ridged", 6, 0.0005, 2.0, 0.4, seed=s0, jitter=0.0)]
    if m == ELEV["legacy"]:
        return [layer("value", "fbm", 4, 0.0005, seed=s0, jitter=0.0), layer("value", "fbm", 4, 1.0, seed=s1, jitter=0.0)]
    if m in (ELEV["lakes"], ELEV["islands"]):
        return [layer("value", "fbm", 4, 0.0013, seed=s0, jitter=0.0),
                layer("gradient", "none", 1, 0.0015, seed=s1, jitter=0.0),
                layer("simplex", "ridged", 4, 0.001, 1.5, -0.4, seed=s2, jitter=0.0),
                layer("gradient", "fbm", 3, 0.002, 2.0, 0.3, seed=s3, jitter=0.0)]
    if m == ELEV["craters"]:
        v = Layer()  # the impact Voronoi: seed, points per cell (octaves), cell size (freq[0])
        v.noise, v.octaves, v.seed, v.weight = NOISE["cellular"], 2, s0 & 0xFFFFFFFF, 1.0
        v.freq[0], v.amp[0] = 300.0, 1.0
        return [v, layer("value", "none", 1, 1.0, seed=s1, jitter=0.0)]
    return []


def spec_elevation_mode(mode, seed: int = 1, shape="sphere", shape_params=(1100.0,), compose="land",
                        elevation_scale=0.1) -> FieldSpec:
    """A reference planet: base shape (metres, the sphere of sources/sdf.cpp:20-23 by
    default) minus elevation / 10 (sources/modesConfiguration.cpp:227, 248-258), with the
    elevation of a fixed mode."""
    s = field_spec(shape, shape_params, elevation_mode_layers(mode, seed), compose, elevation_scale)
    s.elevation_mode = ELEV[mode] if isinstance(mode, str) else int(mode)
    return s


def grid_reference(n: int = 128) -> Grid:
    """The reference's sweep box [-2000, 2000]^3 at n nodes per axis
    (sources/meshGeneration.cpp:19-28)."""
    return uniform_grid(-2000.0, 2000.0, n)


def uniform_grid(lo: float, hi: float, n: int) -> Grid:
    """n^3 nodes over [lo, hi]^3: origin lo, spacing (hi - lo)/(n - 1) in float32."""
    g = Grid()
    sp = f32(f32(f32(hi) - f32(lo)) / f32(n - 1))
    for a in range(3):
        g.origin[a] = lo
        g.spacing[a] = sp
        g.dims[a] = n
    return g


def grid_dims(g: Grid):
    return int(g.dims[0]), int(g.dims[1]), int(g.dims[2])


def copy_spec(s: FieldSpec) -> FieldSpec:
    r = FieldSpec()
    ctypes.memmove(ctypes.byref(r), ctypes.byref(s), ctypes.sizeof(FieldSpec))
    return r


def theta_index(kind: str, layer_idx: int = 0, octave: int = 0) -> int:
    """Index into the 76-float upf_theta_grad vector."""
    if kind == "shape":
        return layer_idx  # shape parameter index
    if kind == "weight":
        return MAX_SHAPE_PARAMS + layer_idx
    if kind == "freq":
        return MAX_SHAPE_PARAMS + MAX_LAYERS + layer_idx * MAX_OCTAVES + octave
    if kind == "amp":
        return MAX_SHAPE_PARAMS + MAX_LAYERS + MAX_LAYERS * MAX_OCTAVES + layer_idx * MAX_OCTAVES + octave
    raise KeyError(kind)


# ---------------------------------------------------------------------------- workloads
def spec_sphere_fbm(seed: int = 1234) -> FieldSpec:
    """Sphere r=1 plus 0.05 x 4-octave gradient fBm (f0 2.0, lacunarity 2.0, gain 0.5)."""
    return field_spec("sphere", (1.0,), [layer("gradient", "fbm", 4, 2.0, 2.0, 0.5, 0.05, seed)])


def grid_sphere_fbm(n: int = 128) -> Grid:
    return uniform_grid(-1.3, 1.3, n)


def spec_torus_ridged_voronoi(seed: int = 42) -> FieldSpec:
    """Torus R=1.0 r=0.4; 6-octave ridged gradient noise and a Voronoi F1 layer (jitter 0.9)
    blended by a 1-octave low-frequency gradient mask."""
    return field_spec("torus", (1.0, 0.4), [
        layer("gradient", "ridged", 6, 2.0, 2.0, 0.5, 0.06, seed),
        layer("gradient", "none", 1, 0.75, 2.0, 0.5, 0.5, seed + 101, blend="mask"),
        layer("cellular", "none", 1, 6.0, 2.0, 0.5, 0.08, seed + 202, jitter=0.9, blend="lerp"),
    ])


def grid_torus(n: int = 512) -> Grid:
    return uniform_grid(-1.6, 1.6, n)


def spec_torus_fbm_voronoi(seed: int = 7) -> FieldSpec:
    """Torus R=1, r=0.4 + 4-octave gradient fBm + cellular F1 layer (all blended by add):
    the 11-parameter θ = {4 amp, 4 freq, Voronoi weight, R, r} of the adjoint workload."""
    return field_spec("torus", (1.0, 0.4), [
        layer("gradient", "fbm", 4, 2.0, 2.0, 0.5, 0.05, seed),
        layer("cellular", "none", 1, 4.0, 2.0, 0.5, 0.03, seed + 101, jitter=0.9),
    ])


# ---------------------------------------------------------------------------- batched chunks
class Chunk(ctypes.Structure):
    _fields_ = [("spec", FieldSpec), ("grid", Grid), ("iso", ctypes.c_float), ("reserved", ctypes.c_int32 * 3)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("field_variant", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("classify_z", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class NbStats(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_int64), ("active_blocks", ctypes.c_int64), ("n_vertices", ctypes.c_int64),
                ("n_triangles", ctypes.c_int64), ("unresolved", ctypes.c_int64), ("reserved", ctypes.c_int64 * 3)]


class ChunkMesh(ctypes.Structure):
    _fields_ = [("vertex_offset", ctypes.c_int64), ("n_vertices", ctypes.c_int64),
                ("triangle_offset", ctypes.c_int64), ("n_triangles", ctypes.c_int64)]


BATCH_SIZES = (32, 48, 64, 96, 128, 192)
BATCH_SHAPES = ("sphere", "box", "torus", "capsule")
BATCH_SHAPE_PARAMS = {"sphere": (1.0,), "box": (0.7, 0.7, 0.7, 0.2), "torus": (1.0, 0.4), "capsule": (1.2, 0.5)}


def batch_manifest(seed: int = 0, n: int = 64):
    """The uneven batch: n chunks, grid size ~ U{32,48,64,96,128,192}^3 over [-1.6,1.6]^3,
    shape ~ U{sphere, rounded cube, torus, capsule}, a gradient fBm layer with
    octaves ~ U{3..6} (f0 2.0, lacunarity 2.0, gain 0.5, weight 0.05) and a per-chunk
    seed, all drawn from numpy's default_rng(seed)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        size = int(rng.choice(BATCH_SIZES))
        shape = BATCH_SHAPES[int(rng.integers(0, 4))]
        octaves = int(rng.integers(3, 7))
        cseed = int(rng.integers(0, 2**31))
        out.append({"n": size, "shape": shape, "octaves": octaves, "seed": cseed})
    return out


def batch_chunks(manifest):
    """ctypes Chunk array for a manifest."""
    arr = (Chunk * len(manifest))()
    for i, m in enumerate(manifest):
        arr[i].spec = field_spec(m["shape"], BATCH_SHAPE_PARAMS[m["shape"]],
                                 [layer("gradient", "fbm", m["octaves"], 2.0, 2.0, 0.5, 0.05, m["seed"])])
        arr[i].grid = uniform_grid(-1.6, 1.6, m["n"])
        arr[i].iso = 0.0
    return arr


def nb_planets():
    """The narrow-band workload's eight planets of mixed shapes and seeds: a 5-octave
    gradient fBm (f0 2.0, lacunarity 2.0, gain 0.5, weight 0.05, seed 1000 + p) on sphere,
    rounded cube, torus and capsule, twice.  Returns [(name, FieldSpec)]."""
    out = []
    for p in range(8):
        shape = BATCH_SHAPES[p % 4]
        out.append((f"{shape}/seed {1000 + p}",
                    field_spec(shape, BATCH_SHAPE_PARAMS[shape], [layer("gradient", "fbm", 5, 2.0, 2.0, 0.5, 0.05,
                                                                        1000 + p)])))
    return out


def nb_band(g: Grid, lipschitz: float = 2.0) -> float:
    """A band that holds the surface: every point of an 8^3-cell block lies within half the
    block diagonal of a corner, so band = L * diagonal / 2 with L >= max |grad f|."""
    d = NB_BLOCK * math.sqrt(float(g.spacing[0]) ** 2 + float(g.spacing[1]) ** 2 + float(g.spacing[2]) ** 2)
    return lipschitz * d / 2


# ---------------------------------------------------------------------------- texel bake
class BakeSpec(ctypes.Structure):
    _fields_ = [("field", FieldSpec), ("detail", Layer), ("bands", ctypes.c_float * 3),
                ("band_width", ctypes.c_float), ("colors", (ctypes.c_float * 3) * 5),
                ("cliff", ctypes.c_float * 2), ("detail_albedo", ctypes.c_float),
                ("roughness", ctypes.c_float * 4), ("bump", ctypes.c_float * 2),
                ("reserved", ctypes.c_float * 2)]


assert ctypes.sizeof(BakeSpec) == ctypes.sizeof(FieldSpec) + 96 + 4 * (3 + 1 + 15 + 2 + 1 + 4 + 2 + 2)

# sand, grass, rock, snow, cliff: linear RGB
DEFAULT_COLORS = ((0.76, 0.70, 0.50), (0.25, 0.45, 0.15), (0.45, 0.40, 0.35), (0.95, 0.95, 0.97),
                  (0.35, 0.30, 0.28))


def bake_spec(field: FieldSpec, detail: Layer | None = None, bands=(-0.01, 0.005, 0.025), band_width=0.005,
              colors=DEFAULT_COLORS, cliff=(0.05, 0.25), detail_albedo=0.3, roughness=(0.6, 0.5, 0.2, -0.3),
              bump=(0.2, 5.0)) -> BakeSpec:
    """Material of the texel bake.  The default bands suit a land field whose surface sits
    within about +-0.05 of the base shape (elev = shape(p)/elevation_scale on the surface
    is the elevation stack's value there)."""
    b = BakeSpec()
    b.field = field
    b.detail = detail if detail is not None else layer("gradient", "fbm", 3, 16.0, 2.0, 0.5, 1.0, 977)
    for i in range(3):
        b.bands[i] = bands[i]
    b.band_width = band_width
    for i in range(5):
        for j in range(3):
            b.colors[i][j] = colors[i][j]
    b.cliff[0], b.cliff[1] = cliff
    b.detail_albedo = detail_albedo
    for i in range(4):
        b.roughness[i] = roughness[i]
    b.bump[0], b.bump[1] = bump
    return b


# ---------------------------------------------------------------------------- material stack
COLORING = {"default": 0, "barren": 1, "debug": 2}
PURPOSE = {"land": 1, "water": 2, "navigation": 3}
BIOMES = ("bare", "tundra", "taiga", "shrubland", "grassland", "temperate_seasonal_forest", "temperate_rain_forest",
          "desert", "savanna", "tropical_seasonal_forest", "tropical_rain_forest", "water")
TYPES = ("road", "flat", "rough", "cliffs", "water")


class MaterialSpec(ctypes.Structure):
    _fields_ = [("field", FieldSpec), ("coloring", ctypes.c_int32), ("purpose", ctypes.c_int32),
                ("poles", ctypes.c_int32), ("flowers", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("reserved", ctypes.c_int32 * 3)]


class Tile(ctypes.Structure):
    _fields_ = [("albedo", ctypes.c_float * 3), ("roughness", ctypes.c_float), ("metallic", ctypes.c_float),
                ("height", ctypes.c_float), ("opacity", ctypes.c_float), ("elevation", ctypes.c_float),
                ("slope", ctypes.c_float), ("temperature", ctypes.c_float), ("precipitation", ctypes.c_float),
                ("biome", ctypes.c_int32), ("type", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


assert ctypes.sizeof(MaterialSpec) == ctypes.sizeof(FieldSpec) + 32
assert ctypes.sizeof(Tile) == 64
TILE_FLOATS = ("r", "g", "b", "roughness", "metallic", "height", "opacity", "elevation", "slope", "temperature",
               "precipitation")


def material_spec(field: FieldSpec, coloring="default", purpose="land", poles=True, flowers=True,
                  seed: int = 1) -> MaterialSpec:
    """A material stack over a planet spec (the reference's coloring mode and the
    poles / flowers configuration switches)."""
    m = MaterialSpec()
    m.field = field
    m.coloring = COLORING[coloring] if isinstance(coloring, str) else int(coloring)
    m.purpose = PURPOSE[purpose] if isinstance(purpose, str) else int(purpose)
    m.poles, m.flowers = int(bool(poles)), int(bool(flowers))
    m.seed = int(seed) & 0xFFFFFFFF
    return m


def tiles_fields(raw):
    """Named columns of a [n, 16] float32 tile array (numpy): floats per TILE_FLOATS, then
    biome and type as int32."""
    import numpy as np
    raw = np.ascontiguousarray(raw, dtype=np.float32).reshape(-1, 16)
    out = {k: raw[:, i] for i, k in enumerate(TILE_FLOATS)}
    out["biome"] = raw[:, 11].view(np.int32)
    out["type"] = raw[:, 12].view(np.int32)
    return out
