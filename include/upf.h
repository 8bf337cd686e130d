/*
 * upf.h — C-ABI of libupf, the MI355X (gfx950) field-evaluation and meshing path
 * of unnatural-planets.
 *
 * The reference hands its engine per-point callbacks; this library takes whole
 * grids, point sets, meshes and atlases instead:
 *
 *   reference boundary                                         replaced by
 *   ---------------------------------------------------------  -----------------------------
 *   MarchingCubes::updateByPosition(Delegate<Real(const Vec3&)>)   upf_field_eval_grid
 *     bound to terrainSdfLand/Water/Navigation                      (compose = LAND/WATER/NAV)
 *     (sources/meshGeneration.cpp:38, modesConfiguration.cpp:248-279)
 *   MarchingCubes::makeMesh()  (sources/meshGeneration.cpp:39)    upf_mc_extract
 *   terrainSdf*(pos) called point-wise (tileProperties.cpp:258,
 *     coloringModes.cpp:51-57, meshGeneration.cpp:68)              upf_field_eval_points
 *   meshGenerateTexture(mesh, {w, h, generator = pixel})
 *     (sources/textures.cpp:65-71, pixel at :24-43)                upf_bake
 *   (no counterpart: gradient of the field w.r.t. position and     upf_field_adjoint,
 *    spec parameters)                                              upf_field_adjoint_grid
 *   (no counterpart: many chunks in one launch)                    upf_batch_*
 *
 * Conventions
 *  - Every device pointer is caller-allocated HIP device memory; the library never
 *    allocates or frees caller memory.  Scratch comes in as an explicit workspace
 *    (upf_*_workspace_bytes tells how much).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).  Calls are
 *    stream-ordered.  Calls that must return counts (upf_mc_extract, upf_mc_count,
 *    upf_batch_totals) synchronise the stream before they return.
 *  - Errors are int status codes; nothing throws across the ABI.
 *  - No global mutable state: seeds live in the spec (the reference draws them from a
 *    process-global RNG, sources/math.cpp:73-77).
 *
 * Field definition (all FP32):
 *   elev(p) = layer stack (see upf_layer)
 *   LAND        f = shape(p) - elevation_scale * elev(p)             (modesConfiguration.cpp:248-258)
 *   WATER       f = shape(p)                                         (:260-267)
 *   NAVIGATION  f = shape(p) - elevation_scale * max(elev(p), 0)     (:269-279)
 *   ELEVATION   f = shape(p) / elevation_scale                       (:230-237, ratio 10)
 *   ELEVATION_RAW f = elev(p)                                        (:239-246)
 * Negative is inside.  The reference uses elevation_scale = 1/10 (meshElevationRatio,
 * modesConfiguration.cpp:227).
 *
 * Grid: node (i,j,k) sits at origin + (i,j,k)*spacing, each coordinate computed as
 * origin[a] + (float)i * spacing[a] with two IEEE roundings (no fused multiply-add);
 * storage is float32 row-major [k][j][i] (x fastest).
 */
#ifndef UPF_H_INCLUDED
#define UPF_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UPF_VERSION 1

#define UPF_MAX_LAYERS 4
#define UPF_MAX_OCTAVES 8
#define UPF_MAX_SHAPE_PARAMS 8

/* ---- status codes ------------------------------------------------------------ */
enum upf_status {
  UPF_OK = 0,
  UPF_ERR_INVALID_VALUE = 1, /* NaN/Inf in a spec parameter or in the produced field
                                (modesConfiguration.cpp:234-277 throws on !valid) */
  UPF_ERR_BAD_SPEC = 2,      /* unknown shape/noise/fractal/blend/compose id, bad counts */
  UPF_ERR_CAPACITY = 3,      /* output buffers too small; required counts written back */
  UPF_ERR_WORKSPACE = 4,     /* workspace missing or too small */
  UPF_ERR_BAD_GRID = 5,      /* dims < 2 or too large, spacing not finite */
  UPF_ERR_HIP = 6,           /* a HIP runtime call failed */
  UPF_ERR_NULL = 7,          /* a required pointer is NULL */
  UPF_ERR_EMPTY_MESH = 8,    /* used by host mirrors: meshGenerateBase* on an empty mesh
                                (meshGeneration.cpp:50-51, 104-105) */
  UPF_ERR_BAND = 9           /* narrow band: the surface leaves the active blocks     */
};

/* ---- spec enums ---------------------------------------------------------------- */
enum upf_shape_id {
  UPF_SHAPE_NONE = 0,    /* shape(p) = 0                                              */
  UPF_SHAPE_SPHERE = 1,  /* params: r                    |p| - r   (sdf.cpp:20-23)        */
  UPF_SHAPE_BOX = 2,     /* params: hx, hy, hz, rounding  box - rounding (sdf.cpp:40-48)  */
  UPF_SHAPE_TORUS = 3,   /* params: R, r   ring in the xz-plane, axis y (sdf.cpp:81-84)   */
  UPF_SHAPE_CAPSULE = 4, /* params: h, r   segment z in [-h/2, h/2] (sdf.cpp:25-28 swaps
                                            y and z before cage's y-axis capsule)          */
  /* Extended shapes (the reference's compositions in sdf.cpp over textbook primitives).
   * The adjoint reports their position gradient but no shape-parameter gradient. */
  UPF_SHAPE_CYLINDER = 5,/* params: h, r   axis z, |z| <= h/2   (sdf.cpp:30-38 swizzled)  */
  UPF_SHAPE_PLANE = 6,   /* params: nx, ny, nz (unit), d     dot(p,n) - d  (sdf.cpp:10-18)*/
  UPF_SHAPE_OCTAHEDRON = 7, /* params: s, rounding     (sdf.cpp:55-58)                    */
  UPF_SHAPE_DOUBLE_TORUS = 8,  /* params: R, r, k   smoothMin of two tori (sdf.cpp:292-299)  */
  UPF_SHAPE_TORUS_CROSS = 9,   /* params: R, r, k   (sdf.cpp:301-308)                      */
  UPF_SHAPE_H2O = 10,          /* no params, metre scale  (sdf.cpp:141-149)                */
  UPF_SHAPE_BOWL = 11,         /* no params, metre scale  (sdf.cpp:310-313)                */
  UPF_SHAPE_INSIDE_CUBE = 12,  /* no params, metre scale  (sdf.cpp:315-318)                */
  UPF_SHAPE_WORMHOLE = 13,     /* no params, metre scale  (sdf.cpp:399-404)                */
  UPF_SHAPE_PIPE = 14,         /* no params, metre scale  (sdf.cpp:355-366)                */
  UPF_SHAPE_HEMISPHERES = 15,  /* no params, metre scale  (sdf.cpp:383-397)                */
  /* Further shapes of sdf.cpp (metre scale).  cage's atan2(a, b) is read as C's atan2(a, b),
   * quaternion rolls as rotations about z; textbook primitives where cage supplies one. */
  UPF_SHAPE_H3O = 16,          /* no params (sdf.cpp:151-163)                               */
  UPF_SHAPE_H4O = 17,          /* no params (sdf.cpp:165-178)                               */
  UPF_SHAPE_GEAR = 18,         /* no params (sdf.cpp:180-204)                               */
  UPF_SHAPE_MANDELBULB = 19,   /* no params, power 3, 20 iterations (sdf.cpp:206-232)       */
  UPF_SHAPE_MOBIUS = 20,       /* no params (sdf.cpp:94-120); on the axis, where the
                                  reference normalizes a zero vector, the projection is 0   */
  UPF_SHAPE_FIBERS = 21,       /* params: gyroid offset ox, oy, oz (the reference draws it
                                  from randomRange3(-100, 100), sdf.cpp:122-139)            */
  UPF_SHAPE_ASTEROID = 22,     /* params: frequency (the reference draws 0.65 .. 0.75), seed
                                  (an integer < 2^24); simplex fBm 2 octaves, gain 0.3, of
                                  the unit direction (sdf.cpp:333-353)                     */
  UPF_SHAPE_TWISTED_PLANE = 23,/* no params (sdf.cpp:368-373)                               */
  UPF_SHAPE_BELT = 24,         /* no params (sdf.cpp:375-381)                               */
  UPF_SHAPE_TRIANGULAR_PRISM = 25, /* no params; exact point-triangle distance (Ericson)
                                      to the reference's triangle (sdf.cpp:62-74)           */
  UPF_SHAPE_HEXAGONAL_PRISM = 26,  /* no params; along y, apothem 400, length 300
                                      (Quilez's hexagonal prism; sdf.cpp:76-79)             */
  UPF_SHAPE_TWISTED_HEX_PRISM = 27,/* no params; apothem 650, length 900 (sdf.cpp:234-242) */
  UPF_SHAPE_TETRAHEDRON = 28,  /* no params; regular, circumradius 600, exact distance,
                                  minus 100 (sdf.cpp:50-53)                                 */
  UPF_SHAPE_SINE_NET = 29,     /* params: scale, d1, d2; weights: shape_weights (below).
                                  The architecture of sdfBunny / sdfMonkeyHead
                                  (sdf.cpp:252-303), whose weight tables the caller passes */
  UPF_SHAPE_KNOT = 30,         /* no params; sdfKnot = knot(p * (1, 1, 0.3)) / |(1, 1, 0.3)|
                                  (sdf.cpp:86-90) with cage's knot(q, 1600, 1.5) taken as
                                  the (3, 2) torus knot of ratio 1.5 about the z axis: core
                                  radius 1600, strands 200 off the core circle at angle
                                  1.5 t, tube radius 100; distance in the half-plane of q's
                                  azimuth phi to the two strands crossing it (t = phi,
                                  phi + 2 pi), minus the tube radius                        */
  UPF_SHAPE_COUNT = 31
};

/* UPF_SHAPE_SINE_NET: a sine network of width 16 evaluated at q = (x, z, y) * scale:
 *   q outside the unit ball: f = (|q| - 0.8) / scale;
 *   h0 = sin(A q.y + B q.z + C q.x + D)                   (16 units, 4 blocks of 4)
 *   h1[o] = sin(sum_b M1[o][b] h0[b] + V1[o]) / d1 + h0[o]   (block o = 0..3)
 *   h2[o] = sin(sum_b M2[o][b] h1[b] + V2[o]) / d2 + h1[o]
 *   f = (sum_b dot(h2[b], U[b]) + c) / scale.
 * shape_weights points to UPF_SINE_NET_WEIGHTS floats in the memory the evaluating call
 * reads (device memory for the GPU entry points), in the order the reference's source
 * writes them:
 *   [0, 64)     layer 0, block b at 16 b: A[4], B[4], C[4], D[4] (sdfBunnyImpl's
 *               "- p[0] * Vec4(...)" is C negated);
 *   [64, 336)   layer 1, block o at 64 + 68 o: M1[o][0..3] (16 floats each, the Mat4
 *               constructor's order, column-major: element 4 col + row), then V1[o][4];
 *   [336, 608)  layer 2, the same layout;
 *   [608, 624)  U[0..3] (4 floats each); [624] c.
 * The reference uses scale 0.0005, d1 1.0, d2 1.4 for both shapes. */
#define UPF_SINE_NET_WEIGHTS 625

enum upf_noise_id {
  UPF_NOISE_GRADIENT = 1, /* 3D lattice gradient noise, quintic fade, 12-edge gradients   */
  UPF_NOISE_VALUE = 2,    /* 3D lattice value noise, quintic fade                         */
  UPF_NOISE_CELLULAR = 3, /* F1 Euclidean distance to jittered feature points, 27 cells   */
  UPF_NOISE_SIMPLEX = 4,  /* 3D simplex noise (Perlin 2001): skew 1/3, unskew 1/6, the 4
                             corners of the containing simplex, kernel (0.5 - r^2)^4 times
                             the corner gradient's dot product, scaled by 72.  The radius
                             0.5 keeps it continuous.  Gradients and hash are the lattice
                             noise's; cage's Simplex (sources/elevationModes.cpp:19) is
                             this noise's role in the reference                          */
  UPF_NOISE_COUNT = 5
};

/* Elevation modes: the reference's fixed elevation functions
 * (sources/elevationModes.cpp:9-198), selected with upf_field_spec.elevation_mode in
 * place of the generic layer stack.  Each mode reads the noise layers it needs from
 * layers[0..n) in the order listed (upf_spec_set_elevation_mode fills them with the
 * reference's configurations); every layer's weight and blend are ignored.  Results are
 * in metres, so elevation_scale = 1/10 gives the reference's meshElevationRatio
 * (sources/modesConfiguration.cpp:227).
 *   NONE     elev = 100                                                        (:9-12)
 *   SIMPLE   1 layer: a = 0.3 - n0; a = (1.3 a - 0.35)^3 + 0.1; 100 - 1000 a   (:14-32)
 *   LEGACY   2 layers: s = n0 * 0.0005 + 0.0015; a = n1(p * s) + 0.11;
 *            a = a < 0 ? -(-a)^0.85 : a^1.7; 2500 a                             (:34-64)
 *   LAKES    4 layers: land = 1 - saturate(0.5 n0 + 0.5)^1.24, then
 *            l = 2 land - 1, land = 150 (l / (|l| + 0.17) + 0.15), then the mountains
 *            of n1 (mask), n2 (ridges), n3 (terraces)                  (:68-123, 129-149)
 *   ISLANDS  as LAKES with exponent 0.83                                   (:151-171)
 *   CRATERS  2 layers: layers[0] is the impact Voronoi (sources/voronoi.cpp:10-80):
 *            seed = its seed, octaves = points per cell (1..8), freq[0] = cell size in
 *            metres; c = the nearest point, s = 100 + 50 n1(c), d = |p - c|,
 *            x = saturate(d / s), h = 10 (x^4 - 0.9 x^5 - 0.1), elev = 200 + 190 h
 *                                                                            (:173-198)
 * The Voronoi's integer hash (cage's hash(), not in the reference checkout) is the
 * murmur3 32-bit finaliser.  pow is the C library's powf.  The parameter gradient is
 * defined for the generic layer stack only: with a mode, upf_field_adjoint with
 * d_dLdtheta and upf_field_adjoint_grid return UPF_ERR_BAD_SPEC, while the position
 * gradient (normals) is available through upf_field_eval_points.  Grids with a mode are
 * evaluated per lane (UPF_EVAL_PER_LANE). */
enum upf_elevation_mode_id {
  UPF_ELEV_LAYERS = 0,  /* the generic layer stack (blend ops below)                    */
  UPF_ELEV_NONE = 1,
  UPF_ELEV_SIMPLE = 2,
  UPF_ELEV_LEGACY = 3,
  UPF_ELEV_LAKES = 4,
  UPF_ELEV_ISLANDS = 5,
  UPF_ELEV_CRATERS = 6,
  UPF_ELEV_COUNT = 7
};

enum upf_fractal_id {
  UPF_FRACTAL_NONE = 0,   /* one octave, amp[0] = 1                                       */
  UPF_FRACTAL_FBM = 1,    /* sum_k amp[k] * n_k                                            */
  UPF_FRACTAL_RIDGED = 2, /* sum_k amp[k] * (1 - 2|n_k|)                                  */
  UPF_FRACTAL_COUNT = 3
};

enum upf_blend_id {
  UPF_BLEND_ADD = 0,  /* elev += weight * v                                               */
  UPF_BLEND_MASK = 1, /* m = saturate(weight * v + 0.5)   (sets the blend mask, elev kept) */
  UPF_BLEND_LERP = 2, /* elev += m * (weight * v - elev)  (lerp towards this layer by m)  */
  UPF_BLEND_COUNT = 3
};

enum upf_compose_id {
  UPF_COMPOSE_LAND = 0,
  UPF_COMPOSE_WATER = 1,
  UPF_COMPOSE_NAVIGATION = 2,
  UPF_COMPOSE_ELEVATION = 3,
  UPF_COMPOSE_ELEVATION_RAW = 4,
  UPF_COMPOSE_COUNT = 5
};

/* One noise layer.  Octave k samples noise with seed (seed + k) at p * freq[k] and is
 * weighted by amp[k]; the per-octave arrays are part of the interface so that the
 * adjoint can report d/d freq[k] and d/d amp[k] and so that a caller can set them
 * directly.  upf_layer_init() fills them from (frequency, lacunarity, gain) the way
 * the reference's NoiseFunctionCreateConfig does (sources/elevationModes.cpp:83-87):
 *   freq[0] = frequency, freq[k] = freq[k-1] * lacunarity        (float, in order)
 *   amp[0]  = 1 / sum_{j<octaves} |gain|^j,  amp[k] = amp[k-1] * gain
 * (gain may be negative, sources/elevationModes.cpp:85). */
typedef struct upf_layer {
  int32_t noise;    /* upf_noise_id   */
  int32_t fractal;  /* upf_fractal_id */
  int32_t octaves;  /* 1..UPF_MAX_OCTAVES */
  int32_t blend;    /* upf_blend_id   */
  uint32_t seed;
  float weight;     /* layer amplitude (the "Voronoi weight" of the adjoint) */
  float jitter;     /* cellular only: feature-point jitter in [0, 1]          */
  float reserved;
  float freq[UPF_MAX_OCTAVES];
  float amp[UPF_MAX_OCTAVES];
} upf_layer;

typedef struct upf_field_spec {
  int32_t shape;    /* upf_shape_id   */
  int32_t compose;  /* upf_compose_id */
  int32_t n_layers; /* 0..UPF_MAX_LAYERS */
  int32_t elevation_mode; /* upf_elevation_mode_id; 0 = the generic layer stack */
  float shape_params[UPF_MAX_SHAPE_PARAMS];
  float elevation_scale;
  int32_t reserved;
  const float* shape_weights; /* UPF_SHAPE_SINE_NET's weights, else NULL */
  upf_layer layers[UPF_MAX_LAYERS];
} upf_field_spec;

typedef struct upf_grid {
  float origin[3];
  float spacing[3];
  int32_t dims[3]; /* nodes per axis, >= 2 */
  int32_t reserved;
} upf_grid;

/* Gradient of a scalar loss w.r.t. the spec's continuous parameters (the θ of the
 * adjoint).  Same indexing as upf_field_spec. */
typedef struct upf_theta_grad {
  float shape_params[UPF_MAX_SHAPE_PARAMS];
  float weight[UPF_MAX_LAYERS];
  float freq[UPF_MAX_LAYERS][UPF_MAX_OCTAVES];
  float amp[UPF_MAX_LAYERS][UPF_MAX_OCTAVES];
} upf_theta_grad; /* 76 floats */
#define UPF_THETA_COUNT 76

/* Marching-cubes output.  positions/normals: float[cap_vertices][3]; indices:
 * uint32[cap_triangles][3].  normals may be NULL.  n_vertices / n_triangles are
 * written back (also when UPF_ERR_CAPACITY is returned, then as the required size). */
typedef struct upf_mesh_buffers {
  float* positions;
  float* normals;
  uint32_t* indices;
  int64_t cap_vertices;
  int64_t cap_triangles;
  int64_t n_vertices;
  int64_t n_triangles;
} upf_mesh_buffers;

/* ---- library ------------------------------------------------------------------- */
int upf_version(void);
const char* upf_status_string(int status);

/* Fill layer->freq/amp from (frequency, lacunarity, gain) as described at upf_layer.
 * Replaces newNoiseFunction(NoiseFunctionCreateConfig) (sources/elevationModes.cpp:16-27). */
int upf_layer_init(upf_layer* layer, int32_t noise, int32_t fractal, int32_t octaves,
                   float frequency, float lacunarity, float gain, float weight,
                   uint32_t seed, float jitter, int32_t blend);

/* Set spec->elevation_mode and fill spec->layers / n_layers with the noise
 * configurations of that mode in the reference (sources/elevationModes.cpp:14-198), the
 * k-th noise seeded with seed + k (the reference draws each from noiseSeed(),
 * sources/math.cpp:73-77).  cage's unset NoiseFunctionCreateConfig fields are taken as
 * lacunarity 2 and gain 0.5.  The shape, composition and elevation_scale are kept. */
int upf_spec_set_elevation_mode(upf_field_spec* spec, int32_t mode, uint32_t seed);

/* Checks ids, counts and that every parameter is finite.  Replaces the choose*Function
 * validation (sources/modesConfiguration.cpp:131-135, 190-194) and the NaN throws. */
int upf_spec_validate(const upf_field_spec* spec);

/* FP32 flops per sample of the forward field, by the counting rule FMA = 2,
 * add/sub/mul/div/sqrt = 1, compare/select/floor/convert/integer = 0. */
double upf_flops_per_sample(const upf_field_spec* spec);
/* Same for the fused forward + adjoint (value, d/dx, d/dθ accumulation). */
double upf_adjoint_flops_per_sample(const upf_field_spec* spec);

/* f at every grid node -> d_field[dims[2]][dims[1]][dims[0]].
 * Replaces cubes->updateByPosition(Delegate(terrainSdf*)) (sources/meshGeneration.cpp:38). */
int upf_field_eval_grid(const upf_field_spec* spec, const upf_grid* grid, float* d_field,
                        void* stream);

/* Launch options of the grid evaluation (the knobs the batch autotuner searches). */
enum upf_eval_variant {
  UPF_EVAL_AUTO = 0,      /* wave-cooperative rows: lattice points hashed once per wave */
  UPF_EVAL_ROWS = 1,      /* same as AUTO                                               */
  UPF_EVAL_PER_LANE = 2   /* every lane hashes its own corners / cells                 */
};
typedef struct upf_eval_opts {
  int32_t variant; /* upf_eval_variant */
  int32_t reserved[3];
} upf_eval_opts;

/* upf_field_eval_grid with explicit launch options (NULL = defaults).  No counterpart
 * in the reference.  Both variants use the same lattice points, hashes and gradients;
 * UPF_EVAL_ROWS regroups the interpolation sums (a few ulps from UPF_EVAL_PER_LANE). */
int upf_field_eval_grid_ex(const upf_field_spec* spec, const upf_grid* grid, float* d_field,
                           const upf_eval_opts* opts, void* stream);

/* f at n points d_xyz[n][3] -> d_f[n]; if d_dfdx != NULL also the analytic gradient
 * d_dfdx[n][3] (exact normals).  Replaces point-wise terrainSdf* calls
 * (sources/coloringModes.cpp:51-57,122-151; sources/tileProperties.cpp:258-273). */
int upf_field_eval_points(const upf_field_spec* spec, const float* d_xyz, int64_t n,
                          float* d_f, float* d_dfdx, void* stream);

/* Fused forward + adjoint at n points.  d_f (nullable) gets f, d_dfdx (nullable) gets
 * df/dx, and d_dLdtheta (a device upf_theta_grad, nullable) gets sum_i dLdf[i] * df_i/dθ
 * ADDED to its current contents (the caller zeroes it).  d_dLdf may be NULL only when
 * d_dLdtheta is NULL.  No counterpart in the reference. */
int upf_field_adjoint(const upf_field_spec* spec, const float* d_xyz, int64_t n,
                      const float* d_dLdf, float* d_f, float* d_dfdx, float* d_dLdtheta,
                      void* stream);

/* The same over grid nodes (positions generated, not read). */
int upf_field_adjoint_grid(const upf_field_spec* spec, const upf_grid* grid,
                           const float* d_dLdf, float* d_f, float* d_dfdx,
                           float* d_dLdtheta, void* stream);

/* Marching cubes over a field produced by upf_field_eval_grid (or any FP32 grid in the
 * same layout).  Case code bit c set <=> f(corner c) < iso.  Vertex = linear zero
 * crossing on the cell edge that owns it; vertices are numbered by (owning node linear
 * index, axis x<y<z), triangles by (cell linear index, table order).  Normals are the
 * normalised grid gradient (central differences, one-sided at the border) interpolated
 * along the edge; they point towards increasing f (outward).  Triangles wind
 * counter-clockwise seen from outside.  Synchronises `stream`.
 * Replaces cubes->makeMesh() (sources/meshGeneration.cpp:39). */
size_t upf_mc_workspace_bytes(const upf_grid* grid);
int upf_mc_extract(const float* d_field, const upf_grid* grid, float iso,
                   upf_mesh_buffers* out, void* d_workspace, size_t workspace_bytes,
                   void* stream);

/* The two phases of upf_mc_extract, for callers that size outputs exactly:
 * upf_mc_count classifies and scans (leaving its state in the workspace), synchronises
 * the stream and returns the vertex/triangle counts; upf_mc_emit then writes the mesh
 * from that state (same field, grid, iso and workspace; asynchronous). */
int upf_mc_count(const float* d_field, const upf_grid* grid, float iso, void* d_workspace,
                 size_t workspace_bytes, int64_t* n_vertices, int64_t* n_triangles, void* stream);
int upf_mc_emit(const float* d_field, const upf_grid* grid, float iso, upf_mesh_buffers* out,
                void* d_workspace, size_t workspace_bytes, void* stream);

/* upf_mc_extract with no host synchronisation, for a stored field: classification, scan
 * and emission in stream order, the mesh into `out` within its capacities and the
 * required counts into d_counts[0] (vertices), d_counts[1] (triangles), device memory,
 * exactly as upf_field_mesh_async does after its field pass (same truncation rule when a
 * count exceeds its capacity).  Mesh and counts equal upf_mc_extract's bit for bit.
 * Replaces cubes->makeMesh() (sources/meshGeneration.cpp:39) for callers that keep
 * buffers sized by an earlier call. */
int upf_mc_extract_async(const float* d_field, const upf_grid* grid, float iso, upf_mesh_buffers* out,
                         int64_t* d_counts, void* d_workspace, size_t workspace_bytes, void* stream);

/* Field evaluation fused with the counting half of marching cubes: evaluates the field
 * into d_field (as upf_field_eval_grid_ex) while the field kernel stores each 64-node
 * group's sign mask, then classifies from the sign masks alone (no second read of the
 * field) and scans.  The counts and the workspace are those of upf_mc_count on the same
 * field, so upf_mc_emit follows as after upf_mc_count.  Replaces the pair
 * updateByPosition + makeMesh of meshGenerateGeneric (sources/meshGeneration.cpp:38-39)
 * up to the emission.  Synchronises the stream (16 B read-back). */
int upf_field_mesh_count(const upf_field_spec* spec, const upf_grid* grid, float iso, float* d_field,
                         const upf_eval_opts* opts, void* d_workspace, size_t workspace_bytes,
                         int64_t* n_vertices, int64_t* n_triangles, void* stream);

/* upf_field_mesh_count followed by upf_mc_emit in one stream-ordered call, with no host
 * synchronisation.  The mesh goes to `out` within its capacities (out->cap_vertices,
 * out->cap_triangles; out->n_vertices / n_triangles are neither read nor written), and
 * the required counts go to d_counts[0] (vertices) and d_counts[1] (triangles), device
 * memory, in stream order.  Where a count exceeds its capacity the writes past it are
 * dropped and the mesh is incomplete: the caller compares after its own synchronisation
 * and calls again with larger buffers.  For loops that keep buffers sized by an earlier
 * call (re-meshing a field whose parameters change a little), this removes the count
 * read-back and the host round trip between classification and emission.  Mesh and
 * counts are those of upf_field_mesh_count + upf_mc_emit bit for bit.  Replaces
 * updateByPosition + makeMesh of meshGenerateGeneric (sources/meshGeneration.cpp:38-40). */
int upf_field_mesh_async(const upf_field_spec* spec, const upf_grid* grid, float iso, float* d_field,
                         const upf_eval_opts* opts, upf_mesh_buffers* out, int64_t* d_counts,
                         void* d_workspace, size_t workspace_bytes, void* stream);

/* upf_field_mesh_async without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The workspace is a library-owned device
 * array of upf_static_scratch_bytes() = 64 MiB per GPU, held until the process ends,
 * which fits the grids whose upf_mc_workspace_bytes is at most that (up to 512^3 nodes);
 * larger grids return UPF_ERR_WORKSPACE.  The call allocates nothing and clears nothing:
 * every word of the workspace is written before it is read.  All calls on one GPU share
 * that GPU's array, so calls that overlap on one GPU (other streams, other threads)
 * race: order them on one stream.  Mesh and counts equal upf_field_mesh_async's. */
size_t upf_static_scratch_bytes(void);
int upf_field_mesh_async_static(const upf_field_spec* spec, const upf_grid* grid, float iso, float* d_field,
                                const upf_eval_opts* opts, upf_mesh_buffers* out, int64_t* d_counts,
                                void* stream);

/* Case code of every cell -> d_codes[dims[2]-1][dims[1]-1][dims[0]-1] (uint8). */
int upf_mc_case_codes(const float* d_field, const upf_grid* grid, float iso,
                      uint8_t* d_codes, void* stream);

/* ---- mesh post-processing -------------------------------------------------------------
 * The steps after marching cubes in meshGenerateGeneric / meshGenerateBaseWater
 * (sources/meshGeneration.cpp:41-42, 62-95).  cage's meshRemoveDisconnected and
 * meshRemoveInvalid are not in the reference checkout; the rules here are documented
 * choices.  Both write an order-preserving compaction of `in` into `out` (kept vertex i
 * keeps its order among the kept vertices, kept triangle t among the kept triangles;
 * normals copied when both normal pointers are set), return UPF_ERR_CAPACITY with the
 * counts in out when out is too small (its buffers then left untouched), and synchronise
 * the stream.  meshFlipNormals
 * (:42) has no counterpart: upf_mc_emit already writes outward normals and CCW-outward
 * winding.  Triangles naming a vertex outside [0, in->n_vertices) are dropped. */
size_t upf_mesh_post_workspace_bytes(int64_t n_vertices, int64_t n_triangles);
/* meshRemoveDisconnected: keep the connected component (vertices joined by triangle
 * edges) with the most triangles; ties go to the component holding the smaller vertex
 * id.  Unreferenced vertices are removed. */
int upf_mesh_remove_disconnected(const upf_mesh_buffers* in, upf_mesh_buffers* out, void* d_workspace,
                                 size_t workspace_bytes, void* stream);
/* Water-mesh vertex mask (meshGeneration.cpp:62-95): a vertex is valid when the raw
 * elevation (spec with compose = ELEVATION_RAW, terrainSdfElevationRaw) at it is below
 * `threshold` (the reference uses 0.1); two passes then make all three vertices of every
 * triangle touching a valid vertex valid; invalid vertices and the triangles naming them
 * are removed (meshRemoveInvalid). */
int upf_mesh_water_mask(const upf_field_spec* spec, float threshold, const upf_mesh_buffers* in,
                        upf_mesh_buffers* out, void* d_workspace, size_t workspace_bytes, void* stream);

/* ---- narrow band ----------------------------------------------------------------------
 * Marching cubes of a fine grid evaluated only near the surface (no counterpart in the
 * reference, which sweeps the dense grid, sources/meshGeneration.cpp:35-39).  The fine
 * grid's cells are cut into blocks of UPF_NB_BLOCK^3 cells ((dims[a] - 1) must be a
 * multiple of UPF_NB_BLOCK).  The field is first evaluated at the block corners (the fine
 * nodes whose indices are multiples of UPF_NB_BLOCK); a block is active when one of its
 * 8 corner values lies within `band` of iso (|f - iso| < band) or its corners straddle
 * iso.  Active blocks are evaluated at every fine node and meshed.
 *
 * The mesh is the dense marching-cubes mesh of the fine grid (same vertex positions, same
 * triangles, per-lane field values) whenever every block the surface passes through is
 * active -- guaranteed when band >= L * (block diagonal) / 2 for a Lipschitz bound L of
 * f (every point of a block lies within half its diagonal of one of its corners).
 * Order differs from upf_mc_extract: vertices by (active block, node within the block's
 * 9^3 brick, axis x<y<z), triangles by (active block, cell).  Normals are the normalized
 * analytic gradient of f at each vertex (the fused forward + adjoint evaluation), not the
 * interpolated central differences of upf_mc_emit.
 *
 * d_vertex_keys (optional): per vertex the edge it lies on, (node linear index) * 3 + axis
 * in the fine grid -- a canonical id, equal for the same vertex in any extraction.
 * Errors: UPF_ERR_WORKSPACE (more active blocks than the workspace holds; stats->
 * active_blocks written), UPF_ERR_CAPACITY (counts written to stats and to out),
 * UPF_ERR_BAND (a crossing edge on an active block's face belongs to an inactive block:
 * band too narrow; stats->unresolved written).  Synchronises the stream twice (16 B
 * and 32 B read-backs). */
#define UPF_NB_BLOCK 8
typedef struct upf_nb_stats {
  int64_t blocks;        /* all blocks of the fine grid                         */
  int64_t active_blocks; /* blocks evaluated at fine resolution                  */
  int64_t n_vertices;
  int64_t n_triangles;
  int64_t unresolved;    /* crossing face edges whose owning block is inactive   */
  int64_t reserved[3];
} upf_nb_stats;

size_t upf_nb_workspace_bytes(const upf_grid* fine, int64_t max_active_blocks);
int upf_nb_extract(const upf_field_spec* spec, const upf_grid* fine, float iso, float band,
                   upf_mesh_buffers* out, uint64_t* d_vertex_keys, void* d_workspace,
                   size_t workspace_bytes, upf_nb_stats* stats, void* stream);

/* ---- texel bake ---------------------------------------------------------------------
 * Replaces meshGenerateTexture(mesh, {w, h, generator = Generator::pixel})
 * (sources/textures.cpp:65-71): texels are rasterized in UV space (a texel belongs to
 * the triangle whose UV triangle holds its centre ((x + 0.5)/w, (y + 0.5)/h); where UV
 * triangles overlap the larger triangle id wins), then each covered texel interpolates
 * position and normal (positionAt / normalAt, textures.cpp:27-28) and evaluates the
 * material; uncovered texels stay NaN (textures.cpp:47-63).
 *
 * Material (a reduced, deterministic form of coloringDefault, coloringModes.cpp:905-930):
 *   elev  = shape(p) / elevation_scale           (terrainSdfElevation, coloringModes.cpp:51-57)
 *   slope = 1 - dot(n, p/|p|)                      (0 flat .. 2 overhang; the reference
 *                                                  measures slope with 8 probes, :122-151)
 *   d     = detail.weight * (the detail layer's fractal value at p)
 *   albedo: sand -> grass -> rock -> snow by smoothstep bands of elev, cliff colour by a
 *           smoothstep of slope, times (1 + detail_albedo * d); saturated (:895-902)
 *   roughness = saturate(r0 + r_slope * slope + r_detail * d + r_snow * snow)
 *   bump      = saturate(0.5 + b_detail * d + b_elev * elev)   (the reference's height map)
 */
typedef struct upf_bake_spec {
  upf_field_spec field; /* shape and elevation_scale give elev */
  upf_layer detail;     /* detail noise: weight * fractal value (blend ignored)             */
  float bands[3];       /* elev of the sand|grass, grass|rock, rock|snow transitions        */
  float band_width;     /* half width of each transition                                    */
  float colors[5][3];   /* sand, grass, rock, snow, cliff (linear RGB)                      */
  float cliff[2];       /* slope where the cliff colour starts and is full                  */
  float detail_albedo;  /* albedo modulation by d                                           */
  float roughness[4];   /* r0, r_slope, r_detail, r_snow                                    */
  float bump[2];        /* b_detail, b_elev                                                 */
  float reserved[2];
} upf_bake_spec;

/* Deterministic per-triangle atlas: triangles 2c and 2c+1 share square cell c (cells in
 * row-major order, side = the largest integer number of texels that fits every cell in
 * width x height); triangle 2c takes the lower-left half (0,0) (s,0) (0,s) of the cell's
 * diagonal split, 2c+1 the upper-right half (s,s) (0,s) (s,0), corners on texel
 * boundaries (texel centres on the diagonal go to 2c+1).
 * d_uv: float[n_triangles][3][2], UV per triangle corner in [0,1].
 * UPF_ERR_CAPACITY if a cell would be smaller than 2 x 2 texels. */
int upf_atlas_per_triangle(int64_t n_triangles, int32_t width, int32_t height, float* d_uv,
                           void* stream);
/* UV rasterization: d_tri[height][width] = covering triangle id or -1. */
int upf_bake_rasterize(const float* d_uv, int64_t n_triangles, int32_t width, int32_t height,
                       int32_t* d_tri, void* stream);
/* The same map, tile-binned: triangles are binned into 64 x 16 texel tiles, each tile is
 * resolved in on-chip memory and every texel is stored once (no global atomics for
 * triangles over at most four tiles while a tile's bin of 1024 has room, no separate -1
 * fill).  d_tri equals upf_bake_rasterize's bit for bit.  The workspace (size from
 * upf_bake_rasterize_workspace_bytes, 0 for invalid sizes) holds the bins, 4 bytes per
 * texel, and a list of 16 bytes per triangle.  n_triangles < 2^30, width and height at
 * most 2^24 and fewer than 2^31 tiles (UPF_ERR_BAD_GRID otherwise).  Overlapping calls
 * must use different workspaces. */
size_t upf_bake_rasterize_workspace_bytes(int64_t n_triangles, int32_t width, int32_t height);
int upf_bake_rasterize_ws(const float* d_uv, int64_t n_triangles, int32_t width, int32_t height,
                          int32_t* d_tri, void* d_ws, size_t ws_bytes, void* stream);
/* Per-texel material.  Mesh: positions/normals float[V][3], indices uint32[T][3]; d_uv
 * and d_tri from above.  Outputs: albedo float[h][w][3], bump and roughness float[h][w].
 * A texel whose triangle id is outside [0, T) or whose triangle names a vertex outside
 * [0, V) is written as uncovered (NaN). */
int upf_bake(const upf_bake_spec* spec, const float* d_positions, const float* d_normals,
             int64_t n_vertices, const uint32_t* d_indices, int64_t n_triangles, const float* d_uv,
             const int32_t* d_tri, int32_t width, int32_t height, float* d_albedo, float* d_bump,
             float* d_roughness, void* stream);

/* ---- the full material stack ------------------------------------------------------
 * terrainTile with the reference's coloring modes coloringDefault / coloringBarren /
 * coloringDebug (sources/coloringModes.cpp:905-966, selected at
 * sources/modesConfiguration.cpp:196-224): elevation, precipitation, temperature, slope
 * from 8 elevation probes at radius 0.5 in the tangent plane, biome, type, and the
 * layers bedrock, cliffs, mica, dirt, sand, grass, boulders, water, flowers, ice, snow.
 *
 * The reference's noise objects (about 40, each with its configuration in
 * coloringModes.cpp) are evaluated with those configurations; object k is seeded
 * seed + 16 k (the table is in upf_material.hpp) where the reference draws noiseSeed().
 * cage's noise kinds are not in the reference's checkout; this library defines them as
 *   Perlin          lattice gradient noise (UPF_NOISE_GRADIENT)
 *   Value           lattice value noise (UPF_NOISE_VALUE)
 *   Cubic           value noise interpolated by the cubic through the 4 lattice values
 *                   around the sample per axis, t^3 p + t^2 ((a - b) - p) + t (c - a) + b
 *                   with p = (d - c) - (a - b), scaled by 1 / 1.5^3
 *   Simplex, SimplexReduced   simplex noise (UPF_NOISE_SIMPLEX)
 *   Cellular        F1 / F2 over the 27 cells around the sample (one point per cell at
 *                   k - 0.5 + r, r in [0, 1]^3), distance Euclidean (squared, rooted for
 *                   the distance operations), Manhattan, or Hybrid (Manhattan +
 *                   squared Euclidean); operation Cell = hash of the nearest cell as
 *                   int32 * 2^-31, Subtract = F2 - F1 - 1, Multiply = F2 F1 / 2 - 1,
 *                   Divide = F1 / F2 - 1
 * with fractal amplitudes bounded by 1 / sum |gain|^k, lacunarity 2 and gain 0.5 where
 * the reference leaves them unset.  The colour helpers: HSV is the hexcone model;
 * colorHueShift is the reference's (sources/math.cpp:47-52); interpolateColor runs in
 * HSV with the hue along the shorter arc; colorDeviation (sources/math.cpp:40-45, a
 * global RNG per texel in the reference) moves the colour in HSV by a vector hashed
 * from the position's float bits.  anyPerpendicular and the Voronoi are the reference's
 * (sources/math.cpp:65-72, sources/voronoi.cpp with the plane projection of :58-66).
 * The beach mask's 1-D noise of the elevation samples the 3-D noise at (e, 0, 0).
 *
 * The spec's field gives terrainSdfElevation = shape / elevation_scale (LAND purpose)
 * and terrainSdfElevationRaw = the elevation stack (other purposes); its compose is
 * ignored. */
enum upf_coloring_id {
  UPF_COLORING_DEFAULT = 0,
  UPF_COLORING_BARREN = 1,
  UPF_COLORING_DEBUG = 2,
  UPF_COLORING_COUNT = 3
};
enum upf_purpose_id {  /* MeshPurposeEnum, sources/planets.h:50-56 */
  UPF_PURPOSE_LAND = 1,
  UPF_PURPOSE_WATER = 2,
  UPF_PURPOSE_NAVIGATION = 3
};
typedef struct upf_material_spec {
  upf_field_spec field;
  int32_t coloring; /* upf_coloring_id */
  int32_t purpose;  /* upf_purpose_id  */
  int32_t poles;    /* config unnatural-planets/poles/enable   (sources/coloringModes.cpp:18) */
  int32_t flowers;  /* config unnatural-planets/flowers/enable (:19): mica, boulders, flowers */
  uint32_t seed;
  int32_t reserved[3];
} upf_material_spec;

/* Tile (sources/planets.h:63-83): the fields the material stack produces.  slope in
 * radians; biome and type in TerrainBiomeEnum / TerrainTypeEnum order (Bare = 0 ..
 * Water = 11; Road = 0, Flat, Rough, Cliffs, Water). */
typedef struct upf_tile {
  float albedo[3];
  float roughness, metallic, height, opacity;
  float elevation, slope, temperature, precipitation;
  int32_t biome, type;
  int32_t reserved[3];
} upf_tile;

/* terrainTile at n points with unit normals (the navmesh loop of
 * sources/tileProperties.cpp:258-273, purpose NAVIGATION there).  d_positions,
 * d_normals: float[n][3]; d_tiles: upf_tile[n]. */
int upf_material_tiles(const upf_material_spec* spec, const float* d_positions, const float* d_normals,
                       int64_t n, upf_tile* d_tiles, void* stream);
/* The reference's texel bake with the full material (Generator::pixel,
 * sources/textures.cpp:24-43): the tile at the texel's barycentric position with the
 * normalized barycentric normal.  d_albedo: float[h][w][3] for LAND, [h][w][4] (albedo,
 * opacity) for WATER; d_special: float[h][w][2] (roughness, metallic); d_height:
 * float[h][w].  Uncovered texels are NaN, as upf_bake. */
int upf_bake_material(const upf_material_spec* spec, const float* d_positions, const float* d_normals,
                      int64_t n_vertices, const uint32_t* d_indices, int64_t n_triangles, const float* d_uv,
                      const int32_t* d_tri, int32_t width, int32_t height, float* d_albedo, float* d_special,
                      float* d_height, void* stream);

/* ---- texture post-processing -----------------------------------------------------
 * The steps after the bake in Generator::generate (sources/textures.cpp:73-81):
 * imageDilation(image, 7, true) and imageConvert(image, U8).  cage's kernels are not in
 * the reference checkout; the rules here are documented choices.
 *
 * Dilation: `rounds` passes; in each pass a texel whose first channel is NaN takes, per
 * channel, the mean of its 8 neighbours whose first channel is not NaN (summed in the
 * order dy = -1..1, dx = -1..1, then divided by their count), or stays NaN when it has
 * none; other texels keep their value.  d_img: float[h][w][channels], in place;
 * d_scratch: the same size. */
int upf_image_dilate(float* d_img, float* d_scratch, int32_t width, int32_t height, int32_t channels,
                     int32_t rounds, void* stream);
/* U8 conversion: out = NaN ? 0 : (uint8)(saturate(v) * 255 + 0.5), n values. */
int upf_image_to_u8(const float* d_img, int64_t n_values, uint8_t* d_out, void* stream);

/* The two conversions a land chunk's U8 maps get before export
 * (LandProcessor::chunkEntry, sources/generator.cpp:136-139).  cage's
 * imageConvertSpecialToGltfPbr and imageConvertHeigthToNormal are not in the reference
 * checkout; the rules here are documented choices, all in FP32.
 *
 * Height to normal, imageConvertHeigthToNormal(heightMap, strength): h = byte / 255 read
 * with clamp-to-edge, Sobel derivatives along x (columns) and y (rows),
 *   gx = (h[y-1][x+1] + 2 h[y][x+1] + h[y+1][x+1]) - (h[y-1][x-1] + 2 h[y][x-1] + h[y+1][x-1])
 *   gy = (h[y+1][x-1] + 2 h[y+1][x] + h[y+1][x+1]) - (h[y-1][x-1] + 2 h[y-1][x] + h[y-1][x+1])
 * summed left to right as written; v = (-strength * gx, -strength * gy, 1),
 * n = v * (1 / sqrt(v.v)) and out = (uint8)((n * 0.5 + 0.5) * 255 + 0.5) per component.
 * d_height: uint8[height][width]; d_out: uint8[height][width][3]. */
int upf_image_height_to_normal(const uint8_t* d_height, int32_t width, int32_t height, float strength,
                               uint8_t* d_out, void* stream);
/* Special to glTF PBR, imageConvertSpecialToGltfPbr(special): the 2-channel
 * (roughness, metallic) image becomes glTF's metallicRoughness layout, roughness in G
 * and metallic in B; R (ambient occlusion in a packed occlusion-roughness-metallic map)
 * is 255, no occlusion.  d_special: uint8[n][2]; d_out: uint8[n][3]. */
int upf_image_special_to_gltf_pbr(const uint8_t* d_special, int64_t n_texels, uint8_t* d_out, void* stream);

/* ---- batched chunks ---------------------------------------------------------------
 * Many chunks (planets) of uneven grid sizes, shapes and seeds evaluated and meshed by
 * one set of launches: the field rows, marching-cubes groups and active groups of all
 * chunks are each spread over one launch, and one scan numbers the
 * vertices and triangles of all chunks (chunk after chunk).  upf_batch_run makes no
 * allocation, no memset and no host synchronisation, so it can be captured into a
 * HIP graph (upf_batch_capture / upf_batch_replay do that).  No counterpart in the
 * reference, whose chunks are meshed one task each (sources/generator.cpp:159). */
typedef struct upf_chunk {
  upf_field_spec spec;
  upf_grid grid;
  float iso;
  int32_t reserved[3];
} upf_chunk;

/* Launch knobs searched by upf_batch_autotune; every setting gives identical results. */
typedef struct upf_batch_opts {
  int32_t field_variant;  /* upf_eval_variant                                  */
  int32_t blocks_per_cu;  /* accepted values 0 .. 32; no effect (the field pass
                             launches one workgroup per 4 field items)      */
  int32_t classify_z;     /* accepted values 0, 16, 32, 64; no effect (the
                             classification reads the sign words the field
                             kernel stores, one lane per group)           */
  int32_t reserved;
} upf_batch_opts;

/* Per-chunk result, written on the device by upf_batch_run: the chunk's vertices are
 * positions[vertex_offset .. vertex_offset + n_vertices), its triangles
 * indices[triangle_offset ..), with vertex indices local to the chunk. */
typedef struct upf_chunk_mesh {
  int64_t vertex_offset;
  int64_t n_vertices;
  int64_t triangle_offset;
  int64_t n_triangles;
} upf_chunk_mesh;

typedef struct upf_batch upf_batch; /* opaque plan: owns small device descriptors */

int upf_batch_create(const upf_chunk* chunks, int32_t n_chunks, const upf_batch_opts* opts,
                     upf_batch** out);
void upf_batch_destroy(upf_batch* b);
/* Field buffer size (floats, all chunks' grids back to back) and workspace size. */
int64_t upf_batch_field_floats(const upf_batch* b);
int64_t upf_batch_field_offset(const upf_batch* b, int32_t chunk);
size_t upf_batch_workspace_bytes(const upf_batch* b);
/* Changing the options drops a graph captured with other ones (upf_batch_replay then
 * returns UPF_ERR_NULL until upf_batch_capture runs again); upf_batch_autotune changes
 * them, so capture after tuning. */
int upf_batch_set_opts(upf_batch* b, const upf_batch_opts* opts);
int upf_batch_get_opts(const upf_batch* b, upf_batch_opts* opts);
/* Evaluate + mesh every chunk.  Writes past cap_vertices / cap_triangles are dropped;
 * the true totals are readable with upf_batch_totals. */
int upf_batch_run(upf_batch* b, float* d_field, void* d_workspace, size_t workspace_bytes,
                  float* d_positions, float* d_normals, uint32_t* d_indices, int64_t cap_vertices,
                  int64_t cap_triangles, upf_chunk_mesh* d_meshes, void* stream);
/* Reads the batch totals from the workspace (synchronises the stream). */
int upf_batch_totals(const upf_batch* b, const void* d_workspace, int64_t* n_vertices,
                     int64_t* n_triangles, void* stream);
/* Capture upf_batch_run with these buffers into a HIP graph owned by the plan, and
 * replay it. */
int upf_batch_capture(upf_batch* b, float* d_field, void* d_workspace, size_t workspace_bytes,
                      float* d_positions, float* d_normals, uint32_t* d_indices, int64_t cap_vertices,
                      int64_t cap_triangles, upf_chunk_mesh* d_meshes, void* stream);
int upf_batch_replay(upf_batch* b, void* stream);
/* Time every knob setting on these buffers (HIP events), keep the fastest in the plan
 * and return it in *best (nullable). */
int upf_batch_autotune(upf_batch* b, float* d_field, void* d_workspace, size_t workspace_bytes,
                       float* d_positions, float* d_normals, uint32_t* d_indices, int64_t cap_vertices,
                       int64_t cap_triangles, upf_chunk_mesh* d_meshes, void* stream,
                       upf_batch_opts* best);

#ifdef __cplusplus
}
#endif

#endif /* UPF_H_INCLUDED */
