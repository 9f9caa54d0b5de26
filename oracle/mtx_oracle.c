/*
 * CPU oracle for the isosurface-extraction path.  TEST INFRASTRUCTURE ONLY:
 * it is the checker for the HIP kernels, loaded by tests/, by the smoke check
 * and by the benchmark's CPU-baseline leg, never by the product library.
 *
 * It restates the reference's semantics in plain scalar C (FP64 where the
 * reference computes geometry):
 *   orc_minmax_legacy   src/testMinMax.h:79-130 (computeMinMaxFromFile loop:
 *                       window [B*b, B*b+B-1], clipped, {0,0} if empty)
 *   orc_minmax          src/shaders/minMaxLeaf.comp.glsl:39-79 (window
 *                       [B*b, min(B*b+B, dim-1)], the live GPU definition)
 *   orc_pyramid         src/minMaxManager.cpp:85-92,145-179 and
 *                       src/shaders/minMaxOctreeReduce.comp.glsl:33-57
 *   orc_active          src/testMinMax.h:220-242 and
 *                       src/blockFilteringTestUtils.h:438-478 (CPU rule) and
 *                       src/shaders/occupiedBlockPrefixSum.comp.glsl:84
 *                       (live GPU rule)
 *   orc_filter          src/blockFilteringTestUtils.h:438-478 (ascending IDs)
 *   orc_extract         marching cubes with PMB edge ownership:
 *                       case index src/shaders/marching_cubes_pmb.task.glsl:81-96,121,
 *                       triangles src/mc_tables.h:17-273,
 *                       edge owners src/shaders/marching_cubes_pmb.mesh.glsl:39-42,
 *                       vertex  src/shaders/marching_cubes_pmb.mesh.glsl:122-141,
 *                       normal  src/shaders/marching_cubes_pmb.mesh.glsl:111-120
 *                       with the zero-gradient guard of
 *                       src/shaders/marching_cubes.mesh.glsl:147-149,181-183
 *   orc_meshlets        meshlet partition of src/shaders/marching_cubes_pmb.task.glsl:182-215
 *                       over the renderer layout (self-contained vertex ranges,
 *                       src/shaders/render.task.glsl:61-63,125-128)
 *   orc_adjoint         d(vertex)/d(voxel), d(vertex)/d(iso) of the edge
 *                       interpolation above, topology frozen
 *
 * Pinning (which check fixes which function) is listed in tests/test_oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "../meshtrex_amd/csrc/mc_cases.h"
#include "../meshtrex_amd/csrc/synth_recipes.h"

#define ORC_U8 0
#define ORC_F32 1
#define ORC_RULE_CPU 0
#define ORC_RULE_GPU 1

int orc_version(void) { return 1; }

/* ---- order-preserving f32 <-> u32 key -------------------------------------- */

uint32_t orc_key_f32(float f)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

float orc_unkey_f32(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static inline uint32_t voxel_key(const void* vol, int dtype, uint64_t i)
{
    if (dtype == ORC_U8) return ((const uint8_t*)vol)[i];
    return orc_key_f32(((const float*)vol)[i]);
}

static inline float voxel_value(const void* vol, int dtype, uint64_t i)
{
    if (dtype == ORC_U8) return (float)((const uint8_t*)vol)[i];
    return ((const float*)vol)[i];
}

static inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

/* ---- generators ------------------------------------------------------------ */

static void blob_field_minmax(const mtx_blob* blobs, int n, const uint32_t d[3], double* fmin,
                              double* fmax, double* field)
{
    double lo = 1e300, hi = -1e300;
    uint64_t i = 0;
    for (uint32_t z = 0; z < d[2]; ++z)
        for (uint32_t y = 0; y < d[1]; ++y)
            for (uint32_t x = 0; x < d[0]; ++x, ++i) {
                double f = mtx_blob_field(blobs, n, x, y, z);
                field[i] = f;
                lo = f < lo ? f : lo;
                hi = f > hi ? f : hi;
            }
    *fmin = lo;
    *fmax = hi;
}

int orc_gen_blobs_u8(uint8_t* out, const uint32_t dims[3], uint64_t seed, int nblobs)
{
    uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    mtx_blob* b = (mtx_blob*)malloc(sizeof(mtx_blob) * (size_t)nblobs);
    double* field = (double*)malloc(sizeof(double) * n);
    if (!b || !field) { free(b); free(field); return -1; }
    double lo, hi;
    mtx_recipe_blobs(seed, dims, nblobs, b);
    blob_field_minmax(b, nblobs, dims, &lo, &hi, field);
    for (uint64_t i = 0; i < n; ++i) out[i] = mtx_blob_quantize(field[i], lo, hi, seed, i);
    free(b);
    free(field);
    return 0;
}

int orc_gen_blobs_f32(float* out, const uint32_t dims[3], uint64_t seed, int nblobs)
{
    uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    mtx_blob* b = (mtx_blob*)malloc(sizeof(mtx_blob) * (size_t)nblobs);
    double* field = (double*)malloc(sizeof(double) * n);
    if (!b || !field) { free(b); free(field); return -1; }
    double lo, hi;
    mtx_recipe_blobs(seed, dims, nblobs, b);
    blob_field_minmax(b, nblobs, dims, &lo, &hi, field);
    for (uint64_t i = 0; i < n; ++i) out[i] = mtx_blob_rescale(field[i], lo, hi);
    free(b);
    free(field);
    return 0;
}

int orc_gen_vessels_u8(uint8_t* out, const uint32_t dims[3], uint64_t seed, int npoly, int nseg)
{
    mtx_segment* s = (mtx_segment*)malloc(sizeof(mtx_segment) * (size_t)npoly * nseg);
    if (!s) return -1;
    mtx_recipe_vessels(seed, dims, npoly, nseg, s);
    uint64_t i = 0;
    for (uint32_t z = 0; z < dims[2]; ++z)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t x = 0; x < dims[0]; ++x, ++i)
                out[i] = mtx_vessel_value(s, npoly * nseg, seed, i, x, y, z);
    free(s);
    return 0;
}

int orc_gen_sdf_f32(float* out, const uint32_t dims[3], uint64_t seed, int nsph, int ntor,
                    double scale, double sigma)
{
    mtx_prim* p = (mtx_prim*)malloc(sizeof(mtx_prim) * (size_t)(nsph + ntor));
    if (!p) return -1;
    mtx_recipe_sdf(seed, dims, nsph, ntor, scale, p);
    uint64_t i = 0;
    for (uint32_t z = 0; z < dims[2]; ++z)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t x = 0; x < dims[0]; ++x, ++i)
                out[i] = mtx_sdf_value(p, nsph + ntor, sigma, seed, i, x, y, z);
    free(p);
    return 0;
}

/* ---- block min/max ----------------------------------------------------------- */

/* Reference CPU oracle semantics (src/testMinMax.h:79-130) generalised to any
 * dims/B: window [B*b, B*b+B-1] clipped to the volume, {0,0} for a block with
 * no voxel.  Pairs are stored as keys (u8: value, f32: order-preserving). */
void orc_minmax_legacy(const void* vol, int dtype, const uint32_t dims[3], uint32_t B,
                       uint32_t* pairs)
{
    uint32_t G[3] = { ceil_div(dims[0], B), ceil_div(dims[1], B), ceil_div(dims[2], B) };
    for (uint32_t bz = 0; bz < G[2]; ++bz)
        for (uint32_t by = 0; by < G[1]; ++by)
            for (uint32_t bx = 0; bx < G[0]; ++bx) {
                uint32_t lo = 0xFFFFFFFFu, hi = 0;
                int any = 0;
                for (uint32_t lz = 0; lz < B; ++lz)
                    for (uint32_t ly = 0; ly < B; ++ly)
                        for (uint32_t lx = 0; lx < B; ++lx) {
                            uint32_t x = bx * B + lx, y = by * B + ly, z = bz * B + lz;
                            if (x < dims[0] && y < dims[1] && z < dims[2]) {
                                uint32_t k = voxel_key(vol, dtype,
                                                       x + (uint64_t)dims[0] * (y + (uint64_t)dims[1] * z));
                                lo = k < lo ? k : lo;
                                hi = k > hi ? k : hi;
                                any = 1;
                            }
                        }
                uint64_t id = bx + (uint64_t)G[0] * (by + (uint64_t)G[1] * bz);
                pairs[2 * id] = any ? lo : 0;
                pairs[2 * id + 1] = any ? hi : 0;
            }
}

/* Live GPU semantics (src/shaders/minMaxLeaf.comp.glsl:39-79): window
 * [B*b, B*b+B] per axis, clipped to dim-1 (one voxel of overlap with the next
 * block so every cell of the block sees all 8 corners). */
void orc_minmax(const void* vol, int dtype, const uint32_t dims[3], uint32_t B, uint32_t* pairs)
{
    uint32_t G[3] = { ceil_div(dims[0], B), ceil_div(dims[1], B), ceil_div(dims[2], B) };
    for (uint32_t bz = 0; bz < G[2]; ++bz)
        for (uint32_t by = 0; by < G[1]; ++by)
            for (uint32_t bx = 0; bx < G[0]; ++bx) {
                uint32_t lo = 0xFFFFFFFFu, hi = 0;
                uint32_t x1 = bx * B + B, y1 = by * B + B, z1 = bz * B + B;
                if (x1 > dims[0] - 1) x1 = dims[0] - 1;
                if (y1 > dims[1] - 1) y1 = dims[1] - 1;
                if (z1 > dims[2] - 1) z1 = dims[2] - 1;
                for (uint32_t z = bz * B; z <= z1; ++z)
                    for (uint32_t y = by * B; y <= y1; ++y)
                        for (uint32_t x = bx * B; x <= x1; ++x) {
                            uint32_t k = voxel_key(vol, dtype,
                                                   x + (uint64_t)dims[0] * (y + (uint64_t)dims[1] * z));
                            lo = k < lo ? k : lo;
                            hi = k > hi ? k : hi;
                        }
                uint64_t id = bx + (uint64_t)G[0] * (by + (uint64_t)G[1] * bz);
                pairs[2 * id] = lo;
                pairs[2 * id + 1] = hi;
            }
}

/* Pyramid geometry (src/minMaxManager.cpp:85-92,147-151): level count
 * 1+floor(log2(max G)), extent_{l+1} = ceil(extent_l/2).  Returns the level
 * count; fills extents and pair offsets (level 0 at offset 0). */
uint32_t orc_pyramid_layout(const uint32_t G[3], uint32_t ext[][3], uint64_t* off)
{
    uint32_t m = G[0] > G[1] ? G[0] : G[1];
    m = m > G[2] ? m : G[2];
    uint32_t L = 1;
    while ((m >> L) > 0) ++L;
    uint64_t o = 0;
    uint32_t e[3] = { G[0], G[1], G[2] };
    for (uint32_t l = 0; l < L; ++l) {
        ext[l][0] = e[0]; ext[l][1] = e[1]; ext[l][2] = e[2];
        off[l] = o;
        o += (uint64_t)e[0] * e[1] * e[2];
        for (int a = 0; a < 3; ++a) e[a] = (e[a] + 1) >> 1;
    }
    off[L] = o;
    return L;
}

/* Levels 1.. from level 0 (src/shaders/minMaxOctreeReduce.comp.glsl:33-57):
 * a texel is the min/max over its existing 2x2x2 children. */
void orc_pyramid(uint32_t* pairs, const uint32_t G[3])
{
    uint32_t ext[32][3];
    uint64_t off[33];
    uint32_t L = orc_pyramid_layout(G, ext, off);
    for (uint32_t l = 1; l < L; ++l) {
        const uint32_t* s = ext[l - 1];
        const uint32_t* d = ext[l];
        for (uint32_t z = 0; z < d[2]; ++z)
            for (uint32_t y = 0; y < d[1]; ++y)
                for (uint32_t x = 0; x < d[0]; ++x) {
                    uint32_t lo = 0xFFFFFFFFu, hi = 0;
                    for (uint32_t c = 0; c < 8; ++c) {
                        uint32_t sx = 2 * x + (c & 1), sy = 2 * y + ((c >> 1) & 1), sz = 2 * z + (c >> 2);
                        if (sx >= s[0] || sy >= s[1] || sz >= s[2]) continue;
                        uint64_t si = off[l - 1] + sx + (uint64_t)s[0] * (sy + (uint64_t)s[1] * sz);
                        lo = pairs[2 * si] < lo ? pairs[2 * si] : lo;
                        hi = pairs[2 * si + 1] > hi ? pairs[2 * si + 1] : hi;
                    }
                    uint64_t di = off[l] + x + (uint64_t)d[0] * (y + (uint64_t)d[1] * z);
                    pairs[2 * di] = lo;
                    pairs[2 * di + 1] = hi;
                }
    }
}

/* ---- activity and compaction -------------------------------------------- */

/* CPU rule (src/testMinMax.h:227-232, src/blockFilteringTestUtils.h:456-468):
 * min != max and round(iso) in [min,max] for u8 (integer compare); for f32
 * there is no rounding (fp32 compare on the decoded values).
 * GPU rule (src/shaders/occupiedBlockPrefixSum.comp.glsl:84): iso >= min and
 * iso <= max in fp32, no min != max exclusion. */
int orc_active(uint32_t lo, uint32_t hi, int dtype, float iso, int rule)
{
    float flo = dtype == ORC_U8 ? (float)lo : orc_unkey_f32(lo);
    float fhi = dtype == ORC_U8 ? (float)hi : orc_unkey_f32(hi);
    if (rule == ORC_RULE_GPU) return iso >= flo && iso <= fhi;
    if (lo == hi) return 0;
    if (dtype == ORC_U8) {
        float r = roundf(iso);
        if (!(r >= 0.0f && r < 4294967296.0f)) return 0;
        uint32_t ir = (uint32_t)r;
        return ir >= lo && ir <= hi;
    }
    return iso >= flo && iso <= fhi;
}

/* Ascending list of active block IDs (src/blockFilteringTestUtils.h:460-474). */
uint64_t orc_filter(const uint32_t* pairs, uint64_t nb, int dtype, float iso, int rule, uint32_t* ids)
{
    uint64_t n = 0;
    for (uint64_t b = 0; b < nb; ++b)
        if (orc_active(pairs[2 * b], pairs[2 * b + 1], dtype, iso, rule)) {
            if (ids) ids[n] = (uint32_t)b;
            ++n;
        }
    return n;
}

/* ---- marching cubes with PMB ownership ------------------------------------ */

typedef struct {
    const void* vol;
    int dtype;
    uint32_t d[3];
    float iso;
} vol_t;

static inline uint64_t lin(const vol_t* v, uint32_t x, uint32_t y, uint32_t z)
{
    return x + (uint64_t)v->d[0] * (y + (uint64_t)v->d[1] * z);
}

static inline float val(const vol_t* v, uint32_t x, uint32_t y, uint32_t z)
{
    return voxel_value(v->vol, v->dtype, lin(v, x, y, z));
}

/* t along an edge, src/shaders/marching_cubes_pmb.mesh.glsl:129-134:
 * mu = 0.5 unless |v1 - v0| > 1e-5, the difference and the bound taken in
 * fp32 as the shader takes them; the quotient in FP64; clamped to [0, 1]. */
static inline double edge_t(float v0, float v1, float iso)
{
    volatile float df = v1 - v0;  /* rounded to fp32, like the shader's subtraction */
    double t = 0.5;
    if (fabsf(df) > 1e-5f) t = ((double)iso - (double)v0) / ((double)v1 - (double)v0);
    return t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
}

/* Whether t depends on the edge's values: the guard above passes. */
static inline int edge_t_live(float v0, float v1)
{
    volatile float df = v1 - v0;
    return fabsf(df) > 1e-5f;
}

/* Inside/outside bit of src/shaders/marching_cubes_pmb.task.glsl:91 (value <= iso). */
static inline int below(const vol_t* v, uint32_t x, uint32_t y, uint32_t z)
{
    return val(v, x, y, z) <= v->iso;
}

static inline int edge_active(const vol_t* v, uint32_t x, uint32_t y, uint32_t z, int axis)
{
    uint32_t q[3] = { x, y, z };
    q[axis] += 1;
    if (q[axis] >= v->d[axis]) return 0;
    return below(v, x, y, z) != below(v, q[0], q[1], q[2]);
}

static inline unsigned cell_case(const vol_t* v, uint32_t x, uint32_t y, uint32_t z)
{
    unsigned c = 0;
    for (unsigned i = 0; i < 8; ++i)
        c |= (unsigned)below(v, x + mtx_corner_x(i), y + mtx_corner_y(i), z + mtx_corner_z(i)) << i;
    return c;
}

/* Case index of every cell of every listed block, B^3 bytes per block in
 * local x-fastest order; cells outside the volume's valid range (any
 * coordinate >= dim-1) read 0. */
void orc_case_indices(const void* vol, int dtype, const uint32_t dims[3], uint32_t B, float iso,
                      const uint32_t* ids, uint64_t na, uint8_t* cases)
{
    vol_t v = { vol, dtype, { dims[0], dims[1], dims[2] }, iso };
    uint32_t G0 = ceil_div(dims[0], B), G1 = ceil_div(dims[1], B);
    for (uint64_t i = 0; i < na; ++i) {
        uint32_t id = ids[i];
        uint32_t bx = id % G0, by = (id / G0) % G1, bz = id / (G0 * G1);
        for (uint32_t lz = 0; lz < B; ++lz)
            for (uint32_t ly = 0; ly < B; ++ly)
                for (uint32_t lx = 0; lx < B; ++lx) {
                    uint32_t x = bx * B + lx, y = by * B + ly, z = bz * B + lz;
                    uint8_t c = 0;
                    if (x + 1 < dims[0] && y + 1 < dims[1] && z + 1 < dims[2])
                        c = (uint8_t)cell_case(&v, x, y, z);
                    cases[i * (uint64_t)B * B * B + lx + B * (ly + B * lz)] = c;
                }
    }
}

/* Open-addressing map edge ID -> vertex index. */
typedef struct { uint64_t* keys; uint32_t* vals; uint64_t mask; } emap_t;

static int emap_init(emap_t* m, uint64_t n)
{
    uint64_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    m->keys = (uint64_t*)malloc(sizeof(uint64_t) * cap);
    m->vals = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    m->mask = cap - 1;
    if (!m->keys || !m->vals) return -1;
    memset(m->keys, 0xFF, sizeof(uint64_t) * cap);
    return 0;
}

static void emap_put(emap_t* m, uint64_t k, uint32_t v)
{
    uint64_t h = mtx_mix64(k) & m->mask;
    while (m->keys[h] != ~0ull) h = (h + 1) & m->mask;
    m->keys[h] = k;
    m->vals[h] = v;
}

static int64_t emap_get(const emap_t* m, uint64_t k)
{
    uint64_t h = mtx_mix64(k) & m->mask;
    while (m->keys[h] != ~0ull) {
        if (m->keys[h] == k) return m->vals[h];
        h = (h + 1) & m->mask;
    }
    return -1;
}

/* Endpoint normal (src/shaders/marching_cubes_pmb.mesh.glsl:111-120):
 * normalize(s(p-x)-s(p+x), s(p-y)-s(p+y), s(p-z)-s(p+z)), samples clamped to
 * the volume; a zero gradient (|g| < 1e-5) gives (0,1,0) as in
 * src/shaders/marching_cubes.mesh.glsl:147-149. */
static void corner_normal(const vol_t* v, const int64_t p[3], double n[3])
{
    for (int a = 0; a < 3; ++a) {
        int64_t lo[3] = { p[0], p[1], p[2] }, hi[3] = { p[0], p[1], p[2] };
        lo[a] -= 1;
        hi[a] += 1;
        for (int b = 0; b < 3; ++b) {
            int64_t m = (int64_t)v->d[b] - 1;
            lo[b] = lo[b] < 0 ? 0 : (lo[b] > m ? m : lo[b]);
            hi[b] = hi[b] < 0 ? 0 : (hi[b] > m ? m : hi[b]);
        }
        n[a] = (double)val(v, (uint32_t)lo[0], (uint32_t)lo[1], (uint32_t)lo[2]) -
               (double)val(v, (uint32_t)hi[0], (uint32_t)hi[1], (uint32_t)hi[2]);
    }
    double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l < 1e-5) { n[0] = 0.0; n[1] = 1.0; n[2] = 0.0; return; }
    n[0] /= l; n[1] /= l; n[2] /= l;
}

/* Visit the vertices in output order: blocks by ascending ID, owner corners
 * in block-local x-fastest order, then axes x, y, z.  Calls fn for each. */
typedef void (*vert_fn)(void* ctx, uint64_t idx, uint32_t x, uint32_t y, uint32_t z, int axis);

static uint64_t for_each_vertex_rows(const vol_t* v, uint32_t B, uint32_t bz0, uint32_t bz1, vert_fn fn,
                                     void* ctx)
{
    uint32_t G[3] = { ceil_div(v->d[0], B), ceil_div(v->d[1], B), ceil_div(v->d[2], B) };
    uint64_t n = 0;
    if (v->d[0] < 2 || v->d[1] < 2 || v->d[2] < 2) return 0;
    if (bz1 > G[2]) bz1 = G[2];
    for (uint32_t bz = bz0; bz < bz1; ++bz)
        for (uint32_t by = 0; by < G[1]; ++by)
            for (uint32_t bx = 0; bx < G[0]; ++bx)
                for (uint32_t lz = 0; lz < B; ++lz)
                    for (uint32_t ly = 0; ly < B; ++ly)
                        for (uint32_t lx = 0; lx < B; ++lx) {
                            uint32_t x = bx * B + lx, y = by * B + ly, z = bz * B + lz;
                            if (x >= v->d[0] || y >= v->d[1] || z >= v->d[2]) continue;
                            for (int a = 0; a < 3; ++a)
                                if (edge_active(v, x, y, z, a)) {
                                    if (fn) fn(ctx, n, x, y, z, a);
                                    ++n;
                                }
                        }
    return n;
}

/* Visit valid cells in output order (blocks by ascending ID, cells in
 * block-local x-fastest order). */
typedef void (*cell_fn)(void* ctx, uint32_t x, uint32_t y, uint32_t z, unsigned c);

static uint64_t for_each_cell_rows(const vol_t* v, uint32_t B, uint32_t bz0, uint32_t bz1, cell_fn fn,
                                   void* ctx)
{
    uint32_t G[3] = { ceil_div(v->d[0], B), ceil_div(v->d[1], B), ceil_div(v->d[2], B) };
    uint64_t ntri = 0;
    if (v->d[0] < 2 || v->d[1] < 2 || v->d[2] < 2) return 0;
    if (bz1 > G[2]) bz1 = G[2];
    for (uint32_t bz = bz0; bz < bz1; ++bz)
        for (uint32_t by = 0; by < G[1]; ++by)
            for (uint32_t bx = 0; bx < G[0]; ++bx)
                for (uint32_t lz = 0; lz < B; ++lz)
                    for (uint32_t ly = 0; ly < B; ++ly)
                        for (uint32_t lx = 0; lx < B; ++lx) {
                            uint32_t x = bx * B + lx, y = by * B + ly, z = bz * B + lz;
                            if (x + 1 >= v->d[0] || y + 1 >= v->d[1] || z + 1 >= v->d[2]) continue;
                            unsigned c = cell_case(v, x, y, z);
                            unsigned nt = mtx_word_ntri(mtx_tri_words[c]);
                            if (nt && fn) fn(ctx, x, y, z, c);
                            ntri += nt;
                        }
    return ntri;
}

static uint64_t for_each_vertex(const vol_t* v, uint32_t B, vert_fn fn, void* ctx)
{
    return for_each_vertex_rows(v, B, 0, UINT32_MAX, fn, ctx);
}

static uint64_t for_each_cell(const vol_t* v, uint32_t B, cell_fn fn, void* ctx)
{
    return for_each_cell_rows(v, B, 0, UINT32_MAX, fn, ctx);
}

/* Vertex and triangle counts of the blocks in block rows bz0 <= bz < bz1.
 * Every vertex belongs to the block of its owning corner and every cell to
 * one block, so the mesh's counts are the sums over any split of the rows. */
void orc_extract_count_rows(const void* vol, int dtype, const uint32_t dims[3], uint32_t B, float iso,
                            uint32_t bz0, uint32_t bz1, uint64_t* nvert, uint64_t* ntri)
{
    vol_t v = { vol, dtype, { dims[0], dims[1], dims[2] }, iso };
    *nvert = for_each_vertex_rows(&v, B, bz0, bz1, NULL, NULL);
    *ntri = for_each_cell_rows(&v, B, bz0, bz1, NULL, NULL);
}

/* Vertex and triangle counts of the mesh. */
void orc_extract_count(const void* vol, int dtype, const uint32_t dims[3], uint32_t B, float iso,
                       uint64_t* nvert, uint64_t* ntri)
{
    vol_t v = { vol, dtype, { dims[0], dims[1], dims[2] }, iso };
    *nvert = for_each_vertex(&v, B, NULL, NULL);
    *ntri = for_each_cell(&v, B, NULL, NULL);
}

typedef struct {
    const vol_t* v;
    emap_t map;
    uint64_t* edge_ids;
    double* t;
    double* ndc;
    double* normal;
    uint32_t* tris;
    uint64_t tri_pos;
    int missing;
} ext_ctx;

static void emit_vertex(void* p, uint64_t idx, uint32_t x, uint32_t y, uint32_t z, int a)
{
    ext_ctx* c = (ext_ctx*)p;
    const vol_t* v = c->v;
    uint64_t eid = 3 * lin(v, x, y, z) + (uint64_t)a;
    emap_put(&c->map, eid, (uint32_t)idx);
    if (c->edge_ids) c->edge_ids[idx] = eid;
    int64_t p0[3] = { x, y, z }, p1[3] = { x, y, z };
    p1[a] += 1;
    double t = edge_t(val(v, x, y, z), val(v, (uint32_t)p1[0], (uint32_t)p1[1], (uint32_t)p1[2]), v->iso);
    if (c->t) c->t[idx] = t;
    if (c->ndc)
        for (int k = 0; k < 3; ++k) {
            double pk = (double)p0[k] + (k == a ? t : 0.0);
            c->ndc[3 * idx + k] = pk / (double)v->d[k] * 2.0 - 1.0;
        }
    if (c->normal) {
        double n0[3], n1[3], m[3];
        corner_normal(v, p0, n0);
        corner_normal(v, p1, n1);
        for (int k = 0; k < 3; ++k) m[k] = n0[k] * (1.0 - t) + n1[k] * t;
        double l = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        if (l < 1e-5) {
            m[0] = n0[0]; m[1] = n0[1]; m[2] = n0[2];
            l = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        }
        for (int k = 0; k < 3; ++k) c->normal[3 * idx + k] = m[k] / l;
    }
}

static void emit_cell(void* p, uint32_t x, uint32_t y, uint32_t z, unsigned cs)
{
    ext_ctx* c = (ext_ctx*)p;
    uint64_t w = mtx_tri_words[cs];
    unsigned nt = mtx_word_ntri(w);
    for (unsigned k = 0; k < nt; ++k)
        for (unsigned j = 0; j < 3; ++j) {
            unsigned o = mtx_edge_owner(mtx_word_edge(w, k, j));
            uint32_t ox = x + (o & 1u), oy = y + ((o >> 1) & 1u), oz = z + ((o >> 2) & 1u);
            uint64_t eid = 3 * lin(c->v, ox, oy, oz) + (o >> 3);
            int64_t vi = emap_get(&c->map, eid);
            if (vi < 0) c->missing = 1;
            if (c->tris) c->tris[c->tri_pos] = (uint32_t)vi;
            c->tri_pos++;
        }
}

/* Full extraction.  Output arrays are sized from orc_extract_count; any may
 * be NULL.  Returns 0, or -1 on allocation failure, -2 if a triangle names an
 * edge with no vertex (cannot happen for a consistent table). */
int orc_extract(const void* vol, int dtype, const uint32_t dims[3], uint32_t B, float iso,
                uint64_t* edge_ids, double* t, double* ndc, double* normal, uint32_t* tris)
{
    vol_t v = { vol, dtype, { dims[0], dims[1], dims[2] }, iso };
    uint64_t nv = for_each_vertex(&v, B, NULL, NULL);
    ext_ctx c;
    memset(&c, 0, sizeof(c));
    c.v = &v;
    if (emap_init(&c.map, nv)) return -1;
    c.edge_ids = edge_ids;
    c.t = t;
    c.ndc = ndc;
    c.normal = normal;
    c.tris = tris;
    for_each_vertex(&v, B, emit_vertex, &c);
    for_each_cell(&v, B, emit_cell, &c);
    free(c.map.keys);
    free(c.map.vals);
    return c.missing ? -2 : 0;
}

/* ---- meshlets ------------------------------------------------------------ */

/* Meshlets of the renderer layout.  Per block (ascending block ID), the
 * occupied cells (valid, at least one triangle) in cell order are cut greedily
 * as src/shaders/marching_cubes_pmb.task.glsl:182-215 cuts them: a cell
 * starts a new meshlet when it is not the meshlet's first and adding it would
 * take the meshlet past 64 vertices or 126 triangles.  A meshlet's vertex
 * count is the number of distinct vertices its triangles use (each meshlet
 * carries its own copy of each, so that every index of a meshlet falls in its
 * own vertex range: src/shaders/render.task.glsl:61-63,125-128), taken in
 * order of first use (cell order, then the cell's edges in ascending owner
 * index mtx_edge_uidx).  Outputs (any may be NULL; sizes from a first call with NULLs):
 * desc[4 m + 0..3] = vertexOffset, indexOffset, vertexCount, primitiveCount;
 * map[j] = index in the shared mesh (orc_extract's vertex order) of meshlet
 * vertex j; mind[3 t + k] = meshlet vertex of vertex k of triangle t (the
 * shared mesh's triangle order). */
#define ORC_MESHLET_VERTS 64u
#define ORC_MESHLET_PRIMS 126u

typedef struct {
    uint64_t edge[ORC_MESHLET_VERTS];
    uint32_t n, prims;
    uint64_t vbase, tbase;
} orc_meshlet_t;

static int mlet_find(const orc_meshlet_t* m, uint64_t eid)
{
    for (uint32_t i = 0; i < m->n; ++i)
        if (m->edge[i] == eid) return (int)i;
    return -1;
}

int orc_meshlets(const void* vol, int dtype, const uint32_t dims[3], uint32_t B, float iso, uint64_t* n_meshlets,
                 uint64_t* n_mverts, uint32_t* desc, uint32_t* map, uint32_t* mind)
{
    vol_t v = { vol, dtype, { dims[0], dims[1], dims[2] }, iso };
    *n_meshlets = 0;
    *n_mverts = 0;
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2) return 0;
    emap_t em;
    const uint64_t nv = for_each_vertex(&v, B, NULL, NULL);
    if (emap_init(&em, nv)) return -1;
    {
        ext_ctx c;
        memset(&c, 0, sizeof(c));
        c.v = &v;
        c.map = em;
        for_each_vertex(&v, B, emit_vertex, &c);
        em = c.map;
    }
    uint32_t G[3] = { ceil_div(dims[0], B), ceil_div(dims[1], B), ceil_div(dims[2], B) };
    uint64_t nm = 0, ns = 0, tri = 0;
    orc_meshlet_t m;
    memset(&m, 0, sizeof m);
    for (uint32_t bz = 0; bz < G[2]; ++bz)
        for (uint32_t by = 0; by < G[1]; ++by)
            for (uint32_t bx = 0; bx < G[0]; ++bx) {
                int open = 0;
                for (uint32_t lz = 0; lz < B; ++lz)
                    for (uint32_t ly = 0; ly < B; ++ly)
                        for (uint32_t lx = 0; lx < B; ++lx) {
                            uint32_t x = bx * B + lx, y = by * B + ly, z = bz * B + lz;
                            if (x + 1 >= dims[0] || y + 1 >= dims[1] || z + 1 >= dims[2]) continue;
                            unsigned cs = cell_case(&v, x, y, z);
                            uint64_t w = mtx_tri_words[cs];
                            unsigned nt = mtx_word_ntri(w);
                            if (!nt) continue;
                            /* the cell's edges, ascending owner index, as global edge IDs */
                            unsigned emask = mtx_case_edge_mask(cs);
                            uint64_t ce[12];
                            unsigned nce = 0, nnew = 0;
                            for (unsigned u = 0; u < 12u; ++u) {
                                unsigned e = mtx_uidx_edge(u);
                                if (!((emask >> e) & 1u)) continue;
                                unsigned o = mtx_edge_owner(e);
                                uint64_t eid = 3 * lin(&v, x + (o & 1u), y + ((o >> 1) & 1u), z + ((o >> 2) & 1u)) +
                                               (o >> 3);
                                ce[nce++] = eid;
                                if (!open || mlet_find(&m, eid) < 0) ++nnew;
                            }
                            if (open && (m.n + nnew > ORC_MESHLET_VERTS || m.prims + nt > ORC_MESHLET_PRIMS)) {
                                if (desc) {
                                    desc[4 * nm + 0] = (uint32_t)m.vbase;
                                    desc[4 * nm + 1] = (uint32_t)(3 * m.tbase);
                                    desc[4 * nm + 2] = m.n;
                                    desc[4 * nm + 3] = m.prims;
                                }
                                ++nm;
                                ns += m.n;
                                open = 0;
                            }
                            if (!open) {
                                m.n = 0;
                                m.prims = 0;
                                m.vbase = ns;
                                m.tbase = tri;
                                open = 1;
                            }
                            for (unsigned i = 0; i < nce; ++i)
                                if (mlet_find(&m, ce[i]) < 0) {
                                    if (map) map[m.vbase + m.n] = (uint32_t)emap_get(&em, ce[i]);
                                    m.edge[m.n++] = ce[i];
                                }
                            for (unsigned k = 0; k < nt; ++k, ++tri)
                                for (unsigned j = 0; j < 3; ++j) {
                                    unsigned o = mtx_edge_owner(mtx_word_edge(w, k, j));
                                    uint64_t eid = 3 * lin(&v, x + (o & 1u), y + ((o >> 1) & 1u), z + ((o >> 2) & 1u)) +
                                                   (o >> 3);
                                    if (mind) mind[3 * tri + j] = (uint32_t)(m.vbase + (uint64_t)mlet_find(&m, eid));
                                }
                            m.prims += nt;
                        }
                if (open) {
                    if (desc) {
                        desc[4 * nm + 0] = (uint32_t)m.vbase;
                        desc[4 * nm + 1] = (uint32_t)(3 * m.tbase);
                        desc[4 * nm + 2] = m.n;
                        desc[4 * nm + 3] = m.prims;
                    }
                    ++nm;
                    ns += m.n;
                }
            }
    free(em.keys);
    free(em.vals);
    *n_meshlets = nm;
    *n_mverts = ns;
    return 0;
}

/* ---- adjoint ------------------------------------------------------------- */

/* Gradient of L through the vertex positions p = owner + t*e_axis (voxel
 * units), t = (iso - v0)/(v1 - v0), topology frozen:
 *   dt/dv0 = (t-1)/(v1-v0), dt/dv1 = -t/(v1-v0), dt/diso = 1/(v1-v0).
 * dLdp holds 3 doubles per vertex; dLdvol (one double per voxel) is
 * overwritten. */
void orc_adjoint(const float* vol, const uint32_t dims[3], float iso, const uint64_t* edge_ids,
                 uint64_t nv, const double* dLdp, double* dLdvol, double* dLdiso)
{
    uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    memset(dLdvol, 0, sizeof(double) * n);
    double gi = 0.0;
    for (uint64_t i = 0; i < nv; ++i) {
        uint64_t vox = edge_ids[i] / 3;
        int a = (int)(edge_ids[i] % 3);
        uint64_t stride = a == 0 ? 1 : (a == 1 ? dims[0] : (uint64_t)dims[0] * dims[1]);
        if (!edge_t_live(vol[vox], vol[vox + stride])) continue;  /* t = 0.5, constant */
        double v0 = vol[vox], v1 = vol[vox + stride];
        double d = v1 - v0;
        double t = ((double)iso - v0) / d;
        double g = dLdp[3 * i + a];
        dLdvol[vox] += g * (t - 1.0) / d;
        dLdvol[vox + stride] += g * (-t) / d;
        gi += g / d;
    }
    *dLdiso = gi;
}

/* Positions in voxel units (FP64) for given edge IDs, recomputing t. */
void orc_positions(const float* vol, const uint32_t dims[3], float iso, const uint64_t* edge_ids,
                   uint64_t nv, double* pos)
{
    for (uint64_t i = 0; i < nv; ++i) {
        uint64_t vox = edge_ids[i] / 3;
        int a = (int)(edge_ids[i] % 3);
        uint64_t stride = a == 0 ? 1 : (a == 1 ? dims[0] : (uint64_t)dims[0] * dims[1]);
        double t = edge_t(vol[vox], vol[vox + stride], iso);
        uint64_t x = vox % dims[0], y = (vox / dims[0]) % dims[1], z = vox / ((uint64_t)dims[0] * dims[1]);
        pos[3 * i + 0] = (double)x + (a == 0 ? t : 0.0);
        pos[3 * i + 1] = (double)y + (a == 1 ? t : 0.0);
        pos[3 * i + 2] = (double)z + (a == 2 ? t : 0.0);
    }
}

/* ---- table views for the tests ------------------------------------------- */

/* The case table unpacked to the reference's int32[256][16] layout (-1 padded). */
void orc_tri_table(int32_t* out)
{
    for (unsigned c = 0; c < 256; ++c) {
        uint64_t w = mtx_tri_words[c];
        unsigned n = mtx_word_ntri(w);
        for (unsigned i = 0; i < 16; ++i)
            out[c * 16 + i] = i < 3 * n ? (int32_t)mtx_word_edge(w, i / 3, i % 3) : -1;
    }
}

void orc_edge_table(int32_t* out)
{
    for (unsigned c = 0; c < 256; ++c) out[c] = (int32_t)mtx_case_edge_mask(c);
}

void orc_edge_owner_table(uint32_t* out)
{
    for (unsigned e = 0; e < 12; ++e) {
        unsigned o = mtx_edge_owner(e);
        out[4 * e + 0] = o & 1u;
        out[4 * e + 1] = (o >> 1) & 1u;
        out[4 * e + 2] = (o >> 2) & 1u;
        out[4 * e + 3] = o >> 3;
    }
}
