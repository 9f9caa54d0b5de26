/*
 * Synthetic volume recipes for the benchmark and test workloads.
 *
 * Every random draw comes from a counter-based generator (a splitmix64 hash
 * of (seed, counter)), so a voxel's value depends only on the seed, the
 * recipe parameters and its own index: generation is order-free and runs the
 * same way on the GPU and on the host.  Geometry is evaluated in FP64.
 *
 * Workloads (seed = 0x6D747278 + k for workload k):
 *   blobs    sum of isotropic Gaussian blobs, centres uniform in the volume,
 *            sigma U[8,32], amplitude U[0.5,1]; u8 form is normalised to
 *            0..255 plus integer noise U{-2..2}, clamped; f32 form is the raw
 *            field rescaled to [-1,1].
 *   vessels  polylines of straight segments (start uniform in the volume,
 *            directions uniform on the sphere, lengths U[20,80], one radius
 *            U[2,6] per polyline); value 255*exp(-d^2/(2r^2)) over a
 *            background U{0..20}, rounded and clamped to u8.
 *   sdf      signed distance to the union of spheres (r U[20,120]) and tori
 *            (R U[40,150], r U[8,30], axis uniform on the sphere), centres
 *            uniform in the volume, radii scaled by `scale`, plus N(0,0.05^2).
 *
 * This header is plain C with inline functions; under hipcc the per-voxel
 * functions are also device functions.
 */
#ifndef MTX_SYNTH_RECIPES_H
#define MTX_SYNTH_RECIPES_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define MTX_SHD __host__ __device__
#else
#define MTX_SHD
#endif

#define MTX_SEED_BASE 0x6D747278ull
/* Counter offset of the per-voxel noise stream (parameters use small counters). */
#define MTX_NOISE_STREAM (1ull << 62)

static inline MTX_SHD uint64_t mtx_mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline MTX_SHD uint64_t mtx_rand64(uint64_t seed, uint64_t k)
{
    return mtx_mix64(mtx_mix64(seed) ^ (k * 0xD1B54A32D192ED03ull));
}

/* Uniform double in [0,1). */
static inline MTX_SHD double mtx_rand01(uint64_t seed, uint64_t k)
{
    return (double)(mtx_rand64(seed, k) >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct { double c[3]; double sigma; double amp; } mtx_blob;
typedef struct { double a[3]; double b[3]; double r; } mtx_segment;
typedef struct { int kind; double c[3]; double r; double R; double axis[3]; } mtx_prim;

/* Unit vector uniform on the sphere from two uniforms. */
static inline void mtx_unit_vector(double u, double v, double out[3])
{
    double z = 2.0 * u - 1.0;
    double s = sqrt(1.0 - z * z > 0.0 ? 1.0 - z * z : 0.0);
    double phi = 6.283185307179586 * v;
    out[0] = s * cos(phi);
    out[1] = s * sin(phi);
    out[2] = z;
}

/* ---- parameters (host only) -------------------------------------------- */

static inline void mtx_recipe_blobs(uint64_t seed, const uint32_t dims[3], int n, mtx_blob* out)
{
    uint64_t k = 0;
    for (int i = 0; i < n; ++i) {
        out[i].c[0] = mtx_rand01(seed, k++) * dims[0];
        out[i].c[1] = mtx_rand01(seed, k++) * dims[1];
        out[i].c[2] = mtx_rand01(seed, k++) * dims[2];
        out[i].sigma = 8.0 + 24.0 * mtx_rand01(seed, k++);
        out[i].amp = 0.5 + 0.5 * mtx_rand01(seed, k++);
    }
}

static inline void mtx_recipe_vessels(uint64_t seed, const uint32_t dims[3], int npoly, int nseg,
                                      mtx_segment* out)
{
    uint64_t k = 0;
    for (int p = 0; p < npoly; ++p) {
        double pos[3];
        pos[0] = mtx_rand01(seed, k++) * dims[0];
        pos[1] = mtx_rand01(seed, k++) * dims[1];
        pos[2] = mtx_rand01(seed, k++) * dims[2];
        double r = 2.0 + 4.0 * mtx_rand01(seed, k++);
        for (int s = 0; s < nseg; ++s) {
            double dir[3];
            double u = mtx_rand01(seed, k++);
            double v = mtx_rand01(seed, k++);
            mtx_unit_vector(u, v, dir);
            double len = 20.0 + 60.0 * mtx_rand01(seed, k++);
            mtx_segment* sg = &out[p * nseg + s];
            for (int d = 0; d < 3; ++d) {
                sg->a[d] = pos[d];
                sg->b[d] = pos[d] + len * dir[d];
                pos[d] = sg->b[d];
            }
            sg->r = r;
        }
    }
}

static inline void mtx_recipe_sdf(uint64_t seed, const uint32_t dims[3], int nsph, int ntor,
                                  double scale, mtx_prim* out)
{
    uint64_t k = 0;
    for (int i = 0; i < nsph + ntor; ++i) {
        mtx_prim* p = &out[i];
        p->kind = i < nsph ? 0 : 1;
        p->c[0] = mtx_rand01(seed, k++) * dims[0];
        p->c[1] = mtx_rand01(seed, k++) * dims[1];
        p->c[2] = mtx_rand01(seed, k++) * dims[2];
        if (p->kind == 0) {
            p->r = (20.0 + 100.0 * mtx_rand01(seed, k++)) * scale;
            p->R = 0.0;
            p->axis[0] = p->axis[1] = 0.0;
            p->axis[2] = 1.0;
        } else {
            p->R = (40.0 + 110.0 * mtx_rand01(seed, k++)) * scale;
            p->r = (8.0 + 22.0 * mtx_rand01(seed, k++)) * scale;
            double u = mtx_rand01(seed, k++);
            double v = mtx_rand01(seed, k++);
            mtx_unit_vector(u, v, p->axis);
        }
    }
}

/* ---- per-voxel values --------------------------------------------------- */

static inline MTX_SHD double mtx_blob_field(const mtx_blob* b, int n, double x, double y, double z)
{
    double f = 0.0;
    for (int i = 0; i < n; ++i) {
        double dx = x - b[i].c[0], dy = y - b[i].c[1], dz = z - b[i].c[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        f += b[i].amp * exp(-d2 / (2.0 * b[i].sigma * b[i].sigma));
    }
    return f;
}

/* Quantise a blob field value to u8 with noise; fmin/fmax are the field's
 * global extrema over the volume. */
static inline MTX_SHD uint8_t mtx_blob_quantize(double f, double fmin, double fmax, uint64_t seed,
                                                uint64_t idx)
{
    double span = fmax - fmin;
    double q = span > 0.0 ? rint(255.0 * (f - fmin) / span) : 0.0;
    int noise = (int)(mtx_rand64(seed, MTX_NOISE_STREAM | idx) % 5ull) - 2;
    int v = (int)q + noise;
    return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

static inline MTX_SHD float mtx_blob_rescale(double f, double fmin, double fmax)
{
    double span = fmax - fmin;
    return (float)(span > 0.0 ? 2.0 * (f - fmin) / span - 1.0 : 0.0);
}

static inline MTX_SHD double mtx_segment_dist2(const mtx_segment* s, double x, double y, double z)
{
    double ab[3] = { s->b[0] - s->a[0], s->b[1] - s->a[1], s->b[2] - s->a[2] };
    double ap[3] = { x - s->a[0], y - s->a[1], z - s->a[2] };
    double l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    double t = l2 > 0.0 ? (ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2]) / l2 : 0.0;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    double d[3] = { ap[0] - t * ab[0], ap[1] - t * ab[1], ap[2] - t * ab[2] };
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

static inline MTX_SHD uint8_t mtx_vessel_value(const mtx_segment* s, int n, uint64_t seed,
                                               uint64_t idx, double x, double y, double z)
{
    double best = (double)(mtx_rand64(seed, MTX_NOISE_STREAM | idx) % 21ull);
    for (int i = 0; i < n; ++i) {
        double r = s[i].r;
        double d2 = mtx_segment_dist2(&s[i], x, y, z);
        /* 255*exp(-d2/(2r^2)) < 0.5 beyond d^2 > 2r^2*ln(510): cannot change the max */
        if (d2 > 12.48 * r * r) continue;
        double v = 255.0 * exp(-d2 / (2.0 * r * r));
        best = v > best ? v : best;
    }
    double q = rint(best);
    return (uint8_t)(q < 0.0 ? 0.0 : (q > 255.0 ? 255.0 : q));
}

static inline MTX_SHD double mtx_prim_sdf(const mtx_prim* p, double x, double y, double z)
{
    double q[3] = { x - p->c[0], y - p->c[1], z - p->c[2] };
    double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    if (p->kind == 0) return sqrt(q2) - p->r;
    double h = q[0] * p->axis[0] + q[1] * p->axis[1] + q[2] * p->axis[2];
    double rho2 = q2 - h * h;
    double rho = sqrt(rho2 > 0.0 ? rho2 : 0.0);
    double a = rho - p->R;
    return sqrt(a * a + h * h) - p->r;
}

/* Standard normal from the voxel's noise stream (Box-Muller). */
static inline MTX_SHD double mtx_gauss(uint64_t seed, uint64_t idx)
{
    uint64_t h = mtx_rand64(seed, MTX_NOISE_STREAM | idx);
    double u1 = ((double)(h >> 32) + 1.0) * (1.0 / 4294967296.0);  /* (0,1] */
    double u2 = (double)(h & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static inline MTX_SHD float mtx_sdf_value(const mtx_prim* p, int n, double sigma, uint64_t seed,
                                          uint64_t idx, double x, double y, double z)
{
    double d = 1e300;
    for (int i = 0; i < n; ++i) {
        double v = mtx_prim_sdf(&p[i], x, y, z);
        d = v < d ? v : d;
    }
    return (float)(d + sigma * mtx_gauss(seed, idx));
}

#endif /* MTX_SYNTH_RECIPES_H */
