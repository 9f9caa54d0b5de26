// GPU generator of the synthetic workloads (blobs, vessels, SDF) described in
// synth_recipes.h.  Separate from the extraction library: it makes inputs for
// the benchmark and the tests, it is not part of the extraction path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "synth_recipes.h"

namespace {

__global__ void k_blob_extrema(const mtx_blob* __restrict__ b, int n, uint32_t X, uint32_t Y,
                               uint32_t Z, unsigned long long* ext)
{
    const uint64_t N = (uint64_t)X * Y * Z;
    // the field is a sum of positive terms, so every value and both start
    // values are >= 0 and IEEE bit order is numeric order (lanes that see no
    // voxel must not start the maximum below zero)
    double lo = 1e300, hi = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(i % X), y = (uint32_t)((i / X) % Y), z = (uint32_t)(i / ((uint64_t)X * Y));
        double f = mtx_blob_field(b, n, x, y, z);
        lo = f < lo ? f : lo;
        hi = f > hi ? f : hi;
    }
    for (int d = 32; d > 0; d >>= 1) {
        double a = __shfl_xor(lo, d, 64), c = __shfl_xor(hi, d, 64);
        lo = a < lo ? a : lo;
        hi = c > hi ? c : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&ext[0], (unsigned long long)__double_as_longlong(lo));
        atomicMax(&ext[1], (unsigned long long)__double_as_longlong(hi));
    }
}

template <bool U8>
__global__ void k_blob_write(const mtx_blob* __restrict__ b, int n, uint32_t X, uint32_t Y, uint32_t Z,
                             const unsigned long long* ext, uint64_t seed, void* out)
{
    const uint64_t N = (uint64_t)X * Y * Z;
    const double lo = __longlong_as_double((long long)ext[0]);
    const double hi = __longlong_as_double((long long)ext[1]);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(i % X), y = (uint32_t)((i / X) % Y), z = (uint32_t)(i / ((uint64_t)X * Y));
        double f = mtx_blob_field(b, n, x, y, z);
        if (U8) static_cast<uint8_t*>(out)[i] = mtx_blob_quantize(f, lo, hi, seed, i);
        else static_cast<float*>(out)[i] = mtx_blob_rescale(f, lo, hi);
    }
}

__global__ void k_vessels(const mtx_segment* __restrict__ s, int n, uint32_t X, uint32_t Y, uint32_t Z,
                          uint64_t seed, uint8_t* out)
{
    const uint64_t N = (uint64_t)X * Y * Z;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(i % X), y = (uint32_t)((i / X) % Y), z = (uint32_t)(i / ((uint64_t)X * Y));
        out[i] = mtx_vessel_value(s, n, seed, i, x, y, z);
    }
}

__global__ void k_sdf(const mtx_prim* __restrict__ p, int n, double sigma, uint32_t X, uint32_t Y,
                      uint32_t Z, uint64_t seed, float* out)
{
    const uint64_t N = (uint64_t)X * Y * Z;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(i % X), y = (uint32_t)((i / X) % Y), z = (uint32_t)(i / ((uint64_t)X * Y));
        out[i] = mtx_sdf_value(p, n, sigma, seed, i, x, y, z);
    }
}

uint32_t grid_for(uint64_t n)
{
    uint64_t g = (n + 255) / 256;
    return (uint32_t)(g > 65536 ? 65536 : (g ? g : 1));
}

template <typename P>
int upload(const P* host, int n, P** dev, hipStream_t s)
{
    if (hipMallocAsync((void**)dev, sizeof(P) * (n > 0 ? n : 1), s) != hipSuccess) return 2;
    if (n > 0 && hipMemcpyAsync(*dev, host, sizeof(P) * n, hipMemcpyHostToDevice, s) != hipSuccess) return 3;
    return 0;
}

int blobs(void* out, const uint32_t dims[3], uint64_t seed, int nblobs, bool u8, hipStream_t s)
{
    if (!out || nblobs < 0 || nblobs > 4096) return 1;
    mtx_blob host[4096];
    mtx_recipe_blobs(seed, dims, nblobs, host);
    mtx_blob* dev = nullptr;
    unsigned long long* ext = nullptr;
    if (int r = upload(host, nblobs, &dev, s)) return r;
    if (hipMallocAsync((void**)&ext, 16, s) != hipSuccess) return 2;
    unsigned long long init[2] = { ~0ull, 0ull };
    hipMemcpyAsync(ext, init, 16, hipMemcpyHostToDevice, s);
    const uint64_t N = (uint64_t)dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(k_blob_extrema, dim3(grid_for(N)), dim3(256), 0, s, dev, nblobs, dims[0],
                       dims[1], dims[2], ext);
    if (u8)
        hipLaunchKernelGGL(k_blob_write<true>, dim3(grid_for(N)), dim3(256), 0, s, dev, nblobs,
                           dims[0], dims[1], dims[2], ext, seed, out);
    else
        hipLaunchKernelGGL(k_blob_write<false>, dim3(grid_for(N)), dim3(256), 0, s, dev, nblobs,
                           dims[0], dims[1], dims[2], ext, seed, out);
    hipFreeAsync(dev, s);
    hipFreeAsync(ext, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : 4;
}

}  // namespace

extern "C" {

// All return 0 on success; `out` is device memory of X*Y*Z elements.
int mtx_synth_blobs_u8(void* out, const uint32_t dims[3], uint64_t seed, int nblobs, void* stream)
{
    return blobs(out, dims, seed, nblobs, true, static_cast<hipStream_t>(stream));
}

int mtx_synth_blobs_f32(void* out, const uint32_t dims[3], uint64_t seed, int nblobs, void* stream)
{
    return blobs(out, dims, seed, nblobs, false, static_cast<hipStream_t>(stream));
}

int mtx_synth_vessels_u8(void* out, const uint32_t dims[3], uint64_t seed, int npoly, int nseg,
                         void* stream)
{
    if (!out || npoly < 0 || nseg < 0 || npoly * nseg > 65536) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    mtx_segment* host = new mtx_segment[(size_t)npoly * nseg + 1];
    mtx_recipe_vessels(seed, dims, npoly, nseg, host);
    mtx_segment* dev = nullptr;
    int r = upload(host, npoly * nseg, &dev, s);
    delete[] host;
    if (r) return r;
    const uint64_t N = (uint64_t)dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(k_vessels, dim3(grid_for(N)), dim3(256), 0, s, dev, npoly * nseg, dims[0],
                       dims[1], dims[2], seed, static_cast<uint8_t*>(out));
    hipFreeAsync(dev, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : 4;
}

int mtx_synth_sdf_f32(void* out, const uint32_t dims[3], uint64_t seed, int nsph, int ntor,
                      double scale, double sigma, void* stream)
{
    if (!out || nsph < 0 || ntor < 0 || nsph + ntor > 65536) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    mtx_prim* host = new mtx_prim[(size_t)(nsph + ntor) + 1];
    mtx_recipe_sdf(seed, dims, nsph, ntor, scale, host);
    mtx_prim* dev = nullptr;
    int r = upload(host, nsph + ntor, &dev, s);
    delete[] host;
    if (r) return r;
    const uint64_t N = (uint64_t)dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(k_sdf, dim3(grid_for(N)), dim3(256), 0, s, dev, nsph + ntor, sigma, dims[0],
                       dims[1], dims[2], seed, static_cast<float*>(out));
    hipFreeAsync(dev, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : 4;
}

}  // extern "C"

// Plain device copy of known size: the calibration kernel for HBM byte
// counters (FETCH_SIZE / WRITE_SIZE) of the benchmark.
namespace {
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                              uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        dst[i] = src[i];
}

// Read-only stream: every thread sums its 16-byte loads (8 in flight per
// iteration) and one value per wave is written, so the rate is the achievable
// HBM read rate the min/max kernel is compared with.  The loads are
// nontemporal: measured on MI355X that is the fastest read stream (6.9 TB/s
// against 6.1 TB/s for plain loads over 4 GiB).
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ src, float* __restrict__ out, uint64_t n)
{
    float acc = 0.0f;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const f4v* __restrict__ s4 = reinterpret_cast<const f4v*>(src);
    for (; i + 7 * stride < n; i += 8 * stride) {
        f4v v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(s4 + i + k * stride);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k].x + v[k].y + v[k].z + v[k].w;
    }
    for (; i < n; i += stride) {
        const float4 v = src[i];
        acc += v.x + v.y + v.z + v.w;
    }
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) out[blockIdx.x * 4 + (threadIdx.x >> 6)] = acc;
}
}  // namespace

// Reads `bytes` (a multiple of 16) once; `sink` holds 4 floats per workgroup
// (16384 workgroups: 256 KiB).
extern "C" int mtx_bench_read(const void* src, uint64_t bytes, void* sink, void* stream)
{
    if (!src || !sink || bytes % 16) return 1;
    hipLaunchKernelGGL(k_read, dim3(16384), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4*>(src), static_cast<float*>(sink), bytes / 16);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

extern "C" int mtx_bench_copy(void* dst, const void* src, uint64_t bytes, void* stream)
{
    if (!dst || !src || bytes % 16) return 1;
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4*>(src), static_cast<float4*>(dst), bytes / 16);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}
