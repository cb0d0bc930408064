// hbm_probe.hip -- streaming kernels of known size for the measurement side of the bench:
// the practical HBM ceilings (a copy and a read of a large buffer, and a copy with the access
// pattern of the one-pass step) and the calibration of the rocprofv3 FETCH_SIZE / WRITE_SIZE
// counters, whose readings are scaled by what they report for the copy.  Not part of the
// solver's C ABI (include/lbm_rb.h).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// Copy and read: each 1024-thread block streams one contiguous range of 16-byte vectors, with
// four (copy) or eight (read) vectors per lane in flight; consecutive lanes touch consecutive
// vectors, so every wave reads and writes whole 1 KiB spans.  Measured on one MI355X over 4 GiB
// (one block per CU): copy 5.47 TB/s and read 6.07 TB/s, against 4.6-4.7 and 5.4-5.5 TB/s for the
// grid-stride forms of the same loops (256-thread blocks, 8 per CU), 6.0 TB/s for the grid-stride
// read with non-temporal loads, and 4.8 TB/s for a grid-stride write of the same size.
template <int U>
__device__ __forceinline__ void k_range(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                size_t n, unsigned *__restrict__ out)
{
    const size_t per = (n + gridDim.x - 1) / gridDim.x;
    const size_t b0 = (size_t)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
    size_t t = b0 + threadIdx.x;
    unsigned acc = 0;
    for (; t + (U - 1) * 1024 < b1; t += U * 1024) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[t + u * 1024];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (dst) dst[t + u * 1024] = v[u];
            else acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
        }
    }
    for (; t < b1; t += 1024) {
        if (dst) dst[t] = src[t];
        else acc ^= src[t].x;
    }
    if (!dst) out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;   // no load is dead
}

__global__ void __launch_bounds__(1024) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    k_range<4>(src, dst, n, nullptr);
}

__global__ void __launch_bounds__(1024) k_read(const uint4 *__restrict__ src, size_t n, unsigned *__restrict__ out)
{
    k_range<8>(src, nullptr, n, out);
}

// The one-pass step's access pattern without its arithmetic: 32 x 13 node tiles of an
// nx x ny x nz field of doubles, each block marching a chunk of 32 planes and copying A arrays
// (plane pitch P elements) node by node, all loads of a plane before its stores, with the step's
// XCD-contiguous block order.  With A = 29 (the FP64 one-pass step reads and writes 19 flow and
// 7 thermal populations and 3 history values per node) a launch moves the step's algorithmic
// bytes.
constexpr int TXN = 32, TYN = 13, KCH = 32, TILE_ARRAYS = 29;
template <int A>
__global__ void __launch_bounds__(TXN *TYN) k_tile_copy(const double *__restrict__ src, double *__restrict__ dst,
                                                        int nx, int ny, int nz, size_t P)
{
    const unsigned ntx = (nx + TXN - 1) / TXN, nty = (ny + TYN - 1) / TYN, nb = gridDim.x;
    const unsigned per = nb / 8, rem = nb % 8, x = blockIdx.x % 8, r = blockIdx.x / 8;
    const unsigned lb = r < per + (x < rem ? 1u : 0u) ? x * per + (x < rem ? x : rem) + r : blockIdx.x;
    const unsigned tx = lb % ntx, ty = (lb / ntx) % nty, ch = lb / (ntx * nty);
    const int i = (int)tx * TXN + (int)threadIdx.x % TXN, j = (int)ty * TYN + (int)threadIdx.x / TXN;
    if (i >= nx || j >= ny) return;
    const int k1 = min(nz, (int)(ch + 1) * KCH);
    for (int k = (int)ch * KCH; k < k1; ++k) {
        const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
        double v[A];
#pragma unroll
        for (int a = 0; a < A; ++a) v[a] = src[(size_t)a * P + c];
#pragma unroll
        for (int a = 0; a < A; ++a) dst[(size_t)a * P + c] = v[a];
    }
}

__global__ void k_fill(uint4 *__restrict__ p, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
        p[t] = make_uint4((unsigned)t, (unsigned)(t >> 32), 0x9E3779B9u, (unsigned)t * 3u);
}

} // namespace

// Copy and read `bytes` (rounded down to 16) `reps` times each on `device` and report the
// bandwidths in GB/s: copy counts the bytes read plus the bytes written.  The buffers are
// allocated and freed by the call.  Returns 0, or the HIP error code.
extern "C" int hbm_probe(int device, uint64_t bytes, int reps, double *copy_gbs, double *read_gbs)
{
    if (reps <= 0 || bytes < 16) return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    const size_t n = bytes / 16;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return (int)e;
    const unsigned blocks = (unsigned)cus, threads = 1024;
    uint4 *a = nullptr, *b = nullptr;
    unsigned *o = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0.f;
    if ((e = hipMalloc((void **)&a, n * 16)) != hipSuccess) goto done;
    if ((e = hipMalloc((void **)&b, n * 16)) != hipSuccess) goto done;
    if ((e = hipMalloc((void **)&o, (size_t)blocks * threads * sizeof(unsigned))) != hipSuccess) goto done;
    if ((e = hipEventCreate(&t0)) != hipSuccess || (e = hipEventCreate(&t1)) != hipSuccess) goto done;
    k_fill<<<blocks * 8, 256>>>(a, n);
    k_copy<<<blocks, threads>>>(a, b, n);   // warm-up
    k_read<<<blocks, threads>>>(a, n, o);
    if ((e = hipDeviceSynchronize()) != hipSuccess) goto done;
    if ((e = hipEventRecord(t0, 0)) != hipSuccess) goto done;
    for (int r = 0; r < reps; ++r) k_copy<<<blocks, threads>>>(a, b, n);
    if ((e = hipEventRecord(t1, 0)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
        (e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess)
        goto done;
    if (copy_gbs) *copy_gbs = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
    if ((e = hipEventRecord(t0, 0)) != hipSuccess) goto done;
    for (int r = 0; r < reps; ++r) k_read<<<blocks, threads>>>(a, n, o);
    if ((e = hipEventRecord(t1, 0)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
        (e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess)
        goto done;
    if (read_gbs) *read_gbs = (double)n * 16 * reps / (ms * 1e-3) / 1e9;
    e = hipGetLastError();
done:
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (o) (void)hipFree(o);
    return (int)e;
}

// The step's access pattern: copy 29 arrays of an nx x ny x nz field of doubles in 32 x 13
// tiles marching 32-plane chunks, `reps` times, and report GB/s (bytes read plus written).
// Allocates and frees 2 x 29 x (nx ny nz + padding) doubles.
extern "C" int hbm_probe_tile(int device, int nx, int ny, int nz, int reps, double *gbs)
{
    if (reps <= 0 || nx <= 0 || ny <= 0 || nz <= 0) return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    const size_t N = (size_t)nx * ny * nz;
    const size_t P = N + 131072 + 128;   // padded plane pitch, as the solver pads its planes
    const size_t bytes = (size_t)TILE_ARRAYS * P * sizeof(double);
    const unsigned nb = (unsigned)(((nx + TXN - 1) / TXN) * ((ny + TYN - 1) / TYN) * ((nz + KCH - 1) / KCH));
    double *a = nullptr, *b = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0.f;
    if ((e = hipMalloc((void **)&a, bytes)) != hipSuccess) goto done;
    if ((e = hipMalloc((void **)&b, bytes)) != hipSuccess) goto done;
    if ((e = hipMemset(a, 0, bytes)) != hipSuccess) goto done;
    if ((e = hipEventCreate(&t0)) != hipSuccess || (e = hipEventCreate(&t1)) != hipSuccess) goto done;
    k_tile_copy<TILE_ARRAYS><<<nb, TXN * TYN>>>(a, b, nx, ny, nz, P);   // warm-up
    if ((e = hipDeviceSynchronize()) != hipSuccess) goto done;
    if ((e = hipEventRecord(t0, 0)) != hipSuccess) goto done;
    for (int r = 0; r < reps; ++r) k_tile_copy<TILE_ARRAYS><<<nb, TXN * TYN>>>(a, b, nx, ny, nz, P);
    if ((e = hipEventRecord(t1, 0)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
        (e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess)
        goto done;
    if (gbs) *gbs = 2.0 * (double)N * TILE_ARRAYS * sizeof(double) * reps / (ms * 1e-3) / 1e9;
    e = hipGetLastError();
done:
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return (int)e;
}
