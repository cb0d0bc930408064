// hbm_probe.hip -- streaming kernels of known size for the measurement side of the bench:
// the practical HBM ceiling (a copy and a read of a large buffer) and the calibration of the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters, whose readings are scaled by what they report
// for these kernels.  Not part of the solver's C ABI (include/lbm_rb.h).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// Each thread moves 16-byte vectors in a grid-stride loop, four of them in flight per
// iteration; consecutive lanes touch consecutive vectors, so every wave reads and writes whole
// 1 KiB spans.
__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                              size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t + 3 * stride < n; t += 4 * stride) {
        const uint4 a = src[t], b = src[t + stride], c = src[t + 2 * stride], d = src[t + 3 * stride];
        dst[t] = a;
        dst[t + stride] = b;
        dst[t + 2 * stride] = c;
        dst[t + 3 * stride] = d;
    }
    for (; t < n; t += stride) dst[t] = src[t];
}

// Read-only stream: an exclusive-or of every word, written once per thread so that no load is
// dead.
__global__ void __launch_bounds__(256) k_read(const uint4 *__restrict__ src, size_t n,
                                              unsigned *__restrict__ out)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned acc = 0;
    for (; t + 3 * stride < n; t += 4 * stride) {
        const uint4 a = src[t], b = src[t + stride], c = src[t + 2 * stride], d = src[t + 3 * stride];
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; t < n; t += stride) acc ^= src[t].x ^ src[t].y ^ src[t].z ^ src[t].w;
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
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
    const unsigned blocks = (unsigned)cus * 8u, threads = 256;
    uint4 *a = nullptr, *b = nullptr;
    unsigned *o = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0.f;
    if ((e = hipMalloc((void **)&a, n * 16)) != hipSuccess) goto done;
    if ((e = hipMalloc((void **)&b, n * 16)) != hipSuccess) goto done;
    if ((e = hipMalloc((void **)&o, (size_t)blocks * threads * sizeof(unsigned))) != hipSuccess) goto done;
    if ((e = hipEventCreate(&t0)) != hipSuccess || (e = hipEventCreate(&t1)) != hipSuccess) goto done;
    k_fill<<<blocks, threads>>>(a, n);
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
