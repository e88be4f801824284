// qnd_probe.hip -- latency probes behind a design decision of libqnd (built
// as libqnd_probe.so, loaded by bench.py; not part of the C ABI).
//
// The 512^2 step is a chain of 12 dependent line passes, each a kernel.  A
// persistent kernel would replace each kernel boundary by a grid-wide
// barrier.  These probes time both on the device:
//   qnd_probe_barrier  : `iters` grid barriers in one cooperative launch of
//                        `blocks` workgroups (a counter advanced with
//                        device-scope release atomics and polled with
//                        device-scope acquire loads, the cache maintenance a
//                        real persistent pass needs);
//   qnd_probe_boundary : `iters` back-to-back launches of an empty kernel of
//                        `blocks` workgroups on one stream.
// Every wait is bounded: a workgroup that polls more than `spin_max` times
// sets the error flag and leaves, so the kernel always ends.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ void k_probe_barrier(unsigned *count, unsigned *err, int iters, unsigned spin_max)
{
    for (int k = 0; k < iters; ++k) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned target = (unsigned)(k + 1) * gridDim.x;
            __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            unsigned spins = 0;
            while (__hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++spins > spin_max) {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
    }
}

__global__ void k_probe_empty(unsigned *sink)
{
    if (sink && threadIdx.x == 0 && blockIdx.x == 0x7fffffff) *sink = 1u;  // never taken
}

}  // namespace

extern "C" {

// Microseconds per grid barrier (*us), or a negative value on error: -1 the
// launch was refused (e.g. the workgroups cannot all be resident), -2 a
// bounded wait expired.
int qnd_probe_barrier(int blocks, int threads, int iters, double *us)
{
    *us = -1.0;
    unsigned *d = nullptr;
    if (hipMalloc(&d, 2 * sizeof(unsigned)) != hipSuccess) return -1;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        hipFree(d);
        return -1;
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int rc = 0;
    for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // the first launch warms up
        hipMemsetAsync(d, 0, 2 * sizeof(unsigned), st);
        unsigned *count = d, *err = d + 1;
        unsigned spin_max = 1u << 20;
        int it = iters;
        void *args[] = {&count, &err, &it, &spin_max};
        hipEventRecord(e0, st);
        if (hipLaunchCooperativeKernel((const void *)k_probe_barrier, dim3(blocks), dim3(threads), args, 0, st) !=
            hipSuccess) {
            (void)hipGetLastError();
            rc = -1;
            break;
        }
        hipEventRecord(e1, st);
        hipStreamSynchronize(st);
        unsigned h[2] = {0, 0};
        hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
        if (h[1] != 0u) rc = -2;
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *us = 1e3 * ms / iters;
    }
    if (rc) *us = rc;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(d);
    return rc;
}

// Microseconds per launch of an empty kernel of `blocks` x `threads`, from
// `iters` back-to-back launches on one stream.
int qnd_probe_boundary(int blocks, int threads, int iters, double *us)
{
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (int k = 0; k < 16; ++k) hipLaunchKernelGGL(k_probe_empty, dim3(blocks), dim3(threads), 0, st, nullptr);
    hipEventRecord(e0, st);
    for (int k = 0; k < iters; ++k) hipLaunchKernelGGL(k_probe_empty, dim3(blocks), dim3(threads), 0, st, nullptr);
    hipEventRecord(e1, st);
    hipStreamSynchronize(st);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    *us = 1e3 * ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    return 0;
}

}  // extern "C"
