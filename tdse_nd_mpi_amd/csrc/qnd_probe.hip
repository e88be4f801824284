// qnd_probe.hip -- latency probes behind a design decision of libqnd (built
// as libqnd_probe.so, loaded by bench.py; not part of the C ABI).
//
// The 512^2 step is a chain of 12 dependent line passes, each a kernel.  A
// persistent kernel would replace each kernel boundary by a grid-wide
// barrier.  These probes time both on the device:
//   qnd_probe_barrier  : `iters` grid barriers in one cooperative launch of
//                        `blocks` workgroups, in one of three forms:
//                        0  one counter, advanced with device-scope release
//                           atomics and polled with device-scope acquire
//                           loads (each poll invalidates the L2);
//                        1  one counter, polled with relaxed device-scope
//                           loads and one acquire fence once it is reached;
//                        2  two levels: a counter per group of workgroups
//                           (blockIdx mod 8, the XCD the dispatcher deals the
//                           workgroup to), whose last arrival advances the
//                           top counter; relaxed polls of the top counter,
//                           then one acquire fence.
//                        3  no shared counter: each workgroup stores its own
//                           arrival flag (one 128-byte line each), the first
//                           workgroup polls all flags with relaxed loads (one
//                           flag per thread) and then stores a release flag,
//                           which the others poll; one acquire fence.
//                        Every form makes the data written before the
//                        barrier visible after it (release before arriving,
//                        acquire after leaving), as a persistent pass needs;
//   qnd_probe_boundary : `iters` back-to-back launches of an empty kernel of
//                        `blocks` workgroups on one stream;
//   qnd_probe_tiles    : the memory side of a 512^3 complex64 line pass alone:
//                        an in-place scaled copy of the state read and
//                        written as tiles of TW lines of 512 points along one
//                        axis, with the passes' lane mapping and XCD tile
//                        order (what the tile shape costs without the
//                        transforms).
// Every wait is bounded: a workgroup that polls more than `spin_max` times
// sets the error flag and leaves, so the kernel always ends.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// counters: [0] top (or the only) counter, [64 + 64 g] group g's counter
// (separate 256-byte lines), err: the error flag
// flags[32 * w]: workgroup w's arrival (form 3); flags[32 * gridDim.x]: the
// release flag
__device__ void barrier_flags(unsigned *flags, unsigned *err, unsigned k, unsigned spin_max)
{
    __shared__ int pending;
    unsigned *rel = flags + 32 * gridDim.x;
    if (threadIdx.x == 0) __hip_atomic_store(flags + 32 * blockIdx.x, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0) {
        unsigned spins = 0;
        for (;;) {
            int mine = 0;
            for (unsigned w = threadIdx.x; w < gridDim.x; w += blockDim.x)
                mine |= __hip_atomic_load(flags + 32 * w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k;
            if (threadIdx.x == 0) pending = 0;
            __syncthreads();
            if (mine) pending = 1;
            __syncthreads();
            const int p = pending;
            __syncthreads();
            if (!p) break;
            if (++spins > spin_max) {
                if (threadIdx.x == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (threadIdx.x == 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(rel, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(rel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
            if (++spins > spin_max) {
                __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
}

__global__ void k_probe_barrier(unsigned *cnt, unsigned *err, int iters, unsigned spin_max, int form)
{
    const unsigned ng = 8, g = blockIdx.x % ng;
    const unsigned members = gridDim.x / ng + (g < gridDim.x % ng ? 1u : 0u);
    for (int k = 0; k < iters; ++k) {
        __syncthreads();
        if (form == 3) {
            barrier_flags(cnt, err, (unsigned)(k + 1), spin_max);
        } else if (threadIdx.x == 0) {
            unsigned *top = cnt;
            unsigned target;
            if (form == 2) {
                unsigned *gc = cnt + 64 + 64 * g;
                const unsigned old = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                if (old + 1u == (unsigned)(k + 1) * members)
                    __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                target = (unsigned)(k + 1) * (gridDim.x < ng ? gridDim.x : ng);
            } else {
                __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                target = (unsigned)(k + 1) * gridDim.x;
            }
            unsigned spins = 0;
            const int order = form == 0 ? __ATOMIC_ACQUIRE : __ATOMIC_RELAXED;
            while (__hip_atomic_load(top, order, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++spins > spin_max) {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (form != 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
    }
}

// Tile pass of TW lines (VPT values per thread, T = 512 / VPT threads per
// line): line L starts at ((L >> log2S) << (9 + log2S)) + (L & (S - 1)) and
// its element i is i * S further (S = 1: contiguous lines at L * 512).  Every
// index is inside the 512^3 grid for L < 512^2.
template <int TW, int VPT>
__global__ void __launch_bounds__(1024) k_probe_tile(float2 *data, int64_t S, int log2S, float scale)
{
    constexpr int n = 512, T = n / VPT;
    constexpr int log2TW = TW == 1 ? 0 : TW == 2 ? 1 : TW == 4 ? 2 : TW == 8 ? 3 : TW == 16 ? 4 : 5;
    constexpr int log2T = T == 32 ? 5 : T == 64 ? 6 : 7;
    const int tid = threadIdx.x, bid = blockIdx.x, nblk = gridDim.x;
    const int tb = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    int l, t;
    if (S == 1) { l = tid >> log2T; t = tid & (T - 1); }
    else { l = tid & (TW - 1); t = tid >> log2TW; }
    const int64_t L = (int64_t)tb * TW + l;
    const int64_t base = S == 1 ? L * n : ((L >> log2S) << (9 + log2S)) + (L & (S - 1));
    float2 x[VPT];
#pragma unroll
    for (int m = 0; m < VPT; ++m) x[m] = data[base + (int64_t)(t + m * T) * S];
#pragma unroll
    for (int m = 0; m < VPT; ++m) {
        x[m].x *= scale;
        x[m].y *= scale;
    }
#pragma unroll
    for (int m = 0; m < VPT; ++m) data[base + (int64_t)(t + m * T) * S] = x[m];
}

__global__ void k_probe_empty(unsigned *sink)
{
    if (sink && threadIdx.x == 0 && blockIdx.x == 0x7fffffff) *sink = 1u;  // never taken
}

}  // namespace

extern "C" {

// Microseconds per grid barrier of form `form` (*us), or a negative value on
// error: -1 the launch was refused (e.g. the workgroups cannot all be
// resident), -2 a bounded wait expired.
#define QND_PROBE_WORDS (32 * 1025 + 64 * 8 + 1)  // form 3's flags for up to 1024 workgroups, or form 2's
int qnd_probe_barrier(int blocks, int threads, int iters, int form, double *us)
{
    *us = -1.0;
    unsigned *d = nullptr;
    if (hipMalloc(&d, QND_PROBE_WORDS * sizeof(unsigned)) != hipSuccess) return -1;
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
        hipMemsetAsync(d, 0, QND_PROBE_WORDS * sizeof(unsigned), st);
        unsigned *count = d, *err = d + QND_PROBE_WORDS - 1;
        unsigned spin_max = 1u << 20;
        int it = iters, fm = form;
        void *args[] = {&count, &err, &it, &spin_max, &fm};
        hipEventRecord(e0, st);
        if (hipLaunchCooperativeKernel((const void *)k_probe_barrier, dim3(blocks), dim3(threads), args, 0, st) !=
            hipSuccess) {
            (void)hipGetLastError();
            rc = -1;
            break;
        }
        hipEventRecord(e1, st);
        hipStreamSynchronize(st);
        unsigned h = 0;
        hipMemcpy(&h, d + QND_PROBE_WORDS - 1, sizeof h, hipMemcpyDeviceToHost);
        if (h != 0u) rc = -2;
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

// Microseconds per tile pass over a 512^3 complex64 grid along `axis` (0, 1
// or 2) with `tw` lines per tile (1 to 32) and `vpt` values per thread (8 or
// 16; tw * 512 / vpt <= 1024 threads), averaged over `iters` passes; -1 on a
// refused shape or a HIP error.
int qnd_probe_tiles(int axis, int tw, int vpt, int iters, double *us)
{
    *us = -1.0;
    const int64_t n = 512, N = n * n * n;
    const int64_t S = axis == 0 ? n * n : axis == 1 ? n : 1;
    const int log2S = axis == 0 ? 18 : axis == 1 ? 9 : 0;
    const int threads = tw * (int)(n / vpt);
    if (threads > 1024 || (vpt != 8 && vpt != 16)) return -1;
    const int blocks = (int)(N / n / tw);
    const void *f = nullptr;
#define QND_TILE(TWc, VPTc) \
    if (tw == TWc && vpt == VPTc) f = (const void *)k_probe_tile<TWc, VPTc>;
    QND_TILE(1, 8) QND_TILE(2, 8) QND_TILE(4, 8) QND_TILE(8, 8) QND_TILE(16, 8)
    QND_TILE(4, 16) QND_TILE(8, 16) QND_TILE(16, 16) QND_TILE(32, 16)
#undef QND_TILE
    if (!f) return -1;
    float2 *d = nullptr;
    if (hipMalloc(&d, N * sizeof(float2)) != hipSuccess) return -1;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        hipFree(d);
        return -1;
    }
    hipMemsetAsync(d, 0, N * sizeof(float2), st);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int64_t Sa = S;
    int l2 = log2S;
    float one = 1.0f;
    void *args[] = {&d, &Sa, &l2, &one};
    int rc = 0;
    for (int k = 0; k < 2 && rc == 0; ++k)  // warm-up
        if (hipLaunchKernel(f, dim3(blocks), dim3(threads), args, 0, st) != hipSuccess) rc = -1;
    hipEventRecord(e0, st);
    for (int k = 0; k < iters && rc == 0; ++k)
        if (hipLaunchKernel(f, dim3(blocks), dim3(threads), args, 0, st) != hipSuccess) rc = -1;
    hipEventRecord(e1, st);
    if (hipStreamSynchronize(st) != hipSuccess) rc = -1;
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (rc == 0) *us = 1e3 * ms / iters;
    (void)hipGetLastError();
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(d);
    return rc;
}

}  // extern "C"
