// probe_grid_sync.hip -- measures what one grid-wide synchronisation costs on the device,
// against what one kernel boundary costs, for a grid of one 256-thread workgroup per CU.
// It decides whether a sequence of small dependent stages (the network's 16^3 and 8^3
// layers, the multigrid's coarse levels) is better run as one persistent kernel with a
// barrier between stages or as one kernel per stage.
//
// Each stage writes 4 KB per workgroup and, after the synchronisation, reads the block
// written by the next workgroup (double-buffered by stage parity, so a workgroup that runs
// ahead never overwrites a block still to be read); a mismatch is counted.  Variants:
//   boundary      one kernel launch per stage (stream order is the synchronisation)
//   cg            cooperative_groups::this_grid().sync() in a cooperative launch
//   flat          one agent-scope atomic counter, every workgroup polls it
//   hier          eight counters (workgroups grouped by blockIdx % 8), the last arrival
//                 of each group bumps a top counter that every workgroup polls
// The atomic barriers bound their polling (2^22 polls) and record a timeout instead of
// waiting on, so no variant can hang.
//
//   make -C caffe_old_amd/csrc _build/probe_grid_sync && caffe_old_amd/csrc/_build/probe_grid_sync
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cstdio>

namespace cg = cooperative_groups;

#define CK(x)                                                                                  \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            std::printf("%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_));              \
            return 1;                                                                          \
        }                                                                                      \
    } while (0)

constexpr int BLK = 1024;        // floats per workgroup block of a buffer
constexpr int MAXG = 1024;       // largest grid the buffers hold
constexpr int POLLS = 1 << 22;   // bounded polling

__device__ __forceinline__ void writeStage(float* buf, int stage) {
    float* b = buf + (size_t)(stage & 1) * MAXG * BLK + (size_t)blockIdx.x * BLK;
    b[threadIdx.x * 4] = (float)stage;
}
__device__ __forceinline__ void checkStage(const float* buf, int stage, int* err) {
    const float* b = buf + (size_t)(stage & 1) * MAXG * BLK + (size_t)((blockIdx.x + 1) % gridDim.x) * BLK;
    if (b[threadIdx.x * 4] != (float)stage) atomicAdd(err, 1);
}
__device__ __forceinline__ void spinUntil(const unsigned* c, unsigned target, int* timeouts) {
    int polls = 0;
    while (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++polls > POLLS) {
            atomicAdd(timeouts, 1);
            break;
        }
    }
}

__global__ void k_stage(float* buf, int stage, int* err) {
    if (stage > 0) checkStage(buf, stage - 1, err);
    writeStage(buf, stage);
}

__global__ void k_cg(float* buf, int nst, int* err) {
    cg::grid_group g = cg::this_grid();
    for (int s = 0; s < nst; ++s) {
        writeStage(buf, s);
        g.sync();
        checkStage(buf, s, err);
    }
}

__global__ void k_flat(unsigned* cnt, float* buf, int nst, int* err, int* timeouts) {
    for (int s = 0; s < nst; ++s) {
        writeStage(buf, s);
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            spinUntil(cnt, (unsigned)(s + 1) * gridDim.x, timeouts);
        }
        __syncthreads();
        checkStage(buf, s, err);
    }
}

// cnt[0]: top counter; cnt[64 * (1 + g)]: group g (own 256-byte line)
__global__ void k_hier(unsigned* cnt, float* buf, int nst, int* err, int* timeouts) {
    const int g = blockIdx.x & 7;
    const unsigned gsize = (gridDim.x - g + 7) / 8;   // workgroups with blockIdx % 8 == g
    for (int s = 0; s < nst; ++s) {
        writeStage(buf, s);
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned old = __hip_atomic_fetch_add(cnt + 64 * (1 + g), 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1 == (unsigned)(s + 1) * gsize)
                __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            spinUntil(cnt, (unsigned)(s + 1) * 8, timeouts);
        }
        __syncthreads();
        checkStage(buf, s, err);
    }
}

int main() {
    int ncu = 0;
    CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
    const int grid = ncu <= MAXG ? ncu : MAXG;
    unsigned* cnt;
    int *err, *tmo;
    float* buf;
    CK(hipMalloc(&cnt, 64 * 9 * sizeof(unsigned)));
    CK(hipMalloc(&err, sizeof(int)));
    CK(hipMalloc(&tmo, sizeof(int)));
    CK(hipMalloc(&buf, 2 * (size_t)MAXG * BLK * sizeof(float)));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    std::printf("CUs %d, grid %d workgroups x 256 threads\n", ncu, grid);
    const int nst = 400;
    for (int variant = 0; variant < 4; ++variant) {
        for (int rep = 0; rep < 2; ++rep) {   // the first repetition warms up
            CK(hipMemset(cnt, 0, 64 * 9 * sizeof(unsigned)));
            CK(hipMemset(err, 0, sizeof(int)));
            CK(hipMemset(tmo, 0, sizeof(int)));
            CK(hipDeviceSynchronize());
            CK(hipEventRecord(e0, 0));
            int n = nst;
            if (variant == 0) {
                for (int s = 0; s < nst; ++s) hipLaunchKernelGGL(k_stage, dim3(grid), dim3(256), 0, 0, buf, s, err);
                CK(hipGetLastError());
            } else if (variant == 1) {
                void* args[] = {&buf, &n, &err};
                CK(hipLaunchCooperativeKernel((const void*)k_cg, dim3(grid), dim3(256), args, 0, 0));
            } else if (variant == 2) {
                void* args[] = {&cnt, &buf, &n, &err, &tmo};
                CK(hipLaunchCooperativeKernel((const void*)k_flat, dim3(grid), dim3(256), args, 0, 0));
            } else {
                void* args[] = {&cnt, &buf, &n, &err, &tmo};
                CK(hipLaunchCooperativeKernel((const void*)k_hier, dim3(grid), dim3(256), args, 0, 0));
            }
            CK(hipEventRecord(e1, 0));
            CK(hipDeviceSynchronize());
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, e0, e1));
            int he = 0, ht = 0;
            CK(hipMemcpy(&he, err, sizeof(int), hipMemcpyDeviceToHost));
            CK(hipMemcpy(&ht, tmo, sizeof(int), hipMemcpyDeviceToHost));
            static const char* names[] = {"boundary", "cg", "flat", "hier"};
            if (rep == 1)
                std::printf("%-8s %4d stages: %8.1f us, %6.2f us per stage, mismatches %d, timeouts %d\n", names[variant],
                            nst, ms * 1e3, ms * 1e3 / nst, he, ht);
        }
    }
    return 0;
}
