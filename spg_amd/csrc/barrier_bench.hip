// Calibration of a grid-wide barrier inside one cooperative launch against back-to-back kernel launches: the
// measurement behind running one launch per colour (replayed from a hipGraph) instead of one persistent launch
// that crosses colours at a barrier. Not part of the solver.
#include <hip/hip_runtime.h>

#include <string>

#include "host_util.h"
#include "spg_vbd.h"

using namespace spgvbd;

namespace {

// Every spin is bounded: a block whose wait exceeds kSpinLimit polls raises the abort flag, and every block
// leaves at its next wait, so the kernel ends whatever happens.
constexpr unsigned kSpinLimit = 1u << 20;

__device__ __forceinline__ unsigned loadRelaxed(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// `rounds` barriers. Blocks are counted in groups of `group` consecutive blocks, each with its own arrival
// counter (group == gridDim.x: one grid-wide barrier); arrival is a release add, the wait spins on relaxed loads
// with an acquire fence after it, as a per-colour barrier of a persistent solve would.
__global__ void barrierRoundsKernel(unsigned *counters, unsigned *abortFlag, int32_t rounds, int32_t group,
                                    float *sink)
{
    __shared__ int ok;
    const unsigned g = blockIdx.x / group;
    unsigned *cnt = counters + 64 * g;  // one 256-byte line per group
    float acc = 0.f;
    for (int32_t r = 1; r <= rounds; ++r) {
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = static_cast<unsigned>(r) * static_cast<unsigned>(group);
            unsigned n = 0;
            int good = 1;
            while (loadRelaxed(cnt) < target) {
                if (++n > kSpinLimit || loadRelaxed(abortFlag)) {
                    __hip_atomic_store(abortFlag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ok = good;
        }
        __syncthreads();
        if (!ok) break;
        acc += 1.f;
    }
    if (threadIdx.x == 0) sink[blockIdx.x] = acc;
}

__global__ void emptyKernel(float *sink)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && sink[0] < -1.f) sink[1] = 0.f;
}

__global__ void zeroKernel(unsigned *p, int32_t n)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

extern "C" int spg_vbd_bench_barrier(int32_t grid, int32_t block, int32_t group, int32_t rounds, int32_t device,
                                     double *us_per_barrier, double *us_per_launch)
{
    try {
        if (grid < 1 || grid > 8192 || block < 64 || block > 1024 || block % 64 || group < 1 || group > grid ||
            grid % group || rounds < 1 || rounds > 100000 || !us_per_barrier || !us_per_launch)
            return SPG_VBD_ERR_INVALID;
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        const int32_t nCounters = 64 * (grid / group) + 64;  // group lines, then the abort flag's line
        DevBuf<unsigned> counters;
        DevBuf<float> sink;
        counters.alloc(nCounters);
        sink.alloc(grid);
        unsigned *abortFlag = counters.p + nCounters - 64;
        hipEvent_t e0, e1;
        hipCheck(hipEventCreate(&e0), "hipEventCreate");
        hipCheck(hipEventCreate(&e1), "hipEventCreate");
        float ms = 0.f;
        // back-to-back launches of an empty kernel of the same shape
        for (int k = 0; k < 10; ++k) emptyKernel<<<grid, block, 0, s>>>(sink.p);
        hipCheck(hipEventRecord(e0, s), "hipEventRecord");
        for (int32_t k = 0; k < rounds; ++k) emptyKernel<<<grid, block, 0, s>>>(sink.p);
        hipCheck(hipEventRecord(e1, s), "hipEventRecord");
        hipCheck(hipEventSynchronize(e1), "hipEventSynchronize");
        hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        *us_per_launch = 1e3 * ms / rounds;
        // the barrier rounds in one cooperative launch (refused by the runtime when the blocks cannot all be
        // resident at once)
        zeroKernel<<<(nCounters + 255) / 256, 256, 0, s>>>(counters.p, nCounters);
        void *args[] = {&counters.p, &abortFlag, &rounds, &group, &sink.p};
        hipCheck(hipEventRecord(e0, s), "hipEventRecord");
        hipCheck(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(barrierRoundsKernel), dim3(grid),
                                            dim3(block), args, 0, s),
                 "cooperative launch");
        hipCheck(hipEventRecord(e1, s), "hipEventRecord");
        hipCheck(hipEventSynchronize(e1), "hipEventSynchronize");
        hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        unsigned aborted = 0;
        hipCheck(hipMemcpy(&aborted, abortFlag, sizeof(aborted), hipMemcpyDeviceToHost), "download");
        *us_per_barrier = aborted ? -1.0 : 1e3 * ms / rounds;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        return SPG_VBD_OK;
    } catch (const VbdError &e) {
        return e.code;
    } catch (...) {
        return SPG_VBD_ERR_HIP;
    }
}
