// The fused per-world step: one workgroup (one 64-lane wavefront) advances one
// world through PlayRho's regular step, Step(AabbTreeWorld&, const StepConf&)
// (AabbTreeWorld.cpp:2144-2219):
//
//   A  proxy creation for new shapes              CreateProxies        :557-579
//   B  contact destruction (fat AABBs apart)       DestroyContacts      :2290-2331
//   C  narrowphase of flagged contacts             UpdateContacts/Update:2333-2421,2587-2721
//   D  pair finding + contact creation             FindContacts/AddContacts :709-751,2424-2547
//   E  islands (DFS), per-island Gauss-Seidel solve, writeback, sleep
//                                                  SolveReg/SolveRegIslandViaGS :1472-1682
//   F  proxy synchronisation, C and D again        SolveReg             :1532-1557
//   G  (doToi) the TOI phase, in a second kernel   SolveToi             :1752-1831
//
// Lane-parallel phases: proxy AABBs, destruction with stable compaction
// (ballot prefix counts), narrowphase (one lane per contact), candidate pair
// tests, key sort, contact creation, per-body adjacency, synchronisation.
// Order-dependent phases replay the reference order exactly: the island DFS
// (one lane) and each island's Gauss-Seidel solve (one lane per island;
// islands share no dynamic body, so they are independent).
//
// A batch of worlds is one launch: workgroup g steps world schedule[g]
// (largest worlds first), so uneven world sizes balance across the 256 CUs.
#include "step_device.hpp"

namespace prxd {

// At most 168 VGPRs, so that three waves share a SIMD.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_step_worlds(DeviceState d,
                                                                                          StepParams c) {
    const int w = d.schedule[blockIdx.x];
    step_world<false>(d, w, c);
}

// The TOI phase (doToi) of every world, after k_step_worlds in the same stream.
__global__ __launch_bounds__(64) void k_toi_worlds(DeviceState d, StepParams c) {
    const int w = d.schedule[blockIdx.x];
    toi_world(d, w, c);
}

// The same step recording its solve for the adjoint (every island swept by
// one lane: coopMin must exceed every island's constraint count).
__global__ __launch_bounds__(64) void k_step_worlds_tape(DeviceState d, StepParams c) {
    const int w = d.schedule[blockIdx.x];
    step_world<true>(d, w, c);
}

}  // namespace prxd

// Host-side launcher (compiled in this translation unit).
extern "C" hipError_t prx_launch_step(const prx::DeviceState* d, const prx::StepParams* c, int first, int count,
                                      hipStream_t stream) {
    prx::DeviceState dd = *d;
    dd.schedule = d->schedule + first;
    hipLaunchKernelGGL(prxd::k_step_worlds, dim3(unsigned(count)), dim3(64), c->ldsBytes, stream, dd, *c);
    if (c->doToi && c->dt != 0.0f) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(prxd::k_toi_worlds, dim3(unsigned(count)), dim3(64), 0, stream, dd, *c);
    }
    return hipGetLastError();
}

extern "C" hipError_t prx_launch_step_tape(const prx::DeviceState* d, const prx::StepParams* c, int count,
                                           hipStream_t stream) {
    hipLaunchKernelGGL(prxd::k_step_worlds_tape, dim3(unsigned(count)), dim3(64), c->ldsBytes, stream, *d, *c);
    return hipGetLastError();
}

// Diagnostic: evaluates the device sine/cosine (UnitVec::Get) on n angles.
namespace prxd {
__global__ void k_sincos(const float* in, float* out, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        const V2 q = unit_from_angle(in[i]);
        out[2 * i] = q.x;
        out[2 * i + 1] = q.y;
    }
}
}  // namespace prxd

extern "C" int prx_selftest_sincos(const float* angles, float* out, int n) {
    if (n <= 0 || !angles || !out) return 1;
    float* d_in = nullptr;
    float* d_out = nullptr;
    if (hipMalloc(&d_in, sizeof(float) * std::size_t(n)) != hipSuccess) return 5;
    if (hipMalloc(&d_out, 2 * sizeof(float) * std::size_t(n)) != hipSuccess) {
        (void)hipFree(d_in);
        return 5;
    }
    hipError_t e = hipMemcpy(d_in, angles, sizeof(float) * std::size_t(n), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(prxd::k_sincos, dim3(unsigned((n + 255) / 256)), dim3(256), 0, 0, d_in, d_out, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, 2 * sizeof(float) * std::size_t(n), hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return e == hipSuccess ? 0 : 5;
}
