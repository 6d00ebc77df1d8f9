// Initial particle state on the device, in the manner of
// KernelFillGridWithParticles (particles/ParticlesInit.kernel:66-302): each
// supercell gets ppc particles per cell, placed round-robin over its 256
// cells (slot = k*256 + localCellIdx), at uniform random in-cell positions
// (startPosition/RandomImpl.hpp:46-56) with weighting density*V/ppc
// (WeightMacroParticles.hpp:30-42) and momentum mass*(sign*u_drift + u_th*N(0,1))
// (manipulators/unary/Drift.hpp:53-76, Temperature.hpp:60-84); the drift sign
// is +1, or set by the cell's y band (shear flow), or alternates with the
// particle's index in its cell (two counter-streaming beams).
#include "pic_kernels.hpp"
#include "pic_rng.hpp"

namespace pic {

__global__ __launch_bounds__(256) void kInit(InitArgs a) {
    int const sc = blockIdx.x;
    int const sx = sc % a.g.nscx;
    int const sy = (sc / a.g.nscx) % a.g.nscy;
    int const sz = sc / (a.g.nscx * a.g.nscy);
    size_t const base = (size_t)sc * (size_t)a.sp.cap;
    int const n = a.ppc * kSCV;
    for (int i = threadIdx.x; i < a.sp.cap; i += blockDim.x) {
        size_t const g = base + i;
        if (i >= n) {
            a.sp.mask[g] = 0;
            continue;
        }
        int const k = i / kSCV;
        int const lc = i % kSCV;
        int const gx = sx * kSCX + lc % kSCX;
        int const gy = sy * kSCY + (lc / kSCX) % kSCY;
        int const gz = sz * kSCZ + lc / (kSCX * kSCY);
        uint64_t const cell = (uint64_t)gx + (uint64_t)a.g.nx * ((uint64_t)gy + (uint64_t)a.g.ny * (uint64_t)gz);
        uint32_t const spc = (uint32_t)a.species;
        a.sp.pos[0][g] = uniformF(rngKey(a.seed, spc, cell, (uint32_t)k, 0u));
        a.sp.pos[1][g] = uniformF(rngKey(a.seed, spc, cell, (uint32_t)k, 1u));
        a.sp.pos[2][g] = uniformF(rngKey(a.seed, spc, cell, (uint32_t)k, 2u));
        double n0, n1, n2, n3;
        normalPair(a.seed, spc, cell, (uint32_t)k, 0u, &n0, &n1);
        normalPair(a.seed, spc, cell, (uint32_t)k, 1u, &n2, &n3);
        double sign = 1.0;
        if (a.drift_band == 1) sign = (gy >= a.g.ny / 4 && gy < 3 * a.g.ny / 4) ? 1.0 : -1.0;
        if (a.drift_band == 2) sign = (k & 1) ? -1.0 : 1.0;  // counter-streaming beams
        double const m = (double)a.mass;
        a.sp.mom[0][g] = (float)(m * (sign * a.u_drift[0] + a.u_th[0] * n0));
        a.sp.mom[1][g] = (float)(m * (sign * a.u_drift[1] + a.u_th[1] * n1));
        a.sp.mom[2][g] = (float)(m * (sign * a.u_drift[2] + a.u_th[2] * n2));
        (void)n3;
        a.sp.w[g] = a.weighting;
        a.sp.cell[g] = (uint16_t)lc;
        a.sp.mask[g] = 1;
    }
    if (threadIdx.x == 0) a.sp.count[sc] = (uint32_t)(n < a.sp.cap ? n : a.sp.cap);
}

__global__ __launch_bounds__(256) void kNoise(float* f, size_t n, uint64_t seed, int comp, float amp) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float const u = uniformF(rngKey(seed, 200u + (uint32_t)comp, (uint64_t)i, 0u, 0u));
    f[i] = f[i] + amp * (2.0f * u - 1.0f);
}

hipError_t launchInit(const InitArgs& a, hipStream_t s) {
    int const nsc = a.g.nscx * a.g.nscy * a.g.nscz;
    hipLaunchKernelGGL(kInit, dim3(nsc), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launchNoise(float* f, size_t n, uint64_t seed, int comp, float amp, hipStream_t s) {
    hipLaunchKernelGGL(kNoise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, n, seed, comp, amp);
    return hipGetLastError();
}

}  // namespace pic
