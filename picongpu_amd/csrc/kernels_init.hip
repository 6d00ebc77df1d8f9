// Initial particle state on the device, in the manner of
// KernelFillGridWithParticles (particles/ParticlesInit.kernel:66-302): each
// cell gets the macro-particle count of WeightMacroParticles
// (startPosition/detail/WeightMacroParticles.hpp:30-42: floor(min(ppc,
// N / MIN_WEIGHTING)) for N real particles in the cell, so that no weighting
// falls below MIN_WEIGHTING = 10; full density gives ppc), placed
// round-robin over the supercell's cells (round k: the cells with more than k
// particles, in cell order), at uniform random in-cell positions
// (startPosition/RandomImpl.hpp:46-56) with weighting N / count (density*V/ppc
// where the count is ppc) and momentum mass*(sign*u_drift + u_th*N(0,1))
// (manipulators/unary/Drift.hpp:53-76, Temperature.hpp:60-84); the drift sign
// is +1, or set by the cell's y band (shear flow), or alternates with the
// particle's index in its cell (two counter-streaming beams).
#include "pic_kernels.hpp"
#include "pic_rng.hpp"

namespace pic {

// density factor of the profile at the origin of global cell (gx, gy, gz)
__device__ __forceinline__ double densityFactor(const InitArgs& a, int gx, int gy, int gz) {
    if (a.profile == 0) return 1.0;
    double const x = (double)(a.axis == 0 ? gx : (a.axis == 1 ? gy : gz));
    double const f = (x - a.rampStart) / a.rampLength;
    return f <= 0.0 ? 0.0 : (f >= 1.0 ? 1.0 : f);
}

constexpr float kInitMinWeighting = 10.0f;  // particle.param MIN_WEIGHTING

// WeightMacroParticles for a cell of density factor f: the macro-particle count
// and (when below ppc) their weighting N / count, N = density*V*f in FP32
__device__ __forceinline__ int macroCount(const InitArgs& a, double f, float* wLow) {
    if (!(f > 0.0)) return 0;
    float const real = (float)(a.weighting * (double)a.ppc * f);
    float const maxPar = real / kInitMinWeighting;
    float const m = fminf((float)a.ppc, maxPar);
    int const n = (int)floorf(m);
    *wLow = n > 0 ? real / (float)n : 0.0f;
    return n;
}

__global__ __launch_bounds__(256) void kInit(InitArgs a) {
    __shared__ uint32_t sWave[4];
    __shared__ int16_t sCellOfRank[kSCV];
    __shared__ uint8_t sCount[kSCV];        // macro particles per cell (<= ppc <= 255)
    __shared__ uint32_t sRound[kMaxInitPpc + 1];  // first slot of each round
    int const sc = blockIdx.x;
    int const sx = sc % a.g.nscx;
    int const sy = (sc / a.g.nscx) % a.g.nscy;
    int const sz = sc / (a.g.nscx * a.g.nscy);
    size_t const base = slotBase(a.sp, sc);
    int const cap = (int)slotCap(a.sp, sc);
    // cells with macro particles, ranked in cell order (256 threads = 256 cells)
    int nActive;
    {
        int const lc = threadIdx.x;
        float wLow;
        int const cnt = macroCount(a, densityFactor(a, sx * kSCX + lc % kSCX, sy * kSCY + (lc / kSCX) % kSCY,
                                                    sz * kSCZ + lc / (kSCX * kSCY)),
                                   &wLow);
        sCount[lc] = (uint8_t)cnt;
        bool const active = cnt > 0;
        for (int k = threadIdx.x; k <= a.ppc; k += blockDim.x) sRound[k] = 0u;
        int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        unsigned long long const b = __ballot(active);
        if (lane == 0) sWave[wave] = (uint32_t)__popcll(b);
        __syncthreads();
        int before = 0;
        nActive = 0;
        for (int w = 0; w < 4; ++w) {
            before += w < wave ? (int)sWave[w] : 0;
            nActive += (int)sWave[w];
        }
        if (active) sCellOfRank[before + __popcll(b & ((1ull << lane) - 1ull))] = (int16_t)lc;
        // round k holds the cells with more than k particles
        for (int k = 0; k < cnt; ++k) atomicAdd(&sRound[k + 1], 1u);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 1; k <= a.ppc; ++k) sRound[k] += sRound[k - 1];
        __syncthreads();
    }
    // slot sRound[k] + j: the j-th cell, in rank order, of round k (with every
    // populated cell at ppc particles: slot = k * nActive + rank)
    int const n = (int)sRound[a.ppc];
    for (int i = threadIdx.x; i < cap; i += blockDim.x) {
        size_t const g = base + i;
        if (i >= n) {
            a.sp.mask[g] = 0;
            continue;
        }
        int k = 0;
        while ((int)sRound[k + 1] <= i) ++k;
        int j = i - (int)sRound[k];
        int lc;
        if ((int)(sRound[k + 1] - sRound[k]) == nActive) {
            lc = sCellOfRank[j];
        } else {
            int r = 0;
            for (;; ++r) {
                int const c = sCellOfRank[r];
                if (sCount[c] > k) {
                    if (j == 0) break;
                    --j;
                }
            }
            lc = sCellOfRank[r];
        }
        int const gx = sx * kSCX + lc % kSCX;
        int const gy = sy * kSCY + (lc / kSCX) % kSCY;
        int const gz = sz * kSCZ + lc / (kSCX * kSCY);
        uint64_t const cell = (uint64_t)gx + (uint64_t)a.g.nx * ((uint64_t)gy + (uint64_t)a.g.ny * (uint64_t)gz);
        uint32_t const spc = (uint32_t)a.species;
        if (a.quiet[0] > 0) {
            // QuietImpl (startPosition/QuietImpl.hpp:34-49): particle k of the
            // cell takes lattice point ppc-1-k, mapToND with x fastest
            int const idx = a.ppc - 1 - k;
            int const q[3] = {idx % a.quiet[0], (idx / a.quiet[0]) % a.quiet[1], idx / (a.quiet[0] * a.quiet[1])};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                float const spacing = 1.0f / (float)a.quiet[d];
                a.sp.pos[d][g] = (float)q[d] * spacing + spacing * 0.5f;
            }
        } else {
            a.sp.pos[0][g] = uniformF(rngKey(a.seed, spc, cell, (uint32_t)k, 0u));
            a.sp.pos[1][g] = uniformF(rngKey(a.seed, spc, cell, (uint32_t)k, 1u));
            a.sp.pos[2][g] = uniformF(rngKey(a.seed, spc, cell, (uint32_t)k, 2u));
        }
        double n0, n1, n2, n3;
        normalPair(a.seed, spc, cell, (uint32_t)k, 0u, &n0, &n1);
        normalPair(a.seed, spc, cell, (uint32_t)k, 1u, &n2, &n3);
        double sign = 1.0;
        if (a.drift_band == 1) sign = (gy >= a.g.ny / 4 && gy < 3 * a.g.ny / 4) ? 1.0 : -1.0;
        if (a.drift_band == 2) sign = (k & 1) ? -1.0 : 1.0;  // counter-streaming beams
        float wLow;
        double const f = densityFactor(a, gx, gy, gz);
        int const cnt = macroCount(a, f, &wLow);
        float const w = cnt == a.ppc ? (float)(a.weighting * f) : wLow;
        double const m = (double)(a.frameMass * w);
        a.sp.mom[0][g] = (float)(m * (sign * a.u_drift[0] + a.u_th[0] * n0));
        a.sp.mom[1][g] = (float)(m * (sign * a.u_drift[1] + a.u_th[1] * n1));
        a.sp.mom[2][g] = (float)(m * (sign * a.u_drift[2] + a.u_th[2] * n2));
        (void)n3;
        a.sp.w[g] = w;
        a.sp.cell[g] = (uint16_t)lc;
        a.sp.mask[g] = 1;
    }
    if (threadIdx.x == 0) a.sp.count[sc] = (uint32_t)(n < cap ? n : cap);
}

// Gaussian pulse added to E and B at the Yee positions (YeeCell.hpp:69-123)
__global__ __launch_bounds__(256) void kLaser(LaserArgs a) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int const idx[3] = {(int)(i % a.g.nx), (int)((i / a.g.nx) % a.g.ny), (int)(i / ((size_t)a.g.nx * a.g.ny))};
    int const P = a.prop, Q = a.pol, R = 3 - a.prop - a.pol;  // R: the magnetic component's axis
    double const sgnB = ((Q + 1) % 3 == P) ? -1.0 : 1.0;      // (k_hat x e_Q) along e_R
    for (int f = 0; f < 2; ++f) {
        int const comp = f == 0 ? Q : R;
        double x[3];
        for (int d = 0; d < 3; ++d) {
            // E_c sits at +1/2 along c; B_c at +1/2 along the two other axes
            double const off = f == 0 ? (d == comp ? 0.5 : 0.0) : (d == comp ? 0.0 : 0.5);
            x[d] = ((double)idx[d] + off) * a.cell[d] - a.center[d];
        }
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d)
            if (d != P) r2 += x[d] * x[d];
        double const xi = x[P];
        double const v = a.e0 * exp(-r2 / (a.w0 * a.w0)) * exp(-xi * xi / (2.0 * a.sigma * a.sigma)) * cos(a.k * xi);
        if (f == 0)
            a.E[comp][i] = a.E[comp][i] + (float)v;
        else
            a.B[comp][i] = a.B[comp][i] + (float)(sgnB * v * a.invc);
    }
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

hipError_t launchLaser(const LaserArgs& a, hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    hipLaunchKernelGGL(kLaser, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launchNoise(float* f, size_t n, uint64_t seed, int comp, float amp, hipStream_t s) {
    hipLaunchKernelGGL(kNoise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, n, seed, comp, amp);
    return hipGetLastError();
}

}  // namespace pic
