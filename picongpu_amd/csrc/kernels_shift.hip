// Particle shift between supercells with periodic wrap.
//
// Replaces KernelShiftParticles / KernelFillGaps / KernelFillGapsLastFrame
// (pmacc/particles/ParticlesBase.kernel:61-630) and the periodic exchange of
// guard particles (ParticlesBase.tpp:61-100, boundary/Apply.hpp:74-116): on
// one periodic GPU a leaving particle is moved straight into the neighbour
// supercell, whose index wraps around the box. Three kernels, each one
// workgroup per supercell, with a canonical particle order that does not
// depend on scheduling:
//   collect: leaver slots of each supercell, grouped by exchange code, each
//            group ascending in slot order;
//   insert:  every supercell appends, for codes 1..26 in order, the leavers
//            its neighbours send towards it behind its own particles;
//   compact: the k-th gap (ascending slot) below the new count is filled with
//            the k-th valid particle (ascending slot) above it.
#include "pic_kernels.hpp"

namespace pic {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ unsigned long long laneMaskLt() {
    return (1ull << (threadIdx.x & 63)) - 1ull;
}

// Block-wide exclusive rank of `pred` among lanes of this chunk, in thread
// order; returns the rank and writes the chunk total into *total.
__device__ __forceinline__ int blockRank(bool pred, uint32_t* sWave, int* total) {
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    unsigned long long const b = __ballot(pred);
    if (lane == 0) sWave[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    int before = 0, sum = 0;
    for (int w = 0; w < kWaves; ++w) {
        int const c = (int)sWave[w];
        before += w < wave ? c : 0;
        sum += c;
    }
    __syncthreads();
    *total = sum;
    return before + __popcll(b & laneMaskLt());
}

__global__ __launch_bounds__(kThreads) void kCollect(SpeciesDev S) {
    __shared__ uint32_t sCnt[kNumDirs];
    __shared__ uint32_t sRun[kNumDirs];
    __shared__ uint32_t sWaveCnt[kWaves][kNumDirs];
    int const sc = blockIdx.x;
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const n = (int)S.count[sc];
    size_t const base = (size_t)sc * (size_t)S.cap;
    uint32_t* off = S.outOff + (size_t)sc * kOffStride;
    int32_t* out = S.outIdx + base;

    if (tid < kNumDirs) sCnt[tid] = 0u;
    __syncthreads();
    for (int i = tid; i < n; i += kThreads) {
        int const m = S.mask[base + i];
        if (m >= 2) atomicAdd(&sCnt[m - 1], 1u);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        off[0] = 0u;
        for (int k = 1; k < kNumDirs; ++k) {
            off[k] = acc;
            sRun[k] = acc;
            acc += sCnt[k];
        }
        off[kNumDirs] = acc;
        sRun[0] = acc;
    }
    __syncthreads();
    if (sRun[0] == 0u) return;  // no leavers in this supercell

    for (int c0 = 0; c0 < n; c0 += kThreads) {
        int const i = c0 + tid;
        int code = 0;
        if (i < n) {
            int const m = S.mask[base + i];
            code = m >= 2 ? m - 1 : 0;
        }
        if (tid < kNumDirs * kWaves) (&sWaveCnt[0][0])[tid] = 0u;
        __syncthreads();
        // rank among lanes of this wave with the same code (slot order)
        int myRank = 0;
        unsigned long long pending = __ballot(code != 0);
        while (pending) {
            int const leader = __ffsll((long long)pending) - 1;
            int const k = __shfl(code, leader);
            unsigned long long const same = __ballot(code == k);
            if (code == k) myRank = __popcll(same & laneMaskLt());
            if (lane == 0) sWaveCnt[wave][k] = (uint32_t)__popcll(same);
            pending &= ~same;
        }
        __syncthreads();
        if (code != 0) {
            uint32_t r = sRun[code] + (uint32_t)myRank;
            for (int w = 0; w < wave; ++w) r += sWaveCnt[w][code];
            out[r] = i;
        }
        __syncthreads();
        if (tid > 0 && tid < kNumDirs) {
            uint32_t add = 0;
            for (int w = 0; w < kWaves; ++w) add += sWaveCnt[w][tid];
            sRun[tid] += add;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void copyParticle(const SpeciesDev& S, size_t dst, size_t src) {
    S.pos[0][dst] = S.pos[0][src];
    S.pos[1][dst] = S.pos[1][src];
    S.pos[2][dst] = S.pos[2][src];
    S.mom[0][dst] = S.mom[0][src];
    S.mom[1][dst] = S.mom[1][src];
    S.mom[2][dst] = S.mom[2][src];
    S.w[dst] = S.w[src];
    S.cell[dst] = S.cell[src];
}

__global__ __launch_bounds__(kThreads) void kInsert(SpeciesDev S, GridDev g, uint32_t* errFlags) {
    int const d = blockIdx.x;
    int const tid = threadIdx.x;
    int const dx = d % g.nscx;
    int const dy = (d / g.nscx) % g.nscy;
    int const dz = d / (g.nscx * g.nscy);
    size_t const dbase = (size_t)d * (size_t)S.cap;
    uint32_t running = S.count[d];
    for (int code = 1; code < kNumDirs; ++code) {
        // exchange code -> step per axis (1 -> +1, 2 -> -1)
        int const ex = code % 3, ey = (code / 3) % 3, ez = code / 9;
        int const stx = ex == 2 ? -1 : ex, sty = ey == 2 ? -1 : ey, stz = ez == 2 ? -1 : ez;
        int sx = dx - stx, sy = dy - sty, sz = dz - stz;
        sx = sx < 0 ? sx + g.nscx : (sx >= g.nscx ? sx - g.nscx : sx);
        sy = sy < 0 ? sy + g.nscy : (sy >= g.nscy ? sy - g.nscy : sy);
        sz = sz < 0 ? sz + g.nscz : (sz >= g.nscz ? sz - g.nscz : sz);
        int const s = sx + g.nscx * (sy + g.nscy * sz);
        const uint32_t* off = S.outOff + (size_t)s * kOffStride;
        uint32_t const beg = off[code];
        uint32_t const end = off[code + 1 < kNumDirs ? code + 1 : kNumDirs];
        uint32_t const m = end - beg;
        if (m == 0u) continue;
        size_t const sbase = (size_t)s * (size_t)S.cap;
        for (uint32_t t = tid; t < m; t += kThreads) {
            uint32_t const slot = running + t;
            if (slot >= (uint32_t)S.cap) {
                atomicOr(errFlags, 1u);
                continue;
            }
            size_t const src = sbase + (size_t)S.outIdx[sbase + beg + t];
            size_t const dst = dbase + slot;
            copyParticle(S, dst, src);
            S.mask[dst] = 1;
        }
        running += m;
    }
    if (tid == 0) S.occ[d] = running < (uint32_t)S.cap ? running : (uint32_t)S.cap;
}

__global__ __launch_bounds__(kThreads) void kCompact(SpeciesDev S) {
    __shared__ uint32_t sWave[kWaves];
    int const d = blockIdx.x;
    int const tid = threadIdx.x;
    size_t const base = (size_t)d * (size_t)S.cap;
    uint32_t const n = S.count[d];
    uint32_t const tot = S.occ[d];
    uint32_t const L = S.outOff[(size_t)d * kOffStride + kNumDirs];
    if (L == 0u) {
        if (tid == 0) S.count[d] = tot;
        return;
    }
    uint32_t const nn = tot > L ? tot - L : 0u;  // tot < L only after a capacity error
    int32_t* gaps = S.scratch + base;
    int32_t* srcs = S.outIdx + base;  // outIdx of this supercell is no longer read
    (void)n;
    // gaps: slots < nn holding a leaver
    int run = 0;
    for (uint32_t c0 = 0; c0 < nn; c0 += kThreads) {
        uint32_t const i = c0 + tid;
        bool const pred = i < nn && S.mask[base + i] >= 2;
        int total;
        int const r = blockRank(pred, sWave, &total);
        if (pred) gaps[run + r] = (int32_t)i;
        run += total;
    }
    int const G = run;
    // sources: valid slots in [nn, tot)
    run = 0;
    for (uint32_t c0 = nn; c0 < tot; c0 += kThreads) {
        uint32_t const i = c0 + tid;
        bool const pred = i < tot && S.mask[base + i] == 1;
        int total;
        int const r = blockRank(pred, sWave, &total);
        if (pred) srcs[run + r] = (int32_t)i;
        run += total;
    }
    __syncthreads();
    int const M = G < run ? G : run;
    for (int k = tid; k < M; k += kThreads) {
        size_t const dst = base + (size_t)gaps[k];
        size_t const src = base + (size_t)srcs[k];
        copyParticle(S, dst, src);
        S.mask[dst] = 1;
    }
    __syncthreads();
    for (uint32_t i = nn + tid; i < tot; i += kThreads) S.mask[base + i] = 0;
    if (tid == 0) S.count[d] = nn;
}

hipError_t launchShift(const GridDev& g, const SpeciesDev& sp, uint32_t* errFlags, hipStream_t s) {
    int const nsc = g.nscx * g.nscy * g.nscz;
    hipLaunchKernelGGL(kCollect, dim3(nsc), dim3(kThreads), 0, s, sp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kInsert, dim3(nsc), dim3(kThreads), 0, s, sp, g, errFlags);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kCompact, dim3(nsc), dim3(kThreads), 0, s, sp);
    return hipGetLastError();
}

}  // namespace pic
