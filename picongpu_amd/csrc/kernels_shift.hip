// Particle shift between supercells with periodic wrap.
//
// Replaces KernelShiftParticles / KernelFillGaps / KernelFillGapsLastFrame
// (pmacc/particles/ParticlesBase.kernel:61-630) and the periodic exchange of
// guard particles (ParticlesBase.tpp:61-100, boundary/Apply.hpp:74-116): on
// one periodic GPU a leaving particle is moved straight into the neighbour
// supercell, whose index wraps around the box. Canonical order, independent
// of scheduling:
//   leaver lists (built by the push kernel's epilogue, or by kCollect):
//            slot-ordered, and grouped by exchange code ascending in slot;
//   insert:  every supercell appends, for codes 1..26 in order, the leavers
//            its neighbours send towards it behind its own particles;
//   compact: the k-th gap (ascending slot) below the new count is filled with
//            the k-th valid particle (ascending slot) above it.
#include "pic_shift_common.hpp"

namespace pic {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Leaver lists from the multiMask of every slot (used when the push did not
// build them, e.g. after frames were written from the host).
__global__ __launch_bounds__(kThreads) void kCollect(SpeciesDev S) {
    __shared__ uint32_t sWave[kWaves];
    __shared__ uint32_t sWaveCnt[kWaves * kNumDirs];
    __shared__ GroupScratch g;
    int const sc = blockIdx.x;
    int const tid = threadIdx.x;
    int const n = (int)S.count[sc];
    size_t const base = (size_t)sc * (size_t)S.cap;
    int run = 0;
    for (int c0 = 0; c0 < n; c0 += kThreads) {
        int const i = c0 + tid;
        bool const pred = i < n && S.mask[base + i] >= 2;
        int total;
        int const r = blockRank<kWaves>(pred, sWave, &total);
        if (pred) S.scratch[base + run + r] = i;
        run += total;
    }
    __syncthreads();
    groupLeavers<kWaves>(S, sc, run, g, sWaveCnt);
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

// One wave per destination supercell, four supercells per workgroup.
constexpr int kScPerBlock = kThreads / 64;

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t v) {
    int const lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t const t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// insert for destination supercell d; called by one whole wave
__device__ __forceinline__ void insertSC(const SpeciesDev& S, const GridDev& g, int d, uint32_t* errFlags) {
    int const lane = threadIdx.x & 63;
    int const dx = d % g.nscx;
    int const dy = (d / g.nscx) % g.nscy;
    int const dz = d / (g.nscx * g.nscy);
    // lane `code` (1..26) looks up what its sender pushes towards d
    uint32_t cnt = 0, beg = 0;
    int src = 0;
    if (lane >= 1 && lane < kNumDirs) {
        int const code = lane;
        int const ex = code % 3, ey = (code / 3) % 3, ez = code / 9;
        int const stx = ex == 2 ? -1 : ex, sty = ey == 2 ? -1 : ey, stz = ez == 2 ? -1 : ez;
        int sx = dx - stx, sy = dy - sty, sz = dz - stz;
        sx = sx < 0 ? sx + g.nscx : (sx >= g.nscx ? sx - g.nscx : sx);
        sy = sy < 0 ? sy + g.nscy : (sy >= g.nscy ? sy - g.nscy : sy);
        sz = sz < 0 ? sz + g.nscz : (sz >= g.nscz ? sz - g.nscz : sz);
        src = sx + g.nscx * (sy + g.nscy * sz);
        const uint32_t* off = S.outOff + (size_t)src * kOffStride;
        beg = off[code];
        cnt = off[code + 1] - beg;
    }
    uint32_t const incl = waveInclusiveScan(cnt);
    uint32_t const start = incl - cnt;  // exclusive prefix: first incoming index of this code
    uint32_t const total = __shfl(incl, 63);
    uint32_t const n = S.count[d];
    size_t const dbase = (size_t)d * (size_t)S.cap;
    // Wave-uniform loop: every lane runs the shuffles of every round (a
    // shuffle reads nothing valid from lanes that have left a divergent
    // loop); only the copy is predicated on k < total.
    for (uint32_t k0 = 0; k0 < total; k0 += 64) {
        uint32_t const k = k0 + lane;
        bool const valid = k < total;
        // largest code with start(code) <= k (codes 1..26 sit in lanes 1..26)
        int lo = 1, hi = kNumDirs - 1;
        for (int step = 0; step < 5; ++step) {  // ceil(log2(26)) rounds, uniform count
            int const mid = (lo + hi + 1) >> 1;
            uint32_t const sm = __shfl(start, mid);
            if (lo < hi) {
                if (sm <= k)
                    lo = mid;
                else
                    hi = mid - 1;
            }
        }
        int const code = lo;
        uint32_t const st = __shfl(start, code);
        uint32_t const bg = __shfl(beg, code);
        int const sc = __shfl(src, code);
        if (!valid) continue;
        uint32_t const slot = n + k;
        if (slot >= (uint32_t)S.cap) {
            atomicOr(errFlags, 1u);
            continue;
        }
        size_t const sbase = (size_t)sc * (size_t)S.cap;
        size_t const from = sbase + (size_t)S.outIdx[sbase + bg + (k - st)];
        size_t const dst = dbase + slot;
        copyParticle(S, dst, from);
        S.mask[dst] = 1;
    }
    if (lane == 0) {
        uint32_t const occ = n + total;
        S.occ[d] = occ < (uint32_t)S.cap ? occ : (uint32_t)S.cap;
    }
}

// compaction of supercell d; called by one whole wave
__device__ __forceinline__ void compactSC(const SpeciesDev& S, int d) {
    int const lane = threadIdx.x & 63;
    size_t const base = (size_t)d * (size_t)S.cap;
    uint32_t const n = S.count[d];
    uint32_t const tot = S.occ[d];
    uint32_t const L = S.outOff[(size_t)d * kOffStride + kNumDirs];
    if (L == 0u) {
        if (lane == 0) S.count[d] = tot;
        return;
    }
    uint32_t const nn = tot > L ? tot - L : 0u;  // tot < L only after a capacity error
    const int32_t* gaps = S.scratch + base;  // slot-ordered leavers, written by an earlier kernel
    // G: leavers below the new count (the gaps)
    uint32_t G = 0;
    for (uint32_t c0 = 0; c0 < L; c0 += 64) {
        uint32_t const k = c0 + lane;
        G += (uint32_t)__popcll(__ballot(k < L && (uint32_t)gaps[k] < nn));
    }
    // sources in order: own non-leavers in [nn, n) ascending, then the
    // incoming particles from max(nn, n); source r fills gap r
    uint32_t a = 0;
    for (uint32_t c0 = nn; c0 < n; c0 += 64) {
        uint32_t const i = c0 + lane;
        bool const pred = i < n && S.mask[base + i] == 1;
        unsigned long long const b = __ballot(pred);
        if (pred) {
            uint32_t const r = a + (uint32_t)__popcll(b & laneMaskLt());
            if (r < G) {
                size_t const dst = base + (size_t)gaps[r];
                copyParticle(S, dst, base + i);
                S.mask[dst] = 1;
            }
        }
        a += (uint32_t)__popcll(b);
    }
    uint32_t const inStart = nn > n ? nn : n;
    for (uint32_t k = a + lane; k < G; k += 64) {
        uint32_t const srcSlot = inStart + (k - a);
        if (srcSlot >= tot) continue;  // only after a capacity error
        size_t const dst = base + (size_t)gaps[k];
        copyParticle(S, dst, base + srcSlot);
        S.mask[dst] = 1;
    }
    // gaps are below nn; the vacated tail [nn, tot) is cleared
    for (uint32_t i = nn + lane; i < tot; i += 64) S.mask[base + i] = 0;
    if (lane == 0) S.count[d] = nn;
}

__global__ __launch_bounds__(kThreads) void kInsert(SpeciesDev S, GridDev g, int nsc, uint32_t* errFlags) {
    int const d = blockIdx.x * kScPerBlock + (threadIdx.x >> 6);
    if (d >= nsc) return;  // whole wave
    insertSC(S, g, d, errFlags);
}

__global__ __launch_bounds__(kThreads) void kCompact(SpeciesDev S, int nsc) {
    int const d = blockIdx.x * kScPerBlock + (threadIdx.x >> 6);
    if (d >= nsc) return;  // whole wave
    compactSC(S, d);
}

// Batched forms for ensembles: work items (table entry, supercell), one wave
// per item; the table holds one entry per (simulation, species).
__global__ __launch_bounds__(kThreads) void kInsertBatch(const ShiftEntry* __restrict__ table,
                                                        const int2* __restrict__ work, int nwork) {
    int const w = blockIdx.x * kScPerBlock + (threadIdx.x >> 6);
    if (w >= nwork) return;  // whole wave
    int2 const it = work[w];
    const ShiftEntry& e = table[it.x];
    insertSC(e.S, e.g, it.y, e.errFlags);
}

__global__ __launch_bounds__(kThreads) void kCompactBatch(const ShiftEntry* __restrict__ table,
                                                         const int2* __restrict__ work, int nwork) {
    int const w = blockIdx.x * kScPerBlock + (threadIdx.x >> 6);
    if (w >= nwork) return;  // whole wave
    int2 const it = work[w];
    compactSC(table[it.x].S, it.y);
}

hipError_t launchShiftBatch(const ShiftEntry* table, const int2* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    int const blocks = (nwork + kScPerBlock - 1) / kScPerBlock;
    hipLaunchKernelGGL(kInsertBatch, dim3(blocks), dim3(kThreads), 0, s, table, work, nwork);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kCompactBatch, dim3(blocks), dim3(kThreads), 0, s, table, work, nwork);
    return hipGetLastError();
}

hipError_t launchCollect(const GridDev& g, const SpeciesDev& sp, hipStream_t s) {
    int const nsc = g.nscx * g.nscy * g.nscz;
    hipLaunchKernelGGL(kCollect, dim3(nsc), dim3(kThreads), 0, s, sp);
    return hipGetLastError();
}

hipError_t launchShift(const GridDev& g, const SpeciesDev& sp, uint32_t* errFlags, hipStream_t s) {
    int const nsc = g.nscx * g.nscy * g.nscz;
    int const blocks = (nsc + kScPerBlock - 1) / kScPerBlock;
    hipLaunchKernelGGL(kInsert, dim3(blocks), dim3(kThreads), 0, s, sp, g, nsc, errFlags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kCompact, dim3(blocks), dim3(kThreads), 0, s, sp, nsc);
    return hipGetLastError();
}

}  // namespace pic
