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
//   stage:   the grouped leavers' records go to the supercell's outbox and the
//            supercell is compacted on its own (stageSC): the k-th leaver slot
//            below count-L is filled with the k-th remaining particle at or
//            above it, ascending; count becomes count-L;
//   insert:  every supercell appends, for codes 1..26 in order, the outbox
//            records its neighbours send towards it behind its own particles.
// In a full step the push epilogue stages (its supercell's lines are still in
// L2), so the shift itself is the insert: coalesced reads of 32-byte records
// and coalesced appends.
#include <algorithm>

#include "pic_shift_common.hpp"

namespace pic {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Leaver lists from the multiMask of every slot (used when the push did not
// build them, e.g. after frames were written from the host).
__global__ __launch_bounds__(kThreads) void kCollect(SpeciesDev S, GridDev grid) {
    __shared__ uint32_t sWave[kWaves];
    __shared__ uint32_t sWaveCnt[kWaves * kNumDirs];
    __shared__ GroupScratch g;
    if (halted(grid)) return;
    int const sc = blockIdx.x;
    int const tid = threadIdx.x;
    int const n = (int)S.count[sc];
    size_t const base = slotBase(S, sc);
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

// Staging of every supercell from the leaver lists (when the push did not stage).
__global__ __launch_bounds__(kThreads) void kStage(SpeciesDev S, GridDev grid) {
    __shared__ uint32_t sWave[kWaves];
    if (halted(grid)) return;
    int const sc = blockIdx.x;
    stageSC<kWaves>(S, sc, (int)S.outOff[(size_t)sc * kOffStride + kNumDirs], sWave);
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

// insert for destination supercell d; called by one whole wave. A record
// whose slot lies beyond the capacity is left in the sender's outbox, the
// count still grows by every incoming record and the overflow flag is set;
// with redo the supercell's count is taken as its count before that insert
// (count - incoming) and the insert runs again into grown supercells.
__device__ __forceinline__ void insertSC(const SpeciesDev& S, const GridDev& g, int d, bool redo) {
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
        // leavers crossing an open border are absorbed (boundary/Absorbing.hpp:63-92)
        bool const absorbed = ((g.open & 1) && (sx < 0 || sx >= g.nscx)) || ((g.open & 2) && (sy < 0 || sy >= g.nscy)) ||
                              ((g.open & 4) && (sz < 0 || sz >= g.nscz));
        sx = sx < 0 ? sx + g.nscx : (sx >= g.nscx ? sx - g.nscx : sx);
        sy = sy < 0 ? sy + g.nscy : (sy >= g.nscy ? sy - g.nscy : sy);
        sz = sz < 0 ? sz + g.nscz : (sz >= g.nscz ? sz - g.nscz : sz);
        src = sx + g.nscx * (sy + g.nscy * sz);
        const uint32_t* off = S.outOff + (size_t)src * kOffStride;
        beg = off[code];
        cnt = absorbed ? 0u : off[code + 1] - beg;
    }
    uint32_t const incl = waveInclusiveScan(cnt);
    uint32_t const start = incl - cnt;  // exclusive prefix: first incoming index of this code
    uint32_t const total = __shfl(incl, 63);
    uint32_t const n = redo ? S.count[d] - total : S.count[d];
    size_t const dbase = slotBase(S, d);
    uint32_t const dcap = slotCap(S, d);
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
        if (slot >= dcap) {
            if (g.flags) atomicOr(&g.flags->overflow, 1u);
            continue;
        }
        size_t const rec = slotBase(S, sc) + bg + (k - st);
        float4 const r0 = S.outbox[2 * rec];
        float4 const r1 = S.outbox[2 * rec + 1];
        size_t const dst = dbase + slot;
        S.pos[0][dst] = r0.x;
        S.pos[1][dst] = r0.y;
        S.pos[2][dst] = r0.z;
        S.mom[0][dst] = r0.w;
        S.mom[1][dst] = r1.x;
        S.mom[2][dst] = r1.y;
        S.w[dst] = r1.z;
        S.cell[dst] = (uint16_t)__float_as_int(r1.w);
        S.mask[dst] = 1;
    }
    if (lane == 0) S.count[d] = n + total;
}

__global__ __launch_bounds__(kThreads) void kInsert(SpeciesDev S, GridDev g, int nsc, int redo) {
    if (!redo && halted(g)) return;
    int const d = blockIdx.x * kScPerBlock + (threadIdx.x >> 6);
    if (d >= nsc) return;  // whole wave
    insertSC(S, g, d, redo != 0);
}

__global__ __launch_bounds__(256) void kClampCounts(SpeciesDev S, int nsc) {
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsc && S.count[i] > slotCap(S, i)) S.count[i] = slotCap(S, i);
}

// Batched forms for ensembles: work items (table entry, supercell), one wave
// per item; the table holds one entry per (simulation, species).
__global__ __launch_bounds__(kThreads) void kInsertBatch(const ShiftEntry* __restrict__ table,
                                                        const int2* __restrict__ work, int nwork) {
    int const w = blockIdx.x * kScPerBlock + (threadIdx.x >> 6);
    if (w >= nwork) return;  // whole wave
    int2 const it = work[w];
    const ShiftEntry& e = table[it.x];
    if (halted(e.g)) return;
    insertSC(e.S, e.g, it.y, false);
}

hipError_t launchShiftBatch(const ShiftEntry* table, const int2* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    int const blocks = (nwork + kScPerBlock - 1) / kScPerBlock;
    hipLaunchKernelGGL(kInsertBatch, dim3(blocks), dim3(kThreads), 0, s, table, work, nwork);
    return hipGetLastError();
}

hipError_t launchCollect(const GridDev& g, const SpeciesDev& sp, hipStream_t s) {
    int const nsc = g.nscx * g.nscy * g.nscz;
    hipLaunchKernelGGL(kCollect, dim3(nsc), dim3(kThreads), 0, s, sp, g);
    return hipGetLastError();
}

hipError_t launchShift(const GridDev& g, const SpeciesDev& sp, bool staged, hipStream_t s) {
    int const nsc = g.nscx * g.nscy * g.nscz;
    if (!staged) {
        hipLaunchKernelGGL(kStage, dim3(nsc), dim3(kThreads), 0, s, sp, g);
        hipError_t const e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    int const blocks = (nsc + kScPerBlock - 1) / kScPerBlock;
    hipLaunchKernelGGL(kInsert, dim3(blocks), dim3(kThreads), 0, s, sp, g, nsc, 0);
    return hipGetLastError();
}

hipError_t launchInsertRedo(const GridDev& g, const SpeciesDev& sp, hipStream_t s) {
    int const nsc = g.nscx * g.nscy * g.nscz;
    int const blocks = (nsc + kScPerBlock - 1) / kScPerBlock;
    hipLaunchKernelGGL(kInsert, dim3(blocks), dim3(kThreads), 0, s, sp, g, nsc, 1);
    return hipGetLastError();
}

// Rows of oldWords 16-byte words from src (row pitch oldWords) to dst (row
// pitch newWords >= oldWords): the slots of every supercell keep their index
// when the supercells grow.
// One workgroup per listed run: `slots` slots from slot `from` to slot `to`
// (16-byte words: runs are whole frames of slots, so multiples of 16 bytes
// for every attribute size).
__global__ __launch_bounds__(256) void kMoveRuns(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                 const RunMove* __restrict__ moves, uint32_t bytesPerSlot) {
    RunMove const m = moves[blockIdx.x];
    size_t const words = (size_t)m.slots * bytesPerSlot / 16;
    const uint4* from = src + (size_t)m.from * bytesPerSlot / 16;
    uint4* to = dst + (size_t)m.to * bytesPerSlot / 16;
    for (size_t t = threadIdx.x; t < words; t += blockDim.x) to[t] = from[t];
}

hipError_t launchMoveRuns(void* dst, const void* src, uint32_t bytesPerSlot, const RunMove* moves, int n,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kMoveRuns, dim3((unsigned)n), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), moves, bytesPerSlot);
    return hipGetLastError();
}

hipError_t launchClampCounts(const SpeciesDev& sp, int nsc, hipStream_t s) {
    hipLaunchKernelGGL(kClampCounts, dim3((nsc + 255) / 256), dim3(256), 0, s, sp, nsc);
    return hipGetLastError();
}

}  // namespace pic
