// Block-level helpers of the supercell shift, shared by the push kernel's
// epilogue and the shift kernels. All orders are fixed by slot index, never
// by scheduling.
#pragma once

#include "pic_kernels.hpp"

namespace pic {

__device__ __forceinline__ unsigned long long laneMaskLt() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// Exclusive rank of `pred` in thread order within one block-wide chunk;
// *total receives the chunk's count. NW = waves per block.
template <int NW>
__device__ __forceinline__ int blockRank(bool pred, uint32_t* sWave, int* total) {
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    unsigned long long const b = __ballot(pred);
    if (lane == 0) sWave[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    int before = 0, sum = 0;
    for (int w = 0; w < NW; ++w) {
        int const c = (int)sWave[w];
        before += w < wave ? c : 0;
        sum += c;
    }
    __syncthreads();
    *total = sum;
    return before + __popcll(b & laneMaskLt());
}

struct GroupScratch {
    uint32_t cnt[kNumDirs];
    uint32_t run[kNumDirs];
};

// Given the slot-ordered leaver list of supercell sc in S.scratch[base, base+L),
// write S.outIdx grouped by exchange code (ascending slot inside a code) and
// the offset table S.outOff[sc] (entry k = start of code k, [27] = L).
// NW waves; sWaveCnt is NW*kNumDirs words of LDS.
template <int NW>
__device__ void groupLeavers(const SpeciesDev& S, int sc, int L, GroupScratch& g, uint32_t* sWaveCnt) {
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    constexpr int NT = NW * 64;
    size_t const base = slotBase(S, sc);
    uint32_t* off = S.outOff + (size_t)sc * kOffStride;
    if (tid < kNumDirs) g.cnt[tid] = 0u;
    __syncthreads();
    for (int k = tid; k < L; k += NT) {
        int const slot = S.scratch[base + k];
        atomicAdd(&g.cnt[S.mask[base + slot] - 1], 1u);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        off[0] = 0u;
        for (int k = 1; k < kNumDirs; ++k) {
            off[k] = acc;
            g.run[k] = acc;
            acc += g.cnt[k];
        }
        off[kNumDirs] = acc;
    }
    __syncthreads();
    for (int c0 = 0; c0 < L; c0 += NT) {
        int const k = c0 + tid;
        int code = 0, slot = 0;
        if (k < L) {
            slot = S.scratch[base + k];
            code = S.mask[base + slot] - 1;
        }
        if (tid < kNumDirs * NW) sWaveCnt[tid] = 0u;
        __syncthreads();
        int myRank = 0;
        unsigned long long pending = __ballot(code != 0);
        while (pending) {
            int const leader = __ffsll((long long)pending) - 1;
            int const kc = __shfl(code, leader);
            unsigned long long const same = __ballot(code == kc);
            if (code == kc) myRank = __popcll(same & laneMaskLt());
            if (lane == 0) sWaveCnt[wave * kNumDirs + kc] = (uint32_t)__popcll(same);
            pending &= ~same;
        }
        __syncthreads();
        if (code != 0) {
            uint32_t r = g.run[code] + (uint32_t)myRank;
            for (int w = 0; w < wave; ++w) r += sWaveCnt[w * kNumDirs + code];
            S.outIdx[base + r] = slot;
        }
        __syncthreads();
        if (tid > 0 && tid < kNumDirs) {
            uint32_t add = 0;
            for (int w = 0; w < NW; ++w) add += sWaveCnt[w * kNumDirs + tid];
            g.run[tid] += add;
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

// The leavers of supercell sc leave it. With the slot-ordered leaver list in
// S.scratch and the grouped one in S.outIdx/S.outOff (L leavers):
//  1. the record of the k-th grouped leaver goes to outbox entry k (position,
//     momentum, weighting, localCellIdx: 32 bytes, two float4);
//  2. the supercell is compacted on its own: with nn = count - L, the r-th
//     leaver slot below nn (ascending) receives the r-th remaining particle at
//     or above nn (ascending); multiMask of [nn, count) is cleared and count
//     becomes nn.
// The insert of the neighbours' outbox records then appends at nn. Called by
// the whole workgroup (NW waves); sWave is NW words of LDS.
template <int NW>
__device__ void stageSC(const SpeciesDev& S, int sc, int L, uint32_t* sWave) {
    if (L == 0) return;  // uniform
    constexpr int NT = NW * 64;
    int const tid = threadIdx.x;
    size_t const base = slotBase(S, sc);
    uint32_t const n = S.count[sc];
    for (int k = tid; k < L; k += NT) {
        size_t const from = base + (size_t)S.outIdx[base + k];
        float4 const r0 = make_float4(S.pos[0][from], S.pos[1][from], S.pos[2][from], S.mom[0][from]);
        float4 const r1 = make_float4(S.mom[1][from], S.mom[2][from], S.w[from], __int_as_float((int)S.cell[from]));
        S.outbox[2 * (base + k)] = r0;
        S.outbox[2 * (base + k) + 1] = r1;
    }
    __syncthreads();  // every leaver record is out before its slot is overwritten
    uint32_t const nn = n - (uint32_t)L;
    const int32_t* leavers = S.scratch + base;
    int run = 0;
    for (uint32_t c0 = nn; c0 < n; c0 += NT) {
        uint32_t const i = c0 + tid;
        bool const src = i < n && S.mask[base + i] == 1;
        int total;
        int const r = blockRank<NW>(src, sWave, &total);
        if (src) {
            size_t const dst = base + (size_t)leavers[run + r];
            copyParticle(S, dst, base + i);
            S.mask[dst] = 1;
        }
        run += total;
    }
    __syncthreads();  // multiMask of [nn, n) has been read
    for (uint32_t i = nn + tid; i < n; i += NT) S.mask[base + i] = 0;
    if (tid == 0) S.count[sc] = nn;
}

}  // namespace pic
