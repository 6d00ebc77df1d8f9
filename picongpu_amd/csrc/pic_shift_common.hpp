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
    size_t const base = (size_t)sc * (size_t)S.cap;
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

}  // namespace pic
