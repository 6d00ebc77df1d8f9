// Fused field gather + particle push + moveParticle + Esirkepov deposit.
//
// One workgroup per supercell (the reference's KernelMoveAndMarkParticles,
// particles/Particles.kernel:169-273, and KernelComputeCurrent,
// fields/FieldJ.kernel:51-107). The E/B tile of the supercell plus the
// interpolation margins is staged in LDS; particles of all species in the
// launch are streamed as coalesced SoA slot reads; each particle's current
// is accumulated into an int64 fixed-point J tile in LDS (order independent)
// that is flushed once to HBM. Launches run over a checkerboard of
// supercells whose J tiles do not overlap, so the flush is a plain
// read-add-write in a fixed phase order.
#include <type_traits>

#include "pic_device.hpp"
#include "pic_kernels.hpp"
#include "pic_shift_common.hpp"

namespace pic {

__device__ __forceinline__ int wrapIdx(int v, int n) {
    v = v < 0 ? v + n : v;
    return v >= n ? v - n : v;
}

constexpr int kPushThreads = kPushThreadsPerBlock;
constexpr int kPushWaves = kPushThreads / 64;

// LDS bytes of one supercell's work: the E/B tile (or the epilogue's
// scratch), then the int64 J tile
// A particle whose move changes its shifted cell on some axis needs the
// general Esirkepov form (up to (SUPP+1)^2*SUPP cells per component); the rest
// need SUPP^2*(SUPP-1). The push deposits the latter at once and queues the
// former per wave: lane l holds queued item l in A (items 64..127 in B), and
// a full A is deposited by all 64 lanes together. An item carries what the
// deposit reads: the tile cell, the position after the move, the move in
// cells (vel * dt / cell) and the charge, so the contribution is the one the
// particle would have made in place.
struct SlowItem {
    int cell;  // tile cell x | y << 8 | z << 16
    float x, y, z, dx, dy, dz, q;
};

__device__ __forceinline__ int permuteLane(int dstLane, int v) { return __builtin_amdgcn_ds_permute(dstLane << 2, v); }
__device__ __forceinline__ float permuteLane(int dstLane, float v) {
    return __int_as_float(__builtin_amdgcn_ds_permute(dstLane << 2, __float_as_int(v)));
}

// Append the items of the lanes with `slow` set after the P queued ones
// (wave-uniform P). Every lane of the wave calls it: the ds_permute moves
// slow lanes to queue positions P.. and the other lanes to the free ones, a
// permutation of the 64 lanes.
__device__ __forceinline__ void queueSlow(bool slow, const SlowItem& mine, SlowItem& A, SlowItem& B, int& P) {
    unsigned long long const m = __ballot(slow);
    int const k = __popcll(m);
    if (k == 0) return;
    int const lane = threadIdx.x & 63;
    int const below = __popcll(m & laneMaskLt());
    int const dst = (P + (slow ? below : k + (lane - below))) & 63;
    int const off = (lane - P) & 63;
    bool const toA = off < k && lane >= P;
    bool const toB = off < k && lane < P;
#define PIC_QUEUE_MOVE(f)                          \
    {                                              \
        auto const r = permuteLane(dst, mine.f);   \
        if (toA) A.f = r;                          \
        if (toB) B.f = r;                          \
    }
    PIC_QUEUE_MOVE(cell)
    PIC_QUEUE_MOVE(x)
    PIC_QUEUE_MOVE(y)
    PIC_QUEUE_MOVE(z)
    PIC_QUEUE_MOVE(dx)
    PIC_QUEUE_MOVE(dy)
    PIC_QUEUE_MOVE(dz)
    PIC_QUEUE_MOVE(q)
#undef PIC_QUEUE_MOVE
    P += k;
}

template <int SUPP, class TG>
__device__ __forceinline__ void depositQueued(unsigned long long* __restrict__ sJ, const SlowItem& t,
                                              const DepositConsts& dk) {
    float const dp[3] = {t.dx, t.dy, t.dz};
    Esirkepov<SUPP>::template depositGeometry<TG, true>(sJ, t.cell & 255, (t.cell >> 8) & 255, t.cell >> 16,
                                                        Esirkepov<SUPP>::geometryDp(Vec3{t.x, t.y, t.z}, dp), t.q, dk);
}

template <int SUPP, int MODE>
struct PushLds {
    static constexpr int GLO = SUPP / 2;
    static constexpr int GUP = (SUPP + 1) / 2;
    static constexpr int GV = (SCX + GLO + GUP) * (SCY + GLO + GUP) * (SCZ + GLO + GUP);
    using TG = JTileGeom<Esirkepov<SUPP>::lowerMargin, Esirkepov<SUPP>::upperMargin>;
    static constexpr int EB_ELEMS = MODE == 2 ? 1 : 6 * GV;
    static constexpr int J_ELEMS = MODE == 0 ? 1 : TG::total();
    static constexpr int EB_BYTES = ((EB_ELEMS * 4 + 15) / 16) * 16;
    static constexpr int EPI_BYTES =
        (kPushMaxIters * kPushWaves + 1) * 4 + kPushWaves * kNumDirs * 4 + (int)sizeof(GroupScratch);
    // 16-byte multiple: the int64 J tile that follows needs 8-byte aligned LDS addresses
    static constexpr int LO_BYTES = (((EB_BYTES > EPI_BYTES ? EB_BYTES : EPI_BYTES) + 15) / 16) * 16;
    static constexpr int BYTES = LO_BYTES + J_ELEMS * 8;
};

// The work of one supercell `sc` of the simulation described by `a`; every
// thread of the workgroup calls it.
template <int SUPP, int PUSHER, int MODE, bool EMZ = false, bool DAMP = false>
__device__ __forceinline__ void pushDepositSC(const PushArgs& a, int sc, unsigned char* smem) {
    constexpr int GLO = SUPP / 2;          // FieldToParticleInterpolation.hpp:49
    constexpr int GUP = (SUPP + 1) / 2;    // FieldToParticleInterpolation.hpp:50
    constexpr int GX = SCX + GLO + GUP, GY = SCY + GLO + GUP, GZ = SCZ + GLO + GUP;
    constexpr int GV = GX * GY * GZ;
    // The J tile is the supercell plus the reference's deposit margins. In
    // the fused form the deposit runs in the frame of the particle's new
    // cell, which may lie one cell outside the supercell, but the cells it
    // touches still lie within those margins of the particle's old cell: its
    // path starts inside the old cell and moves less than one cell per axis
    // (the host requires float(c*dt)/cell < 1).
    constexpr int DLO = Esirkepov<SUPP>::lowerMargin;
    using TG = JTileGeom<DLO, Esirkepov<SUPP>::upperMargin>;
    constexpr int EB_ELEMS = MODE == 2 ? 1 : 6 * GV;
    constexpr int J_ELEMS = MODE == 0 ? 1 : TG::total();

    // One LDS buffer: the E/B tile, then the J tile. After the particle loop
    // the E/B tile is dead and its space holds the epilogue's scratch.
    constexpr int LO_BYTES = PushLds<SUPP, MODE>::LO_BYTES;
    static_assert(LO_BYTES % 16 == 0, "J tile must be 16-byte aligned");
    static_assert(EB_ELEMS == PushLds<SUPP, MODE>::EB_ELEMS && J_ELEMS == PushLds<SUPP, MODE>::J_ELEMS, "LDS plan");
    float* const sEB = reinterpret_cast<float*>(smem);
    unsigned long long* const sJ = reinterpret_cast<unsigned long long*>(smem + LO_BYTES);
    uint32_t* const sPre = reinterpret_cast<uint32_t*>(smem);
    uint32_t* const sWaveCnt = sPre + kPushMaxIters * kPushWaves + 1;
    GroupScratch& sGroup = *reinterpret_cast<GroupScratch*>(sWaveCnt + kPushWaves * kNumDirs);

    if (halted(a.g)) return;  // uniform: a later step of an overflowed simulation
    int const sx = sc % a.g.nscx;
    int const sy = (sc / a.g.nscx) % a.g.nscy;
    int const sz = sc / (a.g.nscx * a.g.nscy);
    int const ox = sx * SCX, oy = sy * SCY, oz = sz * SCZ;
    int const tid = threadIdx.x;

    int total = 0;
    for (int s = 0; s < a.nsp; ++s) total += (int)a.sp[s].count[sc];
    if (total == 0) {  // uniform across the workgroup
        if (MODE != 2 && tid < kOffStride) {
            for (int s = 0; s < a.nsp; ++s)
                if (a.sp[s].lists) a.sp[s].outOff[(size_t)sc * kOffStride + tid] = 0u;
        }
        if (MODE == 1 && a.tileOut)  // an empty tile for the gather flush
            for (int t = tid; t < J_ELEMS; t += blockDim.x) a.tileOut[(size_t)sc * J_ELEMS + t] = 0.0f;
        return;
    }

    // the open-border checks are compiled only into the loop a box with open
    // borders runs (the branch is uniform)
    auto stageEB = [&](auto openBorders) {
        for (int t = tid; t < 6 * GV; t += blockDim.x) {
            int const c = t / GV;
            int const r = t - c * GV;
            int const tx = r % GX;
            int const ty = (r / GX) % GY;
            int const tz = r / (GX * GY);
            int const gx = wrapIdx(ox - GLO + tx, a.g.nx);
            int const gy = wrapIdx(oy - GLO + ty, a.g.ny);
            int const gz = wrapIdx(oz - GLO + tz, a.g.nz);
            size_t const gi = (size_t)gx + (size_t)a.g.nx * ((size_t)gy + (size_t)a.g.ny * gz);
            const float* src = c < 3 ? a.E[c] : a.B[c - 3];
            // fields beyond an open border read zero
            bool const out = decltype(openBorders)::value && outsideOpen(a.g, ox - GLO + tx, oy - GLO + ty, oz - GLO + tz);
            sEB[t] = out ? 0.0f : src[gi];
        }
    };
    if (MODE != 2) {
        if (a.g.open)
            stageEB(std::true_type{});
        else
            stageEB(std::false_type{});
    }
    if (MODE != 0) {
        for (int t = tid; t < J_ELEMS; t += blockDim.x) sJ[t] = 0ull;
    }
    __syncthreads();

    PushConsts pk;
    pk.cell[0] = a.cell[0];
    pk.cell[1] = a.cell[1];
    pk.cell[2] = a.cell[2];
    pk.dt = a.dt;
    pk.c2 = a.c2;
    pk.rc2 = a.rc2;
    pk.c = a.c;
    DepositConsts dk;
    dk.cell[0] = a.cell[0];
    dk.cell[1] = a.cell[1];
    dk.cell[2] = a.cell[2];
    dk.dt = a.dt;
    dk.inv_vdt = a.inv_vdt;
    dk.jscale = a.jscale;

    // whether a particle of this supercell can end in a damping area (a move
    // is less than a cell per axis): the others skip the exchange loop
    bool dampSC = false;
    if (DAMP && a.damp.axes) {
        int const org[3] = {ox, oy, oz}, ext[3] = {SCX, SCY, SCZ};
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if ((a.damp.axes >> d) & 1)
                dampSC = dampSC || org[d] - 1 < a.damp.lo[d] || org[d] + ext[d] >= a.damp.hi[d];
    }

    SlowItem qA{}, qB{};
    int queued = 0;  // wave-uniform
    for (int s = 0; s < a.nsp; ++s) {
        const SpeciesDev& S = a.sp[s];
        int const n = (int)S.count[sc];
        size_t const base = slotBase(S, sc);
        uint32_t leaveBits = 0u;  // bit `it`: the particle of iteration `it` left the supercell
        int const iters = (n + (int)blockDim.x - 1) / (int)blockDim.x;  // uniform across the workgroup
        // the last species runs one iteration more, without particles, that drains the queue
        bool const lastSpecies = s == a.nsp - 1;
        for (int it = 0; it < iters + (MODE == 1 && !EMZ && lastSpecies ? 1 : 0); ++it) {
            int const i = it * (int)blockDim.x + tid;
            bool slow = false;
            SlowItem mine{};
            if (i < n) {
                size_t const gidx = base + i;
                Vec3 const pos{S.pos[0][gidx], S.pos[1][gidx], S.pos[2][gidx]};
                Vec3 mom{S.mom[0][gidx], S.mom[1][gidx], S.mom[2][gidx]};
                float w = S.w[gidx];
                int const lidx = S.cell[gidx];
                int const lx = lidx % SCX, ly = (lidx / SCX) % SCY, lz = lidx / (SCX * SCY);
                float const mass = S.mass * w;     // attribute::getMass (GetMass.hpp:40-44)
                float charge = S.charge * w;       // attribute::getCharge (GetCharge.hpp:45-66)
                if (MODE != 2) {
                    Vec3 const e = gatherE<SUPP, GX, GY, GV>(sEB, lx + GLO, ly + GLO, lz + GLO, pos);
                    Vec3 const b = gatherB<SUPP, GX, GY, GV>(sEB + 3 * GV, lx + GLO, ly + GLO, lz + GLO, pos);
                    Vec3 vel;
                    if (PUSHER == PIC_PUSHER_VAY)
                        pushVay(e, b, mom, vel, mass, charge, pk);
                    else if (PUSHER == PIC_PUSHER_HIGUERA_CARY)
                        pushHigueraCary(e, b, mom, vel, mass, charge, pk);
                    else
                        pushBoris(e, b, mom, vel, mass, charge, pk);
                    float p[3] = {pos.x, pos.y, pos.z};
                    // the move in cells; Esirkepov forms the same quotients (Esirkepov.hpp:82)
                    float dp[3] = {(vel.x * pk.dt) / pk.cell[0], (vel.y * pk.dt) / pk.cell[1],
                                   (vel.z * pk.dt) / pk.cell[2]};
                    p[0] += dp[0];
                    p[1] += dp[1];
                    p[2] += dp[2];
                    int dir[3], newIdx, newMask;
                    moveParticle(p, lidx, dir, newIdx, newMask);
                    if (S.reflect) {  // the species' reflecting borders (boundary::apply after the push)
                        int cellNow[3] = {ox + lx + dir[0], oy + ly + dir[1], oz + lz + dir[2]};
                        int const nGrid[3] = {a.g.nx, a.g.ny, a.g.nz};
                        if (reflectParticle(nGrid, S.reflect, cellNow, p, mom, vel, dir, dp))
                            cellAfterMove(lidx, dir, newIdx, newMask);
                    }
                    if (DAMP && dampSC) {
                        // the absorbing boundary's PML damping, exchange by exchange
                        // (+x, -x, +y, -y, +z, -z); the deposit stage then takes the
                        // velocity and charge of the damped particle. One loop, not
                        // unrolled, with the axis' values selected (no indexed local
                        // arrays, which the compiler would move to LDS)
                        int const c0 = ox + lx + dir[0], c1 = oy + ly + dir[1], c2 = oz + lz + dir[2];
                        bool any = false;
                        if (!(a.g.open && outsideOpen(a.g, c0, c1, c2))) {
#pragma unroll 1
                            for (int e = 0; e < 6; ++e) {
                                int const d = e >> 1;
                                if (!((a.damp.axes >> d) & 1) || ((S.reflect >> d) & 1)) continue;
                                int const cn = d == 0 ? c0 : (d == 1 ? c1 : c2);
                                bool const inArea = (e & 1) == 0 ? cn >= a.damp.hi[d] : cn < a.damp.lo[d];
                                if (!inArea) continue;
                                float const pd = d == 0 ? p[0] : (d == 1 ? p[1] : p[2]);
                                float const cs = d == 0 ? pk.cell[0] : (d == 1 ? pk.cell[1] : pk.cell[2]);
                                pmlDamp(a.damp, d, cn, pd, S.mass, pk.rc2, pk.dt, cs, mom, w);
                                any = true;
                            }
                        }
                        if (any) {
                            S.w[gidx] = w;
                            charge = S.charge * w;
                            vel = velocityOf(mom, S.mass * w, pk.rc2);
                            dp[0] = (vel.x * pk.dt) / pk.cell[0];
                            dp[1] = (vel.y * pk.dt) / pk.cell[1];
                            dp[2] = (vel.z * pk.dt) / pk.cell[2];
                        }
                    }
                    S.pos[0][gidx] = p[0];
                    S.pos[1][gidx] = p[1];
                    S.pos[2][gidx] = p[2];
                    S.mom[0][gidx] = mom.x;
                    S.mom[1][gidx] = mom.y;
                    S.mom[2][gidx] = mom.z;
                    if ((dir[0] | dir[1] | dir[2]) != 0) S.cell[gidx] = (uint16_t)newIdx;
                    if (newMask >= 2) S.mask[gidx] = (uint8_t)newMask;
                    leaveBits |= (newMask >= 2 ? 1u : 0u) << (i / kPushThreads);
                    // a particle leaving through an open border is absorbed by the
                    // boundary before the reference's deposit stage sees it
                    if (MODE == 1 && !(a.g.open && outsideOpen(a.g, ox + lx + dir[0], oy + ly + dir[1], oz + lz + dir[2]))) {
                        int const cx = lx + dir[0] + DLO, cy = ly + dir[1] + DLO, cz = lz + dir[2] + DLO;
                        Vec3 const np{p[0], p[1], p[2]};
                        if (EMZ) {
                            EmZ<SUPP>::template depositDp<TG>(sJ, cx, cy, cz, np, dp, charge, dk);
                        } else {
                            auto const G = Esirkepov<SUPP>::geometryDp(np, dp);
                            if (Esirkepov<SUPP>::changesCell(G)) {
                                slow = true;
                                mine = SlowItem{cx | (cy << 8) | (cz << 16), np.x, np.y, np.z, dp[0], dp[1], dp[2], charge};
                            } else {
                                Esirkepov<SUPP>::template depositStay<TG>(sJ, cx, cy, cz, G, charge, dk);
                            }
                        }
                    }
                } else {
                    Vec3 const vel = velocityOf(mom, mass, a.rc2);
                    if (EMZ)
                        EmZ<SUPP>::template deposit<TG>(sJ, lx + DLO, ly + DLO, lz + DLO, pos, vel, charge, dk);
                    else
                        Esirkepov<SUPP>::template deposit<TG>(sJ, lx + DLO, ly + DLO, lz + DLO, pos, vel, charge,
                                                                      dk);
                }
            }
            if (MODE == 1 && !EMZ) {
                queueSlow(slow, mine, qA, qB, queued);
                if (queued >= 64 || (it == iters && queued > 0)) {
                    if ((int)(tid & 63) < queued) depositQueued<SUPP, TG>(sJ, qA, dk);
                    qA = qB;
                    queued = queued > 64 ? queued - 64 : 0;
                }
            }
        }
        if (MODE != 2 && S.lists) S.leaveBits[(size_t)sc * kPushThreads + tid] = leaveBits;
    }
    __syncthreads();

    auto flushJ = [&](auto openBorders) {
        for (int t = tid; t < J_ELEMS; t += blockDim.x) {
            long long const acc = fixedTileValue(sJ[t]);
            if (acc == 0) continue;
            int const c = t < TG::off(1) ? 0 : (t < TG::off(2) ? 1 : 2);
            int const r = t - (c == 0 ? 0 : TG::off(c));
            int const ex = c == 0 ? TG::ext(0, 0) : (c == 1 ? TG::ext(1, 0) : TG::ext(2, 0));
            int const ey = c == 0 ? TG::ext(0, 1) : (c == 1 ? TG::ext(1, 1) : TG::ext(2, 1));
            int const tx = r % ex;
            int const ty = (r / ex) % ey;
            int const tz = r / (ex * ey);
            if (decltype(openBorders)::value && outsideOpen(a.g, ox - DLO + tx, oy - DLO + ty, oz - DLO + tz))
                continue;  // dropped beyond the border
            int const gx = wrapIdx(ox - DLO + tx, a.g.nx);
            int const gy = wrapIdx(oy - DLO + ty, a.g.ny);
            int const gz = wrapIdx(oz - DLO + tz, a.g.nz);
            size_t const gi = (size_t)gx + (size_t)a.g.nx * ((size_t)gy + (size_t)a.g.ny * gz);
            float const v = (float)((double)acc * a.jinv);
            a.J[c][gi] += v;
        }
    };
    if (MODE == 1 && a.tileOut) {
        // one launch for every supercell: the tile is left for the gather
        // flush as the floats the flush would add, zero where it adds
        // nothing (adding a zero would change nothing either: J never holds
        // -0, its sums start at +0 and round to nearest)
        for (int t = tid; t < J_ELEMS; t += blockDim.x) {
            long long const acc = fixedTileValue(sJ[t]);
            a.tileOut[(size_t)sc * J_ELEMS + t] = acc == 0 ? 0.0f : (float)((double)acc * a.jinv);
        }
    } else if (MODE != 0) {
        if (a.g.open)
            flushJ(std::true_type{});
        else
            flushJ(std::false_type{});
    }

    // Epilogue: slot-ordered leaver list from the ballots, then grouped by
    // exchange code (the input of the supercell shift); in a full step also
    // the staging of the shift (outbox records, compaction of the supercell).
    if (MODE != 2) {
        for (int s = 0; s < a.nsp; ++s) {
            const SpeciesDev& S = a.sp[s];
            if (!S.lists) continue;
            int const n = (int)S.count[sc];
            int const iters = (n + kPushThreads - 1) / kPushThreads;
            int const wave = tid >> 6;
            uint32_t const bits = S.leaveBits[(size_t)sc * kPushThreads + tid];
            __syncthreads();
            // leavers per (iteration, wave) entry, entries in slot order
            for (int it = 0; it < iters; ++it) {
                unsigned long long const m = __ballot((bits >> it) & 1u);
                if ((tid & 63) == 0) sPre[it * kPushWaves + wave] = (uint32_t)__popcll(m);
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t acc = 0;
                for (int e = 0; e < iters * kPushWaves; ++e) {
                    uint32_t const c = sPre[e];
                    sPre[e] = acc;
                    acc += c;
                }
                sPre[kPushMaxIters * kPushWaves] = acc;
            }
            __syncthreads();
            int const L = (int)sPre[kPushMaxIters * kPushWaves];
            size_t const base = slotBase(S, sc);
            for (int it = 0; it < iters; ++it) {
                bool const left = (bits >> it) & 1u;
                unsigned long long const m = __ballot(left);
                if (left)
                    S.scratch[base + sPre[it * kPushWaves + wave] + (uint32_t)__popcll(m & laneMaskLt())] =
                        it * kPushThreads + tid;
            }
            __syncthreads();
            groupLeavers<kPushWaves>(S, sc, L, sGroup, sWaveCnt);
            if (a.stageShift) stageSC<kPushWaves>(S, sc, L, sWaveCnt);
        }
    }
}

// Waves per SIMD the register budget must allow: PQS at most 128 VGPRs, so
// two workgroups share a CU (their LDS fits); TSC at most 80, so three do
// (3 x 53.6 KB of LDS fit the CU's 160 KB): the compiler spills a few values
// and the launch is faster, the LDS bank conflicts of the gather and the
// deposit being hidden better by the extra waves
template <int SUPP, int PUSHER, int MODE, bool EMZ = false>
__global__ __launch_bounds__(kPushThreads, SUPP == 4 ? 4 : (SUPP == 3 ? 6 : 1)) void kPushDeposit(PushArgs a) {
    __shared__ __align__(16) unsigned char smem[PushLds<SUPP, MODE>::BYTES];
    pushDepositSC<SUPP, PUSHER, MODE, EMZ>(a, a.scList[blockIdx.x], smem);
}

// the same with the PML's weight damping of absorbing species (a kernel of
// its own: the FP64 pow and exp of the damping would raise the register
// pressure of every launch)
template <int SUPP, int PUSHER, int MODE, bool EMZ = false>
__global__ __launch_bounds__(kPushThreads, SUPP == 4 ? 4 : (SUPP == 3 ? 6 : 1)) void kPushDepositPml(PushArgs a) {
    __shared__ __align__(16) unsigned char smem[PushLds<SUPP, MODE>::BYTES];
    pushDepositSC<SUPP, PUSHER, MODE, EMZ, true>(a, a.scList[blockIdx.x], smem);
}

// Batched form for ensembles: work items (simulation, supercell) of one
// checkerboard phase of every simulation in one launch; a workgroup takes
// items blockIdx.x, blockIdx.x + gridDim.x, ... so a grid smaller than the
// item count runs persistently.
template <int SUPP, int PUSHER, int MODE, bool EMZ = false>
__global__ __launch_bounds__(kPushThreads) void kPushDepositBatch(const PushArgs* __restrict__ table,
                                                                  const int2* __restrict__ work, int nwork) {
    __shared__ __align__(16) unsigned char smem[PushLds<SUPP, MODE>::BYTES];
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        int2 const it = work[w];
        pushDepositSC<SUPP, PUSHER, MODE>(table[it.x], it.y, smem);
        __syncthreads();  // the next item reuses the LDS
    }
}

hipError_t launchPushBatch(int supp, int pusher, const PushArgs* table, const int2* work, int nwork, int nblocks,
                           hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    if (nblocks <= 0 || nblocks > nwork) nblocks = nwork;
    switch (pusher * 8 + supp) {
#define PIC_BATCH_CASE(P, S)                                                                                        \
    case P * 8 + S:                                                                                                 \
        hipLaunchKernelGGL((kPushDepositBatch<S, P, 1>), dim3(nblocks), dim3(kPushThreads), 0, s, table, work, nwork); \
        break;
        PIC_BATCH_CASE(PIC_PUSHER_BORIS, 2)
        PIC_BATCH_CASE(PIC_PUSHER_BORIS, 3)
        PIC_BATCH_CASE(PIC_PUSHER_BORIS, 4)
        PIC_BATCH_CASE(PIC_PUSHER_VAY, 2)
        PIC_BATCH_CASE(PIC_PUSHER_VAY, 3)
        PIC_BATCH_CASE(PIC_PUSHER_VAY, 4)
        PIC_BATCH_CASE(PIC_PUSHER_HIGUERA_CARY, 2)
        PIC_BATCH_CASE(PIC_PUSHER_HIGUERA_CARY, 3)
        PIC_BATCH_CASE(PIC_PUSHER_HIGUERA_CARY, 4)
#undef PIC_BATCH_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SUPP, int PUSHER, int MODE, bool EMZ>
static hipError_t launchT(const PushArgs& a, int nblocks, hipStream_t s) {
    if (MODE != 2 && a.damp.axes)
        hipLaunchKernelGGL((kPushDepositPml<SUPP, PUSHER, MODE, EMZ>), dim3(nblocks), dim3(kPushThreads), 0, s, a);
    else
        hipLaunchKernelGGL((kPushDeposit<SUPP, PUSHER, MODE, EMZ>), dim3(nblocks), dim3(kPushThreads), 0, s, a);
    return hipGetLastError();
}

template <int SUPP, bool EMZ>
static hipError_t launchS(int pusher, int mode, const PushArgs& a, int nblocks, hipStream_t s) {
    if (mode == 2) return launchT<SUPP, PIC_PUSHER_BORIS, 2, EMZ>(a, nblocks, s);
    if (mode == 0) {  // no deposit: the current solver does not matter
        if (pusher == PIC_PUSHER_VAY) return launchT<SUPP, PIC_PUSHER_VAY, 0, false>(a, nblocks, s);
        if (pusher == PIC_PUSHER_HIGUERA_CARY) return launchT<SUPP, PIC_PUSHER_HIGUERA_CARY, 0, false>(a, nblocks, s);
        return launchT<SUPP, PIC_PUSHER_BORIS, 0, false>(a, nblocks, s);
    }
    if (pusher == PIC_PUSHER_VAY) return launchT<SUPP, PIC_PUSHER_VAY, 1, EMZ>(a, nblocks, s);
    if (pusher == PIC_PUSHER_HIGUERA_CARY) return launchT<SUPP, PIC_PUSHER_HIGUERA_CARY, 1, EMZ>(a, nblocks, s);
    return launchT<SUPP, PIC_PUSHER_BORIS, 1, EMZ>(a, nblocks, s);
}

size_t pushTileElems(int supp) {
    switch (supp) {
        case 2: return PushLds<2, 1>::J_ELEMS;
        case 3: return PushLds<3, 1>::J_ELEMS;
        case 4: return PushLds<4, 1>::J_ELEMS;
        default: return 0;
    }
}

// The supercells whose J tile of component c covers global cell g along axis
// a (tile origin at the supercell's origin minus DLO, extent ext): the
// candidates' supercell index along the axis (wrapped on a periodic axis)
// and the cell's tile coordinate, ordered by the axis' checkerboard colour
// (colour[s]). Returns their number (at most 3).
template <int SUPP>
__device__ __forceinline__ int coveringAlong(int g, int scSize, int ext, int nsc, bool open,
                                             const uint8_t* __restrict__ colour, int* scOut, int* tOut) {
    constexpr int DLO = Esirkepov<SUPP>::lowerMargin;
    int n = 0;
    int col[3];
    int const hi = (g + DLO) >= 0 ? (g + DLO) / scSize : -((-(g + DLO) + scSize - 1) / scSize);  // floor
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int const s = hi - k;
        int const t = g + DLO - s * scSize;
        if (t < 0 || t >= ext) continue;
        int sw = s;
        if (open) {
            if (s < 0 || s >= nsc) continue;
        } else {
            sw = ((s % nsc) + nsc) % nsc;
        }
        int const cs = colour[sw];
        int m = n++;
        for (; m > 0 && col[m - 1] > cs; --m) {
            col[m] = col[m - 1];
            scOut[m] = scOut[m - 1];
            tOut[m] = tOut[m - 1];
        }
        col[m] = cs;
        scOut[m] = sw;
        tOut[m] = t;
    }
    return n;
}

// One thread per J cell: for each component the covering supercells' tile
// values, added as the flush adds them, in ascending checkerboard phase. A
// phase's index is colour_x + ncol_x (colour_y + ncol_y colour_z), so
// visiting z, then y, then x candidates each in ascending colour is
// ascending phase (every supercell covering a cell lies in another phase).
template <int SUPP>
__device__ __forceinline__ void gatherFlushCell(const GridDev& g, const float* __restrict__ tiles,
                                                const uint8_t* __restrict__ colour, float* J0, float* J1, float* J2,
                                                size_t i) {
    using TG = JTileGeom<Esirkepov<SUPP>::lowerMargin, Esirkepov<SUPP>::upperMargin>;
    constexpr int J_ELEMS = PushLds<SUPP, 1>::J_ELEMS;
    int const x = (int)(i % g.nx);
    int const y = (int)((i / g.nx) % g.ny);
    int const z = (int)(i / ((size_t)g.nx * g.ny));
    float* const Jc[3] = {J0, J1, J2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int sx[3], tx[3], sy[3], ty[3], sz[3], tz[3];
        int const nx = coveringAlong<SUPP>(x, SCX, TG::ext(c, 0), g.nscx, (g.open & 1) != 0, colour, sx, tx);
        int const ny = coveringAlong<SUPP>(y, SCY, TG::ext(c, 1), g.nscy, (g.open & 2) != 0, colour + g.nscx, sy, ty);
        int const nz = coveringAlong<SUPP>(z, SCZ, TG::ext(c, 2), g.nscz, (g.open & 4) != 0,
                                           colour + g.nscx + g.nscy, sz, tz);
        float v = Jc[c][i];
        bool any = false;
        for (int a = 0; a < nz; ++a)
            for (int b = 0; b < ny; ++b)
                for (int d = 0; d < nx; ++d) {
                    int const sc = sx[d] + g.nscx * (sy[b] + g.nscy * sz[a]);
                    float const w =
                        tiles[(size_t)sc * J_ELEMS + TG::off(c) + (tz[a] * TG::ext(c, 1) + ty[b]) * TG::ext(c, 0) + tx[d]];
                    if (w == 0.0f) continue;
                    v += w;
                    any = true;
                }
        if (any) Jc[c][i] = v;
    }
}

template <int SUPP>
__global__ __launch_bounds__(256) void kGatherFlush(GridDev g, const float* __restrict__ tiles,
                                                     const uint8_t* __restrict__ colour, float* J0, float* J1,
                                                     float* J2) {
    if (halted(g)) return;
    size_t const n = (size_t)g.nx * g.ny * g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gatherFlushCell<SUPP>(g, tiles, colour, J0, J1, J2, i);
}

template <int SUPP>
__global__ __launch_bounds__(256) void kGatherFlushBatch(const GatherEntry* __restrict__ table,
                                                          const int2* __restrict__ blocks) {
    int2 const it = blocks[blockIdx.x];
    const GatherEntry& e = table[it.x];
    if (halted(e.g)) return;
    size_t const n = (size_t)e.g.nx * e.g.ny * e.g.nz;
    size_t const i = (size_t)it.y * blockDim.x + threadIdx.x;
    if (i < n) gatherFlushCell<SUPP>(e.g, e.tiles, e.colour, e.J[0], e.J[1], e.J[2], i);
}

hipError_t launchGatherFlushBatch(int supp, const GatherEntry* table, const int2* blocks, int nblocks,
                                  hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    switch (supp) {
        case 2: hipLaunchKernelGGL(kGatherFlushBatch<2>, dim3(nblocks), dim3(256), 0, s, table, blocks); break;
        case 3: hipLaunchKernelGGL(kGatherFlushBatch<3>, dim3(nblocks), dim3(256), 0, s, table, blocks); break;
        case 4: hipLaunchKernelGGL(kGatherFlushBatch<4>, dim3(nblocks), dim3(256), 0, s, table, blocks); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchGatherFlush(int supp, const GridDev& g, const float* tiles, const uint8_t* colour,
                             float* const J[3], hipStream_t s) {
    size_t const n = (size_t)g.nx * g.ny * g.nz;
    dim3 const grid((unsigned)((n + 255) / 256));
    switch (supp) {
        case 2: hipLaunchKernelGGL(kGatherFlush<2>, grid, dim3(256), 0, s, g, tiles, colour, J[0], J[1], J[2]); break;
        case 3: hipLaunchKernelGGL(kGatherFlush<3>, grid, dim3(256), 0, s, g, tiles, colour, J[0], J[1], J[2]); break;
        case 4: hipLaunchKernelGGL(kGatherFlush<4>, grid, dim3(256), 0, s, g, tiles, colour, J[0], J[1], J[2]); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchPush(int supp, int pusher, int mode, int currentSolver, const PushArgs& a, int nblocks,
                      hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    bool const emz = currentSolver == PIC_CURRENT_SOLVER_EMZ;
    switch (supp) {
        case 2: return emz ? launchS<2, true>(pusher, mode, a, nblocks, s) : launchS<2, false>(pusher, mode, a, nblocks, s);
        case 3: return emz ? launchS<3, true>(pusher, mode, a, nblocks, s) : launchS<3, false>(pusher, mode, a, nblocks, s);
        case 4: return emz ? launchS<4, true>(pusher, mode, a, nblocks, s) : launchS<4, false>(pusher, mode, a, nblocks, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pic
