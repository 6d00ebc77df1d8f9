// Differentiable form of the PIC step and its hand-written adjoint.
//
// Forward ("differentiable mode"): one mobile CIC species with the Boris
// pusher and the Yee solver, particles kept in identity order with global
// cell indices (no supercell shift), the same FP32 per-particle arithmetic as
// kPushDeposit (gather, Boris, moveParticle, Esirkepov), current accumulated
// as int64 fixed point in global memory (order independent), then the FDTD
// kernels of kernels_field.hip.
//
// Adjoint (FP64): reverse of one step given the adjoints of E, B at step n+1
// and of every particle's position X = cell + pos and momentum p:
//   fields:  transposed curls, (curl_forward)^T = curl_backward and vice versa;
//   current: lambda_J = -(dt/eps0) lambda_E;
//   particles: reverse Esirkepov written as a gather of lambda_J through the
//            derivatives of the CIC weights, reverse velocity, reverse Boris,
//            reverse field gather (scatter of lambda_E, lambda_B with FP64
//            atomics, derivative of the interpolation weights).
#include "pic_device.hpp"
#include "pic_kernels.hpp"

#include <rocprim/device/device_radix_sort.hpp>

namespace pic {

__device__ __forceinline__ int wrapI(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

__device__ __forceinline__ size_t gidx(const GridDev& g, int x, int y, int z) {
    return (size_t)wrapI(x, g.nx) + (size_t)g.nx * ((size_t)wrapI(y, g.ny) + (size_t)g.ny * (size_t)wrapI(z, g.nz));
}

// Signed periodic difference in [-n/2, n/2).
__device__ __forceinline__ int wrapDiff(int d, int n) {
    d %= n;
    if (d >= (n + 1) / 2) d -= n;
    if (d < -(n / 2)) d += n;
    return d;
}

// Window of the grid around one block's particles: accumulation goes to LDS
// when the window holds every node the block's particles touch, and the
// window is added to global memory once at the end of the block.
struct BlockWindow {
    int ref[3];  // reference cell (the block's first particle)
    int lo[3];   // window origin relative to ref
    int ext[3];  // window extent in cells
    int vol;
    bool lds;
};

// Every thread of the block calls this (it synchronises). c: the particle's
// cell, active: whether the thread has a particle; nodes c-LO .. c+HI per
// axis must fit. rel receives c relative to the reference cell.
template <int LO, int HI>
__device__ __forceinline__ BlockWindow blockWindow(const GridDev& g, bool active, const int* c, int* rel, int* sBox,
                                                   int capCells) {
    int const n[3] = {g.nx, g.ny, g.nz};
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sBox[d] = c[d];
            sBox[3 + d] = 0x7fffffff;
            sBox[6 + d] = -0x7fffffff;
        }
    }
    __syncthreads();
    BlockWindow w;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w.ref[d] = sBox[d];
        rel[d] = active ? wrapDiff(c[d] - w.ref[d], n[d]) : 0;
    }
    if (active) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            atomicMin(&sBox[3 + d], rel[d]);
            atomicMax(&sBox[6 + d], rel[d]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w.lo[d] = sBox[3 + d] - LO;
        w.ext[d] = sBox[6 + d] + HI - w.lo[d] + 1;
    }
    // a box larger than the LDS capacity is trimmed, one cell at a time from
    // alternating ends of its longest axis; particles whose nodes then leave
    // the window go to global memory (windowHolds)
    while (w.ext[0] * w.ext[1] * w.ext[2] > capCells) {
        if (w.ext[0] >= w.ext[1] && w.ext[0] >= w.ext[2]) {
            w.ext[0] -= 1;
            w.lo[0] += w.ext[0] & 1;
        } else if (w.ext[1] >= w.ext[2]) {
            w.ext[1] -= 1;
            w.lo[1] += w.ext[1] & 1;
        } else {
            w.ext[2] -= 1;
            w.lo[2] += w.ext[2] & 1;
        }
    }
    w.vol = w.ext[0] * w.ext[1] * w.ext[2];
    w.lds = true;
    __syncthreads();  // sBox is reused by the next call
    return w;
}

// whether the nodes rel-LO .. rel+HI of a particle lie in the window
template <int LO, int HI>
__device__ __forceinline__ bool windowHolds(const BlockWindow& w, const int* rel) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) in = in && rel[d] - LO >= w.lo[d] && rel[d] + HI < w.lo[d] + w.ext[d];
    return in;
}

// global linear index of window element e (x fastest within the window)
__device__ __forceinline__ uint32_t windowToGlobal(const GridDev& g, const BlockWindow& w, int e) {
    int const lx = e % w.ext[0];
    int const ly = (e / w.ext[0]) % w.ext[1];
    int const lz = e / (w.ext[0] * w.ext[1]);
    return (uint32_t)wrapI(w.ref[0] + w.lo[0] + lx, g.nx) +
           (uint32_t)g.nx * ((uint32_t)wrapI(w.ref[1] + w.lo[1] + ly, g.ny) +
                             (uint32_t)g.ny * (uint32_t)wrapI(w.ref[2] + w.lo[2] + lz, g.nz));
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// FieldToParticleInterpolation<CIC> on global memory: same arithmetic as gather1<2>.
__device__ __forceinline__ float gatherGlobalCIC(const float* __restrict__ F, const GridDev& g, int cx, int cy, int cz,
                                                 Vec3 pos, float fx, float fy, float fz) {
    float p[3] = {pos.x, pos.y, pos.z};
    float const f[3] = {fx, fy, fz};
    int is[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float const v_pos = p[d] - f[d] - 0.5f;
        is[d] = v_pos >= -0.5f ? 0 : -1;
        p[d] = v_pos - (float)is[d] + 0.5f;
    }
    float sx[2], sy[2], sz[2];
    Shape<2>::onSupport(p[0], sx);
    Shape<2>::onSupport(p[1], sy);
    Shape<2>::onSupport(p[2], sz);
    // partial sums start at their first term and add the others as
    // multiply-adds, as gather1
    float result_z = 0.0f;
#pragma unroll
    for (int z = 0; z < 2; ++z) {
        float result_y = 0.0f;
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            float const result_x = mad(F[gidx(g, cx + is[0] + 1, cy + is[1] + y, cz + is[2] + z)], sx[1],
                                       F[gidx(g, cx + is[0], cy + is[1] + y, cz + is[2] + z)] * sx[0]);
            result_y = y == 0 ? result_x * sy[0] : mad(result_x, sy[y], result_y);
        }
        result_z = z == 0 ? result_y * sz[0] : mad(result_y, sz[z], result_z);
    }
    return result_z;
}



// Wrapped offsets of the nodes b, b+1, b+2 along one axis, times the axis stride.
__device__ __forceinline__ void axisOffsets(int b, int n, uint32_t stride, uint32_t* o) {
#pragma unroll
    for (int m = 0; m < 3; ++m) o[m] = (uint32_t)wrapI(b + m, n) * stride;
}

// One rotated pass of Esirkepov<CIC> (cptCurrent1D) into global int64
// accumulators: the float arithmetic of Esirkepov<2>::cpt1D; the rotated
// axes i, j, k are the real axes AI, AJ, AK, so every node offset is a
// compile-time pick from the per-axis offset tables.
template <int AI, int AJ, int AK, bool LDS>
__device__ __forceinline__ void cptGlobal(unsigned long long* __restrict__ J, const uint32_t (&off)[3][3],
                                          const float* p0, const float* p1, const int* st, float csd, double jscale) {
    constexpr int SUPP = 2, begin = 0, end = 2;
    float const q0[3] = {p0[AI], p0[AJ], p0[AK]};
    float const q1[3] = {p1[AI], p1[AJ], p1[AK]};
    int const s3[3] = {st[AI], st[AJ], st[AK]};
    if (q0[2] == q1[2]) return;
    float s0j[SUPP + 1], s1j[SUPP + 1], s0k[SUPP + 1], s1k[SUPP + 1];
    cachedShape<SUPP>(q0[1], !(s3[1] & 2), s0j);
    cachedShape<SUPP>(q1[1], !(s3[1] & 4), s1j);
    cachedShape<SUPP>(q0[2], !(s3[2] & 2), s0k);
    cachedShape<SUPP>(q1[2], !(s3[2] & 4), s1k);
    int const leaveI = s3[0] & 1, leaveJ = s3[1] & 1, leaveK = s3[2] & 1;
#pragma unroll
    for (int i = begin; i < end + 1; ++i) {
        if (i < end + leaveI) {
            float const s0i = Shape<SUPP>::assign((float)i - q0[0]);
            float const dsi = Shape<SUPP>::assign((float)i - q1[0]) - s0i;
#pragma unroll
            for (int j = begin; j < end + 1; ++j) {
                if (j < end + leaveJ) {
                    float const sj0 = s0j[j - begin];
                    float const dsj = s1j[j - begin] - sj0;
                    float const tmp = -csd * esirkepovWeight(s0i, dsi, sj0, dsj);
                    float acc = 0.0f;
#pragma unroll
                    for (int k = begin; k < end; ++k) {
                        if (k < end + leaveK - 1) {
                            float const dsk = s1k[k - begin] - s0k[k - begin];
                            acc = k == begin ? dsk * tmp : mad(dsk, tmp, acc);
                            unsigned long long* const at = J + (off[AI][i] + off[AJ][j] + off[AK][k]);
                            if (LDS)
                                __hip_atomic_fetch_add(at, toFixed(acc, jscale), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            else
                                __hip_atomic_fetch_add(at, toFixed(acc, jscale), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                        }
                    }
                }
            }
        }
    }
}

// Esirkepov<CIC> deposit of one particle (Esirkepov.hpp:61-242, the float
// arithmetic of Esirkepov<2>::deposit) into int64 accumulators: the block's
// LDS window (LDS = true; rel is the cell relative to the window reference) or
// global memory. The particle sits in global cell c at in-cell position pos
// after its move.
template <bool LDS>
__device__ __forceinline__ void depositCIC(const DiffArgs& a, unsigned long long* sJ, const BlockWindow& w,
                                           const int* c, const int* rel, Vec3 pos, Vec3 vel, float charge) {
    float const np[3] = {pos.x, pos.y, pos.z};
    float const dp[3] = {vel.x * a.dt / a.cell[0], vel.y * a.dt / a.cell[1], vel.z * a.dt / a.cell[2]};
    float p0[3], p1[3];
    int gs[3], st[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        p0[d] = np[d] - dp[d];
        p1[d] = np[d];
        int const iS = (int)floorf(p0[d]);
        int const iE = (int)floorf(p1[d]);
        gs[d] = iS < iE ? iS : iE;
        st[d] = (gs[d] == iS ? 2 : 0) | (gs[d] == iE ? 4 : 0) | (iS != iE ? 1 : 0);
        p0[d] -= (float)gs[d];
        p1[d] -= (float)gs[d];
    }
    uint32_t off[3][3];
    unsigned long long* J[3];
    if (LDS) {
        uint32_t const ls[3] = {1u, (uint32_t)w.ext[0], (uint32_t)(w.ext[0] * w.ext[1])};
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < 3; ++m) off[d][m] = (uint32_t)(rel[d] + gs[d] + m - w.lo[d]) * ls[d];
#pragma unroll
        for (int d = 0; d < 3; ++d) J[d] = sJ + d * w.vol;
    } else {
        axisOffsets(c[0] + gs[0], a.g.nx, 1u, off[0]);
        axisOffsets(c[1] + gs[1], a.g.ny, (uint32_t)a.g.nx, off[1]);
        axisOffsets(c[2] + gs[2], a.g.nz, (uint32_t)a.g.nx * (uint32_t)a.g.ny, off[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) J[d] = a.Jfix[d];
    }
    // rotations of Esirkepov.hpp:106-124: J_x (i,j,k) = (y,z,x), J_y (z,x,y), J_z (x,y,z)
    cptGlobal<1, 2, 0, LDS>(J[0], off, p0, p1, st, charge * a.inv_vdt * a.cell[0], a.jscale);
    cptGlobal<2, 0, 1, LDS>(J[1], off, p0, p1, st, charge * a.inv_vdt * a.cell[1], a.jscale);
    cptGlobal<0, 1, 2, LDS>(J[2], off, p0, p1, st, charge * a.inv_vdt * a.cell[2], a.jscale);
}

// int64 LDS window of the forward deposit: 3 components of this many cells
constexpr int kDiffLdsCells = 1024;

__global__ __launch_bounds__(256) void kDiffPush(DiffArgs a, DiffParticles P) {
    __shared__ unsigned long long sJ[3 * kDiffLdsCells];
    __shared__ int sBox[9];
    int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool const active = i < P.n;
    int64_t const n = P.n;
    int nc[3] = {0, 0, 0};
    float p[3] = {0.0f, 0.0f, 0.0f};
    Vec3 vel{0.0f, 0.0f, 0.0f};
    float charge = 0.0f;
    if (active) {
        int const cx = P.cell[i], cy = P.cell[n + i], cz = P.cell[2 * n + i];
        Vec3 const pos{P.pos[i], P.pos[n + i], P.pos[2 * n + i]};
        Vec3 mom{P.mom[i], P.mom[n + i], P.mom[2 * n + i]};
        float const w = P.w[i];
        float const mass = a.mass * w;
        charge = a.charge * w;
        Vec3 const e{gatherGlobalCIC(a.E[0], a.g, cx, cy, cz, pos, 0.5f, 0.0f, 0.0f),
                     gatherGlobalCIC(a.E[1], a.g, cx, cy, cz, pos, 0.0f, 0.5f, 0.0f),
                     gatherGlobalCIC(a.E[2], a.g, cx, cy, cz, pos, 0.0f, 0.0f, 0.5f)};
        Vec3 const b{gatherGlobalCIC(a.B[0], a.g, cx, cy, cz, pos, 0.0f, 0.5f, 0.5f),
                     gatherGlobalCIC(a.B[1], a.g, cx, cy, cz, pos, 0.5f, 0.0f, 0.5f),
                     gatherGlobalCIC(a.B[2], a.g, cx, cy, cz, pos, 0.5f, 0.5f, 0.0f)};
        PushConsts pk;
        for (int d = 0; d < 3; ++d) pk.cell[d] = a.cell[d];
        pk.dt = a.dt;
        pk.c2 = a.c2;
        pk.rc2 = a.rc2;
        pk.c = a.c;
        pushBoris(e, b, mom, vel, mass, charge, pk);
        p[0] = pos.x;
        p[1] = pos.y;
        p[2] = pos.z;
        p[0] += (vel.x * a.dt) / a.cell[0];
        p[1] += (vel.y * a.dt) / a.cell[1];
        p[2] += (vel.z * a.dt) / a.cell[2];
        int dir[3], idxUnused, maskUnused;
        moveParticle(p, 0, dir, idxUnused, maskUnused);
        nc[0] = wrapI(cx + dir[0], a.g.nx);
        nc[1] = wrapI(cy + dir[1], a.g.ny);
        nc[2] = wrapI(cz + dir[2], a.g.nz);
        P.cellOut[i] = nc[0];
        P.cellOut[n + i] = nc[1];
        P.cellOut[2 * n + i] = nc[2];
        P.posOut[i] = p[0];
        P.posOut[n + i] = p[1];
        P.posOut[2 * n + i] = p[2];
        P.momOut[i] = mom.x;
        P.momOut[n + i] = mom.y;
        P.momOut[2 * n + i] = mom.z;
    }
    // deposit nodes of a particle in cell c: c-1 .. c+2 per axis
    int rel[3];
    BlockWindow const win = blockWindow<1, 2>(a.g, active, nc, rel, sBox, kDiffLdsCells);
    if (win.lds) {
        for (int e = threadIdx.x; e < 3 * win.vol; e += blockDim.x) sJ[e] = 0ull;
        __syncthreads();
        if (active) {
            // nodes c-1..c+2 lie in the window when the cell is held and the
            // start and end positions lie in (-1,2) and [0,1); any other
            // particle (outside a trimmed window, a position rounded onto the
            // cell edge, non-finite data) deposits to global memory with
            // wrapped indices instead
            bool fits = windowHolds<1, 2>(win, rel);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                float const vd = d == 0 ? vel.x : (d == 1 ? vel.y : vel.z);
                float const p0 = p[d] - vd * a.dt / a.cell[d];
                fits = fits && p[d] >= 0.0f && p[d] < 1.0f && p0 > -1.0f && p0 < 2.0f;
            }
            if (fits)
                depositCIC<true>(a, sJ, win, nc, rel, Vec3{p[0], p[1], p[2]}, vel, charge);
            else
                depositCIC<false>(a, sJ, win, nc, rel, Vec3{p[0], p[1], p[2]}, vel, charge);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 3 * win.vol; e += blockDim.x) {
            unsigned long long const v = sJ[e];
            if (v != 0ull) {
                int const comp = e / win.vol;
                __hip_atomic_fetch_add(a.Jfix[comp] + windowToGlobal(a.g, win, e - comp * win.vol), v,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

__global__ __launch_bounds__(256) void kFixToFloat(unsigned long long* fix, float* J0, float* J1, float* J2, size_t n,
                                                   double jinv) {
    size_t const t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    size_t const c = t / n, i = t - c * n;
    float* const J = c == 0 ? J0 : (c == 1 ? J1 : J2);
    long long const acc = (long long)fix[t];
    J[i] = acc == 0 ? 0.0f : (float)((double)acc * jinv);
    fix[t] = 0ull;
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------

// dst += coef * curl(src): forward differences (fwd = 1) or backward (fwd = 0)
template <int FWD>
__global__ __launch_bounds__(256) void kCurlAdd(AdjFieldArgs a, double* d0, double* d1, double* d2, const double* s0,
                                                const double* s1, const double* s2, double coef, double* lJ,
                                                double jcoef) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int const x = (int)(i % a.g.nx), y = (int)((i / a.g.nx) % a.g.ny), z = (int)(i / ((size_t)a.g.nx * a.g.ny));
    const double* s[3] = {s0, s1, s2};
    auto der = [&](int c, int d) {
        int const ux = d == 0, uy = d == 1, uz = d == 2;
        if (FWD) return (s[c][gidx(a.g, x + ux, y + uy, z + uz)] - s[c][i]) / a.cell[d];
        return (s[c][i] - s[c][gidx(a.g, x - ux, y - uy, z - uz)]) / a.cell[d];
    };
    double const cx = der(2, 1) - der(1, 2);
    double const cy = der(0, 2) - der(2, 0);
    double const cz = der(1, 0) - der(0, 1);
    double const v0 = d0[i] + coef * cx;
    double const v1 = d1[i] + coef * cy;
    double const v2 = d2[i] + coef * cz;
    d0[i] = v0;
    d1[i] = v1;
    d2[i] = v2;
    if (lJ) {  // the current's adjoint, lambda_J = coeff * lambda_E'
        lJ[i] = jcoef * v0;
        lJ[n + i] = jcoef * v1;
        lJ[2 * n + i] = jcoef * v2;
    }
}




template <typename R>
struct V3 {
    R x, y, z;
};
template <typename R>
__device__ __forceinline__ V3<R> crossD(V3<R> a, V3<R> b) {
    return V3<R>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R>
__device__ __forceinline__ R dotD(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R>
__device__ __forceinline__ V3<R> axpy(R s, V3<R> x, V3<R> y) { return V3<R>{s * x.x + y.x, s * x.y + y.y, s * x.z + y.z}; }
template <typename R>
__device__ __forceinline__ V3<R> add3(V3<R> a, V3<R> b) { return V3<R>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R>
__device__ __forceinline__ V3<R> scale3(R s, V3<R> a) { return V3<R>{s * a.x, s * a.y, s * a.z}; }

// Values the compiler must treat as new: recomputing from them is not merged
// with earlier identical work, so intermediates are not kept live across phases.
template <typename R>
__device__ __forceinline__ R opaque(R x) {
    asm volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ int opaqueI(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Phase boundary: no instruction is scheduled across it, so each phase's
// loads and temporaries stay within the phase.
__device__ __forceinline__ void phaseBarrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// CIC assignment 1-|k-p| on |k-p| < 1 and its derivative with respect to p
template <typename R>
__device__ __forceinline__ R cicD(R k, R p) {
    R const u = k - p;
    return fabs(u) < R(1.0) ? R(1.0) - fabs(u) : R(0.0);
}
template <typename R>
__device__ __forceinline__ R cicDp(R k, R p) {
    R const u = k - p;
    return fabs(u) < R(1.0) ? (u > R(0.0) ? R(1.0) : (u < R(0.0) ? -R(1.0) : R(0.0))) : R(0.0);
}

// Yee offsets of the E and B components (YeeCell.hpp:69-123)
template <int FIELD, int COMP>  // FIELD 0: E, 1: B
struct YeeOff {
    static constexpr double f(int d) {
        return FIELD == 0 ? (d == COMP ? 0.5 : 0.0) : (d == COMP ? 0.0 : 0.5);
    }
};

// Node offsets and CIC weights of the interpolation of one Yee component at
// in-cell position r of cell c (ShiftCoordinateSystem for an even support).
template <int FIELD, int COMP, typename R>
__device__ __forceinline__ void interpNodes(const GridDev& g, const int* c, const R* r, uint32_t (&o)[3][2],
                                            R (&w)[3][2]) {
    int const n[3] = {g.nx, g.ny, g.nz};
    uint32_t const stride[3] = {1u, (uint32_t)g.nx, (uint32_t)g.nx * (uint32_t)g.ny};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        R const v = r[d] - R(YeeOff<FIELD, COMP>::f(d)) - R(0.5);
        int const s = v >= -R(0.5) ? 0 : -1;
        R const p = v - (R)s + R(0.5);
        w[d][0] = R(1.0) - p;
        w[d][1] = p;
        o[d][0] = (uint32_t)wrapI(c[d] + s, n[d]) * stride[d];
        o[d][1] = (uint32_t)wrapI(c[d] + s + 1, n[d]) * stride[d];
    }
}

template <int FIELD, int COMP, typename R>
__device__ __forceinline__ R gatherD(const float* __restrict__ F, const GridDev& g, const int* c, const R* r) {
    uint32_t o[3][2];
    R w[3][2];
    interpNodes<FIELD, COMP>(g, c, r, o, w);
    R v = R(0.0);
#pragma unroll
    for (int zz = 0; zz < 2; ++zz)
#pragma unroll
        for (int yy = 0; yy < 2; ++yy)
#pragma unroll
            for (int xx = 0; xx < 2; ++xx)
                v += (R)F[o[0][xx] + o[1][yy] + o[2][zz]] * (w[0][xx] * w[1][yy] * w[2][zz]);
    return v;
}

// Reverse of the interpolation of one component: adds lam * weight to the
// field adjoint at the 8 nodes (the block's LDS window plane L when LDS, with
// rel the cell relative to the window reference; the global plane otherwise)
// and lam * d(value)/dX to gX.
template <int FIELD, int COMP, bool LDS, typename R>
__device__ __forceinline__ void gatherAdj(const float* __restrict__ F, double* __restrict__ L, const GridDev& g,
                                          const BlockWindow& win, const int* c, const int* rel, const R* r,
                                          R lam, R* gX) {
    uint32_t o[3][2];
    R w[3][2];
    interpNodes<FIELD, COMP>(g, c, r, o, w);
    uint32_t lo[3][2];
    if (LDS) {
        uint32_t const ls[3] = {1u, (uint32_t)win.ext[0], (uint32_t)(win.ext[0] * win.ext[1])};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int const sh = (r[d] - R(YeeOff<FIELD, COMP>::f(d)) - R(0.5)) >= -R(0.5) ? 0 : -1;
            lo[d][0] = (uint32_t)(rel[d] + sh - win.lo[d]) * ls[d];
            lo[d][1] = lo[d][0] + ls[d];
        }
    }
    R gx = R(0.0), gy = R(0.0), gz = R(0.0);
#pragma unroll
    for (int zz = 0; zz < 2; ++zz)
#pragma unroll
        for (int yy = 0; yy < 2; ++yy)
#pragma unroll
            for (int xx = 0; xx < 2; ++xx) {
                uint32_t const at = o[0][xx] + o[1][yy] + o[2][zz];
                R const f = (R)F[at];
                R const sx = xx ? R(1.0) : -R(1.0), sy = yy ? R(1.0) : -R(1.0), sz = zz ? R(1.0) : -R(1.0);
                gx += f * sx * w[1][yy] * w[2][zz];
                gy += f * w[0][xx] * sy * w[2][zz];
                gz += f * w[0][xx] * w[1][yy] * sz;
                double const v = (double)(lam * (w[0][xx] * w[1][yy] * w[2][zz]));
                if (LDS)
                    __hip_atomic_fetch_add(L + (lo[0][xx] + lo[1][yy] + lo[2][zz]), v, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    unsafeAtomicAdd(L + at, v);
            }
    gX[0] += lam * gx;
    gX[1] += lam * gy;
    gX[2] += lam * gz;
}

// Reverse of the Esirkepov<CIC> deposit of one particle moving from X0 to X1
// (positions in cells relative to the first node of the stencil): the current
// adjoint lambda_J is gathered through the derivatives of the CIC weights of
// both end points. With S0, S1 the weights at the nodes and DS = S1 - S0, the
// current of component c at node (i, j, k) of its rotated axes is
// -csd W(i,j) C(k), W = S0a S0b + (DSa S0b + S0a DSb)/2 + DSa DSb/3 and C(k)
// the prefix sum of DS along c (Esirkepov.hpp:140-240). Accumulates
// dF/dX0 into l0 and dF/dX1 into l1 per axis.
template <int COMP, typename R>
__device__ __forceinline__ void esirkepovAdjComp(const double* __restrict__ lJ, const GridDev& g, const int* bc,
                                                 const R* X0, const R* X1, R csd, R* l0,
                                                 R* l1) {
    constexpr int A = (COMP + 1) % 3;  // rotated i axis
    constexpr int B = (COMP + 2) % 3;  // rotated j axis
    int const n[3] = {g.nx, g.ny, g.nz};
    uint32_t const stride[3] = {1u, (uint32_t)g.nx, (uint32_t)g.nx * (uint32_t)g.ny};
    uint32_t const k0 = (uint32_t)wrapI(bc[COMP], n[COMP]) * stride[COMP];
    uint32_t const k1 = (uint32_t)wrapI(bc[COMP] + 1, n[COMP]) * stride[COMP];
    uint32_t ob[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) ob[j] = (uint32_t)wrapI(bc[B] + j, n[B]) * stride[B];
    R const C0 = cicD(R(0.0), X1[COMP]) - cicD(R(0.0), X0[COMP]);
    R const C1 = C0 + (cicD(R(1.0), X1[COMP]) - cicD(R(1.0), X0[COMP]));
    R dC0 = R(0.0), dC1 = R(0.0);
    R l0a = R(0.0), l1a = R(0.0), l0b = R(0.0), l1b = R(0.0);
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        uint32_t const oa = (uint32_t)wrapI(bc[A] + i, n[A]) * stride[A];
        // the six current adjoints of this row, loaded together
        R m[3][2];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            m[j][0] = lJ[oa + ob[j] + k0];
            m[j][1] = lJ[oa + ob[j] + k1];
        }
        R const s0a = cicD((R)i, X0[A]);
        R const dsa = cicD((R)i, X1[A]) - s0a;
        R const d0a = cicDp((R)i, X0[A]);
        R const d1a = cicDp((R)i, X1[A]);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            R const s0b = cicD((R)j, X0[B]);
            R const dsb = cicD((R)j, X1[B]) - s0b;
            R const W = s0a * s0b + R(0.5) * (dsa * s0b + s0a * dsb) + (R(1.0) / R(3.0)) * dsa * dsb;
            R const m0 = -csd * m[j][0];
            R const m1 = -csd * m[j][1];
            R const G = m0 * C0 + m1 * C1;
            dC0 += m0 * W;
            dC1 += m1 * W;
            // dW/dS0a = S0b + DSb/2 - dW/dDSa, dW/dS1a = dW/dDSa = S0b/2 + DSb/3 (and a <-> b)
            R const wa = R(0.5) * s0b + (R(1.0) / R(3.0)) * dsb;
            R const wb = R(0.5) * s0a + (R(1.0) / R(3.0)) * dsa;
            l0a += G * (s0b + R(0.5) * dsb - wa) * d0a;
            l1a += G * wa * d1a;
            l0b += G * (s0a + R(0.5) * dsa - wb) * cicDp((R)j, X0[B]);
            l1b += G * wb * cicDp((R)j, X1[B]);
        }
    }
    l0[A] += l0a;
    l1[A] += l1a;
    l0[B] += l0b;
    l1[B] += l1b;
    // C0 = DS_0, C1 = DS_0 + DS_1 along the current's own axis
    l1[COMP] += (dC0 + dC1) * cicDp(R(0.0), X1[COMP]) + dC1 * cicDp(R(1.0), X1[COMP]);
    l0[COMP] -= (dC0 + dC1) * cicDp(R(0.0), X0[COMP]) + dC1 * cicDp(R(1.0), X0[COMP]);
}

template <typename R>
struct BorisFwd {
    V3<R> pm, t, s, pp, p1;
    R g, f, beta, hq;
};

// Boris push in the adjoint's precision (particlePusherBoris.hpp:41-90)
template <typename R>
__device__ __forceinline__ BorisFwd<R> borisFwd(V3<R> p0, V3<R> E, V3<R> B, R m, R q, R dt, R rc2) {
    BorisFwd<R> b;
    b.hq = R(0.5) * q * dt;
    b.pm = axpy(b.hq, E, p0);
    b.g = sqrt(R(1.0) + dotD(b.pm, b.pm) * rc2 / (m * m));
    b.beta = R(0.5) * (q / m) * dt;
    b.t = scale3(b.beta / b.g, B);
    b.f = R(1.0) / (R(1.0) + dotD(b.t, b.t));
    b.s = scale3(R(2.0) * b.f, b.t);
    b.pp = add3(b.pm, crossD(b.pm, b.t));
    b.p1 = axpy(b.hq, E, add3(b.pm, crossD(b.pp, b.s)));
    return b;
}

template <typename R>
__device__ __forceinline__ void adjointParticle(const AdjPartArgs& a, const AdjState& S, int64_t idx, const int* c,
                                                const BlockWindow& win, const int* rel, double* sL, bool lds) {
    int64_t const n = S.n;
    R const r[3] = {(R)S.pos[idx], (R)S.pos[n + idx], (R)S.pos[2 * n + idx]};
    V3<R> const p0{(R)S.mom[idx], (R)S.mom[n + idx], (R)S.mom[2 * n + idx]};
    R const w = (R)S.w[idx];
    R const m = (R)a.mass * w;
    R const q = (R)a.charge * w;
    R const dt = (R)a.dt;
    R const rc2 = (R)a.rc2;
    R const cellR[3] = {(R)a.cell[0], (R)a.cell[1], (R)a.cell[2]};
    R const ivdt = (R)a.inv_vdt;

    // ---- forward recompute: gather, Boris, velocity ----
    V3<R> const E{gatherD<0, 0>(a.E[0], a.g, c, r), gatherD<0, 1>(a.E[1], a.g, c, r), gatherD<0, 2>(a.E[2], a.g, c, r)};
    phaseBarrier();
    V3<R> const B{gatherD<1, 0>(a.B[0], a.g, c, r), gatherD<1, 1>(a.B[1], a.g, c, r), gatherD<1, 2>(a.B[2], a.g, c, r)};
    phaseBarrier();
    V3<R> p1;
    {
        BorisFwd<R> const b = borisFwd(p0, E, B, m, q, dt, rc2);
        p1 = b.p1;
    }
    R const G = sqrt(m * m + dotD(p1, p1) * rc2);
    R const v[3] = {p1.x / G, p1.y / G, p1.z / G};

    // ---- reverse Esirkepov ----
    R lX1[3], lv[3];
    {
        R X0[3], X1[3];
        int bc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            X1[d] = r[d] + v[d] * dt / cellR[d];
            X0[d] = X1[d] - v[d] * dt / cellR[d];
            int const gs = (int)fmin(floor(X0[d]), floor(X1[d]));
            X0[d] -= (R)gs;
            X1[d] -= (R)gs;
            bc[d] = c[d] + gs;
        }
        R l0[3] = {R(0.0), R(0.0), R(0.0)}, l1[3] = {R(0.0), R(0.0), R(0.0)};
        phaseBarrier();
        esirkepovAdjComp<0>(a.lJ[0], a.g, bc, X0, X1, q * ivdt * cellR[0], l0, l1);
        phaseBarrier();
        esirkepovAdjComp<1>(a.lJ[1], a.g, bc, X0, X1, q * ivdt * cellR[1], l0, l1);
        phaseBarrier();
        esirkepovAdjComp<2>(a.lJ[2], a.g, bc, X0, X1, q * ivdt * cellR[2], l0, l1);
        phaseBarrier();
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            // X1 = X + v dt/cell, X0 = X1 - v dt/cell
            lX1[d] = (R)S.lX[d * n + idx] + l0[d] + l1[d];
            lv[d] = (lX1[d] - l0[d]) * dt / cellR[d];
        }
    }

    phaseBarrier();
    // ---- reverse velocity: v = p1 / sqrt(m^2 + |p1|^2 rc2) ----
    V3<R> const LV{lv[0], lv[1], lv[2]};
    R const kv = dotD(p1, LV) * rc2 / (G * G * G);
    V3<R> const l1{(R)S.lP[idx] + LV.x / G - p1.x * kv, (R)S.lP[n + idx] + LV.y / G - p1.y * kv,
                   (R)S.lP[2 * n + idx] + LV.z / G - p1.z * kv};

    // ---- reverse Boris ----
    V3<R> lE, lB, lp0;
    {
        V3<R> const p0o{opaque(p0.x), opaque(p0.y), opaque(p0.z)};
        V3<R> const Eo{opaque(E.x), opaque(E.y), opaque(E.z)};
        V3<R> const Bo{opaque(B.x), opaque(B.y), opaque(B.z)};
        BorisFwd<R> const b = borisFwd(p0o, Eo, Bo, m, q, dt, rc2);
        lE = scale3(b.hq, l1);                   // p1 = pl + hq E
        V3<R> lpm = l1;                             // pl = pm + pp x s
        V3<R> const lpp = crossD(b.s, l1);
        V3<R> const ls = crossD(l1, b.pp);
        lpm = add3(lpm, add3(lpp, crossD(b.t, lpp)));  // pp = pm + pm x t
        V3<R> const lt1 = crossD(lpp, b.pm);
        R const tls = dotD(b.t, ls);              // s = 2 f t, f = 1/(1+|t|^2)
        V3<R> const lt = add3(lt1, add3(scale3(R(2.0) * b.f, ls), scale3(-R(4.0) * b.f * b.f * tls, b.t)));
        lB = scale3(b.beta / b.g, lt);                 // t = beta B / g
        R const lg = -b.beta * dotD(B, lt) / (b.g * b.g);
        lpm = axpy(lg * rc2 / (m * m * b.g), b.pm, lpm);  // g = sqrt(1 + |pm|^2 rc2/m^2)
        lE = axpy(b.hq, lpm, lE);                      // pm = p0 + hq E
        lp0 = lpm;
    }

    phaseBarrier();
    // ---- reverse gather (weights recomputed, not kept from the forward gather) ----
    R lX0[3] = {lX1[0], lX1[1], lX1[2]};
    int const co[3] = {opaqueI(c[0]), opaqueI(c[1]), opaqueI(c[2])};
    R const ro[3] = {opaque(r[0]), opaque(r[1]), opaque(r[2])};
    if (lds) {
        gatherAdj<0, 0, true>(a.E[0], sL, a.g, win, co, rel, ro, lE.x, lX0);
        phaseBarrier();
        gatherAdj<0, 1, true>(a.E[1], sL + win.vol, a.g, win, co, rel, ro, lE.y, lX0);
        phaseBarrier();
        gatherAdj<0, 2, true>(a.E[2], sL + 2 * win.vol, a.g, win, co, rel, ro, lE.z, lX0);
        phaseBarrier();
        gatherAdj<1, 0, true>(a.B[0], sL + 3 * win.vol, a.g, win, co, rel, ro, lB.x, lX0);
        phaseBarrier();
        gatherAdj<1, 1, true>(a.B[1], sL + 4 * win.vol, a.g, win, co, rel, ro, lB.y, lX0);
        phaseBarrier();
        gatherAdj<1, 2, true>(a.B[2], sL + 5 * win.vol, a.g, win, co, rel, ro, lB.z, lX0);
    } else {
        gatherAdj<0, 0, false>(a.E[0], a.lE[0], a.g, win, co, rel, ro, lE.x, lX0);
        phaseBarrier();
        gatherAdj<0, 1, false>(a.E[1], a.lE[1], a.g, win, co, rel, ro, lE.y, lX0);
        phaseBarrier();
        gatherAdj<0, 2, false>(a.E[2], a.lE[2], a.g, win, co, rel, ro, lE.z, lX0);
        phaseBarrier();
        gatherAdj<1, 0, false>(a.B[0], a.lB[0], a.g, win, co, rel, ro, lB.x, lX0);
        phaseBarrier();
        gatherAdj<1, 1, false>(a.B[1], a.lB[1], a.g, win, co, rel, ro, lB.y, lX0);
        phaseBarrier();
        gatherAdj<1, 2, false>(a.B[2], a.lB[2], a.g, win, co, rel, ro, lB.z, lX0);
    }
    S.lX[idx] = lX0[0];
    S.lX[n + idx] = lX0[1];
    S.lX[2 * n + idx] = lX0[2];
    S.lP[idx] = lp0.x;
    S.lP[n + idx] = lp0.y;
    S.lP[2 * n + idx] = lp0.z;
}


// FP64 LDS window of the field adjoint scatter: 6 components of this many cells
constexpr int kAdjLdsCells = 640;

// Precision of the per-particle adjoint arithmetic (Jacobian-vector
// products); the adjoint state (lambda of particles and fields) is always
// accumulated in FP64.
using AdjReal = float;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void kAdjParticles(AdjPartArgs a,
                                                                                                 AdjState S) {
    __shared__ double sL[6 * kAdjLdsCells];
    __shared__ int sBox[9];
    int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool const active = idx < S.n;
    int64_t const n = S.n;
    int c[3] = {0, 0, 0};
    if (active) {
        c[0] = S.cell[idx];
        c[1] = S.cell[n + idx];
        c[2] = S.cell[2 * n + idx];
    }
    // interpolation nodes of a particle in cell c: c-1 .. c+1 per axis
    int rel[3];
    BlockWindow const win = blockWindow<1, 1>(a.g, active, c, rel, sBox, kAdjLdsCells);
    if (win.lds) {
        for (int e = threadIdx.x; e < 6 * win.vol; e += blockDim.x) sL[e] = 0.0;
        __syncthreads();
        if (active) adjointParticle<AdjReal>(a, S, idx, c, win, rel, sL, windowHolds<1, 1>(win, rel));
        __syncthreads();
        for (int e = threadIdx.x; e < 6 * win.vol; e += blockDim.x) {
            double const v = sL[e];
            if (v != 0.0) {
                int const plane = e / win.vol;
                double* const dst = plane < 3 ? a.lE[plane] : a.lB[plane - 3];
                unsafeAtomicAdd(dst + windowToGlobal(a.g, win, e - plane * win.vol), v);
            }
        }
    }
}

__global__ __launch_bounds__(256) void kFillD(double* p, size_t n, double v) {
    size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// re-tiling of the particle order at checkpoints
// ---------------------------------------------------------------------------
// key = tile of 8x4x1 cells of the particle's cell (x-fastest tile order), value = slot
__global__ __launch_bounds__(256) void kTileKeys(const int32_t* __restrict__ cell, int64_t n, GridDev g,
                                                 uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t const ntx = (uint32_t)g.nx / kTileX, nty = (uint32_t)g.ny / kTileY;
    uint32_t const tx = (uint32_t)cell[i] / kTileX, ty = (uint32_t)cell[n + i] / kTileY;
    uint32_t const tz = (uint32_t)cell[2 * n + i] / kTileZ;
    keys[i] = tx + ntx * (ty + nty * tz);
    vals[i] = (int32_t)i;
}

// dst slot j takes the particle of src slot perm[j]; ids follow the particles
__global__ __launch_bounds__(256) void kGatherParticles(const int32_t* __restrict__ perm, DiffParticles src,
                                                        DiffParticles dst, const int32_t* __restrict__ idsIn,
                                                        int32_t* __restrict__ idsOut, float* __restrict__ wOut) {
    int64_t const j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t const n = src.n;
    if (j >= n) return;
    int64_t const i = perm[j];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dst.cellOut[d * n + j] = src.cell[d * n + i];
        dst.posOut[d * n + j] = src.pos[d * n + i];
        dst.momOut[d * n + j] = src.mom[d * n + i];
    }
    wOut[j] = src.w[i];
    idsOut[j] = idsIn[i];
}

// adjoint of the re-tiling: dst[perm[j]] = src[j] for 3 planes of n
__global__ __launch_bounds__(256) void kScatterAdjoint(const int32_t* __restrict__ perm, const double* __restrict__ src,
                                                       double* __restrict__ dst, int64_t n) {
    int64_t const j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int64_t const i = perm[j];
#pragma unroll
    for (int d = 0; d < 3; ++d) dst[d * n + i] = src[d * n + j];
}

// rocPRIM's stable LSD radix sort of (tile key, slot) pairs: equal keys keep
// their slot order, so the permutation is deterministic
hipError_t sortTempBytes(int64_t n, size_t* bytes) {
    return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0u, 32u, nullptr);
}

hipError_t launchTileSort(const int32_t* cell, int64_t n, const GridDev& g, uint32_t* keys[2], int32_t* vals[2],
                          void* temp, size_t tempBytes, int32_t** perm, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kTileKeys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cell, n, g, keys[0], vals[0]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t const tiles = (uint32_t)(g.nx / kTileX) * (uint32_t)(g.ny / kTileY) * (uint32_t)(g.nz / kTileZ);
    int bits = 1;
    while (bits < 32 && (1u << bits) < tiles) ++bits;
    size_t b = tempBytes;
    e = rocprim::radix_sort_pairs(temp, b, keys[0], keys[1], vals[0], vals[1], (size_t)n, 0u, (unsigned)bits, s);
    *perm = vals[1];
    return e;
}

hipError_t launchGatherParticles(const int32_t* perm, const DiffParticles& src, const DiffParticles& dst,
                                 const int32_t* idsIn, int32_t* idsOut, float* wOut, hipStream_t s) {
    if (src.n == 0) return hipSuccess;
    hipLaunchKernelGGL(kGatherParticles, dim3((unsigned)((src.n + 255) / 256)), dim3(256), 0, s, perm, src, dst,
                       idsIn, idsOut, wOut);
    return hipGetLastError();
}

hipError_t launchScatterAdjoint(const int32_t* perm, const double* src, double* dst, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kScatterAdjoint, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, src, dst, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocksOf(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launchDiffPush(const DiffArgs& a, const DiffParticles& P, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(kDiffPush, dim3(blocksOf((size_t)P.n)), dim3(256), 0, s, a, P);
    return hipGetLastError();
}

hipError_t launchFixToFloat(unsigned long long* fix, float* const J[3], size_t n, double jinv, hipStream_t s) {
    hipLaunchKernelGGL(kFixToFloat, dim3(blocksOf(3 * n)), dim3(256), 0, s, fix, J[0], J[1], J[2], n, jinv);
    return hipGetLastError();
}

// reverse of: B' = B - h C+ E; E' = E + dt c^2 C- B' + coeff J; B'' = B' - h C+ E'
// given lE, lB (adjoints of E'', B''), leaves the adjoints of E, B there and
// lambda_J = coeff * (adjoint of E') in lJ.
hipError_t launchAdjFields(const AdjFieldArgs& a, double dt, double c2, double coeff, double* lJ[3], hipStream_t s) {
    size_t const n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    double const h = 0.5 * dt;
    // B'' = B' - h C+ E'  ->  lE' = lE - h C- lB, and lambda_J = coeff lE'
    hipLaunchKernelGGL(kCurlAdd<0>, dim3(blocksOf(n)), dim3(256), 0, s, a, a.lE[0], a.lE[1], a.lE[2], a.lB[0],
                       a.lB[1], a.lB[2], -h, lJ[0], coeff);
    // E' = E + dt c^2 C- B' + coeff J  ->  lB' = lB + dt c^2 C+ lE'
    hipLaunchKernelGGL(kCurlAdd<1>, dim3(blocksOf(n)), dim3(256), 0, s, a, a.lB[0], a.lB[1], a.lB[2], a.lE[0],
                       a.lE[1], a.lE[2], dt * c2, (double*)nullptr, 0.0);
    // B' = B - h C+ E  ->  lE = lE' - h C- lB'
    hipLaunchKernelGGL(kCurlAdd<0>, dim3(blocksOf(n)), dim3(256), 0, s, a, a.lE[0], a.lE[1], a.lE[2], a.lB[0],
                       a.lB[1], a.lB[2], -h, (double*)nullptr, 0.0);
    return hipGetLastError();
}

hipError_t launchFillD(double* p, size_t n, double v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kFillD, dim3(blocksOf(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launchAdjParticles(const AdjPartArgs& a, const AdjState& st, hipStream_t s) {
    if (st.n == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdjParticles, dim3(blocksOf((size_t)st.n)), dim3(256), 0, s, a, st);
    return hipGetLastError();
}

}  // namespace pic
