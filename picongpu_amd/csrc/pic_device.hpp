// Device-side physics of the PIC cycle for gfx950.
//
// Every function here reproduces the floating-point operation order of the
// reference functor it replaces (cited per function); the code is compiled
// with -ffp-contract=off and IEEE division/sqrt (the hipcc defaults we keep),
// so the only multiply-adds are the ones written as `mad` below: the field
// gather's sums and Esirkepov's W and prefix sums, the places where the
// reference's GPU build (nvcc contracts by default) fuses them too. Elsewhere
// the results are bit-identical to the reference's CPU path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_kernels.hpp"

namespace pic {

// Supercell shape (param/memory.param:51 SuperCellSize = 8,8,4) and frame size.
constexpr int SCX = 8;
constexpr int SCY = 8;
constexpr int SCZ = 4;
constexpr int SCV = SCX * SCY * SCZ;
constexpr int FRAME = 256;

struct Vec3 {
    float x, y, z;
};

__device__ __forceinline__ float comp(const Vec3& v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }

// l2norm2: x*x, then += y*y, += z*z (pmacc/math/vector/Vector.tpp:99-110)
__device__ __forceinline__ float l2norm2(Vec3 v) {
    float t = v.x * v.x;
    t += v.y * v.y;
    t += v.z * v.z;
    return t;
}

// a * b + c rounded once: the device order's contracted multiply-add
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Esirkepov's weight S0i S0j + 1/2 (dSi S0j + S0i dSj) + 1/3 dSj dSi
// (Esirkepov.hpp:218) in the reference's order, each addition fused with the
// product it adds (the 1/2 product is exact, so that fusion rounds nothing)
__device__ __forceinline__ float esirkepovWeight(float s0i, float dsi, float s0j, float dsj) {
    float const mixed = mad(dsi, s0j, s0i * dsj);
    float const half = mad(0.5f, mixed, s0i * s0j);
    return mad((1.0f / 3.0f) * dsj, dsi, half);
}

// cross product (pmacc/math/vector/Vector.tpp:71-76)
__device__ __forceinline__ Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------------------
// Shapes: on-support assignment arrays used by the field gather
// (particles/shapes/{CIC,TSC,PQS}.hpp, struct detail::*::shapeArray and
// ChargeAssignmentOnSupport::shapeArray).
// ---------------------------------------------------------------------------
template <int SUPP>
struct Shape;

template <>
struct Shape<2> {  // CIC.hpp:38-161
    static constexpr int begin = 0;      // ChargeAssignment::begin
    static constexpr int end = 2;        // ChargeAssignment::end
    // x in [0,1): {1-x, x}
    __device__ __forceinline__ static void onSupport(float x, float* s) {
        s[0] = 1.0f - x;
        s[1] = x;
    }
    // ChargeAssignment::operator() (CIC.hpp:89-107)
    __device__ __forceinline__ static float assign(float x) {
        float const abs_x = fabsf(x);
        bool const below_1 = abs_x < 1.0f;
        float const onSupport = 1.0f - abs_x;
        float result = 0.0f;
        if (below_1) result = onSupport;
        return result;
    }
    // ChargeAssignmentOnSupport::operator() (CIC.hpp:139-147)
    __device__ __forceinline__ static float onSupportAt(float x) { return 1.0f - fabsf(x); }
    // assign(1 - p) for p in [0,1): 1 - p lies in (0,1], and at 1 both
    // branches give 0, so the value is 1 - |1 - p| without a branch
    __device__ __forceinline__ static float assignAt1(float p) { return 1.0f - fabsf(1.0f - p); }
};

template <>
struct Shape<3> {  // TSC.hpp:39-201
    static constexpr int begin = -1;
    static constexpr int end = 2;
    __device__ __forceinline__ static float ff1(float x) {
        float const square_x = x * x;
        return 0.75f - square_x;
    }
    __device__ __forceinline__ static float ff2(float x) {
        float const tmp = 1.5f - x;  // 3.0_X / 2.0_X
        float const square_tmp = tmp * tmp;
        return 0.5f * square_tmp;
    }
    // x in [-0.5,0.5): grid points -1,0,1
    __device__ __forceinline__ static void onSupport(float x, float* s) {
        s[0] = ff2(fabsf(-1.0f - x));
        s[1] = ff1(fabsf(x));
        s[2] = 1.0f - (s[0] + s[1]);
    }
    __device__ __forceinline__ static float assign(float x) {
        float const abs_x = fabsf(x);
        bool const below_05 = abs_x < 0.5f;
        bool const below_1_5 = abs_x < 1.5f;
        float const rad1 = ff1(abs_x);
        float const rad2 = ff2(abs_x);
        float result = 0.0f;
        if (below_05)
            result = rad1;
        else if (below_1_5)
            result = rad2;
        return result;
    }
    // ChargeAssignmentOnSupport::operator() (TSC.hpp:166-185)
    __device__ __forceinline__ static float onSupportAt(float x) {
        float const abs_x = fabsf(x);
        bool const below_05 = abs_x < 0.5f;
        float const rad1 = ff1(abs_x);
        float const rad2 = ff2(abs_x);
        return below_05 ? rad1 : rad2;
    }
    // assign(1 - p) for p in [-0.5,0.5): |1 - p| lies in [0.5,1.5] (the
    // rounded difference included), where assign takes the ff2 branch below
    // 1.5; at 1.5 its zero branch equals ff2(1.5) = 0
    __device__ __forceinline__ static float assignAt1(float p) { return ff2(fabsf(1.0f - p)); }
};

template <>
struct Shape<4> {  // PQS.hpp:38-205
    static constexpr int begin = -1;
    static constexpr int end = 3;
    __device__ __forceinline__ static float ff1(float x) {
        float const square_x = x * x;
        float const triple_x = square_x * x;
        return (1.0f / 6.0f) * (4.0f - 6.0f * square_x + 3.0f * triple_x);
    }
    __device__ __forceinline__ static float ff2(float x) {
        float const tmp = 2.0f - x;
        float const triple_tmp = tmp * tmp * tmp;
        return (1.0f / 6.0f) * triple_tmp;
    }
    // x in [0,1): grid points -1,0,1,2
    __device__ __forceinline__ static void onSupport(float x, float* s) {
        s[0] = ff2(fabsf(-1.0f - x));
        s[1] = ff1(x);
        s[3] = ff2(2.0f - x);
        s[2] = 1.0f - (s[0] + s[1] + s[3]);
    }
    __device__ __forceinline__ static float assign(float x) {
        float const abs_x = fabsf(x);
        bool const below_1 = abs_x < 1.0f;
        bool const below_2 = abs_x < 2.0f;
        float const rad1 = ff1(abs_x);
        float const rad2 = ff2(abs_x);
        float result = 0.0f;
        if (below_1)
            result = rad1;
        else if (below_2)
            result = rad2;
        return result;
    }
    // ChargeAssignmentOnSupport::operator() (PQS.hpp:166-185)
    __device__ __forceinline__ static float onSupportAt(float x) {
        float const abs_x = fabsf(x);
        bool const below_1 = abs_x < 1.0f;
        float const rad1 = ff1(abs_x);
        float const rad2 = ff2(abs_x);
        return below_1 ? rad1 : rad2;
    }
    // assign(1 - p) for p in [0,1): 1 - p lies in (0,1], where assign takes
    // the ff1 branch; at 1 ff2 gives the same 1/6 (4 - 6 + 3 = 1 exactly)
    __device__ __forceinline__ static float assignAt1(float p) { return ff1(fabsf(1.0f - p)); }
};

// ChargeAssignment::shapeArray(pos, isOutOfRange): SUPP+1 values for grid
// points begin..end (CIC.hpp:115-128, TSC.hpp:141-155, PQS.hpp:143-157).
template <int SUPP>
__device__ __forceinline__ void cachedShape(float pos, bool outOfRange, float* s) {
    float const x = outOfRange ? pos - 1.0f : pos;
    float v[SUPP + 1];
    Shape<SUPP>::onSupport(x, v);
    v[SUPP] = 0.0f;
#pragma unroll
    for (int i = SUPP; i >= 1; --i) s[i] = outOfRange ? v[i - 1] : v[i];
    s[0] = outOfRange ? 0.0f : v[0];
}

// ---------------------------------------------------------------------------
// Field gather: FieldToParticleInterpolation<Shape, AssignedTrilinearInterpolation>
// (algorithms/FieldToParticleInterpolation.hpp:43-131,
// ShiftCoordinateSystem.hpp:36-80, AssignedTrilinearInterpolation.hpp:53-86).
// `tile` is one field component staged as [z][y][x] with row pitch TX and
// plane pitch TX*TY; (cx,cy,cz) is the particle's cell in tile coordinates.
// ---------------------------------------------------------------------------
template <int SUPP, int TX, int TY>
__device__ __forceinline__ float gather1(const float* __restrict__ tile, int cx, int cy, int cz, Vec3 pos,
                                         float fx, float fy, float fz) {
    constexpr bool isEven = (SUPP % 2) == 0;
    constexpr int begin = -SUPP / 2 + (SUPP + 1) % 2;
    float p[3] = {pos.x, pos.y, pos.z};
    float const f[3] = {fx, fy, fz};
    int ishift[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float const v_pos = p[d] - f[d] - 0.5f;
        int s;
        if (isEven)
            s = v_pos >= -0.5f ? 0 : -1;
        else
            s = v_pos >= 0.0f ? 1 : 0;
        ishift[d] = s;
        p[d] = v_pos - (float)s + 0.5f;
    }
    float sx[SUPP], sy[SUPP], sz[SUPP];
    Shape<SUPP>::onSupport(p[0], sx);
    Shape<SUPP>::onSupport(p[1], sy);
    Shape<SUPP>::onSupport(p[2], sz);
    int const bx = cx + ishift[0] + begin;
    int const by = cy + ishift[1] + begin;
    int const bz = cz + ishift[2] + begin;
    // each partial sum starts at its first term; the reference adds that term
    // to 0.0, which differs only in the sign of an all-zero sum (a zero field
    // value adds nothing to a momentum, whatever its sign). Every further
    // term is a multiply-add.
    float result_z = 0.0f;
#pragma unroll
    for (int z = 0; z < SUPP; ++z) {
        float result_y = 0.0f;
#pragma unroll
        for (int y = 0; y < SUPP; ++y) {
            const float* row = tile + ((bz + z) * TY + (by + y)) * TX + bx;
            float result_x = row[0] * sx[0];
#pragma unroll
            for (int x = 1; x < SUPP; ++x) result_x = mad(row[x], sx[x], result_x);
            result_y = y == 0 ? result_x * sy[0] : mad(result_x, sy[y], result_y);
        }
        result_z = z == 0 ? result_y * sz[0] : mad(result_y, sz[z], result_z);
    }
    return result_z;
}

// E components sit at (1/2,0,0),(0,1/2,0),(0,0,1/2); B at (0,1/2,1/2),
// (1/2,0,1/2),(1/2,1/2,0) (fields/YeeCell.hpp:69-123).
template <int SUPP, int TX, int TY, int TV>
__device__ __forceinline__ Vec3 gatherE(const float* __restrict__ t, int cx, int cy, int cz, Vec3 pos) {
    return Vec3{gather1<SUPP, TX, TY>(t, cx, cy, cz, pos, 0.5f, 0.0f, 0.0f),
                gather1<SUPP, TX, TY>(t + TV, cx, cy, cz, pos, 0.0f, 0.5f, 0.0f),
                gather1<SUPP, TX, TY>(t + 2 * TV, cx, cy, cz, pos, 0.0f, 0.0f, 0.5f)};
}
template <int SUPP, int TX, int TY, int TV>
__device__ __forceinline__ Vec3 gatherB(const float* __restrict__ t, int cx, int cy, int cz, Vec3 pos) {
    return Vec3{gather1<SUPP, TX, TY>(t, cx, cy, cz, pos, 0.0f, 0.5f, 0.5f),
                gather1<SUPP, TX, TY>(t + TV, cx, cy, cz, pos, 0.5f, 0.0f, 0.5f),
                gather1<SUPP, TX, TY>(t + 2 * TV, cx, cy, cz, pos, 0.5f, 0.5f, 0.0f)};
}

// ---------------------------------------------------------------------------
// Gamma and Velocity (algorithms/Gamma.hpp:28-38, Velocity.hpp:28-39).
// Velocity uses 1/sqrt like the reference CPU build
// (thirdParty/alpaka/include/alpaka/math/Traits.hpp:729-734), not rsqrtf.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gammaOf(Vec3 mom, float mass, float c2) {
    float const fMom2 = l2norm2(mom);
    float const m2_c2_reci = 1.0f / (mass * mass * c2);
    return sqrtf(1.0f + fMom2 * m2_c2_reci);
}

__device__ __forceinline__ Vec3 velocityOf(Vec3 mom, float mass, float rc2) {
    float const m0_2 = mass * mass;
    float const fMom2 = l2norm2(mom);
    float const t = 1.0f / sqrtf(m0_2 + fMom2 * rc2);
    return Vec3{t * mom.x, t * mom.y, t * mom.z};
}

// ---------------------------------------------------------------------------
// PML weight damping (boundary/Absorbing.hpp:95-200 dampWeightsInPml and
// DampWeightsInPml, fields/absorber/pml/Pml.kernel:116-136 getRelativeDepth,
// :196-216 getNormalizedSigma, :326-386 getNormalizedSigmaIntegral,
// traits/attribute/DampedWeighting.hpp:67-90 dampWeighting). The FP32 pow
// and exp the reference calls are taken as FP64 results rounded to FP32,
// from the FP64 routines below (the platform's FP64 pow and exp would spill
// registers in the push kernel): IEEE add, multiply and divide in a fixed
// order plus exact scalings by powers of two, so every platform gets the
// same bits.
// ---------------------------------------------------------------------------

// e^x for x <= 0: x = k ln2 + r, |r| <= ln2 / 2, with ln2 in two parts (Cody
// and Waite's reduction: k ln2_hi is exact for |k| < 2^20), e^r by its
// Taylor polynomial of degree 13 (remainder below 1e-17), scaled by 2^k
__device__ __forceinline__ double expNegF64(double x) {
    if (x < -745.0) return 0.0;
    double const k = rint(x * 1.4426950408889634);
    double const r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double p = 1.0 / 6227020800.0;  // 1/13!
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, static_cast<int>(k));
}

// ln(d) for 0 < d <= 1: d = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m =
// 2 atanh(s), s = (m - 1) / (m + 1), |s| < 0.172, by its series to s^21
__device__ __forceinline__ double logUnitF64(double d) {
    int e;
    double m = frexp(d, &e);  // [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    double const s = (m - 1.0) / (m + 1.0);
    double const s2 = s * s;
    double q = 1.0 / 21.0;
    q = q * s2 + 1.0 / 19.0;
    q = q * s2 + 1.0 / 17.0;
    q = q * s2 + 1.0 / 15.0;
    q = q * s2 + 1.0 / 13.0;
    q = q * s2 + 1.0 / 11.0;
    q = q * s2 + 1.0 / 9.0;
    q = q * s2 + 1.0 / 7.0;
    q = q * s2 + 1.0 / 5.0;
    q = q * s2 + 1.0 / 3.0;
    q = q * s2 + 1.0;
    double const lnm = 2.0 * s * q;
    double const de = static_cast<double>(e);
    return (de * 6.93147180369123816490e-01 + lnm) + de * 1.90821492927058770002e-10;
}

constexpr float kMinWeighting = 10.0f;  // particle.param MIN_WEIGHTING

// normalised sigma at a position along axis d (cells, no guard)
__device__ __forceinline__ float pmlSigmaAt(const PmlDampArgs& p, int d, float pos) {
    // the reference adds the guard cells and subtracts them again
    float const withGuard = pos + static_cast<float>(p.guard[d]);
    float const zero = withGuard - static_cast<float>(p.guard[d]);
    float depth = 0.0f;
    if (zero < p.negB[d]) {
        depth = (p.negB[d] - zero) / p.negB[d];
    } else {
        float const rightStart = p.nCells[d] - p.posB[d];
        if (zero > rightStart) depth = (zero - rightStart) / p.posB[d];
    }
    if (depth == 0.0f) return 0.0f;
    // depth^grading in FP64: repeated products for an integer order, else
    // e^(grading ln depth) (depth in (0, 1]; a particle reconstructed beyond
    // the border may give more, where both still hold)
    double powered;
    if (p.gradingInt >= 0) {
        powered = 1.0;
        for (int i = 0; i < p.gradingInt; ++i) powered = powered * static_cast<double>(depth);
    } else {
        powered = expNegF64(static_cast<double>(p.grading) * logUnitF64(static_cast<double>(depth)));
    }
    return p.sigmaMax[d] * static_cast<float>(powered);
}

// integral of the normalised sigma over the straight path start -> finish
// taking dt (cells along axis d)
__device__ __forceinline__ float pmlSigmaIntegral(const PmlDampArgs& p, int d, float start, float finish, float dt) {
    float const vel = (finish - start) / dt;
    if (static_cast<double>(fabsf(vel)) < p.slow[d]) {
        float const s0 = pmlSigmaAt(p, d, start), s1 = pmlSigmaAt(p, d, finish);
        return dt * 0.5f * (s0 + s1);
    }
    float as = start, af = finish;
    float const negEnd = p.negB[d];
    if (finish < negEnd && start >= negEnd) as = negEnd;
    if (start < negEnd && finish >= negEnd) af = negEnd;
    float const posStart = p.nCells[d] - p.posB[d];
    if (finish > posStart && start <= posStart) as = posStart;
    if (start > posStart && finish <= posStart) af = posStart;
    float const adjT = (af - as) / vel;
    float depthCoeff = 0.0f;
    if (as <= negEnd) depthCoeff = as - negEnd;
    if (as >= posStart) depthCoeff = as - posStart;
    float const s0 = pmlSigmaAt(p, d, as), s1 = pmlSigmaAt(p, d, af);
    return (s1 * adjT + (s1 - s0) * depthCoeff / vel) / (p.grading + 1.0f);
}

// one exchange's damping along axis d of a particle at in-cell position pos
// of global cell `cell`: the path over the last step from its velocity,
// weighting and momentum scaled by exp(-integral), the weighting kept at
// MIN_WEIGHTING or above
__device__ __forceinline__ void pmlDamp(const PmlDampArgs& p, int d, int cell, float pos, float massPerWeight,
                                        float rc2, float dt, float cellSize, Vec3& mom, float& w) {
    Vec3 const vel = velocityOf(mom, massPerWeight * w, rc2);
    float const v = d == 0 ? vel.x : (d == 1 ? vel.y : vel.z);
    float const current = static_cast<float>(cell) + pos;
    float const previous = current - v * dt / cellSize;
    float const integral = pmlSigmaIntegral(p, d, previous, current, dt);
    float const multiplier = static_cast<float>(expNegF64(static_cast<double>(-integral)));
    float const attempted = w * multiplier;
    float const damped = fmaxf(attempted, kMinWeighting);
    float const ratio = damped / w;
    mom.x = mom.x * ratio;
    mom.y = mom.y * ratio;
    mom.z = mom.z * ratio;
    w = damped;
}

struct PushConsts {
    float cell[3];
    float dt;
    float c2;         // speedOfLight^2
    float rc2;        // mue0*eps0
    float c;          // speedOfLight
};

// Boris (particles/pusher/particlePusherBoris.hpp:41-90). Returns new momentum
// and velocity; pos is advanced in place.
__device__ __forceinline__ void pushBoris(Vec3 e, Vec3 b, Vec3& mom, Vec3& vel, float mass, float charge,
                                          const PushConsts& k) {
    float const QoM = charge / mass;
    float const deltaT = k.dt;
    float const hq = 0.5f * charge;
    Vec3 const mom_minus{mom.x + hq * e.x * deltaT, mom.y + hq * e.y * deltaT, mom.z + hq * e.z * deltaT};
    float const gamma_reci = 1.0f / gammaOf(mom_minus, mass, k.c2);
    float const hqm = 0.5f * QoM;
    Vec3 const t{hqm * b.x * gamma_reci * deltaT, hqm * b.y * gamma_reci * deltaT, hqm * b.z * gamma_reci * deltaT};
    float const sfac = 1.0f / (1.0f + l2norm2(t));
    Vec3 const s{2.0f * t.x * sfac, 2.0f * t.y * sfac, 2.0f * t.z * sfac};
    Vec3 const cp = cross(mom_minus, t);
    Vec3 const mom_prime{mom_minus.x + cp.x, mom_minus.y + cp.y, mom_minus.z + cp.z};
    Vec3 const cs = cross(mom_prime, s);
    Vec3 const mom_plus{mom_minus.x + cs.x, mom_minus.y + cs.y, mom_minus.z + cs.z};
    mom = Vec3{mom_plus.x + hq * e.x * deltaT, mom_plus.y + hq * e.y * deltaT, mom_plus.z + hq * e.z * deltaT};
    vel = velocityOf(mom, mass, k.rc2);
}

// Vay (particles/pusher/particlePusherVay.hpp:42-111); the gamma root section
// in FP64 as param/pusher.param:62 selects (sqrt_Vay = precision64Bit).
__device__ __forceinline__ void pushVay(Vec3 e, Vec3 b, Vec3& mom, Vec3& vel, float mass, float charge,
                                        const PushConsts& k) {
    float const deltaT = k.dt;
    float const factor = (float)(0.5 * (double)charge * (double)deltaT);
    Vec3 const v_mh = velocityOf(mom, mass, k.rc2);
    Vec3 const cvb = cross(v_mh, b);
    Vec3 const mom0{mom.x + factor * (e.x + cvb.x), mom.y + factor * (e.y + cvb.y), mom.z + factor * (e.z + cvb.z)};
    Vec3 const mp{mom0.x + factor * e.x, mom0.y + factor * e.y, mom0.z + factor * e.z};
    float const gamma_prime = gammaOf(mp, mass, k.c2);
    float const fm = factor / mass;
    double const tx = (double)(fm * b.x), ty = (double)(fm * b.y), tz = (double)(fm * b.z);
    double dotp = (double)mp.x * tx;
    dotp += (double)mp.y * ty;
    dotp += (double)mp.z * tz;
    double const u_star = dotp / (double)(k.c * mass);
    double tau2 = tx * tx;
    tau2 += ty * ty;
    tau2 += tz * tz;
    double const sigma = (double)(gamma_prime * gamma_prime) - tau2;
    double const gamma_ph = sqrt(0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))));
    double const rg = 1.0 / gamma_ph;
    Vec3 const t{(float)(tx * rg), (float)(ty * rg), (float)(tz * rg)};
    float const s = 1.0f / (1.0f + l2norm2(t));
    float dpt = mp.x * t.x;
    dpt += mp.y * t.y;
    dpt += mp.z * t.z;
    Vec3 const cpt = cross(mp, t);
    mom = Vec3{s * (mp.x + dpt * t.x + cpt.x), s * (mp.y + dpt * t.y + cpt.y), s * (mp.z + dpt * t.z + cpt.z)};
    vel = velocityOf(mom, mass, k.rc2);
}

// ---------------------------------------------------------------------------
// Higuera-Cary (particlePusherHigueraCary.hpp:61-142, with the WarpX
// correction mom_minus -> mom_plus in the second half acceleration), the
// auxiliary quantities in FP64 (param/pusher.param:71 sqrt_HigueraCary =
// precision64Bit); Gamma<> stays FP32 (Gamma.def:31, Gamma.hpp:28-38).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pushHigueraCary(Vec3 e, Vec3 b, Vec3& mom, Vec3& vel, float mass, float charge,
                                                const PushConsts& k) {
    float const deltaT = k.dt;
    float const hq = 0.5f * charge;
    Vec3 const de{hq * e.x * deltaT, hq * e.y * deltaT, hq * e.z * deltaT};
    Vec3 const mmf{mom.x + de.x, mom.y + de.y, mom.z + de.z};
    double const mx = mmf.x, my = mmf.y, mz = mmf.z;
    double const gamma_minus = (double)gammaOf(mmf, mass, k.c2);
    double const tx = (double)(0.5f * b.x * charge * deltaT / mass);
    double const ty = (double)(0.5f * b.y * charge * deltaT / mass);
    double const tz = (double)(0.5f * b.z * charge * deltaT / mass);
    double tau2 = tx * tx;
    tau2 += ty * ty;
    tau2 += tz * tz;
    double const sigma = gamma_minus * gamma_minus - tau2;
    double dmt = mx * tx;
    dmt += my * ty;
    dmt += mz * tz;
    double const u_star = dmt / (double)(mass * k.c);
    double const gamma_plus = sqrt(0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))));
    double const t0 = tx / gamma_plus, t1 = ty / gamma_plus, t2 = tz / gamma_plus;
    double tt = t0 * t0;
    tt += t1 * t1;
    tt += t2 * t2;
    double const s = 1.0 / (1.0 + tt);
    double dm = mx * t0;
    dm += my * t1;
    dm += mz * t2;
    double const px = s * (mx + dm * t0 + (my * t2 - mz * t1));
    double const py = s * (my + dm * t1 + (mz * t0 - mx * t2));
    double const pz = s * (mz + dm * t2 + (mx * t1 - my * t0));
    Vec3 const d2{(float)(py * t2 - pz * t1), (float)(pz * t0 - px * t2), (float)(px * t1 - py * t0)};
    mom = Vec3{(float)px + (de.x + d2.x), (float)py + (de.y + d2.y), (float)pz + (de.z + d2.z)};
    vel = velocityOf(mom, mass, k.rc2);
}

// ---------------------------------------------------------------------------
// moveParticle (particles/MoveParticle.hpp:47-160). pos in [-1,2) per axis on
// entry, [0,1) on exit; dir receives the cell step per axis; returns the new
// local cell index (wrapped into the destination supercell) and multiMask.
// ---------------------------------------------------------------------------
// the position part of moveParticle: pos in [-1,2) per axis to [0,1), the
// cell step per axis into dir
__device__ __forceinline__ void moveAxes(float* pos, int* dir) {
    // The reference marks `shift` volatile on HIP to keep (p-0.5)+0.5 from
    // being folded; the asm barrier below has the same effect.
    float shift = 0.5f;
    asm volatile("" : "+v"(shift));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float p = pos[i] - shift;
        float moveDir = 0.0f;
        if (p < -0.5f) moveDir = -1.0f;
        if (p >= 0.5f) moveDir = 1.0f;
        p -= moveDir;
        pos[i] = p + shift;
        dir[i] = (int)moveDir;
    }
}

// the cell part of moveParticle: local cell index after the steps dir
// (wrapped into the destination supercell) and the multiMask towards it
__device__ __forceinline__ void cellAfterMove(int localIdx, const int* dir, int& newIdx, int& newMask) {
    const int sc[3] = {SCX, SCY, SCZ};
    newIdx = localIdx;
    newMask = 1;
    if ((dir[0] | dir[1] | dir[2]) != 0) {
        int lc[3] = {localIdx % SCX, (localIdx / SCX) % SCY, localIdx / (SCX * SCY)};
        int ex[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            lc[i] += dir[i];
            ex[i] = (unsigned)lc[i] >= (unsigned)sc[i] ? dir[i] : 0;
            lc[i] -= ex[i] * sc[i];
        }
        newIdx = lc[0] + SCX * (lc[1] + SCY * lc[2]);
        int exchangeType = 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            newMask += (ex[i] == -1 ? 2 : ex[i]) * exchangeType;
            exchangeType *= 3;
        }
    }
}

__device__ __forceinline__ void moveParticle(float* pos, int localIdx, int* dir, int& newIdx, int& newMask) {
    moveAxes(pos, dir);
    cellAfterMove(localIdx, dir, newIdx, newMask);
}

// boundary::ReflectParticleIfOutside (boundary/Reflecting.hpp:38-66) with
// offset 0, as boundary::apply runs it over the exchanges +x, -x, +y, -y, +z,
// -z (Apply.hpp, Utility.hpp getAllAxisAlignedExchanges): for each axis in
// `reflect` whose border the particle's cell `cell` (after the move) lies
// beyond, the in-cell position is mirrored (-pos - epsilon above the upper
// border, 2 - pos below the lower one), the momentum and velocity components
// are negated (Velocity is odd in each momentum component) and moveParticle's
// position arithmetic runs on all three axes; dir accumulates the steps.
// n: the grid's cells per axis. Returns whether anything was reflected.
// dp, the move in cells per axis (vel * dt / cell), is negated with the
// velocity: (-v dt) / cell is -(v dt / cell) exactly.
__device__ __forceinline__ bool reflectParticle(const int* n, int reflect, int* cell, float* pos, Vec3& mom, Vec3& vel,
                                                int* dir, float* dp) {
    bool any = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!((reflect >> d) & 1)) continue;
        bool const above = cell[d] >= n[d], below = cell[d] < 0;
        if (!(above || below)) continue;
        pos[d] = above ? -pos[d] - 1.1920929e-7f : 2.0f - pos[d];  // std::numeric_limits<float>::epsilon()
        dp[d] = -dp[d];
        if (d == 0) {
            mom.x = -mom.x;
            vel.x = -vel.x;
        } else if (d == 1) {
            mom.y = -mom.y;
            vel.y = -vel.y;
        } else {
            mom.z = -mom.z;
            vel.z = -vel.z;
        }
        int step[3];
        moveAxes(pos, step);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            dir[i] += step[i];
            cell[i] += step[i];
        }
        any = true;
    }
    return any;
}

// ---------------------------------------------------------------------------
// Esirkepov<Shape> 3-D current deposition
// (fields/currentDeposition/Esirkepov/Esirkepov.hpp:61-242, relayPoint.hpp:47-64,
// bitPacking.hpp:37-90, TrajectoryAssignmentShapeFunction.hpp, Line.hpp,
// PermutatedFieldValueAccess.hpp). Shape strategy (Strategy.def detail::ShapeStrategy):
// outer loop uses the exact ("Jit") shape, middle/inner loops the cached arrays.
// The per-cell value added is the FP32 prefix sum `accumulated_J` of the
// reference; it is accumulated here as an int64 fixed-point number so the sum
// over particles does not depend on their order.
// ---------------------------------------------------------------------------
struct DepositConsts {
    float cell[3];
    float dt;
    float inv_vdt;  // 1 / float(V*dt)
    double jscale;  // 2^e
};

// Round-to-nearest-even integer of acc*2^e as two's complement int64, for
// |acc*2^e| < 2^51: the sum with 1.5*2^52 puts the integer in the low mantissa
// bits (one rounding, ties to even), and removing the constant only touches
// the high word because its low word is zero.
__device__ __forceinline__ unsigned long long toFixed(float acc, double jscale) {
    double const d = __fma_rn((double)acc, jscale, 6755399441055744.0);
    unsigned long long const b = (unsigned long long)__double_as_longlong(d);
    return b - 0x4338000000000000ull;
}

// The bits of the same double, magic constant included: for |acc*2^e| < 2^51
// they are 0x4338000000000000 + round(acc*2^e) as 64-bit integers. A tile
// element adds N of them; the constant is a multiple of 2^51, so the sum's
// low 51 bits are those of the sum S of the rounded values, and S follows
// from them (fixedTileValue) while |S| < 2^50. The exponent e keeps one
// contribution below 2^30; a supercell holds at most 255 frames (below 2^16
// particles) of each of at most 4 species, and a particle adds at most two
// values to one element (EmZ's two virtual particles), so |S| < 2^49 always;
// one instruction per contribution less than subtracting the constant.
__device__ __forceinline__ unsigned long long toFixedBits(float acc, double jscale) {
    return (unsigned long long)__double_as_longlong(__fma_rn((double)acc, jscale, 6755399441055744.0));
}

// S from the sum of N toFixedBits values: the low 51 bits, sign-extended
__device__ __forceinline__ long long fixedTileValue(unsigned long long bits) {
    return (long long)(bits << 13) >> 13;
}

// Geometry of a current tile: the supercell plus LO cells below and UP
// above on every axis, one box per component (planes [z][y][x], the
// components one after another). A component reaches one cell less above
// along its own axis: its innermost deposit loop (the prefix sum) stops one
// grid point short of the others.
template <int LO, int UP>
struct JTileGeom {
    static constexpr int ext(int c, int a) { return (a == 0 ? SCX : (a == 1 ? SCY : SCZ)) + LO + (a == c ? UP - 1 : UP); }
    static constexpr int py(int c) { return ext(c, 0); }
    static constexpr int pz(int c) { return ext(c, 0) * ext(c, 1); }
    static constexpr int vol(int c) { return ext(c, 0) * ext(c, 1) * ext(c, 2); }
    static constexpr int off(int c) { return c == 0 ? 0 : (c == 1 ? vol(0) : vol(0) + vol(1)); }
    static constexpr int total() { return vol(0) + vol(1) + vol(2); }
};

template <int SUPP>
struct Esirkepov {
    static constexpr int begin = Shape<SUPP>::begin;
    static constexpr int end = begin + SUPP;
    static constexpr int lowerMargin = SUPP / 2 + 1 - (SUPP + 1) % 2;  // Esirkepov.hpp:42
    static constexpr int upperMargin = (SUPP + 1) / 2 + 1;             // Esirkepov.hpp:43

    // One rotated 1-D pass (cptCurrent1D). p0/p1 are the rotated start/end
    // positions, st the rotated status bits. `jt` points at the tile element
    // of rotated offset (begin,begin,begin); SI/SJ/SK are the tile strides of
    // the rotated i/j/k axes, so every address is jt + a compile-time offset.
    // GENERAL = false is the form for a particle that stays in its (shifted)
    // cell on every axis: then st = 6 on every axis, the leaveCell rows are
    // absent and the cached shapes need no shift, so the loops visit exactly
    // the cells the general form gives a contribution, with the same values.
    template <int SI, int SJ, int SK, bool GENERAL>
    __device__ __forceinline__ static void cpt1D(unsigned long long* __restrict__ jt, const float* p0, const float* p1,
                                                 const int* st, float csd, double jscale) {
        if (p0[2] == p1[2]) return;
        float s0j[SUPP + 1], s1j[SUPP + 1], s0k[SUPP + 1], s1k[SUPP + 1];
        cachedShape<SUPP>(p0[1], GENERAL && !(st[1] & 2), s0j);
        cachedShape<SUPP>(p1[1], GENERAL && !(st[1] & 4), s1j);
        cachedShape<SUPP>(p0[2], GENERAL && !(st[2] & 2), s0k);
        cachedShape<SUPP>(p1[2], GENERAL && !(st[2] & 4), s1k);
        int const leaveCellI = GENERAL ? st[0] & 1 : 0;
        int const leaveCellJ = GENERAL ? st[1] & 1 : 0;
        int const leaveCellK = GENERAL ? st[2] & 1 : 0;
#pragma unroll
        for (int i = begin; i < end + 1; ++i) {
            if (i < end + leaveCellI) {
                float const s0i = Shape<SUPP>::assign((float)i - p0[0]);
                float const dsi = Shape<SUPP>::assign((float)i - p1[0]) - s0i;
#pragma unroll
                for (int j = begin; j < end + 1; ++j) {
                    if (j < end + leaveCellJ) {
                        float const sj0 = s0j[j - begin];
                        float const dsj = s1j[j - begin] - sj0;
                        float const tmp = -csd * esirkepovWeight(s0i, dsi, sj0, dsj);
                        float accumulated_J = 0.0f;
#pragma unroll
                        for (int k = begin; k < end; ++k) {
                            if (k < end + leaveCellK - 1) {
                                float const dsk = s1k[k - begin] - s0k[k - begin];
                                // 0 + W differs from W only in the sign of a zero, which the
                                // fixed-point conversion does not see: start the sum at W
                                // and add every further W = dsk * tmp as a multiply-add
                                accumulated_J = k == begin ? dsk * tmp : mad(dsk, tmp, accumulated_J);
                                __hip_atomic_fetch_add(jt + ((i - begin) * SI + (j - begin) * SJ + (k - begin) * SK),
                                                       toFixedBits(accumulated_J, jscale), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                }
            }
        }
    }

    // Start/end points of the move relative to the grid shift, and the status
    // bits per axis (bit 0: the shifted cell changes, bits 1/2: start/end lie
    // in the shifted cell gs).
    struct Geometry {
        float p0[3], p1[3];
        int gs[3], st[3];
    };

    // dp: the move in cells per axis, vel * dt / cell as Esirkepov.hpp:82
    // forms it (the push forms the same quotient for the position update)
    __device__ __forceinline__ static Geometry geometryDp(Vec3 pos, const float* dp) {
        constexpr bool isEven = (SUPP % 2) == 0;
        float const np[3] = {pos.x, pos.y, pos.z};
        Geometry G;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            G.p0[d] = np[d] - dp[d];
            G.p1[d] = np[d];
            int iS, iE;
            if (isEven) {
                iS = (int)floorf(G.p0[d]);
                iE = (int)floorf(G.p1[d]);
            } else {
                iS = (int)floorf(G.p0[d] + 0.5f);
                iE = (int)floorf(G.p1[d] + 0.5f);
            }
            G.gs[d] = iS < iE ? iS : iE;
            G.st[d] = (G.gs[d] == iS ? 2 : 0) | (G.gs[d] == iE ? 4 : 0) | (iS != iE ? 1 : 0);
            G.p0[d] -= (float)G.gs[d];
            G.p1[d] -= (float)G.gs[d];
        }
        return G;
    }

    __device__ __forceinline__ static Geometry geometry(Vec3 pos, Vec3 vel, const DepositConsts& k) {
        float const dp[3] = {vel.x * k.dt / k.cell[0], vel.y * k.dt / k.cell[1], vel.z * k.dt / k.cell[2]};
        return geometryDp(pos, dp);
    }

    // true when the move changes the shifted cell on some axis (the general form is needed)
    __device__ __forceinline__ static bool changesCell(const Geometry& G) {
        return ((G.st[0] | G.st[1] | G.st[2]) & 1) != 0;
    }

    // Shape values of one axis for a particle whose start and end lie in the
    // shifted cell (st = 6): the cached arrays (onSupport at the start and
    // end position, cachedShape without shift) and the outer loop's exact
    // ("Jit") values assign(i - p). For such positions assign(i - p) takes a
    // fixed branch for every i in [begin, end) and equals the cached value,
    // the same operations on the same float, for every i but i = 1, whose
    // value is Shape::assignAt1(p). So an axis needs its two cached arrays
    // and two assignAt1 values, shared by the three components.
    struct StayAxis {
        float s0[SUPP], s1[SUPP];  // grid points begin..end-1
        float j0, j1;              // assign(1 - p0), assign(1 - p1)
    };

    __device__ __forceinline__ static StayAxis stayAxis(float p0, float p1) {
        StayAxis a;
        Shape<SUPP>::onSupport(p0, a.s0);
        Shape<SUPP>::onSupport(p1, a.s1);
        a.j0 = Shape<SUPP>::assignAt1(p0);
        a.j1 = Shape<SUPP>::assignAt1(p1);
        return a;
    }

    // cpt1D for a particle that stays in its shifted cell, with the rotated
    // axes' shape values given: the cells and values of cpt1D<..., false>.
    template <int SI, int SJ, int SK>
    __device__ __forceinline__ static void cpt1DStay(unsigned long long* __restrict__ jt, const StayAxis& ai,
                                                     const StayAxis& aj, const StayAxis& ak, float p0k, float p1k,
                                                     float csd, double jscale) {
        if (p0k == p1k) return;
        float dsk[SUPP - 1];
#pragma unroll
        for (int k = 0; k < SUPP - 1; ++k) dsk[k] = ak.s1[k] - ak.s0[k];
#pragma unroll
        for (int i = begin; i < end; ++i) {
            float const s0i = i == 1 ? ai.j0 : ai.s0[i - begin];
            float const dsi = (i == 1 ? ai.j1 : ai.s1[i - begin]) - s0i;
#pragma unroll
            for (int j = begin; j < end; ++j) {
                float const sj0 = aj.s0[j - begin];
                float const dsj = aj.s1[j - begin] - sj0;
                float const tmp = -csd * esirkepovWeight(s0i, dsi, sj0, dsj);
                float accumulated_J = 0.0f;
#pragma unroll
                for (int k = begin; k < end - 1; ++k) {
                    accumulated_J = k == begin ? dsk[k - begin] * tmp : mad(dsk[k - begin], tmp, accumulated_J);
                    __hip_atomic_fetch_add(jt + ((i - begin) * SI + (j - begin) * SJ + (k - begin) * SK),
                                           toFixedBits(accumulated_J, jscale), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }

    // Deposit of a particle with changesCell(G) == false: the contributions
    // of depositGeometry<..., false>, bit for bit, with each axis' shape
    // values formed once.
    // tile element of component c for the cell (x, y, z) in tile coordinates
    template <class TG>
    __device__ __forceinline__ static unsigned long long* cornerOf(unsigned long long* jt, int c, int x, int y, int z) {
        return jt + TG::off(c) + (z * TG::ext(c, 1) + y) * TG::ext(c, 0) + x;
    }

    template <class TG>
    __device__ __forceinline__ static void depositStay(unsigned long long* __restrict__ jt, int cx, int cy, int cz,
                                                       const Geometry& G, float charge, const DepositConsts& k) {
        int const x = cx + G.gs[0] + begin, y = cy + G.gs[1] + begin, z = cz + G.gs[2] + begin;
        StayAxis const ax = stayAxis(G.p0[0], G.p1[0]);
        StayAxis const ay = stayAxis(G.p0[1], G.p1[1]);
        StayAxis const az = stayAxis(G.p0[2], G.p1[2]);
        float const cq = charge * k.inv_vdt;
        // J_x: i->y, j->z, k->x; J_y: i->z, j->x, k->y; J_z: i->x, j->y, k->z
        cpt1DStay<TG::py(0), TG::pz(0), 1>(cornerOf<TG>(jt, 0, x, y, z), ay, az, ax, G.p0[0], G.p1[0],
                                           cq * k.cell[0], k.jscale);
        cpt1DStay<TG::pz(1), 1, TG::py(1)>(cornerOf<TG>(jt, 1, x, y, z), az, ax, ay, G.p0[1], G.p1[1],
                                           cq * k.cell[1], k.jscale);
        cpt1DStay<1, TG::py(2), TG::pz(2)>(cornerOf<TG>(jt, 2, x, y, z), ax, ay, az, G.p0[2], G.p1[2],
                                           cq * k.cell[2], k.jscale);
    }

    // Deposit one particle. (cx,cy,cz) is the particle's cell in J-tile
    // coordinates, G its move geometry. Tile layout: component-major planes
    // of TV elements, [z][y][x] with pitches TX, TX*TY. GENERAL = false only
    // for a geometry with changesCell(G) == false.
    template <class TG, bool GENERAL>
    __device__ __forceinline__ static void depositGeometry(unsigned long long* __restrict__ jt, int cx, int cy, int cz,
                                                           const Geometry& G, float charge, const DepositConsts& k) {
        // tile element of cell (gridShift + begin) in every axis
        int const x = cx + G.gs[0] + begin, y = cy + G.gs[1] + begin, z = cz + G.gs[2] + begin;
        // J_x: rotation (1,2,0): i->y, j->z, k->x
        {
            float const q0[3] = {G.p0[1], G.p0[2], G.p0[0]};
            float const q1[3] = {G.p1[1], G.p1[2], G.p1[0]};
            int const s3[3] = {G.st[1], G.st[2], G.st[0]};
            float const csd = charge * k.inv_vdt * k.cell[0];
            cpt1D<TG::py(0), TG::pz(0), 1, GENERAL>(cornerOf<TG>(jt, 0, x, y, z), q0, q1, s3, csd, k.jscale);
        }
        // J_y: rotation (2,0,1): i->z, j->x, k->y
        {
            float const q0[3] = {G.p0[2], G.p0[0], G.p0[1]};
            float const q1[3] = {G.p1[2], G.p1[0], G.p1[1]};
            int const s3[3] = {G.st[2], G.st[0], G.st[1]};
            float const csd = charge * k.inv_vdt * k.cell[1];
            cpt1D<TG::pz(1), 1, TG::py(1), GENERAL>(cornerOf<TG>(jt, 1, x, y, z), q0, q1, s3, csd, k.jscale);
        }
        // J_z: identity: i->x, j->y, k->z
        {
            float const csd = charge * k.inv_vdt * k.cell[2];
            cpt1D<1, TG::py(2), TG::pz(2), GENERAL>(cornerOf<TG>(jt, 2, x, y, z), G.p0, G.p1, G.st, csd, k.jscale);
        }
    }

    // Deposit one particle. (cx,cy,cz) is the particle's cell in J-tile
    // coordinates, pos its in-cell position after the move, vel its velocity.
    template <class TG>
    __device__ __forceinline__ static void deposit(unsigned long long* __restrict__ jt, int cx, int cy, int cz, Vec3 pos,
                                                   Vec3 vel, float charge, const DepositConsts& k) {
        depositGeometry<TG, true>(jt, cx, cy, cz, geometry(pos, vel, k), charge, k);
    }
};

// ---------------------------------------------------------------------------
// EmZ current deposition (currentDeposition/EmZ/EmZ.hpp:61-137,
// EmZ/DepositCurrent.hpp:46-112): the move splits at the relay point
// (relayPoint.hpp:47-64) into one or two virtual particles, each inside one
// assignment cell, deposited with on-support shapes (Jit outer, Cached
// middle and inner) over SUPP x SUPP x (SUPP-1) cells; same fixed-point tile
// as Esirkepov.
// ---------------------------------------------------------------------------
template <int SUPP>
struct EmZ {
    static constexpr int begin = Shape<SUPP>::begin;
    static constexpr int end = begin + SUPP;

    template <int SI, int SJ, int SK>
    __device__ __forceinline__ static void cpt1D(unsigned long long* __restrict__ jt, const float* p0, const float* p1,
                                                 float csd, double jscale) {
        if (p0[2] == p1[2]) return;
        float s0j[SUPP], s1j[SUPP], s0k[SUPP], s1k[SUPP];
        Shape<SUPP>::onSupport(p0[1], s0j);
        Shape<SUPP>::onSupport(p1[1], s1j);
        Shape<SUPP>::onSupport(p0[2], s0k);
        Shape<SUPP>::onSupport(p1[2], s1k);
#pragma unroll
        for (int i = begin; i < end; ++i) {
            float const s0i = Shape<SUPP>::onSupportAt((float)i - p0[0]);
            float const dsi = Shape<SUPP>::onSupportAt((float)i - p1[0]) - s0i;
#pragma unroll
            for (int j = begin; j < end; ++j) {
                float const sj0 = s0j[j - begin];
                float const dsj = s1j[j - begin] - sj0;
                float const tmp = -csd * esirkepovWeight(s0i, dsi, sj0, dsj);
                float accumulated_J = 0.0f;
#pragma unroll
                for (int k = begin; k < end - 1; ++k) {
                    float const dsk = s1k[k - begin] - s0k[k - begin];
                    accumulated_J = k == begin ? dsk * tmp : mad(dsk, tmp, accumulated_J);  // as in Esirkepov::cpt1D
                    __hip_atomic_fetch_add(jt + ((i - begin) * SI + (j - begin) * SJ + (k - begin) * SK),
                                           toFixedBits(accumulated_J, jscale), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }

    // one virtual particle from a to b (positions relative to shifted cell s)
    template <class TG>
    __device__ __forceinline__ static void virtualParticle(unsigned long long* __restrict__ jt, int cx, int cy, int cz,
                                                           const int* s, const float* a, const float* b,
                                                           float chargeDensity, const DepositConsts& k) {
        int const x = cx + s[0] + begin, y = cy + s[1] + begin, z = cz + s[2] + begin;
        {  // J_x: rotation (1,2,0)
            float const q0[3] = {a[1], a[2], a[0]};
            float const q1[3] = {b[1], b[2], b[0]};
            cpt1D<TG::py(0), TG::pz(0), 1>(Esirkepov<SUPP>::template cornerOf<TG>(jt, 0, x, y, z), q0, q1,
                                           k.cell[0] * chargeDensity / k.dt, k.jscale);
        }
        {  // J_y: rotation (2,0,1)
            float const q0[3] = {a[2], a[0], a[1]};
            float const q1[3] = {b[2], b[0], b[1]};
            cpt1D<TG::pz(1), 1, TG::py(1)>(Esirkepov<SUPP>::template cornerOf<TG>(jt, 1, x, y, z), q0, q1,
                                           k.cell[1] * chargeDensity / k.dt, k.jscale);
        }
        cpt1D<1, TG::py(2), TG::pz(2)>(Esirkepov<SUPP>::template cornerOf<TG>(jt, 2, x, y, z), a, b,
                                       k.cell[2] * chargeDensity / k.dt, k.jscale);  // J_z
    }

    // Deposit one particle: (cx,cy,cz) its cell in J-tile coordinates, pos
    // its in-cell position after the move, vel its velocity.
    template <class TG>
    __device__ __forceinline__ static void deposit(unsigned long long* __restrict__ jt, int cx, int cy, int cz, Vec3 pos,
                                                   Vec3 vel, float charge, const DepositConsts& k) {
        float const dp[3] = {vel.x * k.dt / k.cell[0], vel.y * k.dt / k.cell[1], vel.z * k.dt / k.cell[2]};
        depositDp<TG>(jt, cx, cy, cz, pos, dp, charge, k);
    }

    // the same with the move in cells dp = vel * dt / cell given
    template <class TG>
    __device__ __forceinline__ static void depositDp(unsigned long long* __restrict__ jt, int cx, int cy, int cz,
                                                     Vec3 pos, const float* dp, float charge, const DepositConsts& k) {
        constexpr bool isEven = (SUPP % 2) == 0;
        float const p1[3] = {pos.x, pos.y, pos.z};
        float p0[3], relay[3];
        int s0[3], s1[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            p0[d] = p1[d] - dp[d];
            if (isEven) {
                s0[d] = (int)floorf(p0[d]);
                s1[d] = (int)floorf(p1[d]);
                relay[d] = s0[d] == s1[d] ? p1[d] : (float)(s0[d] > s1[d] ? s0[d] : s1[d]);
            } else {
                s0[d] = (int)floorf(p0[d] + 0.5f);
                s1[d] = (int)floorf(p1[d] + 0.5f);
                relay[d] = s0[d] == s1[d] ? p1[d] : (float)(s0[d] + s1[d]) / 2.0f;
            }
        }
        float const chargeDensity = charge / (k.cell[0] * k.cell[1] * k.cell[2]);
        float a[3], b[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a[d] = p0[d] - (float)s0[d];
            b[d] = relay[d] - (float)s0[d];
        }
        virtualParticle<TG>(jt, cx, cy, cz, s0, a, b, chargeDensity, k);
        if (s0[0] != s1[0] || s0[1] != s1[1] || s0[2] != s1[2]) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                b[d] = p1[d] - (float)s1[d];
                a[d] = relay[d] - (float)s1[d];
            }
            virtualParticle<TG>(jt, cx, cy, cz, s1, a, b, chargeDensity, k);
        }
    }
};

}  // namespace pic
