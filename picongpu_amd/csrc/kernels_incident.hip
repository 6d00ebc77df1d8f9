// Incident field through a Huygens surface: ApplyIncidentFieldKernel with
// UpdateFunctor (fields/incidentField/Solver.kernel:60-330) and the
// GaussianPulse / PlaneWave functors (profiles/GaussianPulse.hpp:80-250,
// profiles/PlaneWave.hpp:60-140, Functors.hpp:60-570), FP32 in the
// reference's operation order.
//
// One thread per updated cell of a face; the cells of one launch are
// distinct, so each adds its correction to E or B without atomics. The
// launches of a step run in the reference's face order on one stream.
//
// exp, cos and atan of FP32 arguments are evaluated as the FP64 function
// rounded to FP32 (within half an ulp of the exact value but for rare
// double-rounding cases); the reference calls the platform's FP32 functions
// here. The PlaneWave envelope is FP64 as in the reference, whose RAMP_INIT
// is a float_64 (PlaneWave.def:44).
#include <hip/hip_runtime.h>

#include "pic_incident.hpp"

namespace pic {
namespace {

__device__ __forceinline__ float fexp(float x) { return static_cast<float>(exp(static_cast<double>(x))); }
__device__ __forceinline__ float fcos(float x) { return static_cast<float>(cos(static_cast<double>(x))); }
__device__ __forceinline__ float fatan(float x) { return static_cast<float>(atan(static_cast<double>(x))); }

constexpr float kPiF = 3.14159265358979323846f;  // pmacc Pi<float_X>::value
constexpr float kTwoPiF = 2.0f * kPiF;            // Pi<float_X>::doubleValue
constexpr float kHalfPiF = kPiF / 2.0f;           // Pi<float_X>::halfValue

// pmacc::math::dot (Vector.tpp:80-93): a0 b0, then += a1 b1, += a2 b2
__device__ __forceinline__ float dot3(const float* a, const float* b) {
    float r = a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}

// BaseFunctorE::getCurrentTime (Functors.hpp:103-110)
__device__ __forceinline__ float currentTime(const IncidentProfile& P, float tOrigin, const float* idx) {
    float s[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] = idx[d] * P.cell[d] - P.origin[d];
    float const distance = dot3(s, P.axis0);
    float const timeDelay = distance / P.phaseVelocity + P.timeDelay;
    return tOrigin - timeDelay;
}

// GaussianPulseFunctorIncidentE::simpleLaguerre (GaussianPulse.hpp:230-250)
__device__ __forceinline__ float simpleLaguerre(int n, float x) {
    if (n == 0) return 1.0f;
    int cur = 1;
    float lm1 = 1.0f;
    float l = 1.0f - x;
    float lp1 = 0.0f;
    while (cur < n) {
        lp1 = ((2.0f * static_cast<float>(cur) + 1.0f - x) * l - static_cast<float>(cur) * lm1) /
              static_cast<float>(cur + 1);
        lm1 = l;
        l = lp1;
        ++cur;
    }
    return l;
}

// GaussianPulseFunctorIncidentE::getValue (GaussianPulse.hpp:130-228), 3-D,
// no pulse-front tilt, GaussianPulseEnvelope (GaussianPulse.hpp:270-290)
__device__ float gaussianValue(const IncidentProfile& P, float tOrigin, const float* idx, float phaseShift) {
    float s[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] = idx[d] * P.cell[d] - P.origin[d];
    float pos[3] = {dot3(s, P.axis0), dot3(s, P.axis1), dot3(s, P.axis2)};
    float time = currentTime(P, tOrigin, idx);
    if (time < 0.0f) return 0.0f;
    time += P.timeShift;
    float fr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) fr[d] = P.focus[d] - P.origin[d];
    float const distanceFocus = dot3(fr, P.axis0);
    float const focusPos = distanceFocus - pos[0];
    float const w = P.w0 * sqrtf(1.0f + (focusPos / P.rayleighLength) * (focusPos / P.rayleighLength));
    float const phase = P.omega * (time - focusPos / P.c) + P.laserPhase + phaseShift;
    float const q0 = pos[0] * 0.0f, q1 = pos[1] * 1.0f, q2 = pos[2] * 1.0f;
    float trans = q0 * q0;
    trans += q1 * q1;
    trans += q2 * q2;
    float const rInv = -focusPos / (P.rayleighLength * P.rayleighLength + focusPos * focusPos);
    float const xi = fatan(-focusPos / P.rayleighLength);
    float etrans = 0.0f;
    float const r2OverW2 = trans / w / w;
    float const r = 0.5f * trans * rInv;
    for (int m = 0; m < P.numModes; ++m) {
        float const arg = kTwoPiF / P.waveLength * focusPos - kTwoPiF / P.waveLength * r +
                          (2.0f * static_cast<float>(m) + 1.0f) * xi + phase + P.phases[m];
        etrans += P.modes[m] * simpleLaguerre(m, 2.0f * r2OverW2) * fexp(-r2OverW2) * fcos(arg);
    }
    float const shiftedTime = time - r / P.c;
    float const exponent = shiftedTime / (2.0f * P.pulseDuration);
    etrans *= fexp(-exponent * exponent);
    float norm = 0.0f;
    for (int m = 0; m < P.numModes; ++m) norm += P.modes[m];
    float envelope = P.amplitude;
    envelope *= P.w0 / w;
    return envelope * etrans / norm;
}

// PlaneWaveFunctorIncidentE::getLongitudinal (PlaneWave.hpp:90-125)
__device__ float planeLongitudinal(const IncidentProfile& P, float time, float phaseShift) {
    float envelope = P.amplitude;
    double const mue = 0.5 * P.rampInit * static_cast<double>(P.pulseDuration);
    float const tau = P.pulseDuration * sqrtf(2.0f);
    double const endUpramp = mue;
    double const startDownramp = mue + static_cast<double>(P.nofocus);
    float icf = 0.0f;
    double const t = static_cast<double>(time);
    if (t > startDownramp) {
        double const exponent = (t - startDownramp) / static_cast<double>(tau);
        envelope = static_cast<float>(static_cast<double>(envelope) * exp(-0.5 * exponent * exponent));
        icf = static_cast<float>((t - startDownramp) / static_cast<double>(P.omega * tau * tau));
    } else if (t < endUpramp) {
        double const exponent = (t - endUpramp) / static_cast<double>(tau);
        envelope = static_cast<float>(static_cast<double>(envelope) * exp(-0.5 * exponent * exponent));
        icf = static_cast<float>((t - endUpramp) / static_cast<double>(P.omega * tau * tau));
    }
    double const timeOszi = t - endUpramp;
    double const phase = static_cast<double>(P.omega) * timeOszi + static_cast<double>(P.laserPhase) +
                         static_cast<double>(phaseShift);
    return static_cast<float>((sin(phase) + cos(phase) * static_cast<double>(icf)) * static_cast<double>(envelope));
}

// E_inc at a total cell position (GaussianPulse operator(), GaussianPulse.hpp:112-126;
// BaseSeparableFunctorE operator() for the plane wave, Functors.hpp:262-283)
__device__ void incidentE(const IncidentProfile& P, float tOrigin, const float* idx, float* e) {
    bool const linear = P.polarisation == PIC_POLARISATION_LINEAR;
    float cp1[3], cp2[3];
    if (!linear) {
        float const s2 = sqrtf(2.0f);
#pragma unroll
        for (int d = 0; d < 3; ++d) cp1[d] = P.axis1[d] / s2;
        cp2[0] = P.axis0[1] * cp1[2] - P.axis0[2] * cp1[1];
        cp2[1] = P.axis0[2] * cp1[0] - P.axis0[0] * cp1[2];
        cp2[2] = P.axis0[0] * cp1[1] - P.axis0[1] * cp1[0];
    }
    if (P.kind == PIC_INCIDENT_GAUSSIAN_PULSE) {
        if (linear) {
            float const v = gaussianValue(P, tOrigin, idx, 0.0f);
#pragma unroll
            for (int d = 0; d < 3; ++d) e[d] = P.axis1[d] * v;
        } else {
            float const v1 = gaussianValue(P, tOrigin, idx, kHalfPiF);
            float const v2 = gaussianValue(P, tOrigin, idx, 0.0f);
#pragma unroll
            for (int d = 0; d < 3; ++d) e[d] = cp1[d] * v1 + cp2[d] * v2;
        }
    } else {
        float const time = currentTime(P, tOrigin, idx);
        if (time < 0.0f) {
            e[0] = e[1] = e[2] = 0.0f;
            return;
        }
        float const transversal = 1.0f;
        if (linear) {
            float const v = planeLongitudinal(P, time, 0.0f) * transversal;
#pragma unroll
            for (int d = 0; d < 3; ++d) e[d] = P.axis1[d] * v;
        } else {
            float const v1 = planeLongitudinal(P, time, kHalfPiF) * transversal;
            float const v2 = planeLongitudinal(P, time, 0.0f) * transversal;
#pragma unroll
            for (int d = 0; d < 3; ++d) e[d] = cp1[d] * v1 + cp2[d] * v2;
        }
    }
}

// component `comp` of the incident field: E_inc, or B_inc = cross(direction, E_inc) / c
// (ApproximateIncidentB, Functors.hpp:540-570)
__device__ float incidentComponent(const IncidentUpdate& u, float tOrigin, const float* idx, int comp) {
    float e[3];
    incidentE(u.prof, tOrigin, idx, e);
    if (!u.incidentIsB) return e[comp];
    const float* a = u.prof.axis0;
    float const b[3] = {a[1] * e[2] - a[2] * e[1], a[2] * e[0] - a[0] * e[2], a[0] * e[1] - a[1] * e[0]};
    return b[comp] / u.prof.c;
}

__global__ __launch_bounds__(256) void kIncident(IncidentUpdate u, GridDev g, float* F0, float* F1, float* F2,
                                                 const int64_t* step, float stepOffset) {
    if (halted(g)) return;
    int const ex = u.end[0] - u.begin[0], ey = u.end[1] - u.begin[1], ez = u.end[2] - u.begin[2];
    long long const n = static_cast<long long>(ex) * ey * ez;
    long long const t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int idx[3];
    idx[0] = u.begin[0] + static_cast<int>(t % ex);
    idx[1] = u.begin[1] + static_cast<int>((t / ex) % ey);
    idx[2] = u.begin[2] + static_cast<int>(t / (static_cast<long long>(ex) * ey));
    (void)ez;
    // BaseFunctorE::currentTimeOrigin = currentStep * dt
    float const currentStep = static_cast<float>(*step) + stepOffset;
    float const tOrigin = currentStep * u.prof.dt;
    int const axis = u.axis, d1 = (axis + 1) % 3, d2 = (axis + 2) % 3;
    // ApplyIncidentFieldKernel (Solver.kernel:300-325)
    bool last[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) last[d] = u.isLast[d] && idx[d] == u.end[d] - 1;
    // UpdateFunctor::operator() and getUpdatedFieldCorrection (Solver.kernel:95-235)
    bool const total = u.updatedTotal != 0;
    bool const apply1 = (total && !last[d1]) || (!total && !last[d2]);
    bool const apply2 = (total && !last[d2]) || (!total && !last[d1]);
    float base[3] = {static_cast<float>(idx[0]), static_cast<float>(idx[1]), static_cast<float>(idx[2])};
    base[axis] = static_cast<float>(u.huygens);
    float s1[3], s2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s1[d] = base[d] + u.inCellShift1[d];
        s2[d] = base[d] + u.inCellShift2[d];
    }
    int const ufs = idx[axis] - u.huygens;
    // getInitialIncidentFieldShift / getIncidentFieldShiftIncrement (Solver.kernel:237-262)
    float init = static_cast<float>(-ufs);
    if (total && static_cast<float>(ufs) * u.direction > 0.0f) init += u.direction;
    if (!total && static_cast<float>(ufs) * u.direction < 0.0f) init -= u.direction;
    float const incr = ufs < 0 ? 1.0f : (ufs > 0 ? -1.0f : (total ? -u.direction : u.direction));
    float res[3] = {0.0f, 0.0f, 0.0f};
    float shift[3] = {0.0f, 0.0f, 0.0f};
    shift[axis] = init;
    for (int as = ufs < 0 ? -ufs : ufs; as < u.margin; ++as) {
        int ci[3] = {0, 0, 0};
        ci[axis] = as;
        shift[d1] = static_cast<float>(1 - u.size1);
        for (int a1 = 1 - u.size1; a1 < u.size1; ++a1) {
            ci[d1] = a1 < 0 ? -a1 : a1;
            shift[d2] = static_cast<float>(1 - u.size2);
            for (int a2 = 1 - u.size2; a2 < u.size2; ++a2) {
                ci[d2] = a2 < 0 ? -a2 : a2;
                float const dc = u.value[ci[0]][ci[1]][ci[2]];
                if (apply1) {
                    float const p[3] = {s1[0] + shift[0], s1[1] + shift[1], s1[2] + shift[2]};
                    res[d1] += dc * incidentComponent(u, tOrigin, p, d2);
                }
                if (apply2) {
                    float const p[3] = {s2[0] + shift[0], s2[1] + shift[1], s2[2] + shift[2]};
                    res[d2] += dc * incidentComponent(u, tOrigin, p, d1);
                }
                shift[d2] += 1.0f;
            }
            shift[d1] += 1.0f;
        }
        shift[axis] += incr;
    }
    res[d1] *= u.baseCoefficient;
    res[d2] *= -u.baseCoefficient;
    size_t const at = static_cast<size_t>(idx[0]) + static_cast<size_t>(g.nx) *
                      (static_cast<size_t>(idx[1]) + static_cast<size_t>(g.ny) * static_cast<size_t>(idx[2]));
    F0[at] += res[0];
    F1[at] += res[1];
    F2[at] += res[2];
}

// end of a step: the field step counter advances; a halted simulation counts
// the step as skipped instead, and an overflow in this step halts the next ones
__global__ void kAdvanceStep(int64_t* step, StepFlags* flags) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (flags && flags->halt) {
        flags->skipped += 1u;
        return;
    }
    step[0] += 1;
    if (flags && flags->overflow) flags->halt = 1u;
}

}  // namespace

hipError_t launchIncident(const IncidentUpdate& u, const GridDev& g, float* const F[3], const int64_t* step,
                          float stepOffset, hipStream_t s) {
    long long const n = static_cast<long long>(u.end[0] - u.begin[0]) * (u.end[1] - u.begin[1]) *
                        (u.end[2] - u.begin[2]);
    if (n <= 0) return hipSuccess;
    unsigned const blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(kIncident, dim3(blocks), dim3(256), 0, s, u, g, F[0], F[1], F[2], step, stepOffset);
    return hipGetLastError();
}

hipError_t launchAdvanceStep(int64_t* step, StepFlags* flags, hipStream_t s) {
    hipLaunchKernelGGL(kAdvanceStep, dim3(1), dim3(64), 0, s, step, flags);
    return hipGetLastError();
}

}  // namespace pic
