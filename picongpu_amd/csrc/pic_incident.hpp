// Incident field through a Huygens surface (total-field / scattered-field):
// the launch descriptions shared by the host set-up and the kernel.
//
// The reference adds, around every field update, the curl of the analytic
// incident field taken across the Huygens surface to the cells next to it
// (fields/incidentField/Solver.hpp:120-258, Solver.kernel:60-330):
// update_beforeCurrent calls updateBHalf(step) after the second B half step
// and updateE(step + 0.5) before the E update (FDTDBase.hpp:97-121),
// update_afterCurrent updateBHalf(step + 1) before the first B half step
// (FDTDBase.hpp:151-183). One IncidentUpdate is one ApplyIncidentFieldKernel
// launch: one face (axis, side) and one updated field.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pic_kernels.hpp"

namespace pic {

// Constants of one profile in PIC units, rounded to FP32 as the reference's
// Unitless structs hold them (profiles/BaseParam.hpp:40-140,
// GaussianPulse.hpp:60-75, PlaneWave.hpp:60-75, Functors.hpp:60-230).
struct IncidentProfile {
    int kind;          // PIC_INCIDENT_*
    int polarisation;  // PIC_POLARISATION_*
    float origin[3];   // BaseFunctorE::origin (Functors.hpp:177-205)
    float focus[3];    // BaseFunctorE::focus (Functors.hpp:158-175)
    float axis0[3];    // propagation direction
    float axis1[3];    // polarisation direction
    float axis2[3];    // axis0 x axis1, componentwise as Functors.hpp:218-225
    float cell[3];
    float c;
    float dt;
    float phaseVelocity;  // calculatePhaseVelocity.hpp:31-40, in units of c
    float timeDelay;      // TIME_DELAY
    float waveLength, omega, amplitude, pulseDuration, laserPhase;
    // GaussianPulse
    float w0, rayleighLength, timeShift;  // timeShift = GaussianPulseEnvelope::TIME_SHIFT
    int numModes;                         // number of Laguerre modes (numModes + 1 of the param)
    float modes[PIC_INCIDENT_MAX_MODES];
    float phases[PIC_INCIDENT_MAX_MODES];
    // PlaneWave
    double rampInit;  // RAMP_INIT (FP64 in the reference, which makes the envelope FP64)
    float nofocus;    // LASER_NOFOCUS_CONSTANT
};

// One ApplyIncidentFieldKernel launch (Solver.hpp:120-258).
struct IncidentUpdate {
    IncidentProfile prof;
    int axis;              // T_axis of the face
    float direction;       // +1 at the negative border's surface, -1 at the positive one
    int updatedTotal;      // isUpdatedFieldTotal: 1 for E, 0 for B (Solver.hpp:127-128)
    int incidentIsB;       // 1: updating E from B_inc = ApproximateIncidentB (Functors.hpp:540-570)
    int margin;            // coefficient extent along the axis (Solver.kernel:123)
    int size1, size2;      // extents along dir1 = (axis+1)%3 and dir2 = (axis+2)%3
    float value[2][2][2];  // DerivativeCoefficients value[x][y][z] (DerivativeCoefficients.hpp)
    float baseCoefficient; // curlCoefficient / cell[axis] * sign(direction) (Solver.kernel:150-155)
    float inCellShift1[3]; // base shift + Yee position of incident component (axis+2)%3
    float inCellShift2[3]; // base shift + Yee position of incident component (axis+1)%3
    int begin[3], end[3];  // updated cells [begin, end), total cell indices
    int huygens;           // Huygens-surface cell along the axis: begin[axis] + margin - 1
    int isLast[3];         // isLastLocalDomain per axis
};

// F += the correction of u over its cells; the source time is
// (float(*step) + stepOffset) * dt (BaseFunctorE::currentTimeOrigin)
hipError_t launchIncident(const IncidentUpdate& u, const GridDev& g, float* const F[3], const int64_t* step,
                          float stepOffset, hipStream_t s);
// end of a step: *step += 1 (the field step counter) and the step flags
// (a halted simulation counts a skipped step; an overflow halts the next steps)
hipError_t launchAdvanceStep(int64_t* step, StepFlags* flags, hipStream_t s);

}  // namespace pic
