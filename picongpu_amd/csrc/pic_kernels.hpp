// Kernel argument structures and launcher declarations shared by the host
// runtime (pic_api.cpp) and the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/picongpu_amd.h"

namespace pic {

constexpr int kSCX = 8, kSCY = 8, kSCZ = 4, kSCV = 256;
constexpr int kNumDirs = 27;      // exchange codes 0..26 (0 = stays)
constexpr int kOffStride = 28;    // per-supercell offset table entries

struct SpeciesDev {
    float* pos[3];
    float* mom[3];
    float* w;
    uint16_t* cell;
    uint8_t* mask;
    uint32_t* count;     // particles per supercell
    float4* outbox;      // per supercell: records of the leavers in outIdx order, 2 float4 each
    int32_t* outIdx;     // per supercell: leaver slots grouped by exchange code
    int32_t* scratch;    // per supercell: gap/source lists of the compaction
    uint32_t* outOff;    // per supercell: 28 offsets into outIdx (code 1..26, [27]=total)
    uint32_t* leaveBits; // per supercell and push thread: leaver flags per loop iteration
    // supercell sc owns the run of slots [base[sc], base[sc] + capOf[sc]) of
    // every per-slot array (whole frames): supercells grow one by one, as the
    // reference's take frames from its heap (ParticlesBox.hpp:96-122)
    const uint64_t* base;
    const uint32_t* capOf;
    int32_t cap;         // the largest capacity of a supercell (slots)
    int32_t lists;       // 1: the push kernel builds the leaver lists of this species
    int32_t reflect;     // bit d: particles crossing an open border of axis d are reflected
    float mass;          // frame::getMass()
    float charge;        // frame::getCharge()
};

__device__ __forceinline__ size_t slotBase(const SpeciesDev& S, int sc) { return (size_t)S.base[sc]; }
__device__ __forceinline__ uint32_t slotCap(const SpeciesDev& S, int sc) { return S.capOf[sc]; }

// Step state of one simulation on the device. A shift that finds a supercell
// full keeps the true count, leaves the records it could not place in the
// sender's outbox and sets `overflow`; the end-of-step kernel then sets
// `halt`, and every kernel of the following steps returns at once (counting
// the skipped steps), so no particle is dropped and the host can grow the
// supercells, place the waiting records and run the skipped steps.
struct StepFlags {
    uint32_t overflow;  // a supercell of some species overflowed its capacity in a shift
    uint32_t halt;      // steps after the overflowing one are skipped
    uint32_t skipped;   // steps skipped while halted
    uint32_t pad;
};

struct GridDev {
    int nx, ny, nz;
    int nscx, nscy, nscz;
    int open;  // bit d: axis d has open borders (fields read zero beyond, particles leaving are absorbed)
    StepFlags* flags;  // step state (null: never halted)
};

// every kernel of a step returns at once while the simulation is halted
__device__ __forceinline__ bool halted(const GridDev& g) { return g.flags != nullptr && g.flags->halt != 0u; }

// (x, y, z) lies beyond the border of an open axis
__host__ __device__ __forceinline__ bool outsideOpen(const GridDev& g, int x, int y, int z) {
    return ((g.open & 1) && (x < 0 || x >= g.nx)) || ((g.open & 2) && (y < 0 || y >= g.ny)) ||
           ((g.open & 4) && (z < 0 || z >= g.nz));
}

// Weight damping of particles in the PML layers by an absorbing particle
// boundary (boundary/Absorbing.hpp dampWeightsInPml, Lehe et al. 2022):
// per open axis with layers, the damping areas (the supercells the layers
// reach) and the layer parameters of getNormalizedSigma (Pml.kernel:196-216)
struct PmlDampArgs {
    int axes;              // bit d: axis d is open and has a PML layer
    int lo[3], hi[3];      // areas: cells below lo[d] (min side), at or above hi[d] (max side)
    float negB[3], posB[3];  // layer thickness at the min and max side
    float nCells[3];       // cells along each axis
    int guard[3];          // the reference's guard cells (one supercell)
    float sigmaMax[3];     // NORMALIZED_SIGMA_MAX
    float grading;         // SIGMA_KAPPA_GRADING_ORDER
    int gradingInt;        // the order when it is an integer in [0, 32], else -1
    double slow[3];        // 1e-5 * cell size: the integral's bound for a slow particle
};

struct PushArgs {
    GridDev g;
    const float* E[3];
    const float* B[3];
    float* J[3];
    float cell[3];
    float dt, c, c2, rc2;
    float inv_vdt;
    double jscale;    // 2^e
    double jinv;      // 2^-e
    int nsp;
    SpeciesDev sp[PIC_MAX_SPECIES];
    const int32_t* scList;  // supercells of this launch, in launch order
    int stageShift;  // 1: the epilogue also moves the leavers to the outbox and compacts the supercell
    // non-null (fused push + deposit only): every supercell in one launch,
    // each leaving its J tile at tileOut[sc * tile elements], as the floats
    // the flush would add to J (zero where it adds none), for
    // launchGatherFlush instead of adding it to J
    float* tileOut;
    PmlDampArgs damp;  // axes == 0: no PML damping
};

struct FieldArgs {
    GridDev g;
    float* E[3];
    float* B[3];
    float* J[3];
    float cell[3];
    float dt, c2, coeff;
    float lehe_alpha, lehe_delta, lehe_beta1, lehe_beta2;
    int lehe_dir;
    // inner != 0: the cells in [in0, in1) per axis lie outside every PML
    // layer, and the PML kernels take the plain update there without looking
    // up the coefficients
    int inner;
    int in0[3], in1[3];
};

// launchers (return hipError_t of the launch)
// mode 0: push only, 1: push + fused deposit, 2: deposit only
// currentSolver: PIC_CURRENT_SOLVER_* of the launch's species (one per launch)
hipError_t launchPush(int supp, int pusher, int mode, int currentSolver, const PushArgs& a, int nblocks,
                      hipStream_t s);
// elements of the fused kernel's J tile for a shape support
size_t pushTileElems(int supp);
// J += the tiles of a one-launch fused push (PushArgs::tileOut), each J cell
// adding its covering supercells' tile values in the order of their
// checkerboard phases: the float additions, in the same order, of the
// phase-by-phase flush, so J is the same bit for bit. colour: the
// checkerboard colour of every supercell index along x, then y, then z
// (nscx + nscy + nscz entries).
hipError_t launchGatherFlush(int supp, const GridDev& g, const float* tiles, const uint8_t* colour,
                             float* const J[3], hipStream_t s);
// shift of one species: `staged` = the push epilogue already moved the leavers
// to the outbox and compacted every supercell (then only the insert runs).
// redo: the insert again after the supercells grew, from count - incoming
// (the count an overflowing insert left) and regardless of the halt.
hipError_t launchShift(const GridDev& g, const SpeciesDev& sp, bool staged, hipStream_t s);
hipError_t launchInsertRedo(const GridDev& g, const SpeciesDev& sp, hipStream_t s);
// count[sc] = min(count[sc], cap): the last resort when a supercell cannot grow
hipError_t launchClampCounts(const SpeciesDev& sp, int nsc, hipStream_t s);
// the runs of the supercells listed in `moves` (one entry per supercell:
// slots oldBase, count slots, newBase), copied from src to dst (may be the
// same array when no run overlaps another)
struct RunMove {
    uint64_t from, to;
    uint32_t slots, pad;
};
hipError_t launchMoveRuns(void* dst, const void* src, uint32_t bytesPerSlot, const RunMove* moves, int n,
                          hipStream_t s);
hipError_t launchCollect(const GridDev& g, const SpeciesDev& sp, hipStream_t s);
// the push kernel builds leaver lists for species with at most this many slots per supercell
constexpr int kPushThreadsPerBlock = 512;
constexpr int kPushMaxIters = 32;  // bits of a leaveBits word
constexpr int kListMaxSlots = kPushThreadsPerBlock * kPushMaxIters;
hipError_t launchFill(float* p, size_t n, float v, hipStream_t s);
// J = 0 at the start of a step (skipped while halted)
hipError_t launchCurrentReset(const GridDev& g, float* const J[3], size_t n, hipStream_t s);
hipError_t launchBHalf(int solver, const FieldArgs& a, hipStream_t s);
// PML (absorber/pml/Pml.kernel:190-590): the convolution fields psi of E and
// B (components yx, zx, xy, zy, xz, yz, one value per cell) and the
// coefficient tables of E and B positions (b, c, kappa per axis and index:
// b of x, y, z, then c, then kappa)
struct PmlArgs {
    float* psiE[6];
    float* psiB[6];
    const float* tabE;
    const float* tabB;
    // the z-marching B half step's tiles: the nPlain tiles whose cells all
    // lie outside the layers, then the nLayer others (null: every tile takes
    // the PML kernel)
    const int* tiles;
    int nPlain, nLayer;
};
// tiles of the z-marching B half step on this grid (0: the per-cell form)
int pmlTileCount(const GridDev& g);
// their split: out[0, *nPlain) the tiles inside [in0, in1) on every axis, the
// rest after them
void pmlTileSplit(const GridDev& g, const int in0[3], const int in1[3], int* out, int* nPlain);
// UpdateBHalfFunctor with PML: cells inside the layer update psi of B when
// updatePsi and add the stretched curl; the others B -= curl E * 0.5 * dt
// (the z-marching form when the grid fills the chip, as launchBHalf)
hipError_t launchBHalfPml(int solver, const FieldArgs& a, const PmlArgs& p, bool updatePsi, hipStream_t s);
// UpdateEFunctor with PML (E only; J is added by launchAddCurrent)
hipError_t launchEFullPml(const FieldArgs& a, const PmlArgs& p, hipStream_t s);
// the PML's E full step and J -> E (None) per cell in one pass into Enew
hipError_t launchEFullPmlJ(const FieldArgs& a, const PmlArgs& p, float* const Enew[3], hipStream_t s);
// E full step and J -> E (None) per cell in one pass into Enew (may be a.E)
hipError_t launchEFullJ(const FieldArgs& a, float* const Enew[3], hipStream_t s);
hipError_t launchEFull(const FieldArgs& a, hipStream_t s);
// binomial: the 27-point filter of currentInterpolation/Binomial.hpp (periodic box)
hipError_t launchAddCurrent(const FieldArgs& a, bool binomial, hipStream_t s);
hipError_t launchBinomialJ(const FieldArgs& a, float* const out[3], hipStream_t s);
// fused update_beforeCurrent + addCurrent + update_afterCurrent; the new E is
// written to Enew (the caller swaps E and Enew afterwards); the grid must be a
// multiple of 8x8x4 cells
// layers of the exponential absorber, in the reference's exchange order
struct AbsorbLayer {
    int axis, minSide;
    int depth;            // damped cells: distance 0 .. depth-1 from the border (thickness - 1)
    const float* table;   // table[f] = exp(-strength * f), f = 1 .. depth
};
struct AbsorbPlan {
    int n;
    AbsorbLayer layer[6];
};
hipError_t launchAbsorb(const GridDev& g, float* const F[3], const AbsorbPlan& p, hipStream_t s);
hipError_t launchFieldStep(int solver, const FieldArgs& a, float* const Enew[3], const AbsorbPlan* ab, hipStream_t s);
// tiles of the fused field step of one grid (work items of the batched form)
int fieldStepTiles(int solver, const GridDev& g);
// per component (B x, y, z, E x, y, z) the FP64 sum of the FP32 squares:
// nblocks partial sums (6 * nblocks doubles of scratch), then out[0..5]
hipError_t launchEnergy(const FieldArgs& a, double* partials, int nblocks, double* out, hipStream_t s);
hipError_t launchCountSum(const uint32_t* counts, int n, unsigned long long* out, hipStream_t s);
hipError_t launchGaussResidual(const FieldArgs& f, const SpeciesDev* sp, const int* supp, int nsp, double qScale,
                               double eps0, double* buf, hipStream_t s);
// The allocation-free Gauss residual's scratch: a file-scope device array of
// kNoAllocMaxCells doubles (16 MiB) per device. noAllocScratch returns its
// address on `device`, which must be the current device; the address is
// looked up once per device and kept.
constexpr size_t kNoAllocMaxCells = size_t(1) << 21;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noAllocScratch(int device, double** out);

struct InitArgs {
    GridDev g;
    SpeciesDev sp;
    uint64_t seed;
    int32_t species;
    int32_t ppc;
    int32_t drift_band;
    double weighting;    // density*V/ppc at full density
    float frameMass;     // frame::getMass()
    double u_drift[3];
    double u_th[3];
    int32_t profile;     // 0 uniform, 1 linear ramp
    int32_t axis;
    double rampStart;    // in cells
    double rampLength;   // in cells
    int32_t quiet[3];    // particles per dimension of the quiet start, 0: random positions
};
hipError_t launchInit(const InitArgs& a, hipStream_t s);
constexpr int kMaxInitPpc = 255;  // macro particles per cell of an initialisation
struct LaserArgs {
    GridDev g;
    float* E[3];
    float* B[3];
    double e0;           // peak field in PIC units
    double k;            // 2 pi / wavelength, per cell of the propagation axis... in PIC length units
    double w0, sigma;    // in PIC length units
    double center[3];    // in PIC length units
    double cell[3];      // in PIC length units
    double invc;         // 1 / c
    int32_t prop, pol;
};
hipError_t launchLaser(const LaserArgs& a, hipStream_t s);
hipError_t launchNoise(float* f, size_t n, uint64_t seed, int comp, float amp, hipStream_t s);

// ---------------------------------------------------------------------------
// batched ensembles: one launch per phase over every simulation
// ---------------------------------------------------------------------------
struct ShiftEntry {
    SpeciesDev S;
    GridDev g;
};
// items (simulation, supercell) into a table of PushArgs (fused push + deposit)
hipError_t launchPushBatch(int supp, int pusher, const PushArgs* table, const int2* work, int nwork, int nblocks,
                           hipStream_t s);
// items (entry, supercell) into a table of (species, grid); the batched push
// staged the shift, so this is the insert
hipError_t launchShiftBatch(const ShiftEntry* table, const int2* work, int nwork, hipStream_t s);
// items (simulation, tile) for the tiled kernel and (simulation, block of 256
// cells) for the per-cell kernels, into tables of FieldArgs and new-E planes
struct FieldBatchEntry {
    FieldArgs a;
    float* Enew[3];
};
hipError_t launchFieldStepBatch(int solver, const FieldBatchEntry* table, const int2* tiles, int ntiles,
                                const int2* blocks, int nblocks, hipStream_t s);
// J = 0 for items (simulation, block of 256 cells)
hipError_t launchCurrentResetBatch(const FieldBatchEntry* table, const int2* blocks, int nblocks, hipStream_t s);
// the gather flush (launchGatherFlush) of every member of a batch that ran
// its fused push in one launch, for items (simulation, block of 256 cells)
struct GatherEntry {
    GridDev g;
    const float* tiles;
    const uint8_t* colour;
    float* J[3];
};
hipError_t launchGatherFlushBatch(int supp, const GatherEntry* table, const int2* blocks, int nblocks,
                                  hipStream_t s);
// end of a batched step: every member's step flags as the end-of-step kernel of pic_step sets them
hipError_t launchEndStepBatch(const FieldBatchEntry* table, int n, hipStream_t s);

// ---------------------------------------------------------------------------
// differentiable form (one CIC/Boris species in identity order) and its adjoint
// ---------------------------------------------------------------------------
struct DiffParticles {
    const int32_t* cell;  // global cell index, 3 planes of n
    const float* pos;     // in-cell position, 3 planes of n
    const float* mom;     // 3 planes of n
    const float* w;       // n
    int32_t* cellOut;     // the state after the step (may equal the input)
    float* posOut;
    float* momOut;
    int64_t n;
};

struct DiffArgs {
    GridDev g;
    const float* E[3];
    const float* B[3];
    unsigned long long* Jfix[3];  // int64 fixed-point current, zero on entry
    float cell[3];
    float dt, c2, rc2, c;
    float inv_vdt;
    double jscale;
    float mass;    // frame mass
    float charge;  // frame charge
};

struct AdjFieldArgs {
    GridDev g;
    double* lE[3];
    double* lB[3];
    double cell[3];
};

struct AdjPartArgs {
    GridDev g;
    const float* E[3];  // fields at the start of the step
    const float* B[3];
    const double* lJ[3];
    double* lE[3];  // accumulated with atomics
    double* lB[3];
    double cell[3];
    double dt;
    double rc2;
    double mass;    // frame mass
    double charge;  // frame charge
    double inv_vdt;
};

struct AdjState {
    const int32_t* cell;  // particle state at the start of the step
    const float* pos;
    const float* mom;
    const float* w;
    double* lX;  // adjoint of the global position (cells), 3 planes, in/out
    double* lP;  // adjoint of the momentum, 3 planes, in/out
    int64_t n;
};

hipError_t launchDiffPush(const DiffArgs& a, const DiffParticles& P, hipStream_t s);
hipError_t launchFixToFloat(unsigned long long* fix, float* const J[3], size_t n, double jinv, hipStream_t s);
hipError_t launchAdjFields(const AdjFieldArgs& a, double dt, double c2, double coeff, double* lJ[3], hipStream_t s);
hipError_t launchAdjParticles(const AdjPartArgs& a, const AdjState& st, hipStream_t s);
hipError_t launchFillD(double* p, size_t n, double v, hipStream_t s);
// particle order of the differentiable form: tiles of 8x4x1 cells (grids are
// multiples of the supercell, so of the tile)
constexpr int kTileX = 8, kTileY = 4, kTileZ = 1;
hipError_t sortTempBytes(int64_t n, size_t* bytes);
// stable sort of the slots by tile; *perm receives vals[1] (new slot -> old slot)
hipError_t launchTileSort(const int32_t* cell, int64_t n, const GridDev& g, uint32_t* keys[2], int32_t* vals[2],
                          void* temp, size_t tempBytes, int32_t** perm, hipStream_t s);
hipError_t launchGatherParticles(const int32_t* perm, const DiffParticles& src, const DiffParticles& dst,
                                 const int32_t* idsIn, int32_t* idsOut, float* wOut, hipStream_t s);
hipError_t launchScatterAdjoint(const int32_t* perm, const double* src, double* dst, int64_t n, hipStream_t s);

}  // namespace pic
