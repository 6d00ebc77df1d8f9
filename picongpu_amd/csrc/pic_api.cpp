// Host runtime of the C ABI declared in include/picongpu_amd.h.
//
// Owns the device buffers of one simulation, derives the FP32 constants in
// PIC units the way unitless/simulation.unitless does, plans the supercell
// checkerboard phases of the deposit, and launches the kernels of each stage
// on the context's HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <numeric>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/picongpu_amd.h"
#include "pic_incident.hpp"
#include "pic_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define PIC_HIP(expr)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(PIC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// param/physicalConstants.param:31-48 (2022 CODATA values)
constexpr double kSpeedOfLightSI = 2.99792458e8;
constexpr double kMue0SI = 1.25663706127e-6;
constexpr double kElectronMassSI = 9.1093837139e-31;
constexpr double kElectronChargeSI = -1.602176634e-19;
constexpr double kPi = 3.141592653589793238462643383279502884197169399;

int supportOf(int shape) { return shape == PIC_SHAPE_CIC ? 2 : (shape == PIC_SHAPE_TSC ? 3 : 4); }

// Esirkepov current margins (Esirkepov.hpp:42-43)
int depLower(int supp) { return supp / 2 + 1 - (supp + 1) % 2; }
int depUpper(int supp) { return (supp + 1) / 2 + 1; }

int validateIncident(const pic_config& c);

int validateConfig(const pic_config& c) {
    const int sc[3] = {pic::kSCX, pic::kSCY, pic::kSCZ};
    for (int d = 0; d < 3; ++d) {
        if (c.grid[d] <= 0 || c.grid[d] % sc[d] != 0)
            return fail(PIC_E_INVALID, "grid must be a positive multiple of the supercell 8x8x4 per axis");
        if (!(c.cell_si[d] > 0.0)) return fail(PIC_E_INVALID, "cell_si must be positive");
    }
    if (!(c.dt_si > 0.0)) return fail(PIC_E_INVALID, "dt_si must be positive");
    if (!(c.base_density_si > 0.0)) return fail(PIC_E_INVALID, "base_density_si must be positive");
    if (c.typical_ppc <= 0) return fail(PIC_E_INVALID, "typical_ppc must be positive");
    if (c.n_species < 0 || c.n_species > PIC_MAX_SPECIES) return fail(PIC_E_INVALID, "n_species out of range");
    if (c.solver != PIC_SOLVER_YEE && c.solver != PIC_SOLVER_LEHE) return fail(PIC_E_INVALID, "unknown solver");
    if (c.solver == PIC_SOLVER_LEHE && (c.lehe_dir < 0 || c.lehe_dir > 2))
        return fail(PIC_E_INVALID, "lehe_dir must be 0, 1 or 2");
    for (int s = 0; s < c.n_species; ++s) {
        const pic_species_desc& d = c.species[s];
        if (d.shape < PIC_SHAPE_CIC || d.shape > PIC_SHAPE_PQS) return fail(PIC_E_INVALID, "unknown shape");
        if (d.pusher < PIC_PUSHER_NONE || d.pusher > PIC_PUSHER_HIGUERA_CARY) return fail(PIC_E_INVALID, "unknown pusher");
        if (d.current_solver != PIC_CURRENT_SOLVER_ESIRKEPOV && d.current_solver != PIC_CURRENT_SOLVER_EMZ)
            return fail(PIC_E_INVALID, "unknown current solver");
        if (d.frames_per_sc <= 0 || d.frames_per_sc > 255)
            return fail(PIC_E_INVALID, "frames_per_sc must be in [1,255]");
        if (!(d.mass_ratio > 0.0)) return fail(PIC_E_INVALID, "mass_ratio must be positive");
        for (int a = 0; a < 3; ++a) {
            if (d.particle_boundary[a] != PIC_PARTICLE_ABSORBING && d.particle_boundary[a] != PIC_PARTICLE_REFLECTING)
                return fail(PIC_E_INVALID, "particle_boundary must be PIC_PARTICLE_ABSORBING or _REFLECTING");
            if (d.particle_boundary[a] == PIC_PARTICLE_REFLECTING && c.boundary[a] != PIC_BOUNDARY_OPEN)
                return fail(PIC_E_INVALID, "a reflecting particle boundary needs an open axis");
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (c.boundary[d] != PIC_BOUNDARY_PERIODIC && c.boundary[d] != PIC_BOUNDARY_OPEN)
            return fail(PIC_E_INVALID, "boundary must be PIC_BOUNDARY_PERIODIC or PIC_BOUNDARY_OPEN");
        for (int side = 0; side < 2; ++side) {
            if (c.absorber_cells[d][side] < 0 || c.absorber_cells[d][side] > PIC_ABSORBER_MAX_CELLS)
                return fail(PIC_E_INVALID, "absorber_cells must be in [0, PIC_ABSORBER_MAX_CELLS]");
            if (!(c.absorber_strength[d][side] >= 0.0f) || !std::isfinite(c.absorber_strength[d][side]))
                return fail(PIC_E_INVALID, "absorber_strength must be finite and >= 0");
        }
        if (c.boundary[d] == PIC_BOUNDARY_OPEN && c.absorber != PIC_ABSORBER_NONE &&
            c.absorber_cells[d][0] + c.absorber_cells[d][1] > c.grid[d])
            return fail(PIC_E_INVALID, "absorber layers are thicker than the grid");
    }
    if (c.absorber != PIC_ABSORBER_NONE && c.absorber != PIC_ABSORBER_EXPONENTIAL && c.absorber != PIC_ABSORBER_PML)
        return fail(PIC_E_INVALID, "absorber must be PIC_ABSORBER_NONE, PIC_ABSORBER_EXPONENTIAL or PIC_ABSORBER_PML");
    if (c.absorber == PIC_ABSORBER_PML && c.pml.custom != 0) {
        const pic_pml_params& q = c.pml;
        bool ok = q.sigma_kappa_grading_order >= 0.0 && q.alpha_grading_order >= 0.0;
        for (int d = 0; d < 3; ++d)
            ok = ok && q.sigma_max_si[d] >= 0.0 && q.kappa_max[d] >= 1.0 && q.alpha_max_si[d] >= 0.0;
        if (!ok)
            return fail(PIC_E_INVALID, "PML: grading orders, SIGMA_MAX_SI and ALPHA_MAX_SI must be >= 0, "
                                       "KAPPA_MAX >= 1 (fieldAbsorber.unitless)");
    }
    if (c.current_interpolation != PIC_CURRENT_NONE && c.current_interpolation != PIC_CURRENT_BINOMIAL)
        return fail(PIC_E_INVALID, "current_interpolation must be PIC_CURRENT_NONE or PIC_CURRENT_BINOMIAL");
    if (c.current_interpolation == PIC_CURRENT_BINOMIAL &&
        (c.boundary[0] != PIC_BOUNDARY_PERIODIC || c.boundary[1] != PIC_BOUNDARY_PERIODIC ||
         c.boundary[2] != PIC_BOUNDARY_PERIODIC))
        return fail(PIC_E_INVALID, "the binomial current filter needs a periodic box");
    return validateIncident(c);
}

bool allPeriodic(const pic_config& c) {
    return c.boundary[0] == PIC_BOUNDARY_PERIODIC && c.boundary[1] == PIC_BOUNDARY_PERIODIC &&
           c.boundary[2] == PIC_BOUNDARY_PERIODIC;
}

void computeConstants(const pic_config& c, pic_constants& k) {
    std::memset(&k, 0, sizeof(k));
    // unitless/simulation.unitless:428-479
    double const unitTime = c.dt_si;
    double const unitSpeed = kSpeedOfLightSI;
    double const unitLength = unitTime * unitSpeed;
    double const cellVolume = c.cell_si[0] * c.cell_si[1] * c.cell_si[2];
    double const wTyp = (c.base_density_si * cellVolume) / static_cast<double>(c.typical_ppc);
    double const unitMass = kElectronMassSI * wTyp;
    double const unitCharge = -1.0 * kElectronChargeSI * wTyp;
    k.unit_length = unitLength;
    k.unit_time = unitTime;
    k.unit_mass = unitMass;
    k.unit_charge = unitCharge;
    k.typical_weighting = wTyp;
    for (int d = 0; d < 3; ++d) k.cell[d] = static_cast<float>(c.cell_si[d] / unitLength);
    k.dt = static_cast<float>(c.dt_si / unitTime);
    k.c = static_cast<float>(kSpeedOfLightSI / unitSpeed);
    k.mue0 = static_cast<float>(kMue0SI / unitLength / unitMass * unitCharge * unitCharge);
    k.eps0 = static_cast<float>(1. / static_cast<double>(k.mue0) / static_cast<double>(k.c) / static_cast<double>(k.c));
    k.mue0_eps0 = static_cast<float>(1. / static_cast<double>(k.c) / static_cast<double>(k.c));
    // FDTD.hpp:84
    k.add_current_coeff = -(1.0f / k.eps0) * k.dt;
    // Esirkepov.hpp:190-192: 1/float(cellSize.productOfComponents()*dt)
    float const vdt = k.cell[0] * k.cell[1] * k.cell[2] * k.dt;
    k.inv_cell_volume_dt = 1.0f / vdt;
    float const baseMass = static_cast<float>(kElectronMassSI / unitMass);
    float const baseCharge = static_cast<float>(kElectronChargeSI / unitCharge);
    for (int s = 0; s < c.n_species; ++s) {
        k.species_mass[s] = baseMass * static_cast<float>(c.species[s].mass_ratio);
        k.species_charge[s] = baseCharge * static_cast<float>(c.species[s].charge_ratio);
    }
    // Lehe/Derivative.hpp:90-111
    if (c.solver == PIC_SOLVER_LEHE) {
        int const d0 = c.lehe_dir, d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        double const stepRatio = static_cast<double>(k.cell[d0] / (k.c * k.dt));
        double const coeff = stepRatio * std::sin(kPi / 2.0 * static_cast<double>(k.c) * static_cast<double>(k.dt) /
                                                  static_cast<double>(k.cell[d0]));
        k.lehe_delta = static_cast<float>(0.25 * (1.0 - coeff * coeff));
        double const stepRatio1 = static_cast<double>(k.cell[d0] / k.cell[d1]);
        double const stepRatio2 = static_cast<double>(k.cell[d0] / k.cell[d2]);
        double const betaDir1 = 0.125 * stepRatio1 * stepRatio1;
        double const betaDir2 = 0.125 * stepRatio2 * stepRatio2;
        k.lehe_alpha =
            static_cast<float>(1.0 - 2.0 * betaDir1 - 2.0 * betaDir2 - 3.0 * static_cast<double>(k.lehe_delta));
    }
}

// CFL checks: Yee/Yee.hpp:43-65, Lehe/Lehe.hpp:47-69
int checkCfl(const pic_config& c, const pic_constants& k) {
    // moveParticle and the deposit's J tile assume a move of less than one
    // cell per axis: dp = float(float(v*dt)/cell) with |v| < c is below 1 on
    // every axis when the same quotient of c is
    for (int d = 0; d < 3; ++d) {
        float const cdt = k.c * k.dt;
        if (!(cdt / k.cell[d] < 1.0f))
            return fail(PIC_E_INVALID, "c*dt reaches a cell size: a particle could move a whole cell per step");
    }
    if (c.solver == PIC_SOLVER_YEE) {
        float inv = 0.0f;
        for (int d = 0; d < 3; ++d) inv += 1.0f / (k.cell[d] * k.cell[d]);
        if (k.c * k.c * k.dt * k.dt * inv > 1.0f)
            return fail(PIC_E_INVALID, "Courant-Friedrichs-Lewy condition failure (Yee)");
    } else {
        if (k.c * k.dt > k.cell[c.lehe_dir])
            return fail(PIC_E_INVALID, "Courant-Friedrichs-Lewy condition failure (Lehe)");
    }
    return PIC_OK;
}

// ---------------------------------------------------------------------------
// PML coefficients (absorber/pml/Pml.kernel:60-270): per axis and index the
// b, c and kappa of getCoefficients, for E (at the cell index, half = false)
// or B (at index + 1/2), in the reference's float_X arithmetic. The relative
// depth is getRelativeDepth's, the layer thickness absorber_cells on open
// axes and 0 on periodic ones (a periodic border has a neighbour). The
// normalised conductivity and frequency shift are fieldAbsorber.unitless's
// SIGMA_MAX_SI / eps0 * unit time. Layout: b of x, y, z (nx + ny + nz
// values), then c, then kappa.
// fieldAbsorber.param's namespace pml (its defaults unless custom) and the
// normalised maxima of fieldAbsorber.unitless (SI value / eps0 * unit time)
struct PmlParameters {
    double gradingSK = 4.0, gradingA = 1.0;  // fieldAbsorber.param:97,143
    double sigmaMax[3], kappaMax[3] = {1.0, 1.0, 1.0}, alphaMax[3] = {0.2, 0.2, 0.2};
    float nSigma[3], nAlpha[3];
};

PmlParameters pmlParameters(const pic_config& cfg, const pic_constants& k) {
    PmlParameters p;
    double const z0 = kMue0SI * kSpeedOfLightSI;                        // simulation.unitless:277-280
    double const eps0 = 1.0 / kMue0SI / kSpeedOfLightSI / kSpeedOfLightSI;  // :264-267
    for (int d = 0; d < 3; ++d) p.sigmaMax[d] = 0.8 * (p.gradingSK + 1.0) / (z0 * cfg.cell_si[d]);  // SIGMA_OPT_SI
    if (cfg.pml.custom != 0) {
        p.gradingSK = cfg.pml.sigma_kappa_grading_order;
        p.gradingA = cfg.pml.alpha_grading_order;
        for (int d = 0; d < 3; ++d) {
            p.sigmaMax[d] = cfg.pml.sigma_max_si[d];
            p.kappaMax[d] = cfg.pml.kappa_max[d];
            p.alphaMax[d] = cfg.pml.alpha_max_si[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        p.nSigma[d] = static_cast<float>(p.sigmaMax[d] / eps0 * k.unit_time);  // NORMALIZED_SIGMA_MAX
        p.nAlpha[d] = static_cast<float>(p.alphaMax[d] / eps0 * k.unit_time);  // NORMALIZED_ALPHA_MAX
    }
    return p;
}

void pmlTables(const pic_config& cfg, const pic_constants& k, bool half, std::vector<float>& out) {
    int const n[3] = {cfg.grid[0], cfg.grid[1], cfg.grid[2]};
    int const total = n[0] + n[1] + n[2];
    out.assign(3 * static_cast<size_t>(total), 0.0f);
    PmlParameters const P = pmlParameters(cfg, k);
    float const m = static_cast<float>(P.gradingSK);
    float const ma = static_cast<float>(P.gradingA);
    float const dt = k.dt;
    int off = 0;
    for (int d = 0; d < 3; ++d) {
        bool const open = cfg.boundary[d] == PIC_BOUNDARY_OPEN;
        float const negB = static_cast<float>(open ? cfg.absorber_cells[d][0] : 0);
        float const posB = static_cast<float>(open ? cfg.absorber_cells[d][1] : 0);
        float const nSigma = P.nSigma[d];
        float const nAlpha = P.nAlpha[d];
        float const kMax = static_cast<float>(P.kappaMax[d]);
        for (int i = 0; i < n[d]; ++i) {
            float const zero = static_cast<float>(i) + (half ? 0.5f : 0.0f);
            float depth = 0.0f;
            if (zero < negB) {
                depth = (negB - zero) / negB;
            } else {
                float const rightStart = static_cast<float>(n[d]) - posB;
                if (zero > rightStart) depth = (zero - rightStart) / posB;
            }
            float sigma = 0.0f, kappa = 1.0f, alpha = 0.0f;
            if (depth != 0.0f) {
                float const g = std::pow(depth, m);
                sigma = nSigma * g;
                kappa = 1.0f + (kMax - 1.0f) * g;
                float const ga = std::pow(1.0f - depth, ma);
                alpha = nAlpha * ga;
            }
            float const b = std::exp(-(sigma / kappa + alpha) * dt);
            float const den = kappa * (sigma + alpha * kappa);
            out[off + i] = b;
            out[total + off + i] = den != 0.0f ? sigma * (b - 1.0f) / den : 0.0f;
            out[2 * total + off + i] = kappa;
        }
        off += n[d];
    }
}

// The PML weight damping of absorbing species (boundary/Absorbing.hpp:95-130
// dampWeightsInPml): on each open axis with a layer on a side, the area of
// that side's exchange is the supercells the layer reaches (particle boundary
// offset 0), the layer parameters those of getNormalizedSigma.
pic::PmlDampArgs pmlDampArgs(const pic_config& cfg, const pic_constants& k) {
    pic::PmlDampArgs a{};
    PmlParameters const P = pmlParameters(cfg, k);
    int const scs[3] = {pic::kSCX, pic::kSCY, pic::kSCZ};
    a.grading = static_cast<float>(P.gradingSK);
    a.gradingInt = (a.grading >= 0.0f && a.grading <= 32.0f && a.grading == std::floor(a.grading))
                       ? static_cast<int>(a.grading)
                       : -1;
    for (int d = 0; d < 3; ++d) {
        int const n = cfg.grid[d];
        bool const open = cfg.boundary[d] == PIC_BOUNDARY_OPEN;
        int const neg = open ? cfg.absorber_cells[d][0] : 0, pos = open ? cfg.absorber_cells[d][1] : 0;
        a.negB[d] = static_cast<float>(neg);
        a.posB[d] = static_cast<float>(pos);
        a.nCells[d] = static_cast<float>(n);
        a.guard[d] = scs[d];  // memory.param GuardSize: one supercell
        a.sigmaMax[d] = P.nSigma[d];
        a.slow[d] = 1e-5 * static_cast<double>(k.cell[d]);
        a.lo[d] = neg > 0 ? scs[d] * ((neg + scs[d] - 1) / scs[d]) : 0;
        a.hi[d] = pos > 0 ? n - scs[d] * ((pos + scs[d] - 1) / scs[d]) : n;
        if (open && (neg > 0 || pos > 0)) a.axes |= 1 << d;
    }
    return a;
}

// The box of cells whose E and B coefficients are neutral (b == 1 along all
// three axes, where the PML kernels take the plain update): the indices
// along each axis that are not neutral form a prefix and a suffix (the
// layers at both borders). Should an axis have a neutral index between
// layer indices, its range is left empty (every cell looks its
// coefficients up).
void pmlInnerBox(const pic_config& cfg, const std::vector<float>& tabE, const std::vector<float>& tabB, int in0[3],
                 int in1[3]) {
    int const n[3] = {cfg.grid[0], cfg.grid[1], cfg.grid[2]};
    int off = 0;
    for (int d = 0; d < 3; ++d) {
        auto neutral = [&](int i) { return tabE[off + i] == 1.0f && tabB[off + i] == 1.0f; };
        int lo = 0, hi = n[d];
        while (lo < n[d] && !neutral(lo)) ++lo;
        while (hi > lo && !neutral(hi - 1)) --hi;
        for (int i = lo; i < hi; ++i)
            if (!neutral(i)) lo = hi = 0;
        in0[d] = lo;
        in1[d] = hi;
        off += n[d];
    }
}

// Incident field set-up (fields/incidentField/Solver.hpp, Functors.hpp,
// profiles/*, DerivativeCoefficients.hpp, calculatePhaseVelocity.hpp)
// ---------------------------------------------------------------------------
constexpr float kPiF = 3.14159265358979323846f;  // pmacc Pi<float_X>::value
constexpr float kTwoPiF = 2.0f * kPiF;            // Pi<float_X>::doubleValue
// Yee-cell positions of the E and B components (fields/YeeCell.hpp:69-141)
constexpr float kYeeE[3][3] = {{0.5f, 0.0f, 0.0f}, {0.0f, 0.5f, 0.0f}, {0.0f, 0.0f, 0.5f}};
constexpr float kYeeB[3][3] = {{0.0f, 0.5f, 0.5f}, {0.5f, 0.0f, 0.5f}, {0.5f, 0.5f, 0.0f}};

// Coefficients of the derivative along `axis` in the curl of the field the
// update reads: DerivativeCoefficients (DerivativeCoefficients.hpp:35-200).
// curlE: the B update (the Lehe derivative under Lehe), else the E update
// (Yee's backward derivative, also under Lehe: Lehe.def:73).
void incidentCoefficients(const pic_config& cfg, const pic_constants& k, bool curlE, int axis, float value[2][2][2],
                          int size[3]) {
    std::memset(value, 0, sizeof(float) * 8);
    size[0] = size[1] = size[2] = 1;
    if (!(curlE && cfg.solver == PIC_SOLVER_LEHE)) {
        value[0][0][0] = 1.0f;
        return;
    }
    int const dir = cfg.lehe_dir;
    if (dir == axis) {
        // DerivativeCoefficients<lehe::DerivativeFunctor<axis, axis>> (:95-165)
        int const d0 = axis, d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        double const stepRatio = static_cast<double>(k.cell[d0] / (k.c * k.dt));
        double const coeff = stepRatio * std::sin(kPi / 2.0 * static_cast<double>(k.c) * static_cast<double>(k.dt) /
                                                  static_cast<double>(k.cell[d0]));
        float const delta = static_cast<float>(0.25 * (1.0 - coeff * coeff));
        double const stepRatio1 = static_cast<double>(k.cell[d0] / k.cell[d1]);
        double const stepRatio2 = static_cast<double>(k.cell[d0] / k.cell[d2]);
        double const betaDir1 = 0.125 * stepRatio1 * stepRatio1;
        double const betaDir2 = 0.125 * stepRatio2 * stepRatio2;
        float const alpha = static_cast<float>(1.0 - 2.0 * betaDir1 - 2.0 * betaDir2 - 3.0 * delta);
        size[0] = size[1] = size[2] = 2;
        value[0][0][0] = alpha;
        int i0[3] = {0, 0, 0}, i1[3] = {0, 0, 0}, i2[3] = {0, 0, 0};
        i0[d0] = 1;
        i1[d1] = 1;
        i2[d2] = 1;
        value[i0[0]][i0[1]][i0[2]] = delta;
        value[i1[0]][i1[1]][i1[2]] = static_cast<float>(betaDir1);
        value[i2[0]][i2[1]][i2[2]] = static_cast<float>(betaDir2);
    } else {
        // DerivativeCoefficients<lehe::DerivativeFunctor<dir, axis>> (:167-200)
        float const beta = 0.125f;
        float const alpha = 1.0f - 2.0f * beta;
        size[dir] = 2;
        value[0][0][0] = alpha;
        int i[3] = {0, 0, 0};
        i[dir] = 1;
        value[i[0]][i[1]][i[2]] = beta;
    }
}

// margin of the B and E updates along `axis` (Solver.kernel:123)
int incidentMargin(const pic_config& cfg, bool curlE, int axis) {
    return (curlE && cfg.solver == PIC_SOLVER_LEHE && cfg.lehe_dir == axis) ? 2 : 1;
}

// DispersionRelation<Solver>::relation / relationDerivative (DispersionRelation.hpp:60-100,
// Yee.hpp:68-119, Lehe.hpp:77-200) in FP64
double dispersion(const pic_config& cfg, const pic_constants& k, double omega, const double* dir, double absK,
                  bool derivative) {
    double const timeStep = static_cast<double>(k.dt);
    double const step[3] = {static_cast<double>(k.cell[0]), static_cast<double>(k.cell[1]),
                            static_cast<double>(k.cell[2])};
    double const cc = static_cast<double>(k.c);
    double const lhsTerm = std::sin(0.5 * omega * timeStep) / (cc * timeStep);
    if (cfg.solver == PIC_SOLVER_YEE) {
        double r = 0.0;
        for (int d = 0; d < 3; ++d) {
            double const arg = 0.5 * absK * dir[d] * step[d];
            double const term = std::sin(arg) / step[d];
            if (derivative) {
                double const termDerivative = 0.5 * dir[d] * std::cos(arg);
                r += 2.0 * term * termDerivative;
            } else {
                r += term * term;
            }
        }
        return derivative ? r : r - lhsTerm * lhsTerm;
    }
    int const d0 = cfg.lehe_dir, d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
    double const stepSq[3] = {step[0] * step[0], step[1] * step[1], step[2] * step[2]};
    double const stepRatio = step[d0] / (cc * timeStep);
    double const coeff = stepRatio * std::sin(kPi / 2.0 * (cc * timeStep) / step[d0]);
    double const delta = 0.25 * (1.0 - coeff * coeff);
    double const stepRatio1 = step[d0] / step[d1];
    double const stepRatio2 = step[d0] / step[d2];
    double const betaDir0 = 0.125;
    double const betaDir1 = 0.125 * stepRatio1 * stepRatio1;
    double const betaDir2 = 0.125 * stepRatio2 * stepRatio2;
    double sSq[3];
    for (int d = 0; d < 3; ++d) {
        double const s = std::sin(0.5 * absK * dir[d] * step[d]);
        sSq[d] = s * s;
    }
    if (!derivative) {
        double rhs = 0.0;
        for (int d = 0; d < 3; ++d) rhs += sSq[d] / stepSq[d];
        rhs -= 4.0 * delta * sSq[d0] * sSq[d0] / stepSq[d0];
        rhs -= 4.0 * (betaDir1 / stepSq[d1] + betaDir0 / stepSq[d0]) * sSq[d0] * sSq[d1];
        rhs -= 4.0 * (betaDir2 / stepSq[d2] + betaDir0 / stepSq[d0]) * sSq[d0] * sSq[d2];
        return rhs - lhsTerm * lhsTerm;
    }
    double sSqD[3];
    for (int d = 0; d < 3; ++d) {
        double const arg = 0.5 * absK * dir[d] * step[d];
        double const s = std::sin(arg);
        double const sD = 0.5 * dir[d] * step[d] * std::cos(arg);
        sSqD[d] = 2.0 * s * sD;
    }
    double r = 0.0;
    for (int d = 0; d < 3; ++d) r += sSqD[d] / stepSq[d];
    r -= 4.0 * delta * sSq[d0] * sSqD[d0] / stepSq[d0];
    r -= 4.0 * (betaDir1 / stepSq[d1] + betaDir0 / stepSq[d0]) * (sSq[d0] * sSqD[d1] + sSqD[d0] * sSq[d1]);
    // the reference takes betaDir1 in this term as well (Lehe.hpp:190-191)
    r -= 4.0 * (betaDir1 / stepSq[d1] + betaDir0 / stepSq[d0]) * (sSq[d0] * sSqD[d2] + sSqD[d0] * sSq[d2]);
    return r;
}

// calculatePhaseVelocity (traits/GetPhaseVelocity.hpp:31-45) with
// DispersionRelationSolver (DispersionRelationSolver.hpp:36-55): Newton from
// |k| = omega * c, at most 100 steps, stopping on a flat derivative
float incidentPhaseVelocity(const pic_config& cfg, const pic_constants& k, float waveLength, const float* dirF) {
    double const omega = 2.0 * kPi * static_cast<double>(k.c / waveLength);
    double const dir[3] = {static_cast<double>(dirF[0]), static_cast<double>(dirF[1]), static_cast<double>(dirF[2])};
    double absK = omega * static_cast<double>(k.c);
    for (int i = 0; i < 100; ++i) {
        double const der = dispersion(cfg, k, omega, dir, absK, true);
        if (std::abs(der) < 1e-5) break;
        absK = absK - dispersion(cfg, k, omega, dir, absK, false) / der;
    }
    return static_cast<float>(omega / absK / static_cast<double>(k.c));
}

// Huygens surface cell index along `axis` at the negative (side 0) and
// positive (side 1) border (Solver.hpp:343-354)
int huygensBorder(const pic_config& cfg, int axis, int side) {
    if (side == 0) return cfg.incident_position[axis][0];
    int const p = cfg.incident_position[axis][1];
    return p > 0 ? p : cfg.grid[axis] + p;
}

int validateIncident(const pic_config& c) {
    bool any = false;
    for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) {
            const pic_incident_profile& p = c.incident[axis][side];
            if (p.kind == PIC_INCIDENT_NONE) continue;
            if (p.kind != PIC_INCIDENT_GAUSSIAN_PULSE && p.kind != PIC_INCIDENT_PLANE_WAVE)
                return fail(PIC_E_INVALID, "unknown incident-field profile kind");
            any = true;
            if (p.polarisation != PIC_POLARISATION_LINEAR && p.polarisation != PIC_POLARISATION_CIRCULAR)
                return fail(PIC_E_INVALID, "polarisation must be PIC_POLARISATION_LINEAR or _CIRCULAR");
            if (!(p.wavelength_si > 0.0) || !(p.pulse_duration_si > 0.0))
                return fail(PIC_E_INVALID, "incident field: wavelength and pulse duration must be positive");
            if (!(p.time_delay_si >= 0.0))
                return fail(PIC_E_INVALID, "incident field: the time delay must be >= 0 (BaseParam.hpp:90-93)");
            // BaseParam.hpp:95-120: unit direction, unit polarisation, orthogonal
            float const dir[3] = {static_cast<float>(p.direction[0]), static_cast<float>(p.direction[1]),
                                  static_cast<float>(p.direction[2])};
            float const pol[3] = {static_cast<float>(p.polarisation_direction[0]),
                                  static_cast<float>(p.polarisation_direction[1]),
                                  static_cast<float>(p.polarisation_direction[2])};
            double const n2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
            double const p2 = pol[0] * pol[0] + pol[1] * pol[1] + pol[2] * pol[2];
            double const dp = dir[0] * pol[0] + dir[1] * pol[1] + dir[2] * pol[2];
            if (!(n2 > 0.9999 && n2 < 1.0001)) return fail(PIC_E_INVALID, "incident field: direction must be a unit vector");
            if (!(p2 > 0.9999 && p2 < 1.0001))
                return fail(PIC_E_INVALID, "incident field: polarisation direction must be a unit vector");
            if (!(dp > -0.0001 && dp < 0.0001))
                return fail(PIC_E_INVALID, "incident field: polarisation must be orthogonal to the direction");
            for (int d = 0; d < 3; ++d)
                if (p.focus_origin[d] != PIC_ORIGIN_ZERO && p.focus_origin[d] != PIC_ORIGIN_CENTER)
                    return fail(PIC_E_INVALID, "incident field: focus origin must be PIC_ORIGIN_ZERO or _CENTER");
            if (p.kind == PIC_INCIDENT_GAUSSIAN_PULSE) {
                if (!(p.w0_si > 0.0)) return fail(PIC_E_INVALID, "incident field: W0 must be positive");
                if (p.num_modes < 0 || p.num_modes >= PIC_INCIDENT_MAX_MODES)
                    return fail(PIC_E_INVALID, "incident field: num_modes must be in [0, PIC_INCIDENT_MAX_MODES)");
                float norm = 0.0f;
                for (int m = 0; m <= p.num_modes; ++m) norm += p.laguerre_modes[m];
                if (!(std::abs(norm) > std::numeric_limits<float>::epsilon()))
                    return fail(PIC_E_INVALID, "Sum of laguerreModes can not be 0.");
            }
            // checkRequirements (Solver.hpp:60-100): the surface keeps the
            // absorber plus the update margin from the border, and fits the domain
            int const margin = std::max(incidentMargin(c, true, axis), incidentMargin(c, false, axis));
            int const th = c.absorber != PIC_ABSORBER_NONE ? c.absorber_cells[axis][side] : 0;
            int offset;
            if (side == 0)
                offset = c.incident_position[axis][0];
            else if (c.incident_position[axis][1] > 0)
                offset = c.grid[axis] - c.incident_position[axis][1];
            else
                offset = -c.incident_position[axis][1];
            if (offset < th + margin - 1)
                return fail(PIC_E_INVALID, "Incident field POSITION[" + std::to_string(axis) + "][" +
                                               std::to_string(side) +
                                               "] is too close to the boundary for used field solver and absorber, "
                                               "must be at least " + std::to_string(th + margin - 1) + " cells away");
            for (int updateB = 0; updateB < 2; ++updateB) {
                int const m = incidentMargin(c, updateB != 0, axis);
                int begin = huygensBorder(c, axis, 0) + 1, end = huygensBorder(c, axis, 1);
                if (updateB && side == 0) begin -= 1;
                if (side == 0) {
                    begin -= m - 1;
                } else {
                    end += m - 1;
                    begin = end - (2 * m - 1);
                }
                begin = std::max(begin, 0);
                if (begin + 1 < m || begin + m > c.grid[axis])
                    return fail(PIC_E_INVALID, "The Huygens surface for incident field generation is too close to a "
                                               "local domain border. Adjust POSITION.");
            }
        }
    // an empty volume inside the surface generates no field (the reference
    // only warns, Solver.hpp:380-390): its launches have no cells
    (void)any;
    return PIC_OK;
}

// Profile constants in PIC units (BaseParamUnitless, BaseParam.hpp:40-140;
// GaussianPulseUnitless, GaussianPulse.hpp:60-75; PlaneWaveUnitless,
// PlaneWave.hpp:60-75; BaseFunctorE, Functors.hpp:80-230)
pic::IncidentProfile incidentProfile(const pic_config& cfg, const pic_constants& k, const pic_incident_profile& p) {
    pic::IncidentProfile P{};
    double const uL = k.unit_length, uT = k.unit_time, uM = k.unit_mass, uQ = k.unit_charge;
    double const unitE = 1.0 / (uT * uT / uM / uL * uQ);  // simulation.unitless:461-464
    P.kind = p.kind;
    P.polarisation = p.polarisation;
    for (int d = 0; d < 3; ++d) P.cell[d] = k.cell[d];
    P.c = k.c;
    P.dt = k.dt;
    P.waveLength = static_cast<float>(p.wavelength_si / uL);
    float const f = static_cast<float>(k.c / P.waveLength);
    P.omega = kTwoPiF * f;
    P.amplitude = static_cast<float>(p.amplitude_si / unitE);
    P.pulseDuration = static_cast<float>(p.pulse_duration_si / uT);
    P.laserPhase = p.laser_phase;
    float const delaySi = static_cast<float>(p.time_delay_si);  // GetTimeDelay::value is a float_X
    P.timeDelay = static_cast<float>(static_cast<double>(delaySi) / uT);
    float dir[3], pol[3], focus[3];
    for (int d = 0; d < 3; ++d) {
        dir[d] = static_cast<float>(p.direction[d]);
        pol[d] = static_cast<float>(p.polarisation_direction[d]);
        focus[d] = static_cast<float>(p.focus_position_si[d] / uL);
        P.axis0[d] = dir[d];
        P.axis1[d] = pol[d];
    }
    P.axis2[0] = dir[1] * pol[2] - dir[2] * pol[1];
    P.axis2[1] = dir[2] * pol[0] - dir[0] * pol[2];
    P.axis2[2] = dir[0] * pol[1] - dir[1] * pol[0];
    // getFocus (Functors.hpp:158-175)
    for (int d = 0; d < 3; ++d)
        if (p.focus_origin[d] == PIC_ORIGIN_CENTER)
            focus[d] += static_cast<float>(static_cast<uint32_t>(cfg.grid[d]) / 2u) * k.cell[d];
    // getOrigin (Functors.hpp:177-205)
    float originP = -std::numeric_limits<float>::infinity();
    for (int axis = 0; axis < 3; ++axis) {
        if (std::abs(dir[axis]) > std::numeric_limits<float>::epsilon()) {
            float const minPosition = (static_cast<float>(cfg.incident_position[axis][0]) + 0.75f) * k.cell[axis];
            int const maxIdx = cfg.incident_position[axis][1] > 0 ? cfg.incident_position[axis][1]
                                                                  : cfg.grid[axis] + cfg.incident_position[axis][1];
            float const maxPosition = (static_cast<float>(maxIdx) - 0.75f) * k.cell[axis];
            float const axisP = std::min((minPosition - focus[axis]) / dir[axis], (maxPosition - focus[axis]) / dir[axis]);
            originP = std::max(originP, axisP);
        }
    }
    for (int d = 0; d < 3; ++d) {
        P.focus[d] = focus[d];
        P.origin[d] = focus[d] + originP * dir[d];
    }
    P.phaseVelocity = incidentPhaseVelocity(cfg, k, P.waveLength, dir);
    if (p.kind == PIC_INCIDENT_GAUSSIAN_PULSE) {
        P.w0 = static_cast<float>(p.w0_si / uL);
        P.rayleighLength = kPiF * P.w0 * P.w0 / P.waveLength;
        P.timeShift = static_cast<float>(-0.5 * p.pulse_init * static_cast<double>(P.pulseDuration));
        P.numModes = p.num_modes + 1;
        for (int m = 0; m < P.numModes; ++m) {
            P.modes[m] = p.laguerre_modes[m];
            P.phases[m] = p.laguerre_phases[m];
        }
    } else {
        P.rampInit = p.ramp_init;
        P.nofocus = static_cast<float>(p.laser_nofocus_constant_si / uT);
    }
    return P;
}

// The launches of updateBHalf (updateB = true) or updateE in the reference's
// order: axes x, y, z, each the negative border's profile (direction +1)
// then the positive one's (updateField, Solver.hpp:120-258)
std::vector<pic::IncidentUpdate> incidentUpdates(const pic_config& cfg, const pic_constants& k, bool updateB) {
    std::vector<pic::IncidentUpdate> out;
    // UpdateB: curlCoefficient = -timeIncrement with timeIncrement = 0.5 dt;
    // UpdateE: timeIncrement * c^2 with timeIncrement = dt (Solver.hpp:290-330, 380-420)
    float const c2 = k.c * k.c;
    float const curlCoefficient = updateB ? -(0.5f * k.dt) : k.dt * c2;
    for (int axis = 0; axis < 3; ++axis) {
        for (int side = 0; side < 2; ++side) {
            const pic_incident_profile& p = cfg.incident[axis][side];
            if (p.kind == PIC_INCIDENT_NONE) continue;
            pic::IncidentUpdate u{};
            u.prof = incidentProfile(cfg, k, p);
            u.axis = axis;
            u.direction = side == 0 ? 1.0f : -1.0f;
            u.updatedTotal = updateB ? 0 : 1;  // E_x sits at x = 0.5, B_x at x = 0
            u.incidentIsB = updateB ? 0 : 1;
            int size[3];
            incidentCoefficients(cfg, k, updateB, axis, u.value, size);
            u.margin = size[axis];
            u.size1 = size[(axis + 1) % 3];
            u.size2 = size[(axis + 2) % 3];
            float const sign = u.direction > 0.0f ? 1.0f : -1.0f;
            u.baseCoefficient = curlCoefficient / k.cell[axis] * sign;
            // updated cells (Solver.hpp:130-200)
            int begin[3], end[3];
            for (int d = 0; d < 3; ++d) {
                begin[d] = huygensBorder(cfg, d, 0) + 1;
                end[d] = huygensBorder(cfg, d, 1);
            }
            if (!u.updatedTotal && u.direction > 0.0f) begin[axis] -= 1;
            bool const contiguous = p.kind == PIC_INCIDENT_PLANE_WAVE;  // PlaneWave.def:70-74
            for (int d = 0; d < 3; ++d) {
                if (contiguous && cfg.boundary[d] == PIC_BOUNDARY_PERIODIC && d != axis) {
                    begin[d] = 0;
                    end[d] = cfg.grid[d];
                    u.isLast[d] = 0;
                } else {
                    u.isLast[d] = 1;
                }
            }
            int const sizeAlongAxis = 2 * u.margin - 1;
            if (u.direction > 0.0f) {
                begin[axis] -= u.margin - 1;
                end[axis] = begin[axis] + sizeAlongAxis;
            } else {
                end[axis] += u.margin - 1;
                begin[axis] = end[axis] - sizeAlongAxis;
            }
            bool anyCells = true;
            for (int d = 0; d < 3; ++d) {
                u.begin[d] = std::max(begin[d], 0);
                u.end[d] = std::min(end[d], cfg.grid[d]);
                anyCells = anyCells && u.begin[d] < u.end[d];
            }
            if (!anyCells) continue;
            u.huygens = u.begin[axis] + u.margin - 1;
            // incidentFieldBaseShift + the incident component's Yee position (Solver.hpp:226-240)
            float baseShift[3] = {0.0f, 0.0f, 0.0f};
            if (u.direction > 0.0f) baseShift[axis] = u.updatedTotal ? -1.0f : 1.0f;
            const float(*pos)[3] = updateB ? kYeeE : kYeeB;
            int const c1 = (axis + 2) % 3, c2i = (axis + 1) % 3;
            for (int d = 0; d < 3; ++d) {
                u.inCellShift1[d] = baseShift[d] + pos[c1][d];
                u.inCellShift2[d] = baseShift[d] + pos[c2i][d];
            }
            out.push_back(u);
        }
    }
    return out;
}

// Colours of a periodic ring of n supercells such that equal colours are at
// least `s` apart: either the smallest divisor of n not below s, or
// i % s for the first n - n%s supercells plus one colour each for the rest.
int ringColours(int n, int s, std::vector<int>& colour) {
    colour.assign(n, 0);
    int div = s;
    while (n % div != 0) ++div;
    int const r = n % s;
    if (div <= s + r) {
        for (int i = 0; i < n; ++i) colour[i] = i % div;
        return div;
    }
    for (int i = 0; i < n; ++i) colour[i] = i < n - r ? i % s : s + (i - (n - r));
    return s + r;
}

struct PhasePlan {
    bool valid = false;
    int ncol[3] = {0, 0, 0};
    std::vector<int> offset;  // per phase into list
    std::vector<int> count;   // per phase
    std::vector<int> host;    // the list on the host, launch order per phase
    int32_t* dList = nullptr;
    // fused plans: every supercell in one launch, the tiles added to J by the
    // gather flush in phase order (same J, bit for bit), instead of one
    // launch per phase; chosen when the phases are too small to fill the GPU
    bool single = false;
    float* tiles = nullptr;               // nsc tiles
    uint8_t* colour = nullptr;            // checkerboard colour per supercell index along x, y, z
};

struct SpeciesMem {
    pic_species_desc desc{};
    // supercell sc owns the run of slots [base[sc], base[sc] + capSc[sc])
    // (whole frames) of every per-slot array, a pool of `slots` slots whose
    // runs end at `used`; a growing supercell gets a new run at the end. `cap`
    // is the largest run: the pitch of the frame view pic_read_frames and
    // pic_write_frames hand across the ABI.
    std::vector<uint32_t> capSc;
    std::vector<uint64_t> base;
    uint64_t* dBase = nullptr;
    uint32_t* dCap = nullptr;
    size_t slots = 0;
    size_t used = 0;
    int cap = 0;
    float mass = 0.0f;
    float charge = 0.0f;
    double wmax = 0.0;
    // what the last push left for the shift: 0 nothing (the shift collects the
    // leaver lists itself), 1 the leaver lists in scratch/outIdx/outOff, 2 the
    // lists plus the staged outbox and compaction (only the insert remains)
    int shiftReady = 0;
    pic::SpeciesDev dev{};
};

// one stored state of the differentiable form
struct DiffState {
    int32_t* cell = nullptr;  // 3 planes of n
    float* pos = nullptr;     // 3 planes of n
    float* mom = nullptr;     // 3 planes of n
    float* E = nullptr;       // 3 planes of ncells
    float* B = nullptr;       // 3 planes of ncells
    float* w = nullptr;       // checkpoints: weightings in this state's particle order
    int32_t* perm = nullptr;  // checkpoints: slot -> slot of the previous order (the re-tiling)
};

struct AdjointMem {
    int species = -1;
    int64_t n = 0;
    float mass = 0.0f;    // frame mass
    float charge = 0.0f;  // frame charge
    int jexp = 40;
    float* w = nullptr;      // weightings in the order of `work`
    float* wInit = nullptr;  // weightings as loaded
    int32_t* ids = nullptr;  // original index of each slot of `work`
    int32_t* idsTmp = nullptr;
    int32_t* iota = nullptr;
    uint32_t* keys[2] = {nullptr, nullptr};  // tile sort
    int32_t* vals[2] = {nullptr, nullptr};
    void* sortTemp = nullptr;
    size_t sortTempBytes = 0;
    double* lXtmp = nullptr;  // adjoint state while it is permuted
    double* lPtmp = nullptr;
    DiffState init;   // particle part: loaded initial particles
    DiffState work;   // particle part: particles of the running forward
    DiffState scratch;              // particle part and w: target of a re-tiling gather
    std::vector<const float*> tapeW;  // weightings of each tape state's particle order
    std::vector<DiffState> checkpoints;  // state at steps 0, K, 2K, ...
    std::vector<DiffState> tape;         // states of the segment being reversed
    unsigned long long* Jfix = nullptr;  // 3 planes of ncells
    double* lE = nullptr;  // 3 planes of ncells each
    double* lB = nullptr;
    double* lJ = nullptr;
    double* lX = nullptr;  // 3 planes of n each
    double* lP = nullptr;
    int nsteps = 0;
    int every = 0;
    bool forwardValid = false;
};

}  // namespace

struct pic_ctx {
    pic_config cfg{};
    pic_constants k{};
    pic::GridDev g{};
    int nsc = 0;
    size_t ncells = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    float* E[3] = {nullptr, nullptr, nullptr};
    float* Ealt[3] = {nullptr, nullptr, nullptr};  // receives the new E of the fused field step
    int eParity = 0;                               // flips with every swap of E and Ealt
    pic::AbsorbPlan absorb{};                      // exponential absorber layers, exchange order
    pic::PmlArgs pml{};                            // PML: convolution fields and coefficient tables
    bool hasPml = false;
    pic::PmlDampArgs pmlDamp{};                    // the PML's weight damping of absorbing species
    int pmlIn0[3] = {0, 0, 0}, pmlIn1[3] = {0, 0, 0};  // cells outside every PML layer
    float* Jf[3] = {nullptr, nullptr, nullptr};    // binomially filtered J for the fused field step
    std::vector<pic::IncidentUpdate> incB, incE;   // updateBHalf / updateE launches, face order
    int64_t* stepDev = nullptr;                    // field step counter (the solver's currentStep)
    float* B[3] = {nullptr, nullptr, nullptr};
    float* J[3] = {nullptr, nullptr, nullptr};
    std::vector<SpeciesMem> sp;
    pic::StepFlags* flags = nullptr;  // device step state (capacity overflow, halt)
    uint32_t hostFlags = 0;           // PIC_FLAG_* raised since the last pic_error_flags
    bool broken = false;              // particles were lost: the context refuses further steps
    double* energyPartials = nullptr;
    int energyBlocks = 0;
    unsigned long long* countOut = nullptr;
    double* diagBuf = nullptr;  // ncells doubles, allocated on first use
    int jexp = 40;
    // [supp-2][fused]
    PhasePlan plans[3][2];
    int cus = 256;          // compute units of the device
    int pushSchedule = 0;   // PIC_PUSH_SCHEDULE_*
    bool member = false;    // owned by an ensemble, whose device tables hold its plans' scratch
    int32_t* dAll = nullptr;
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[PIC_TIMER_COUNT];
    std::vector<hipEvent_t> eventPool;
    double totalMs[PIC_TIMER_COUNT] = {0, 0, 0};
    int64_t launches[PIC_TIMER_COUNT] = {0, 0, 0};
    std::vector<void*> allocations;
    // captured steps for pic_step_graph, one per E buffer parity
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graphExec[2] = {nullptr, nullptr};
    int graphJexp[2] = {0, 0};
    // differentiable form
    AdjointMem adj;
};

namespace {

template <typename T>
int devAlloc(pic_ctx* c, T** p, size_t n, bool zero = true) {
    void* q = nullptr;
    size_t const bytes = std::max<size_t>(n * sizeof(T), 16);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(PIC_E_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    if (zero) {
        e = hipMemsetAsync(q, 0, bytes, c->stream);
        if (e != hipSuccess) return fail(PIC_E_HIP, std::string("hipMemsetAsync failed: ") + hipGetErrorString(e));
    }
    c->allocations.push_back(q);
    *p = static_cast<T*>(q);
    return PIC_OK;
}

hipEvent_t takeEvent(pic_ctx* c) {
    if (!c->eventPool.empty()) {
        hipEvent_t e = c->eventPool.back();
        c->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct TimerScope {
    pic_ctx* c;
    int t;
    hipEvent_t a = nullptr, b = nullptr;
    TimerScope(pic_ctx* ctx, int timer) : c(ctx), t(timer) {
        if (c->timing) {
            a = takeEvent(c);
            b = takeEvent(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~TimerScope() {
        if (c->timing && a && b) {
            (void)hipEventRecord(b, c->stream);
            c->pending[t].push_back({a, b});
        }
    }
};

int resolveTimers(pic_ctx* c) {
    PIC_HIP(hipStreamSynchronize(c->stream));
    for (int t = 0; t < PIC_TIMER_COUNT; ++t) {
        for (auto& pr : c->pending[t]) {
            float ms = 0.0f;
            PIC_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            c->totalMs[t] += ms;
            c->launches[t] += 1;
            c->eventPool.push_back(pr.first);
            c->eventPool.push_back(pr.second);
        }
        c->pending[t].clear();
    }
    return PIC_OK;
}

// The J tile of a supercell spans the reference's deposit margins in both
// forms: a fused deposit runs in the frame of the particle's new cell, but
// its cells lie within the margins of the old one (the path starts in the
// old cell and moves less than a cell per axis).
int buildPlan(pic_ctx* c, int supp, bool fused, PhasePlan& plan) {
    (void)fused;
    int const lo = depLower(supp);
    int const up = depUpper(supp);
    const int scs[3] = {pic::kSCX, pic::kSCY, pic::kSCZ};
    const int nsc[3] = {c->g.nscx, c->g.nscy, c->g.nscz};
    std::vector<int> col[3];
    for (int d = 0; d < 3; ++d) {
        int const stride = (lo + up + scs[d] - 1) / scs[d] + 1;
        if (nsc[d] < stride)
            return fail(PIC_E_INVALID, "grid has fewer supercells along an axis than the current stencil needs");
        plan.ncol[d] = ringColours(nsc[d], stride, col[d]);
    }
    int const nph = plan.ncol[0] * plan.ncol[1] * plan.ncol[2];
    std::vector<std::vector<int>> members(nph);
    for (int z = 0; z < nsc[2]; ++z)
        for (int y = 0; y < nsc[1]; ++y)
            for (int x = 0; x < nsc[0]; ++x) {
                int const ph = col[0][x] + plan.ncol[0] * (col[1][y] + plan.ncol[1] * col[2][z]);
                members[ph].push_back(x + nsc[0] * (y + nsc[1] * z));
            }
    // launch order: workgroup b runs on XCD b % 8; give each XCD a contiguous
    // run of spatially ordered supercells so neighbouring tiles share its L2
    std::vector<int> list;
    plan.offset.clear();
    plan.count.clear();
    for (int ph = 0; ph < nph; ++ph) {
        const std::vector<int>& m = members[ph];
        int const n = static_cast<int>(m.size());
        std::vector<int> order(n);
        int pos = 0;
        for (int x = 0; x < 8; ++x) {
            for (int b = x; b < n; b += 8) order[b] = m[pos++];
        }
        plan.offset.push_back(static_cast<int>(list.size()));
        plan.count.push_back(n);
        list.insert(list.end(), order.begin(), order.end());
    }
    plan.host = list;
    if (devAlloc(c, &plan.dList, list.size()) != PIC_OK) return PIC_E_HIP;
    PIC_HIP(hipMemcpyAsync(plan.dList, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    // one launch for every supercell when the largest phase would leave more
    // than half of the CUs' workgroup slots idle (small boxes)
    int const most = *std::max_element(plan.count.begin(), plan.count.end());
    plan.single = fused && (c->pushSchedule == PIC_PUSH_SCHEDULE_SINGLE ||
                            (c->pushSchedule == PIC_PUSH_SCHEDULE_AUTO && most < 2 * c->cus));
    if (plan.single) {
        std::vector<uint8_t> colour;
        for (int d = 0; d < 3; ++d) {
            if (plan.ncol[d] > 255) return fail(PIC_E_INVALID, "more than 255 checkerboard colours");
            for (int v : col[d]) colour.push_back(static_cast<uint8_t>(v));
        }
        if (devAlloc(c, &plan.tiles, static_cast<size_t>(c->nsc) * pic::pushTileElems(supp), false) != PIC_OK ||
            devAlloc(c, &plan.colour, colour.size(), false) != PIC_OK)
            return PIC_E_HIP;
        PIC_HIP(hipMemcpyAsync(plan.colour, colour.data(), colour.size(), hipMemcpyHostToDevice, c->stream));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    plan.valid = true;
    return PIC_OK;
}

int getPlan(pic_ctx* c, int supp, bool fused, PhasePlan** out) {
    PhasePlan& p = c->plans[supp - 2][fused ? 1 : 0];
    if (!p.valid) {
        int const r = buildPlan(c, supp, fused, p);
        if (r != PIC_OK) return r;
    }
    *out = &p;
    return PIC_OK;
}

void updateJexp(pic_ctx* c) {
    if (c->cfg.j_scale_exp != 0) {
        c->jexp = c->cfg.j_scale_exp;
        return;
    }
    // largest magnitude one particle can add to a J cell: |q| * 1/(V dt) * max cell edge
    double bound = 0.0;
    float const maxCell = std::max(c->k.cell[0], std::max(c->k.cell[1], c->k.cell[2]));
    for (const SpeciesMem& s : c->sp) {
        if (s.desc.pusher == PIC_PUSHER_NONE) continue;
        double const b = std::fabs((double)s.charge) * s.wmax * (double)c->k.inv_cell_volume_dt * (double)maxCell;
        bound = std::max(bound, b);
    }
    if (!(bound > 0.0)) {
        c->jexp = 30;
        return;
    }
    // one contribution below 2^30: a tile element's sum over the particles of
    // a supercell (at most 4 species x 255 frames, two values per particle)
    // stays below 2^49, which the tile's decoding needs (below 2^50)
    int const e = 30 - static_cast<int>(std::ceil(std::log2(bound)));
    c->jexp = std::max(-60, std::min(100, e));
}

pic::PushArgs makePushArgs(pic_ctx* c) {
    pic::PushArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) {
        a.E[d] = c->E[d];
        a.B[d] = c->B[d];
        a.J[d] = c->J[d];
        a.cell[d] = c->k.cell[d];
    }
    a.dt = c->k.dt;
    a.c = c->k.c;
    a.c2 = c->k.c * c->k.c;
    a.rc2 = c->k.mue0_eps0;
    a.inv_vdt = c->k.inv_cell_volume_dt;
    a.jscale = std::ldexp(1.0, c->jexp);
    a.jinv = std::ldexp(1.0, -c->jexp);
    a.damp = c->pmlDamp;
    return a;
}

int applyIncident(pic_ctx* c, const std::vector<pic::IncidentUpdate>& list, float* const F[3], float stepOffset) {
    for (const pic::IncidentUpdate& u : list)
        PIC_HIP(pic::launchIncident(u, c->g, F, c->stepDev, stepOffset, c->stream));
    return PIC_OK;
}

pic::FieldArgs makeFieldArgs(pic_ctx* c) {
    pic::FieldArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) {
        a.E[d] = c->E[d];
        a.B[d] = c->B[d];
        a.J[d] = c->J[d];
        a.cell[d] = c->k.cell[d];
    }
    a.dt = c->k.dt;
    a.c2 = c->k.c * c->k.c;
    a.coeff = c->k.add_current_coeff;
    a.lehe_alpha = c->k.lehe_alpha;
    a.lehe_delta = c->k.lehe_delta;
    a.lehe_dir = c->cfg.lehe_dir;
    if (c->cfg.solver == PIC_SOLVER_LEHE) {
        int const d0 = c->cfg.lehe_dir, d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        float const r1 = c->k.cell[d0] / c->k.cell[d1];
        float const r2 = c->k.cell[d0] / c->k.cell[d2];
        a.lehe_beta1 = 0.125f * r1 * r1;
        a.lehe_beta2 = 0.125f * r2 * r2;
    }
    return a;
}

int checkSpecies(pic_ctx* c, int32_t s) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (s < 0 || s >= static_cast<int>(c->sp.size())) return fail(PIC_E_INVALID, "species index out of range");
    return PIC_OK;
}

// Mobile species grouped by (support, pusher, current solver), groups ordered by their first species.
std::vector<std::vector<int>> pushGroups(pic_ctx* c, int32_t species, bool byPusher) {
    std::vector<std::vector<int>> groups;
    std::vector<std::array<int, 3>> keys;
    for (int s = 0; s < static_cast<int>(c->sp.size()); ++s) {
        if (species >= 0 && s != species) continue;
        const SpeciesMem& m = c->sp[s];
        if (m.desc.pusher == PIC_PUSHER_NONE) continue;
        std::array<int, 3> key{supportOf(m.desc.shape), byPusher ? m.desc.pusher : 0, m.desc.current_solver};
        size_t gi = 0;
        for (; gi < keys.size(); ++gi)
            if (keys[gi] == key) break;
        if (gi == keys.size()) {
            keys.push_back(key);
            groups.emplace_back();
        }
        groups[gi].push_back(s);
    }
    return groups;
}

int runPush(pic_ctx* c, int32_t species, int mode, bool stageShift = false) {
    bool const byPusher = mode != 2;
    auto groups = pushGroups(c, species, byPusher);
    for (const auto& grp : groups) {
        int const supp = supportOf(c->sp[grp[0]].desc.shape);
        int const pusher = c->sp[grp[0]].desc.pusher;
        int const scheme = c->sp[grp[0]].desc.current_solver;
        pic::PushArgs a = makePushArgs(c);
        a.nsp = static_cast<int>(grp.size());
        for (size_t i = 0; i < grp.size(); ++i) a.sp[i] = c->sp[grp[i]].dev;
        a.stageShift = stageShift ? 1 : 0;
        if (mode != 2)
            for (int s : grp) c->sp[s].shiftReady = c->sp[s].dev.lists != 0 ? (stageShift ? 2 : 1) : 0;
        if (mode == 0) {
            a.scList = c->dAll;
            TimerScope ts(c, PIC_TIMER_PUSH);
            PIC_HIP(pic::launchPush(supp, pusher, 0, scheme, a, c->nsc, c->stream));
            continue;
        }
        PhasePlan* plan = nullptr;
        int const r = getPlan(c, supp, mode == 1, &plan);
        if (r != PIC_OK) return r;
        if (plan->single) {
            // every supercell in one launch (phase-ordered list), then the
            // tiles gathered into J in phase order
            a.scList = plan->dList;
            a.tileOut = plan->tiles;
            TimerScope ts(c, PIC_TIMER_PUSH);
            PIC_HIP(pic::launchPush(supp, pusher, mode, scheme, a, c->nsc, c->stream));
            PIC_HIP(pic::launchGatherFlush(supp, c->g, plan->tiles, plan->colour, c->J, c->stream));
            continue;
        }
        for (size_t ph = 0; ph < plan->count.size(); ++ph) {
            a.scList = plan->dList + plan->offset[ph];
            TimerScope ts(c, PIC_TIMER_PUSH);
            PIC_HIP(pic::launchPush(supp, pusher, mode, scheme, a, plan->count[ph], c->stream));
        }
    }
    return PIC_OK;
}

// EnergyFields::getEnergyFields (plugins/EnergyFields.x.cpp:183-237): the
// per-component sums of squares times 0.5/mue0*V (B) and eps0*V*0.5 (E) in
// the reference's float/double mix (mue0, eps0 and the cell volume are
// float_X), added per field and converted to joules with unit.energy()
// (unitless/simulation.unitless:455-458). sums: B x, y, z, E x, y, z.
// out: total, B x, y, z, E x, y, z.
void fieldEnergyFromSums(const pic_ctx* c, const double sums[6], double out[7]) {
    float const volume = c->k.cell[0] * c->k.cell[1] * c->k.cell[2];  // productOfComponents
    double const bFactor = 0.5 / static_cast<double>(c->k.mue0) * static_cast<double>(volume);
    float const eV = c->k.eps0 * volume;
    double const eFactor = static_cast<double>(eV) * 0.5;
    double const unitEnergy = c->k.unit_mass * c->k.unit_length * c->k.unit_length / (c->k.unit_time * c->k.unit_time);
    double eB = 0.0, eE = 0.0;
    for (int d = 0; d < 3; ++d) {
        double const b = sums[d] * bFactor;
        double const e = sums[3 + d] * eFactor;
        eB += b;
        eE += e;
        out[1 + d] = b * unitEnergy;
        out[4 + d] = e * unitEnergy;
    }
    out[0] = (eE + eB) * unitEnergy;
}

// the box of cells outside every PML layer, for the PML kernels
pic::FieldArgs withPmlInner(const pic_ctx* c, pic::FieldArgs a) {
    a.inner = 1;
    for (int d = 0; d < 3; ++d) a.in0[d] = c->pmlIn0[d], a.in1[d] = c->pmlIn1[d];
    return a;
}

pic::FieldArgs pmlFieldArgs(pic_ctx* c) { return withPmlInner(c, makeFieldArgs(c)); }

// the field step runs as the two fused kernels (no incident field between its stages)
bool fusedFieldStep(const pic_ctx* c) { return c->incB.empty() && c->incE.empty() && !c->hasPml; }
// the PML's field step without an incident field: B half step, E full step
// with J into Ealt, B half step
bool pmlFieldStep(const pic_ctx* c) { return c->incB.empty() && c->incE.empty() && c->hasPml; }
// the field step leaves the new E in Ealt and swaps the buffers
bool fieldStepSwapsE(const pic_ctx* c) { return fusedFieldStep(c) || pmlFieldStep(c); }

// B half step of update_beforeCurrent (second half: PML's psi of B stays) or
// of update_afterCurrent (first half: psi of B advances), FDTDBase.hpp:
// updateBSecondHalf / updateBFirstHalf
int bHalfStage(pic_ctx* c, const pic::FieldArgs& a, bool firstHalf) {
    if (c->hasPml)
        PIC_HIP(pic::launchBHalfPml(c->cfg.solver, withPmlInner(c, a), c->pml, firstHalf, c->stream));
    else
        PIC_HIP(pic::launchBHalf(c->cfg.solver, a, c->stream));
    return PIC_OK;
}

void dropGraphs(pic_ctx* c) {
    for (int p = 0; p < 2; ++p) {
        if (c->graphExec[p]) (void)hipGraphExecDestroy(c->graphExec[p]);
        if (c->graph[p]) (void)hipGraphDestroy(c->graph[p]);
        c->graphExec[p] = nullptr;
        c->graph[p] = nullptr;
    }
}

void devFree(pic_ctx* c, void* p) {
    if (!p) return;
    auto it = std::find(c->allocations.begin(), c->allocations.end(), p);
    if (it != c->allocations.end()) c->allocations.erase(it);
    (void)hipFree(p);
}

int grownFrames(int frames, uint32_t count, int maxFrames);

// Contiguous runs of the given lengths in supercell order, in a pool of
// `pool` slots: the host tables (base, capSc, used, cap; slots = pool).
void contiguousRuns(SpeciesMem& m, const std::vector<uint32_t>& caps, size_t pool) {
    m.capSc = caps;
    m.base.assign(caps.size(), 0);
    uint64_t at = 0;
    for (size_t i = 0; i < caps.size(); ++i) {
        m.base[i] = at;
        at += caps[i];
    }
    m.used = static_cast<size_t>(at);
    m.slots = pool;
    m.cap = static_cast<int>(*std::max_element(caps.begin(), caps.end()));
}

// the run tables to the device
int uploadRuns(pic_ctx* c, SpeciesMem& m) {
    PIC_HIP(hipMemcpyAsync(m.dBase, m.base.data(), m.base.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                           c->stream));
    PIC_HIP(hipMemcpyAsync(m.dCap, m.capSc.data(), m.capSc.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           c->stream));
    return PIC_OK;
}

// Sets the slot layout of a new species: contiguous runs of `caps` slots.
int setLayout(pic_ctx* c, SpeciesMem& m, const std::vector<uint32_t>& caps) {
    uint64_t const total = std::accumulate(caps.begin(), caps.end(), uint64_t(0));
    // an eighth more slots than the runs: room for the first supercells to grow
    contiguousRuns(m, caps, static_cast<size_t>(total + total / 8));
    if (devAlloc(c, &m.dBase, caps.size()) != PIC_OK || devAlloc(c, &m.dCap, caps.size()) != PIC_OK)
        return PIC_E_HIP;
    return uploadRuns(c, m);
}

// Grows the supercells of species s to the run lengths `caps` (slots, at
// least the present ones), as the reference takes further frames from its
// heap for a supercell that needs them (ParticlesBox.hpp:96-122). A growing
// supercell's run moves, slot order kept, to a new run at the end of the
// pool; its old run is left unused. The outbox moves along, so records an
// overflowing insert left there can still be placed. The arrays stay where
// they are, so captured steps stay valid (they read the run tables, which
// change in place). Only when the pool is full are the arrays reallocated
// with every run packed and a quarter more slots (the captured steps are
// then dropped).
int growSpecies(pic_ctx* c, int s, const std::vector<uint32_t>& caps) {
    SpeciesMem& m = c->sp[s];
    std::vector<pic::RunMove> moves;
    uint64_t grow = 0;
    for (int i = 0; i < c->nsc; ++i) {
        if (caps[i] < m.capSc[i]) return fail(PIC_E_INVALID, "a supercell cannot shrink");
        if (caps[i] > m.capSc[i]) grow += caps[i];
    }
    if (grow == 0) return PIC_OK;
    pic::SpeciesDev& d = m.dev;
    int r = PIC_OK;
    pic::RunMove* dMoves = nullptr;
    auto moveAll = [&](pic::SpeciesDev& to, const pic::SpeciesDev& from) -> int {
        if (devAlloc(c, &dMoves, moves.size(), false) != PIC_OK) return PIC_E_HIP;
        PIC_HIP(hipMemcpyAsync(dMoves, moves.data(), moves.size() * sizeof(pic::RunMove), hipMemcpyHostToDevice,
                               c->stream));
        int const n = static_cast<int>(moves.size());
        for (int a = 0; a < 3; ++a) {
            PIC_HIP(pic::launchMoveRuns(to.pos[a], from.pos[a], 4, dMoves, n, c->stream));
            PIC_HIP(pic::launchMoveRuns(to.mom[a], from.mom[a], 4, dMoves, n, c->stream));
        }
        PIC_HIP(pic::launchMoveRuns(to.w, from.w, 4, dMoves, n, c->stream));
        PIC_HIP(pic::launchMoveRuns(to.cell, from.cell, 2, dMoves, n, c->stream));
        PIC_HIP(pic::launchMoveRuns(to.mask, from.mask, 1, dMoves, n, c->stream));
        PIC_HIP(pic::launchMoveRuns(to.outbox, from.outbox, 32, dMoves, n, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        devFree(c, dMoves);
        return PIC_OK;
    };
    if (m.used + grow <= m.slots) {
        // new runs at the end of the pool, in supercell order
        for (int i = 0; i < c->nsc; ++i) {
            if (caps[i] == m.capSc[i]) continue;
            moves.push_back(pic::RunMove{m.base[i], m.used, m.capSc[i], 0});
            m.base[i] = m.used;
            m.used += caps[i];
        }
        if ((r = moveAll(d, d)) != PIC_OK) return r;
        m.capSc = caps;
        m.cap = static_cast<int>(*std::max_element(caps.begin(), caps.end()));
        m.desc.frames_per_sc = m.cap / pic::kSCV;
        c->cfg.species[s].frames_per_sc = m.desc.frames_per_sc;
        d.cap = m.cap;
        int const lists = m.cap <= pic::kListMaxSlots ? 1 : 0;
        if (lists != d.lists) {  // captured steps hold the old flag
            d.lists = lists;
            dropGraphs(c);
        }
        return uploadRuns(c, m);
    }
    // repack into fresh arrays with a quarter more slots
    uint64_t const total = std::accumulate(caps.begin(), caps.end(), uint64_t(0));
    size_t const slots = static_cast<size_t>(total + total / 4);
    pic::SpeciesDev n = d;
    std::vector<void*> fresh;
    // the kernels read a supercell's slots below its count and the lists
    // they wrote themselves; only the multiMask of new slots is cleared
    auto alloc = [&](auto** p, size_t count, bool zero) {
        if (devAlloc(c, p, count, zero) != PIC_OK) return false;
        fresh.push_back(*p);
        return true;
    };
    bool ok = true;
    for (int a = 0; a < 3 && ok; ++a) ok = alloc(&n.pos[a], slots, false) && alloc(&n.mom[a], slots, false);
    ok = ok && alloc(&n.w, slots, false) && alloc(&n.cell, slots, false) && alloc(&n.mask, slots, true) &&
         alloc(&n.outbox, 2 * slots, false) && alloc(&n.outIdx, slots, false) && alloc(&n.scratch, slots, false);
    if (!ok) {
        for (void* p : fresh) devFree(c, p);
        return fail(PIC_E_CAPACITY, "supercells of species " + std::to_string(s) + " cannot grow to " +
                                        std::to_string(slots) + " slots: " + g_err);
    }
    std::vector<uint64_t> const oldBase = m.base;
    std::vector<uint32_t> const oldCap = m.capSc;
    contiguousRuns(m, caps, slots);
    for (int i = 0; i < c->nsc; ++i) moves.push_back(pic::RunMove{oldBase[i], m.base[i], oldCap[i], 0});
    if ((r = moveAll(n, d)) != PIC_OK) return r;
    for (int a = 0; a < 3; ++a) {
        devFree(c, d.pos[a]);
        devFree(c, d.mom[a]);
    }
    devFree(c, d.w);
    devFree(c, d.cell);
    devFree(c, d.mask);
    devFree(c, d.outbox);
    devFree(c, d.outIdx);
    devFree(c, d.scratch);
    n.cap = m.cap;
    n.lists = m.cap <= pic::kListMaxSlots ? 1 : 0;
    d = n;
    m.desc.frames_per_sc = m.cap / pic::kSCV;
    c->cfg.species[s].frames_per_sc = m.desc.frames_per_sc;
    dropGraphs(c);  // captured steps hold the old arrays
    return uploadRuns(c, m);
}

// capacities with every supercell that must hold more than its slots
// (`need`) grown to grownFrames; false when nothing changes
bool grownCaps(const SpeciesMem& m, const std::vector<uint32_t>& need, int maxFrames, std::vector<uint32_t>& caps) {
    caps = m.capSc;
    bool any = false;
    for (size_t i = 0; i < caps.size(); ++i) {
        if (need[i] <= caps[i]) continue;
        caps[i] = static_cast<uint32_t>(grownFrames(static_cast<int>(caps[i] / pic::kSCV), need[i], maxFrames)) *
                  pic::kSCV;
        any = true;
    }
    return any;
}

// the most frames a supercell of a species may have (the configuration's
// bound); the current tile's fixed-point range assumes it
constexpr int kMaxFrames = 255;

// grows the supercells of species s that must hold more than their slots
// (`need` per supercell), at most to maxFrames frames
int growFor(pic_ctx* c, int s, const std::vector<uint32_t>& need, int maxFrames) {
    std::vector<uint32_t> caps;
    if (!grownCaps(c->sp[s], need, maxFrames, caps)) return PIC_OK;
    for (size_t i = 0; i < caps.size(); ++i)
        if (need[i] > caps[i])
            return fail(PIC_E_CAPACITY, "a supercell of species " + std::to_string(s) + " would need " +
                                            std::to_string(need[i]) + " slots, more than " +
                                            std::to_string(maxFrames) + " frames");
    return growSpecies(c, s, caps);
}

// frames for a supercell that must hold `count` particles: geometric growth
// (half again) with at least two spare frames, at most maxFrames (0: no limit)
int grownFrames(int frames, uint32_t count, int maxFrames) {
    int const need = static_cast<int>((count + pic::kSCV - 1) / pic::kSCV);
    int f = std::max(need + 2, frames + (frames + 1) / 2);
    if (maxFrames > 0) f = std::min(f, maxFrames);
    return f;
}

// After steps: when a shift found a supercell full, the insert kept the true
// count and left the records it could not place in the senders' outboxes, and
// the steps after it were skipped. Grow the supercells of every species that
// needs it, insert again (the records land in the slots a large enough
// capacity would have given them, so the result is the one of that capacity
// bit for bit), clear the flags and return the number of skipped steps in
// *skipped. Each skipped fused field step swapped E and Ealt on the host
// without writing, so an odd count is swapped back. maxFrames bounds the
// growth (0: none); beyond it the counts are clamped, particles are lost and
// the context refuses further steps.
int recoverCapacity(pic_ctx* c, int* skipped, int maxFrames) {
    *skipped = 0;
    pic::StepFlags f{};
    PIC_HIP(hipMemcpyAsync(&f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    if (f.overflow == 0) return PIC_OK;
    int r = PIC_OK;
    std::vector<uint32_t> counts(c->nsc);
    for (int s = 0; s < static_cast<int>(c->sp.size()); ++s) {
        SpeciesMem& m = c->sp[s];
        if (m.desc.pusher == PIC_PUSHER_NONE) continue;
        PIC_HIP(hipMemcpyAsync(counts.data(), m.dev.count, c->nsc * 4, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        // supercells already filled to 7/8 grow along with the overflowing
        // ones: the next overflows (a moving compression front) come later
        std::vector<uint32_t> need(counts);
        bool over = false;
        for (int i = 0; i < c->nsc; ++i) {
            over = over || counts[i] > m.capSc[i];
            if (counts[i] <= m.capSc[i] && counts[i] * 8ull >= m.capSc[i] * 7ull) need[i] = m.capSc[i] + 1;
        }
        std::vector<uint32_t> caps;
        if (!over || !grownCaps(m, need, maxFrames, caps)) continue;
        for (int i = 0; i < c->nsc; ++i) caps[i] = std::max(caps[i], m.capSc[i]);
        uint32_t bad = 0;
        for (int i = 0; i < c->nsc; ++i)
            if (counts[i] > caps[i]) bad = std::max(bad, counts[i]);
        if (bad)
            r = fail(PIC_E_CAPACITY, "a supercell of species " + std::to_string(s) + " needs " + std::to_string(bad) +
                                         " slots, more than the " + std::to_string(maxFrames) + " frames allowed here");
        else
            r = growSpecies(c, s, caps);
        if (r != PIC_OK) {
            // particles beyond the capacity are lost: keep the kernels in bounds
            for (SpeciesMem& q : c->sp)
                if (q.desc.pusher != PIC_PUSHER_NONE) PIC_HIP(pic::launchClampCounts(q.dev, c->nsc, c->stream));
            c->broken = true;
            break;
        }
        PIC_HIP(pic::launchInsertRedo(c->g, m.dev, c->stream));
    }
    PIC_HIP(hipMemsetAsync(c->flags, 0, sizeof(pic::StepFlags), c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->hostFlags |= PIC_FLAG_CAPACITY_GREW;
    if ((f.skipped & 1u) && fieldStepSwapsE(c)) {
        for (int d = 0; d < 3; ++d) std::swap(c->E[d], c->Ealt[d]);
        c->eParity ^= 1;
    }
    *skipped = static_cast<int>(f.skipped);
    return r;
}

}  // namespace

extern "C" {

const char* pic_last_error(void) { return g_err.c_str(); }

int32_t pic_abi_version(void) { return PIC_ABI_VERSION; }

int pic_constants_from_config(const pic_config* cfg, pic_constants* out) {
    if (!cfg || !out) return fail(PIC_E_INVALID, "null argument");
    int const r = validateConfig(*cfg);
    if (r != PIC_OK) return r;
    computeConstants(*cfg, *out);
    return PIC_OK;
}

int pic_create(const pic_config* cfg, pic_ctx** out) {
    if (!cfg || !out) return fail(PIC_E_INVALID, "null argument");
    *out = nullptr;
    int r = validateConfig(*cfg);
    if (r != PIC_OK) return r;
    pic_ctx* c = new pic_ctx();
    c->cfg = *cfg;
    computeConstants(*cfg, c->k);
    r = checkCfl(*cfg, c->k);
    if (r != PIC_OK) {
        delete c;
        return r;
    }
    c->g.nx = cfg->grid[0];
    c->g.ny = cfg->grid[1];
    c->g.nz = cfg->grid[2];
    c->g.nscx = cfg->grid[0] / pic::kSCX;
    c->g.nscy = cfg->grid[1] / pic::kSCY;
    c->g.nscz = cfg->grid[2] / pic::kSCZ;
    for (int d = 0; d < 3; ++d)
        if (cfg->boundary[d] == PIC_BOUNDARY_OPEN) c->g.open |= 1 << d;
    c->nsc = c->g.nscx * c->g.nscy * c->g.nscz;
    c->ncells = static_cast<size_t>(cfg->grid[0]) * cfg->grid[1] * cfg->grid[2];
    c->device = cfg->device;
    auto bail = [&](int code) {
        pic_destroy(c);
        return code;
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return bail(fail(PIC_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(fail(PIC_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    {
        hipDeviceProp_t prop{};
        if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
            c->cus = prop.multiProcessorCount;
    }
    for (int d = 0; d < 3; ++d) {
        if (devAlloc(c, &c->E[d], c->ncells) || devAlloc(c, &c->B[d], c->ncells) || devAlloc(c, &c->J[d], c->ncells) ||
            devAlloc(c, &c->Ealt[d], c->ncells))
            return bail(PIC_E_HIP);
    }
    if (devAlloc(c, &c->flags, 1)) return bail(PIC_E_HIP);
    c->g.flags = c->flags;
    if (cfg->current_interpolation == PIC_CURRENT_BINOMIAL)
        for (int d = 0; d < 3; ++d)
            if (devAlloc(c, &c->Jf[d], c->ncells)) return bail(PIC_E_HIP);
    if (cfg->absorber == PIC_ABSORBER_EXPONENTIAL) {
        // ExponentialImpl::run (Exponential.hpp:52-95): plane exchanges 1, 2,
        // 3, 6, 9, 18 in this order (RIGHT +x, LEFT -x, BOTTOM +y, TOP -y,
        // BACK +z, FRONT -z: pmacc/type/Exchange.hpp:48-53, an odd code is
        // the max side, pmacc/boundary/Utility.hpp isMaxSide), on borders
        // without a neighbour (open axes), thickness and strength read at
        // [axis][exchange % 2] ({negative, positive} of fieldAbsorber.param);
        // the factors exp(-strength * f) of Exponential.kernel:99 in float
        constexpr int kExchanges[6] = {1, 2, 3, 6, 9, 18};
        constexpr int kStride = PIC_ABSORBER_MAX_CELLS + 1;
        std::vector<float> table(6 * kStride, 0.0f);
        float* dTable = nullptr;
        if (devAlloc(c, &dTable, table.size())) return bail(PIC_E_HIP);
        for (int l = 0; l < 6; ++l) {
            int const ex = kExchanges[l];
            int const axis = ex >= 9 ? 2 : (ex >= 3 ? 1 : 0);
            int const th = cfg->absorber_cells[axis][ex % 2];
            float const strength = cfg->absorber_strength[axis][ex % 2];
            if (cfg->boundary[axis] != PIC_BOUNDARY_OPEN || th <= 1) continue;
            for (int f = 1; f < th; ++f) table[l * kStride + f] = std::exp(-strength * static_cast<float>(f));
            pic::AbsorbLayer& L = c->absorb.layer[c->absorb.n++];
            L.axis = axis;
            L.minSide = ex % 2 == 0 ? 1 : 0;
            L.depth = th - 1;
            L.table = dTable + l * kStride;
        }
        e = hipMemcpyAsync(dTable, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(e)));
    }
    if (cfg->absorber == PIC_ABSORBER_PML) {
        std::vector<float> tabE, tabB;
        pmlTables(*cfg, c->k, false, tabE);
        pmlTables(*cfg, c->k, true, tabB);
        float* dE = nullptr;
        float* dB = nullptr;
        if (devAlloc(c, &dE, tabE.size()) || devAlloc(c, &dB, tabB.size())) return bail(PIC_E_HIP);
        for (int q = 0; q < 6; ++q)
            if (devAlloc(c, &c->pml.psiE[q], c->ncells) || devAlloc(c, &c->pml.psiB[q], c->ncells))
                return bail(PIC_E_HIP);
        e = hipMemcpyAsync(dE, tabE.data(), tabE.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dB, tabB.data(), tabB.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(e)));
        c->pml.tabE = dE;
        c->pml.tabB = dB;
        c->hasPml = true;
        pmlInnerBox(*cfg, tabE, tabB, c->pmlIn0, c->pmlIn1);
        c->pmlDamp = pmlDampArgs(*cfg, c->k);
        int const tiles = pic::pmlTileCount(c->g);
        if (tiles > 0) {
            std::vector<int> list(static_cast<size_t>(tiles));
            int nPlain = 0;
            pic::pmlTileSplit(c->g, c->pmlIn0, c->pmlIn1, list.data(), &nPlain);
            int* dList = nullptr;
            if (devAlloc(c, &dList, list.size())) return bail(PIC_E_HIP);
            e = hipMemcpyAsync(dList, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(e)));
            c->pml.tiles = dList;
            c->pml.nPlain = nPlain;
            c->pml.nLayer = tiles - nPlain;
        }
    }
    if (devAlloc(c, &c->stepDev, 1)) return bail(PIC_E_HIP);
    c->incB = incidentUpdates(*cfg, c->k, true);
    c->incE = incidentUpdates(*cfg, c->k, false);
    c->energyBlocks = 1024;
    if (devAlloc(c, &c->energyPartials, 6 * c->energyBlocks + 6)) return bail(PIC_E_HIP);
    if (devAlloc(c, &c->countOut, 1)) return bail(PIC_E_HIP);
    std::vector<int32_t> all(c->nsc);
    for (int i = 0; i < c->nsc; ++i) all[i] = i;
    if (devAlloc(c, &c->dAll, all.size())) return bail(PIC_E_HIP);
    e = hipMemcpyAsync(c->dAll, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(e)));
    for (int s = 0; s < cfg->n_species; ++s) {
        SpeciesMem m;
        m.desc = cfg->species[s];
        m.mass = c->k.species_mass[s];
        m.charge = c->k.species_charge[s];
        if (setLayout(c, m, std::vector<uint32_t>(c->nsc, static_cast<uint32_t>(cfg->species[s].frames_per_sc) *
                                                             pic::kSCV)) != PIC_OK)
            return bail(PIC_E_HIP);
        size_t const slots = m.slots;
        pic::SpeciesDev& d = m.dev;
        for (int a = 0; a < 3; ++a) {
            if (devAlloc(c, &d.pos[a], slots) || devAlloc(c, &d.mom[a], slots)) return bail(PIC_E_HIP);
        }
        if (devAlloc(c, &d.w, slots) || devAlloc(c, &d.cell, slots) || devAlloc(c, &d.mask, slots) ||
            devAlloc(c, &d.count, c->nsc) || devAlloc(c, &d.outbox, 2 * slots) || devAlloc(c, &d.outIdx, slots) ||
            devAlloc(c, &d.scratch, slots) || devAlloc(c, &d.outOff, static_cast<size_t>(c->nsc) * pic::kOffStride) ||
            devAlloc(c, &d.leaveBits, static_cast<size_t>(c->nsc) * pic::kPushThreadsPerBlock))
            return bail(PIC_E_HIP);
        d.base = m.dBase;
        d.capOf = m.dCap;
        d.cap = m.cap;
        d.lists = m.cap <= pic::kListMaxSlots ? 1 : 0;
        d.reflect = 0;
        for (int a = 0; a < 3; ++a)
            if (cfg->species[s].particle_boundary[a] == PIC_PARTICLE_REFLECTING) d.reflect |= 1 << a;
        d.mass = m.mass;
        d.charge = m.charge;
        c->sp.push_back(m);
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(e)));
    updateJexp(c);
    *out = c;
    return PIC_OK;
}

int pic_destroy(pic_ctx* c) {
    if (!c) return PIC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int t = 0; t < PIC_TIMER_COUNT; ++t)
        for (auto& pr : c->pending[t]) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : c->eventPool) (void)hipEventDestroy(e);
    for (int p = 0; p < 2; ++p) {
        if (c->graphExec[p]) (void)hipGraphExecDestroy(c->graphExec[p]);
        if (c->graph[p]) (void)hipGraphDestroy(c->graph[p]);
    }
    for (void* p : c->allocations) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PIC_OK;
}

int pic_get_constants(const pic_ctx* c, pic_constants* out) {
    if (!c || !out) return fail(PIC_E_INVALID, "null argument");
    *out = c->k;
    return PIC_OK;
}

int pic_sync(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int32_t pic_num_supercells(const pic_ctx* c) { return c ? c->nsc : -1; }

int32_t pic_slots_per_supercell(const pic_ctx* c, int32_t s) {
    if (!c || s < 0 || s >= static_cast<int>(c->sp.size())) return -1;
    return c->sp[s].cap;
}

int pic_supercell_slots(const pic_ctx* c, int32_t s, uint32_t* out) {
    if (!c || !out) return fail(PIC_E_INVALID, "null argument");
    if (s < 0 || s >= static_cast<int>(c->sp.size())) return fail(PIC_E_INVALID, "species index out of range");
    std::copy(c->sp[s].capSc.begin(), c->sp[s].capSc.end(), out);
    return PIC_OK;
}

int32_t pic_j_scale_exp(const pic_ctx* c) { return c ? c->jexp : 0; }

int pic_set_push_schedule(pic_ctx* c, int32_t schedule) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (schedule < PIC_PUSH_SCHEDULE_AUTO || schedule > PIC_PUSH_SCHEDULE_SINGLE)
        return fail(PIC_E_INVALID, "unknown push schedule");
    if (schedule == c->pushSchedule) return PIC_OK;
    if (c->member) return fail(PIC_E_INVALID, "an ensemble member keeps the ensemble's push schedule");
    PIC_HIP(hipSetDevice(c->device));
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->pushSchedule = schedule;
    dropGraphs(c);  // captured launches follow the old plans
    for (int supp = 0; supp < 3; ++supp) {
        PhasePlan& p = c->plans[supp][1];
        if (!p.valid) continue;
        devFree(c, p.dList);
        devFree(c, p.tiles);
        devFree(c, p.colour);
        p = PhasePlan{};
    }
    return PIC_OK;
}

int32_t pic_push_schedule(const pic_ctx* c, int32_t support) {
    if (!c || support < 2 || support > 4) return -1;
    PhasePlan* p = nullptr;
    if (getPlan(const_cast<pic_ctx*>(c), support, true, &p) != PIC_OK) return -1;
    return p->single ? PIC_PUSH_SCHEDULE_SINGLE : PIC_PUSH_SCHEDULE_PHASES;
}

int32_t pic_num_push_phases(const pic_ctx* c, int32_t fused) {
    if (!c) return -1;
    pic_ctx* m = const_cast<pic_ctx*>(c);
    int maxPh = 0;
    for (const SpeciesMem& s : m->sp) {
        if (s.desc.pusher == PIC_PUSHER_NONE) continue;
        PhasePlan* p = nullptr;
        if (getPlan(m, supportOf(s.desc.shape), fused != 0, &p) != PIC_OK) return -1;
        maxPh = std::max(maxPh, static_cast<int>(p->count.size()));
    }
    return maxPh;
}

int pic_load_particles(pic_ctx* c, int32_t s, int64_t n, const int32_t* cell, const float* pos, const float* mom,
                       const float* w) {
    int r = checkSpecies(c, s);
    if (r != PIC_OK) return r;
    if (n < 0 || (n > 0 && (!cell || !pos || !mom || !w))) return fail(PIC_E_INVALID, "null particle arrays");
    SpeciesMem& m = c->sp[s];
    std::vector<uint32_t> counts(c->nsc, 0);
    for (int64_t i = 0; i < n; ++i) {
        int const x = cell[3 * i], y = cell[3 * i + 1], z = cell[3 * i + 2];
        if (x < 0 || y < 0 || z < 0 || x >= c->g.nx || y >= c->g.ny || z >= c->g.nz)
            return fail(PIC_E_INVALID, "particle cell outside the grid");
        for (int a = 0; a < 3; ++a) {
            float const p = pos[3 * i + a];
            if (!(p >= 0.0f && p < 1.0f)) return fail(PIC_E_INVALID, "in-cell position outside [0,1)");
            if (!std::isfinite(mom[3 * i + a])) return fail(PIC_E_INVALID, "non-finite momentum");
        }
        if (!std::isfinite(w[i])) return fail(PIC_E_INVALID, "non-finite weighting");
        int const sc = x / pic::kSCX + c->g.nscx * (y / pic::kSCY + c->g.nscy * (z / pic::kSCZ));
        counts[sc] += 1;
    }
    // supercells that receive more particles than they hold grow first
    {
        if ((r = growFor(c, s, counts, kMaxFrames)) != PIC_OK) return r;
    }
    size_t const slots = m.slots;
    std::vector<float> hp[3], hm[3];
    for (int a = 0; a < 3; ++a) {
        hp[a].assign(slots, 0.0f);
        hm[a].assign(slots, 0.0f);
    }
    std::vector<float> hw(slots, 0.0f);
    std::vector<uint16_t> hc(slots, 0);
    std::vector<uint8_t> hk(slots, 0);
    double wmax = 0.0;
    std::fill(counts.begin(), counts.end(), 0u);
    for (int64_t i = 0; i < n; ++i) {
        int const x = cell[3 * i], y = cell[3 * i + 1], z = cell[3 * i + 2];
        int const sc = x / pic::kSCX + c->g.nscx * (y / pic::kSCY + c->g.nscy * (z / pic::kSCZ));
        int const lc = x % pic::kSCX + pic::kSCX * (y % pic::kSCY + pic::kSCY * (z % pic::kSCZ));
        size_t const g = static_cast<size_t>(m.base[sc]) + counts[sc]++;
        for (int a = 0; a < 3; ++a) {
            hp[a][g] = pos[3 * i + a];
            hm[a][g] = mom[3 * i + a];
        }
        hw[g] = w[i];
        hc[g] = static_cast<uint16_t>(lc);
        hk[g] = 1;
        wmax = std::max(wmax, std::fabs(static_cast<double>(w[i])));
    }
    r = pic_write_frames(c, s, counts.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    if (r != PIC_OK) return r;
    pic::SpeciesDev& d = m.dev;
    for (int a = 0; a < 3; ++a) {
        PIC_HIP(hipMemcpyAsync(d.pos[a], hp[a].data(), slots * 4, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(d.mom[a], hm[a].data(), slots * 4, hipMemcpyHostToDevice, c->stream));
    }
    PIC_HIP(hipMemcpyAsync(d.w, hw.data(), slots * 4, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(d.cell, hc.data(), slots * 2, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(d.mask, hk.data(), slots, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    m.wmax = wmax;
    updateJexp(c);
    return PIC_OK;
}

}  // extern "C"

namespace {

// whether every supercell of the species has `pitch` slots (the device
// layout is then the frame view itself)
bool uniformLayout(const SpeciesMem& m, int pitch) {
    for (size_t i = 0; i < m.capSc.size(); ++i)
        if (m.capSc[i] != static_cast<uint32_t>(pitch) || m.base[i] != i * static_cast<uint64_t>(pitch)) return false;
    return true;
}

// device array (the species' layout) -> frame view of `pitch` slots per
// supercell; slots a supercell does not have read as zero
template <typename T>
int layoutToView(pic_ctx* c, const SpeciesMem& m, int pitch, const T* dev, T* view) {
    size_t const n = static_cast<size_t>(c->nsc) * pitch;
    if (uniformLayout(m, pitch)) {
        PIC_HIP(hipMemcpyAsync(view, dev, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        return PIC_OK;
    }
    std::vector<T> tmp(m.slots);
    PIC_HIP(hipMemcpyAsync(tmp.data(), dev, m.slots * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    std::fill(view, view + n, T(0));
    for (int sc = 0; sc < c->nsc; ++sc) {
        size_t const k = std::min<size_t>(m.capSc[sc], static_cast<size_t>(pitch));
        std::copy(tmp.begin() + m.base[sc], tmp.begin() + m.base[sc] + k, view + static_cast<size_t>(sc) * pitch);
    }
    return PIC_OK;
}

// frame view of `pitch` slots per supercell -> device array (the species'
// layout); slots beyond the view's pitch are left empty
template <typename T>
int viewToLayout(pic_ctx* c, const SpeciesMem& m, int pitch, const T* view, T* dev) {
    if (uniformLayout(m, pitch)) {
        size_t const n = static_cast<size_t>(c->nsc) * pitch;
        PIC_HIP(hipMemcpyAsync(dev, view, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
        return PIC_OK;
    }
    std::vector<T> tmp(m.slots, T(0));
    for (int sc = 0; sc < c->nsc; ++sc) {
        size_t const k = std::min<size_t>(m.capSc[sc], static_cast<size_t>(pitch));
        const T* src = view + static_cast<size_t>(sc) * pitch;
        std::copy(src, src + k, tmp.begin() + m.base[sc]);
    }
    PIC_HIP(hipMemcpyAsync(dev, tmp.data(), m.slots * sizeof(T), hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_write_frames(pic_ctx* c, int32_t s, const uint32_t* counts, const float* pos, const float* mom, const float* w,
                     const uint16_t* lc, const uint8_t* mask) {
    int r = checkSpecies(c, s);
    if (r != PIC_OK) return r;
    SpeciesMem& m = c->sp[s];
    int const pitch = m.cap;  // the frame view's slots per supercell
    size_t const slots = static_cast<size_t>(c->nsc) * pitch;
    if (counts) {
        for (int i = 0; i < c->nsc; ++i)
            if (counts[i] > static_cast<uint32_t>(pitch)) return fail(PIC_E_CAPACITY, "count above frame capacity");
    }
    // the occupied slots of what is written must hold a particle the kernels
    // can take: local cell in the supercell, position in [0,1), finite
    // momentum and weighting, multiMask 1 or an exchange code 2..27
    if (pos || mom || w || lc || mask) {
        std::vector<uint32_t> n(c->nsc);
        if (counts) {
            std::copy(counts, counts + c->nsc, n.begin());
        } else {
            PIC_HIP(hipMemcpyAsync(n.data(), m.dev.count, c->nsc * 4, hipMemcpyDeviceToHost, c->stream));
            PIC_HIP(hipStreamSynchronize(c->stream));
        }
        for (int sc = 0; sc < c->nsc; ++sc)
            for (uint32_t k = 0; k < n[sc]; ++k) {
                size_t const i = static_cast<size_t>(sc) * pitch + k;
                if (lc && lc[i] >= pic::kSCV) return fail(PIC_E_INVALID, "localCellIdx outside the supercell");
                if (mask && (mask[i] < 1 || mask[i] > 27))
                    return fail(PIC_E_INVALID, "multiMask of an occupied slot outside 1..27");
                if (w && !std::isfinite(w[i])) return fail(PIC_E_INVALID, "non-finite weighting");
                for (int a = 0; a < 3; ++a) {
                    if (pos && !(pos[a * slots + i] >= 0.0f && pos[a * slots + i] < 1.0f))
                        return fail(PIC_E_INVALID, "in-cell position outside [0,1)");
                    if (mom && !std::isfinite(mom[a * slots + i])) return fail(PIC_E_INVALID, "non-finite momentum");
                }
            }
    }
    // supercells whose new count exceeds their slots grow first
    if (counts) {
        std::vector<uint32_t> need(counts, counts + c->nsc);
        if ((r = growFor(c, s, need, kMaxFrames)) != PIC_OK) return r;
    }
    pic::SpeciesDev& d = m.dev;
    m.shiftReady = 0;
    if (counts) PIC_HIP(hipMemcpyAsync(d.count, counts, c->nsc * 4, hipMemcpyHostToDevice, c->stream));
    for (int a = 0; a < 3; ++a) {
        if (pos && (r = viewToLayout(c, m, pitch, pos + a * slots, d.pos[a])) != PIC_OK) return r;
        if (mom && (r = viewToLayout(c, m, pitch, mom + a * slots, d.mom[a])) != PIC_OK) return r;
    }
    if (w) {
        if ((r = viewToLayout(c, m, pitch, w, d.w)) != PIC_OK) return r;
        double wmax = 0.0;
        for (size_t i = 0; i < slots; ++i) wmax = std::max(wmax, std::fabs(static_cast<double>(w[i])));
        m.wmax = wmax;
        updateJexp(c);
    }
    if (lc && (r = viewToLayout(c, m, pitch, lc, d.cell)) != PIC_OK) return r;
    if (mask && (r = viewToLayout(c, m, pitch, mask, d.mask)) != PIC_OK) return r;
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_read_frames(pic_ctx* c, int32_t s, uint32_t* counts, float* pos, float* mom, float* w, uint16_t* lc,
                    uint8_t* mask) {
    int r = checkSpecies(c, s);
    if (r != PIC_OK) return r;
    SpeciesMem& m = c->sp[s];
    pic::SpeciesDev& d = m.dev;
    int const pitch = m.cap;
    size_t const slots = static_cast<size_t>(c->nsc) * pitch;
    if (counts) PIC_HIP(hipMemcpyAsync(counts, d.count, c->nsc * 4, hipMemcpyDeviceToHost, c->stream));
    for (int a = 0; a < 3; ++a) {
        if (pos && (r = layoutToView(c, m, pitch, d.pos[a], pos + a * slots)) != PIC_OK) return r;
        if (mom && (r = layoutToView(c, m, pitch, d.mom[a], mom + a * slots)) != PIC_OK) return r;
    }
    if (w && (r = layoutToView(c, m, pitch, d.w, w)) != PIC_OK) return r;
    if (lc && (r = layoutToView(c, m, pitch, d.cell, lc)) != PIC_OK) return r;
    if (mask && (r = layoutToView(c, m, pitch, d.mask, mask)) != PIC_OK) return r;
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int64_t pic_particle_count(pic_ctx* c, int32_t s) {
    if (checkSpecies(c, s) != PIC_OK) return -1;
    unsigned long long total = 0;
    if (pic::launchCountSum(c->sp[s].dev.count, c->nsc, c->countOut, c->stream) != hipSuccess) return -1;
    if (hipMemcpyAsync(&total, c->countOut, sizeof(total), hipMemcpyDeviceToHost, c->stream) != hipSuccess) return -1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    return static_cast<int64_t>(total);
}

int pic_read_particles(pic_ctx* c, int32_t s, int64_t capacity, int64_t* n_out, int32_t* cell, float* pos,
                       float* mom, float* w) {
    int r = checkSpecies(c, s);
    if (r != PIC_OK) return r;
    SpeciesMem& m = c->sp[s];
    size_t const slots = static_cast<size_t>(c->nsc) * m.cap;
    std::vector<uint32_t> counts(c->nsc);
    std::vector<float> hp(3 * slots), hm(3 * slots), hw(slots);
    std::vector<uint16_t> hc(slots);
    r = pic_read_frames(c, s, counts.data(), hp.data(), hm.data(), hw.data(), hc.data(), nullptr);
    if (r != PIC_OK) return r;
    int64_t total = 0;
    for (uint32_t v : counts) total += v;
    if (n_out) *n_out = total;
    if (total > capacity && (cell || pos || mom || w)) return fail(PIC_E_INVALID, "output capacity too small");
    int64_t o = 0;
    for (int sc = 0; sc < c->nsc; ++sc) {
        int const sx = sc % c->g.nscx, sy = (sc / c->g.nscx) % c->g.nscy, sz = sc / (c->g.nscx * c->g.nscy);
        for (uint32_t i = 0; i < counts[sc]; ++i, ++o) {
            size_t const g = static_cast<size_t>(sc) * m.cap + i;
            int const lc = hc[g];
            if (cell) {
                cell[3 * o] = sx * pic::kSCX + lc % pic::kSCX;
                cell[3 * o + 1] = sy * pic::kSCY + (lc / pic::kSCX) % pic::kSCY;
                cell[3 * o + 2] = sz * pic::kSCZ + lc / (pic::kSCX * pic::kSCY);
            }
            for (int a = 0; a < 3; ++a) {
                if (pos) pos[3 * o + a] = hp[a * slots + g];
                if (mom) mom[3 * o + a] = hm[a * slots + g];
            }
            if (w) w[o] = hw[g];
        }
    }
    return PIC_OK;
}

int pic_init_species(pic_ctx* c, int32_t s, const pic_init_desc* desc) {
    int r = checkSpecies(c, s);
    if (r != PIC_OK) return r;
    if (!desc || desc->ppc <= 0) return fail(PIC_E_INVALID, "invalid init description");
    if (desc->ppc > pic::kMaxInitPpc) return fail(PIC_E_INVALID, "at most 255 particles per cell");
    SpeciesMem& m = c->sp[s];
    {  // every supercell must hold ppc particles in each of its cells
        std::vector<uint32_t> need(c->nsc, static_cast<uint32_t>(desc->ppc) * pic::kSCV);
        if ((r = growFor(c, s, need, kMaxFrames)) != PIC_OK) return r;
    }
    double const vol = c->cfg.cell_si[0] * c->cfg.cell_si[1] * c->cfg.cell_si[2];
    pic::InitArgs a{};
    a.g = c->g;
    a.sp = m.dev;
    a.seed = desc->seed;
    a.species = s;
    a.ppc = desc->ppc;
    if (desc->drift_band < 0 || desc->drift_band > 2) return fail(PIC_E_INVALID, "drift_band must be 0, 1 or 2");
    if (desc->profile < 0 || desc->profile > 1) return fail(PIC_E_INVALID, "unknown density profile");
    if (desc->profile == 1 && (desc->profile_axis < 0 || desc->profile_axis > 2 || !(desc->ramp_length_si > 0.0)))
        return fail(PIC_E_INVALID, "a ramp needs an axis in 0..2 and a positive length");
    a.drift_band = desc->drift_band;
    a.weighting = desc->density_si * vol / static_cast<double>(desc->ppc);
    a.frameMass = m.mass;
    for (int d = 0; d < 3; ++d) {
        a.u_drift[d] = desc->u_drift[d];
        a.u_th[d] = desc->u_th[d];
    }
    bool const quiet = desc->quiet[0] != 0 || desc->quiet[1] != 0 || desc->quiet[2] != 0;
    if (quiet) {
        if (desc->quiet[0] <= 0 || desc->quiet[1] <= 0 || desc->quiet[2] <= 0 ||
            desc->quiet[0] * desc->quiet[1] * desc->quiet[2] != desc->ppc)
            return fail(PIC_E_INVALID, "quiet start: positive particles per dimension whose product is ppc");
        for (int d = 0; d < 3; ++d) a.quiet[d] = desc->quiet[d];
    }
    a.profile = desc->profile;
    a.axis = desc->profile == 1 ? desc->profile_axis : 0;
    if (desc->profile == 1) {
        a.rampStart = desc->ramp_start_si / c->cfg.cell_si[a.axis];
        a.rampLength = desc->ramp_length_si / c->cfg.cell_si[a.axis];
    }
    m.shiftReady = 0;
    PIC_HIP(pic::launchInit(a, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    // the largest weighting: density*V/ppc, or, where WeightMacroParticles
    // takes fewer particles than ppc, N / count (below 2 MIN_WEIGHTING on a
    // ramp's thin cells)
    m.wmax = std::fabs(static_cast<double>(static_cast<float>(a.weighting)));
    {
        double const full = a.weighting * static_cast<double>(a.ppc);
        double const nFull = std::floor(std::min(static_cast<double>(a.ppc), full / 10.0));
        if (nFull >= 1.0 && nFull < a.ppc) m.wmax = std::max(m.wmax, full / nFull);
        if (a.profile == 1) m.wmax = std::max(m.wmax, 20.0);
    }
    updateJexp(c);
    return PIC_OK;
}

}  // extern "C"

namespace {

// the device planes of a field for pic_load_field / pic_read_field: E, B, J
// (3) or the PML's psi (6, in NodeValues order xy, xz, yx, yz, zx, zy; the
// kernels keep them as yx, zx, xy, zy, xz, yz)
int fieldPlanes(pic_ctx* c, int32_t field, float* planes[6], int* n) {
    if (field == PIC_FIELD_E || field == PIC_FIELD_B || field == PIC_FIELD_J) {
        float** f = field == PIC_FIELD_E ? c->E : (field == PIC_FIELD_B ? c->B : c->J);
        for (int a = 0; a < 3; ++a) planes[a] = f[a];
        *n = 3;
        return PIC_OK;
    }
    if (field == PIC_FIELD_PML_E || field == PIC_FIELD_PML_B) {
        if (!c->hasPml) return fail(PIC_E_INVALID, "the PML fields exist only with absorber PIC_ABSORBER_PML");
        float** psi = field == PIC_FIELD_PML_E ? c->pml.psiE : c->pml.psiB;
        constexpr int kNodeOrder[6] = {2, 4, 0, 5, 1, 3};
        for (int q = 0; q < 6; ++q) planes[q] = psi[kNodeOrder[q]];
        *n = 6;
        return PIC_OK;
    }
    return fail(PIC_E_INVALID, "unknown field");
}

}  // namespace

extern "C" {

int pic_load_field(pic_ctx* c, int32_t field, const float* data) {
    if (!c || !data) return fail(PIC_E_INVALID, "null argument");
    float* f[6];
    int n = 0;
    int const r = fieldPlanes(c, field, f, &n);
    if (r != PIC_OK) return r;
    for (int a = 0; a < n; ++a)
        PIC_HIP(hipMemcpyAsync(f[a], data + a * c->ncells, c->ncells * 4, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_read_field(pic_ctx* c, int32_t field, float* data) {
    if (!c || !data) return fail(PIC_E_INVALID, "null argument");
    float* f[6];
    int n = 0;
    int const r = fieldPlanes(c, field, f, &n);
    if (r != PIC_OK) return r;
    for (int a = 0; a < n; ++a)
        PIC_HIP(hipMemcpyAsync(data + a * c->ncells, f[a], c->ncells * 4, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_add_field_noise(pic_ctx* c, int32_t field, uint64_t seed, float amp) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    float** f = field == PIC_FIELD_E ? c->E : (field == PIC_FIELD_B ? c->B : (field == PIC_FIELD_J ? c->J : nullptr));
    if (!f) return fail(PIC_E_INVALID, "unknown field");
    for (int a = 0; a < 3; ++a) PIC_HIP(pic::launchNoise(f[a], c->ncells, seed, a, amp, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_add_laser(pic_ctx* c, const pic_laser_desc* l) {
    if (!c || !l) return fail(PIC_E_INVALID, "null argument");
    if (l->propagation_axis < 0 || l->propagation_axis > 2 || l->polarisation_axis < 0 || l->polarisation_axis > 2 ||
        l->propagation_axis == l->polarisation_axis)
        return fail(PIC_E_INVALID, "propagation and polarisation must be two different axes");
    if (!(l->wavelength_si > 0.0) || !(l->waist_si > 0.0) || !(l->fwhm_si > 0.0))
        return fail(PIC_E_INVALID, "wavelength, waist and duration must be positive");
    double const cSi = 2.99792458e8, meSi = 9.1093837139e-31, qeSi = 1.602176634e-19;
    double const L = c->k.unit_length;
    double const unitE = c->k.unit_mass * L / (c->k.unit_time * c->k.unit_time * c->k.unit_charge);
    double const omega = 2.0 * 3.14159265358979323846 * cSi / l->wavelength_si;
    pic::LaserArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) {
        a.E[d] = c->E[d];
        a.B[d] = c->B[d];
        a.cell[d] = static_cast<double>(c->k.cell[d]);
        a.center[d] = l->center_si[d] / L;
    }
    a.e0 = l->a0 * meSi * cSi * omega / qeSi / unitE;
    a.k = 2.0 * 3.14159265358979323846 / (l->wavelength_si / L);
    a.w0 = l->waist_si / L;
    a.sigma = cSi * l->fwhm_si / (2.0 * std::sqrt(std::log(2.0))) / L;
    a.invc = 1.0 / static_cast<double>(c->k.c);
    a.prop = l->propagation_axis;
    a.pol = l->polarisation_axis;
    PIC_HIP(pic::launchLaser(a, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // extern "C"

namespace {

int runShift(pic_ctx* c, int32_t species) {
    for (int s = 0; s < static_cast<int>(c->sp.size()); ++s) {
        if (species >= 0 && s != species) continue;
        if (c->sp[s].desc.pusher == PIC_PUSHER_NONE) continue;
        TimerScope ts(c, PIC_TIMER_SHIFT);
        int const ready = c->sp[s].shiftReady;
        if (ready == 0) PIC_HIP(pic::launchCollect(c->g, c->sp[s].dev, c->stream));
        PIC_HIP(pic::launchShift(c->g, c->sp[s].dev, ready == 2, c->stream));
        c->sp[s].shiftReady = 0;
    }
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_current_reset(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    TimerScope ts(c, PIC_TIMER_FIELD);
    PIC_HIP(pic::launchCurrentReset(c->g, c->J, c->ncells, c->stream));
    return PIC_OK;
}

int pic_push(pic_ctx* c, int32_t species, int32_t fuse) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (species >= static_cast<int>(c->sp.size())) return fail(PIC_E_INVALID, "species index out of range");
    if (species >= 0 && c->sp[species].desc.pusher == PIC_PUSHER_NONE)
        return fail(PIC_E_INVALID, "species is immobile");
    return runPush(c, species, fuse ? 1 : 0);
}

int pic_deposit(pic_ctx* c, int32_t species) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (species >= static_cast<int>(c->sp.size())) return fail(PIC_E_INVALID, "species index out of range");
    return runPush(c, species, 2);
}

int pic_shift(pic_ctx* c, int32_t species) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (species >= static_cast<int>(c->sp.size())) return fail(PIC_E_INVALID, "species index out of range");
    if (c->broken) return fail(PIC_E_STATE, "particles were lost to a failed capacity growth");
    int r = runShift(c, species);
    if (r != PIC_OK) return r;
    // a supercell that overflowed grows now, so the next stage sees every particle
    int skipped = 0;
    return recoverCapacity(c, &skipped, kMaxFrames);
}

int pic_fdtd_before_current(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    pic::FieldArgs a = makeFieldArgs(c);
    TimerScope ts(c, PIC_TIMER_FIELD);
    // FDTDBase.hpp:97-121: B half step, incident B at currentStep, incident E
    // at currentStep + 0.5 * timeStep / dt, E full step
    int r;
    if ((r = bHalfStage(c, a, false)) != PIC_OK) return r;
    if ((r = applyIncident(c, c->incB, c->B, 0.0f)) != PIC_OK) return r;
    if ((r = applyIncident(c, c->incE, c->E, 0.5f * c->k.dt / c->k.dt)) != PIC_OK) return r;
    if (c->hasPml)
        PIC_HIP(pic::launchEFullPml(withPmlInner(c, a), c->pml, c->stream));
    else
        PIC_HIP(pic::launchEFull(a, c->stream));
    return PIC_OK;
}

int pic_fold_current(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    return PIC_OK;
}

int pic_add_current(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    pic::FieldArgs a = makeFieldArgs(c);
    TimerScope ts(c, PIC_TIMER_FIELD);
    PIC_HIP(pic::launchAddCurrent(a, c->cfg.current_interpolation == PIC_CURRENT_BINOMIAL, c->stream));
    return PIC_OK;
}

int pic_fdtd_after_current(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    pic::FieldArgs a = makeFieldArgs(c);
    TimerScope ts(c, PIC_TIMER_FIELD);
    // FDTDBase.hpp:151-183: absorber on E, incident B at currentStep +
    // timeStep / dt, B half step, absorber on B
    if (c->absorb.n > 0) PIC_HIP(pic::launchAbsorb(c->g, c->E, c->absorb, c->stream));
    int r = applyIncident(c, c->incB, c->B, c->k.dt / c->k.dt);
    if (r != PIC_OK) return r;
    if ((r = bHalfStage(c, a, true)) != PIC_OK) return r;
    if (c->absorb.n > 0) PIC_HIP(pic::launchAbsorb(c->g, c->B, c->absorb, c->stream));
    PIC_HIP(pic::launchAdvanceStep(c->stepDev, c->flags, c->stream));
    return PIC_OK;
}

int pic_set_step(pic_ctx* c, int64_t step) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (step < 0) return fail(PIC_E_INVALID, "the step counter must be >= 0");
    PIC_HIP(hipMemcpyAsync(c->stepDev, &step, sizeof(step), hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_get_step(pic_ctx* c, int64_t* step) {
    if (!c || !step) return fail(PIC_E_INVALID, "null argument");
    PIC_HIP(hipMemcpyAsync(step, c->stepDev, sizeof(*step), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // extern "C"

namespace {

// update_beforeCurrent + addCurrent + update_afterCurrent as the two fused
// kernels; the new E lands in Ealt and the buffers swap

int fieldStep(pic_ctx* c) {
    if (pmlFieldStep(c)) {
        // update_beforeCurrent, AddCurrentDensity, update_afterCurrent with
        // the PML: the E full step and J in one pass into Ealt
        pic::FieldArgs a = pmlFieldArgs(c);
        TimerScope ts(c, PIC_TIMER_FIELD);
        if (c->cfg.current_interpolation == PIC_CURRENT_BINOMIAL) {
            PIC_HIP(pic::launchBinomialJ(a, c->Jf, c->stream));
            for (int d = 0; d < 3; ++d) a.J[d] = c->Jf[d];
        }
        PIC_HIP(pic::launchBHalfPml(c->cfg.solver, a, c->pml, false, c->stream));
        PIC_HIP(pic::launchEFullPmlJ(a, c->pml, c->Ealt, c->stream));
        for (int d = 0; d < 3; ++d) a.E[d] = c->Ealt[d];
        PIC_HIP(pic::launchBHalfPml(c->cfg.solver, a, c->pml, true, c->stream));
        PIC_HIP(pic::launchAdvanceStep(c->stepDev, c->flags, c->stream));
        for (int d = 0; d < 3; ++d) std::swap(c->E[d], c->Ealt[d]);
        c->eParity ^= 1;
        return PIC_OK;
    }
    if (!fusedFieldStep(c)) {
        // the incident field adds between the half steps the fused kernels
        // join: the stage kernels in place, E full step and J in one pass
        pic::FieldArgs a = makeFieldArgs(c);
        TimerScope ts(c, PIC_TIMER_FIELD);
        int r;
        if ((r = bHalfStage(c, a, false)) != PIC_OK) return r;
        if ((r = applyIncident(c, c->incB, c->B, 0.0f)) != PIC_OK) return r;
        if ((r = applyIncident(c, c->incE, c->E, 0.5f * c->k.dt / c->k.dt)) != PIC_OK) return r;
        pic::FieldArgs aj = a;
        if (c->cfg.current_interpolation == PIC_CURRENT_BINOMIAL) {
            PIC_HIP(pic::launchBinomialJ(a, c->Jf, c->stream));
            for (int d = 0; d < 3; ++d) aj.J[d] = c->Jf[d];
        }
        // E full step and J (AddCurrentDensity) in one pass, in place: a cell
        // reads only its own E and J besides B
        if (c->hasPml)
            PIC_HIP(pic::launchEFullPmlJ(withPmlInner(c, aj), c->pml, c->E, c->stream));
        else
            PIC_HIP(pic::launchEFullJ(aj, c->E, c->stream));
        if (c->absorb.n > 0) PIC_HIP(pic::launchAbsorb(c->g, c->E, c->absorb, c->stream));
        if ((r = applyIncident(c, c->incB, c->B, c->k.dt / c->k.dt)) != PIC_OK) return r;
        if ((r = bHalfStage(c, a, true)) != PIC_OK) return r;
        if (c->absorb.n > 0) PIC_HIP(pic::launchAbsorb(c->g, c->B, c->absorb, c->stream));
        PIC_HIP(pic::launchAdvanceStep(c->stepDev, c->flags, c->stream));
        return PIC_OK;
    }
    pic::FieldArgs a = makeFieldArgs(c);
    {
        TimerScope ts(c, PIC_TIMER_FIELD);
        if (c->cfg.current_interpolation == PIC_CURRENT_BINOMIAL) {
            PIC_HIP(pic::launchBinomialJ(a, c->Jf, c->stream));
            for (int d = 0; d < 3; ++d) a.J[d] = c->Jf[d];
        }
        PIC_HIP(pic::launchFieldStep(c->cfg.solver, a, c->Ealt, &c->absorb, c->stream));
        PIC_HIP(pic::launchAdvanceStep(c->stepDev, c->flags, c->stream));
    }
    for (int d = 0; d < 3; ++d) std::swap(c->E[d], c->Ealt[d]);
    c->eParity ^= 1;
    return PIC_OK;
}

// one full step in the reference stage order, push and deposit fused, the
// field stages fused
int oneStep(pic_ctx* c) {
    int r;
    if ((r = pic_current_reset(c)) != PIC_OK) return r;
    if ((r = runPush(c, -1, 1, true)) != PIC_OK) return r;
    if ((r = runShift(c, -1)) != PIC_OK) return r;
    return fieldStep(c);
}

int stepGraphs(pic_ctx* c, int32_t nsteps) {
    for (int p = 0; p < 2; ++p) {
        if (c->graphExec[p] && c->graphJexp[p] != c->jexp) {
            (void)hipGraphExecDestroy(c->graphExec[p]);
            (void)hipGraphDestroy(c->graph[p]);
            c->graphExec[p] = nullptr;
            c->graph[p] = nullptr;
        }
    }
    for (int32_t i = 0; i < nsteps; ++i) {
        int const p = c->eParity;
        if (!c->graphExec[p]) {
            // every launch plan must exist before capture (building one copies and synchronises)
            for (const SpeciesMem& m : c->sp) {
                if (m.desc.pusher == PIC_PUSHER_NONE) continue;
                PhasePlan* plan = nullptr;
                int const r = getPlan(c, supportOf(m.desc.shape), true, &plan);
                if (r != PIC_OK) return r;
            }
            PIC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int const r = oneStep(c);  // records the step and swaps E/Ealt as its replay will
            hipGraph_t g = nullptr;
            hipError_t const e = hipStreamEndCapture(c->stream, &g);
            if (r != PIC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            if (e != hipSuccess) return fail(PIC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            c->graph[p] = g;
            PIC_HIP(hipGraphInstantiate(&c->graphExec[p], c->graph[p], nullptr, nullptr, 0));
            c->graphJexp[p] = c->jexp;
        } else if (fieldStepSwapsE(c)) {
            // the replayed step writes the new E into the other buffer
            for (int d = 0; d < 3; ++d) std::swap(c->E[d], c->Ealt[d]);
            c->eParity ^= 1;
        }
        PIC_HIP(hipGraphLaunch(c->graphExec[p], c->stream));
    }
    for (SpeciesMem& m : c->sp) m.shiftReady = 0;
    return PIC_OK;
}


}  // namespace

extern "C" int pic_fdtd_step(pic_ctx* c) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    return fieldStep(c);
}

extern "C" {

int pic_step(pic_ctx* c, int32_t nsteps) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (nsteps < 0) return fail(PIC_E_INVALID, "nsteps must be >= 0");
    if (c->broken) return fail(PIC_E_STATE, "particles were lost to a failed capacity growth");
    // steps after a capacity overflow are skipped on the device; they run
    // again once the supercells have grown
    int todo = nsteps;
    do {
        for (int32_t i = 0; i < todo; ++i) {
            int const r = oneStep(c);
            if (r != PIC_OK) return r;
        }
        int const r = recoverCapacity(c, &todo, kMaxFrames);
        if (r != PIC_OK) return r;
    } while (todo > 0);
    return PIC_OK;
}

int pic_step_graph(pic_ctx* c, int32_t nsteps) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (nsteps < 0) return fail(PIC_E_INVALID, "nsteps must be >= 0");
    if (c->timing) return fail(PIC_E_STATE, "pic_step_graph cannot run with kernel timing enabled");
    if (c->broken) return fail(PIC_E_STATE, "particles were lost to a failed capacity growth");
    int todo = nsteps;
    do {
        int const r = stepGraphs(c, todo);
        if (r != PIC_OK) return r;
        // a repacking growth drops the graphs: recaptured on the next replay
        int const rr = recoverCapacity(c, &todo, kMaxFrames);
        if (rr != PIC_OK) return rr;
    } while (todo > 0);
    return PIC_OK;
}

int pic_field_energy(pic_ctx* c, double out[7]) {
    if (!c || !out) return fail(PIC_E_INVALID, "null argument");
    pic::FieldArgs a = makeFieldArgs(c);
    PIC_HIP(pic::launchEnergy(a, c->energyPartials, c->energyBlocks, c->energyPartials + 6 * c->energyBlocks,
                              c->stream));
    double sums[6];
    PIC_HIP(hipMemcpyAsync(sums, c->energyPartials + 6 * c->energyBlocks, sizeof(sums), hipMemcpyDeviceToHost,
                           c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    fieldEnergyFromSums(c, sums, out);
    return PIC_OK;
}

int pic_gauss_residual(pic_ctx* c, double* out) {
    if (!c || !out) return fail(PIC_E_INVALID, "null argument");
    if (!c->diagBuf && devAlloc(c, &c->diagBuf, c->ncells) != PIC_OK) return PIC_E_HIP;
    std::vector<pic::SpeciesDev> sp;
    std::vector<int> supp;
    for (const SpeciesMem& m : c->sp) {
        sp.push_back(m.dev);
        supp.push_back(supportOf(m.desc.shape));
    }
    pic::FieldArgs a = makeFieldArgs(c);
    double const qScale = static_cast<double>(c->k.inv_cell_volume_dt) * static_cast<double>(c->k.dt);
    PIC_HIP(pic::launchGaussResidual(a, sp.data(), supp.data(), static_cast<int>(sp.size()), qScale,
                                     static_cast<double>(c->k.eps0), c->diagBuf, c->stream));
    PIC_HIP(hipMemcpyAsync(out, c->diagBuf, c->ncells * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_gauss_residual_noalloc(pic_ctx* c, double* out) {
    if (!c || !out) return fail(PIC_E_INVALID, "null argument");
    if (c->ncells > pic::kNoAllocMaxCells)
        return fail(PIC_E_INVALID, "pic_gauss_residual_noalloc: the grid has " + std::to_string(c->ncells) +
                                       " cells, more than its scratch holds (" +
                                       std::to_string(pic::kNoAllocMaxCells) + ")");
    PIC_HIP(hipSetDevice(c->device));
    double* buf = nullptr;
    PIC_HIP(pic::noAllocScratch(c->device, &buf));
    std::array<pic::SpeciesDev, PIC_MAX_SPECIES> sp;
    std::array<int, PIC_MAX_SPECIES> supp;
    int nsp = 0;
    for (const SpeciesMem& m : c->sp) {
        sp[nsp] = m.dev;
        supp[nsp++] = supportOf(m.desc.shape);
    }
    pic::FieldArgs a = makeFieldArgs(c);
    double const qScale = static_cast<double>(c->k.inv_cell_volume_dt) * static_cast<double>(c->k.dt);
    // the kernels clear the scratch first (it keeps the last call's values)
    PIC_HIP(pic::launchGaussResidual(a, sp.data(), supp.data(), nsp, qScale, static_cast<double>(c->k.eps0), buf,
                                     c->stream));
    PIC_HIP(hipMemcpyAsync(out, buf, c->ncells * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_error_flags(pic_ctx* c, uint32_t* flags) {
    if (!c || !flags) return fail(PIC_E_INVALID, "null argument");
    PIC_HIP(hipStreamSynchronize(c->stream));
    *flags = c->hostFlags;
    c->hostFlags = 0;
    return PIC_OK;
}

int pic_timing_enable(pic_ctx* c, int32_t enable) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    int const r = resolveTimers(c);
    if (r != PIC_OK) return r;
    c->timing = enable != 0;
    for (int t = 0; t < PIC_TIMER_COUNT; ++t) {
        c->totalMs[t] = 0.0;
        c->launches[t] = 0;
    }
    return PIC_OK;
}

int pic_kernel_timer(pic_ctx* c, int32_t timer, double* total_ms, int64_t* launches) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    if (timer < 0 || timer >= PIC_TIMER_COUNT) return fail(PIC_E_INVALID, "unknown timer");
    int const r = resolveTimers(c);
    if (r != PIC_OK) return r;
    if (total_ms) *total_ms = c->totalMs[timer];
    if (launches) *launches = c->launches[timer];
    return PIC_OK;
}


// ---------------------------------------------------------------------------
// differentiable form and adjoint
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

int allocParticles(pic_ctx* c, DiffState& st, int64_t n) {
    if (devAlloc(c, &st.cell, 3 * (size_t)n) != PIC_OK) return PIC_E_HIP;
    if (devAlloc(c, &st.pos, 3 * (size_t)n) != PIC_OK) return PIC_E_HIP;
    if (devAlloc(c, &st.mom, 3 * (size_t)n) != PIC_OK) return PIC_E_HIP;
    return PIC_OK;
}

int allocState(pic_ctx* c, DiffState& st, int64_t n) {
    if (allocParticles(c, st, n) != PIC_OK) return PIC_E_HIP;
    if (devAlloc(c, &st.E, 3 * c->ncells) != PIC_OK) return PIC_E_HIP;
    if (devAlloc(c, &st.B, 3 * c->ncells) != PIC_OK) return PIC_E_HIP;
    return PIC_OK;
}

int copyParticles(pic_ctx* c, DiffState& dst, const DiffState& src, int64_t n) {
    size_t const b = 3 * (size_t)n * 4;
    PIC_HIP(hipMemcpyAsync(dst.cell, src.cell, b, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(dst.pos, src.pos, b, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(dst.mom, src.mom, b, hipMemcpyDeviceToDevice, c->stream));
    return PIC_OK;
}

int copyFields(pic_ctx* c, float* dst, const float* src) {
    PIC_HIP(hipMemcpyAsync(dst, src, 3 * c->ncells * 4, hipMemcpyDeviceToDevice, c->stream));
    return PIC_OK;
}

// ctx E/B are three separate planes; states keep them in one buffer
int gatherPlanes(pic_ctx* c, float* dst, float* const planes[3]) {
    for (int a = 0; a < 3; ++a)
        PIC_HIP(hipMemcpyAsync(dst + a * c->ncells, planes[a], c->ncells * 4, hipMemcpyDeviceToDevice, c->stream));
    return PIC_OK;
}

int scatterPlanes(pic_ctx* c, float* const planes[3], const float* src) {
    for (int a = 0; a < 3; ++a)
        PIC_HIP(hipMemcpyAsync(planes[a], src + a * c->ncells, c->ncells * 4, hipMemcpyDeviceToDevice, c->stream));
    return PIC_OK;
}

// one step of the differentiable form on particles P and fields E[3], B[3]
int diffStep(pic_ctx* c, const pic::DiffParticles& P, float* const E[3], float* const B[3]) {
    AdjointMem& m = c->adj;
    pic::DiffArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) {
        a.E[d] = E[d];
        a.B[d] = B[d];
        a.Jfix[d] = m.Jfix + d * c->ncells;
        a.cell[d] = c->k.cell[d];
    }
    a.dt = c->k.dt;
    a.c = c->k.c;
    a.c2 = c->k.c * c->k.c;
    a.rc2 = c->k.mue0_eps0;
    a.inv_vdt = c->k.inv_cell_volume_dt;
    a.jscale = std::ldexp(1.0, m.jexp);
    a.mass = m.mass;
    a.charge = m.charge;
    PIC_HIP(pic::launchDiffPush(a, P, c->stream));
    double const jinv = std::ldexp(1.0, -m.jexp);
    PIC_HIP(pic::launchFixToFloat(m.Jfix, c->J, c->ncells, jinv, c->stream));
    pic::FieldArgs f = makeFieldArgs(c);
    for (int d = 0; d < 3; ++d) {
        f.E[d] = E[d];
        f.B[d] = B[d];
    }
    PIC_HIP(pic::launchBHalf(PIC_SOLVER_YEE, f, c->stream));
    PIC_HIP(pic::launchEFull(f, c->stream));
    PIC_HIP(pic::launchAddCurrent(f, false, c->stream));
    PIC_HIP(pic::launchBHalf(PIC_SOLVER_YEE, f, c->stream));
    return PIC_OK;
}

// particles of `from` stepped into `to` (the same state for an in-place step)
pic::DiffParticles particlesOf(const AdjointMem& m, const DiffState& from, const DiffState& to, const float* w) {
    pic::DiffParticles P{};
    P.cell = from.cell;
    P.pos = from.pos;
    P.mom = from.mom;
    P.w = w;
    P.cellOut = to.cell;
    P.posOut = to.pos;
    P.momOut = to.mom;
    P.n = m.n;
    return P;
}

// state `to` = one step of state `from`: particles out of place, fields
// copied and then updated in place
int stepState(pic_ctx* c, const DiffState& from, const DiffState& to, const float* w) {
    int r;
    if ((r = copyFields(c, to.E, from.E)) != PIC_OK) return r;
    if ((r = copyFields(c, to.B, from.B)) != PIC_OK) return r;
    float* const E[3] = {to.E, to.E + c->ncells, to.E + 2 * c->ncells};
    float* const B[3] = {to.B, to.B + c->ncells, to.B + 2 * c->ncells};
    return diffStep(c, particlesOf(c->adj, from, to, w), E, B);
}

// re-tiling at a checkpoint: sort the running particles by tile into `cp`
// (particles, weightings, the permutation), continue from the sorted order
// Particles are sorted by tile again every kRetileEvery steps (and at every
// checkpoint) so the particles of a thread block stay in a compact box of
// cells while they drift.
constexpr int kRetileEvery = 4;

// sorts the particles of `st` (weightings w, identities ids if not null) by
// tile, in place through the scratch state; the permutation (new slot -> old
// slot) is copied to permOut if not null
int retileInPlace(pic_ctx* c, DiffState& st, float* w, int32_t* ids, int32_t* permOut) {
    AdjointMem& m = c->adj;
    int32_t* perm = nullptr;
    PIC_HIP(pic::launchTileSort(st.cell, m.n, c->g, m.keys, m.vals, m.sortTemp, m.sortTempBytes, &perm, c->stream));
    pic::DiffParticles src = particlesOf(m, st, st, w);
    pic::DiffParticles dst = particlesOf(m, m.scratch, m.scratch, m.scratch.w);
    const int32_t* idsIn = ids ? ids : m.iota;
    PIC_HIP(pic::launchGatherParticles(perm, src, dst, idsIn, m.idsTmp, m.scratch.w, c->stream));
    int r;
    if ((r = copyParticles(c, st, m.scratch, m.n)) != PIC_OK) return r;
    PIC_HIP(hipMemcpyAsync(w, m.scratch.w, (size_t)m.n * 4, hipMemcpyDeviceToDevice, c->stream));
    if (ids) PIC_HIP(hipMemcpyAsync(ids, m.idsTmp, (size_t)m.n * 4, hipMemcpyDeviceToDevice, c->stream));
    if (permOut) PIC_HIP(hipMemcpyAsync(permOut, perm, (size_t)m.n * 4, hipMemcpyDeviceToDevice, c->stream));
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_adjoint_load_particles(pic_ctx* c, int32_t s, int64_t n, const int32_t* cell, const float* pos,
                               const float* mom, const float* w) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    int r = checkSpecies(c, s);
    if (r != PIC_OK) return r;
    if (n <= 0 || !cell || !pos || !mom || !w) return fail(PIC_E_INVALID, "particle arrays and n > 0 are required");
    if (c->cfg.solver != PIC_SOLVER_YEE) return fail(PIC_E_INVALID, "the differentiable form uses the Yee solver");
    if (!allPeriodic(c->cfg) || c->cfg.absorber != PIC_ABSORBER_NONE || c->cfg.current_interpolation != PIC_CURRENT_NONE ||
        !c->incB.empty() || !c->incE.empty())
        return fail(PIC_E_INVALID,
                    "the differentiable form needs a periodic box without absorber, current filter or incident field");
    if (c->ncells > (size_t)INT32_MAX) return fail(PIC_E_INVALID, "the differentiable form needs fewer than 2^31 cells");
    const SpeciesMem& sm = c->sp[s];
    if (sm.desc.shape != PIC_SHAPE_CIC || sm.desc.pusher != PIC_PUSHER_BORIS ||
        sm.desc.current_solver != PIC_CURRENT_SOLVER_ESIRKEPOV)
        return fail(PIC_E_INVALID,
                    "the differentiable form supports the CIC shape with the Boris pusher and Esirkepov current");
    int const ng[3] = {c->g.nx, c->g.ny, c->g.nz};
    double wmax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            int32_t const ci = cell[a * n + i];
            float const p = pos[a * n + i];
            if (ci < 0 || ci >= ng[a] || !(p >= 0.0f && p < 1.0f))
                return fail(PIC_E_INVALID, "particle cell or in-cell position out of range");
        }
        if (!(w[i] > 0.0f)) return fail(PIC_E_INVALID, "weightings must be positive");
        wmax = std::max(wmax, (double)w[i]);
    }
    AdjointMem& m = c->adj;
    if (m.n != n) {
        if (m.n != 0) return fail(PIC_E_STATE, "the differentiable particle count is fixed once loaded");
        m.n = n;
        if (devAlloc(c, &m.w, (size_t)n) != PIC_OK) return PIC_E_HIP;
        if (allocParticles(c, m.init, n) != PIC_OK || allocParticles(c, m.work, n) != PIC_OK) return PIC_E_HIP;
        if (devAlloc(c, &m.Jfix, 3 * c->ncells) != PIC_OK) return PIC_E_HIP;
        if (devAlloc(c, &m.lE, 3 * c->ncells) != PIC_OK || devAlloc(c, &m.lB, 3 * c->ncells) != PIC_OK ||
            devAlloc(c, &m.lJ, 3 * c->ncells) != PIC_OK)
            return PIC_E_HIP;
        if (devAlloc(c, &m.lX, 3 * (size_t)n) != PIC_OK || devAlloc(c, &m.lP, 3 * (size_t)n) != PIC_OK ||
            devAlloc(c, &m.lXtmp, 3 * (size_t)n) != PIC_OK || devAlloc(c, &m.lPtmp, 3 * (size_t)n) != PIC_OK)
            return PIC_E_HIP;
        if (devAlloc(c, &m.wInit, (size_t)n) != PIC_OK || devAlloc(c, &m.ids, (size_t)n) != PIC_OK ||
            devAlloc(c, &m.idsTmp, (size_t)n) != PIC_OK || devAlloc(c, &m.iota, (size_t)n) != PIC_OK)
            return PIC_E_HIP;
        for (int k = 0; k < 2; ++k)
            if (devAlloc(c, &m.keys[k], (size_t)n) != PIC_OK || devAlloc(c, &m.vals[k], (size_t)n) != PIC_OK)
                return PIC_E_HIP;
        if (allocParticles(c, m.scratch, n) != PIC_OK || devAlloc(c, &m.scratch.w, (size_t)n) != PIC_OK)
            return PIC_E_HIP;
        PIC_HIP(pic::sortTempBytes(n, &m.sortTempBytes));
        unsigned char* t = nullptr;
        if (devAlloc(c, &t, m.sortTempBytes) != PIC_OK) return PIC_E_HIP;
        m.sortTemp = t;
        std::vector<int32_t> iota((size_t)n);
        for (int64_t i = 0; i < n; ++i) iota[(size_t)i] = (int32_t)i;
        PIC_HIP(hipMemcpyAsync(m.iota, iota.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipStreamSynchronize(c->stream));
    }
    m.species = s;
    m.mass = sm.mass;
    m.charge = sm.charge;
    if (c->cfg.j_scale_exp != 0) {
        m.jexp = c->cfg.j_scale_exp;
    } else {
        float const maxCell = std::max(c->k.cell[0], std::max(c->k.cell[1], c->k.cell[2]));
        double const bound = std::fabs((double)sm.charge) * wmax * (double)c->k.inv_cell_volume_dt * (double)maxCell;
        m.jexp = bound > 0.0 ? std::max(-60, std::min(100, 40 - (int)std::ceil(std::log2(bound)))) : 40;
    }
    size_t const b = 3 * (size_t)n * 4;
    PIC_HIP(hipMemcpyAsync(m.init.cell, cell, b, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.init.pos, pos, b, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.init.mom, mom, b, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.wInit, w, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.w, m.wInit, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.ids, m.iota, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    r = copyParticles(c, m.work, m.init, n);
    if (r != PIC_OK) return r;
    PIC_HIP(hipStreamSynchronize(c->stream));
    m.forwardValid = false;
    return PIC_OK;
}

int pic_adjoint_read_particles(pic_ctx* c, int32_t* cell, float* pos, float* mom) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    AdjointMem& m = c->adj;
    if (m.n == 0) return fail(PIC_E_STATE, "no differentiable particles loaded");
    // the running order is tiled; hand the particles back in the loaded order
    size_t const n = (size_t)m.n;
    std::vector<int32_t> ids(n), hc(3 * n);
    std::vector<float> hp(3 * n), hm(3 * n);
    PIC_HIP(hipMemcpyAsync(ids.data(), m.ids, n * 4, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipMemcpyAsync(hc.data(), m.work.cell, 3 * n * 4, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipMemcpyAsync(hp.data(), m.work.pos, 3 * n * 4, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipMemcpyAsync(hm.data(), m.work.mom, 3 * n * 4, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    for (size_t j = 0; j < n; ++j) {
        size_t const i = (size_t)ids[j];
        for (int d = 0; d < 3; ++d) {
            if (cell) cell[d * n + i] = hc[d * n + j];
            if (pos) pos[d * n + i] = hp[d * n + j];
            if (mom) mom[d * n + i] = hm[d * n + j];
        }
    }
    return PIC_OK;
}

int pic_adjoint_forward(pic_ctx* c, const float* E0, const float* B0, int32_t nsteps, int32_t every, float* E_out,
                        float* B_out) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    AdjointMem& m = c->adj;
    if (m.n == 0) return fail(PIC_E_STATE, "no differentiable particles loaded");
    if (nsteps < 0 || every < 1) return fail(PIC_E_INVALID, "nsteps >= 0 and checkpoint_every >= 1 are required");
    int r;
    if (E0 && (r = scatterPlanes(c, c->E, E0)) != PIC_OK) return r;
    if (B0 && (r = scatterPlanes(c, c->B, B0)) != PIC_OK) return r;
    if ((r = copyParticles(c, m.work, m.init, m.n)) != PIC_OK) return r;
    PIC_HIP(hipMemcpyAsync(m.w, m.wInit, (size_t)m.n * 4, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.ids, m.iota, (size_t)m.n * 4, hipMemcpyDeviceToDevice, c->stream));
    size_t const ncp = (size_t)(nsteps + every - 1) / (size_t)every;
    while (m.checkpoints.size() < ncp) {
        DiffState st;
        if (allocState(c, st, m.n) != PIC_OK) return PIC_E_HIP;
        if (devAlloc(c, &st.w, (size_t)m.n) != PIC_OK || devAlloc(c, &st.perm, (size_t)m.n) != PIC_OK)
            return PIC_E_HIP;
        m.checkpoints.push_back(st);
    }
    pic::DiffParticles const P = particlesOf(m, m.work, m.work, m.w);
    for (int32_t i = 0; i < nsteps; ++i) {
        if (i % every == 0) {
            DiffState& cp = m.checkpoints[(size_t)(i / every)];
            if ((r = retileInPlace(c, m.work, m.w, m.ids, cp.perm)) != PIC_OK) return r;
            if ((r = copyParticles(c, cp, m.work, m.n)) != PIC_OK) return r;
            PIC_HIP(hipMemcpyAsync(cp.w, m.w, (size_t)m.n * 4, hipMemcpyDeviceToDevice, c->stream));
            if ((r = gatherPlanes(c, cp.E, c->E)) != PIC_OK) return r;
            if ((r = gatherPlanes(c, cp.B, c->B)) != PIC_OK) return r;
        } else if (i % kRetileEvery == 0) {
            if ((r = retileInPlace(c, m.work, m.w, m.ids, nullptr)) != PIC_OK) return r;
        }
        if ((r = diffStep(c, P, c->E, c->B)) != PIC_OK) return r;
    }
    if (E_out && (r = gatherPlanes(c, E_out, c->E)) != PIC_OK) return r;
    if (B_out && (r = gatherPlanes(c, B_out, c->B)) != PIC_OK) return r;
    PIC_HIP(hipStreamSynchronize(c->stream));
    m.nsteps = nsteps;
    m.every = every;
    m.forwardValid = true;
    return PIC_OK;
}

int pic_adjoint_backward(pic_ctx* c, const double* dE, const double* dB, double* dE0, double* dB0, double* dX0,
                         double* dP0) {
    if (!c) return fail(PIC_E_INVALID, "null context");
    AdjointMem& m = c->adj;
    if (!m.forwardValid) return fail(PIC_E_STATE, "pic_adjoint_backward needs a preceding pic_adjoint_forward");
    if (!dE || !dB || !dE0 || !dB0) return fail(PIC_E_INVALID, "dE, dB, dE0 and dB0 are required");
    size_t const nf = 3 * c->ncells;
    size_t const np = 3 * (size_t)m.n;
    PIC_HIP(hipMemcpyAsync(m.lE, dE, nf * 8, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(m.lB, dB, nf * 8, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(pic::launchFillD(m.lX, np, 0.0, c->stream));
    PIC_HIP(pic::launchFillD(m.lP, np, 0.0, c->stream));
    int const K = m.every;
    while (m.tape.size() < (size_t)std::min(K, std::max(m.nsteps, 1))) {
        DiffState st;
        if (allocState(c, st, m.n) != PIC_OK) return PIC_E_HIP;
        if (devAlloc(c, &st.w, (size_t)m.n) != PIC_OK || devAlloc(c, &st.perm, (size_t)m.n) != PIC_OK)
            return PIC_E_HIP;
        m.tape.push_back(st);
    }
    m.tapeW.assign(m.tape.size(), nullptr);
    pic::AdjFieldArgs fa{};
    fa.g = c->g;
    for (int d = 0; d < 3; ++d) {
        fa.lE[d] = m.lE + d * c->ncells;
        fa.lB[d] = m.lB + d * c->ncells;
        fa.cell[d] = (double)c->k.cell[d];
    }
    double* lJ[3] = {m.lJ, m.lJ + c->ncells, m.lJ + 2 * c->ncells};
    pic::AdjPartArgs pa{};
    pa.g = c->g;
    for (int d = 0; d < 3; ++d) {
        pa.lJ[d] = lJ[d];
        pa.lE[d] = fa.lE[d];
        pa.lB[d] = fa.lB[d];
        pa.cell[d] = (double)c->k.cell[d];
    }
    pa.dt = (double)c->k.dt;
    pa.rc2 = (double)c->k.mue0_eps0;
    pa.mass = (double)m.mass;
    pa.charge = (double)m.charge;
    pa.inv_vdt = (double)c->k.inv_cell_volume_dt;
    double const dt = (double)c->k.dt;
    double const c2 = (double)(c->k.c * c->k.c);
    double const coeff = (double)c->k.add_current_coeff;
    int r;
    int const nseg = (m.nsteps + K - 1) / K;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        int const len = std::min(K, m.nsteps - seg * K);
        const DiffState& cp = m.checkpoints[(size_t)seg];
        if ((r = copyParticles(c, m.tape[0], cp, m.n)) != PIC_OK) return r;
        if ((r = copyFields(c, m.tape[0].E, cp.E)) != PIC_OK) return r;
        if ((r = copyFields(c, m.tape[0].B, cp.B)) != PIC_OK) return r;
        // replay the segment with the forward's re-tilings (deterministic), keeping their permutations
        m.tapeW[0] = cp.w;
        for (int t = 1; t < len; ++t) {
            DiffState& st = m.tape[(size_t)t];
            if ((r = stepState(c, m.tape[(size_t)t - 1], st, m.tapeW[(size_t)t - 1])) != PIC_OK) return r;
            m.tapeW[(size_t)t] = m.tapeW[(size_t)t - 1];
            if ((seg * K + t) % kRetileEvery == 0) {
                PIC_HIP(hipMemcpyAsync(st.w, m.tapeW[(size_t)t - 1], (size_t)m.n * 4, hipMemcpyDeviceToDevice,
                                       c->stream));
                if ((r = retileInPlace(c, st, st.w, nullptr, st.perm)) != PIC_OK) return r;
                m.tapeW[(size_t)t] = st.w;
            }
        }
        for (int t = len - 1; t >= 0; --t) {
            const DiffState& st = m.tape[(size_t)t];
            PIC_HIP(pic::launchAdjFields(fa, dt, c2, coeff, lJ, c->stream));
            for (int d = 0; d < 3; ++d) {
                pa.E[d] = st.E + d * c->ncells;
                pa.B[d] = st.B + d * c->ncells;
            }
            pic::AdjState as{};
            as.cell = st.cell;
            as.pos = st.pos;
            as.mom = st.mom;
            as.w = m.tapeW[(size_t)t];
            as.lX = m.lX;
            as.lP = m.lP;
            as.n = m.n;
            PIC_HIP(pic::launchAdjParticles(pa, as, c->stream));
            if (t > 0 && (seg * K + t) % kRetileEvery == 0) {  // undo the re-tiling of state t
                PIC_HIP(pic::launchScatterAdjoint(st.perm, m.lX, m.lXtmp, m.n, c->stream));
                PIC_HIP(pic::launchScatterAdjoint(st.perm, m.lP, m.lPtmp, m.n, c->stream));
                std::swap(m.lX, m.lXtmp);
                std::swap(m.lP, m.lPtmp);
            }
        }
        if (seg > 0) {  // undo this checkpoint's re-tiling: the previous segment's order
            PIC_HIP(pic::launchScatterAdjoint(cp.perm, m.lX, m.lXtmp, m.n, c->stream));
            PIC_HIP(pic::launchScatterAdjoint(cp.perm, m.lP, m.lPtmp, m.n, c->stream));
            std::swap(m.lX, m.lXtmp);
            std::swap(m.lP, m.lPtmp);
        }
    }
    PIC_HIP(hipMemcpyAsync(dE0, m.lE, nf * 8, hipMemcpyDeviceToDevice, c->stream));
    PIC_HIP(hipMemcpyAsync(dB0, m.lB, nf * 8, hipMemcpyDeviceToDevice, c->stream));
    if (nseg == 0) {  // no step: the loaded particles are the final ones
        if (dX0) PIC_HIP(hipMemcpyAsync(dX0, m.lX, np * 8, hipMemcpyDeviceToDevice, c->stream));
        if (dP0) PIC_HIP(hipMemcpyAsync(dP0, m.lP, np * 8, hipMemcpyDeviceToDevice, c->stream));
    } else {  // the first checkpoint's permutation maps to the loaded order
        const int32_t* perm0 = m.checkpoints[0].perm;
        if (dX0) PIC_HIP(pic::launchScatterAdjoint(perm0, m.lX, dX0, m.n, c->stream));
        if (dP0) PIC_HIP(pic::launchScatterAdjoint(perm0, m.lP, dP0, m.n, c->stream));
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched ensembles of independent simulations
// ---------------------------------------------------------------------------
struct pic_ensemble {
    std::vector<pic_ctx*> sims;
    int device = 0;
    hipStream_t stream = nullptr;
    int supp = 0, pusher = 0, solver = 0;
    int parity = 0;  // every member's E/Ealt parity
    // device tables; PushArgs and field entries per parity (E and Ealt swap)
    pic::PushArgs* dPush[2] = {nullptr, nullptr};
    pic::FieldBatchEntry* dField[2] = {nullptr, nullptr};
    pic::ShiftEntry* dShift = nullptr;
    int nShiftEntries = 0;
    // work lists
    std::vector<int> phaseOffset, phaseCount;
    int2* dPhaseWork = nullptr;
    int2* dShiftWork = nullptr;
    int nShiftWork = 0;
    int2* dTiles = nullptr;
    int nTiles = 0;
    int2* dBlocks = nullptr;
    int nBlocks = 0;
    // every member's fused push in one launch with its tiles gathered
    // (all members' plans are one-launch plans), or one launch per phase
    bool single = false;
    pic::GatherEntry* dGather = nullptr;
    // schedule
    int order = 1;         // PIC_ENSEMBLE_ORDER_*
    int blocksPerCu = 0;   // workgroups per CU of the batched push (0: one per work item)
    int cus = 256;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graphExec[2] = {nullptr, nullptr};
    std::vector<void*> allocations;
};

namespace {

template <typename T>
int ensAlloc(pic_ensemble* e, T** p, size_t n) {
    void* q = nullptr;
    hipError_t const r = hipMalloc(&q, std::max<size_t>(n * sizeof(T), 16));
    if (r != hipSuccess) return fail(PIC_E_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(r));
    e->allocations.push_back(q);
    *p = static_cast<T*>(q);
    return PIC_OK;
}

void ensDropGraphs(pic_ensemble* e) {
    for (int p = 0; p < 2; ++p) {
        if (e->graphExec[p]) (void)hipGraphExecDestroy(e->graphExec[p]);
        if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
        e->graphExec[p] = nullptr;
        e->graph[p] = nullptr;
    }
}

// per-simulation tables for both parities (contents change with the current
// scale exponent; the buffers stay, so captured graphs stay valid)
int ensUploadTables(pic_ensemble* e) {
    int const n = static_cast<int>(e->sims.size());
    for (int par = 0; par < 2; ++par) {
        std::vector<pic::PushArgs> push(n);
        std::vector<pic::FieldBatchEntry> field(n);
        for (int i = 0; i < n; ++i) {
            pic_ctx* c = e->sims[i];
            // members share one parity; parity `par` has E in E if par == parity, else in Ealt
            bool const swap = par != e->parity;
            float* E[3];
            float* Ealt[3];
            for (int d = 0; d < 3; ++d) {
                E[d] = swap ? c->Ealt[d] : c->E[d];
                Ealt[d] = swap ? c->E[d] : c->Ealt[d];
            }
            pic::PushArgs a = makePushArgs(c);
            for (int d = 0; d < 3; ++d) a.E[d] = E[d];
            a.stageShift = 1;
            int k = 0;
            for (const SpeciesMem& m : c->sp)
                if (m.desc.pusher != PIC_PUSHER_NONE) a.sp[k++] = m.dev;
            a.nsp = k;
            a.scList = nullptr;
            if (e->single) a.tileOut = c->plans[e->supp - 2][1].tiles;
            push[i] = a;
            pic::FieldArgs f = makeFieldArgs(c);
            for (int d = 0; d < 3; ++d) {
                f.E[d] = E[d];
                field[i].Enew[d] = Ealt[d];
            }
            field[i].a = f;
        }
        PIC_HIP(hipMemcpyAsync(e->dPush[par], push.data(), n * sizeof(pic::PushArgs), hipMemcpyHostToDevice, e->stream));
        PIC_HIP(hipMemcpyAsync(e->dField[par], field.data(), n * sizeof(pic::FieldBatchEntry), hipMemcpyHostToDevice,
                               e->stream));
    }
    if (e->single) {
        std::vector<pic::GatherEntry> gather(n);
        for (int i = 0; i < n; ++i) {
            pic_ctx* c = e->sims[i];
            const PhasePlan& plan = c->plans[e->supp - 2][1];
            gather[i] = pic::GatherEntry{c->g, plan.tiles, plan.colour, {c->J[0], c->J[1], c->J[2]}};
        }
        if (!e->dGather && ensAlloc(e, &e->dGather, static_cast<size_t>(n)) != PIC_OK) return PIC_E_HIP;
        PIC_HIP(hipMemcpyAsync(e->dGather, gather.data(), n * sizeof(pic::GatherEntry), hipMemcpyHostToDevice,
                               e->stream));
    }
    std::vector<pic::ShiftEntry> shift;
    for (pic_ctx* c : e->sims)
        for (const SpeciesMem& m : c->sp)
            if (m.desc.pusher != PIC_PUSHER_NONE) shift.push_back(pic::ShiftEntry{m.dev, c->g});
    PIC_HIP(hipMemcpyAsync(e->dShift, shift.data(), shift.size() * sizeof(pic::ShiftEntry), hipMemcpyHostToDevice,
                           e->stream));
    PIC_HIP(hipStreamSynchronize(e->stream));
    return PIC_OK;
}

int64_t ensParticles(pic_ensemble* e, int i) {
    pic_ctx* c = e->sims[i];
    int64_t n = 0;
    for (const SpeciesMem& m : c->sp)
        if (m.desc.pusher != PIC_PUSHER_NONE) n += (int64_t)m.slots;
    return n;
}

// work lists of the batched push in the chosen order, XCD-interleaved per phase
int ensBuildWork(pic_ensemble* e) {
    int const n = static_cast<int>(e->sims.size());
    std::vector<PhasePlan*> plans(n);
    size_t nph = 0;
    for (int i = 0; i < n; ++i) {
        int const r = getPlan(e->sims[i], e->supp, true, &plans[i]);
        if (r != PIC_OK) return r;
        nph = std::max(nph, plans[i]->count.size());
    }
    std::vector<int> simOrder(n);
    for (int i = 0; i < n; ++i) simOrder[i] = i;
    if (e->order == PIC_ENSEMBLE_ORDER_HEAVY_FIRST)
        std::stable_sort(simOrder.begin(), simOrder.end(),
                         [&](int a, int b) { return ensParticles(e, a) > ensParticles(e, b); });
    // one launch for every member's supercells when every member runs its
    // own push in one launch: the items of all phases in one list
    e->single = true;
    for (int i = 0; i < n; ++i) e->single = e->single && plans[i]->single;
    std::vector<int2> work;
    size_t const lists = e->single ? 1 : nph;
    e->phaseOffset.assign(lists, 0);
    e->phaseCount.assign(lists, 0);
    std::vector<int2> items;
    for (size_t ph = 0; ph < nph; ++ph) {
        if (!e->single) items.clear();
        if (e->order == PIC_ENSEMBLE_ORDER_INTERLEAVED) {
            size_t longest = 0;
            for (int i = 0; i < n; ++i)
                if (ph < plans[i]->count.size()) longest = std::max(longest, (size_t)plans[i]->count[ph]);
            for (size_t k = 0; k < longest; ++k)
                for (int i : simOrder)
                    if (ph < plans[i]->count.size() && k < (size_t)plans[i]->count[ph])
                        items.push_back(int2{i, plans[i]->host[plans[i]->offset[ph] + k]});
        } else {
            for (int i : simOrder)
                if (ph < plans[i]->count.size())
                    for (int k = 0; k < plans[i]->count[ph]; ++k)
                        items.push_back(int2{i, plans[i]->host[plans[i]->offset[ph] + k]});
        }
        if (e->single && ph + 1 < nph) continue;  // one list of every phase's items
        // workgroup b runs on XCD b % 8: contiguous runs of items per XCD
        int const m = static_cast<int>(items.size());
        std::vector<int2> ordered(m);
        int pos = 0;
        for (int x = 0; x < 8; ++x)
            for (int b = x; b < m; b += 8) ordered[b] = items[pos++];
        size_t const l = e->single ? 0 : ph;
        e->phaseOffset[l] = static_cast<int>(work.size());
        e->phaseCount[l] = m;
        work.insert(work.end(), ordered.begin(), ordered.end());
    }
    if (!e->dPhaseWork && ensAlloc(e, &e->dPhaseWork, work.size()) != PIC_OK) return PIC_E_HIP;
    PIC_HIP(hipMemcpyAsync(e->dPhaseWork, work.data(), work.size() * sizeof(int2), hipMemcpyHostToDevice, e->stream));
    PIC_HIP(hipStreamSynchronize(e->stream));
    return PIC_OK;
}

int ensSyncMembers(pic_ensemble* e) {
    for (pic_ctx* c : e->sims) PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

// one batched step: J reset, push + deposit per phase over all members, shift,
// fused field step; the members' E/Ealt swap as in pic_step
int ensOneStep(pic_ensemble* e) {
    int const p = e->parity;
    PIC_HIP(pic::launchCurrentResetBatch(e->dField[p], e->dBlocks, e->nBlocks, e->stream));
    for (size_t ph = 0; ph < e->phaseCount.size(); ++ph) {
        int const m = e->phaseCount[ph];
        int const grid = e->blocksPerCu > 0 ? std::min(m, e->blocksPerCu * e->cus) : m;
        PIC_HIP(pic::launchPushBatch(e->supp, e->pusher, e->dPush[p], e->dPhaseWork + e->phaseOffset[ph], m, grid,
                                     e->stream));
    }
    if (e->single) PIC_HIP(pic::launchGatherFlushBatch(e->supp, e->dGather, e->dBlocks, e->nBlocks, e->stream));
    PIC_HIP(pic::launchShiftBatch(e->dShift, e->dShiftWork, e->nShiftWork, e->stream));
    PIC_HIP(pic::launchFieldStepBatch(e->solver, e->dField[p], e->dTiles, e->nTiles, e->dBlocks, e->nBlocks,
                                      e->stream));
    PIC_HIP(pic::launchEndStepBatch(e->dField[p], static_cast<int>(e->sims.size()), e->stream));
    for (pic_ctx* c : e->sims) {
        for (int d = 0; d < 3; ++d) std::swap(c->E[d], c->Ealt[d]);
        c->eParity ^= 1;
    }
    e->parity ^= 1;
    return PIC_OK;
}

// After batched steps: a member whose shift overflowed skipped the steps
// after it; it grows (within the frames the batched push can list), and runs
// them on its own (its own steps equal its batched ones bit for bit). The
// members' buffers may have moved, so the device tables are written again
// (the captured ensemble graphs read them, so they stay valid).
int ensFinish(pic_ensemble* e) {
    PIC_HIP(hipStreamSynchronize(e->stream));
    bool moved = false;
    for (pic_ctx* c : e->sims) {
        for (SpeciesMem& m : c->sp) m.shiftReady = 0;
        int skipped = 0;
        uint32_t const raised = c->hostFlags;
        c->hostFlags = 0;
        int r = recoverCapacity(c, &skipped, pic::kListMaxSlots / pic::kSCV);
        if (c->hostFlags & PIC_FLAG_CAPACITY_GREW) moved = true;
        c->hostFlags |= raised;
        if (r != PIC_OK) return r;
        while (skipped > 0) {
            for (int i = 0; i < skipped; ++i)
                if ((r = oneStep(c)) != PIC_OK) return r;
            if ((r = recoverCapacity(c, &skipped, pic::kListMaxSlots / pic::kSCV)) != PIC_OK) return r;
        }
    }
    if (moved) return ensUploadTables(e);
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_ensemble_create(const pic_config* cfgs, int32_t n, pic_ensemble** out) {
    if (!cfgs || !out || n <= 0) return fail(PIC_E_INVALID, "need n > 0 configurations and an output pointer");
    *out = nullptr;
    pic_ensemble* e = new pic_ensemble();
    auto bail = [&](int code) {
        pic_ensemble_destroy(e);
        return code;
    };
    e->device = cfgs[0].device;
    for (int32_t i = 0; i < n; ++i) {
        const pic_config& c = cfgs[i];
        if (c.device != e->device) return bail(fail(PIC_E_INVALID, "ensemble members must share one device"));
        bool incident = false;
        for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) incident = incident || c.incident[a][side].kind != PIC_INCIDENT_NONE;
        if (!allPeriodic(c) || c.absorber != PIC_ABSORBER_NONE || c.current_interpolation != PIC_CURRENT_NONE || incident)
            return bail(fail(PIC_E_INVALID,
                             "ensemble members need periodic boxes without absorber, current filter or incident field"));
        int supp = 0, pusher = 0;
        for (int s = 0; s < c.n_species; ++s) {
            if (c.species[s].pusher == PIC_PUSHER_NONE) continue;
            int const sp = supportOf(c.species[s].shape);
            if ((supp && sp != supp) || (pusher && c.species[s].pusher != pusher))
                return bail(fail(PIC_E_INVALID, "ensemble mobile species must share one shape and pusher"));
            if (c.species[s].current_solver != PIC_CURRENT_SOLVER_ESIRKEPOV)
                return bail(fail(PIC_E_INVALID, "ensemble species use the Esirkepov current solver"));
            supp = sp;
            pusher = c.species[s].pusher;
            if (c.species[s].frames_per_sc * pic::kSCV > pic::kListMaxSlots)
                return bail(fail(PIC_E_INVALID, "ensemble species need at most 64 frames per supercell"));
        }
        if (i == 0) {
            e->supp = supp;
            e->pusher = pusher;
            e->solver = c.solver;
        } else if (supp != e->supp || pusher != e->pusher || c.solver != e->solver) {
            return bail(fail(PIC_E_INVALID, "ensemble members must share shape, pusher and solver"));
        }
        if (supp == 0) return bail(fail(PIC_E_INVALID, "every ensemble member needs a mobile species"));
        pic_ctx* ctx = nullptr;
        int const r = pic_create(&c, &ctx);
        if (r != PIC_OK) return bail(r);
        ctx->member = true;
        e->sims.push_back(ctx);
    }
    hipError_t he = hipSetDevice(e->device);
    if (he != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(he)));
    he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(he)));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) == hipSuccess && prop.multiProcessorCount > 0)
        e->cus = prop.multiProcessorCount;
    int nEntries = 0;
    for (pic_ctx* c : e->sims)
        for (const SpeciesMem& m : c->sp)
            if (m.desc.pusher != PIC_PUSHER_NONE) ++nEntries;
    e->nShiftEntries = nEntries;
    if (ensAlloc(e, &e->dPush[0], n) || ensAlloc(e, &e->dPush[1], n) || ensAlloc(e, &e->dField[0], n) ||
        ensAlloc(e, &e->dField[1], n) || ensAlloc(e, &e->dShift, nEntries))
        return bail(PIC_E_HIP);
    // shift items (entry, supercell), field tiles and 256-cell blocks per member
    std::vector<int2> shiftWork, tiles, blocks;
    int entry = 0;
    for (int32_t i = 0; i < n; ++i) {
        pic_ctx* c = e->sims[i];
        for (const SpeciesMem& m : c->sp) {
            if (m.desc.pusher == PIC_PUSHER_NONE) continue;
            for (int sc = 0; sc < c->nsc; ++sc) shiftWork.push_back(int2{entry, sc});
            ++entry;
        }
        int const nt = pic::fieldStepTiles(c->cfg.solver, c->g);
        for (int t = 0; t < nt; ++t) tiles.push_back(int2{i, t});
        int const nb = static_cast<int>((c->ncells + 255) / 256);
        for (int b = 0; b < nb; ++b) blocks.push_back(int2{i, b});
    }
    e->nShiftWork = static_cast<int>(shiftWork.size());
    e->nTiles = static_cast<int>(tiles.size());
    e->nBlocks = static_cast<int>(blocks.size());
    if (ensAlloc(e, &e->dShiftWork, shiftWork.size()) || ensAlloc(e, &e->dTiles, tiles.size()) ||
        ensAlloc(e, &e->dBlocks, blocks.size()))
        return bail(PIC_E_HIP);
    he = hipMemcpyAsync(e->dShiftWork, shiftWork.data(), shiftWork.size() * sizeof(int2), hipMemcpyHostToDevice,
                        e->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(e->dTiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(e->dBlocks, blocks.data(), blocks.size() * sizeof(int2), hipMemcpyHostToDevice,
                            e->stream);
    if (he != hipSuccess) return bail(fail(PIC_E_HIP, hipGetErrorString(he)));
    int r = ensBuildWork(e);
    if (r != PIC_OK) return bail(r);
    r = ensUploadTables(e);
    if (r != PIC_OK) return bail(r);
    *out = e;
    return PIC_OK;
}

int pic_ensemble_destroy(pic_ensemble* e) {
    if (!e) return PIC_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ensDropGraphs(e);
    for (void* p : e->allocations) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    for (pic_ctx* c : e->sims) pic_destroy(c);
    delete e;
    return PIC_OK;
}

int32_t pic_ensemble_size(const pic_ensemble* e) { return e ? static_cast<int32_t>(e->sims.size()) : -1; }

pic_ctx* pic_ensemble_member(pic_ensemble* e, int32_t i) {
    if (!e || i < 0 || i >= static_cast<int32_t>(e->sims.size())) {
        fail(PIC_E_INVALID, "ensemble member index out of range");
        return nullptr;
    }
    return e->sims[i];
}

int pic_ensemble_step(pic_ensemble* e, int32_t nsteps) {
    if (!e) return fail(PIC_E_INVALID, "null ensemble");
    if (nsteps < 0) return fail(PIC_E_INVALID, "nsteps must be >= 0");
    for (pic_ctx* c : e->sims)
        if (c->eParity != e->parity) return fail(PIC_E_STATE, "an ensemble member was stepped on its own");
    int r = ensSyncMembers(e);
    if (r != PIC_OK) return r;
    for (pic_ctx* c : e->sims) updateJexp(c);
    if ((r = ensUploadTables(e)) != PIC_OK) return r;
    for (int32_t i = 0; i < nsteps; ++i)
        if ((r = ensOneStep(e)) != PIC_OK) return r;
    return ensFinish(e);
}

int pic_ensemble_step_graph(pic_ensemble* e, int32_t nsteps) {
    if (!e) return fail(PIC_E_INVALID, "null ensemble");
    if (nsteps < 0) return fail(PIC_E_INVALID, "nsteps must be >= 0");
    for (pic_ctx* c : e->sims)
        if (c->eParity != e->parity) return fail(PIC_E_STATE, "an ensemble member was stepped on its own");
    int r = ensSyncMembers(e);
    if (r != PIC_OK) return r;
    for (pic_ctx* c : e->sims) updateJexp(c);
    if ((r = ensUploadTables(e)) != PIC_OK) return r;
    for (int32_t i = 0; i < nsteps; ++i) {
        int const p = e->parity;
        if (!e->graphExec[p]) {
            PIC_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            r = ensOneStep(e);  // records the step and swaps parity as its replay will
            hipGraph_t g = nullptr;
            hipError_t const he = hipStreamEndCapture(e->stream, &g);
            if (r != PIC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            if (he != hipSuccess) return fail(PIC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
            e->graph[p] = g;
            PIC_HIP(hipGraphInstantiate(&e->graphExec[p], e->graph[p], nullptr, nullptr, 0));
        } else {
            for (pic_ctx* c : e->sims) {
                for (int d = 0; d < 3; ++d) std::swap(c->E[d], c->Ealt[d]);
                c->eParity ^= 1;
            }
            e->parity ^= 1;
        }
        PIC_HIP(hipGraphLaunch(e->graphExec[p], e->stream));
    }
    return ensFinish(e);
}

int pic_ensemble_set_schedule(pic_ensemble* e, int32_t order, int32_t blocks_per_cu) {
    if (!e) return fail(PIC_E_INVALID, "null ensemble");
    if (order < PIC_ENSEMBLE_ORDER_SIM_MAJOR || order > PIC_ENSEMBLE_ORDER_HEAVY_FIRST || blocks_per_cu < 0 ||
        blocks_per_cu > 64)
        return fail(PIC_E_INVALID, "unknown schedule");
    PIC_HIP(hipStreamSynchronize(e->stream));
    e->order = order;
    e->blocksPerCu = blocks_per_cu;
    ensDropGraphs(e);
    return ensBuildWork(e);
}

int pic_ensemble_get_schedule(const pic_ensemble* e, int32_t* order, int32_t* blocks_per_cu) {
    if (!e) return fail(PIC_E_INVALID, "null ensemble");
    if (order) *order = e->order;
    if (blocks_per_cu) *blocks_per_cu = e->blocksPerCu;
    return PIC_OK;
}

int pic_ensemble_autotune(pic_ensemble* e, int32_t trial_steps, double* best_ms_per_step, int32_t* steps_taken) {
    if (!e) return fail(PIC_E_INVALID, "null ensemble");
    if (trial_steps < 1) return fail(PIC_E_INVALID, "trial_steps must be >= 1");
    int const orders[3] = {PIC_ENSEMBLE_ORDER_SIM_MAJOR, PIC_ENSEMBLE_ORDER_INTERLEAVED,
                           PIC_ENSEMBLE_ORDER_HEAVY_FIRST};
    int const grids[4] = {0, 1, 2, 4};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    PIC_HIP(hipEventCreate(&t0));
    PIC_HIP(hipEventCreate(&t1));
    double best = 1e300;
    int bestOrder = e->order, bestGrid = e->blocksPerCu;
    int taken = 0;
    int r = PIC_OK;
    for (int o : orders) {
        for (int gpc : grids) {
            if ((r = pic_ensemble_set_schedule(e, o, gpc)) != PIC_OK) break;
            if ((r = pic_ensemble_step(e, 1)) != PIC_OK) break;  // warm-up of this schedule
            ++taken;
            if ((r = ensSyncMembers(e)) != PIC_OK) break;
            for (pic_ctx* c : e->sims) updateJexp(c);
            if ((r = ensUploadTables(e)) != PIC_OK) break;
            if (hipEventRecord(t0, e->stream) != hipSuccess) {
                r = fail(PIC_E_HIP, "hipEventRecord");
                break;
            }
            for (int k = 0; k < trial_steps && r == PIC_OK; ++k) r = ensOneStep(e);
            if (r != PIC_OK) break;
            taken += trial_steps;
            float ms = 0.0f;
            if (hipEventRecord(t1, e->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&ms, t0, t1) != hipSuccess) {
                r = fail(PIC_E_HIP, "autotune timing");
                break;
            }
            if ((r = ensFinish(e)) != PIC_OK) break;
            double const per = ms / trial_steps;
            if (per < best) {
                best = per;
                bestOrder = o;
                bestGrid = gpc;
            }
        }
        if (r != PIC_OK) break;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (r != PIC_OK) return r;
    if ((r = pic_ensemble_set_schedule(e, bestOrder, bestGrid)) != PIC_OK) return r;
    if (best_ms_per_step) *best_ms_per_step = best;
    if (steps_taken) *steps_taken = taken;
    return PIC_OK;
}

}  // extern "C"
