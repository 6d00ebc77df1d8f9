// CPU oracle of the PIC-cycle hot path.
//
// TEST INFRASTRUCTURE ONLY: this file is a scalar C++ restatement of the
// reference's algorithms, used by tests/ to check the HIP library, by
// __graft_entry__.smoke() as the checker and by bench.py as the timed CPU
// baseline. The product never links, loads or calls it.
//
// Every function follows the reference file:line it names, in the same
// floating-point operation order. The state type T is float (the FP32 oracle,
// bit-comparable with the device) or double (the FP64 twin used to state the
// error of FP32 results). Built with -ffp-contract=off, no -ffast-math.
//
// Two current-deposition orders are available:
//  order 0 ("device"): current of each particle added as int64 fixed point
//          into a per-supercell tile during the push (deposit fused with the
//          push in the frame of the particle's new cell), tiles flushed into
//          J in the checkerboard phase order of the device; the field
//          gather's sums and Esirkepov's W and prefix sums are multiply-adds
//          rounded once (std::fma), as the device computes them. Bit-identical
//          to the HIP library by construction of the same sums.
//  order 1 ("sequential"): the order and arithmetic of the reference's CPU
//          (omp2b) path, no multiply-add contracted:
//          push and shift first, then per supercell a float cache filled
//          particle by particle in reverse frame order and flushed with a
//          stride-(skip+1) checkerboard (Deposit.hpp:63-84, ForEach.hpp:314-324).
// Supercell shifts follow the device's canonical order (see pic_shift) so
// that the frame layouts can be compared slot by slot.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int SCX = 8, SCY = 8, SCZ = 4, SCV = 256;

// param/physicalConstants.param:31-48
constexpr double kC = 2.99792458e8;
constexpr double kMue0 = 1.25663706127e-6;
constexpr double kMe = 9.1093837139e-31;
constexpr double kQe = -1.602176634e-19;
constexpr double kPi = 3.141592653589793238462643383279502884197169399;

}  // namespace

extern "C" {

typedef struct oracle_species {
    double mass_ratio;
    double charge_ratio;
    int32_t shape;   // 1 CIC, 2 TSC, 3 PQS
    int32_t pusher;  // 0 none, 1 Boris, 2 Vay, 3 Higuera-Cary
    int32_t frames_per_sc;
    int32_t current_solver;  // 0 Esirkepov, 1 EmZ
    int32_t particle_boundary[3];  // 0 absorbing, 1 reflecting (open axes)
} oracle_species;

typedef struct oracle_incident_profile {
    int32_t kind;  // 0 none, 1 GaussianPulse, 2 PlaneWave
    int32_t polarisation;  // 0 linear, 1 circular
    double wavelength_si;
    double amplitude_si;
    double pulse_duration_si;
    double time_delay_si;
    float laser_phase;
    double direction[3];
    double polarisation_direction[3];
    double focus_position_si[3];
    int32_t focus_origin[3];  // 0 zero, 1 center
    double w0_si;
    double pulse_init;
    int32_t num_modes;
    float laguerre_modes[16];
    float laguerre_phases[16];
    double laser_nofocus_constant_si;
    double ramp_init;
} oracle_incident_profile;

typedef struct oracle_pml_params {
    int32_t custom;                    // 0: the fieldAbsorber.param defaults
    double sigma_kappa_grading_order;  // SIGMA_KAPPA_GRADING_ORDER
    double sigma_max_si[3];            // SIGMA_MAX_SI [S/m]
    double kappa_max[3];               // KAPPA_MAX
    double alpha_grading_order;        // ALPHA_GRADING_ORDER
    double alpha_max_si[3];            // ALPHA_MAX_SI [S/m]
} oracle_pml_params;

typedef struct oracle_config {
    int32_t grid[3];
    double cell_si[3];
    double dt_si;
    double base_density_si;
    int32_t typical_ppc;
    int32_t solver;  // 0 Yee, 1 Lehe
    int32_t lehe_dir;
    int32_t n_species;
    oracle_species species[4];
    int32_t j_scale_exp;
    int32_t device;
    int32_t boundary[3];            // 0 periodic, 1 open (absorbing particles, zero fields beyond)
    int32_t absorber;               // 0 none, 1 exponential
    int32_t absorber_cells[3][2];   // fieldAbsorber.param NUM_CELLS [axis][negative, positive]
    float absorber_strength[3][2];  // fieldAbsorber.param exponential::STRENGTH
    int32_t current_interpolation;  // 0 None, 1 Binomial
    oracle_incident_profile incident[3][2];  // incidentField.param XMin..ZMax as [axis][side]
    int32_t incident_position[3][2];         // incidentField.param POSITION
    oracle_pml_params pml;                   // fieldAbsorber.param namespace pml (absorber 2)
} oracle_config;

typedef struct oracle_constants {
    float cell[3];
    float dt;
    float c;
    float eps0;
    float mue0;
    float mue0_eps0;
    float add_current_coeff;
    float inv_cell_volume_dt;
    float lehe_alpha;
    float lehe_delta;
    float species_mass[4];
    float species_charge[4];
    double unit_length, unit_time, unit_mass, unit_charge;
    double typical_weighting;
} oracle_constants;

}  // extern "C"

namespace {

// Constants in PIC units, computed in type T as unitless/simulation.unitless
// does for T = float_X; for T = double the same formulas without the casts.
template <typename T>
struct Consts {
    T cell[3];
    T dt, c, eps0, mue0, mue0_eps0, coeff, inv_vdt;
    T lehe_alpha, lehe_delta, lehe_beta1, lehe_beta2;
    T mass[4], charge[4];
    double unit_length, unit_time, unit_mass, unit_charge, wtyp;
    // the device order's multiply-adds (order 0): the field gather's sums and
    // Esirkepov's W and prefix sums rounded once; the reference's CPU order
    // rounds the product and the sum separately
    bool contract = false;
};

// a * b + c, rounded once when C (std::fma) and twice otherwise
template <bool C, typename T>
inline T madd(T a, T b, T c) {
    if constexpr (C)
        return std::fma(a, b, c);
    else
        return a * b + c;
}

// Esirkepov's weight S0i S0j + 1/2 (dSi S0j + S0i dSj) + 1/3 dSj dSi
// (Esirkepov.hpp:218) in the reference's order; when C every addition is
// fused with the product it adds
template <bool C, typename T>
inline T esirkepovWeight(T s0i, T dsi, T sj, T dsj) {
    if constexpr (C) {
        T const mixed = std::fma(dsi, sj, s0i * dsj);
        T const half = std::fma(T(0.5), mixed, s0i * sj);
        return std::fma((T(1.0) / T(3.0)) * dsj, dsi, half);
    } else {
        return s0i * sj + T(0.5) * (dsi * sj + s0i * dsj) + (T(1.0) / T(3.0)) * dsj * dsi;
    }
}

template <typename T>
Consts<T> makeConsts(const oracle_config& cfg) {
    Consts<T> k{};
    double const unitTime = cfg.dt_si;                  // simulation.unitless:435-438
    double const unitSpeed = kC;                        // :440-443
    double const unitLength = unitTime * unitSpeed;     // :430-433
    double const vol = cfg.cell_si[0] * cfg.cell_si[1] * cfg.cell_si[2];
    double const wTyp = (cfg.base_density_si * vol) / static_cast<double>(cfg.typical_ppc);  // :474-478
    double const unitMass = kMe * wTyp;                 // :445-448
    double const unitCharge = -1.0 * kQe * wTyp;        // :450-453
    k.unit_length = unitLength;
    k.unit_time = unitTime;
    k.unit_mass = unitMass;
    k.unit_charge = unitCharge;
    k.wtyp = wTyp;
    for (int d = 0; d < 3; ++d) k.cell[d] = static_cast<T>(cfg.cell_si[d] / unitLength);  // :36-40
    k.dt = static_cast<T>(cfg.dt_si / unitTime);                                         // :42-46
    k.c = static_cast<T>(kC / unitSpeed);                                                // :48-52
    k.mue0 = static_cast<T>(kMue0 / unitLength / unitMass * unitCharge * unitCharge);    // :84-88
    k.eps0 = static_cast<T>(1. / static_cast<double>(k.mue0) / static_cast<double>(k.c) / static_cast<double>(k.c));
    k.mue0_eps0 = static_cast<T>(1. / static_cast<double>(k.c) / static_cast<double>(k.c));
    k.coeff = -(T(1.0) / k.eps0) * k.dt;                 // FDTD.hpp:84
    T const vdt = k.cell[0] * k.cell[1] * k.cell[2] * k.dt;
    k.inv_vdt = T(1.0) / vdt;                            // Esirkepov.hpp:190-192
    T const baseMass = static_cast<T>(kMe / unitMass);
    T const baseCharge = static_cast<T>(kQe / unitCharge);
    for (int s = 0; s < cfg.n_species; ++s) {
        // speciesDefinition.unitless:45-66: base * ratio in float_X
        k.mass[s] = baseMass * static_cast<T>(cfg.species[s].mass_ratio);
        k.charge[s] = baseCharge * static_cast<T>(cfg.species[s].charge_ratio);
    }
    if (cfg.solver == 1) {  // Lehe/Derivative.hpp:90-111 (FP64 set-up), :124-130 (FP32 betas)
        int const d0 = cfg.lehe_dir, d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        double const stepRatio = static_cast<double>(k.cell[d0] / (k.c * k.dt));
        double const co = stepRatio * std::sin(kPi / 2.0 * static_cast<double>(k.c) * static_cast<double>(k.dt) /
                                               static_cast<double>(k.cell[d0]));
        k.lehe_delta = static_cast<T>(0.25 * (1.0 - co * co));
        double const sr1 = static_cast<double>(k.cell[d0] / k.cell[d1]);
        double const sr2 = static_cast<double>(k.cell[d0] / k.cell[d2]);
        double const b1 = 0.125 * sr1 * sr1;
        double const b2 = 0.125 * sr2 * sr2;
        k.lehe_alpha = static_cast<T>(1.0 - 2.0 * b1 - 2.0 * b2 - 3.0 * static_cast<double>(k.lehe_delta));
        T const r1 = k.cell[d0] / k.cell[d1];
        T const r2 = k.cell[d0] / k.cell[d2];
        k.lehe_beta1 = T(0.125) * r1 * r1;
        k.lehe_beta2 = T(0.125) * r2 * r2;
    }
    return k;
}

template <typename T>
struct V3 {
    T x, y, z;
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

// pmacc/math/vector/Vector.tpp:99-110
template <typename T>
T l2norm2(const V3<T>& v) {
    T t = v.x * v.x;
    t += v.y * v.y;
    t += v.z * v.z;
    return t;
}

// pmacc/math/vector/Vector.tpp:71-76
template <typename T>
V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---- shapes: particles/shapes/CIC.hpp, TSC.hpp, PQS.hpp -------------------
template <typename T>
struct ShapeFns {
    int supp;
    int begin() const { return supp == 2 ? 0 : -1; }  // ChargeAssignment::begin
    // detail::X::shapeArray (on support) -> supp values
    void onSupport(T x, T* s) const {
        if (supp == 2) {  // CIC.hpp:56-66
            s[0] = T(1.0) - x;
            s[1] = x;
        } else if (supp == 3) {  // TSC.hpp:75-86
            s[0] = tscFF2(std::fabs(T(-1.0) - x));
            s[1] = tscFF1(std::fabs(x));
            s[2] = T(1.0) - (s[0] + s[1]);
        } else {  // PQS.hpp:78-89
            s[0] = pqsFF2(std::fabs(T(-1.0) - x));
            s[1] = pqsFF1(x);
            s[3] = pqsFF2(T(2.0) - x);
            s[2] = T(1.0) - (s[0] + s[1] + s[3]);
        }
    }
    static T tscFF1(T x) {  // TSC.hpp:47-54
        T const sq = x * x;
        return T(0.75) - sq;
    }
    static T tscFF2(T x) {  // TSC.hpp:56-64
        T const tmp = T(3.0) / T(2.0) - x;
        T const sq = tmp * tmp;
        return T(0.5) * sq;
    }
    static T pqsFF1(T x) {  // PQS.hpp:47-55
        T const sq = x * x;
        T const tr = sq * x;
        return T(1.0) / T(6.0) * (T(4.0) - T(6.0) * sq + T(3.0) * tr);
    }
    static T pqsFF2(T x) {  // PQS.hpp:57-65
        T const tmp = T(2.0) - x;
        T const tr = tmp * tmp * tmp;
        return T(1.0) / T(6.0) * tr;
    }
    // ChargeAssignment::operator() (CIC.hpp:89-107, TSC.hpp:109-132, PQS.hpp:110-133)
    T assign(T x) const {
        T const a = std::fabs(x);
        T r = T(0.0);
        if (supp == 2) {
            if (a < T(1.0)) r = T(1.0) - a;
        } else if (supp == 3) {
            T const r1 = tscFF1(a), r2 = tscFF2(a);
            if (a < T(0.5))
                r = r1;
            else if (a < T(1.5))
                r = r2;
        } else {
            T const r1 = pqsFF1(a), r2 = pqsFF2(a);
            if (a < T(1.0))
                r = r1;
            else if (a < T(2.0))
                r = r2;
        }
        return r;
    }
    // ChargeAssignmentOnSupport::operator() (CIC.hpp:139-147, TSC.hpp:166-185, PQS.hpp:166-185)
    T onSupportAt(T x) const {
        T const a = std::fabs(x);
        if (supp == 2) return T(1.0) - a;
        if (supp == 3) return a < T(0.5) ? tscFF1(a) : tscFF2(a);
        return a < T(1.0) ? pqsFF1(a) : pqsFF2(a);
    }
    // ChargeAssignment::shapeArray(pos, isOutOfRange) -> supp+1 values
    void cached(T pos, bool out, T* s) const {
        T const x = out ? pos - T(1.0) : pos;
        T v[5];
        onSupport(x, v);
        v[supp] = T(0.0);
        for (int i = supp; i >= 1; --i) s[i] = out ? v[i - 1] : v[i];
        s[0] = out ? T(0.0) : v[0];
    }
};

// ---- state ------------------------------------------------------------------
template <typename T>
struct Species {
    oracle_species desc;
    int supp = 2;
    int cap = 0;
    T mass = 0, charge = 0;
    std::vector<T> pos[3], mom[3], w;
    std::vector<uint16_t> cell;
    std::vector<uint8_t> mask;
    std::vector<uint32_t> count;
    std::vector<uint32_t> id;  // particle identity, follows the particle through shifts
};

template <typename T>
struct State {
    oracle_config cfg;
    Consts<T> k;
    int nx, ny, nz, nscx, nscy, nscz, nsc;
    size_t ncells;
    std::vector<T> E[3], B[3], J[3];
    std::vector<Species<T>> sp;
    int jexp = 40;
    int order = 0;
    bool capacityError = false;

    size_t idx(int x, int y, int z) const {
        x = ((x % nx) + nx) % nx;
        y = ((y % ny) + ny) % ny;
        z = ((z % nz) + nz) % nz;
        return static_cast<size_t>(x) + static_cast<size_t>(nx) * (static_cast<size_t>(y) + static_cast<size_t>(ny) * z);
    }
    // beyond the border of an open axis: the reference's guard cells there are
    // never written nor exchanged, so fields read zero and current deposited
    // there never reaches the core (FieldJ.x.cpp:156-174 exchanges with
    // neighbours only)
    bool anyOpen = false;
    bool outside(int x, int y, int z) const {
        if (!anyOpen) return false;
        return (cfg.boundary[0] && (x < 0 || x >= nx)) || (cfg.boundary[1] && (y < 0 || y >= ny)) ||
               (cfg.boundary[2] && (z < 0 || z >= nz));
    }
    T at(const std::vector<T>& f, int x, int y, int z) const { return outside(x, y, z) ? T(0) : f[idx(x, y, z)]; }
    int64_t step = 0;  // currentStep of the field solver calls
    // PML (absorber 2): convolution fields of E and B (yx, zx, xy, zy, xz,
    // yz) and the coefficient tables of E and B positions (pmlTables)
    bool pml = false;
    std::vector<T> psiE[6], psiB[6];
    std::vector<float> pmlE, pmlB;
    // the PML's weight damping of absorbing species (pmlDampSetup)
    int dampAxes = 0;
    int dampLo[3] = {0, 0, 0}, dampHi[3] = {0, 0, 0};
    float dampNeg[3] = {0, 0, 0}, dampPos[3] = {0, 0, 0}, dampN[3] = {0, 0, 0}, dampSigma[3] = {0, 0, 0};
    float dampGrading = 0.0f;
    int dampGuard[3] = {SCX, SCY, SCZ};
};

// PML coefficients (absorber/pml/Pml.kernel:60-270: getRelativeDepth,
// getAbsorptionParameters, getCoefficients) per axis and index, for E (cell
// index) or B (index + 1/2), in float_X arithmetic; the layer is
// absorber_cells thick on open axes, 0 on periodic ones; the maxima are
// fieldAbsorber.param's (SIGMA_OPT_SI = 0.8 (m + 1) / (Z0 dx), KAPPA_MAX 1,
// ALPHA_MAX_SI 0.2, orders 4 and 1) unless given, normalised as
// fieldAbsorber.unitless:72-100 (/ eps0 * unit time). Layout: b of x, y, z,
// then c, then kappa.
void pmlTables(const oracle_config& cfg, const Consts<float>& k, bool half, std::vector<float>& out) {
    int const n[3] = {cfg.grid[0], cfg.grid[1], cfg.grid[2]};
    int const total = n[0] + n[1] + n[2];
    out.assign(3 * static_cast<size_t>(total), 0.0f);
    double const z0 = kMue0 * kC;
    double const eps0 = 1.0 / kMue0 / kC / kC;
    double mSK = 4.0, mA = 1.0;
    double sigmaMax[3], kappaMax[3] = {1.0, 1.0, 1.0}, alphaMax[3] = {0.2, 0.2, 0.2};
    for (int d = 0; d < 3; ++d) sigmaMax[d] = 0.8 * (mSK + 1.0) / (z0 * cfg.cell_si[d]);
    if (cfg.pml.custom != 0) {
        mSK = cfg.pml.sigma_kappa_grading_order;
        mA = cfg.pml.alpha_grading_order;
        for (int d = 0; d < 3; ++d) {
            sigmaMax[d] = cfg.pml.sigma_max_si[d];
            kappaMax[d] = cfg.pml.kappa_max[d];
            alphaMax[d] = cfg.pml.alpha_max_si[d];
        }
    }
    float const m = static_cast<float>(mSK), ma = static_cast<float>(mA);
    int off = 0;
    for (int d = 0; d < 3; ++d) {
        bool const open = cfg.boundary[d] != 0;
        float const negB = static_cast<float>(open ? cfg.absorber_cells[d][0] : 0);
        float const posB = static_cast<float>(open ? cfg.absorber_cells[d][1] : 0);
        float const nSigma = static_cast<float>(sigmaMax[d] / eps0 * k.unit_time);
        float const nAlpha = static_cast<float>(alphaMax[d] / eps0 * k.unit_time);
        float const kMax = static_cast<float>(kappaMax[d]);
        for (int i = 0; i < n[d]; ++i) {
            float const zero = static_cast<float>(i) + (half ? 0.5f : 0.0f);  // cellIdx - numGuardCells
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
                alpha = nAlpha * std::pow(1.0f - depth, ma);
            }
            float const b = std::exp(-(sigma / kappa + alpha) * k.dt);
            float const den = kappa * (sigma + alpha * kappa);
            out[off + i] = b;
            out[total + off + i] = den != 0.0f ? sigma * (b - 1.0f) / den : 0.0f;
            out[2 * total + off + i] = kappa;
        }
        off += n[d];
    }
}

// the damping areas and layer parameters of dampWeightsInPml per open axis
// with a layer: the supercells the layer reaches at each side, the
// thicknesses, NORMALIZED_SIGMA_MAX and the grading order (as pmlTables)
template <typename T>
void pmlDampSetup(const oracle_config& cfg, const Consts<float>& k, State<T>& st) {
    double const z0 = kMue0 * kC;
    double const eps0 = 1.0 / kMue0 / kC / kC;
    double mSK = cfg.pml.custom ? cfg.pml.sigma_kappa_grading_order : 4.0;
    st.dampGrading = static_cast<float>(mSK);
    int const scs[3] = {SCX, SCY, SCZ};
    for (int d = 0; d < 3; ++d) {
        double const sigmaMax = cfg.pml.custom ? cfg.pml.sigma_max_si[d] : 0.8 * (4.0 + 1.0) / (z0 * cfg.cell_si[d]);
        bool const open = cfg.boundary[d] != 0;
        int const n = cfg.grid[d];
        int const neg = open ? cfg.absorber_cells[d][0] : 0, pos = open ? cfg.absorber_cells[d][1] : 0;
        st.dampNeg[d] = static_cast<float>(neg);
        st.dampPos[d] = static_cast<float>(pos);
        st.dampN[d] = static_cast<float>(n);
        st.dampSigma[d] = static_cast<float>(sigmaMax / eps0 * k.unit_time);
        st.dampLo[d] = neg > 0 ? scs[d] * ((neg + scs[d] - 1) / scs[d]) : 0;
        st.dampHi[d] = pos > 0 ? n - scs[d] * ((pos + scs[d] - 1) / scs[d]) : n;
        if (open && (neg > 0 || pos > 0)) st.dampAxes |= 1 << d;
    }
}

int supportOf(int shape) { return shape == 1 ? 2 : (shape == 2 ? 3 : 4); }

// ---- field gather: FieldToParticleInterpolation.hpp:100-124,
// ShiftCoordinateSystem.hpp:53-79, AssignedTrilinearInterpolation.hpp:53-86
template <typename T, bool OPEN, bool C>
T gather1Impl(const State<T>& st, const std::vector<T>& F, int supp, int gx, int gy, int gz, const T* pos,
              const T* fpos) {
    ShapeFns<T> sh{supp};
    bool const isEven = supp % 2 == 0;
    int const begin = -supp / 2 + (supp + 1) % 2;  // FieldToParticleInterpolation.hpp:60
    T p[3];
    int is[3];
    for (int d = 0; d < 3; ++d) {
        T const v_pos = pos[d] - fpos[d] - T(0.5);
        if (isEven)
            is[d] = v_pos >= T(-0.5) ? 0 : -1;
        else
            is[d] = v_pos >= T(0.0) ? 1 : 0;
        p[d] = v_pos - T(is[d]) + T(0.5);
    }
    T sx[4], sy[4], sz[4];
    sh.onSupport(p[0], sx);
    sh.onSupport(p[1], sy);
    sh.onSupport(p[2], sz);
    // AssignedTrilinearInterpolation.hpp:53-86 sums into zero-initialised
    // partial sums; here each starts at its first term, which differs only in
    // the sign of an all-zero sum
    T rz = T(0.0);
    for (int z = 0; z < supp; ++z) {
        T ry = T(0.0);
        for (int y = 0; y < supp; ++y) {
            T rx = T(0.0);
            for (int x = 0; x < supp; ++x) {
                T const f = OPEN ? st.at(F, gx + is[0] + begin + x, gy + is[1] + begin + y, gz + is[2] + begin + z)
                                 : F[st.idx(gx + is[0] + begin + x, gy + is[1] + begin + y, gz + is[2] + begin + z)];
                rx = x == 0 ? f * sx[x] : madd<C>(f, sx[x], rx);
            }
            ry = y == 0 ? rx * sy[y] : madd<C>(rx, sy[y], ry);
        }
        rz = z == 0 ? ry * sz[z] : madd<C>(ry, sz[z], rz);
    }
    return rz;
}

// the periodic form is kept free of border checks (they slow the loop down)
template <typename T>
T gather1(const State<T>& st, const std::vector<T>& F, int supp, int gx, int gy, int gz, const T* pos, const T* fpos) {
    if (st.k.contract)
        return st.anyOpen ? gather1Impl<T, true, true>(st, F, supp, gx, gy, gz, pos, fpos)
                          : gather1Impl<T, false, true>(st, F, supp, gx, gy, gz, pos, fpos);
    return st.anyOpen ? gather1Impl<T, true, false>(st, F, supp, gx, gy, gz, pos, fpos)
                      : gather1Impl<T, false, false>(st, F, supp, gx, gy, gz, pos, fpos);
}

// ---- pushers ------------------------------------------------------------------
// Gamma.hpp:28-38
template <typename T>
T gammaOf(const V3<T>& mom, T mass, T c) {
    T const fMom2 = l2norm2(mom);
    T const c2 = c * c;
    T const m2c2r = T(1.0) / (mass * mass * c2);
    return std::sqrt(T(1.0) + fMom2 * m2c2r);
}

// Velocity.hpp:28-39 with the CPU rsqrt = 1/sqrt (alpaka math/Traits.hpp:729-734)
template <typename T>
V3<T> velocityOf(const V3<T>& mom, T mass, T rc2) {
    T const m0_2 = mass * mass;
    T const fMom2 = l2norm2(mom);
    T const t = T(1.0) / std::sqrt(m0_2 + fMom2 * rc2);
    return {t * mom.x, t * mom.y, t * mom.z};
}

// particlePusherBoris.hpp:41-90
template <typename T>
void boris(const V3<T>& e, const V3<T>& b, V3<T>& mom, V3<T>& vel, T mass, T charge, const Consts<T>& k) {
    T const QoM = charge / mass;
    T const deltaT = k.dt;
    V3<T> const mm{mom.x + T(0.5) * charge * e.x * deltaT, mom.y + T(0.5) * charge * e.y * deltaT,
                   mom.z + T(0.5) * charge * e.z * deltaT};
    T const gamma_reci = T(1.0) / gammaOf(mm, mass, k.c);
    V3<T> const t{T(0.5) * QoM * b.x * gamma_reci * deltaT, T(0.5) * QoM * b.y * gamma_reci * deltaT,
                  T(0.5) * QoM * b.z * gamma_reci * deltaT};
    T const f = T(1.0) / (T(1.0) + l2norm2(t));
    V3<T> const s{T(2.0) * t.x * f, T(2.0) * t.y * f, T(2.0) * t.z * f};
    V3<T> const c1 = cross(mm, t);
    V3<T> const mp{mm.x + c1.x, mm.y + c1.y, mm.z + c1.z};
    V3<T> const c2 = cross(mp, s);
    V3<T> const mpl{mm.x + c2.x, mm.y + c2.y, mm.z + c2.z};
    mom = {mpl.x + T(0.5) * charge * e.x * deltaT, mpl.y + T(0.5) * charge * e.y * deltaT,
           mpl.z + T(0.5) * charge * e.z * deltaT};
    vel = velocityOf(mom, mass, k.mue0_eps0);
}

// particlePusherVay.hpp:42-111, sqrt_Vay = precision64Bit (param/pusher.param:62)
template <typename T>
void vay(const V3<T>& e, const V3<T>& b, V3<T>& mom, V3<T>& vel, T mass, T charge, const Consts<T>& k) {
    T const deltaT = k.dt;
    T const factor = static_cast<T>(0.5 * static_cast<double>(charge) * static_cast<double>(deltaT));
    V3<T> const vmh = velocityOf(mom, mass, k.mue0_eps0);
    V3<T> const cvb = cross(vmh, b);
    V3<T> const m0{mom.x + factor * (e.x + cvb.x), mom.y + factor * (e.y + cvb.y), mom.z + factor * (e.z + cvb.z)};
    V3<T> const mp{m0.x + factor * e.x, m0.y + factor * e.y, m0.z + factor * e.z};
    T const gp = gammaOf(mp, mass, k.c);
    T const fm = factor / mass;
    double const tx = static_cast<double>(fm * b.x), ty = static_cast<double>(fm * b.y),
                 tz = static_cast<double>(fm * b.z);
    double dot = static_cast<double>(mp.x) * tx;
    dot += static_cast<double>(mp.y) * ty;
    dot += static_cast<double>(mp.z) * tz;
    double const u_star = dot / static_cast<double>(k.c * mass);
    double tau2 = tx * tx;
    tau2 += ty * ty;
    tau2 += tz * tz;
    double const sigma = static_cast<double>(gp * gp) - tau2;
    double const gph = std::sqrt(0.5 * (sigma + std::sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))));
    double const rg = 1.0 / gph;
    V3<T> const t{static_cast<T>(tx * rg), static_cast<T>(ty * rg), static_cast<T>(tz * rg)};
    T const s = T(1.0) / (T(1.0) + l2norm2(t));
    T d = mp.x * t.x;
    d += mp.y * t.y;
    d += mp.z * t.z;
    V3<T> const cpt = cross(mp, t);
    mom = {s * (mp.x + d * t.x + cpt.x), s * (mp.y + d * t.y + cpt.y), s * (mp.z + d * t.z + cpt.z)};
    vel = velocityOf(mom, mass, k.mue0_eps0);
}

// particlePusherHigueraCary.hpp:61-142 (second half acceleration with mom_plus),
// auxiliary quantities in FP64 (param/pusher.param:71), Gamma<> in T
template <typename T>
void higueraCary(const V3<T>& e, const V3<T>& b, V3<T>& mom, V3<T>& vel, T mass, T charge, const Consts<T>& k) {
    T const deltaT = k.dt;
    T const hq = T(0.5) * charge;
    V3<T> const de{hq * e.x * deltaT, hq * e.y * deltaT, hq * e.z * deltaT};
    V3<T> const mm{mom.x + de.x, mom.y + de.y, mom.z + de.z};
    double const mx = static_cast<double>(mm.x), my = static_cast<double>(mm.y), mz = static_cast<double>(mm.z);
    double const gm = static_cast<double>(gammaOf(mm, mass, k.c));
    double const tx = static_cast<double>(T(0.5) * b.x * charge * deltaT / mass);
    double const ty = static_cast<double>(T(0.5) * b.y * charge * deltaT / mass);
    double const tz = static_cast<double>(T(0.5) * b.z * charge * deltaT / mass);
    double tau2 = tx * tx;
    tau2 += ty * ty;
    tau2 += tz * tz;
    double const sigma = gm * gm - tau2;
    double dmt = mx * tx;
    dmt += my * ty;
    dmt += mz * tz;
    double const u_star = dmt / static_cast<double>(mass * k.c);
    double const gp = std::sqrt(0.5 * (sigma + std::sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))));
    double const t0 = tx / gp, t1 = ty / gp, t2 = tz / gp;
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
    V3<T> const d2{static_cast<T>(py * t2 - pz * t1), static_cast<T>(pz * t0 - px * t2),
                   static_cast<T>(px * t1 - py * t0)};
    mom = {static_cast<T>(px) + (de.x + d2.x), static_cast<T>(py) + (de.y + d2.y), static_cast<T>(pz) + (de.z + d2.z)};
    vel = velocityOf(mom, mass, k.mue0_eps0);
}

// ---- moveParticle: particles/MoveParticle.hpp:47-160 -------------------------
// the cell part of moveParticle (MoveParticle.hpp:139-157): localCellIdx and
// multiMask after the cell steps dir; returns the multiMask value
int cellAfterSteps(const int* dirIn, uint16_t& localCellIdx, uint8_t& multiMask) {
    int const sc[3] = {SCX, SCY, SCZ};
    int dir[3] = {dirIn[0], dirIn[1], dirIn[2]};
    int newMultimask = 1;
    if (dir[0] != 0 || dir[1] != 0 || dir[2] != 0) {
        int const idx = localCellIdx;
        int lc[3] = {idx % SCX, (idx / SCX) % SCY, idx / (SCX * SCY)};
        for (int i = 0; i < 3; ++i) lc[i] += dir[i];
        for (int i = 0; i < 3; ++i)
            dir[i] = static_cast<uint32_t>(lc[i]) >= static_cast<uint32_t>(sc[i]) ? dir[i] : 0;
        for (int i = 0; i < 3; ++i) lc[i] -= dir[i] * sc[i];
        localCellIdx = static_cast<uint16_t>(lc[0] + SCX * (lc[1] + SCY * lc[2]));
        int exchangeType = 1;
        for (int i = 0; i < 3; ++i) {
            newMultimask += (dir[i] == -1 ? 2 : dir[i]) * exchangeType;
            exchangeType *= 3;
        }
        if (newMultimask >= 2) multiMask = static_cast<uint8_t>(newMultimask);
    }
    return newMultimask;
}

template <typename T>
int moveParticle(T* newPos, uint16_t& localCellIdx, uint8_t& multiMask, int* cellStep) {
    volatile T shift = T(0.5);
    int dir[3];
    for (int i = 0; i < 3; ++i) {
        T pos = newPos[i] - shift;
        T moveDir = T(0.0);
        if (pos < T(-0.5)) moveDir = T(-1.0);
        if (pos >= T(0.5)) moveDir = T(1.0);
        pos -= moveDir;
        newPos[i] = pos + shift;
        dir[i] = static_cast<int>(moveDir);
        cellStep[i] = dir[i];
    }
    return cellAfterSteps(dir, localCellIdx, multiMask);
}


// ---- Esirkepov: fields/currentDeposition/Esirkepov/Esirkepov.hpp:61-242 ----
// `add(comp, x, y, z, value)` receives each accumulated_J at cell (x,y,z)
// relative to the particle's cell.
template <bool C, typename T, typename Add>
void cptCurrent1D(int supp, const int* st, const T* p0, const T* p1, T csd, int comp, const int* perm, Add add) {
    if (p0[2] == p1[2]) return;  // :157
    ShapeFns<T> sh{supp};
    int const begin = sh.begin();
    int const end = begin + supp;
    T s0j[5], s1j[5], s0k[5], s1k[5];
    sh.cached(p0[1], !(st[1] & 2), s0j);  // ShapeMiddleLoop = Cached (Strategy.def:52-61)
    sh.cached(p1[1], !(st[1] & 4), s1j);
    sh.cached(p0[2], !(st[2] & 2), s0k);  // ShapeInnerLoop = Cached
    sh.cached(p1[2], !(st[2] & 4), s1k);
    int const leaveI = st[0] & 1, leaveJ = st[1] & 1, leaveK = st[2] & 1;
    for (int i = begin; i < end + 1; ++i) {
        if (!(i < end + leaveI)) continue;
        T const s0i = sh.assign(T(i) - p0[0]);  // ShapeOuterLoop = Jit
        T const dsi = sh.assign(T(i) - p1[0]) - s0i;
        for (int j = begin; j < end + 1; ++j) {
            if (!(j < end + leaveJ)) continue;
            T const sj = s0j[j - begin];
            T const dsj = s1j[j - begin] - sj;
            T const tmp = -csd * esirkepovWeight<C>(s0i, dsi, sj, dsj);
            T acc = T(0.0);
            for (int kk = begin; kk < end; ++kk) {
                if (!(kk < end + leaveK - 1)) continue;
                T const dsk = s1k[kk - begin] - s0k[kk - begin];
                // accumulated_J += W, W = dsk * tmp (Esirkepov.hpp:233-236); the
                // contracted sum starts at the first W
                acc = C ? (kk == begin ? dsk * tmp : madd<C>(dsk, tmp, acc)) : acc + dsk * tmp;
                int const ijk[3] = {i, j, kk};
                // permuted access (PermutatedFieldValueAccess.hpp:87-108)
                int rel[3];
                for (int a = 0; a < 3; ++a) rel[perm[a]] = ijk[a];
                add(comp, rel[0], rel[1], rel[2], acc);
            }
        }
    }
}

// EmZ, one virtual particle (EmZ/DepositCurrent.hpp:46-112): on-support
// shapes (Jit outer, Cached middle and inner), supp x supp x (supp-1) cells
// of its own shifted cell, no leave rows
template <bool C, typename T, typename Add>
void emzCpt1D(int supp, const T* p0, const T* p1, T csd, int comp, const int* perm, Add add) {
    if (p0[2] == p1[2]) return;
    ShapeFns<T> sh{supp};
    int const begin = sh.begin();
    int const end = begin + supp;
    T s0j[4], s1j[4], s0k[4], s1k[4];
    sh.onSupport(p0[1], s0j);
    sh.onSupport(p1[1], s1j);
    sh.onSupport(p0[2], s0k);
    sh.onSupport(p1[2], s1k);
    for (int i = begin; i < end; ++i) {
        T const s0i = sh.onSupportAt(T(i) - p0[0]);
        T const dsi = sh.onSupportAt(T(i) - p1[0]) - s0i;
        for (int j = begin; j < end; ++j) {
            T const sj = s0j[j - begin];
            T const dsj = s1j[j - begin] - sj;
            T const tmp = -csd * esirkepovWeight<C>(s0i, dsi, sj, dsj);
            T acc = T(0.0);
            for (int kk = begin; kk < end - 1; ++kk) {
                T const dsk = s1k[kk - begin] - s0k[kk - begin];
                acc = C ? (kk == begin ? dsk * tmp : madd<C>(dsk, tmp, acc)) : acc + dsk * tmp;
                int const ijk[3] = {i, j, kk};
                int rel[3];
                for (int a = 0; a < 3; ++a) rel[perm[a]] = ijk[a];
                add(comp, rel[0], rel[1], rel[2], acc);
            }
        }
    }
}

template <typename T, typename Add>
void emzCpt1DAs(bool contract, int supp, const T* p0, const T* p1, T csd, int comp, const int* perm, Add add) {
    if (contract)
        emzCpt1D<true>(supp, p0, p1, csd, comp, perm, add);
    else
        emzCpt1D<false>(supp, p0, p1, csd, comp, perm, add);
}

template <typename T, typename Add>
void cptCurrent1DAs(bool contract, int supp, const int* st, const T* p0, const T* p1, T csd, int comp,
                    const int* perm, Add add) {
    if (contract)
        cptCurrent1D<true>(supp, st, p0, p1, csd, comp, perm, add);
    else
        cptCurrent1D<false>(supp, st, p0, p1, csd, comp, perm, add);
}

// EmZ (currentDeposition/EmZ/EmZ.hpp:61-137): the move splits at the relay
// point (relayPoint.hpp:47-64) into one or two virtual particles
template <typename T, typename Add>
void emz(int supp, const T* pos, const V3<T>& vel, T charge, const Consts<T>& k, Add add) {
    bool const isEven = supp % 2 == 0;
    T const dp[3] = {vel.x * k.dt / k.cell[0], vel.y * k.dt / k.cell[1], vel.z * k.dt / k.cell[2]};
    T p0[3], relay[3];
    int s0[3], s1[3];
    for (int d = 0; d < 3; ++d) {
        p0[d] = pos[d] - dp[d];
        if (isEven) {
            s0[d] = static_cast<int>(std::floor(p0[d]));
            s1[d] = static_cast<int>(std::floor(pos[d]));
            relay[d] = s0[d] == s1[d] ? pos[d] : static_cast<T>(std::max(s0[d], s1[d]));
        } else {
            s0[d] = static_cast<int>(std::floor(p0[d] + T(0.5)));
            s1[d] = static_cast<int>(std::floor(pos[d] + T(0.5)));
            relay[d] = s0[d] == s1[d] ? pos[d] : static_cast<T>(s0[d] + s1[d]) / T(2.0);
        }
    }
    T const chargeDensity = charge / (k.cell[0] * k.cell[1] * k.cell[2]);
    auto virtualParticle = [&](const int* sh, const T* a, const T* b) {
        auto shifted = [&](int comp, int x, int y, int z, T v) { add(comp, x + sh[0], y + sh[1], z + sh[2], v); };
        {  // J_x: rotateOrigin<1,2,0>
            int const perm[3] = {1, 2, 0};
            T const q0[3] = {a[1], a[2], a[0]}, q1[3] = {b[1], b[2], b[0]};
            emzCpt1DAs(k.contract, supp, q0, q1, k.cell[0] * chargeDensity / k.dt, 0, perm, shifted);
        }
        {  // J_y: rotateOrigin<2,0,1>
            int const perm[3] = {2, 0, 1};
            T const q0[3] = {a[2], a[0], a[1]}, q1[3] = {b[2], b[0], b[1]};
            emzCpt1DAs(k.contract, supp, q0, q1, k.cell[1] * chargeDensity / k.dt, 1, perm, shifted);
        }
        {  // J_z
            int const perm[3] = {0, 1, 2};
            emzCpt1DAs(k.contract, supp, a, b, k.cell[2] * chargeDensity / k.dt, 2, perm, shifted);
        }
    };
    T a[3], b[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = p0[d] - T(s0[d]);
        b[d] = relay[d] - T(s0[d]);
    }
    virtualParticle(s0, a, b);
    if (s0[0] != s1[0] || s0[1] != s1[1] || s0[2] != s1[2]) {
        for (int d = 0; d < 3; ++d) {
            b[d] = pos[d] - T(s1[d]);
            a[d] = relay[d] - T(s1[d]);
        }
        virtualParticle(s1, a, b);
    }
}

// the species' current solver: Esirkepov or EmZ
template <typename T, typename Add>
void depositParticle(int scheme, int supp, const T* pos, const V3<T>& vel, T charge, const Consts<T>& k, Add add);

template <typename T, typename Add>
void esirkepov(int supp, const T* pos, const V3<T>& vel, T charge, const Consts<T>& k, Add add) {
    bool const isEven = supp % 2 == 0;
    T const dp[3] = {vel.x * k.dt / k.cell[0], vel.y * k.dt / k.cell[1], vel.z * k.dt / k.cell[2]};
    T p0[3], p1[3];
    int gs[3], st[3];
    for (int d = 0; d < 3; ++d) {
        p0[d] = pos[d] - dp[d];
        p1[d] = pos[d];
        int iS, iE;  // relayPoint.hpp:47-64
        if (isEven) {
            iS = static_cast<int>(std::floor(p0[d]));
            iE = static_cast<int>(std::floor(p1[d]));
        } else {
            iS = static_cast<int>(std::floor(p0[d] + T(0.5)));
            iE = static_cast<int>(std::floor(p1[d] + T(0.5)));
        }
        gs[d] = iS < iE ? iS : iE;
        st[d] = 0;
        st[d] |= gs[d] == iS ? 2 : 0;  // START_PARTICLE_IN_ASSIGNMENT_CELL
        st[d] |= gs[d] == iE ? 4 : 0;  // END_PARTICLE_IN_ASSIGNMENT_CELL
        st[d] |= iS != iE ? 1 : 0;     // LEAVE_CELL
        p0[d] -= T(gs[d]);
        p1[d] -= T(gs[d]);
    }
    auto shifted = [&](int comp, int x, int y, int z, T v) { add(comp, x + gs[0], y + gs[1], z + gs[2], v); };
    {  // J_x: rotateOrigin<1,2,0>, permutation (1,2,0)
        int const perm[3] = {1, 2, 0};
        T const q0[3] = {p0[1], p0[2], p0[0]}, q1[3] = {p1[1], p1[2], p1[0]};
        int const s3[3] = {st[1], st[2], st[0]};
        T const csd = charge * k.inv_vdt * k.cell[0];
        cptCurrent1DAs(k.contract, supp, s3, q0, q1, csd, 0, perm, shifted);
    }
    {  // J_y: rotateOrigin<2,0,1>
        int const perm[3] = {2, 0, 1};
        T const q0[3] = {p0[2], p0[0], p0[1]}, q1[3] = {p1[2], p1[0], p1[1]};
        int const s3[3] = {st[2], st[0], st[1]};
        T const csd = charge * k.inv_vdt * k.cell[1];
        cptCurrent1DAs(k.contract, supp, s3, q0, q1, csd, 1, perm, shifted);
    }
    {  // J_z
        int const perm[3] = {0, 1, 2};
        T const csd = charge * k.inv_vdt * k.cell[2];
        cptCurrent1DAs(k.contract, supp, st, p0, p1, csd, 2, perm, shifted);
    }
}

template <typename T, typename Add>
void depositParticle(int scheme, int supp, const T* pos, const V3<T>& vel, T charge, const Consts<T>& k, Add add) {
    if (scheme == 1)
        emz(supp, pos, vel, charge, k, add);
    else
        esirkepov(supp, pos, vel, charge, k, add);
}

int depLower(int supp) { return supp / 2 + 1 - (supp + 1) % 2; }  // Esirkepov.hpp:42
int depUpper(int supp) { return (supp + 1) / 2 + 1; }             // Esirkepov.hpp:43

// Checkerboard colouring of a periodic ring (same rule as the device plan).
int ringColours(int n, int s, std::vector<int>& colour) {
    colour.assign(n, 0);
    if (n <= s) {  // ring shorter than the stencil: every supercell its own colour
        for (int i = 0; i < n; ++i) colour[i] = i;
        return n;
    }
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

// Supercells grouped by phase, phases ascending, x fastest inside a phase.
template <typename T>
std::vector<std::vector<int>> phases(const State<T>& st, const int* stride) {
    std::vector<int> col[3];
    int const n[3] = {st.nscx, st.nscy, st.nscz};
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = ringColours(n[d], stride[d], col[d]);
    std::vector<std::vector<int>> ph(nc[0] * nc[1] * nc[2]);
    for (int z = 0; z < n[2]; ++z)
        for (int y = 0; y < n[1]; ++y)
            for (int x = 0; x < n[0]; ++x)
                ph[col[0][x] + nc[0] * (col[1][y] + nc[1] * col[2][z])].push_back(x + n[0] * (y + n[1] * z));
    return ph;
}

template <typename T>
void scCoords(const State<T>& st, int sc, int& sx, int& sy, int& sz) {
    sx = sc % st.nscx;
    sy = (sc / st.nscx) % st.nscy;
    sz = sc / (st.nscx * st.nscy);
}

// One particle: gather, push, move (PushParticlePerFrame, Particles.kernel:275-316).
// Returns the velocity and writes the cell step of the move.
// boundary::ReflectParticleIfOutside (boundary/Reflecting.hpp:38-66, offset
// 0) as boundary::apply runs it after the push, exchange by exchange in the
// order +x, -x, +y, -y, +z, -z (Apply.hpp, Utility.hpp:35-48): a particle
// whose cell lies beyond a reflecting border of an open axis gets the in-cell
// position -pos - epsilon (upper border) or 2 - pos (lower border) on that
// axis, the momentum component negated, and moveParticle (all three axes)
// moves it back. The deposit stage then sees the reflected particle, whose
// velocity component is the negated one. cell: global cell after the push;
// step: accumulated cell steps. Returns whether anything was reflected.
template <typename T>
bool reflectParticle(const State<T>& st, const Species<T>& S, int* cell, T* pos, V3<T>& mom, V3<T>& vel, int* step) {
    int const n[3] = {st.nx, st.ny, st.nz};
    bool any = false;
    for (int d = 0; d < 3; ++d) {
        if (S.desc.particle_boundary[d] != 1 || st.cfg.boundary[d] != 1) continue;
        bool const beyondEnd = cell[d] >= n[d];
        bool const beforeBegin = cell[d] < 0;
        if (!beyondEnd && !beforeBegin) continue;
        if (beyondEnd)
            pos[d] = -pos[d] - std::numeric_limits<T>::epsilon();
        else
            pos[d] = T(2.0) - pos[d];
        T* const m = d == 0 ? &mom.x : (d == 1 ? &mom.y : &mom.z);
        T* const v = d == 0 ? &vel.x : (d == 1 ? &vel.y : &vel.z);
        *m = -*m;
        *v = -*v;
        // moveParticle (MoveParticle.hpp:47-74) on the whole position
        volatile T shift = T(0.5);
        for (int i = 0; i < 3; ++i) {
            T p = pos[i] - shift;
            T moveDir = T(0.0);
            if (p < T(-0.5)) moveDir = T(-1.0);
            if (p >= T(0.5)) moveDir = T(1.0);
            p -= moveDir;
            pos[i] = p + shift;
            step[i] += static_cast<int>(moveDir);
            cell[i] += static_cast<int>(moveDir);
        }
        any = true;
    }
    return any;
}

// ---- PML weight damping of an absorbing particle boundary ----
// boundary/Absorbing.hpp:95-200 (dampWeightsInPml, DampWeightsInPml) with
// fields/absorber/pml/Pml.kernel:116-136 (getRelativeDepth), :196-216
// (getNormalizedSigma), :326-386 (getNormalizedSigmaIntegral) and
// traits/attribute/DampedWeighting.hpp:67-90 (dampWeighting, MIN_WEIGHTING
// 10 of particle.param). After boundary::apply removed the particles beyond
// the border, each exchange +x, -x, +y, -y, +z, -z of an open axis on which
// the species is absorbing damps the particles in the supercells its layer
// reaches (particle boundary offset 0): the integral of sigma / eps0 over the
// last step's straight path, back from the particle's position by its
// velocity, scales the weighting (kept at MIN_WEIGHTING or above) and the
// momentum. The reference's FP32 pow and exp are taken, as on the device,
// as FP64 results rounded to FP32 from the same FP64 routines: e^x by Cody
// and Waite's reduction x = k ln2 + r (ln2 in the two parts of fdlibm) and
// the Taylor polynomial of e^r to degree 13; ln d from d = m 2^e, m in
// [sqrt(1/2), sqrt(2)), as 2 atanh((m - 1) / (m + 1)) by its series to the
// 21st power; pow(d, g) = e^(g ln d), or repeated products for an integer
// g. They are pinned against the C library's exp and log in the damping's
// tests.
double expNegF64(double x) {
    if (x < -745.0) return 0.0;
    double const k = std::nearbyint(x * 1.4426950408889634);
    double const r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    static const double inv[14] = {1.0,
                                   1.0,
                                   0.5,
                                   1.0 / 6.0,
                                   1.0 / 24.0,
                                   1.0 / 120.0,
                                   1.0 / 720.0,
                                   1.0 / 5040.0,
                                   1.0 / 40320.0,
                                   1.0 / 362880.0,
                                   1.0 / 3628800.0,
                                   1.0 / 39916800.0,
                                   1.0 / 479001600.0,
                                   1.0 / 6227020800.0};
    double p = inv[13];
    for (int i = 12; i >= 0; --i) p = p * r + inv[i];
    return std::ldexp(p, static_cast<int>(k));
}

double logUnitF64(double d) {
    int e;
    double m = std::frexp(d, &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    double const s = (m - 1.0) / (m + 1.0);
    double const s2 = s * s;
    double q = 1.0 / 21.0;
    for (int n = 19; n >= 1; n -= 2) q = q * s2 + 1.0 / static_cast<double>(n);
    double const lnm = 2.0 * s * q;
    double const de = static_cast<double>(e);
    return (de * 6.93147180369123816490e-01 + lnm) + de * 1.90821492927058770002e-10;
}

// x^y: repeated FP64 products for an integer order in [0, 32], else
// e^(y ln x)
template <typename T>
T powOf(T x, T y) {
    if (y >= T(0) && y <= T(32) && y == std::floor(y)) {
        double r = 1.0;
        for (int i = 0; i < static_cast<int>(y); ++i) r = r * static_cast<double>(x);
        return static_cast<T>(r);
    }
    return static_cast<T>(expNegF64(static_cast<double>(y) * logUnitF64(static_cast<double>(x))));
}

template <typename T>
T expOf(T x) {
    return static_cast<T>(expNegF64(static_cast<double>(x)));
}

template <typename T>
T dampSigma(const State<T>& st, int d, T pos) {
    // getNormalizedSigma adds the guard cells, getRelativeDepth takes them off again
    T const withGuard = pos + static_cast<T>(st.dampGuard[d]);
    T const zero = withGuard - static_cast<T>(st.dampGuard[d]);
    T const neg = static_cast<T>(st.dampNeg[d]), posB = static_cast<T>(st.dampPos[d]);
    T depth = T(0);
    if (zero < neg) {
        depth = (neg - zero) / neg;
    } else {
        T const rightStart = static_cast<T>(st.dampN[d]) - posB;
        if (zero > rightStart) depth = (zero - rightStart) / posB;
    }
    if (depth == T(0)) return T(0);
    return static_cast<T>(st.dampSigma[d]) * powOf(depth, static_cast<T>(st.dampGrading));
}

template <typename T>
T dampIntegral(const State<T>& st, int d, T start, T finish, T dt) {
    T const vel = (finish - start) / dt;
    if (static_cast<double>(std::abs(vel)) < 1e-5 * static_cast<double>(st.k.cell[d])) {
        T const s0 = dampSigma(st, d, start), s1 = dampSigma(st, d, finish);
        return dt * T(0.5) * (s0 + s1);
    }
    T as = start, af = finish;
    T const negEnd = static_cast<T>(st.dampNeg[d]);
    if (finish < negEnd && start >= negEnd) as = negEnd;
    if (start < negEnd && finish >= negEnd) af = negEnd;
    T const posStart = static_cast<T>(st.dampN[d]) - static_cast<T>(st.dampPos[d]);
    if (finish > posStart && start <= posStart) as = posStart;
    if (start > posStart && finish <= posStart) af = posStart;
    T const adjT = (af - as) / vel;
    T depthCoeff = T(0);
    if (as <= negEnd) depthCoeff = as - negEnd;
    if (as >= posStart) depthCoeff = as - posStart;
    T const s0 = dampSigma(st, d, as), s1 = dampSigma(st, d, af);
    return (s1 * adjT + (s1 - s0) * depthCoeff / vel) / (static_cast<T>(st.dampGrading) + T(1));
}

// returns whether any exchange damped the particle
template <typename T>
bool dampInPml(const State<T>& st, const Species<T>& S, const int* cell, const T* pos, V3<T>& mom, T& w) {
    if (!st.dampAxes || st.outside(cell[0], cell[1], cell[2])) return false;
    bool any = false;
    for (int d = 0; d < 3; ++d) {
        if (!((st.dampAxes >> d) & 1) || S.desc.particle_boundary[d] == 1) continue;
        for (int side = 0; side < 2; ++side) {  // the max side's exchange comes first
            bool const inArea = side == 0 ? cell[d] >= st.dampHi[d] : cell[d] < st.dampLo[d];
            if (!inArea) continue;
            V3<T> const vel = velocityOf(mom, S.mass * w, st.k.mue0_eps0);
            T const current = static_cast<T>(cell[d]) + pos[d];
            T const previous = current - vel[d] * st.k.dt / st.k.cell[d];
            T const integral = dampIntegral(st, d, previous, current, st.k.dt);
            T const multiplier = expOf(-integral);
            T const attempted = w * multiplier;
            T const damped = std::max(attempted, T(10));
            T const ratio = damped / w;
            mom.x = mom.x * ratio;
            mom.y = mom.y * ratio;
            mom.z = mom.z * ratio;
            w = damped;
            any = true;
        }
    }
    return any;
}

template <typename T>
V3<T> pushOne(State<T>& st, Species<T>& S, size_t g, int ox, int oy, int oz, int* step) {
    int const lc = S.cell[g];
    int const cx = ox + lc % SCX, cy = oy + (lc / SCX) % SCY, cz = oz + lc / (SCX * SCY);
    T pos[3] = {S.pos[0][g], S.pos[1][g], S.pos[2][g]};
    V3<T> mom{S.mom[0][g], S.mom[1][g], S.mom[2][g]};
    T w = S.w[g];
    T const mass = S.mass * w;
    T const charge = S.charge * w;
    // YeeCell.hpp:69-123 staggering
    T const fE[3][3] = {{T(0.5), T(0), T(0)}, {T(0), T(0.5), T(0)}, {T(0), T(0), T(0.5)}};
    T const fB[3][3] = {{T(0), T(0.5), T(0.5)}, {T(0.5), T(0), T(0.5)}, {T(0.5), T(0.5), T(0)}};
    V3<T> const e{gather1(st, st.E[0], S.supp, cx, cy, cz, pos, fE[0]), gather1(st, st.E[1], S.supp, cx, cy, cz, pos, fE[1]),
                  gather1(st, st.E[2], S.supp, cx, cy, cz, pos, fE[2])};
    V3<T> const b{gather1(st, st.B[0], S.supp, cx, cy, cz, pos, fB[0]), gather1(st, st.B[1], S.supp, cx, cy, cz, pos, fB[1]),
                  gather1(st, st.B[2], S.supp, cx, cy, cz, pos, fB[2])};
    V3<T> vel;
    if (S.desc.pusher == 2)
        vay(e, b, mom, vel, mass, charge, st.k);
    else if (S.desc.pusher == 3)
        higueraCary(e, b, mom, vel, mass, charge, st.k);
    else
        boris(e, b, mom, vel, mass, charge, st.k);
    for (int d = 0; d < 3; ++d) pos[d] += (vel[d] * st.k.dt) / st.k.cell[d];
    uint16_t const lc0 = S.cell[g];
    uint8_t const mask0 = S.mask[g];
    moveParticle(pos, S.cell[g], S.mask[g], step);
    int cellNow[3] = {cx + step[0], cy + step[1], cz + step[2]};
    if (reflectParticle(st, S, cellNow, pos, mom, vel, step)) {
        // the particle's cell and multiMask after both moves, from where it started
        S.cell[g] = lc0;
        S.mask[g] = mask0;
        cellAfterSteps(step, S.cell[g], S.mask[g]);
    }
    if (dampInPml(st, S, cellNow, pos, mom, w)) {
        // the deposit stage takes the velocity and charge of the damped particle
        S.w[g] = w;
        vel = velocityOf(mom, S.mass * w, st.k.mue0_eps0);
    }
    for (int d = 0; d < 3; ++d) S.pos[d][g] = pos[d];
    S.mom[0][g] = mom.x;
    S.mom[1][g] = mom.y;
    S.mom[2][g] = mom.z;
    return vel;
}

// current value -> int64 fixed point: round-to-nearest-even of v * 2^e,
// computed exactly in FP64
template <typename T>
int64_t toFixed(T v, double jscale) {
    return static_cast<int64_t>(std::nearbyint(static_cast<double>(v) * jscale));
}

// fixed-point tile of one supercell (order 0)
struct Tile {
    int lo, up, tx, ty, tz;
    std::vector<int64_t> acc;
    void reset(int l, int u) {
        lo = l;
        up = u;
        tx = SCX + l + u;
        ty = SCY + l + u;
        tz = SCZ + l + u;
        acc.assign(static_cast<size_t>(3) * tx * ty * tz, 0);
    }
};

template <typename T>
void flushTile(State<T>& st, const Tile& t, int ox, int oy, int oz) {
    double const inv = std::ldexp(1.0, -st.jexp);
    size_t const tv = static_cast<size_t>(t.tx) * t.ty * t.tz;
    for (int c = 0; c < 3; ++c)
        for (int z = 0; z < t.tz; ++z)
            for (int y = 0; y < t.ty; ++y)
                for (int x = 0; x < t.tx; ++x) {
                    int64_t const a = t.acc[c * tv + (static_cast<size_t>(z) * t.ty + y) * t.tx + x];
                    if (a == 0 || st.outside(ox - t.lo + x, oy - t.lo + y, oz - t.lo + z)) continue;
                    size_t const gi = st.idx(ox - t.lo + x, oy - t.lo + y, oz - t.lo + z);
                    st.J[c][gi] += static_cast<T>(static_cast<double>(a) * inv);
                }
}

template <typename T>
std::vector<std::vector<int>> groupsOf(State<T>& st, int species, bool byPusher) {
    std::vector<std::vector<int>> groups;
    // the device runs one launch per (support, pusher, current solver): the
    // species of a launch share one tile per supercell, so the grouping sets
    // the order of the float adds into J
    std::vector<std::array<int, 3>> keys;
    for (int s = 0; s < static_cast<int>(st.sp.size()); ++s) {
        if (species >= 0 && s != species) continue;
        if (st.sp[s].desc.pusher == 0) continue;
        std::array<int, 3> key{st.sp[s].supp, byPusher ? st.sp[s].desc.pusher : 0, st.sp[s].desc.current_solver};
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

// ParticlePush with (fused) or without the deposit.
template <typename T>
void push(State<T>& st, int species, bool fused) {
    bool const fixedPoint = fused && std::is_same<T, float>::value;
    auto groups = groupsOf(st, species, true);
    for (const auto& grp : groups) {
        int const supp = st.sp[grp[0]].supp;
        if (!fused) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int sc = 0; sc < st.nsc; ++sc) {
                int sx, sy, sz;
                scCoords(st, sc, sx, sy, sz);
                for (int s : grp) {
                    Species<T>& S = st.sp[s];
                    size_t const base = static_cast<size_t>(sc) * S.cap;
                    int step[3];
                    for (uint32_t i = 0; i < S.count[sc]; ++i)
                        pushOne(st, S, base + i, sx * SCX, sy * SCY, sz * SCZ, step);
                }
            }
            continue;
        }
        // The device's tile: the deposit margins around the supercell. The
        // deposit runs in the frame of the new cell, whose cells lie within
        // the margins of the old one (a move is less than a cell per axis).
        int const lo = depLower(supp), up = depUpper(supp);
        int const scs[3] = {SCX, SCY, SCZ};
        int stride[3];
        for (int d = 0; d < 3; ++d) stride[d] = (lo + up + scs[d] - 1) / scs[d] + 1;
        auto ph = phases(st, stride);
        for (const auto& members : ph) {
#pragma omp parallel for schedule(dynamic, 1)
            for (size_t m = 0; m < members.size(); ++m) {
                int const sc = members[m];
                int sx, sy, sz;
                scCoords(st, sc, sx, sy, sz);
                int const ox = sx * SCX, oy = sy * SCY, oz = sz * SCZ;
                Tile tile;
                tile.reset(lo, up);
                size_t const tv = static_cast<size_t>(tile.tx) * tile.ty * tile.tz;
                double const jscale = std::ldexp(1.0, st.jexp);
                bool any = false;
                for (int s : grp) {
                    Species<T>& S = st.sp[s];
                    size_t const base = static_cast<size_t>(sc) * S.cap;
                    for (uint32_t i = 0; i < S.count[sc]; ++i) {
                        any = true;
                        size_t const g = base + i;
                        int const lc0 = S.cell[g];
                        int step[3];
                        V3<T> const vel = pushOne(st, S, g, ox, oy, oz, step);
                        // new cell relative to the supercell origin, in tile coordinates
                        int const nx = lc0 % SCX + step[0] + lo;
                        int const ny = (lc0 / SCX) % SCY + step[1] + lo;
                        int const nz = lc0 / (SCX * SCY) + step[2] + lo;
                        T const pos[3] = {S.pos[0][g], S.pos[1][g], S.pos[2][g]};
                        T const charge = S.charge * S.w[g];
                        // a particle that left through an open border is absorbed by the
                        // boundary before the reference's deposit stage sees it
                        if (st.outside(ox - lo + nx, oy - lo + ny, oz - lo + nz)) continue;
                        if (fixedPoint) {
                            depositParticle(S.desc.current_solver, supp, pos, vel, charge, st.k,
                                            [&](int c, int x, int y, int z, T v) {
                                int64_t const q = toFixed(v, jscale);
                                int const tx = nx + x, ty = ny + y, tz = nz + z;
                                if (tx < 0 || ty < 0 || tz < 0 || tx >= tile.tx || ty >= tile.ty || tz >= tile.tz) {
                                    std::fprintf(stderr, "oracle: deposit outside the supercell tile\n");
                                    std::abort();
                                }
                                tile.acc[c * tv + (static_cast<size_t>(tz) * tile.ty + ty) * tile.tx + tx] += q;
                            });
                        } else {
                            // FP64 twin: plain sums straight into J (each supercell of a phase is disjoint)
                            depositParticle(S.desc.current_solver, supp, pos, vel, charge, st.k,
                                            [&](int c, int x, int y, int z, T v) {
                                if (st.outside(ox - lo + nx + x, oy - lo + ny + y, oz - lo + nz + z)) return;
                                st.J[c][st.idx(ox - lo + nx + x, oy - lo + ny + y, oz - lo + nz + z)] += v;
                            });
                        }
                    }
                }
                if (fixedPoint && any) flushTile(st, tile, ox, oy, oz);
            }
        }
    }
}

// CurrentDeposition from the present particle state (KernelComputeCurrent,
// FieldJ.kernel:51-107, ComputePerFrame :109-142).
// order 0: fixed-point tiles in the device's checkerboard; otherwise the
// reference's CPU order: float cache per supercell, particles in reverse frame
// order (ForEach.hpp:314-324), strided flush (Deposit.hpp:63-84).
template <typename T>
void deposit(State<T>& st, int species) {
    bool const fixedPoint = st.order == 0 && std::is_same<T, float>::value;
    auto groups = groupsOf(st, species, false);
    for (const auto& grp : groups) {
        int const supp = st.sp[grp[0]].supp;
        int const lo = depLower(supp), up = depUpper(supp);
        int const scs[3] = {SCX, SCY, SCZ};
        int stride[3];
        if (fixedPoint) {
            for (int d = 0; d < 3; ++d) stride[d] = (lo + up + scs[d] - 1) / scs[d] + 1;
        } else {
            int const skip = (lo + up + SCZ - 1) / SCZ;  // min supercell edge is 4
            for (int d = 0; d < 3; ++d) stride[d] = skip + 1;
        }
        auto ph = phases(st, stride);
        for (const auto& members : ph) {
#pragma omp parallel for schedule(dynamic, 1)
            for (size_t m = 0; m < members.size(); ++m) {
                int const sc = members[m];
                int sx, sy, sz;
                scCoords(st, sc, sx, sy, sz);
                int const ox = sx * SCX, oy = sy * SCY, oz = sz * SCZ;
                int const tx = SCX + lo + up, ty = SCY + lo + up, tz = SCZ + lo + up;
                size_t const tv = static_cast<size_t>(tx) * ty * tz;
                std::vector<int64_t> acc;
                std::vector<T> cache;
                if (fixedPoint)
                    acc.assign(3 * tv, 0);
                else
                    cache.assign(3 * tv, T(0));
                double const jscale = std::ldexp(1.0, st.jexp);
                bool any = false;
                for (int s : grp) {
                    Species<T>& S = st.sp[s];
                    size_t const base = static_cast<size_t>(sc) * S.cap;
                    uint32_t const n = S.count[sc];
                    if (n == 0) continue;
                    any = true;
                    // reverse frame order, ascending slots inside a frame
                    uint32_t const nFrames = (n + SCV - 1) / SCV;
                    for (int f = static_cast<int>(nFrames) - 1; f >= 0; --f) {
                        for (uint32_t l = 0; l < SCV; ++l) {
                            uint32_t const i = static_cast<uint32_t>(f) * SCV + l;
                            if (i >= n) break;
                            size_t const g = base + i;
                            int const lc = S.cell[g];
                            int const cx = lc % SCX + lo, cy = (lc / SCX) % SCY + lo, cz = lc / (SCX * SCY) + lo;
                            T const pos[3] = {S.pos[0][g], S.pos[1][g], S.pos[2][g]};
                            T const w = S.w[g];
                            V3<T> const mom{S.mom[0][g], S.mom[1][g], S.mom[2][g]};
                            V3<T> const vel = velocityOf(mom, S.mass * w, st.k.mue0_eps0);
                            T const charge = S.charge * w;
                            depositParticle(S.desc.current_solver, supp, pos, vel, charge, st.k,
                                            [&](int c, int x, int y, int z, T v) {
                                size_t const ti = c * tv + (static_cast<size_t>(cz + z) * ty + (cy + y)) * tx + (cx + x);
                                if (fixedPoint)
                                    acc[ti] += toFixed(v, jscale);
                                else
                                    cache[ti] += v;
                            });
                        }
                    }
                }
                if (!any) continue;
                if (fixedPoint) {
                    Tile t;
                    t.lo = lo;
                    t.up = up;
                    t.tx = tx;
                    t.ty = ty;
                    t.tz = tz;
                    t.acc.swap(acc);
                    flushTile(st, t, ox, oy, oz);
                } else {
                    for (int c = 0; c < 3; ++c)
                        for (int z = 0; z < tz; ++z)
                            for (int y = 0; y < ty; ++y)
                                for (int x = 0; x < tx; ++x) {
                                    if (st.outside(ox - lo + x, oy - lo + y, oz - lo + z)) continue;
                                    st.J[c][st.idx(ox - lo + x, oy - lo + y, oz - lo + z)] +=
                                        cache[c * tv + (static_cast<size_t>(z) * ty + y) * tx + x];
                                }
                }
            }
        }
    }
}

// Canonical supercell shift (same order as the device):
// 1. per supercell, the records of its leavers, grouped by exchange code and
//    ascending in slot inside a code, go to its outbox;
// 2. every supercell is compacted on its own: with nn = count - L (L leavers),
//    the r-th leaver slot below nn (ascending) receives the r-th remaining
//    particle at or above nn (ascending); count becomes nn;
// 3. every supercell appends, for codes 1..26 in order, the outbox records of
//    the neighbour that sends towards it with that code.
template <typename T>
void shift(State<T>& st, int species) {
    struct Rec {
        T pos[3], mom[3], w;
        uint16_t cell;
        uint32_t id;
    };
    for (int s = 0; s < static_cast<int>(st.sp.size()); ++s) {
        if (species >= 0 && s != species) continue;
        Species<T>& S = st.sp[s];
        if (S.desc.pusher == 0) continue;
        // per supercell: leaver records grouped by code (offsets off[sc][code .. code+1])
        std::vector<std::vector<Rec>> out(st.nsc);
        std::vector<std::array<uint32_t, 28>> off(st.nsc);
#pragma omp parallel for schedule(static)
        for (int sc = 0; sc < st.nsc; ++sc) {
            size_t const base = static_cast<size_t>(sc) * S.cap;
            uint32_t const n = S.count[sc];
            std::array<uint32_t, 28>& o = off[sc];
            o.fill(0);
            std::vector<uint32_t> leavers;
            for (uint32_t i = 0; i < n; ++i)
                if (S.mask[base + i] >= 2) {
                    ++o[S.mask[base + i]];  // count of code m-1 at o[m]
                    leavers.push_back(i);
                }
            uint32_t const L = static_cast<uint32_t>(leavers.size());
            for (int code = 1; code < 28; ++code) o[code] += o[code - 1];  // o[code] = start of code
            out[sc].resize(L);
            std::array<uint32_t, 28> run = o;
            for (uint32_t i : leavers) {
                size_t const b = base + i;
                Rec& r = out[sc][run[S.mask[b] - 1]++];
                for (int c = 0; c < 3; ++c) {
                    r.pos[c] = S.pos[c][b];
                    r.mom[c] = S.mom[c][b];
                }
                r.w = S.w[b];
                r.cell = S.cell[b];
                r.id = S.id[b];
            }
            if (L == 0) continue;
            uint32_t const nn = n - L;
            size_t q = 0;
            for (uint32_t i = nn; i < n; ++i) {
                if (S.mask[base + i] != 1) continue;
                size_t const a = base + leavers[q++], b = base + i;
                for (int c = 0; c < 3; ++c) {
                    S.pos[c][a] = S.pos[c][b];
                    S.mom[c][a] = S.mom[c][b];
                }
                S.w[a] = S.w[b];
                S.cell[a] = S.cell[b];
                S.id[a] = S.id[b];
                S.mask[a] = 1;
            }
            for (uint32_t i = nn; i < n; ++i) S.mask[base + i] = 0;
            S.count[sc] = nn;
        }
        bool overflow = false;
#pragma omp parallel for schedule(static) reduction(|| : overflow)
        for (int d = 0; d < st.nsc; ++d) {
            int dx, dy, dz;
            scCoords(st, d, dx, dy, dz);
            size_t const dbase = static_cast<size_t>(d) * S.cap;
            uint32_t run = S.count[d];
            for (int code = 1; code < 27; ++code) {
                int const ex = code % 3, ey = (code / 3) % 3, ez = code / 9;
                int const tx = ex == 2 ? -1 : ex, ty = ey == 2 ? -1 : ey, tz = ez == 2 ? -1 : ez;
                // leavers crossing an open border are absorbed (boundary/Absorbing.hpp:63-92)
                if ((st.cfg.boundary[0] && (dx - tx < 0 || dx - tx >= st.nscx)) ||
                    (st.cfg.boundary[1] && (dy - ty < 0 || dy - ty >= st.nscy)) ||
                    (st.cfg.boundary[2] && (dz - tz < 0 || dz - tz >= st.nscz)))
                    continue;
                int const sx = ((dx - tx) % st.nscx + st.nscx) % st.nscx;
                int const sy = ((dy - ty) % st.nscy + st.nscy) % st.nscy;
                int const sz = ((dz - tz) % st.nscz + st.nscz) % st.nscz;
                int const src = sx + st.nscx * (sy + st.nscy * sz);
                for (uint32_t q = off[src][code]; q < off[src][code + 1]; ++q) {
                    const Rec& r = out[src][q];
                    if (run >= static_cast<uint32_t>(S.cap)) {
                        overflow = true;
                        ++run;
                        continue;
                    }
                    size_t const a = dbase + run;
                    for (int c = 0; c < 3; ++c) {
                        S.pos[c][a] = r.pos[c];
                        S.mom[c][a] = r.mom[c];
                    }
                    S.w[a] = r.w;
                    S.cell[a] = r.cell;
                    S.id[a] = r.id;
                    S.mask[a] = 1;
                    ++run;
                }
            }
            S.count[d] = std::min<uint32_t>(run, static_cast<uint32_t>(S.cap));
        }
        if (overflow) st.capacityError = true;
    }
}

// ---- FDTD: FDTDBase.kernel:50-124, Curl.hpp:89, Forward/BackwardDerivative.hpp,
// Lehe/Derivative.hpp:65-228, AddCurrentDensity + None.hpp:56-65
template <typename T>
T fwd(const State<T>& st, const std::vector<T>& f, int d, int x, int y, int z) {
    int const ux = d == 0, uy = d == 1, uz = d == 2;
    return (st.at(f, x + ux, y + uy, z + uz) - st.at(f, x, y, z)) / st.k.cell[d];
}

template <typename T>
T derivE(const State<T>& st, const std::vector<T>& f, int d, int x, int y, int z) {
    if (st.cfg.solver != 1) return fwd(st, f, d, x, y, z);
    int const L = st.cfg.lehe_dir;
    if (d == L) {
        int const d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        int const u1[3] = {d1 == 0, d1 == 1, d1 == 2};
        int const u2[3] = {d2 == 0, d2 == 1, d2 == 2};
        int const u0[3] = {d == 0, d == 1, d == 2};
        T r = st.k.lehe_alpha * fwd(st, f, d, x, y, z);
        r = r + st.k.lehe_beta1 * fwd(st, f, d, x + u1[0], y + u1[1], z + u1[2]);
        r = r + st.k.lehe_beta1 * fwd(st, f, d, x - u1[0], y - u1[1], z - u1[2]);
        r = r + st.k.lehe_beta2 * fwd(st, f, d, x + u2[0], y + u2[1], z + u2[2]);
        r = r + st.k.lehe_beta2 * fwd(st, f, d, x - u2[0], y - u2[1], z - u2[2]);
        r = r + st.k.lehe_delta *
                    (st.at(f, x + 2 * u0[0], y + 2 * u0[1], z + 2 * u0[2]) - st.at(f, x - u0[0], y - u0[1], z - u0[2])) /
                    st.k.cell[d];
        return r;
    }
    int const ul[3] = {L == 0, L == 1, L == 2};
    T const beta = T(0.125);
    T const alpha = T(1.0) - T(2.0) * beta;
    T r = alpha * fwd(st, f, d, x, y, z);
    r = r + beta * fwd(st, f, d, x + ul[0], y + ul[1], z + ul[2]);
    r = r + beta * fwd(st, f, d, x - ul[0], y - ul[1], z - ul[2]);
    return r;
}

template <typename T>
T bwd(const State<T>& st, const std::vector<T>& f, int d, int x, int y, int z) {
    int const ux = d == 0, uy = d == 1, uz = d == 2;
    return (st.at(f, x, y, z) - st.at(f, x - ux, y - uy, z - uz)) / st.k.cell[d];
}

// coefficients b, c, kappa of a cell from a PML table (pmlTables layout)
template <typename T>
void pmlCell(const State<T>& st, const std::vector<float>& tab, int x, int y, int z, T* b, T* c, T* kappa) {
    int const total = st.nx + st.ny + st.nz;
    int const at[3] = {x, st.nx + y, st.nx + st.ny + z};
    for (int d = 0; d < 3; ++d) {
        b[d] = static_cast<T>(tab[at[d]]);
        c[d] = static_cast<T>(tab[total + at[d]]);
        kappa[d] = static_cast<T>(tab[2 * total + at[d]]);
    }
}

// UpdateBHalfFunctor (FDTDBase.kernel:93-124) or, with the PML, the one of
// Pml.kernel:487-590 inside the layer (isInPML: b_x b_y b_z != 1): psi of B
// advances in the first half step of update_afterCurrent (updatePsiB) and
// stays in the second, the curl's derivatives are divided by kappa
template <typename T>
void bHalf(State<T>& st, bool updatePsiB) {
    std::vector<T> nb[3];
    for (int c = 0; c < 3; ++c) nb[c] = st.B[c];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z)
        for (int y = 0; y < st.ny; ++y)
            for (int x = 0; x < st.nx; ++x) {
                size_t const i = st.idx(x, y, z);
                if (st.pml) {
                    T b[3], c[3], kappa[3];
                    pmlCell(st, st.pmlB, x, y, z, b, c, kappa);
                    if (b[0] * b[1] * b[2] != T(1.0)) {
                        T const dEdxY = derivE(st, st.E[1], 0, x, y, z), dEdxZ = derivE(st, st.E[2], 0, x, y, z);
                        T const dEdyX = derivE(st, st.E[0], 1, x, y, z), dEdyZ = derivE(st, st.E[2], 1, x, y, z);
                        T const dEdzX = derivE(st, st.E[0], 2, x, y, z), dEdzY = derivE(st, st.E[1], 2, x, y, z);
                        T* const psi[6] = {&st.psiB[0][i], &st.psiB[1][i], &st.psiB[2][i],
                                           &st.psiB[3][i], &st.psiB[4][i], &st.psiB[5][i]};
                        if (updatePsiB) {
                            *psi[0] = b[0] * *psi[0] + c[0] * dEdxZ;  // yx
                            *psi[1] = b[0] * *psi[1] + c[0] * dEdxY;  // zx
                            *psi[2] = b[1] * *psi[2] + c[1] * dEdyZ;  // xy
                            *psi[3] = b[1] * *psi[3] + c[1] * dEdyX;  // zy
                            *psi[4] = b[2] * *psi[4] + c[2] * dEdzY;  // xz
                            *psi[5] = b[2] * *psi[5] + c[2] * dEdzX;  // yz
                        }
                        T const halfDt = T(0.5) * st.k.dt;
                        nb[0][i] = st.B[0][i] + halfDt * (dEdzY / kappa[2] - dEdyZ / kappa[1] + *psi[4] - *psi[2]);
                        nb[1][i] = st.B[1][i] + halfDt * (dEdxZ / kappa[0] - dEdzX / kappa[2] + *psi[0] - *psi[5]);
                        nb[2][i] = st.B[2][i] + halfDt * (dEdyX / kappa[1] - dEdxY / kappa[0] + *psi[3] - *psi[1]);
                        continue;
                    }
                }
                T const cx = derivE(st, st.E[2], 1, x, y, z) - derivE(st, st.E[1], 2, x, y, z);
                T const cy = derivE(st, st.E[0], 2, x, y, z) - derivE(st, st.E[2], 0, x, y, z);
                T const cz = derivE(st, st.E[1], 0, x, y, z) - derivE(st, st.E[0], 1, x, y, z);
                nb[0][i] = st.B[0][i] - cx * T(0.5) * st.k.dt;
                nb[1][i] = st.B[1][i] - cy * T(0.5) * st.k.dt;
                nb[2][i] = st.B[2][i] - cz * T(0.5) * st.k.dt;
            }
    for (int c = 0; c < 3; ++c) st.B[c].swap(nb[c]);
}

// UpdateEFunctor (FDTDBase.kernel:50-91) or, with the PML, the one of
// Pml.kernel:385-485 inside the layer: psi of E advances, the curl's
// derivatives are divided by kappa
template <typename T>
void eFull(State<T>& st) {
    T const c2 = st.k.c * st.k.c;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z)
        for (int y = 0; y < st.ny; ++y)
            for (int x = 0; x < st.nx; ++x) {
                size_t const i = st.idx(x, y, z);
                if (st.pml) {
                    T b[3], c[3], kappa[3];
                    pmlCell(st, st.pmlE, x, y, z, b, c, kappa);
                    if (b[0] * b[1] * b[2] != T(1.0)) {
                        T const dBdxY = bwd(st, st.B[1], 0, x, y, z), dBdxZ = bwd(st, st.B[2], 0, x, y, z);
                        T const dBdyX = bwd(st, st.B[0], 1, x, y, z), dBdyZ = bwd(st, st.B[2], 1, x, y, z);
                        T const dBdzX = bwd(st, st.B[0], 2, x, y, z), dBdzY = bwd(st, st.B[1], 2, x, y, z);
                        T& yx = st.psiE[0][i];
                        T& zx = st.psiE[1][i];
                        T& xy = st.psiE[2][i];
                        T& zy = st.psiE[3][i];
                        T& xz = st.psiE[4][i];
                        T& yz = st.psiE[5][i];
                        yx = b[0] * yx + c[0] * dBdxZ;
                        zx = b[0] * zx + c[0] * dBdxY;
                        xy = b[1] * xy + c[1] * dBdyZ;
                        zy = b[1] * zy + c[1] * dBdyX;
                        xz = b[2] * xz + c[2] * dBdzY;
                        yz = b[2] * yz + c[2] * dBdzX;
                        T const c2dt = c2 * st.k.dt;
                        st.E[0][i] = st.E[0][i] + c2dt * (dBdyZ / kappa[1] - dBdzY / kappa[2] + xy - xz);
                        st.E[1][i] = st.E[1][i] + c2dt * (dBdzX / kappa[2] - dBdxZ / kappa[0] + yz - yx);
                        st.E[2][i] = st.E[2][i] + c2dt * (dBdxY / kappa[0] - dBdyX / kappa[1] + zx - zy);
                        continue;
                    }
                }
                T const cx = bwd(st, st.B[2], 1, x, y, z) - bwd(st, st.B[1], 2, x, y, z);
                T const cy = bwd(st, st.B[0], 2, x, y, z) - bwd(st, st.B[2], 0, x, y, z);
                T const cz = bwd(st, st.B[1], 0, x, y, z) - bwd(st, st.B[0], 1, x, y, z);
                st.E[0][i] = st.E[0][i] + cx * c2 * st.k.dt;
                st.E[1][i] = st.E[1][i] + cy * c2 * st.k.dt;
                st.E[2][i] = st.E[2][i] + cz * c2 * st.k.dt;
            }
}

// E += coeff * J (None.hpp:56-65), or with the 3D binomial filter
// (Binomial.hpp:40-104): corner, edge, face sums in the reference's order
// with weights 1, 2, 4, the centre with 8, times 1/64, on a periodic box
template <typename T>
void addCurrent(State<T>& st) {
    if (st.cfg.current_interpolation != 1) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < st.ncells; ++i)
            for (int c = 0; c < 3; ++c) st.E[c][i] = st.E[c][i] + st.k.coeff * st.J[c][i];
        return;
    }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z)
        for (int y = 0; y < st.ny; ++y)
            for (int x = 0; x < st.nx; ++x)
                for (int c = 0; c < 3; ++c) {
                    auto J = [&](int dx, int dy, int dz) { return st.J[c][st.idx(x + dx, y + dy, z + dz)]; };
                    T const corner = J(-1, -1, -1) + J(+1, -1, -1) + J(-1, +1, -1) + J(+1, +1, -1) + J(-1, -1, +1) +
                                     J(+1, -1, +1) + J(-1, +1, +1) + J(+1, +1, +1);
                    T const edge = J(-1, -1, 0) + J(+1, -1, 0) + J(-1, +1, 0) + J(+1, +1, 0) + J(-1, 0, -1) +
                                   J(+1, 0, -1) + J(-1, 0, +1) + J(+1, 0, +1) + J(0, -1, -1) + J(0, +1, -1) +
                                   J(0, -1, +1) + J(0, +1, +1);
                    T const face = J(-1, 0, 0) + J(+1, 0, 0) + J(0, -1, 0) + J(0, +1, 0) + J(0, 0, -1) + J(0, 0, +1);
                    T avg = T(1.0) * corner + T(2.0) * edge + T(4.0) * face + T(8.0) * J(0, 0, 0);
                    avg = avg * (T(1.0) / (T(8.0) + T(6.0) * T(4.0) + T(12.0) * T(2.0) + T(8.0) * T(1.0)));
                    size_t const i = st.idx(x, y, z);
                    st.E[c][i] = st.E[c][i] + st.k.coeff * avg;
                }
}

// Exponential absorber (absorber/exponential/Exponential.hpp:52-95,
// Exponential.kernel:63-112): for the plane exchanges 1, 2, 3, 6, 9, 18 in
// this order (RIGHT +x, LEFT -x, BOTTOM +y, TOP -y, BACK +z, FRONT -z,
// pmacc/type/Exchange.hpp:48-53; an odd code is the max side,
// pmacc/boundary/Utility.hpp isMaxSide), on open axes, the cell at distance
// d from that border (d = 0 outermost) is multiplied by
// exp(-strength * (T-1-d)) for T-1-d > 0. Thickness and strength are read
// at [axis][exchange % 2]: entry 0 on the negative side, entry 1 on the
// positive one, fieldAbsorber.param's {negative, positive}.
template <typename T>
void absorb(State<T>& st, std::vector<T>* F) {
    if (st.cfg.absorber != 1) return;
    int const n[3] = {st.nx, st.ny, st.nz};
    int const exchanges[6] = {1, 2, 3, 6, 9, 18};
    for (int e : exchanges) {
        int const axis = e >= 9 ? 2 : (e >= 3 ? 1 : 0);
        bool const minSide = e % 2 == 0;
        int const th = st.cfg.absorber_cells[axis][e % 2];
        T const strength = static_cast<T>(st.cfg.absorber_strength[axis][e % 2]);
        if (!st.cfg.boundary[axis] || th <= 1) continue;
#pragma omp parallel for schedule(static)
        for (int z = 0; z < st.nz; ++z)
            for (int y = 0; y < st.ny; ++y)
                for (int x = 0; x < st.nx; ++x) {
                    int const c[3] = {x, y, z};
                    int const dist = minSide ? c[axis] : n[axis] - 1 - c[axis];
                    int const factor = th - 1 - dist;
                    if (factor <= 0) continue;
                    T const a = std::exp(-strength * static_cast<T>(factor));
                    size_t const i = st.idx(x, y, z);
                    for (int k = 0; k < 3; ++k) F[k][i] = F[k][i] * a;
                }
    }
}

// ---- incident field through the Huygens surface (fields/incidentField) ----
// exp, cos and atan of the FP32 functors are taken as the FP64 function
// rounded to T (the reference calls the platform's float_X functions, which
// agree with this to an ulp or two).
template <typename T>
T incExp(T x) { return static_cast<T>(std::exp(static_cast<double>(x))); }
template <typename T>
T incCos(T x) { return static_cast<T>(std::cos(static_cast<double>(x))); }
template <typename T>
T incAtan(T x) { return static_cast<T>(std::atan(static_cast<double>(x))); }

// DispersionRelation<Yee> (Yee.hpp:68-119) / <Lehe> (Lehe.hpp:77-215):
// f(|k|) and df/d|k|, FP64
template <typename T>
struct Dispersion {
    int solver, d0;
    double dt, c, step[3], omega, dir[3];
    double relation(double absK) const {
        double const lhsTerm = std::sin(0.5 * omega * dt) / (c * dt);
        double const lhs = lhsTerm * lhsTerm;
        if (solver == 0) {
            double rhs = 0.0;
            for (int d = 0; d < 3; ++d) {
                double const term = std::sin(0.5 * absK * dir[d] * step[d]) / step[d];
                rhs += term * term;
            }
            return rhs - lhs;
        }
        double sq[3], st2[3];
        for (int d = 0; d < 3; ++d) {
            double const s = std::sin(0.5 * absK * dir[d] * step[d]);
            sq[d] = s * s;
            st2[d] = step[d] * step[d];
        }
        int const d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        double const ratio = step[d0] / (c * dt);
        double const co = ratio * std::sin(kPi / 2.0 * (c * dt) / step[d0]);
        double const delta = 0.25 * (1.0 - co * co);
        double const r1 = step[d0] / step[d1], r2 = step[d0] / step[d2];
        double const b0 = 0.125, b1 = 0.125 * r1 * r1, b2 = 0.125 * r2 * r2;
        double rhs = 0.0;
        for (int d = 0; d < 3; ++d) rhs += sq[d] / st2[d];
        rhs -= 4.0 * delta * sq[d0] * sq[d0] / st2[d0];
        rhs -= 4.0 * (b1 / st2[d1] + b0 / st2[d0]) * sq[d0] * sq[d1];
        rhs -= 4.0 * (b2 / st2[d2] + b0 / st2[d0]) * sq[d0] * sq[d2];
        return rhs - lhs;
    }
    double derivative(double absK) const {
        if (solver == 0) {
            double r = 0.0;
            for (int d = 0; d < 3; ++d) {
                double const arg = 0.5 * absK * dir[d] * step[d];
                double const term = std::sin(arg) / step[d];
                r += 2.0 * term * (0.5 * dir[d] * std::cos(arg));
            }
            return r;
        }
        double sq[3], sqd[3], st2[3];
        for (int d = 0; d < 3; ++d) {
            double const arg = 0.5 * absK * dir[d] * step[d];
            double const s = std::sin(arg);
            sqd[d] = 2.0 * s * (0.5 * dir[d] * step[d] * std::cos(arg));
            double const s2 = std::sin(0.5 * absK * dir[d] * step[d]);
            sq[d] = s2 * s2;
            st2[d] = step[d] * step[d];
        }
        int const d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        double const ratio = step[d0] / (c * dt);
        double const co = ratio * std::sin(kPi / 2.0 * (c * dt) / step[d0]);
        double const delta = 0.25 * (1.0 - co * co);
        double const r1 = step[d0] / step[d1];
        double const b0 = 0.125, b1 = 0.125 * r1 * r1;
        double r = 0.0;
        for (int d = 0; d < 3; ++d) r += sqd[d] / st2[d];
        r -= 4.0 * delta * sq[d0] * sqd[d0] / st2[d0];
        r -= 4.0 * (b1 / st2[d1] + b0 / st2[d0]) * (sq[d0] * sqd[d1] + sqd[d0] * sq[d1]);
        // Lehe.hpp:190-191 uses the dir1 coefficient for the dir2 term too
        r -= 4.0 * (b1 / st2[d1] + b0 / st2[d0]) * (sq[d0] * sqd[d2] + sqd[d0] * sq[d2]);
        return r;
    }
};

// One profile: BaseFunctorE (Functors.hpp:80-230) with the constants of
// BaseParamUnitless (BaseParam.hpp:40-140), GaussianPulseUnitless
// (GaussianPulse.hpp:60-75) and PlaneWaveUnitless (PlaneWave.hpp:60-75)
template <typename T>
struct IncidentFunctor {
    int kind = 0, polarisation = 0;
    T cell[3], c, dt;
    T WAVE_LENGTH, w, AMPLITUDE, PULSE_DURATION, LASER_PHASE, TIME_DELAY;
    T DIR[3], POL[3], AX2[3];
    T focus[3], origin[3];
    T phaseVelocity;
    T W0, rayleighLength, TIME_SHIFT;
    int nModes = 1;
    T modes[16], modePhases[16];
    double RAMP_INIT = 0.0;
    T LASER_NOFOCUS_CONSTANT;

    void init(const oracle_config& cfg, const Consts<T>& k, const oracle_incident_profile& p) {
        kind = p.kind;
        polarisation = p.polarisation;
        for (int d = 0; d < 3; ++d) cell[d] = k.cell[d];
        c = k.c;
        dt = k.dt;
        double const uT = k.unit_time, uL = k.unit_length;
        double const eUnit = 1.0 / (uT * uT / k.unit_mass / uL * k.unit_charge);  // simulation.unitless:461-464
        WAVE_LENGTH = static_cast<T>(p.wavelength_si / uL);
        T const f = static_cast<T>(c / WAVE_LENGTH);
        w = T(2) * static_cast<T>(kPi) * f;
        AMPLITUDE = static_cast<T>(p.amplitude_si / eUnit);
        PULSE_DURATION = static_cast<T>(p.pulse_duration_si / uT);
        LASER_PHASE = static_cast<T>(p.laser_phase);
        T const delayParam = static_cast<T>(p.time_delay_si);
        TIME_DELAY = static_cast<T>(static_cast<double>(delayParam) / uT);
        for (int d = 0; d < 3; ++d) {
            DIR[d] = static_cast<T>(p.direction[d]);
            POL[d] = static_cast<T>(p.polarisation_direction[d]);
            focus[d] = static_cast<T>(p.focus_position_si[d] / uL);
        }
        AX2[0] = DIR[1] * POL[2] - DIR[2] * POL[1];
        AX2[1] = DIR[2] * POL[0] - DIR[0] * POL[2];
        AX2[2] = DIR[0] * POL[1] - DIR[1] * POL[0];
        for (int d = 0; d < 3; ++d)  // getFocus, Functors.hpp:158-175
            if (p.focus_origin[d] == 1) focus[d] += static_cast<T>(static_cast<uint32_t>(cfg.grid[d]) / 2u) * cell[d];
        T originP = -std::numeric_limits<T>::infinity();  // getOrigin, Functors.hpp:177-205
        for (int a = 0; a < 3; ++a) {
            if (std::abs(DIR[a]) > std::numeric_limits<T>::epsilon()) {
                T const minPos = (static_cast<T>(cfg.incident_position[a][0]) + T(0.75)) * cell[a];
                int const maxIdx = cfg.incident_position[a][1] > 0 ? cfg.incident_position[a][1]
                                                                   : cfg.grid[a] + cfg.incident_position[a][1];
                T const maxPos = (static_cast<T>(maxIdx) - T(0.75)) * cell[a];
                T const axisP = std::min((minPos - focus[a]) / DIR[a], (maxPos - focus[a]) / DIR[a]);
                originP = std::max(originP, axisP);
            }
        }
        for (int d = 0; d < 3; ++d) origin[d] = focus[d] + originP * DIR[d];
        // calculatePhaseVelocity + DispersionRelationSolver (FP64 Newton)
        Dispersion<T> dr;
        dr.solver = cfg.solver;
        dr.d0 = cfg.lehe_dir;
        dr.dt = static_cast<double>(dt);
        dr.c = static_cast<double>(c);
        for (int d = 0; d < 3; ++d) {
            dr.step[d] = static_cast<double>(cell[d]);
            dr.dir[d] = static_cast<double>(DIR[d]);
        }
        dr.omega = 2.0 * kPi * static_cast<double>(c / WAVE_LENGTH);
        double absK = dr.omega * static_cast<double>(c);
        for (int it = 0; it < 100; ++it) {
            double const der = dr.derivative(absK);
            if (std::abs(der) < 1e-5) break;
            absK = absK - dr.relation(absK) / der;
        }
        phaseVelocity = static_cast<T>(dr.omega / absK / static_cast<double>(c));
        if (kind == 1) {
            W0 = static_cast<T>(p.w0_si / uL);
            rayleighLength = static_cast<T>(kPi) * W0 * W0 / WAVE_LENGTH;
            TIME_SHIFT = static_cast<T>(-0.5 * p.pulse_init * static_cast<double>(PULSE_DURATION));
            nModes = p.num_modes + 1;
            for (int m = 0; m < nModes; ++m) {
                modes[m] = static_cast<T>(p.laguerre_modes[m]);
                modePhases[m] = static_cast<T>(p.laguerre_phases[m]);
            }
        } else {
            RAMP_INIT = p.ramp_init;
            LASER_NOFOCUS_CONSTANT = static_cast<T>(p.laser_nofocus_constant_si / uT);
        }
    }

    static T dot(const T* a, const T* b) {
        T r = a[0] * b[0];
        r += a[1] * b[1];
        r += a[2] * b[2];
        return r;
    }
    // getCurrentTime (Functors.hpp:103-110)
    T currentTime(T timeOrigin, const T* x) const {
        T sh[3];
        for (int d = 0; d < 3; ++d) sh[d] = x[d] * cell[d] - origin[d];
        T const distance = dot(sh, DIR);
        T const delay = distance / phaseVelocity + TIME_DELAY;
        return timeOrigin - delay;
    }
    // simpleLaguerre (GaussianPulse.hpp:230-250)
    static T laguerre(int n, T x) {
        if (n == 0) return T(1);
        T prev = T(1), cur = T(1) - x, next = T(0);
        for (int k = 1; k < n; ++k) {
            next = ((T(2) * static_cast<T>(k) + T(1) - x) * cur - static_cast<T>(k) * prev) / static_cast<T>(k + 1);
            prev = cur;
            cur = next;
        }
        return cur;
    }
    // GaussianPulseFunctorIncidentE::getValue (GaussianPulse.hpp:130-228)
    T gaussian(T timeOrigin, const T* x, T phaseShift) const {
        T sh[3];
        for (int d = 0; d < 3; ++d) sh[d] = x[d] * cell[d] - origin[d];
        T pos[3] = {dot(sh, DIR), dot(sh, POL), dot(sh, AX2)};  // getInternalCoordinates
        T time = currentTime(timeOrigin, x);
        if (time < T(0)) return T(0);
        time += TIME_SHIFT;
        T const fo[3] = {focus[0] - origin[0], focus[1] - origin[1], focus[2] - origin[2]};
        T const focusPos = dot(fo, DIR) - pos[0];
        T const beamW = W0 * std::sqrt(T(1) + (focusPos / rayleighLength) * (focusPos / rayleighLength));
        T const phase = w * (time - focusPos / c) + LASER_PHASE + phaseShift;
        T const t0 = pos[0] * T(0), t1 = pos[1] * T(1), t2 = pos[2] * T(1);  // pos * planeNoNormal
        T rho2 = t0 * t0;
        rho2 += t1 * t1;
        rho2 += t2 * t2;
        T const Rinv = -focusPos / (rayleighLength * rayleighLength + focusPos * focusPos);
        T const xi = incAtan(-focusPos / rayleighLength);
        T const r2OverW2 = rho2 / beamW / beamW;
        T const r = T(0.5) * rho2 * Rinv;
        T const twoPi = T(2) * static_cast<T>(kPi);
        T etrans = T(0);
        for (int m = 0; m < nModes; ++m)
            etrans += modes[m] * laguerre(m, T(2) * r2OverW2) * incExp(-r2OverW2) *
                      incCos(twoPi / WAVE_LENGTH * focusPos - twoPi / WAVE_LENGTH * r +
                             (T(2) * static_cast<T>(m) + T(1)) * xi + phase + modePhases[m]);
        T const shiftedTime = time - r / c;
        T const ex = shiftedTime / (T(2) * PULSE_DURATION);  // GaussianPulseEnvelope::getEnvelope
        etrans *= incExp(-ex * ex);
        T norm = T(0);
        for (int m = 0; m < nModes; ++m) norm += modes[m];
        T env = AMPLITUDE;
        env *= W0 / beamW;
        return env * etrans / norm;
    }
    // PlaneWaveFunctorIncidentE::getLongitudinal (PlaneWave.hpp:90-125); the
    // FP64 RAMP_INIT makes the ramp arithmetic FP64
    T planeLongitudinal(T time, T phaseShift) const {
        T env = AMPLITUDE;
        double const mue = 0.5 * RAMP_INIT * static_cast<double>(PULSE_DURATION);
        T const tau = PULSE_DURATION * std::sqrt(T(2));
        double const up = mue, down = mue + static_cast<double>(LASER_NOFOCUS_CONSTANT);
        T corr = T(0);
        double const t = static_cast<double>(time);
        if (t > down) {
            double const e = (t - down) / static_cast<double>(tau);
            env = static_cast<T>(static_cast<double>(env) * std::exp(-0.5 * e * e));
            corr = static_cast<T>((t - down) / static_cast<double>(w * tau * tau));
        } else if (t < up) {
            double const e = (t - up) / static_cast<double>(tau);
            env = static_cast<T>(static_cast<double>(env) * std::exp(-0.5 * e * e));
            corr = static_cast<T>((t - up) / static_cast<double>(w * tau * tau));
        }
        double const ph = static_cast<double>(w) * (t - up) + static_cast<double>(LASER_PHASE) +
                          static_cast<double>(phaseShift);
        return static_cast<T>((std::sin(ph) + std::cos(ph) * static_cast<double>(corr)) * static_cast<double>(env));
    }
    // E_inc (GaussianPulse operator(), GaussianPulse.hpp:112-126; separable
    // functor operator(), Functors.hpp:262-283)
    void fieldE(T timeOrigin, const T* x, T* e) const {
        T const halfPi = static_cast<T>(kPi) / T(2);
        T cp1[3], cp2[3];
        T const root2 = std::sqrt(T(2));
        for (int d = 0; d < 3; ++d) cp1[d] = POL[d] / root2;
        cp2[0] = DIR[1] * cp1[2] - DIR[2] * cp1[1];
        cp2[1] = DIR[2] * cp1[0] - DIR[0] * cp1[2];
        cp2[2] = DIR[0] * cp1[1] - DIR[1] * cp1[0];
        T v1, v2;
        if (kind == 1) {
            if (polarisation == 0) {
                v1 = gaussian(timeOrigin, x, T(0));
                for (int d = 0; d < 3; ++d) e[d] = POL[d] * v1;
                return;
            }
            v1 = gaussian(timeOrigin, x, halfPi);
            v2 = gaussian(timeOrigin, x, T(0));
        } else {
            T const time = currentTime(timeOrigin, x);
            if (time < T(0)) {
                e[0] = e[1] = e[2] = T(0);
                return;
            }
            if (polarisation == 0) {
                v1 = planeLongitudinal(time, T(0)) * T(1);
                for (int d = 0; d < 3; ++d) e[d] = POL[d] * v1;
                return;
            }
            v1 = planeLongitudinal(time, halfPi) * T(1);
            v2 = planeLongitudinal(time, T(0)) * T(1);
        }
        for (int d = 0; d < 3; ++d) e[d] = cp1[d] * v1 + cp2[d] * v2;
    }
    // ApproximateIncidentB: cross(direction, E_inc) / c (Functors.hpp:540-570)
    void fieldB(T timeOrigin, const T* x, T* b) const {
        T e[3];
        fieldE(timeOrigin, x, e);
        b[0] = (DIR[1] * e[2] - DIR[2] * e[1]) / c;
        b[1] = (DIR[2] * e[0] - DIR[0] * e[2]) / c;
        b[2] = (DIR[0] * e[1] - DIR[1] * e[0]) / c;
    }
};

// DerivativeCoefficients (DerivativeCoefficients.hpp:35-200) of the curl the
// updated field's solver step uses along `axis`: the E update reads curl B
// (Yee's derivative for both solvers, Lehe.def:73), the B update curl E
// (Lehe's derivative under Lehe)
template <typename T>
void derivativeCoefficients(const State<T>& st, bool curlE, int axis, T value[2][2][2], int size[3]) {
    for (int i = 0; i < 8; ++i) (&value[0][0][0])[i] = T(0);
    size[0] = size[1] = size[2] = 1;
    if (!curlE || st.cfg.solver != 1) {
        value[0][0][0] = T(1);
        return;
    }
    int const dir = st.cfg.lehe_dir;
    const Consts<T>& k = st.k;
    if (dir == axis) {
        int const d0 = axis, d1 = (axis + 1) % 3, d2 = (axis + 2) % 3;
        double const stepRatio = static_cast<double>(k.cell[d0] / (k.c * k.dt));
        double const coeff = stepRatio * std::sin(kPi / 2.0 * static_cast<double>(k.c) * static_cast<double>(k.dt) /
                                                  static_cast<double>(k.cell[d0]));
        T const delta = static_cast<T>(0.25 * (1.0 - coeff * coeff));
        double const sr1 = static_cast<double>(k.cell[d0] / k.cell[d1]);
        double const sr2 = static_cast<double>(k.cell[d0] / k.cell[d2]);
        double const betaDir1 = 0.125 * sr1 * sr1;
        double const betaDir2 = 0.125 * sr2 * sr2;
        size[0] = size[1] = size[2] = 2;
        value[0][0][0] = static_cast<T>(1.0 - 2.0 * betaDir1 - 2.0 * betaDir2 - 3.0 * static_cast<double>(delta));
        int a[3] = {0, 0, 0};
        a[d0] = 1;
        value[a[0]][a[1]][a[2]] = delta;
        a[d0] = 0;
        a[d1] = 1;
        value[a[0]][a[1]][a[2]] = static_cast<T>(betaDir1);
        a[d1] = 0;
        a[d2] = 1;
        value[a[0]][a[1]][a[2]] = static_cast<T>(betaDir2);
    } else {
        T const beta = T(0.125);
        size[dir] = 2;
        value[0][0][0] = T(1) - T(2) * beta;
        int a[3] = {0, 0, 0};
        a[dir] = 1;
        value[a[0]][a[1]][a[2]] = beta;
    }
}

// updateField (Solver.hpp:120-258) with UpdateFunctor (Solver.kernel:60-262)
// for one face; updateB: the B half update from E_inc, else the E update from B_inc
template <typename T>
void incidentFace(State<T>& st, int axis, int side, bool updateB, T currentStep, T timeIncrement) {
    const oracle_config& cfg = st.cfg;
    const oracle_incident_profile& prof = cfg.incident[axis][side];
    if (prof.kind == 0) return;
    IncidentFunctor<T> F;
    F.init(cfg, st.k, prof);
    T const direction = side == 0 ? T(1) : T(-1);
    T const curlCoefficient = updateB ? -timeIncrement : timeIncrement * (st.k.c * st.k.c);
    bool const isTotal = !updateB;  // E_x at (0.5, 0, 0) is a total field, B_x at (0, .5, .5) is not
    int n[3] = {st.nx, st.ny, st.nz};
    int minB[3], maxB[3];
    for (int d = 0; d < 3; ++d) {
        minB[d] = cfg.incident_position[d][0];
        maxB[d] = cfg.incident_position[d][1] > 0 ? cfg.incident_position[d][1] : n[d] + cfg.incident_position[d][1];
    }
    int beginUser[3], endUser[3];
    for (int d = 0; d < 3; ++d) {
        beginUser[d] = minB[d] + 1;
        endUser[d] = maxB[d];
    }
    if (!isTotal && direction > T(0)) beginUser[axis] -= 1;
    bool lastDomain[3];
    bool const contiguous = prof.kind == 2;  // MakePeriodicTransversalHuygensSurfaceContiguous<PlaneWave>
    for (int d = 0; d < 3; ++d) {
        if (contiguous && cfg.boundary[d] == 0 && d != axis) {
            beginUser[d] = 0;
            endUser[d] = n[d];
            lastDomain[d] = false;
        } else {
            lastDomain[d] = true;
        }
    }
    T coeff[2][2][2];
    int size[3];
    derivativeCoefficients(st, updateB, axis, coeff, size);
    int const margin = size[axis];
    if (direction > T(0)) {
        beginUser[axis] -= margin - 1;
        endUser[axis] = beginUser[axis] + 2 * margin - 1;
    } else {
        endUser[axis] += margin - 1;
        beginUser[axis] = endUser[axis] - (2 * margin - 1);
    }
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(beginUser[d], 0);
        hi[d] = std::min(endUser[d], n[d]);
        if (lo[d] >= hi[d]) return;
    }
    T baseShift[3] = {T(0), T(0), T(0)};
    if (direction > T(0)) baseShift[axis] = isTotal ? T(-1) : T(1);
    // Yee positions of the incident field's components (YeeCell.hpp:69-141)
    static const double yeeE[3][3] = {{0.5, 0, 0}, {0, 0.5, 0}, {0, 0, 0.5}};
    static const double yeeB[3][3] = {{0, 0.5, 0.5}, {0.5, 0, 0.5}, {0.5, 0.5, 0}};
    const double(*ypos)[3] = updateB ? yeeE : yeeB;
    int const comp1 = (axis + 2) % 3, comp2 = (axis + 1) % 3;  // incidentComponent1 / 2
    int const dir1 = comp2, dir2 = comp1;
    T inCell1[3], inCell2[3];
    for (int d = 0; d < 3; ++d) {
        inCell1[d] = baseShift[d] + static_cast<T>(ypos[comp1][d]);
        inCell2[d] = baseShift[d] + static_cast<T>(ypos[comp2][d]);
    }
    T const sign = direction > T(0) ? T(1) : T(-1);
    T const baseCoefficient = curlCoefficient / st.k.cell[axis] * sign;
    T const timeOrigin = currentStep * st.k.dt;
    std::vector<T>* dest = updateB ? st.B : st.E;
    int const huygens = lo[axis] + margin - 1;
    int const sizeDir1 = size[dir1], sizeDir2 = size[dir2];
    auto incident = [&](const T* x, int comp) {
        T v[3];
        if (updateB)
            F.fieldE(timeOrigin, x, v);
        else
            F.fieldB(timeOrigin, x, v);
        return v[comp];
    };
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = lo[2]; z < hi[2]; ++z)
        for (int y = lo[1]; y < hi[1]; ++y)
            for (int x = lo[0]; x < hi[0]; ++x) {
                int const cellIdx[3] = {x, y, z};
                bool last[3];
                for (int d = 0; d < 3; ++d) last[d] = lastDomain[d] && cellIdx[d] == hi[d] - 1;
                int surf[3] = {x, y, z};
                surf[axis] = huygens;
                T s1[3], s2[3];
                for (int d = 0; d < 3; ++d) {
                    s1[d] = static_cast<T>(surf[d]) + inCell1[d];
                    s2[d] = static_cast<T>(surf[d]) + inCell2[d];
                }
                int const shiftAlong = cellIdx[axis] - huygens;
                bool const apply1 = (isTotal && !last[comp2]) || (!isTotal && !last[comp1]);
                bool const apply2 = (isTotal && !last[comp1]) || (!isTotal && !last[comp2]);
                T result[3] = {T(0), T(0), T(0)};
                T idxShift[3] = {T(0), T(0), T(0)};
                T initial = static_cast<T>(-shiftAlong);
                if (isTotal && static_cast<T>(shiftAlong) * direction > T(0)) initial += direction;
                if (!isTotal && static_cast<T>(shiftAlong) * direction < T(0)) initial -= direction;
                idxShift[axis] = initial;
                T const increment = shiftAlong < 0 ? T(1) : (shiftAlong > 0 ? T(-1) : (isTotal ? -direction : direction));
                for (int as = std::abs(shiftAlong); as < margin; ++as) {
                    int ci[3] = {0, 0, 0};
                    ci[axis] = as;
                    idxShift[dir1] = static_cast<T>(1 - sizeDir1);
                    for (int a1 = 1 - sizeDir1; a1 < sizeDir1; ++a1) {
                        ci[dir1] = std::abs(a1);
                        idxShift[dir2] = static_cast<T>(1 - sizeDir2);
                        for (int a2 = 1 - sizeDir2; a2 < sizeDir2; ++a2) {
                            ci[dir2] = std::abs(a2);
                            T const dc = coeff[ci[0]][ci[1]][ci[2]];
                            if (apply1) {
                                T const p[3] = {s1[0] + idxShift[0], s1[1] + idxShift[1], s1[2] + idxShift[2]};
                                result[dir1] += dc * incident(p, comp1);
                            }
                            if (apply2) {
                                T const p[3] = {s2[0] + idxShift[0], s2[1] + idxShift[1], s2[2] + idxShift[2]};
                                result[dir2] += dc * incident(p, comp2);
                            }
                            idxShift[dir2] += T(1);
                        }
                        idxShift[dir1] += T(1);
                    }
                    idxShift[axis] += increment;
                }
                result[dir1] *= baseCoefficient;
                result[dir2] *= -baseCoefficient;
                size_t const at = st.idx(x, y, z);
                for (int d = 0; d < 3; ++d) dest[d][at] += result[d];
            }
}

// Solver::updateBHalf / updateE (Solver.hpp:270-330): faces x, y, z, the
// negative border's profile before the positive one's
template <typename T>
void incidentUpdate(State<T>& st, bool updateB, T currentStep) {
    T const timeIncrement = updateB ? T(0.5) * st.k.dt : st.k.dt;
    for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) incidentFace(st, axis, side, updateB, currentStep, timeIncrement);
}

// update_beforeCurrent (FDTDBase.hpp:97-121)
template <typename T>
void beforeCurrent(State<T>& st) {
    bHalf(st, false);  // updateBSecondHalf
    T const cs = static_cast<T>(st.step);
    incidentUpdate(st, true, cs);
    incidentUpdate(st, false, cs + T(0.5) * st.k.dt / st.k.dt);
    eFull(st);
}

// update_afterCurrent (FDTDBase.hpp:151-183): absorber on E, incident B,
// B half step, absorber on B; the step counter then advances
template <typename T>
void afterCurrent(State<T>& st) {
    absorb(st, st.E);
    incidentUpdate(st, true, static_cast<T>(st.step) + st.k.dt / st.k.dt);
    bHalf(st, true);  // updateBFirstHalf
    absorb(st, st.B);
    ++st.step;
}

template <typename T>
void step(State<T>& st, int n) {
    for (int it = 0; it < n; ++it) {
        for (int c = 0; c < 3; ++c) std::fill(st.J[c].begin(), st.J[c].end(), T(0));  // CurrentReset
        bool const fused = st.order == 0 && std::is_same<T, float>::value;
        push(st, -1, fused);  // ParticlePush
        shift(st, -1);
        beforeCurrent(st);  // update_beforeCurrent
        if (!fused) deposit(st, -1);  // CurrentDeposition
        addCurrent(st);  // CurrentInterpolationAndAdditionToEMF
        afterCurrent(st);  // update_afterCurrent
    }
}

// charge density at the cell origins from the B-spline assignment of every
// species (immobile ones included), FP64 accumulation
template <typename T>
void chargeDensity(const State<T>& st, double* rho) {
    std::fill(rho, rho + st.ncells, 0.0);
    for (const Species<T>& S : st.sp) {
        ShapeFns<T> sh{S.supp};
        int const begin = sh.begin();
        for (int sc = 0; sc < st.nsc; ++sc) {
            int sx, sy, sz;
            scCoords(st, sc, sx, sy, sz);
            size_t const base = static_cast<size_t>(sc) * S.cap;
            for (uint32_t i = 0; i < S.count[sc]; ++i) {
                size_t const g = base + i;
                int const lc = S.cell[g];
                int const c[3] = {sx * SCX + lc % SCX, sy * SCY + (lc / SCX) % SCY, sz * SCZ + lc / (SCX * SCY)};
                double const q = static_cast<double>(S.charge) * static_cast<double>(S.w[g]) *
                                 static_cast<double>(st.k.inv_vdt) * static_cast<double>(st.k.dt);
                double wgt[3][5];
                for (int d = 0; d < 3; ++d)
                    for (int o = 0; o <= S.supp; ++o)
                        wgt[d][o] = static_cast<double>(sh.assign(T(begin + o) - S.pos[d][g]));
                for (int oz = 0; oz <= S.supp; ++oz)
                    for (int oy = 0; oy <= S.supp; ++oy)
                        for (int ox = 0; ox <= S.supp; ++ox) {
                            double const v = q * wgt[0][ox] * wgt[1][oy] * wgt[2][oz];
                            if (v != 0.0 && !st.outside(c[0] + begin + ox, c[1] + begin + oy, c[2] + begin + oz))
                                rho[st.idx(c[0] + begin + ox, c[1] + begin + oy, c[2] + begin + oz)] += v;
                        }
            }
        }
    }
}

template <typename T>
State<T>* create(const oracle_config& cfg, int order) {
    State<T>* st = new State<T>();
    st->cfg = cfg;
    st->k = makeConsts<T>(cfg);
    st->nx = cfg.grid[0];
    st->ny = cfg.grid[1];
    st->nz = cfg.grid[2];
    st->nscx = st->nx / SCX;
    st->nscy = st->ny / SCY;
    st->nscz = st->nz / SCZ;
    st->nsc = st->nscx * st->nscy * st->nscz;
    st->ncells = static_cast<size_t>(st->nx) * st->ny * st->nz;
    st->order = order;
    st->k.contract = order == 0;
    if (cfg.absorber == 2) {
        st->pml = true;
        Consts<float> const kf = makeConsts<float>(cfg);
        pmlTables(cfg, kf, false, st->pmlE);
        pmlTables(cfg, kf, true, st->pmlB);
        pmlDampSetup(cfg, kf, *st);
        size_t const n = static_cast<size_t>(cfg.grid[0]) * cfg.grid[1] * cfg.grid[2];
        for (int q = 0; q < 6; ++q) {
            st->psiE[q].assign(n, T(0));
            st->psiB[q].assign(n, T(0));
        }
    }
    st->anyOpen = cfg.boundary[0] || cfg.boundary[1] || cfg.boundary[2];
    for (int c = 0; c < 3; ++c) {
        st->E[c].assign(st->ncells, T(0));
        st->B[c].assign(st->ncells, T(0));
        st->J[c].assign(st->ncells, T(0));
    }
    for (int s = 0; s < cfg.n_species; ++s) {
        Species<T> S;
        S.desc = cfg.species[s];
        S.supp = supportOf(S.desc.shape);
        S.cap = S.desc.frames_per_sc * SCV;
        S.mass = st->k.mass[s];
        S.charge = st->k.charge[s];
        size_t const slots = static_cast<size_t>(st->nsc) * S.cap;
        for (int c = 0; c < 3; ++c) {
            S.pos[c].assign(slots, T(0));
            S.mom[c].assign(slots, T(0));
        }
        S.w.assign(slots, T(0));
        S.cell.assign(slots, 0);
        S.mask.assign(slots, 0);
        S.id.assign(slots, 0);
        S.count.assign(st->nsc, 0);
        st->sp.push_back(std::move(S));
    }
    return st;
}

struct Handle {
    int precision;
    void* st;
};

template <typename R, typename F32, typename F64>
R dispatch(Handle* h, F32 f32, F64 f64) {
    if (h->precision == 64) return f64(static_cast<State<double>*>(h->st));
    return f32(static_cast<State<float>*>(h->st));
}

}  // namespace

extern "C" {

void* oracle_create(const oracle_config* cfg, int32_t precision, int32_t order) {
    Handle* h = new Handle();
    h->precision = precision == 64 ? 64 : 32;
    if (h->precision == 64)
        h->st = create<double>(*cfg, order);
    else
        h->st = create<float>(*cfg, order);
    return h;
}

void oracle_destroy(void* p) {
    Handle* h = static_cast<Handle*>(p);
    if (!h) return;
    if (h->precision == 64)
        delete static_cast<State<double>*>(h->st);
    else
        delete static_cast<State<float>*>(h->st);
    delete h;
}

int32_t oracle_get_constants(const oracle_config* cfg, oracle_constants* out) {
    Consts<float> k = makeConsts<float>(*cfg);
    std::memset(out, 0, sizeof(*out));
    for (int d = 0; d < 3; ++d) out->cell[d] = k.cell[d];
    out->dt = k.dt;
    out->c = k.c;
    out->eps0 = k.eps0;
    out->mue0 = k.mue0;
    out->mue0_eps0 = k.mue0_eps0;
    out->add_current_coeff = k.coeff;
    out->inv_cell_volume_dt = k.inv_vdt;
    out->lehe_alpha = k.lehe_alpha;
    out->lehe_delta = k.lehe_delta;
    for (int s = 0; s < cfg->n_species && s < 4; ++s) {
        out->species_mass[s] = k.mass[s];
        out->species_charge[s] = k.charge[s];
    }
    out->unit_length = k.unit_length;
    out->unit_time = k.unit_time;
    out->unit_mass = k.unit_mass;
    out->unit_charge = k.unit_charge;
    out->typical_weighting = k.wtyp;
    return 0;
}

void oracle_set_jexp(void* p, int32_t e) {
    dispatch<void>(static_cast<Handle*>(p), [&](State<float>* s) { s->jexp = e; },
                   [&](State<double>* s) { s->jexp = e; });
}

// The incident field of face (axis, side) at n total-cell positions pos[3n]
// and source step `current_step` (time current_step * dt), FP64 functor:
// which = 0 E_inc, 1 B_inc (ApproximateIncidentB); out[3n]
void oracle_incident_field(const oracle_config* cfg, int32_t axis, int32_t side, int32_t which, double current_step,
                           const double* pos, int64_t n, double* out) {
    Consts<double> k = makeConsts<double>(*cfg);
    IncidentFunctor<double> F;
    F.init(*cfg, k, cfg->incident[axis][side]);
    double const t0 = current_step * k.dt;
    for (int64_t i = 0; i < n; ++i) {
        if (which == 0)
            F.fieldE(t0, pos + 3 * i, out + 3 * i);
        else
            F.fieldB(t0, pos + 3 * i, out + 3 * i);
    }
}

// phase velocity (units of c) and origin (PIC units) of face (axis, side), FP64
void oracle_incident_info(const oracle_config* cfg, int32_t axis, int32_t side, double* out) {
    Consts<double> k = makeConsts<double>(*cfg);
    IncidentFunctor<double> F;
    F.init(*cfg, k, cfg->incident[axis][side]);
    out[0] = F.phaseVelocity;
    for (int d = 0; d < 3; ++d) out[1 + d] = F.origin[d];
    for (int d = 0; d < 3; ++d) out[4 + d] = F.focus[d];
}

void oracle_set_step(void* p, int64_t step) {
    dispatch<void>(static_cast<Handle*>(p), [&](State<float>* s) { s->step = step; },
                   [&](State<double>* s) { s->step = step; });
}

int64_t oracle_get_step(void* p) {
    return dispatch<int64_t>(static_cast<Handle*>(p), [&](State<float>* s) { return s->step; },
                             [&](State<double>* s) { return s->step; });
}

int32_t oracle_set_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// frames in the device layout: counts[nsc], pos/mom 3 planes of nsc*cap
int32_t oracle_set_frames(void* p, int32_t sp, const uint32_t* counts, const float* pos, const float* mom,
                          const float* w, const uint16_t* cell, const uint8_t* mask) {
    auto impl = [&](auto* st) -> int32_t {
        auto& S = st->sp[sp];
        size_t const slots = static_cast<size_t>(st->nsc) * S.cap;
        for (int i = 0; i < st->nsc; ++i) S.count[i] = counts[i];
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < slots; ++i) {
                S.pos[c][i] = pos[c * slots + i];
                S.mom[c][i] = mom[c * slots + i];
            }
        uint32_t next = 0;
        for (size_t i = 0; i < slots; ++i) {
            S.w[i] = w[i];
            S.cell[i] = cell[i];
            S.mask[i] = mask[i];
            S.id[i] = 0xFFFFFFFFu;
        }
        for (int sc = 0; sc < st->nsc; ++sc)
            for (uint32_t i = 0; i < S.count[sc]; ++i) S.id[static_cast<size_t>(sc) * S.cap + i] = next++;
        return 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

// read frames back; float outputs (FP64 twin values rounded) or double outputs
int32_t oracle_get_frames(void* p, int32_t sp, uint32_t* counts, double* pos, double* mom, double* w, uint16_t* cell,
                          uint8_t* mask) {
    auto impl = [&](auto* st) -> int32_t {
        auto& S = st->sp[sp];
        size_t const slots = static_cast<size_t>(st->nsc) * S.cap;
        for (int i = 0; i < st->nsc; ++i) counts[i] = S.count[i];
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < slots; ++i) {
                pos[c * slots + i] = static_cast<double>(S.pos[c][i]);
                mom[c * slots + i] = static_cast<double>(S.mom[c][i]);
            }
        for (size_t i = 0; i < slots; ++i) {
            w[i] = static_cast<double>(S.w[i]);
            cell[i] = S.cell[i];
            mask[i] = S.mask[i];
        }
        return 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

// particle identities of the frame pool (assigned in slot order by oracle_set_frames)
int32_t oracle_get_ids(void* p, int32_t sp, uint32_t* ids) {
    auto impl = [&](auto* st) -> int32_t {
        auto& S = st->sp[sp];
        std::copy(S.id.begin(), S.id.end(), ids);
        return 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

int32_t oracle_set_field(void* p, int32_t which, const float* data) {
    auto impl = [&](auto* st) -> int32_t {
        auto* f = which == 0 ? st->E : (which == 1 ? st->B : st->J);
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < st->ncells; ++i) f[c][i] = data[c * st->ncells + i];
        return 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

int32_t oracle_get_field(void* p, int32_t which, double* data) {
    auto impl = [&](auto* st) -> int32_t {
        auto* f = which == 0 ? st->E : (which == 1 ? st->B : st->J);
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < st->ncells; ++i) data[c * st->ncells + i] = static_cast<double>(f[c][i]);
        return 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

// the damping's FP64 e^x and ln x (test hooks)
double oracle_exp_f64(double x) { return expNegF64(x); }
double oracle_log_f64(double x) { return logUnitF64(x); }

// the PML's psi of E (which 0) or B (1): six planes in the reference's
// NodeValues order xy, xz, yx, yz, zx, zy (absorber/pml/Field.hpp); -1
// without the PML
int32_t oracle_get_pml(void* p, int32_t which, double* data) {
    auto impl = [&](auto* st) -> int32_t {
        if (!st->pml) return -1;
        auto* psi = which == 0 ? st->psiE : st->psiB;
        constexpr int kNodeOrder[6] = {2, 4, 0, 5, 1, 3};  // psi[] holds yx, zx, xy, zy, xz, yz
        for (int q = 0; q < 6; ++q)
            for (size_t i = 0; i < st->ncells; ++i) data[q * st->ncells + i] = static_cast<double>(psi[kNodeOrder[q]][i]);
        return 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

void oracle_current_reset(void* p) {
    auto impl = [&](auto* st) {
        for (int c = 0; c < 3; ++c) std::fill(st->J[c].begin(), st->J[c].end(), 0);
    };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

void oracle_push(void* p, int32_t species, int32_t fused) {
    auto impl = [&](auto* st) { push(*st, species, fused != 0); };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

// the field gather of pushOne for one particle: E and B of the state at the
// particle in global cell `cell` with in-cell position `pos`, interpolated
// with the shape's on-support assignment (out: Ex, Ey, Ez, Bx, By, Bz)
void oracle_gather(void* p, int32_t shape, const int32_t* cell, const float* pos, double* out) {
    auto impl = [&](auto* st) {
        using T = typename std::remove_reference<decltype(st->k.dt)>::type;
        T const q[3] = {static_cast<T>(pos[0]), static_cast<T>(pos[1]), static_cast<T>(pos[2])};
        T const fE[3][3] = {{T(0.5), T(0), T(0)}, {T(0), T(0.5), T(0)}, {T(0), T(0), T(0.5)}};
        T const fB[3][3] = {{T(0), T(0.5), T(0.5)}, {T(0.5), T(0), T(0.5)}, {T(0.5), T(0.5), T(0)}};
        int const supp = supportOf(shape);
        for (int c = 0; c < 3; ++c) {
            out[c] = static_cast<double>(gather1(*st, st->E[c], supp, cell[0], cell[1], cell[2], q, fE[c]));
            out[3 + c] = static_cast<double>(gather1(*st, st->B[c], supp, cell[0], cell[1], cell[2], q, fB[c]));
        }
    };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

void oracle_deposit(void* p, int32_t species) {
    auto impl = [&](auto* st) { deposit(*st, species); };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

int32_t oracle_shift(void* p, int32_t species) {
    auto impl = [&](auto* st) -> int32_t {
        shift(*st, species);
        return st->capacityError ? -3 : 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

void oracle_fdtd_before_current(void* p) {
    auto impl = [&](auto* st) { beforeCurrent(*st); };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

void oracle_add_current(void* p) {
    auto impl = [&](auto* st) { addCurrent(*st); };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

void oracle_fdtd_after_current(void* p) {
    auto impl = [&](auto* st) { afterCurrent(*st); };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

int32_t oracle_step(void* p, int32_t n) {
    auto impl = [&](auto* st) -> int32_t {
        step(*st, n);
        return st->capacityError ? -3 : 0;
    };
    return dispatch<int32_t>(static_cast<Handle*>(p), impl, impl);
}

// EnergyFields::getEnergyFields (plugins/EnergyFields.x.cpp:183-247): per
// component the FP64 sum of the squares taken in T (squareComponentWise, then
// cast64Bit, :61-69, :244-247), times 0.5/mue0*V for B and eps0*V*0.5 for E
// (:210-220; mue0, eps0 and the cell volume in T), in joules (unit.energy(),
// unitless/simulation.unitless:455-458). out: total, B x, y, z, E x, y, z.
void oracle_field_energy(void* p, double* out) {
    auto impl = [&](auto* st) {
        using T = typename std::remove_reference<decltype(st->k.dt)>::type;
        double sb[3] = {0.0, 0.0, 0.0}, se[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < st->ncells; ++i) {
                T const bv = st->B[c][i], ev = st->E[c][i];
                T const b2 = bv * bv, e2 = ev * ev;
                sb[c] += static_cast<double>(b2);
                se[c] += static_cast<double>(e2);
            }
        T const volume = st->k.cell[0] * st->k.cell[1] * st->k.cell[2];
        double const bFactor = 0.5 / static_cast<double>(st->k.mue0) * static_cast<double>(volume);
        T const eV = st->k.eps0 * volume;
        double const eFactor = static_cast<double>(eV) * 0.5;
        double const unitEnergy =
            st->k.unit_mass * st->k.unit_length * st->k.unit_length / (st->k.unit_time * st->k.unit_time);
        double eB = 0.0, eE = 0.0;
        for (int d = 0; d < 3; ++d) {
            double const b = sb[d] * bFactor, e = se[d] * eFactor;
            eB += b;
            eE += e;
            out[1 + d] = b * unitEnergy;
            out[4 + d] = e * unitEnergy;
        }
        out[0] = (eE + eB) * unitEnergy;
    };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

void oracle_charge_density(void* p, double* rho) {
    auto impl = [&](auto* st) { chargeDensity(*st, rho); };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

// backward-difference divergence of E at the cell origins (ChargeConservation.tpp:118-136), FP64
void oracle_div_e(void* p, double* div) {
    auto impl = [&](auto* st) {
        for (int z = 0; z < st->nz; ++z)
            for (int y = 0; y < st->ny; ++y)
                for (int x = 0; x < st->nx; ++x) {
                    double r = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        int const ux = d == 0, uy = d == 1, uz = d == 2;
                        r += (static_cast<double>(st->at(st->E[d], x, y, z)) -
                              static_cast<double>(st->at(st->E[d], x - ux, y - uy, z - uz))) /
                             static_cast<double>(st->k.cell[d]);
                    }
                    div[st->idx(x, y, z)] = r;
                }
    };
    dispatch<void>(static_cast<Handle*>(p), impl, impl);
}

// single-particle Esirkepov deposit into a caller array J (3 planes, nx*ny*nz)
// for a particle at global cell `cell`, in-cell `pos`, velocity `vel`, charge q
void oracle_esirkepov_single(const oracle_config* cfg, int32_t shape, const int32_t* cell, const float* pos,
                             const float* vel, float charge, double* J) {
    State<float> tmp;
    tmp.cfg = *cfg;
    tmp.k = makeConsts<float>(*cfg);
    tmp.nx = cfg->grid[0];
    tmp.ny = cfg->grid[1];
    tmp.nz = cfg->grid[2];
    size_t const n = static_cast<size_t>(tmp.nx) * tmp.ny * tmp.nz;
    float const p[3] = {pos[0], pos[1], pos[2]};
    V3<float> const v{vel[0], vel[1], vel[2]};
    esirkepov(supportOf(shape), p, v, charge, tmp.k, [&](int c, int x, int y, int z, float val) {
        J[c * n + tmp.idx(cell[0] + x, cell[1] + y, cell[2] + z)] += static_cast<double>(val);
    });
}

// the same with the current solver chosen: 0 Esirkepov, 1 EmZ
void oracle_deposit_single(const oracle_config* cfg, int32_t shape, int32_t scheme, const int32_t* cell,
                           const float* pos, const float* vel, float charge, double* J) {
    State<float> tmp;
    tmp.cfg = *cfg;
    tmp.k = makeConsts<float>(*cfg);
    tmp.nx = cfg->grid[0];
    tmp.ny = cfg->grid[1];
    tmp.nz = cfg->grid[2];
    size_t const n = static_cast<size_t>(tmp.nx) * tmp.ny * tmp.nz;
    float const p[3] = {pos[0], pos[1], pos[2]};
    V3<float> const v{vel[0], vel[1], vel[2]};
    depositParticle(scheme, supportOf(shape), p, v, charge, tmp.k, [&](int c, int x, int y, int z, float val) {
        J[c * n + tmp.idx(cell[0] + x, cell[1] + y, cell[2] + z)] += static_cast<double>(val);
    });
}

// moveParticle on one particle (for the reference unit test cases)
int32_t oracle_move_particle(float* pos, uint16_t* localCellIdx, uint8_t* multiMask) {
    int step[3];
    return moveParticle(pos, *localCellIdx, *multiMask, step);
}

// shape values: kind 0 = ChargeAssignment(x), 1 = on-support array at x
// (supp values), 2 = cached array (supp+1 values, out-of-range flag in `flag`)
void oracle_shape(int32_t shape, int32_t kind, float x, int32_t flag, float* out) {
    ShapeFns<float> sh{supportOf(shape)};
    if (kind == 0)
        out[0] = sh.assign(x);
    else if (kind == 1)
        sh.onSupport(x, out);
    else
        sh.cached(x, flag != 0, out);
}

}  // extern "C"
