/* picongpu_amd.h -- C ABI of the MI355X-native PIC-cycle hot path.
 *
 * One context (pic_ctx) is one periodic 3-D simulation box on one HIP device.
 * Every call is serialised on the context's HIP stream; functions return 0 on
 * success and a negative PIC_E_* code on failure, with a message available from
 * pic_last_error() (thread-local). No C++ exception crosses this boundary.
 * Memory passed in is owned by the caller and only read or written during the
 * call; device buffers are owned by the context.
 *
 * Units: every float handed across the ABI is in PIConGPU's internal unit
 * system (include/picongpu/unitless/simulation.unitless:31-500): time unit dt,
 * speed unit c, length unit c*dt, mass unit BASE_MASS_SI*W_typ, charge unit
 * -BASE_CHARGE_SI*W_typ with W_typ = BASE_DENSITY_SI*cellVolume/TYPICAL_PARTICLES_PER_CELL.
 * Particle positions are cell-relative in [0,1), momenta are macro-particle
 * momenta, exactly as the reference stores them
 * (include/picongpu/param/speciesDefinition.param:49).
 *
 * Storage layout (the reference's supercell frame semantics,
 * include/pmacc/particles/memory/dataTypes/SuperCell.hpp:31-119): the box is
 * tiled by supercells of 8x8x4 cells; every supercell owns `frames_per_sc`
 * frames of 256 slots at first, and a supercell that a shift or a load would
 * overfill grows by whole frames on its own, so no particle is ever dropped
 * (the counterpart of the frames the reference takes on demand from its
 * mallocMC heap, include/pmacc/particles/memory/boxes/ParticlesBox.hpp:96-122);
 * its N particles occupy slots [0,N) so all frames are
 * full except the last, which holds (N-1)%256+1. localCellIdx is the
 * x-fastest linear cell index inside the supercell; multiMask is 0 (empty),
 * 1 (valid) or 2..27 (leaving towards exchange direction multiMask-1, with
 * RIGHT=1, LEFT=2, BOTTOM=3, TOP=6, BACK=9, FRONT=18 summed over axes,
 * include/picongpu/particles/MoveParticle.hpp:139-157).
 * Fields are three SoA planes (x, y, z component) of nx*ny*nz floats,
 * x fastest, Yee-staggered as include/picongpu/fields/YeeCell.hpp:69-141.
 */
#ifndef PICONGPU_AMD_H
#define PICONGPU_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_ABI_VERSION 1
#define PIC_MAX_SPECIES 4

/* error codes */
#define PIC_OK 0
#define PIC_E_INVALID -1     /* bad argument or configuration */
#define PIC_E_HIP -2         /* HIP runtime error */
#define PIC_E_CAPACITY -3    /* a supercell could not grow to hold its particles (device memory) */
#define PIC_E_STATE -4       /* call not valid in the current state */

/* particle shapes: include/picongpu/particles/shapes/{CIC,TSC,PQS}.hpp */
#define PIC_SHAPE_CIC 1
#define PIC_SHAPE_TSC 2
#define PIC_SHAPE_PQS 3

/* pushers: particlePusherBoris.hpp:41-90, particlePusherVay.hpp:42-111;
 * NONE = immobile species (not pushed, no current) */
#define PIC_PUSHER_NONE 0
#define PIC_PUSHER_BORIS 1
#define PIC_PUSHER_VAY 2
#define PIC_PUSHER_HIGUERA_CARY 3 /* particlePusherHigueraCary.hpp:49-142 */

/* current deposition per species (species.param UsedParticleCurrentSolver):
 * Esirkepov (currentDeposition/Esirkepov/Esirkepov.hpp:61-242) or EmZ
 * (currentDeposition/EmZ/EmZ.hpp:61-137, DepositCurrent.hpp:46-112) */
#define PIC_CURRENT_SOLVER_ESIRKEPOV 0
#define PIC_CURRENT_SOLVER_EMZ 1

/* field solvers: fields/MaxwellSolver/Yee/Yee.def:36-47, Lehe/Lehe.def:57-73 */
#define PIC_SOLVER_YEE 0
#define PIC_SOLVER_LEHE 1

/* outer boundary per axis. PERIODIC: fields and particles wrap. OPEN: the
 * reference with that axis not periodic and every species' particle boundary
 * Kind::Absorbing with offset 0 (particles/boundary/Description.hpp:36-44):
 * particles crossing the border are removed (boundary/Absorbing.hpp:63-92,
 * before the deposit sees them); fields read zero beyond it (the guard cells
 * of a non-periodic border are never written nor exchanged) and current
 * deposited beyond it is dropped (FieldJ.x.cpp:156-174 folds guards only
 * between neighbours). */
#define PIC_BOUNDARY_PERIODIC 0
#define PIC_BOUNDARY_OPEN 1

/* particle boundary of a species on an open axis (particles/boundary/Kind.hpp,
 * the species' boundaryDescription()[axis].kind with offset 0): ABSORBING
 * (boundary/Absorbing.hpp:63-92, the default) or REFLECTING
 * (boundary/Reflecting.hpp:38-66: the in-cell position is mirrored at the
 * border and the momentum component negated) */
#define PIC_PARTICLE_ABSORBING 0
#define PIC_PARTICLE_REFLECTING 1

/* field absorber at open borders (fields/absorber/Absorber.hpp): NONE, the
 * exponential damping of absorber/exponential/Exponential.kernel:63-112, or
 * the convolutional PML of absorber/pml/Pml.kernel:190-590 (the reference's
 * default kind, simulation/stage/FieldAbsorber.hpp:88) */
#define PIC_ABSORBER_NONE 0
#define PIC_ABSORBER_EXPONENTIAL 1
#define PIC_ABSORBER_PML 2
#define PIC_ABSORBER_MAX_CELLS 64

/* current interpolation before J is added to E (fields/currentInterpolation):
 * NONE (None.hpp:56-65) or the 27-point BINOMIAL filter (Binomial.hpp:40-104),
 * the latter on periodic boxes (the reference reads neighbour-refreshed J
 * guards, which an open border does not have) */
#define PIC_CURRENT_NONE 0
#define PIC_CURRENT_BINOMIAL 1

/* incident field (fields/incidentField): a profile per face of the Huygens
 * surface (param/incidentField.param XMin..ZMax), NONE (profiles/None.def),
 * GAUSSIAN_PULSE (profiles/GaussianPulse.hpp with GaussianPulseEnvelope) or
 * PLANE_WAVE (profiles/PlaneWave.hpp); polarisation and focus origin as
 * profiles/BaseParam.def:39-56 */
#define PIC_INCIDENT_NONE 0
#define PIC_INCIDENT_GAUSSIAN_PULSE 1
#define PIC_INCIDENT_PLANE_WAVE 2
#define PIC_POLARISATION_LINEAR 0
#define PIC_POLARISATION_CIRCULAR 1
#define PIC_ORIGIN_ZERO 0
#define PIC_ORIGIN_CENTER 1
#define PIC_INCIDENT_MAX_MODES 16

/* fields */
#define PIC_FIELD_E 0
#define PIC_FIELD_B 1
#define PIC_FIELD_J 2
/* the PML's convolution fields of E and B (absorber/pml/Field.hpp: "Convolutional
 * PML E" / "Convolutional PML B", checkpointed by the reference): six planes
 * of nx*ny*nz floats in NodeValues' component order xy, xz, yx, yz, zx, zy,
 * zero outside the layers; pic_load_field / pic_read_field only, and only
 * with absorber PIC_ABSORBER_PML */
#define PIC_FIELD_PML_E 3
#define PIC_FIELD_PML_B 4

/* per-kernel timers (pic_kernel_timer) */
#define PIC_TIMER_PUSH 0      /* fused gather+push+move+Esirkepov launches */
#define PIC_TIMER_SHIFT 1
#define PIC_TIMER_FIELD 2
#define PIC_TIMER_COUNT 3

/* Species description: the flags of a PIConGPU species
 * (param/speciesDefinition.param:49-101): massRatio<>, chargeRatio<>,
 * shape<>, particlePusher<> and the frame capacity of its supercells. */
typedef struct pic_species_desc {
    double mass_ratio;      /* relative to BASE_MASS_SI (electron mass) */
    double charge_ratio;    /* relative to BASE_CHARGE_SI (electron charge) */
    int32_t shape;          /* PIC_SHAPE_* */
    int32_t pusher;         /* PIC_PUSHER_* */
    int32_t frames_per_sc;  /* frames of 256 slots preallocated per supercell; a
                               supercell that needs more grows on its own
                               (pic_supercell_slots reports every capacity) */
    int32_t current_solver; /* PIC_CURRENT_SOLVER_*: the species' current<> flag */
    int32_t particle_boundary[3]; /* PIC_PARTICLE_* per axis, used on open axes */
} pic_species_desc;

/* One incident-field profile: the members of profiles::BaseParam
 * (profiles/BaseParam.def:60-200) and of the GaussianPulse and PlaneWave
 * parameter structs (GaussianPulse.def:36-70, PlaneWave.def:36-49), in SI.
 * Zero-initialised = PIC_INCIDENT_NONE. */
typedef struct pic_incident_profile {
    int32_t kind;                      /* PIC_INCIDENT_* */
    int32_t polarisation;              /* Polarisation: PIC_POLARISATION_* */
    double wavelength_si;              /* WAVE_LENGTH_SI [m] */
    double amplitude_si;               /* AMPLITUDE_SI [V/m] */
    double pulse_duration_si;          /* PULSE_DURATION_SI [s] (sigma of the intensity) */
    double time_delay_si;              /* TIME_DELAY_SI [s], >= 0 */
    float laser_phase;                 /* LASER_PHASE [rad] */
    double direction[3];               /* DIRECTION_X/Y/Z, unit vector */
    double polarisation_direction[3];  /* POLARISATION_DIRECTION_X/Y/Z, unit, orthogonal to direction */
    double focus_position_si[3];       /* FOCUS_POSITION_X/Y/Z_SI [m] */
    int32_t focus_origin[3];           /* FOCUS_ORIGIN_X/Y/Z: PIC_ORIGIN_* */
    /* GaussianPulse */
    double w0_si;                      /* W0_SI [m] */
    double pulse_init;                 /* PULSE_INIT */
    int32_t num_modes;                 /* numModes: Laguerre modes 0..num_modes */
    float laguerre_modes[PIC_INCIDENT_MAX_MODES];
    float laguerre_phases[PIC_INCIDENT_MAX_MODES];
    /* PlaneWave */
    double laser_nofocus_constant_si;  /* LASER_NOFOCUS_CONSTANT_SI [s] */
    double ramp_init;                  /* RAMP_INIT */
} pic_incident_profile;

/* PML parameters (param/fieldAbsorber.param namespace pml). With custom == 0
 * the reference's defaults apply: SIGMA_KAPPA_GRADING_ORDER 4, SIGMA_MAX_SI =
 * SIGMA_OPT_SI = 0.8 (order + 1) / (Z0 cell size) per axis, KAPPA_MAX 1,
 * ALPHA_GRADING_ORDER 1, ALPHA_MAX_SI 0.2 S/m. The layer thickness is
 * pic_config.absorber_cells, on open axes only. */
typedef struct pic_pml_params {
    int32_t custom;                    /* 0: the defaults above; 1: the fields below */
    double sigma_kappa_grading_order;  /* SIGMA_KAPPA_GRADING_ORDER (>= 0) */
    double sigma_max_si[3];            /* SIGMA_MAX_SI [S/m] (>= 0) */
    double kappa_max[3];               /* KAPPA_MAX (>= 1) */
    double alpha_grading_order;        /* ALPHA_GRADING_ORDER (>= 0) */
    double alpha_max_si[3];            /* ALPHA_MAX_SI [S/m] (>= 0) */
} pic_pml_params;

/* Simulation description: the .param surface of the hot path
 * (param/simulation.param DELTA_T_SI/CELL_*_SI, param/density.param
 * BASE_DENSITY_SI, param/particle.param TYPICAL_PARTICLES_PER_CELL,
 * param/fieldSolver.param Solver). */
typedef struct pic_config {
    int32_t grid[3];          /* global cells, multiples of 8,8,4 */
    double cell_si[3];        /* CELL_WIDTH_SI, CELL_HEIGHT_SI, CELL_DEPTH_SI [m] */
    double dt_si;             /* DELTA_T_SI [s] */
    double base_density_si;   /* BASE_DENSITY_SI [1/m^3] */
    int32_t typical_ppc;      /* TYPICAL_PARTICLES_PER_CELL */
    int32_t solver;           /* PIC_SOLVER_* */
    int32_t lehe_dir;         /* Cherenkov-free axis of the Lehe solver (0,1,2) */
    int32_t n_species;
    pic_species_desc species[PIC_MAX_SPECIES];
    int32_t j_scale_exp;      /* fixed-point exponent of the current tile, 0 = automatic */
    int32_t device;           /* HIP device ordinal */
    int32_t boundary[3];      /* PIC_BOUNDARY_* per axis; zero-initialised = periodic */
    int32_t absorber;         /* PIC_ABSORBER_* (acts on open axes only) */
    int32_t absorber_cells[3][2];  /* param/fieldAbsorber.param NUM_CELLS [axis][negative, positive],
                                    * at most PIC_ABSORBER_MAX_CELLS */
    float absorber_strength[3][2]; /* param/fieldAbsorber.param exponential::STRENGTH */
    int32_t current_interpolation; /* PIC_CURRENT_* */
    /* incidentField.param XMin, XMax, YMin, YMax, ZMin, ZMax as [axis][negative, positive] */
    pic_incident_profile incident[3][2];
    /* incidentField.param POSITION[axis][negative, positive]: Huygens surface
     * cell index from the negative border, and from the positive border when
     * the entry is <= 0 (a positive entry is an absolute index) */
    int32_t incident_position[3][2];
    pic_pml_params pml;       /* PIC_ABSORBER_PML parameters */
} pic_config;

/* FP32 constants in PIC units derived from a pic_config, computed as
 * unitless/simulation.unitless:31-110,428-479 and the solver/pusher files do. */
typedef struct pic_constants {
    float cell[3];            /* sim.pic.getCellSize() */
    float dt;                 /* sim.pic.getDt() (== 1) */
    float c;                  /* sim.pic.getSpeedOfLight() (== 1) */
    float eps0;               /* sim.pic.getEps0() */
    float mue0;               /* sim.pic.getMue0() */
    float mue0_eps0;          /* sim.pic.getMue0Eps0() */
    float add_current_coeff;  /* -(1/eps0)*dt, FDTD.hpp:84 */
    float inv_cell_volume_dt; /* 1/float(V*dt), Esirkepov.hpp:190-192 */
    float lehe_alpha;         /* Lehe/Derivative.hpp:90-111 (FP64 -> FP32) */
    float lehe_delta;
    float species_mass[PIC_MAX_SPECIES];   /* frame::getMass() */
    float species_charge[PIC_MAX_SPECIES]; /* frame::getCharge() */
    double unit_length, unit_time, unit_mass, unit_charge; /* SI per PIC unit */
    double typical_weighting; /* W_typ */
} pic_constants;

/* Device-side initial state in the manner of KernelFillGridWithParticles +
 * startPosition::Random (or Quiet) + Drift + Temperature (particles/ParticlesInit.kernel:66-302,
 * startPosition/RandomImpl.hpp:46-56, manipulators/unary/Drift.hpp:53-76,
 * Temperature.hpp:60-84): ppc particles per cell at uniform random in-cell
 * positions, weighting density*V/ppc, momentum mass*(u_drift + u_th*N(0,1)).
 * The drift may flip sign inside a y band (Kelvin-Helmholtz shear). */
typedef struct pic_init_desc {
    uint64_t seed;
    int32_t ppc;              /* macro-particles per cell */
    int32_t drift_band;       /* 0: uniform drift; 1: drift sign +1 for y in [ny/4,3ny/4), -1 outside;
                                 2: sign +1 for even, -1 for odd particle index k in the cell (two
                                 counter-streaming beams) */
    double density_si;        /* physical density [1/m^3] */
    double u_drift[3];        /* gamma*beta of the drift */
    double u_th[3];           /* thermal spread of gamma*beta per component */
    int32_t profile;          /* 0: uniform; 1: linear ramp along profile_axis, evaluated at the cell
                                 origin: density_si * clamp((x - ramp_start_si)/ramp_length_si, 0, 1);
                                 cells at zero density get no particles, the others ppc particles of
                                 weighting density*V/ppc (densityProfiles, WeightMacroParticles.hpp:30-42) */
    int32_t profile_axis;
    double ramp_start_si;
    double ramp_length_si;
    /* start position: all zero = startPosition::Random; otherwise
     * startPosition::Quiet (startPosition/QuietImpl.hpp:34-76) with these
     * particles per dimension (their product = ppc): the k-th particle of a
     * cell sits at lattice point ppc-1-k (x fastest), position (i + 1/2)/n_i */
    int32_t quiet[3];
} pic_init_desc;

/* Gaussian laser pulse set as the initial field (the reference injects it
 * through a Huygens surface, incidentField/profiles/GaussianPulse.hpp; here
 * the pulse starts inside the box at its focus): E along polarisation_axis
 * = E0 exp(-r^2/w0^2) exp(-xi^2/(2 sigma^2)) cos(k xi), B = (1/c) k_hat x E,
 * xi = distance from center_si along propagation_axis, r the distance from
 * that axis, E0 = a0 m_e c omega / e, sigma = c * fwhm_si / (2 sqrt(ln 2))
 * (intensity FWHM), each component sampled at its Yee position. */
typedef struct pic_laser_desc {
    double wavelength_si;
    double a0;
    double waist_si;
    double fwhm_si;
    double center_si[3];
    int32_t propagation_axis;
    int32_t polarisation_axis;
} pic_laser_desc;

typedef struct pic_ctx pic_ctx;

/* --- lifecycle ---------------------------------------------------------- */
const char* pic_last_error(void);
int32_t pic_abi_version(void);
/* unitless/simulation.unitless: constants of a configuration (host only, no device) */
int pic_constants_from_config(const pic_config* cfg, pic_constants* out);
/* Simulation::pluginLoad + Simulation::init (simulation/control/Simulation.hpp:176-436):
 * allocates fields (zeroed) and frame pools on the device */
int pic_create(const pic_config* cfg, pic_ctx** out);
int pic_destroy(pic_ctx* ctx);
int pic_get_constants(const pic_ctx* ctx, pic_constants* out);
int pic_sync(pic_ctx* ctx);

/* --- state in and out --------------------------------------------------- */
/* Load n particles given by global cell coordinates (3n int32, x,y,z), in-cell
 * positions (3n), momenta (3n) and weightings (n). They are binned into their
 * supercells in input order (replaces particle init / restart LoadSpecies). */
int pic_load_particles(pic_ctx* ctx, int32_t species, int64_t n, const int32_t* cell,
                       const float* pos, const float* mom, const float* weighting);
/* Read all particles of a species in frame-canonical order (supercells in
 * x-fastest order, slots ascending). `capacity` is the length of w; n_out gets
 * the particle count. Any output pointer may be NULL. */
int pic_read_particles(pic_ctx* ctx, int32_t species, int64_t capacity, int64_t* n_out,
                       int32_t* cell, float* pos, float* mom, float* weighting);
int64_t pic_particle_count(pic_ctx* ctx, int32_t species);
int32_t pic_num_supercells(const pic_ctx* ctx);
/* The largest capacity (slots) of a supercell of the species: the number of
 * slots per supercell of the frame view pic_read_frames/pic_write_frames use.
 * Supercells grow one by one (by whole frames) when they need more slots. */
int32_t pic_slots_per_supercell(const pic_ctx* ctx, int32_t species);
/* The capacity (slots, a multiple of 256) of every supercell: out[nsc]. */
int pic_supercell_slots(const pic_ctx* ctx, int32_t species, uint32_t* out);
/* Frame view of the pool: per-supercell counts plus the slot arrays of
 * nsc*slots entries each, slots = pic_slots_per_supercell at the call (a
 * supercell with fewer slots reads zero beyond them). pos/mom are 3 planes of
 * nsc*slots floats. */
int pic_read_frames(pic_ctx* ctx, int32_t species, uint32_t* counts, float* pos, float* mom,
                    float* weighting, uint16_t* local_cell, uint8_t* multi_mask);
int pic_write_frames(pic_ctx* ctx, int32_t species, const uint32_t* counts, const float* pos,
                     const float* mom, const float* weighting, const uint16_t* local_cell,
                     const uint8_t* multi_mask);
int pic_init_species(pic_ctx* ctx, int32_t species, const pic_init_desc* desc);
int pic_load_field(pic_ctx* ctx, int32_t field, const float* data);
int pic_read_field(pic_ctx* ctx, int32_t field, float* data);
/* B += amp*U(-1,1) per component, seeded (initial noise of the shear-flow setup) */
int pic_add_field_noise(pic_ctx* ctx, int32_t field, uint64_t seed, float amp);
/* adds the laser pulse to E and B */
int pic_add_laser(pic_ctx* ctx, const pic_laser_desc* laser);

/* --- stages of one PIC step (Simulation::runOneStep, Simulation.hpp:522-542) --- */
/* CurrentReset (simulation/stage/CurrentReset.hpp:39-53): J = 0 */
int pic_current_reset(pic_ctx* ctx);
/* ParticlePush (stage/ParticlePush.x.cpp:141-148 -> KernelMoveAndMarkParticles,
 * Particles.kernel:169-316): field gather, push, moveParticle, mark leavers.
 * species < 0 pushes every mobile species. With fuse_deposit != 0 the same
 * launch also runs Esirkepov<Shape> (CurrentDeposition, stage/CurrentDeposition.x.cpp:49-117)
 * on each particle's move and adds the current to J. */
int pic_push(pic_ctx* ctx, int32_t species, int32_t fuse_deposit);
/* shiftBetweenSupercells + periodic boundary (Particles.tpp:370-377,
 * pmacc/particles/ParticlesBase.kernel:61-630, boundary/Apply.hpp:74-116);
 * a supercell that overflows grows before the call returns (it synchronises) */
int pic_shift(pic_ctx* ctx, int32_t species);
/* CurrentDeposition from the current particle state (unfused form) */
int pic_deposit(pic_ctx* ctx, int32_t species);
/* fields::Solver::update_beforeCurrent (FDTDBase.hpp:97-121): B half step,
 * incident field on B (at the field step counter) and on E (counter + 0.5),
 * E full step */
int pic_fdtd_before_current(pic_ctx* ctx);
/* FieldJ::asyncCommunication (FieldJ.x.cpp:156-174): the fold of the J
 * guards into the opposite border. The deposit's tile flush already adds
 * every contribution at its wrapped cell on a periodic axis and drops it
 * beyond an open border, so J is complete after pic_deposit / a fused
 * pic_push; this call only checks the context. It is the counterpart for a
 * caller that keeps the reference's stage sequence. */
int pic_fold_current(pic_ctx* ctx);
/* CurrentInterpolationAndAdditionToEMF with currentInterpolation::None
 * (CurrentInterpolationAndAdditionToEMF.hpp:99-148, FDTD.hpp:78-86): E += -(dt/eps0) J */
int pic_add_current(pic_ctx* ctx);
/* fields::Solver::update_afterCurrent (FDTDBase.hpp:151-183): absorber on E,
 * incident field on B (counter + 1), B half step, absorber on B; then the
 * field step counter advances by one */
int pic_fdtd_after_current(pic_ctx* ctx);
/* update_beforeCurrent, addCurrent and update_afterCurrent back to back, as
 * Simulation::runOneStep calls them (Simulation.hpp:522-542), in two fused
 * LDS-tiled kernels (the stage kernels when an incident field is set); E, B
 * bit-identical to the three calls above */
int pic_fdtd_step(pic_ctx* ctx);
/* the field step counter: the currentStep argument of update_beforeCurrent /
 * update_afterCurrent (FDTDBase.hpp:97,151), which times the incident field
 * (incidentField/Functors.hpp BaseFunctorE: t = currentStep * dt). It starts
 * at 0 and lives on the device, so captured steps advance it too. */
int pic_set_step(pic_ctx* ctx, int64_t step);
int pic_get_step(pic_ctx* ctx, int64_t* step);
/* nsteps full PIC steps in the reference stage order, push and deposit fused.
 * When a shift finds a supercell full, the records it cannot place wait in
 * the senders' outboxes and the following steps are skipped on the device;
 * before returning, the supercells of that species grow, the waiting records
 * are placed where a larger capacity would have put them and the skipped
 * steps run, so the result equals that of a larger frames_per_sc bit for bit
 * (PIC_FLAG_CAPACITY_GREW is raised). Only a growth that cannot get device
 * memory loses particles: then PIC_E_CAPACITY, and every later step call
 * returns PIC_E_STATE. */
int pic_step(pic_ctx* ctx, int32_t nsteps);
/* the same steps replayed from one captured hipGraph of a step (captured on
 * first use, recaptured when the current's fixed-point scale changes);
 * not available while kernel timing is enabled */
int pic_step_graph(pic_ctx* ctx, int32_t nsteps);

/* --- diagnostics ------------------------------------------------------- */
/* EnergyFields::getEnergyFields (plugins/EnergyFields.x.cpp:183-247): per
 * component the FP64 sum over cells of the FP32 square, times 0.5/mue0*V for
 * B and eps0*V*0.5 for E (the reference's float/double mix), in joules.
 * out[0] = total, out[1..3] = B x, y, z, out[4..6] = E x, y, z: the columns of
 * the reference's output line after the step number. */
int pic_field_energy(pic_ctx* ctx, double out[7]);
/* Gauss-law residual eps0*div(E) - rho at the cell origins, FP64, nx*ny*nz
 * values x fastest; rho from every species' charge assignment shape
 * (ChargeConservation.tpp:181-260). Charge conservation keeps it constant
 * in time up to round-off. */
int pic_gauss_residual(pic_ctx* ctx, double* out);
/* pic_gauss_residual without any device allocation. It has no counterpart in
 * the reference; it is for callers that cannot allocate device memory (tight
 * monitoring loops, harnesses that forbid hipMalloc). The per-cell charge
 * density goes to a file-scope device array of 2^21 doubles (16 MiB) that
 * each device holds until the process ends, so grids above 2,097,152 cells
 * are rejected with PIC_E_INVALID. Every call on one device shares that
 * array: calls that overlap on one device race. Same output as
 * pic_gauss_residual (FP64 atomics: equal up to FP64 round-off). */
int pic_gauss_residual_noalloc(pic_ctx* ctx, double* out);
/* PIC_FLAG_* bits raised since the last call */
#define PIC_FLAG_CAPACITY_GREW 1u /* a supercell overflowed and the species' supercells grew (nothing lost) */
int pic_error_flags(pic_ctx* ctx, uint32_t* flags);
/* per-kernel timing with HIP events on the context stream */
int pic_timing_enable(pic_ctx* ctx, int32_t enable);
int pic_kernel_timer(pic_ctx* ctx, int32_t timer, double* total_ms, int64_t* launches);
int32_t pic_j_scale_exp(const pic_ctx* ctx);
int32_t pic_num_push_phases(const pic_ctx* ctx, int32_t fused);
/* Launch schedule of the fused push/deposit of a step (no counterpart in the
 * reference, whose deposit strategy launches per checkerboard colour
 * (Strategy.def:196-200): an execution knob of this library). PHASES: one
 * launch per checkerboard phase, each workgroup adding its supercell's tile
 * to J. SINGLE: one launch for every supercell, the tiles left in a scratch
 * of one tile per supercell and added to J by a gather launch, every J cell
 * adding its covering tiles in phase order. Both give the same J, bit for
 * bit; SINGLE fills the GPU when a phase holds fewer supercells than the
 * device has workgroup slots. AUTO (the default) takes SINGLE when the
 * largest phase has fewer supercells than twice the compute units. An
 * ensemble member's schedule cannot change (PIC_E_INVALID): the ensemble's
 * batched launches use its plans. pic_push_schedule reports the schedule of
 * a shape support (2, 3, 4). */
#define PIC_PUSH_SCHEDULE_AUTO 0
#define PIC_PUSH_SCHEDULE_PHASES 1
#define PIC_PUSH_SCHEDULE_SINGLE 2
int pic_set_push_schedule(pic_ctx* ctx, int32_t schedule);
int32_t pic_push_schedule(const pic_ctx* ctx, int32_t support);

/* --- differentiable form and its adjoint ------------------------------- */
/* The reference has no adjoint; these entry points expose the same step
 * (Simulation.hpp:522-542 stage order: ParticlePush with Boris and CIC
 * gather, Esirkepov<CIC> deposit, Yee update_beforeCurrent, addCurrent,
 * update_afterCurrent) for one species with particles kept in identity
 * order (global cell index and in-cell position, no supercell shift), and its
 * hand-written FP64 reverse: transposed curl stencils, the Esirkepov transpose
 * as a gather of the current's adjoint, reverse Boris and reverse gather.
 * The forward arithmetic per particle is the FP32 arithmetic of pic_push.
 * The configuration must use the Yee solver; the species must be CIC/Boris.
 * Field and gradient buffers of pic_adjoint_forward/backward are DEVICE
 * pointers on the context's device, 3 component planes of nx*ny*nz values
 * (x fastest); the calls synchronise the context stream before returning. */

/* initial particles of the differentiable form (host pointers): cell = global
 * cell index, pos = in-cell position in [0,1), 3 planes of n values each */
int pic_adjoint_load_particles(pic_ctx* ctx, int32_t species, int64_t n, const int32_t* cell, const float* pos,
                               const float* mom, const float* w);
/* particles after the last pic_adjoint_forward (host pointers, n as loaded) */
int pic_adjoint_read_particles(pic_ctx* ctx, int32_t* cell, float* pos, float* mom);
/* nsteps steps from the loaded particles and the initial fields E0, B0
 * (float, or null to start from the context's E and B), storing the state
 * every checkpoint_every steps; the final fields are left in the context and
 * copied to E_out, B_out when they are not null */
int pic_adjoint_forward(pic_ctx* ctx, const float* E0, const float* B0, int32_t nsteps, int32_t checkpoint_every,
                        float* E_out, float* B_out);
/* reverse of the last pic_adjoint_forward: given dF/dE and dF/dB of the
 * final fields (double), writes dF/dE0 and dF/dB0 (double, 3*ncells each)
 * and, when not null, dF/dX0 (cells) and dF/dp0 of the initial particles
 * (double, 3 planes of n); recomputes each checkpoint segment forward */
int pic_adjoint_backward(pic_ctx* ctx, const double* dE, const double* dB, double* dE0, double* dB0, double* dX0,
                         double* dP0);

/* --- batched ensembles -------------------------------------------------- */
/* Many independent periodic simulations (each a pic_ctx, grids and particle
 * counts may differ) advanced together: one launch per checkerboard phase of
 * the fused push/deposit over every member's supercells, one launch each for
 * the shift's insert and compact passes, the fused field step and the current
 * reset, optionally replayed from a captured hipGraph. The reference runs one
 * simulation per process; there is no counterpart interface. Members are
 * created from the configurations and are then initialised, loaded, read and
 * diagnosed through the single-simulation calls on pic_ensemble_member();
 * stepping a member on its own is allowed between ensemble steps only in even
 * numbers (the fused field step alternates two E buffers). Every member must
 * use the same shape, pusher and solver, at most 64 frames per supercell (a
 * member that overflows grows up to that and runs its skipped steps on its own). */
typedef struct pic_ensemble pic_ensemble;

/* order of the work items of a batched phase */
#define PIC_ENSEMBLE_ORDER_SIM_MAJOR 0    /* member by member */
#define PIC_ENSEMBLE_ORDER_INTERLEAVED 1  /* one supercell of every member in turn */
#define PIC_ENSEMBLE_ORDER_HEAVY_FIRST 2  /* member by member, largest particle capacity first */

int pic_ensemble_create(const pic_config* cfgs, int32_t n, pic_ensemble** out);
int pic_ensemble_destroy(pic_ensemble* ens);
int32_t pic_ensemble_size(const pic_ensemble* ens);
/* the member context i (owned by the ensemble), or null with pic_last_error set */
pic_ctx* pic_ensemble_member(pic_ensemble* ens, int32_t i);
/* nsteps batched steps of every member; each member's results equal its own pic_step */
int pic_ensemble_step(pic_ensemble* ens, int32_t nsteps);
/* the same, replayed from captured graphs (one per E buffer parity) */
int pic_ensemble_step_graph(pic_ensemble* ens, int32_t nsteps);
/* schedule of the batched push: work-item order and workgroups per CU
 * (0 = one workgroup per item; k > 0 = k*CUs persistent workgroups looping over items) */
int pic_ensemble_set_schedule(pic_ensemble* ens, int32_t order, int32_t blocks_per_cu);
int pic_ensemble_get_schedule(const pic_ensemble* ens, int32_t* order, int32_t* blocks_per_cu);
/* times every (order, blocks per CU) candidate over trial_steps batched steps
 * (after one warm-up step each) and keeps the fastest. The schedule does not
 * change results, so these are ordinary steps of every member: steps_taken
 * reports how many were run. best_ms_per_step receives the winner's time. */
int pic_ensemble_autotune(pic_ensemble* ens, int32_t trial_steps, double* best_ms_per_step, int32_t* steps_taken);

#ifdef __cplusplus
}
#endif

#endif /* PICONGPU_AMD_H */
