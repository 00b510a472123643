/*
 * mpsr.h -- C ABI of the MI355X-native MPS rotating-wall timestep.
 *
 * This header replaces the reference's per-kernel launch ABI
 * (solver/MPS_SOLVER_WITH_ROTATION/PMMultiGPU/Source/DeviceProcess.h:34-89, implemented in
 * DeviceProcess.cu:29-304) and the per-device manager that drives it
 * (CCellSet, Source/CGPUCellSet.cpp:1055-1459).  Plain C: POD structs, plain pointers and
 * sizes, a status return on every call, no exceptions across the boundary.
 *
 * Ownership: the handle owns every device allocation.  The caller owns every host buffer it
 * passes in or out.  A handle is not re-entrant; separate handles may run concurrently.
 */
#ifndef MPSR_H
#define MPSR_H

#ifdef __cplusplus
extern "C" {
#endif

/* CCTStatusType, same order (solver/Common/inc/DataType.h:252-261). */
typedef enum mpsr_status {
    MPSR_NOERR = 0,
    MPSR_FILEERR = 1,
    MPSR_PARAMERR = 2,
    MPSR_CALCERR = 3,
    MPSR_MEMERR = 4,
    MPSR_GPUERR = 5, /* CCT_CUDAERR */
    MPSR_ETCERR = 6
} mpsr_status;

typedef struct mpsr_float3 { float x, y, z; } mpsr_float3;
typedef struct mpsr_int3 { int x, y, z; } mpsr_int3;

/* CParameter field for field (solver/Common/inc/DataType.h:426-627).  C++ bool members are
 * carried as int.  The reference spells several names this way; they are kept so a binding
 * copies field by field. */
typedef struct mpsr_params {
    int Dimension;
    mpsr_float3 GravityConstant;
    float Density;
    float ConstantParticleDensity;
    float SurfaceTensionCoefficient;
    float InitialParticleDistance;
    float FreeSurfaceThreshold;
    float GradientInfluenceRadiusCoefficient;
    float LaplacianInfluenceRadiusCoefficient;
    float TemperatureInfluenceRadiusCoefficient;
    float FruidInitializeialTemperature;
    float WallInitializeialTemperature;
    float FruidThermalResistance;
    float WallThermalResistance;
    float FruidThermalConductivity;
    float WallThermalConductivity;
    float FruidSpecificHeat;
    float FruidDensity;
    float Dt;
    int AnalysisStep;
    int ThermalStep;
    int NegativeCalculationStep;
    int OutputStep;
    int LambdaID;
    float SORJacobianRelaxationCoefficient;
    int SORLoopMax;
    float SORConvergenceConstant;
    int bUpdateTemperature;
    float LambdaValue;
    mpsr_float3 CellSize;
    mpsr_float3 MvVelocity;
    float CourantNumber;
    float ViscosityUpdateParameterA;
    float ViscosityUpdateParameterB;
    float MaxWallTemperature;
    float MinWallTemperature;
    float OffSet;
    int ParticleGenerationStep;
    int TotalGenerationStep;
    float Alpha;
    float AlphaLightWeight;
    int MaxNeighborhoodNum;
    int bExplicitCollisonCheck;
    int bExplicitCollisionOut;
    int bImplicitCollisonCheck;
    float CollisionCheckCoefficient;
    float MinDistanceCoefficient;
    int bFreeSurfacePressureUpdate;
    float BasePressure;
    float MaxPressure;
    float Time1;
    int TriggerStep;
    int EndStep;
    float AnalysisArea;
    int bExplicitOut;
    float WallFrictionCoeffieicnt;
    float DragCoefficient;
    int bTwoPhase;
    int StopWallStep;
    float Tolerance;
    float ImplicitSurfaceThreshold;
    int bCGSolver;
    int bUpdateInnerPressure;
    float SurfaceInfluenceRadiusCoefficient;
    float SurfaceThreshold;
    float GradientForce;
    int StartStep;
    float InitialForce;
    float VelocityThreshold;
    int bExplicitPressure;
    float SonicVelocity;
    float ArtificialPressure;
    float r_coeff;
    int bImplicitVelocity;
    int bNJPCGSolver;
    float PressureMax;
    float RLimitCoefficient;
    int bTurbulance;
    float SmagorinskyConstant;
    float FilterWidth;
    float InitialTimeForZeroAngle;
    float FinalTimeToReachAngle;
    float DragPermiabilityConstant;
} mpsr_params;

/* CBox / CGridBox (DataType.h:397-410). */
typedef struct mpsr_box { mpsr_float3 m_MinBound, m_MaxBound; } mpsr_box;
typedef struct mpsr_gridbox {
    mpsr_box m_BufferedZone;
    mpsr_box m_ComputeZone;
    mpsr_box m_PressureZone;
    mpsr_box m_ParentZone;
    mpsr_int3 m_GridSize;
    float m_CellSize;
} mpsr_gridbox;

/* CTriangle (DataType.h:296-301). */
typedef struct mpsr_triangle {
    int ModelIDNumber;
    mpsr_float3 Normal;
    mpsr_float3 Vertex[3];
} mpsr_triangle;

/* CTriangleParameters (DataType.h:327-352), one per STL model. */
typedef struct mpsr_tri_params {
    mpsr_float3 Velocity;
    float Temperature;
    int LinearMovableOrNot;
    int LinearResetInterval;
    int RotatingOrNot;
    float AngleOfRotationInDegree;
    float RotationStartTime;
    float TimeToReachRPM;
    mpsr_float3 CenterOfRotation;
    mpsr_float3 SecondPointOfRotation;
    float WallFrictionCoefficient;
    float WallThermalResistivity;
    float WallThermalConductivity;
    int isDrag;
    int DragTriangleID;
} mpsr_tri_params;

/* DragParameter (DataType.h:302-326), one per drag STL. */
typedef struct mpsr_drag_params {
    int DragID;
    char DragToken;            /* 'E' exponential, 'C' constant velocity, 'F' velocity factor */
    char InputVelocityToken1;  /* 'V', 'T' or 'B' */
    char OutputVelocityToken2;
    float CoefficientofInflueneRegion;
    float VelocityIndex;
    float VelocityExponential;
    float VelocityDragPermiabilityConstant;
    float TempIndex;
    float TempExponential;
    float TempDragPermiabilityConstant;
    float ConstantVx;
    float ConstantVy;
    float ConstantVz;
    float VelocityMagnitueFactor;
} mpsr_drag_params;

/* Host view of the particle state, one row per slot (CParticle, DataType.h:368-390).  Vector
 * fields are xyz-interleaved like the reference's Scalar3 arrays.  Scalars are double on the
 * host side for both builds; an FP32 handle rounds them to float on upload, which is exact for
 * values that came from float. */
typedef struct mpsr_particles {
    int n;
    int *id;
    double *position;          /* [3n] */
    double *velocity;          /* [3n] */
    double *pressure;          /* [n] */
    double *density;           /* [n] particle number density */
    double *temperature;       /* [n] */
    double *kinetic_viscosity; /* [n] */
    double *solid_phase_rate;  /* [n] */
    int *type;                 /* [n] ParticleType: 0 heavy, 1 heavy free surface, 2 light */
} mpsr_particles;

typedef enum mpsr_precision { MPSR_FP32 = 0, MPSR_FP64 = 1 } mpsr_precision;

/* How the pressure of the corrector is obtained.  EOS is the reference's live path
 * (CalcPressureExplicit_Kernel, DeviceProcess_Kernel.h:783-787).  PPE fills the slot the
 * reference declares but never defines (CalcPoisson, DeviceProcess.h:50-53). */
typedef enum mpsr_pressure_mode { MPSR_PRESSURE_EOS = 0, MPSR_PRESSURE_PPE = 1 } mpsr_pressure_mode;

typedef enum mpsr_pcg_kind {
    MPSR_PCG_FP64 = 0,  /* FP64 CSR values, FP64 vectors, Jacobi preconditioner */
    MPSR_PCG_MIXED = 1  /* FP32 CSR values and SpMV, FP64 dots and norms, FP64 refinement */
} mpsr_pcg_kind;

typedef enum mpsr_warm_start {
    MPSR_WARM_ZERO = 0,     /* x0 = 0 */
    MPSR_WARM_PREVIOUS = 1, /* x0 = the previous step's pressure */
    MPSR_WARM_LEARNED = 2   /* x0 and preconditioner from the message-passing network */
} mpsr_warm_start;

typedef enum mpsr_precond {
    MPSR_PRECOND_JACOBI = 0,     /* z = D^-1 r */
    MPSR_PRECOND_TWO_LEVEL = 1,  /* z = D^-1 r + P Ac^-1 P^T r, P = aggregates of cells */
    MPSR_PRECOND_CELL_VCYCLE = 2 /* z = D^-1 r + P1 V(P1^T r): P1 = cells, V = one symmetric V-cycle on
                                    P1^T A P1 (weighted Jacobi, the aggregate correction of
                                    TWO_LEVEL, weighted Jacobi) */
} mpsr_precond;

typedef struct mpsr_ppe_options {
    int solver;        /* mpsr_pcg_kind */
    double rel_tol;    /* stop when ||b - A x||_2 / ||b||_2 <= rel_tol (FP64 residual) */
    int max_iter;      /* per solve, over all refinement passes.  A solve that ends above rel_tol
                          (this cap, the 16-pass refinement cap), breaks down or turns non-finite
                          fails the step with MPSR_CALCERR; no pressure is written back */
    int warm_start;    /* mpsr_warm_start */
    double n0_lap;     /* PND of the ideal lattice at the Laplacian radius; 0 = computed */
    double lambda_lap; /* Laplacian model constant at the Laplacian radius; 0 = computed */
    int precond;       /* mpsr_precond */
    int coarse_cells;  /* aggregate edge in cells for the two-level preconditioner, 1..8; 0 = 2 (the
                          edge grows while the aggregates outnumber 8192) */
    int coarse_rebuild;/* rebuild the coarse operator every this many solves, or sooner when the
                          iteration count grows by a quarter over the solve that rebuilt it (not
                          within 10 solves of a rebuild; a growth rebuild that leaves the
                          iterations within 10 % of the solve before it stops these until the
                          next scheduled one); 0 = 200, 1 = every solve.  In between, the
                          last coarse inverse is applied to the current aggregates: the
                          preconditioner stays symmetric positive definite.  With
                          MPSR_PRECOND_CELL_VCYCLE the cell operator is also rebuilt with the
                          aggregates, every 50 solves, and from 5 solves after its build when a
                          solve takes over 1.2 x the fewest iterations since (stopped the same way
                          when such a rebuild does not help). */
    int learned_omega; /* with MPSR_WARM_LEARNED: 1 = also scale the preconditioner's diagonal part
                          by the network's omega (clamped to [0.5, 1.5]), 0 = warm start only */
    double source_relaxation; /* gamma in the source gamma * alpha * max(n* - n0s, 0) / n0s: the share
                          of the density error corrected per step; 0 = 0.1.  With 1 (all of it per
                          step) the source is stiff enough to amplify PND noise into unphysical
                          motion in the drum at 1M particles and beyond */
} mpsr_ppe_options;

typedef struct mpsr_create_info {
    mpsr_params params;          /* LambdaValue must already be set (mpsr_calculate_lambda) */
    mpsr_box compute_zone;       /* as CMultiGPUHandler::m_ComputeZone (mpsr_auto_compute_zone) */
    int max_particles;           /* slots to allocate, >= particles->n */
    const mpsr_particles *particles;
    const mpsr_triangle *triangles;
    int num_triangles;
    const mpsr_tri_params *models;
    int num_models;
    const mpsr_drag_params *drag;
    int num_drag;
    int precision;               /* mpsr_precision */
    int pressure_mode;           /* mpsr_pressure_mode */
    mpsr_ppe_options ppe;
    int device;
} mpsr_create_info;

/* Wall-clock of the last step's phases, measured with HIP events on the handle's stream. */
typedef struct mpsr_timings {
    double step_ms;
    double bin_ms;        /* both particle binnings: hash + radix sort + cell ranges + gather */
    double tri_ms;        /* triangle binning */
    double stl_ms;        /* both STL distance passes */
    double drag_ms;
    double explicit_ms;
    double relocate_ms;   /* both relocations */
    double pnd_ms;        /* PND + EOS, or PND alone on the PPE path */
    double ppe_ms;        /* PPE assembly + solve (0 on the EOS path) */
    double gradient_ms;
    double wall_ms;       /* translation + rotation */
    int ppe_iterations;   /* PCG iterations of the last solve, summed over refinement passes */
    int ppe_refinements;
    double ppe_rel_residual; /* FP64 ||b - A x|| / ||b|| after the last solve */
    double ppe_initial_rel_residual;
    /* PCG SpMV launches of the last call (profiling on): mean duration from HIP events around
     * each launch, launch count, and algorithmic bytes per launch of its last solve (matrix
     * entries and the row-wise vectors once; neighbour gathers not counted). */
    double spmv_ms;
    int spmv_launches;
    double spmv_bytes;
    long long ppe_nnz;       /* off-diagonal entries of the last PPE matrix */
    int ppe_rows;            /* non-Dirichlet rows of the last PPE matrix */
    int ppe_fallbacks;       /* mixed solves that finished with the plain FP64 PCG */
    int ppe_coarse_rows;     /* aggregates of the two-level preconditioner, 0 = Jacobi was used */
    int ppe_coarse_rebuilt;  /* 1 when this solve rebuilt the coarse operator */
    double ppe_setup_ms;     /* coarse operator set-up of this solve (0 when reused; profiling on) */
    int ppe_learned;         /* 1 when the network gave the warm start and preconditioner weights */
    int ppe_learned_fallback;/* 1 when it lost to the plain solver after 5 iterations (plain restart) */
    double sphere_ms;        /* sphere phase (profiling on) */
    long long ppe_stored;    /* stored entries of the last PPE matrix (SELL padding included) */
    double thermal_ms;       /* CalcTemperatureFactor (profiling on) */
    double turbulence_ms;    /* CalcTurbulenceViscosity (profiling on) */
    /* over all steps of the last mpsr_step / mpsr_step_many call (the fields above that say "the
     * last" describe its last step; the SpMV launch statistics above cover the whole call) */
    int steps;                    /* steps of the call */
    long long ppe_iterations_total; /* PCG iterations summed over its solves */
    int ppe_rebuilds;             /* coarse-operator rebuilds among its solves */
    double ppe_rel_residual_max;  /* the largest final FP64 relative residual of its solves */
} mpsr_timings;

typedef struct mpsr_handle mpsr_handle;

/* ---- set-up helpers (host only, no GPU) ---------------------------------------------------- */

/* CMultiGPUHandler::CalculateLambdaValueLocal (Source/MultiGPUHandler.cpp:1132-1296). */
float mpsr_calculate_lambda(float initial_particle_distance);

/* CMultiGPUHandler::SetComputeZone from particles and triangles (MultiGPUHandler.cpp:111-150). */
mpsr_status mpsr_auto_compute_zone(const mpsr_params *params, const mpsr_particles *particles,
                                   const mpsr_triangle *triangles, int num_triangles,
                                   mpsr_box *out);

/* GetSectionComputeZone(zone, (1,1,1), 0,0,0) then CCellSet::SetComputeZone
 * (solver/Common/inc/Utility_Inline.h:1977-2006, Source/CGPUCellSet.cpp:209-253). */
mpsr_status mpsr_make_gridbox(const mpsr_params *params, const mpsr_box *compute_zone,
                              mpsr_gridbox *out);

/* ---- lifetime ------------------------------------------------------------------------------ */

/* CCellSet::InitializeGPU + InitializeDeviceMemConst (CGPUCellSet.cpp:439-659, DeviceProcess.cu:220-264). */
mpsr_status mpsr_create(const mpsr_create_info *info, mpsr_handle **out);
void mpsr_destroy(mpsr_handle *h);
const char *mpsr_last_error(void);

/* ---- the step ------------------------------------------------------------------------------ */

/* One body of CMultiGPUHandler::CalculateNew's loop (MultiGPUHandler.cpp:362-492) without the
 * output and bucket branches: bin, bin walls, STL distance, drag, predictor, relocate, re-bin,
 * STL distance, pressure (EOS or PPE), corrector, relocate, translate and rotate walls. */
mpsr_status mpsr_step(mpsr_handle *h, int step_index);
/* count consecutive steps first..first+count-1 on the handle's stream, one sync at the end. */
mpsr_status mpsr_step_many(mpsr_handle *h, int first, int count);
mpsr_status mpsr_synchronize(mpsr_handle *h);

/* ---- one launcher per reference launcher (DeviceProcess.h:34-89) ---------------------------- */

mpsr_status mpsr_register_particle_topology(mpsr_handle *h);  /* RegisterParticleTopology :36 */
mpsr_status mpsr_register_triangle_topology(mpsr_handle *h);  /* RegisterTriangleTopology :54 */
mpsr_status mpsr_calc_stl_distance(mpsr_handle *h);           /* CalcSTLDistance :62 */
mpsr_status mpsr_calc_drag_effect(mpsr_handle *h);            /* CalcDragEffect :88 */
mpsr_status mpsr_calc_explicitly(mpsr_handle *h);             /* CalcExplicitly :48 */
mpsr_status mpsr_relocate_particle_data(mpsr_handle *h);      /* RelocateParticleData :64 */
mpsr_status mpsr_calc_pressure_explicit(mpsr_handle *h);      /* CalcPressureExplicit :70 */
mpsr_status mpsr_calc_pressure_poisson(mpsr_handle *h);       /* CalcPoisson :50 (declared only there) */
mpsr_status mpsr_calc_implicit_explicitly(mpsr_handle *h);    /* CalcImplicitExplicitly :49 */
mpsr_status mpsr_update_triangle_position(mpsr_handle *h);    /* UpdateTrianglePosition :58 */
mpsr_status mpsr_calc_temperature_factor(mpsr_handle *h);     /* CalcTemperatureFactor :60 */
mpsr_status mpsr_rotate_triangle_position(mpsr_handle *h, int analysis_step); /* RotateTrianglePosition :65 */

/* ---- state in and out ---------------------------------------------------------------------- */

/* which = 0: the "in" arrays (c_daParticle*), 1: the "out" arrays (c_daOutputParticle*).
 * = CCellSet::ParticleDeviceToHost (CGPUCellSet.cpp:833-892). */
mpsr_status mpsr_get_particles(mpsr_handle *h, int which, mpsr_particles *out);
/* = ParticleHostToDevice + ParticleNumberToConst (CGPUCellSet.cpp:893-952, DeviceProcess.cu:265-271). */
mpsr_status mpsr_set_particles(mpsr_handle *h, const mpsr_particles *in);
mpsr_status mpsr_get_triangles(mpsr_handle *h, mpsr_triangle *out, int n);
/* Rigid spheres carried by the fluid (no counterpart in the reference): diameter d, density
 * rho_s, Schiller-Naumann drag F = 3 pi mu d (u_f - u_s)(1 + 0.15 Re^0.687) (Re < 1000, else
 * Cd = 0.44) with u_f the MPS-weighted fluid velocity within re_lap of the centre and
 * mu = fluid density (params.Density) x fluid_viscosity.  The reaction -F is spread over the fluid
 * particles in reach with the same weights (momentum-conserving), the spheres move under drag,
 * gravity and buoyancy and are kept d/2 off the non-drag walls.  Runs every step between the
 * drag and the predictor.  position/velocity/force: [3n], xyz interleaved. */
typedef struct mpsr_spheres {
    int n;
    double *position;
    double *velocity;
    double *force;          /* output only: the drag of the last step */
    double diameter;
    double density;
    double fluid_viscosity;
} mpsr_spheres;
mpsr_status mpsr_set_spheres(mpsr_handle *h, const mpsr_spheres *s);
/* Fills the non-NULL arrays of out for min(out->n, sphere count) spheres; out->n = sphere count. */
mpsr_status mpsr_get_spheres(mpsr_handle *h, mpsr_spheres *out);
/* The sphere phase alone, for drivers that call the phases one by one. */
mpsr_status mpsr_calc_spheres(mpsr_handle *h);
/* Current slot count (m_ParticleNum): grows when bucket particles are appended. */
mpsr_status mpsr_get_particle_count(mpsr_handle *h, int *n);
/* Bucket inflow = the reference's multi-bucket set-up (LoadMultiBucket,
 * Src/MultiGPUHandler.cpp:1068-1083): bucket j (buckets[j], slot rows as in mpsr_create) is added
 * after step i when (i + 1) % interval[j] == 0, at most total[j] times.  An addition
 * (BucketAddition + InsertBucket, MultiGPUHandler.cpp:943-981, 1085-1123) fills the first dead
 * slots (ID < 0) in ascending slot order, appends the rest after the last slot while the capacity
 * given at mpsr_create (max_particles) allows and drops it otherwise; ID = slot index, the pressure
 * of a reused slot is kept and that of an appended slot is 0.  In a step with an addition the
 * walls of linearly moving models are reset (ResetWallPosition_Kernel, Kern :913-925), the
 * particles are relocated again and the wall translation is skipped, as in CalculateNew
 * (MultiGPUHandler.cpp:462-484).  Replaces any earlier buckets and restarts their counters. */
mpsr_status mpsr_set_buckets(mpsr_handle *h, const mpsr_particles *buckets, const int *interval,
                             const int *total, int count);
/* The addition of one step alone, for drivers that call the phases one by one
 * (CMultiGPUHandler::BucketAddition, MultiGPUHandler.cpp:1085-1123); *added = 1 when a bucket
 * was due. */
mpsr_status mpsr_bucket_addition(mpsr_handle *h, int step, int *added);
mpsr_status mpsr_get_gridbox(mpsr_handle *h, mpsr_gridbox *out);
/* = CheckAndUpdateDt's constant upload (CGPUCellSet.cpp:1278). */
mpsr_status mpsr_set_dt(mpsr_handle *h, float dt, float courant);
mpsr_status mpsr_get_timings(mpsr_handle *h, mpsr_timings *out);
mpsr_status mpsr_set_ppe_options(mpsr_handle *h, const mpsr_ppe_options *opt);

/* ---- inspection of intermediate results (parity tests) ---------------------------------------
 * Binning after the last mpsr_register_particle_topology: sorted keys and sorted-to-slot
 * permutation ([n] each; key == cell count marks a slot that is not binned), and cellStart /
 * cellEnd ([cells] each, -1 = empty), as radix DeviceProcess_Kernel.h:39-124. */
mpsr_status mpsr_get_binning(mpsr_handle *h, int *sorted_keys, int *perm, int *cell_start,
                             int *cell_end);

/* ---- binning that allocates nothing ---------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the binning of
 * mpsr_register_particle_topology (cell key per slot, stable radix sort, cellStart/cellEnd), with
 * the same results as mpsr_get_binning returns after it, on caller-owned device arrays.
 *   gb          the handle's grid (mpsr_make_gridbox / mpsr_get_gridbox)
 *   id, px..pz  device arrays [n]: particle id (< 0 = dead) and FP32 position, slot order
 *   sorted_keys device [n]: cell key per sorted index (the cell count marks a slot not binned)
 *   perm        device [n]: sorted index -> slot
 *   cell_start, cell_end  device [cell_capacity >= cells of gb]: first / end sorted index, -1 empty
 *   stream      a hipStream_t (NULL: the null stream); the call only queues work on it
 * The sort's scratch is a fixed device area inside the library (MPSR_NOALLOC_MAX_PARTICLES
 * particles, about 9 MB per GPU, held until the process ends): n above it is MPSR_PARAMERR.  No
 * hipMalloc, hipFree or other allocation happens in this call.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU race. */
#define MPSR_NOALLOC_MAX_PARTICLES (1 << 20)
mpsr_status mpsr_register_particle_topology_noalloc(const mpsr_gridbox *gb, int n, const int *id,
                                                    const float *px, const float *py, const float *pz,
                                                    unsigned int *sorted_keys, int *perm, int *cell_start,
                                                    int *cell_end, int cell_capacity, int device,
                                                    void *stream);
/* Triangle lists per cell after mpsr_register_triangle_topology: start/count per cell and the
 * ascending triangle ids (capacity cap; *total receives the entry count). */
mpsr_status mpsr_get_triangle_cells(mpsr_handle *h, int *cell_start, int *cell_count,
                                    int *tri_ids, int cap, int *total);
/* Nearest-wall results of the last mpsr_calc_stl_distance, slot order: magnitude [n],
 * direction [3n], triangle id [n] (CDistance / STLID, DeviceProcess_Kernel.h:635-699). */
mpsr_status mpsr_get_stl_distance(mpsr_handle *h, double *magnitude, double *direction,
                                  int *triangle_id);
/* Neighbour lists in the canonical order (stencil z, y, x with x innermost, then ascending
 * sorted index) for radius r in units of l0: offsets [n+1] and neighbour slots (cap). */
mpsr_status mpsr_get_neighbors(mpsr_handle *h, double radius_over_l0, int *offsets,
                               int *neighbor_slots, int cap, int *total);
/* The last PPE system in slot order: rhs [n], solution [n], Dirichlet flag [n]. */
mpsr_status mpsr_get_ppe(mpsr_handle *h, double *rhs, double *solution, int *dirichlet);
/* The last PPE matrix in slot order: row pointers [n+1], column slots and off-diagonal values
 * (capacity cap), diagonal [n]; *nnz receives the off-diagonal count. */
mpsr_status mpsr_get_ppe_matrix(mpsr_handle *h, int *row_ptr, int *col, double *val, double *diag,
                                int cap, int *nnz);
/* The inverse of a dense symmetric positive definite matrix (n x n, 1 <= n <= 16384, host
 * arrays, column-major) by the blocked FP64 Gauss-Jordan sweep that inverts the V-cycle's coarse
 * operator.  No counterpart in the reference (its solver has no PPE); it exposes that sweep for
 * checking.  MPSR_CALCERR when a pivot is not positive (the matrix is not SPD). */
mpsr_status mpsr_spd_inverse(int n, const double *a, double *inv);
/* HIP-event timing of the following steps (mpsr_get_timings); off (0) by default.  1: every phase
 * and every PPE SpMV launch.  2: the PPE phase and the SpMV launches of every fourth PCG chunk
 * (each event delays the stream by a few microseconds; 2 keeps timing runs close to untimed ones). */
mpsr_status mpsr_set_profiling(mpsr_handle *h, int on);

/* ---- turbulence and negative pressure (the radix solver variant's terms) ------------------------
 * Fields of the radix variant's CParameter (radix/SPVSOLVER/MultiGPUSolver/DataType.h:510-520,
 * read by IO.cpp:393-431), in their order.  This path's CParameter carries bTurbulance,
 * SmagorinskyConstant and FilterWidth too (mpsr_params), read from the parameter file but used by
 * none of its kernels; the terms run only when set here.  Zeroed = both terms off. */
typedef struct mpsr_radix_terms {
    int bTurbulance;              /* Smagorinsky viscosity added to the particle viscosity */
    float SmagorinskyConstant;
    float FilterWidth;
    float TurbulanViscosityRatio; /* upper clamp of the turbulent viscosity */
    float Cnue;
    float Cepsilon;
    float TurbulanceIntensity;    /* upper clamp of the turbulent kinetic energy */
    int bNegativePressure;        /* gas-pressure term near rotating or flagged walls */
} mpsr_radix_terms;

/* Per-model fields of the radix CTriangleParameters (DataType.h:160-168, Configuration.cpp:
 * 346-360, MultiGPUHandler.cpp:482-527) that this path's mpsr_tri_params lacks. */
typedef struct mpsr_negative_pressure_model {
    float RotationInRPM;          /* the term applies when != 0 or NegativePressureCheck */
    int NegativePressureCheck;
    float InfluenceRegion;        /* in units of InitialParticleDistance */
    mpsr_float3 DirectionVector;
    float PressureGas;
} mpsr_negative_pressure_model;

/* Sets both terms for the following steps; models may be NULL when bNegativePressure is 0, else
 * it holds num_models entries (the handle's model count).  With bTurbulance the step runs
 * mpsr_calc_turbulence_viscosity after the temperature phase and the predictor multiplies its
 * Laplacian by (nu + nu_t) (radix DeviceProcess_Kernel.h:545-552, MultiGPUHandler.cpp:1421-1427);
 * with bNegativePressure the gradient corrector adds the gas-pressure term of the nearest wall
 * (radix DeviceProcess_Kernel.h:1110-1134, 1214-1245). */
mpsr_status mpsr_set_radix_terms(mpsr_handle *h, const mpsr_radix_terms *terms,
                                 const mpsr_negative_pressure_model *models, int num_models);
/* CalcTurbulenceViscosity (radix DeviceProcess.h:80, DeviceProcess_Kernel.h:1273-1476): the
 * turbulent viscosity and kinetic energy of every particle from the current binning. */
mpsr_status mpsr_calc_turbulence_viscosity(mpsr_handle *h);
/* The last mpsr_calc_turbulence_viscosity results in slot order, n rows each (either may be NULL):
 * c_daParticleTurbulaceViscosity and c_daParticleStrainTensorProduct. */
mpsr_status mpsr_get_turbulence(mpsr_handle *h, double *viscosity, double *kinetic_energy, int n);

/* Launch shape of the per-particle kernels (block: a multiple of 64 in [64, 512]) and of the
 * particle sort (tile = 256 * sort_items keys, sort_items 4, 8 or 16).  Results do not depend on
 * it.  No counterpart in the reference (its launchers fix 256 threads, DeviceProcess.cu:29-304). */
mpsr_status mpsr_set_launch(mpsr_handle *h, int block, int sort_items);

/* ---- batched buckets (no counterpart in the reference) ------------------------------------------
 * Many independent domains ("buckets": each its own particles, walls, grid and precision) advanced
 * together.  The reference steps one domain per CalculateNew call (Src/MultiGPUHandler.cpp:
 * 362-492); here every kernel of a step is one launch over a work list of (bucket, block) items
 * of all buckets of a precision, and the step is recorded once into a HIP graph (about 46
 * launches per precision whatever the number of buckets), so each step is one graph launch.
 * A bucket uses the explicit (EOS) pressure, or the PPE with the FP64 Jacobi-PCG (solver
 * MPSR_PCG_FP64, preconditioner MPSR_PRECOND_JACOBI, not the learned start); the PCG of the PPE
 * buckets runs in chunks of iterations that return at once after convergence, so a step with PPE
 * buckets is three graph launches (up to the PCG with its first 32 iterations, 16-iteration
 * chunks until every bucket's PCG has stopped, the rest of the step) with one flag read per
 * chunk.  Buckets carry no spheres or bucket inflow; other PPE solvers, spheres and inflow are
 * rejected with MPSR_PARAMERR naming the bucket; every bucket must name the same device.  A
 * bucket's results are those of mpsr_step_many on the same handle, bit for bit.  A failed PPE
 * solve (breakdown, not converged, more than 192 matrix entries a row) is the call's
 * MPSR_CALCERR, naming the bucket; that bucket's pressure is then not written back. */
typedef struct mpsr_batch mpsr_batch;
#define MPSR_TUNE_BLOCKS 4   /* blocks 64, 128, 256, 512 */
#define MPSR_TUNE_ITEMS 3    /* sort items 4, 8, 16 */
mpsr_status mpsr_create_batch(const mpsr_create_info *infos, int nbuckets, mpsr_batch **out);
void mpsr_destroy_batch(mpsr_batch *b);
/* The handle of bucket i for the per-domain calls (particles, triangles, timings...); owned by
 * the batch: mpsr_destroy on it does nothing. */
mpsr_status mpsr_batch_handle(mpsr_batch *b, int bucket, mpsr_handle **h);
/* Steps first .. first+count-1 of every bucket: one graph launch per step (the graph is recorded
 * on the first call and again after a launch-shape change). */
mpsr_status mpsr_step_batch(mpsr_batch *b, int first, int count);
/* The same steps with the same launches made directly (no graph). */
mpsr_status mpsr_step_batch_eager(mpsr_batch *b, int first, int count);
mpsr_status mpsr_batch_set_launch(mpsr_batch *b, int block, int sort_items);
/* Times `steps` graph-launched steps from `first` for every launch shape (MPSR_TUNE_BLOCKS x
 * MPSR_TUNE_ITEMS, ms per step into ms_table row-major when non-NULL), each from the same saved
 * state, restores that state and keeps the fastest shape. */
mpsr_status mpsr_batch_autotune(mpsr_batch *b, int first, int steps, int *best_block, int *best_items,
                                float *ms_table);
/* ms per step of the last mpsr_step_batch(_eager) call, nodes of the recorded graph, launch shape. */
mpsr_status mpsr_batch_info(mpsr_batch *b, float *ms_per_step, int *graph_nodes, int *block, int *sort_items);

/* ---- learned PPE warm start and preconditioner (no counterpart in the reference) ---------------
 * A three-layer message-passing network over the PPE matrix graph (hidden width 128, message
 * width 32, 16 input features), BF16 operands on MFMA with FP32 accumulation.  With
 * ppe.warm_start = MPSR_WARM_LEARNED and weights set, every solve starts from x0 = s y0 and
 * scales the diagonal part of the preconditioner by omega = 2 sigmoid(y1); after 5 iterations a
 * residual larger than the plain solver's after 5 iterations (plain start, plain preconditioner,
 * measured every 16 solves) restarts the solve the plain way (timings.ppe_learned_fallback).
 * Parameters: mpsr_mpnn_param_count() FP32 values, flat, order U1 W1 V1 c1 U2 W2 V2 c2 U3 W3 V3
 * c3 (row-major; U: F x 32, W: F x out, V: 32 x out, c: out).  The handle must be in PPE mode. */
int mpsr_mpnn_param_count(void);
mpsr_status mpsr_mpnn_set_weights(mpsr_handle *h, const float *w, int n);
mpsr_status mpsr_mpnn_get_weights(mpsr_handle *h, float *w, int n);
/* One training sample from the last solve (its system, its solution as the target): loss parts
 * {total, supervised, residual, preconditioner} into parts[4], gradients into grads (may be
 * NULL), then one Adam step of rate lr when update != 0. */
mpsr_status mpsr_mpnn_train_step(mpsr_handle *h, double lr, double beta, int update, float *grads,
                                 double *parts);
/* Features [n][16] and outputs [n][2] of the network on the last system, slot order, and the
 * pressure scale s. */
mpsr_status mpsr_mpnn_rows(mpsr_handle *h, float *features, float *outputs, double *scale);

/* ---- input decks and output frames (host only, no device use) --------------------------------
 * A deck is the reference's configuration file and the files it names, parsed as
 *   CConfiguration::ReadConfigurationFileWithDrag (Src/Configuration.cpp:21-421): input directory,
 *     parameter file, particle file, bucket lines, STL lines (moving walls or drag walls D/DE/DC/DF),
 *     output directory and name, output interval, compute zone, section division, start step,
 *     ASCII/BINARY output;
 *   IO::Load of the positional parameter file with its defaults (Src/IO.cpp:66-432);
 *   IO::Load of the 12-column particle file (Src/IO.cpp:7-64);
 *   STL models, ASCII or binary (solver/Common/inc/STLSerializer.cpp:226-351), normals recomputed
 *     as CalcTriangleNormal (Src/IO.cpp:435-453), ModelIDNumber = model index;
 *   the model set-up of CMultiGPUHandler::Run (Src/MultiGPUHandler.cpp:640-786): model offsets
 *     added to the vertices, degrees per step = RPM * Dt * 6, lambda from l0, compute zone from
 *     the configuration or computed from particles and triangles, capacity = particles + the sum
 *     over bucket lines of bucket particles * total additions (the multi-bucket build,
 *     Src/MultiGPUHandler.cpp:758-765).
 * More than one section is rejected (MPSR_PARAMERR): one GPU holds the whole domain. */
typedef struct mpsr_deck mpsr_deck;
mpsr_status mpsr_deck_load(const char *config_path, mpsr_deck **out);
void mpsr_deck_free(mpsr_deck *d);
const char *mpsr_deck_last_error(void);
/* Fills *info for mpsr_create (FP32, EOS pressure); its pointers point into the deck, which must
 * outlive the mpsr_create call. */
mpsr_status mpsr_deck_create_info(const mpsr_deck *d, mpsr_create_info *info);
/* Bucket lines (LoadMultiBucket, Src/MultiGPUHandler.cpp:1068-1083): particle count, addition
 * interval in steps and total number of additions of bucket j. */
int mpsr_deck_num_buckets(const mpsr_deck *d);
mpsr_status mpsr_deck_bucket(const mpsr_deck *d, int j, int *count, int *interval, int *total);
/* Rows of bucket j into caller arrays of out->n >= count slots (out->n is set to count). */
mpsr_status mpsr_deck_get_bucket(const mpsr_deck *d, int j, mpsr_particles *out);
int mpsr_deck_start_step(const mpsr_deck *d);
int mpsr_deck_output_interval(const mpsr_deck *d);
int mpsr_deck_output_ascii(const mpsr_deck *d);
const char *mpsr_deck_output_name(const mpsr_deck *d);
/* Per-model offsets from the configuration (cap entries); returns the model count. */
int mpsr_deck_model_offsets(const mpsr_deck *d, mpsr_float3 *out, int cap);
/* One output frame as IO::SaveAscii / IO::SaveBinary (Src/IO.cpp:635-742): particle count,
 * 12 values per particle (position, velocity, pressure, density, temperature, viscosity, solid
 * rate, type), time, model count, per model position and rotation angle. */
mpsr_status mpsr_write_frame(const char *path, int ascii, const mpsr_particles *p, float time,
                             int nmodels, const mpsr_float3 *model_pos, const float *angles);

/* One record of the time-step schedule file (DeltaTime, Source/DtFrame.cpp:11-35): from step
 * `step` on, Dt, CourantNumber, ParticleGenerationStep and OutputStep take these values. */
typedef struct mpsr_dt_entry {
    int step;
    float dt;
    float courant;
    int inflow_step;
    int output_step;
} mpsr_dt_entry;
/* DtFrame::Load: reads the records of `path` in file order into out[0 .. cap-1] and their number
 * into *count (which may exceed cap).  As the reference's stream reading does, a record counts only
 * when the stream is still good after its last value (a last record ending the file without a line
 * break is dropped) and the first unreadable value ends the list.  The driver applies one due record
 * per step, as CMultiGPUHandler::CheckAndUpdateDt (Source/MultiGPUHandler.cpp:1048-1066), through
 * mpsr_set_dt. */
mpsr_status mpsr_dt_schedule_load(const char *path, mpsr_dt_entry *out, int cap, int *count);

#ifdef __cplusplus
}
#endif
#endif /* MPSR_H */
