/*
 * mph_abi.h -- C ABI of the MI355X MPH-I time stepper (libmph_hip.so).
 *
 * The reference (Masahiro-Kondo-AIST/MphImplicit) has no library interface: its hot
 * path is a set of file-static functions over global arrays called in a fixed order
 * from main() (source/main.cpp:174-211 prototypes, 386-474 time loop).  This header
 * is the drop-in boundary that replaces that path: one call per reference entry
 * point or phase, plain pointers and sizes, status codes instead of exit()/assert.
 *
 * Conventions
 *  - Every function returns MPH_OK (0) or a negative MPH_E_* code; the message of
 *    the last failure on the calling thread is available from mph_last_error().
 *  - A scene owns all of its device state; host arrays passed in are copied.
 *  - A scene handle is not thread-safe.  Distinct scenes may be driven from
 *    distinct host threads.  Each scene runs on its own HIP stream.
 *  - Arrays read back are in the current sorted particle order, exactly as the
 *    reference keeps them after calculateCellParticle (source/main.cpp:1315-1378).
 */
#ifndef MPH_ABI_H
#define MPH_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPH_TYPE_COUNT 4 /* types 0,1 fluid; 2,3 wall (source/main.cpp:52-56) */

#define MPH_OK 0
#define MPH_E_ARG (-1)     /* invalid argument or size                         */
#define MPH_E_IO (-2)      /* file could not be opened or parsed               */
#define MPH_E_DEVICE (-3)  /* HIP runtime error                                */
#define MPH_E_NOMEM (-4)   /* device or host allocation failed                 */
#define MPH_E_STATE (-5)   /* scene state the reference cannot handle either   */
#define MPH_E_FIELD (-6)   /* unknown field name                               */
#define MPH_E_NOTIMPL (-7) /* mode not available in this build                 */

/* Dot-product summation order used by the conjugate-residual solver. */
#define MPH_DOT_TREE 1 /* fixed two-level tree over 256-particle chunks */

/* Precision of the solver's inner products with the particle matrix. */
#define MPH_PRECISION_FP64 0  /* plain FP64, bit-compatible with the reference   */
#define MPH_PRECISION_MIXED 1 /* FP32 SpMV + multigrid inside FP64 refinement    */

/* How the solve applies the particle matrix (mph_params.operator_form).
 * ASSEMBLED stores the 3x3 blocks of A and multiplies by them, the reference's CSR
 * product (main.cpp:4033-4050), bit for bit.
 * FACTORED multiplies by A = viscous part + C Lambda C^T, rebuilt from one scalar per
 * fluid-neighbour pair and per RadiusP pair plus the positions.  A product moves about
 * a fifth of the assembled form's bytes (0.21 on a 1M-particle 3D scene) and equals
 * the assembled product up to rounding, not bit for bit.  Such a scene assembles no
 * 3x3 blocks: the multigrid level 0 and InvDiagA come from the same scalars, except
 * when the level-0 grid spans fewer than 3 cells along an axis (scalar
 * "AssembledBlocks" = 1 then); reading CsrCofA otherwise fails with MPH_E_STATE.
 * A batch may hold scenes of both forms. */
#define MPH_OPERATOR_ASSEMBLED 0
#define MPH_OPERATOR_FACTORED 1

/*
 * Parameters of a run: the keywords of the reference's *.data file
 * (readDataFile, source/main.cpp:523-563) plus the compile-time switches of the
 * reference turned into run-time fields (TWO_DIMENSIONAL, MULTIGRID_SOLVER,
 * the CR stop rule rr/rr0 < 1e-12 at source/main.cpp:4346).
 */
typedef struct mph_params {
    int dim; /* 2 (TWO_DIMENSIONAL) or 3 */
    double dt;
    double output_interval;
    double vtk_output_interval;
    double end_time;
    double radius_ratio_a;
    double radius_ratio_p;
    double radius_ratio_v;
    double density[MPH_TYPE_COUNT];
    double bulk_modulus[MPH_TYPE_COUNT];
    double bulk_viscosity[MPH_TYPE_COUNT];
    double shear_viscosity[MPH_TYPE_COUNT];
    double surface_tension[MPH_TYPE_COUNT];
    double interaction_ratio[MPH_TYPE_COUNT][MPH_TYPE_COUNT];
    double gravity[3];
    double wall_center[MPH_TYPE_COUNT][3];
    double wall_velocity[MPH_TYPE_COUNT][3];
    double wall_omega[MPH_TYPE_COUNT][3];
    /* solver controls; mph_params_default() reproduces the reference build */
    double cr_tolerance; /* stop when rr/rr0 < cr_tolerance (reference: 1e-12)  */
    int multigrid;       /* 1: MG-preconditioned CR (MULTIGRID_SOLVER), 0: CR    */
    int precision;       /* MPH_PRECISION_*                                      */
    int max_iterations;  /* <=0: 3*fluid count, as the reference                 */
    int operator_form;   /* MPH_OPERATOR_*                                       */
} mph_params;

/* Particles and domain, the content of a *.grid / *.prof file
 * (readGridFile, source/main.cpp:565-694). */
typedef struct mph_grid {
    double time;
    int np;
    double spacing;
    double domain_min[3];
    double domain_max[3];
    const int* type;   /* [np]     */
    const double* pos; /* [np][3]  */
    const double* vel; /* [np][3]  */
} mph_grid;

/* Per-call statistics of mph_scene_step. */
typedef struct mph_step_stats {
    int steps_done;
    int cr_iterations; /* of the last step (source/main.cpp:4381 "iter")          */
    int cr_iterations_total;
    double rr0, rr, rs0, rs; /* CR summary of the last step (main.cpp:4381)         */
    int vtk_steps;           /* steps on which the VTK diagnostics were computed    */
    double ms_phase[12];     /* summed device time per MPH_PHASE_* (if timing on)   */
    double ms_total;
    int status;              /* MPH_OK, or why this batch member failed (batches)  */
} mph_step_stats;

/* Phases of one time step, in the order of source/main.cpp:386-474. */
#define MPH_PHASE_KINEMATICS 1  /* calculateConvection/Wall/PeriodicBoundary, resetForce */
#define MPH_PHASE_CELLS 2       /* calculateCellParticle                                  */
#define MPH_PHASE_NEIGHBOR 3    /* calculateNeighbor                                      */
#define MPH_PHASE_EXPLICIT 4    /* DensityA ... calculateAcceleration                     */
#define MPH_PHASE_MATRIX_A 5    /* calculateMatrixA                                       */
#define MPH_PHASE_MATRIX_C 6    /* calculateMatrixC                                       */
#define MPH_PHASE_MULTIPLY_C 7  /* multiplyMatrixC                                        */
#define MPH_PHASE_MULTIGRID 8   /* calculateMultiGridMatrix                               */
#define MPH_PHASE_SOLVE 9       /* solveWithConjugatedGradient                            */
#define MPH_PHASE_DIAGNOSTICS 10 /* DivergenceP, PressureP, ViscosityV, VirialStress      */
#define MPH_PHASE_ADVANCE 11    /* Time += Dt and output bookkeeping                      */

/* Parameters equal to the reference build's defaults: 3D, all .data fields zero
 * except Dt=1e100 (source/main.cpp:75), tolerance 1e-12, multigrid on, FP64. */
void mph_params_default(mph_params* p);

/* readDataFile (source/main.cpp:523-563): keyword parser, unknown lines ignored. */
int mph_read_data_file(const char* path, mph_params* p);

/* readGridFile (source/main.cpp:565-694), in two calls: the header gives np, then
 * the caller supplies arrays of np entries. */
int mph_read_grid_header(const char* path, int* np);
int mph_read_grid_file(const char* path, int np, double* time, double* spacing,
                       double domain_min[3], double domain_max[3], int* type,
                       double* pos, double* vel);

typedef struct mph_scene mph_scene;

/* Create a scene on HIP device `device` and run the reference's initialisation
 * (initializeWeight/Fluid/Wall/Domain and the first cell sort, neighbour search and
 * densities, source/main.cpp:353-375). */
int mph_scene_create(const mph_params* p, const mph_grid* g, int device, mph_scene** out);

/* Advance nsteps iterations of the reference's time loop (source/main.cpp:386-474),
 * including the output-step diagnostics on VTK steps.  `st` may be NULL.
 * MPH_E_STATE when the CR residual is not finite (NaN or infinity in the system):
 * the reference would iterate to its cap of 3*Nf (main.cpp:4344-4346) instead. */
int mph_scene_step(mph_scene* s, int nsteps, mph_step_stats* st);

/* Run one MPH_PHASE_* of the time step on the current state (parity tests). */
int mph_scene_run_phase(mph_scene* s, int phase);

/* Copy a named field to or from the host.  Sizes are in bytes and must equal the
 * field's current size (mph_scene_field_size).  Field names follow the
 * reference's globals:
 *   particles   Position Velocity Property ParticleIndex Force Mass Kappa Lambda Mu
 *               DensityA GravityCenter PressureA VolStrainP DivergenceP PressureP
 *               VirialPressureAtParticle VirialStressAtParticle
 *   cells       CellIndex CellParticle
 *   neighbours  NeighborCount NeighborFluidCount NeighborCountP NeighborPtr
 *               NeighborPtrP NeighborInd NeighborIndP
 *   matrices    CsrPtrA CsrIndA CsrCofA VectorB CsrPtrC CsrIndC CsrCofC VectorP
 *               InvDiagA, the multigrid levels MGCofA_L<l> / MGInvD_L<l>, and the layout of
 *               A on the device as ACol / AVal
 * Scalars (mph_scene_scalar): Time iStep Dt OutputNext VtkOutputNext VtkStep N0a N0p
 * R2g Swa Swg Swp Swv MaxRadius ParticleCount FluidCount PowerParticleCount
 * CellCount0..2 TotalCellCount NeighborIndCount NeighborIndCountP MultiGridDepth
 * CRIterations rr0 rr rs0 rs RefinementSteps ABlocks ASlots CBlocks, the timing
 * scalars of mph_scene_set_timing, TunedSpmv64 TunedSpmv32 TunedCac TunedFuse
 * TunedGalerkin (the variants the run-time tuner kept, -1 while still exploring),
 * and for the factored operator FactoredBytes and FactoredBytesF32 (algorithmic
 * bytes of one product with FP64 / FP32 scalars) and AssembledBlocks (1 when the
 * step assembled the 3x3 blocks: always for the assembled form; for the factored
 * form only when its multigrid level 0 spans fewer than 3 cells along an axis,
 * otherwise level 0 and InvDiagA come from the factored scalars and CsrCofA is not
 * available), FactoredCellsUsed (1 when the FP64 second pass of the factored
 * product runs staged per cell, option FactoredCells), and NeighborOneScan (1 when
 * the last neighbour search built the lists in one scan, 0 when it ran the count and
 * fill passes: a scene's first search, or a list longer than the last search's
 * largest plus a margin; the lists are the same either way, alone or in a batch). */
int mph_scene_field_size(mph_scene* s, const char* field, size_t* bytes, int* elem_bytes);
int mph_scene_read(mph_scene* s, const char* field, void* host, size_t bytes);
int mph_scene_write(mph_scene* s, const char* field, const void* host, size_t bytes);

/* Scalars of the scene: "Time", "Dt", "ParticleCount", "FluidCount", ... */
int mph_scene_scalar(mph_scene* s, const char* name, double* value);

/* Particle spacing and domain box of the scene (the *.grid header). */
int mph_scene_domain(mph_scene* s, double* spacing, double domain_min[3], double domain_max[3]);

/* writeProfFile (source/main.cpp:696-720) and writeVtkFile (722-883). */
int mph_scene_write_prof(mph_scene* s, const char* path);
int mph_scene_write_vtk(mph_scene* s, const char* path);

/* The same two writers over caller arrays (host only, no device involved):
 * byte-identical output for identical values. */
int mph_write_prof_arrays(const char* path, double time, int n, double spacing, const double domain_min[3],
                          const double domain_max[3], const int* type, const double* pos, const double* vel);
int mph_write_vtk_arrays(const char* path, int n, const int* type, const double* pos, const double* vel,
                         const double* force, const double* vol_strain_p, const double* divergence_p,
                         const double* pressure_p, const double* virial_pressure, const int* neighbor_count);

/* Timing with HIP events on the scene stream.  Bit 0: per-phase device time
 * (adds one sync per phase, summed into mph_step_stats.ms_phase).  Bit 1: every
 * launch of the particle-matrix SpMV is bracketed by events; the summed time and
 * the launch count are the scalars "SpmvMs" and "SpmvLaunches" (FP64 matrix) and
 * "SpmvMsF32" and "SpmvLaunchesF32" (FP32 matrix, mixed precision only). */
int mph_scene_set_timing(mph_scene* s, int on);

/* Options of a scene.  "ConvergenceCheck" is the reference's compile-time switch
 * CONVERGENCE_CHECK (source/main.cpp:46, 4348-4351, 4383-4397):
 *   "ConvergenceCheck"  1: each FP64 solve keeps rr and r.s of every CR iteration that
 *                  did not stop (field "ConvergenceHistory", CRIterations rows of
 *                  two doubles, the values of main.cpp:4350's log line; at most 2^22
 *                  rows) and, after the loop, r.r and r.s of the true residual
 *                  r = b - A x with s = M^-1 r, or s = b without multigrid as
 *                  main.cpp:4391 has it (scalars "CheckRR", "CheckRS").  It adds one
 *                  reduction per iteration and a product and a V-cycle per solve;
 *                  the solution is unchanged.  0 (default): nothing is kept.
 *                  MPH_E_STATE for a mixed-precision scene or a scene in a batch.
 * The others are schedule options (no counterpart in the reference); they never
 * change a result, only which of two bit-identical kernels runs.
 *   "LaneRowsMax"  multigrid levels with fewer rows split each row of a smoothing
 *                  sweep over 4 lanes, larger ones take one thread per row
 *                  (default 32768).  Applies from the next multigrid build.
 *   "FactoredCells"     1: the FP64 factored product's second pass runs one
 *                  workgroup per cell with the visited cells' records staged in
 *                  LDS; 0: the chunk kernel, every neighbour gathered through the
 *                  caches; 2 (default): the cells in 3D scenes and in scenes of at
 *                  most 16384 fluid particles, the chunk kernel otherwise (larger 2D
 *                  scenes, whose cells hold too few rows for a workgroup).  Alone
 *                  and in a batch; the same bits either way.
 *   "FactoredCellsXcd"  1 (default): the cells of that pass are dealt to the XCDs in
 *                  contiguous ranges; 0: round robin.
 *   "CrStop"       where the FP64 CR loop tests rr/rr0 < tol (main.cpp:4344-4346):
 *                  0 on the host, reading rr back every iteration; 1 (default) on
 *                  the device, in the r.r reduction, with the host one iteration
 *                  ahead (the kernels of an iteration enqueued after the stop return
 *                  at once); 2 as 1, with iterations 1, 2, ... replayed from captured
 *                  HIP graphs.  The same iterations and bits in every mode.  Scalars
 *                  "CrGraphLaunches", "CrCaptures", "CrInstantiations", "CrCaptureMs"
 *                  and "CrGraphError" count the graph work of mode 2.
 * MPH_E_FIELD for an unknown name, MPH_E_ARG for a value out of range. */
int mph_scene_set_option(mph_scene* s, const char* name, double value);

/* y = A x with the particle matrix of the last assembly, in the scene's operator
 * form (myDcsrmvForA, source/main.cpp:4033-4050, as called with alpha 1, beta 0).
 * x and y are host arrays of 3*FluidCount doubles in the current particle order.
 * precision MPH_PRECISION_MIXED applies the FP32-stored values that a
 * mixed-precision scene's inner solve uses.  MPH_E_STATE: no matrix since the
 * particles last moved, or FP32 values asked of an FP64 scene. */
int mph_scene_apply_matrix(mph_scene* s, int precision, const double* x, double* y);

/* mph_scene_apply_matrix for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); it has
 * no counterpart in the reference.  x and y are DEVICE arrays of 3*FluidCount
 * doubles on the scene's device, owned by the caller; the result is bit-identical
 * to mph_scene_apply_matrix.  The call allocates nothing: the factored product's
 * records (x, t and the positions) go to a fixed 96 MiB scratch array that the
 * library holds on every device it runs on, so a scene of the factored form may
 * have at most 1,797,558 particles (FP64) or 6,291,456 (FP32 values); larger ones
 * are rejected with MPH_E_ARG.  The scratch is shared by every call on a device:
 * calls that overlap on one device race.  Runs on the scene's stream and returns
 * after it. */
int mph_scene_apply_matrix_noalloc(mph_scene* s, int precision, const double* x, double* y);

int mph_scene_destroy(mph_scene* s);

/* ---- Batched executor -------------------------------------------------------
 * No counterpart in the reference, which advances one scene per process
 * (main.cpp:386-474); a batch replaces running many such processes side by side.
 * A batch holds up to 256 caller-owned scenes of one device (they stay owned by the
 * caller; mph_scene_destroy returns MPH_E_STATE while the scene is in a batch:
 * remove it or destroy the batch first).
 * mph_batch_step advances every member whose time is below its end_time by one
 * time step: the phases before the solve on each scene's stream, then the CR
 * solves of all members in lockstep (one launch per kernel for the whole batch,
 * one host synchronisation per iteration), then diagnostics and the advance.
 * Each scene's results are bit-identical to mph_scene_step on that scene.
 * `stats` (may be NULL) receives one record per member, in mph_batch_add order;
 * ms_total is the wall time of the whole batch step, and ms_phase the batch's wall
 * time per group of phases, the same in every record: [MPH_PHASE_NEIGHBOR]
 * kinematics, cells and neighbour lists; [MPH_PHASE_MULTIGRID] explicit forces
 * through the multigrid build; [MPH_PHASE_SOLVE] the lockstep solves;
 * [MPH_PHASE_ADVANCE] diagnostics and advance.
 * A member whose CR residual is not finite leaves the lockstep solve at once and
 * does not advance; its record has status MPH_E_STATE, the other members finish the
 * step, and the call returns MPH_E_STATE (remove the failed member before the next
 * step).  Any other failure returns its code and leaves the members in an
 * unspecified state. */
typedef struct mph_batch mph_batch;
int mph_batch_create(int device, mph_batch** out);
int mph_batch_add(mph_batch* b, mph_scene* s);
int mph_batch_remove(mph_batch* b, mph_scene* s);
int mph_batch_size(const mph_batch* b);
int mph_batch_step(mph_batch* b, mph_step_stats* stats);
int mph_batch_destroy(mph_batch* b);

/* Measured HBM ceilings of `device`, for the roofline (no counterpart in the
 * reference): the best of `reps` timed passes of a streaming read of `bytes` and of a
 * device-to-device copy of `bytes` (read + write counted), with the 16-byte
 * nontemporal loads the SpMV uses.  Allocates and frees 2*bytes of device memory. */
int mph_probe_bandwidth(int device, unsigned long long bytes, int reps, double* read_gbs, double* copy_gbs);

/* Device bytes held by the scene (all persistent buffers, as now grown). */
int mph_scene_device_bytes(mph_scene* s, unsigned long long* bytes);

/* Message of the last failing call on this thread ("" if none). */
const char* mph_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
