/*
 * snac_hip.h -- C ABI of the MI355X (gfx950) implementation of SNAC's explicit
 * time step.
 *
 * The library (libsnac_hip.so) owns one rank's subdomain on one GPU: node
 * coordinates and velocities, the per-tetrahedron stress tensors and the
 * per-tetrahedron plastic strain / viscosity extensions.  The host keeps the
 * reference's StGermain context, entry points and plugin registration; a
 * "SnacHip" plugin replaces the default hooks of the loop-level entry points
 * with the calls below.
 *
 * Reference interfaces each entry point stands in for are cited per function
 * (paths relative to the geodynamics/snac tree).  All functions return an int
 * status: SNH_OK (0) or a negative SNH_E* code; snh_last_error() gives the
 * message of the last failure on the calling thread.  Nothing here aborts the
 * process (the reference's Journal_Firewall does; StGermain/Base/IO/src/
 * Journal.c:568-660).
 *
 * Threading: a context is used by one host thread at a time.  All device work
 * of a context is issued on the context's stream (snh_set_stream); every entry
 * point that returns host-visible values synchronises that stream.
 */
#ifndef SNAC_HIP_H
#define SNAC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNH_ABI_VERSION 6

/* ---- status codes ------------------------------------------------------- */
#define SNH_OK            0
#define SNH_EINVAL       -1   /* bad argument or configuration            */
#define SNH_ENOMEM       -2   /* device or host allocation failed          */
#define SNH_EHIP         -3   /* a HIP runtime call failed                 */
#define SNH_ESTATE       -4   /* call out of order (e.g. step before zero) */
#define SNH_EUNSUPPORTED -5   /* option the device path does not implement */
#define SNH_ENUMERIC     -6   /* NaN/Inf firewall tripped                   */

/* ---- fixed sizes (Snac/libSnac/src/TetrahedraTables.h:47-49) ---------- */
#define SNH_TETRA_COUNT        10
#define SNH_ELEMENT_NODE_COUNT  8
#define SNH_MAX_SEGMENTS        8
#define SNH_MAX_MATERIALS      16
#define SNH_STRESS_COMPONENTS   6   /* order: 00, 11, 22, 01, 02, 12 */

/* ---- rheology bits (Snac/libSnac/src/Material.h:46-50) ----------------- */
#define SNH_RHEOLOGY_ELASTIC      0x1u
#define SNH_RHEOLOGY_PLASTIC      0x2u
#define SNH_RHEOLOGY_MAXWELL      0x4u
#define SNH_RHEOLOGY_VISCOPLASTIC 0x8u

/* ---- dt types (Snac/libSnac/src/Context.c:345-391) --------------------- */
#define SNH_DT_CONSTANT 0
#define SNH_DT_DYNAMIC  1
#define SNH_DT_WAVE     2

/* ---- velocity-BC node sets (StGermain/Discretisation/Utils/src/WallVC.c:581-670,
 *      AllNodesVC).  I = x, J = y (vertical, top = max J), K = z.            */
#define SNH_WALL_FRONT  0   /* k == 0     */
#define SNH_WALL_BACK   1   /* k == K-1   */
#define SNH_WALL_LEFT   2   /* i == 0     */
#define SNH_WALL_RIGHT  3   /* i == I-1   */
#define SNH_WALL_TOP    4   /* j == J-1   */
#define SNH_WALL_BOTTOM 5   /* j == 0     */
#define SNH_WALL_ALL    6   /* every node */

/* component mask bits of a velocity condition */
#define SNH_VX 0x1u
#define SNH_VY 0x2u
#define SNH_VZ 0x4u

/* ---- arithmetic of the step -------------------------------------------- */
#define SNH_PRECISION_FP64  0   /* FP64 throughout, reference-identical order */
#define SNH_PRECISION_MIXED 1   /* FP32 strain-rate/constitutive/force, FP64 state */

/* ---- per-element input fields (generalisations of per-material values) - */
#define SNH_FIELD_REFVISC          0  /* reference viscosity [Pa s], replaces material->refvisc */
#define SNH_FIELD_COHESION_OFFSET  1  /* [Pa], added to the softened cohesion                   */

/* One material phase: the subset of Snac_Material (Snac/libSnac/src/Material.h:53-100)
 * the explicit step reads.  Units SI; angles in degrees. */
typedef struct snh_material {
    uint32_t rheology;            /* SNH_RHEOLOGY_* bits                      */
    double   phsDensity;
    double   alpha, beta;         /* density( T, p ) coefficients             */
    double   lambda, mu;          /* Lame parameters                          */
    uint32_t nsegments;           /* piecewise-linear softening segments      */
    double   plstrain[SNH_MAX_SEGMENTS + 1];
    double   frictionAngle[SNH_MAX_SEGMENTS + 1];
    double   dilationAngle[SNH_MAX_SEGMENTS + 1];
    double   cohesion[SNH_MAX_SEGMENTS + 1];
    double   ten_off;             /* tension cut-off; 0 => cohesion/tan(phi)  */
    double   vis_min, vis_max, refvisc, refsrate, reftemp;
    double   activationE, srexponent;
    double   thermal_conduct;     /* W/m/K, read by the energy entry points    */
    double   heatCapacity;        /* J/kg/K                                     */
} snh_material;

/* One velocity condition of the velocityBCs composite (StGermain/Base/Automation/
 * src/CompositeVC.c).  Later entries win where two constrain one component
 * (_CompositeVC_GetValueIndex searches items last to first). */
typedef struct snh_velocity_bc {
    int32_t  wall;                /* SNH_WALL_*                               */
    uint32_t mask;                /* SNH_VX | SNH_VY | SNH_VZ                 */
    double   value[3];            /* [m/s] for the masked components          */
} snh_velocity_bc;

/* Problem description: the mesh, phase-independent constants and BCs that
 * _Snac_Context_Init / _Snac_Context_InitialConditions read from the
 * dictionary (Snac/libSnac/src/Context.c:155-204,345-391,742-824). */
typedef struct snh_config {
    uint32_t meshSizeI, meshSizeJ, meshSizeK;   /* NODE counts per axis (mesh struct) */
    double   minX, minY, minZ, maxX, maxY, maxZ; /* BlockGeometry extent               */
    int32_t  dtType;              /* SNH_DT_*                                  */
    double   timeStep;            /* initial dt (constant: the dt)              */
    double   gravity;             /* acts on -y (Snac/libSnac/src/Force.c:173)  */
    double   demf;                /* local damping factor                       */
    double   strain_inert;        /* dynamic mass-scaling factor                */
    double   topo_kappa;          /* top-surface diffusivity (0 = off; the
                                     reference's default is 5e-6), Cartesian
                                     meshes; two nodes wide in x (z) it
                                     diffuses along z (x) only, as
                                     UpdateNode.c:494-499 does             */
    double   topoGradCriterion;   /* diffusion runs once the running maximum of
                                     |grad y| over the top surface reaches it
                                     (the reference's default 1e-16)        */
    uint32_t nmaterials;
    const snh_material* materials;
    uint32_t nbcs;
    const snh_velocity_bc* bcs;
    int32_t  precision;           /* SNH_PRECISION_*                            */
} snh_config;

/* Host-side view of the step state, in the reference's index order.
 * Any pointer may be NULL to skip that field.
 *   coord          [N][3]        mesh->nodeCoord
 *   velocity       [N][3]        Snac_Node.velocity
 *   stress         [E][10][6]    Snac_Element_Tetrahedra.stress (00,11,22,01,02,12)
 *   plastic_strain [E][10]       SnacPlastic_Element.plasticStrain
 *   viscosity      [E][10]       SnacMaxwell_Element.viscosity (download only)
 *   volume         [E][10]       tetra volume of the current geometry (download only)
 *   temperature    [N]           SnacTemperature_Node.temperature (contexts
 *                                with snh_set_thermal)
 * Node index n = i + I*(j + J*k); element index e = i + (I-1)*(j + (J-1)*k)
 * (StGermain/Discretisation/Mesh/src/HexaEL.c:400-426). */
typedef struct snh_state {
    double* coord;
    double* velocity;
    double* stress;
    double* plastic_strain;
    double* viscosity;
    double* volume;
    double* temperature;
} snh_state;

/* Step scalars kept by the context (Snac_Context fields). */
typedef struct snh_scalars {
    double   dt;                  /* dt of the next step                       */
    double   speedOfSound;
    double   minLengthScale;
    double   vmax;                /* max |v| over constrained nodes, floored    */
    uint32_t timeStep;            /* steps completed                           */
    double   topoGradMax;         /* running max of |grad y| on the top surface
                                     (topography diffusion; 0 when off)     */
} snh_scalars;

typedef struct snh_context snh_context;

/* Library version / last error. */
int         snh_abi_version(void);
const char* snh_last_error(void);

/* Create a context for one subdomain and allocate its device state.  Builds
 * the mesh (BlockGeometry.c:326-349 coordinates, zero velocity/stress) and the
 * per-node BC table.  Replaces: Snac_Context_New + _Snac_Context_Build
 * (Snac/libSnac/src/Context.c:90-100,707-740). */
int snh_context_create(const snh_config* cfg, snh_context** out);
int snh_context_destroy(snh_context* ctx);

/* Use an existing HIP stream (hipStream_t passed as void*); NULL = the
 * context's own stream. */
int snh_set_stream(snh_context* ctx, void* stream);

/* Per-element material index (element->material_I).  nelem entries.  For
 * the wave dt type the speed of sound is the largest Vp over the materials
 * in use; after snh_timestep_zero it is a running maximum, as the
 * reference's (Context.c:392, UpdateElement.c:62-67): a later call can
 * raise it, never lower it. */
int snh_set_element_material(snh_context* ctx, const uint8_t* material, size_t nelem);

/* Per-element input field (SNH_FIELD_*), nelem doubles. */
int snh_set_element_field(snh_context* ctx, int field, const double* values, size_t nelem);

/* Copy state in/out (host memory).  Upload also re-applies nothing: the
 * caller passes the state as the reference holds it after
 * _Snac_Context_InitialConditions (Context.c:742-774).  An upload clears the
 * device firewall flags; a restart is an upload followed by
 * snh_timestep_zero. */
int snh_upload_state(snh_context* ctx, const snh_state* state);
int snh_download_state(snh_context* ctx, snh_state* state);

/* Initial isotropic lithostatic stress of the SnacHydroStaticIC plugin
 * (Snac/plugins/hydroStaticIC/VariableConditions.c:43-160, hooked before
 * SnacTimeStepZero at Register.c:127-132), for one rank: per element column,
 * top down, P = -(rogh + dP/2) on the diagonal of every tet stress.  Host-side
 * computation on the current coordinates, then uploaded. */
int snh_hydrostatic_ic(snh_context* ctx);

/* Remesher handoff: the remesher stays on the host (remesher_BI,
 * hooked on AbstractContext_EP_Sync, Snac/plugins/remesher_BI/Register.c:
 * 127-131; _SnacRemesher_Remesh, Remesh.c:46-186).  Between steps the host
 * downloads the state (snh_download_state: stress with the last rotation
 * applied), computes new coordinates and interpolates node and element
 * values onto them for the same topology, then hands them back here; NULL
 * fields are kept.  Replaces the copy-back and _SnacRemesher_UpdateElements
 * (RemeshElements.c:133-151): the coordinates before the remesh become the
 * previous coordinates (the old_volume of Maxwell's next volumetric
 * strain), the new mesh's Jaumann rotation with the dt of the last Sync is
 * applied at the next CalcStresses, and minLengthScale becomes the minimum
 * of its current value and the new mesh's; dt is not recomputed.  FP64 and
 * mixed-precision contexts (both hold the state in FP64; a pending rotation
 * is applied by the FP64 kernel). */
int snh_remesh(snh_context* ctx, const snh_state* remeshed);

/* One condition of the temperatureBCs composite (a composite of wall
 * conditions, like velocityBCs; later entries win on shared nodes). */
typedef struct snh_temperature_bc {
    int32_t  wall;                /* SNH_WALL_*                               */
    double   value;               /* temperature on that wall                  */
} snh_temperature_bc;

/* The SnacTemperature plugin (Snac/plugins/temperature/Register.c:77-140):
 * the energy entry points run at the start of every step
 * (_Snac_Context_Solve, Context.c:1073-1074):
 *   Snac_EP_LoopElementsEnergy -> SnacTemperature_LoopElements
 *       per-tet heat flux k grad T from the current geometry, the heat-flux
 *       condition on the vertical walls, thermal stress
 *       (HeatFlux.c:47-155, VariableConditions.c:227-353)
 *   Snac_EP_LoopNodesEnergy -> SnacTemperature_LoopNodes
 *       nodal temperature update, tet average temperature, temperature BCs
 *       (Heat.c:47-172)
 * The nodal temperature then enters the Arrhenius viscosity of the Maxwell
 * and visco-plastic laws, the tet density (alpha) and, with
 * computeThermalStress, the stress of those laws.  Call after create and
 * before snh_timestep_zero; temperature[N] is the initial nodal field (the
 * nodeICs), to which _SnacTemperature_InitialConditions (VariableConditions.c:
 * 155-212) applies the BCs.  FP64 and mixed-precision contexts (the heat
 * passes run in FP64 in both); not with the adjoint. */
typedef struct snh_thermal {
    int32_t  computeThermalStress;   /* Context.c:756-757                   */
    uint32_t nbcs;
    const snh_temperature_bc* bcs;
    const double* temperature;       /* [N] initial nodal temperature         */
} snh_thermal;
int snh_set_thermal(snh_context* ctx, const snh_thermal* thermal);

/* Snac_Context_TimeStepZero (Snac/libSnac/src/Context.c:978-1036): geometry of
 * the initial mesh, minLengthScale and the first Sync (dt). */
int snh_timestep_zero(snh_context* ctx);

/* The loop-level entry points of one step, callable one by one from the
 * host's hooks:
 *   snh_calc_stresses            Snac_EP_CalcStresses      (Context.c:1102-1117)
 *   snh_loop_nodes_momentum      Snac_EP_LoopNodesMomentum (Context.c:1120-1154)
 *   snh_loop_elements_momentum   Snac_EP_LoopElementsMomentum (Context.c:1156-1184)
 *   snh_sync                     AbstractContext_EP_Sync -> _Snac_Context_Sync
 *                                                          (Context.c:1208-1268)
 * The Jaumann rotation computed by LoopElementsMomentum is applied to the
 * stored stress at the start of the next CalcStresses (same values, same
 * order); snh_download_state returns the rotated stress in between. */
int snh_loop_elements_energy(snh_context* ctx);   /* Snac_EP_LoopElementsEnergy (Context.c:1073) */
int snh_loop_nodes_energy(snh_context* ctx);      /* Snac_EP_LoopNodesEnergy (Context.c:1074) */
int snh_calc_stresses(snh_context* ctx);
int snh_loop_nodes_momentum(snh_context* ctx);
int snh_loop_elements_momentum(snh_context* ctx);
int snh_sync(snh_context* ctx);

/* nsteps whole steps (AbstractContext_Step, StGermain/Base/Context/src/
 * AbstractContext.c:921-944) with dt kept on the device.  Asynchronous:
 * returns after enqueueing; snh_synchronize or any download waits. */
int snh_step(snh_context* ctx, int nsteps);

int snh_get_scalars(snh_context* ctx, snh_scalars* out);
/* Waits for the context's stream, then reports a device firewall raised by
 * any step since the last check (SNH_ENUMERIC: a non-finite velocity or dt,
 * the stress firewall of Stress.c:161-166; or the thermal dt bound of
 * temperature/Heat.c:121-126), once. */
int snh_synchronize(snh_context* ctx);

/* Per-kernel durations measured with HIP events recorded on the context's
 * stream around every launch while profiling is on (no counterpart in the
 * reference; its TAU/CallGraph instrumentation, README_PROFILING and
 * Snac/libSnac/src/EntryPoint.c:162-173, plays this role on the CPU).
 * snh_profile(ctx, 1) clears the accumulators and starts recording;
 * snh_get_kernel_times waits for the stream and returns totals in milliseconds. */
typedef struct snh_kernel_times {
    double   elem_ms, node_ms, geo_ms, fin_ms;
    uint32_t elem_n, node_n, geo_n, fin_n;
    double   heat_elem_ms, heat_node_ms;     /* energy entry points */
    uint32_t heat_elem_n, heat_node_n;
} snh_kernel_times;
int snh_profile(snh_context* ctx, int enable);

/* Kernel-variant autotuner (no counterpart in the reference).  Times each
 * candidate implementation of the element pass (one tet per lane, 6
 * elements per wave; one element per thread keeping the tet kinematics in
 * registers; the same recomputing them; one wave per tet index over 64
 * elements) on a scratch copy of the current state, a few launches
 * each, and keeps the fastest for this context.  Call after
 * snh_timestep_zero; the state is not modified.  The choice can be forced
 * with the environment variable SNH_ELEM_VARIANT=0|1|2|3. */
#define SNH_ELEM_TET_PER_LANE            0
#define SNH_ELEM_ELEMENT_PER_LANE        1
#define SNH_ELEM_ELEMENT_PER_LANE_LOWREG 2   /* kinematics recomputed, not kept */
#define SNH_ELEM_TET_PER_WAVE            3   /* 64 elements per block, one wave per tet */
#define SNH_ELEM_VARIANTS                4
#define SNH_NODE_TUNES 4   /* node pass: per node; tiled with 4, 8, 16 planes */
typedef struct snh_tune_report {
    double   elem_ms[SNH_ELEM_VARIANTS];  /* best launch time per variant, ms */
    int32_t  elem_variant;                /* chosen */
    double   node_ms[SNH_NODE_TUNES];     /* FP64 node pass: per-node gather, tiled
                                             with 4, 8 and 16 node planes per
                                             workgroup; 0 when not timed */
    int32_t  node_variant;                /* chosen SNH_NODE_* */
    int32_t  node_kc;                     /* and its node planes per workgroup */
} snh_tune_report;
/* Time the element-pass layouts and (FP64 contexts) the node-pass variants
 * and tile heights on the context's own mesh, law and state, outputs to
 * scratch, and keep the fastest of each; every choice gives the same bits.
 * SNH_ELEM_VARIANT / SNH_NODE_VARIANT in the environment pin the choice. */
int snh_autotune(snh_context* ctx, snh_tune_report* out);
/* Force one element-pass variant (SNH_ELEM_*); both give identical results. */
int snh_set_elem_variant(snh_context* ctx, int32_t variant);

/* Node-pass variants of the FP64 path (no counterpart in the reference;
 * both sum every node's tet-face shares in the order of Force.c:87-183 and
 * give identical results):
 *   SNH_NODE_PER_NODE  one thread per node, the 8 surrounding elements read
 *                      from global memory (the default)
 *   SNH_NODE_TILED     a workgroup per 16 x 8 node columns walking up in k,
 *                      node coordinates and each element plane's tet volumes
 *                      staged in LDS, each tet volume computed once per tile
 * The environment variables SNH_NODE_VARIANT=0|1 and SNH_NODE_KC (node
 * planes per workgroup of the tiled pass, default 8) set them at creation. */
#define SNH_NODE_PER_NODE 0
#define SNH_NODE_TILED    1
int snh_set_node_variant(snh_context* ctx, int32_t variant);

/* ---- dump diagnostics -----------------------------------------------------
 * While enabled, every step also records what the reference's dumps hold
 * besides the state (Context.c:1311-1489, plastic_BI/Output.c:60-80): the
 * element strain-rate invariant (StrainRate.c:272-304), stress invariant and
 * hydrostatic pressure (Stress.c:140-156), the volume-averaged plastic strain
 * (plastic_BI/Constitutive.c:232-236) and the damped nodal force
 * (UpdateNode.c:102-104).  Enable it for the steps that end on a dump.
 * On a mixed-precision context the element quantities come from the FP64
 * element pass run on the same state into scratch before the mixed pass
 * (one more element pass per recorded step), the force from the mixed node
 * pass itself. */
typedef struct snh_diagnostics {
    double* strain_rate;      /* [E] or NULL */
    double* stress;           /* [E] or NULL */
    double* hydro_pressure;   /* [E] or NULL */
    double* plastic_strain;   /* [E] or NULL (0 for laws without one) */
    double* force;            /* [N][3] or NULL */
} snh_diagnostics;
int snh_set_diagnostics(snh_context* ctx, int enable);
/* values of the last step run with diagnostics enabled (host arrays) */
int snh_get_diagnostics(snh_context* ctx, const snh_diagnostics* out);

/* Prescribe, for the next nsteps calls of the step (snh_step; FP64 and mixed),
 * dt and speed of sound of each step and optionally the damping branch
 * sign(v) sign(f) of every node component (damp: [nsteps][N][3] in
 * {-1, 0, 1}, or NULL): sign(v) in the damping term is replaced by
 * code * sign(f).  A schedule recorded from a run reproduces that run bit
 * for bit; with perturbed inputs it holds the run's discrete decisions (the
 * map the adjoint below differentiates).  nsteps = 0 clears it.  No
 * counterpart in the reference. */
int snh_set_schedule(snh_context* ctx, int32_t nsteps, const double* dt, const double* speed_of_sound,
                     const int8_t* damp);

/* Run nsteps steps (as snh_step) and record, per step, what
 * snh_set_schedule prescribes: dt[nsteps] and speed_of_sound[nsteps] the
 * step ran with and, when damp is not NULL, the damping branch
 * sign(v) sign(f) of every node component (damp: [nsteps][N][3]).  Feeding
 * the record to snh_set_schedule of another context with the same inputs
 * replays the run bit for bit; a context of the other precision then
 * follows the recorded run's discrete decisions.  Synchronous.  No
 * counterpart in the reference. */
int snh_record_schedule(snh_context* ctx, int32_t nsteps, double* dt, double* speed_of_sound, int8_t* damp);

/* ---- adjoint of the explicit step ----------------------------------------
 * No counterpart in the reference (SNAC has no adjoint).  Reverse-mode
 * derivative of n explicit steps with respect to a per-element parameter
 * field, for misfits of the final coordinates and velocities.  The
 * derivative is that of the step with its discrete decisions frozen at the
 * recorded trajectory: dt, rotation dt and speed of sound of every step, the
 * damping branch sign(v) sign(f) of every node component, the plastic branch
 * of every tet (elastic / shear / tension) and the viscosity clamp.
 *
 * Schedule.  The steps fall into blocks of checkpoint_interval steps; a
 * block is replayed from the state entering it into a tape (stress, plastic
 * strain, coordinates, velocity, undamped force, mass of each step, about
 * 750 B per element and step) and then reversed.  With snapshots == 0
 * (uniform two-level checkpointing) snh_adjoint_forward keeps the full
 * state entering every block (about 560 B per element each) and every step
 * runs twice in total.  With snapshots > 0 it keeps only the state it
 * starts from, and the reverse sweep follows the binomial checkpointing
 * schedule (Griewank's revolve) over the blocks with at most `snapshots`
 * further states at a time, replaying the fewest blocks that budget allows:
 * memory is the tape of one block and snapshots + 1 states, whatever the
 * number of steps.  Both schedules replay every step from the same state
 * with the same kernels, so they give the same gradient bits.
 *
 * Supported: FP64, dynamic dt, elastic, plastic (plastic_BI semantics,
 * friction and dilation angles constant along the softening table) and
 * Maxwell with a linear viscosity law (srexponent 1, activationE 0). */
#define SNH_PARAM_LOG_VISCOSITY 0   /* theta_e = ln(refvisc_e); needs SNH_FIELD_REFVISC */
#define SNH_PARAM_COHESION      1   /* theta_e = cohesion offset of element e (Pa)      */

/* Mohr-Coulomb corner regularisation.  The reference's single-surface return
 * (plastic_BI/Constitutive.c:159-223) can move two principal stresses past
 * each other; there the returned stress jumps with the eigenvectors of the
 * trial stress and its exact derivative carries the divided difference
 * (g_k - g_l) / (s_k - s_l) < 0, unbounded as the gap closes, which makes
 * long reverse sweeps through plastic flow grow without bound.  With this
 * flag that divided difference is clamped to [0, 1], the value of a return
 * that stops at the corner.  Off: the exact derivative. */
#define SNH_ADJ_CORNER_CLAMP 0x1

typedef struct snh_adjoint_config {
    int32_t param;                /* SNH_PARAM_*                                  */
    int32_t checkpoint_interval;  /* steps per block (tape length), >= 1          */
    int32_t flags;                /* SNH_ADJ_*                                    */
    int32_t snapshots;            /* 0: a state kept per block; > 0: binomial     */
                                  /* schedule with this many states beside the    */
                                  /* first, allocated by snh_adjoint_begin        */
} snh_adjoint_config;

typedef struct snh_adjoint_times {
    double forward_ms;    /* snh_adjoint_forward, checkpoint copies included */
    double replay_ms;     /* forward replays of the reverse sweeps           */
    double reverse_ms;    /* reverse kernels                                 */
    int32_t steps;        /* steps recorded                                  */
    int32_t checkpoints;  /* states held: blocks (uniform) or snapshots + 1  */
    int64_t replay_steps; /* steps replayed by the reverse sweeps            */
} snh_adjoint_times;

/* Start recording at the current state (after snh_timestep_zero). */
int snh_adjoint_begin(snh_context* ctx, const snh_adjoint_config* cfg);
/* Run nsteps explicit steps (same results as snh_step), storing checkpoints;
 * may be called several times before snh_adjoint_reverse, not after it. */
int snh_adjoint_forward(snh_context* ctx, int32_t nsteps);
/* Reverse sweep over all recorded steps.  coord_bar / vel_bar: d(misfit)/d
 * (final coordinates / velocities), [N][3] node-major like snh_state, either
 * may be NULL (zero).  grad: d(misfit)/d(theta), [E] doubles, written.  The
 * context's state is left at the end of the forward run. */
int snh_adjoint_reverse(snh_context* ctx, const double* coord_bar, const double* vel_bar, double* grad);
int snh_adjoint_get_times(snh_context* ctx, snh_adjoint_times* out);
/* Release checkpoints, tape and adjoint buffers. */
int snh_adjoint_end(snh_context* ctx);

/* ---- ensembles -----------------------------------------------------------
 * No counterpart in the reference.  Steps several contexts (independent
 * models of any size, each with its own materials, BCs and element-pass
 * variant, e.g. from snh_autotune) together: a HIP graph holds graph_steps
 * steps of every member, each member's launches on its own stream as a
 * parallel branch, so small meshes share the GPU.  One graph launch then
 * advances every member graph_steps steps with no host work in between.
 * Results are bit-identical to stepping each member with snh_step.
 * graph_steps must be even (the coordinate buffers alternate every step);
 * members must have no schedule and profiling off while in an ensemble.
 *
 * Batched launches.  When the members share one precision and one single
 * law (elastic, Mohr-Coulomb or Maxwell), the same material layout, the
 * per-node node pass and no dump diagnostics or tet densities, and average
 * fewer than 16384 elements (where launches, not kernels, set the time),
 * the graph instead holds, per step, one launch of each pass over all
 * members' blocks (the element pass in the tet-per-lane layout whatever the
 * members' autotuned variants, which give the same bits): 4 launches per
 * step for the whole ensemble instead of 4 per member.  Environment at
 * creation: SNH_ENSEMBLE_BATCH=1 batches whatever the sizes, =0 never;
 * SNH_ENSEMBLE_TW=1 takes the tet-per-wave layout; SNH_ENSEMBLE_GROUPS=k
 * splits the members into k groups of about equal size, batched on k
 * streams. */
typedef struct snh_ensemble snh_ensemble;
int snh_ensemble_create(snh_context* const* members, int32_t n, int32_t graph_steps, snh_ensemble** out);
/* 1 when the ensemble steps its members by batched launches, 0 when by
 * per-member branches; negative on error. */
int snh_ensemble_batched(const snh_ensemble* ens);
/* nsteps steps of every member: whole graph launches, the rest with snh_step */
int snh_ensemble_step(snh_ensemble* ens, int32_t nsteps);
int snh_ensemble_synchronize(snh_ensemble* ens);
int snh_ensemble_destroy(snh_ensemble* ens);

/* ---- allocation-free Sync -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the Sync of a mesh state held in the caller's device
 * memory, with no allocation and no context.  It computes what
 * snh_timestep_zero computes for a context holding that state: the minimum
 * length scale, min V/A over every tet face (UpdateElement.c:108-139), the
 * largest |v| over the nodes on the walls of cfg's velocity conditions,
 * floored at 3e-11f, and the dt of cfg->dtType with its speed of sound
 * (_Snac_Context_Sync, Context.c:1208-1268; the constant type keeps
 * cfg->timeStep; the wave type takes the largest Vp over cfg's materials).
 *   coord, velocity  device memory, component rows: x[N] y[N] z[N], node
 *                    n = i + I (j + J k) of cfg's mesh
 *   out              device memory, 4 doubles: dt, speedOfSound,
 *                    minLengthScale, vmax
 *   stream           hipStream_t as void*, NULL for the default stream
 * Enqueues its kernels and returns without waiting.  The block partials and
 * the scalars live in one fixed device array of the library, 393 KB on each
 * GPU it runs on, held until the process ends: meshes of at most 4194304
 * nodes and 4194304 elements (SNH_EINVAL beyond).  Every call on a GPU uses
 * that GPU's array, so calls that overlap on one GPU race. */
int snh_sync_state(const snh_config* cfg, const double* coord, const double* velocity, double* out, void* stream);

/* Device-to-device copy bandwidth of the current GPU (GB/s, read + write
 * bytes), from a grid-stride copy of nbytes timed with HIP events over iters
 * launches: the measured HBM peak reported beside the 8 TB/s specification.
 * Utility with no counterpart in the reference. */
int snh_measure_copy_bandwidth(size_t nbytes, int iters, double* gbps);
int snh_get_kernel_times(snh_context* ctx, snh_kernel_times* out);

#ifdef __cplusplus
}
#endif
#endif /* SNAC_HIP_H */
