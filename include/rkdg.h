/*
 * rkdg.h -- C-ABI of the MI355X RKDG shallow-water stage engine (librkdg.so).
 *
 * The reference exposes this path as a compile-time C++ "Problem" concept, not as an
 * FFI: SWE::RKDG::Problem (source/problem/SWE/discretization_RKDG/rkdg_swe_problem.hpp:34-194)
 * driven by Serial::Simulation<P>::Run (source/simulation/serial/simulation.hpp:68-82).
 * Each entry point below names the reference hook(s) it replaces.  All calls are
 * stream-ordered on the context's own HIP stream; a context is owned by one host thread
 * (the reference's one-owner-per-submesh rule, simulation/ompi/simulation_ompi.hpp:63).
 * No C++ exceptions cross this boundary: every call returns an rkdg_status and the
 * message of the last failure is available from rkdg_last_error().
 *
 * Layout conventions (host arrays passed in and out):
 *   modal state q is [3][ndof][n_elements], variable-major then mode then element
 *   (ze = 0, qx = 1, qy = 2; modes in the reference's Dubiner order, basis_dubiner_2D.cpp:18-23);
 *   flags are one byte per element.
 */
#ifndef RKDG_H
#define RKDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rkdg_ctx rkdg_ctx;

typedef enum {
    RKDG_OK = 0,
    RKDG_ERR_ARG = 1,         /* invalid argument / mesh (reference: std::logic_error at setup) */
    RKDG_ERR_HIP = 2,         /* a HIP runtime call failed */
    RKDG_ERR_NAN = 3,         /* NaN in the state (reference: scrutinize_solution -> abort()) */
    RKDG_ERR_UNSUPPORTED = 4, /* configuration not built (e.g. p outside 1..3 on the device) */
    RKDG_ERR_NO_DEVICE = 5    /* no HIP device visible */
} rkdg_status;

typedef enum {
    RKDG_FP64_PLAIN = 0, /* FP64 in the reference's operation order: bit-exact parity form */
    RKDG_FP32_BULK = 1,  /* FP32 state/flux/accumulation; FP64 geometry precompute and wet/dry */
    RKDG_FP64_FAST = 2   /* FP64 with fused multiply-adds, reordered sums and Newton reciprocals:
                            within 1e-12 relative per-step L2 of the plain form */
} rkdg_precision;

/* Boundary condition codes (reference: SWE::BoundaryTypes, problem/SWE/swe_definitions.hpp:61-69). */
enum {
    RKDG_BC_LAND = 0,
    RKDG_BC_TIDE = 1,
    RKDG_BC_FLOW = 2,
    RKDG_BC_FUNCTION = 3,
    RKDG_BC_LEVEE = 129     /* internal barrier side (INTERNAL + 1, utilities/edge_types.hpp:4) */
};

/* Closed-form problem function sets (reference: compile-time ic_q/source_q/true_q files,
 * problem/SWE/problem_function_files and examples/; here a runtime choice). */
enum {
    RKDG_PROBLEM_NONE = 0,          /* constant initial state, no function source, no truth */
    RKDG_PROBLEM_MANUFACTURED = 1,  /* examples/swe_manufactured_solution */
    RKDG_PROBLEM_PARABOLIC_BOWL = 2,/* examples/swe_parabolic_bowl */
    RKDG_PROBLEM_DAM_BREAK = 3      /* step in free surface at x = x0 (problem_param[0..2]) */
};

enum { RKDG_FRICTION_NONE = 0, RKDG_FRICTION_CHEZY = 1, RKDG_FRICTION_MANNING = 2 };

/*
 * Mesh description, in the reference's ADCIRC model (preprocessor/ADCIRC_reader/adcirc_format.cpp:3-…):
 * nodes carry (x, y, depth), depth positive below the datum; element e is row e of
 * elt_nodes and its ID is e (elements iterate in ascending ID, heterogeneous_containers.hpp:76-78).
 * Every mesh edge with one adjacent element must be listed in bound_edge_nodes with its
 * boundary code (reference: AdcircFormat::get_ibtype, adcirc_format.cpp:214).
 * Tide/flow forcing: n_constituents (any number) harmonic constituents shared by all forced nodes
 * (reference .bcis data, problem_input/swe_inputs.cpp:212-…), amplitude/phase per node.
 */
typedef struct rkdg_mesh_desc {
    int32_t n_nodes;
    int32_t n_elements;
    const double* node_xyz;          /* [n_nodes*3] */
    const int32_t* elt_nodes;        /* [n_elements*3] */
    int32_t n_bound_edges;
    const int32_t* bound_edge_nodes; /* [n_bound_edges*2] */
    const uint8_t* bound_edge_type;  /* [n_bound_edges], RKDG_BC_* */
    int32_t n_constituents;
    const double* frequency;         /* [n_constituents] rad/s */
    const double* forcing_fact;      /* [n_constituents] */
    const double* equilib_arg;       /* [n_constituents] degrees */
    const double* node_amplitude;    /* [n_nodes*n_constituents] or NULL */
    const double* node_phase;        /* [n_nodes*n_constituents] degrees, or NULL */
    const double* node_manning_n;    /* [n_nodes] or NULL */
    /* Internal barriers (levees).  Both node strings of a barrier are boundary edges of code
     * RKDG_BC_LEVEE; pair i joins front node levee_nodes[2i] with back node levee_nodes[2i+1]
     * (reference: IBTYPE %10 == 4 NBVV/IBCONN in the .14, adcirc_format.cpp:89-102) and
     * carries the barrier height and the sub-/supercritical weir coefficients at that node
     * pair (the .bcis levee block, swe_inputs.cpp:298-310). */
    int32_t n_levee_pairs;
    const int32_t* levee_nodes;      /* [n_levee_pairs*2] front, back */
    const double* levee_param;       /* [n_levee_pairs*3] H_barrier, C_subcritical, C_supercritical */
} rkdg_mesh_desc;

/* Run parameters (reference: InputParameters + SWE::Inputs, preprocessor/input_parameters.hpp:96-261,
 * problem_input/swe_inputs.hpp:126-155, and the globals set by
 * problem_preprocessor/swe_pre_init_prob_params.hpp:5-47). */
typedef struct rkdg_params {
    int32_t p;               /* polynomial order: 1, 2 or 3 */
    int32_t nstages;         /* SSP-RK stages; (nstages, order) is one of the reference's twelve   */
    int32_t order;           /* schemes: (1,1) (2,2) (3,3) (4..8,3) (5..8,4); others: RKDG_ERR_ARG */
    double dt;
    double g;
    double ramp_duration;    /* days; 0 = no ramp */
    int32_t problem;         /* RKDG_PROBLEM_* : initial condition / function BC / truth */
    int32_t ic_constant;     /* 1: constant IC (ze0,qx0,qy0) projected nodally; 0: ic_q(0,x) */
    double ze0, qx0, qy0;
    int32_t function_source; /* 1: add source_q(t,x) (manufactured) */
    int32_t friction;        /* RKDG_FRICTION_* */
    double friction_cf;      /* Chezy coefficient / Manning floor */
    int32_t wetting_drying;
    double h_o;
    int32_t slope_limiting;  /* Cockburn-Shu */
    double sl_M;
    double sl_nu;
    double problem_param[4]; /* problem-specific constants (dam break: x0, h_left, h_right) */
    /* Spherical coordinates (SWE::SphericalProjection, swe_inputs.hpp:10-15).  spherical_mesh:
     * node x / y are longitude / latitude in degrees, projected onto a cylinder before anything
     * else, x = R cos(lat_o) (lon - lon_o) pi / 180, y = R lat pi / 180
     * (swe_pre_preprocess_mesh.hpp:5-27).  sp_correction: the x-flux correction factor
     * sp = cos(lat_o) / cos(y / R) at every Gauss point (swe_pre_init_data.hpp:31-48, 80-160),
     * applied to the volume x-fluxes and in the LLF flux (rkdg_swe_LLF_flux.hpp:14-45); otherwise
     * sp = 1.  Zero-initialised parameters keep both off; the reference's defaults are R =
     * 6378200 m, lon_o = lat_o = 0. */
    int32_t spherical_mesh;
    int32_t sp_correction;
    double sph_R, sph_lon_o, sph_lat_o;
} rkdg_params;

typedef struct rkdg_perf_report {
    int64_t n_elements;
    int32_t p, ndof, nstages;
    int32_t precision;
    int64_t stages_timed;        /* stages executed since the last rkdg_reset_timers */
    double stage_kernel_ms;      /* summed HIP-event time of the fused stage kernel */
    double limiter_kernel_ms;    /* summed HIP-event time of the limiter kernels */
    double total_ms;             /* HIP-event time across whole rkdg_step calls */
    double stage_bytes_alg;      /* algorithmic HBM bytes per stage launch */
    int32_t regs_per_thread;     /* unused (0) unless filled by a profiler */
} rkdg_perf_report;

/* ---- context lifecycle ---------------------------------------------------------- */

/* Builds mesh connectivity (interfaces, boundaries), FP64 geometry, master tables and the
 * initial state, then uploads everything to device HBM.
 * Replaces: initialize_mesh (preprocessor/initialize_mesh.hpp:19-77),
 *   SWE::RKDG::Problem::create_interfaces/create_boundaries (rkdg_swe_problem.hpp:57-113),
 *   preprocessor_serial + initialize_data_serial (kernels_preprocessor/rkdg_swe_pre_serial.hpp:8-16,
 *   problem_preprocessor/swe_pre_init_data.hpp:8-327) and the ESSPRKStepper ctor
 *   (simulation/stepper/explicit_ssp_rk_stepper.cpp:3-14).
 * RKDG_ERR_UNSUPPORTED for a polynomial order the kernels are not built for (p = 0 or p > 3),
 * RKDG_ERR_ARG for any other invalid description (message from rkdg_create_error): negative
 * counts, positive counts with NULL arrays, node indices out of range, non-finite coordinates or
 * depths, zero-area elements, edges of more than two elements or without a boundary type, unknown
 * problem, friction law or SSP-RK scheme.  The description is checked before the device is looked
 * for, so these errors come first on any machine; RKDG_ERR_NO_DEVICE after them. */
rkdg_status rkdg_create(const rkdg_mesh_desc* mesh, const rkdg_params* params, rkdg_precision precision,
                        rkdg_ctx** out);
void rkdg_destroy(rkdg_ctx* ctx);
const char* rkdg_last_error(const rkdg_ctx* ctx);
/* Message of the last failure of a call without a context: rkdg_create, the mesh readers, writers
 * and partitioner. */
const char* rkdg_create_error(void);

/* ---- state ------------------------------------------------------------------------ */

/* Overwrites the current stage's state (state[stage].q: state[0].q at a step boundary, the one
 * rkdg_get_state reads) and the wet flags if wet != NULL; replaces direct writes into
 * SWE::Data::state (problem_data_structure/swe_data.hpp:13-70).  Writing q also clears the NaN
 * record, so a context that returned RKDG_ERR_NAN steps again from the new state (rkdg_nan_info
 * reports none until a new NaN is found). */
rkdg_status rkdg_set_state(rkdg_ctx* ctx, const double* q, const uint8_t* wet);
/* Reads the current-stage state.  limited: [3][n_elements] per-variable slope-limiter
 * write-back flags of the last limiter pass (swe_CS_slope_limiter.hpp:180-187);
 * troubled: slope_limit_state.troubled (swe_trouble_check.hpp:106-115).  Any pointer may be NULL. */
rkdg_status rkdg_get_state(rkdg_ctx* ctx, double* q, uint8_t* wet, uint8_t* limited, uint8_t* troubled);

/* ---- time stepping ---------------------------------------------------------------- */

/* nsteps full SSP-RK steps (Problem::step_serial, kernels_processor/rkdg_swe_proc_serial_step.hpp:9-26),
 * each stage = volume, source, interfaces, boundaries, M^-1 + update, ++stepper, wetting-drying,
 * slope limiter, NaN scan (stage_serial, :28-67).  Returns RKDG_ERR_NAN if a NaN was found;
 * rkdg_nan_info then names where. */
rkdg_status rkdg_step(rkdg_ctx* ctx, int32_t nsteps);
/* One stage (stage_serial), for per-stage parity checks. */
rkdg_status rkdg_stage(rkdg_ctx* ctx);
/* Stepper bookkeeping: t, step, stage (ESSPRKStepper, explicit_ssp_rk_stepper.hpp:42-47). */
rkdg_status rkdg_get_time(rkdg_ctx* ctx, double* t, int32_t* step, int32_t* stage);
/* Sets the stepper's time and step counter at a step boundary (stage 0), for restarts from a
 * saved state; no reference counterpart (the reference has no resume, its modal output
 * swe_post_write_modal.hpp:5-90 has no reader). */
rkdg_status rkdg_set_time(rkdg_ctx* ctx, double t, int32_t step);
/* First NaN found: element ID, step and stage (scrutinize_solution, swe_post_scrutinize.hpp:5-41). */
rkdg_status rkdg_nan_info(rkdg_ctx* ctx, int32_t* element, int32_t* step, int32_t* stage);
/* Waits for all work queued on the context's stream. */
rkdg_status rkdg_sync(rkdg_ctx* ctx);

/* ---- readouts --------------------------------------------------------------------- */

/* sqrt(sum_e int_e |true_q(t) - q|^2) on Dunavant rule 20 at the current stage time.
 * Replaces Simulation::ComputeL2Residual (simulation/serial/simulation.hpp:85-93) and
 * Element::ComputeResidualL2 (geometry/element.hpp:547-592).  Host-side FP64. */
rkdg_status rkdg_l2_residual(rkdg_ctx* ctx, double* l2);
/* Per-kernel HIP-event times and the algorithmic byte count of one stage launch. */
rkdg_status rkdg_report(rkdg_ctx* ctx, rkdg_perf_report* report);
rkdg_status rkdg_reset_timers(rkdg_ctx* ctx);
/* Starts a timing window as rkdg_reset_timers does; with per_kernel = 0 no event is recorded
 * between kernel launches (the launches of a step stay back to back on the stream) and only the
 * HIP events around each rkdg_step call are kept (rkdg_perf_report.total_ms); rkdg_reset_timers
 * is rkdg_set_timing(ctx, 1). */
rkdg_status rkdg_set_timing(rkdg_ctx* ctx, int32_t per_kernel);

/* ---- host-side helpers (no device work) ------------------------------------------- */

/* Reference-element tables of the device path for order p (phi, dphi at the rule-2p points,
 * Dunavant weights), written into caller arrays sized [ndof*ngp] / [ngp].  Lets tests compare
 * the engine's basis against the oracle bit for bit (Master::Triangle ctor, master_triangle.tpp:5-87). */
rkdg_status rkdg_master_tables(int32_t p, int32_t* ndof, int32_t* ngp, double* phi, double* dphi_z1,
                               double* dphi_z2, double* w);

/* Shu-Osher coefficients ark/brk ([nstages][nstages], zero outside the scheme) and stage-time
 * offsets drk of an explicit SSP-RK scheme (ESSPRKStepper::InitializeCoefficients,
 * simulation/stepper/explicit_ssp_rk_stepper.cpp:16-276). */
rkdg_status rkdg_ssprk_coeffs(int32_t nstages, int32_t order, double* ark, double* brk, double* drk);

/* Structured quad-split triangle mesh over a rectangle, the reference's quad_mesh_generator
 * (mesh_generators/quad_mesh_generator.cpp:91-279, patterns :464-541) extended with seeded
 * vertex jitter and seeded Fourier-mode bathymetry.  Coordinates are rounded through "%.12f"
 * text exactly as a written/read .14 file would be (quad_mesh_generator.cpp:164). */
typedef struct rkdg_quad_mesh_spec {
    double x1, y1, x2, y2;      /* rectangle corners (lower-left, upper-right) */
    int32_t m, n;               /* quads in x, y */
    int32_t pattern;            /* 0 simple, 1 zigzag, 2 checker */
    int32_t bc[4];              /* RKDG_BC_* for sides: y=y1, x=x2, y=y2, x=x1 */
    int32_t bathy_kind;         /* 0 constant(b0), 1 bowl -alpha*(x^2+y^2), 2 Fourier modes, 3 step */
    double b0, alpha;           /* constant depth / bowl coefficient */
    double bmin, bmax;          /* Fourier clamp range */
    int32_t n_modes;            /* Fourier modes (<= 0: 16; at most 4096) */
    double jitter;              /* interior vertex jitter as a fraction of spacing */
    uint64_t seed;
    double step_x, step_b_left, step_b_right; /* bathy_kind 3: depth b_left for x<step_x else b_right */
    /* one tidal/flow constituent for tide/flow sides */
    double frequency, forcing_fact, equilib_arg, amplitude, phase;
    double manning_n;           /* >0: uniform nodal Manning n */
} rkdg_quad_mesh_spec;

typedef struct rkdg_mesh rkdg_mesh; /* owns the arrays a rkdg_mesh_desc points into */
/* RKDG_ERR_ARG for m or n < 1, an empty rectangle, an unknown pattern, bathymetry kind or side
 * type (land, tide, flow, function), or more than 4096 Fourier modes. */
rkdg_status rkdg_mesh_generate(const rkdg_quad_mesh_spec* spec, rkdg_mesh** out, rkdg_mesh_desc* desc);

/*
 * Ensembles: several independent meshes stepped together (no counterpart in the reference, which
 * runs one mesh per process).  Instances of equal polynomial order are merged into one element
 * range, each padded to whole 256-element tiles, and stepped by one stage launch per order;
 * each keeps its own dt, time, ramp and forcing.  Instances of one order need equal Runge-Kutta
 * scheme (at least 2 stages), gravity, friction, wetting-drying, limiter and problem options and
 * one constituent list.  A batch context accepts rkdg_step, rkdg_sync, rkdg_reset_timers,
 * rkdg_report and the rkdg_batch_* calls; instance state access needs a step boundary.
 */
typedef struct rkdg_schedule {
    int32_t block;          /* work-items per block: 64, 128 or 256 */
    int32_t groups_per_xcd; /* 0: one work-group per element tile; > 0: persistent stage kernel with
                               this many work-groups on each of the 8 XCDs; -1: persistent stage
                               kernel on as many work-groups as are resident at once on every CU
                               (from the kernel's occupancy at this block size, per order) */
    int32_t graph;          /* 1: each step replays one captured hipGraph (all stages, all orders) */
} rkdg_schedule;

rkdg_status rkdg_create_batch(int32_t n, const rkdg_mesh_desc* meshes, const rkdg_params* params,
                              rkdg_precision precision, rkdg_ctx** out);
int32_t rkdg_batch_size(const rkdg_ctx* ctx);
rkdg_status rkdg_batch_get_state(rkdg_ctx* ctx, int32_t instance, double* q, uint8_t* wet, uint8_t* limited,
                                 uint8_t* troubled);
rkdg_status rkdg_batch_set_state(rkdg_ctx* ctx, int32_t instance, const double* q, const uint8_t* wet);
rkdg_status rkdg_batch_get_time(rkdg_ctx* ctx, int32_t instance, double* t, int32_t* step, int32_t* stage);
/* First NaN found (instance, element within it, step, stage), or all -1. */
rkdg_status rkdg_batch_nan_info(rkdg_ctx* ctx, int32_t* instance, int32_t* element, int32_t* step, int32_t* stage);
rkdg_status rkdg_batch_set_schedule(rkdg_ctx* ctx, const rkdg_schedule* schedule);
rkdg_status rkdg_batch_get_schedule(const rkdg_ctx* ctx, rkdg_schedule* schedule);
/* Times steps_per_trial steps under each candidate schedule (block 64/128/256 x one work-group
 * per tile or the persistent kernel on the resident work-groups (groups_per_xcd = -1) x graph
 * off/on) from the same state, keeps the fastest, and restores the state; best_ms_per_step may
 * be NULL. */
rkdg_status rkdg_batch_autotune(rkdg_ctx* ctx, int32_t steps_per_trial, rkdg_schedule* best,
                                double* best_ms_per_step);

/* Reads an ADCIRC mesh (.14) and, if bcis_path is not NULL, its boundary-condition /
 * interface-specialisation data (.bcis) into a mesh description.  Replaces AdcircFormat
 * (preprocessor/ADCIRC_reader/adcirc_format.cpp:3-130) with its boundary typing
 * AdcircFormat::get_ibtype / get_internal_node_pair (:214-336) and MeshMetaData
 * (preprocessor/mesh_metadata.cpp:3-85), and SWE::Inputs::read_bcis (swe_inputs.cpp:212-312).
 * Node and element IDs become their rank among the file's IDs (an order-preserving relabel).
 * Tide and flow blocks must share one constituent list.  Errors: RKDG_ERR_ARG with the
 * message in rkdg_create_error(). */
rkdg_status rkdg_mesh_read_adcirc(const char* fort14_path, const char* bcis_path, rkdg_mesh** out,
                                  rkdg_mesh_desc* desc);

/* Frees a mesh from rkdg_mesh_generate or rkdg_mesh_read_adcirc. */
void rkdg_mesh_free(rkdg_mesh* mesh);

/* Writes a mesh description as an ADCIRC mesh (.14) and, if bcis_path is not NULL, its .bcis,
 * in the layout AdcircFormat::write_to (adcirc_format.cpp:151-212) and SWE::Inputs::read_bcis
 * (swe_inputs.cpp:212-312) use; `name` (NULL: "mesh") is the title line.  Node and element IDs
 * are their indices; reals are printed with 17 significant digits, so rkdg_mesh_read_adcirc
 * returns the same nodes, elements, boundary codes, levee pairs and forcing.  Boundary edges of
 * one code are chained into node strings: tide -> open boundaries, land / flow / function /
 * outflow (code 4) -> IBTYPE 0 / 2 / 77 / 88, levee pairs -> IBTYPE 4 barriers (split where
 * consecutive pairs' front or back nodes are not joined by a levee edge; every levee edge must
 * join two consecutive pairs).  The .bcis holds one tide and one flow block (the nodes of those
 * edges, all constituents) and the levee block.  Node Manning n has no place in these files and
 * is not written.  Errors: RKDG_ERR_ARG with the message in rkdg_create_error(). */
rkdg_status rkdg_mesh_write_adcirc(const rkdg_mesh_desc* mesh, const char* name, const char* fort14_path,
                                   const char* bcis_path);

/* MeshMetaData files (.meta, the partitioner's per-submesh mesh format): the layout of
 * MeshMetaData::write_to and MeshMetaData(const std::string&) (preprocessor/mesh_metadata.cpp:
 * 87-137, records mesh_metadata.hpp:11-72): element records "ID 3 nodes neighbours types" with
 * face k opposite node k, neighbour 4294967295 on boundary faces (the element across the barrier
 * on levee faces) and each face type one raw byte (SWE::BoundaryTypes, 128 = interior), then node
 * records "ID x y z".  rkdg_mesh_write_meta prints reals with 17 significant digits; IDs are
 * indices.  rkdg_mesh_read_meta relabels IDs by rank (as rkdg_mesh_read_adcirc), lists boundary
 * faces in element and face order, and takes forcing and levee pairs from the .bcis (NULL: none;
 * levee faces then are an error).  Errors: RKDG_ERR_ARG with the message in rkdg_create_error(). */
rkdg_status rkdg_mesh_write_meta(const rkdg_mesh_desc* mesh, const char* name, const char* meta_path);
rkdg_status rkdg_mesh_read_meta(const char* meta_path, const char* bcis_path, rkdg_mesh** out, rkdg_mesh_desc* desc);

/* ---- partitions and halo exchange (the distributed-boundary path) ----------------------
 * Replaces the reference's partitioned run (partitioner/partition.cpp; per-submesh .meta +
 * .dbmd files, mesh_metadata.cpp:139-223; the per-stage exchange of distributed-boundary data,
 * rkdg_swe_proc_ompi_step.hpp:38-102) for element-centric stages: a submesh holds its part's
 * elements and ghost copies of the neighbouring elements of other parts, all in global element
 * order, so each of its elements sees exactly the operands it sees in the whole mesh.  Before
 * every stage the ghosts must hold their owners' current state: after rkdg_stage on every part,
 * each part packs the elements other parts hold as ghosts (rkdg_halo_pack) and each part unpacks
 * the buffers of its owners (rkdg_halo_unpack).  A partitioned run then equals the whole-mesh
 * run bit for bit in every precision.  With the slope limiter the exchange also runs between the
 * stage kernel and the limiter (rkdg_stage_begin / rkdg_stage_finish); the state then still equals
 * the whole-mesh run's, while the `troubled` output flags of elements near a cut may differ (the
 * trouble check's dilation does not cross parts; RKDG never reads those flags).  Levees may cross
 * parts: the element across a levee face is a ghost like an interface neighbour (the reference's
 * distributed levee, rkdg_swe_dbc_distributed_levee.hpp).  Not for ensembles. */

/* One stage in two halves, for a partitioned run with the slope limiter: rkdg_stage_begin runs
 * the stage kernel (the state then is the new stage's, before limiting), rkdg_stage_finish the
 * limiter (with its NaN scan); rkdg_stage is both.  A partitioned run exchanges its halo after
 * each half. */
rkdg_status rkdg_stage_begin(rkdg_ctx* ctx);
rkdg_status rkdg_stage_finish(rkdg_ctx* ctx);
/* Recursive coordinate bisection of the element barycentres into nparts parts of equal size
 * (+-1), along the longer extent at each level: part[e] in [0, nparts). */
rkdg_status rkdg_partition_rcb(const rkdg_mesh_desc* mesh, int32_t nparts, int32_t* part);
/* Submesh of part p: its elements and their neighbours in other parts (ghosts; the element across
 * a levee face counts as a neighbour), in global element order; nodes in global order; ghost faces
 * leading out of the submesh are land (levee faces included); the barrier node pairs whose nodes
 * are both in the submesh are kept.
 * *global_element (may be NULL) receives the global index of each local element (owned by the
 * returned mesh; free with rkdg_mesh_free). */
rkdg_status rkdg_submesh(const rkdg_mesh_desc* mesh, const int32_t* part, int32_t p, rkdg_mesh** out,
                         rkdg_mesh_desc* sub, const int32_t** global_element);
/* Locality renumbering: elements in the order of a Hilbert curve through their barycentres, nodes
 * in order of first use (unused nodes last), every per-node and per-edge array carried along, so
 * neighbouring elements sit close in memory (the stage kernel's neighbour gathers then hit in L2
 * on meshes with an arbitrary file order).  *old_element (may be NULL) receives, per new element,
 * its index in `mesh` (owned by the returned mesh).  The renumbered mesh is the same
 * discretisation; results differ from the original order only by rounding (which side of an
 * interface is the lower ID, and face order, follow the numbering). */
rkdg_status rkdg_mesh_renumber(const rkdg_mesh_desc* mesh, rkdg_mesh** out, rkdg_mesh_desc* renumbered,
                               const int32_t** old_element);
/* Halo slots (up to 64 per context): slot s lists n local elements.  The slot's device buffer
 * holds [3 ndof + 1][n] values of the context's precision (double for the FP64 forms, float for
 * FP32): the modes of the current state and the wet flag (0/1).  pack: current state -> buffer;
 * unpack: a buffer of the same layout (another context's slot, same precision and order) ->
 * current state of the listed elements (`bytes`: the source buffer's size, at least the slot's).
 * Both are stream-ordered on ctx's stream; order them across contexts with rkdg_sync. */
rkdg_status rkdg_halo_define(rkdg_ctx* ctx, int32_t slot, const int32_t* local_elements, int32_t n);
rkdg_status rkdg_halo_buffer(rkdg_ctx* ctx, int32_t slot, void** device_buffer, int64_t* bytes);
rkdg_status rkdg_halo_pack(rkdg_ctx* ctx, int32_t slot);
rkdg_status rkdg_halo_unpack(rkdg_ctx* ctx, int32_t slot, const void* device_buffer, int64_t bytes);

/* ---- output ----------------------------------------------------------------------------
 * The reference's writers, at a step boundary, single-mesh contexts.
 * rkdg_write_modal: write_modal_data (problem_postprocessor/swe_post_write_modal.hpp:5-90):
 *   <prefix>_modal_{ze,qx,qy,bath}.txt, created at step 0 and appended to afterwards; each record
 *   is the time, then "ID value" per element and mode.
 * rkdg_write_vtu: write_VTU_data + the geometry of Writer::InitializeMeshGeometryVTU
 *   (swe_post_write_vtu.hpp:5-103, simulation/writer.hpp:260-350): <prefix>_data_<step>.vtu.
 * rkdg_write_vtk: write_VTK_data + the geometry of Writer::InitializeMeshGeometryVTK
 *   (swe_post_write_vtk.hpp:5-95, simulation/writer.hpp:195-258): <prefix>_data_<step>.vtk.
 * prefix = output path + mesh name, as the reference composes it. */
rkdg_status rkdg_write_modal(rkdg_ctx* ctx, const char* prefix);
rkdg_status rkdg_write_vtu(rkdg_ctx* ctx, const char* prefix);
rkdg_status rkdg_write_vtk(rkdg_ctx* ctx, const char* prefix);

/* ---- learned troubled-cell indicator ------------------------------------------------
 * No counterpart in the reference.  A 36 -> 64 -> 64 -> 64 -> 1 ReLU MLP predicts, per element,
 * whether the Cockburn-Shu limiter (slope_limiting_kernel, swe_CS_slope_limiter.hpp:88-189) would
 * write back a change that matters: larger than 1e-8 of the element's physical scale (FP64 forms)
 * or 1e-6 (FP32 bulk, that form's own rounding level per step); with an indicator set, the
 * limiter's trouble check and dilation passes are skipped and its limit pass runs only where the
 * network's logit exceeds `threshold` (elements where it runs are limited exactly as by the plain
 * limiter).  Inputs: modes 0-2 of ze/qx/qy of the element and its three neighbours (features as
 * in dgswemv2_amd/csrc/indicator.hpp).  Inference and training run as BF16 MFMA products with
 * FP32 accumulation. */
#define RKDG_MLP_IN 36
#define RKDG_MLP_HIDDEN 64
typedef struct rkdg_mlp_weights {
    float w1[RKDG_MLP_HIDDEN * RKDG_MLP_IN]; /* row-major [out][in] */
    float b1[RKDG_MLP_HIDDEN];
    float w2[RKDG_MLP_HIDDEN * RKDG_MLP_HIDDEN];
    float b2[RKDG_MLP_HIDDEN];
    float w3[RKDG_MLP_HIDDEN * RKDG_MLP_HIDDEN];
    float b3[RKDG_MLP_HIDDEN];
    float w4[RKDG_MLP_HIDDEN];
    float b4;
    float threshold; /* gate opens where logit > threshold */
} rkdg_mlp_weights;

/* Counters since the indicator was set (or since the last read with reset): element-stages. */
typedef struct rkdg_indicator_stats {
    int64_t stages;             /* limiter stages run with the indicator */
    int64_t candidates;         /* element and neighbours wet (the limiter may act) */
    int64_t flagged;            /* gate open */
    /* with judging on (the plain limiter's decision evaluated on the same state): */
    int64_t write_backs;        /* the plain limiter writes back */
    int64_t significant;        /* ... a change that matters (above 1e-8 / 1e-6 of the physical scale) */
    int64_t missed;             /* write-backs with the gate closed */
    int64_t missed_significant; /* significant write-backs with the gate closed */
    double max_missed_change;   /* largest relative change among missed write-backs */
} rkdg_indicator_stats;

/* Sets (w != NULL) or removes (w == NULL: the plain limiter again) the indicator of a single-mesh
 * context with the slope limiter on.  judge != 0 also runs the plain limiter's decision as a dry
 * run each stage and counts misses (slower; for validation). */
rkdg_status rkdg_set_indicator(rkdg_ctx* ctx, const rkdg_mlp_weights* w, int32_t judge);
rkdg_status rkdg_get_indicator_stats(rkdg_ctx* ctx, rkdg_indicator_stats* stats, int32_t reset);
/* With capture on, every limiter stage records, for its state before limiting, the features
 * ([36][n_elements]), the label (1 where the plain limiter writes back a change that matters, as
 * above) and the size of that change (relative to the physical scale, 0 where the limiter does not
 * write back), whether the limiter may act (element and neighbours wet), and with an indicator
 * set the logits.  rkdg_get_capture reads the last stage's record (NULL pointers are skipped). */
rkdg_status rkdg_capture(rkdg_ctx* ctx, int32_t on);
rkdg_status rkdg_get_capture(rkdg_ctx* ctx, float* features, uint8_t* label, uint8_t* may_act, float* logit,
                             float* change);

/* Trainer: the same network trained on the GPU (weighted binary cross-entropy on the logit,
 * mean over the batch, weight pos_weight on positive samples; Adam on FP32 master weights). */
typedef struct rkdg_mlp_trainer rkdg_mlp_trainer;
typedef struct rkdg_train_params {
    int32_t batch;     /* samples per step (any size >= 1) */
    int32_t epochs;
    float lr, beta1, beta2, eps;
    float pos_weight;
    uint64_t seed;     /* sample order */
} rkdg_train_params;
/* capacity: the largest sample count rkdg_mlp_set_data / forward / gradient will be given. */
rkdg_status rkdg_mlp_trainer_create(const rkdg_mlp_weights* init, int32_t capacity, rkdg_mlp_trainer** out);
/* Training set: x [36][n] (feature-major), y [n] in {0, 1}. */
rkdg_status rkdg_mlp_set_data(rkdg_mlp_trainer* t, const float* x, const float* y, int32_t n);
/* Runs epochs over the training set; epoch_loss[epochs] (may be NULL): mean batch loss per epoch. */
rkdg_status rkdg_mlp_train(rkdg_mlp_trainer* t, const rkdg_train_params* prm, float* epoch_loss);
/* Current weights (threshold as given at create). */
rkdg_status rkdg_mlp_get_weights(rkdg_mlp_trainer* t, rkdg_mlp_weights* out);
/* Logits of x [36][n] under the current weights (the trainer's forward kernel). */
rkdg_status rkdg_mlp_forward(rkdg_mlp_trainer* t, const float* x, int32_t n, float* logit);
/* Loss and gradient (weights layout; threshold 0) of one batch x [36][n], y [n], n <= capacity. */
rkdg_status rkdg_mlp_gradient(rkdg_mlp_trainer* t, const float* x, const float* y, int32_t n, float pos_weight,
                              float* loss, rkdg_mlp_weights* grad);
/* Allocation-free variant of rkdg_mlp_forward: logits of x [36][n] (host) under the weights w
 * (host), with no trainer.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc / hipFree: its staging (parameters, packed weights, features,
 * logits; about 2.5 MB) is a fixed device array that each GPU holds for the life of the process.
 * n <= RKDG_MLP_FORWARD_NOALLOC_MAX, else RKDG_ERR_ARG.  Calls in one process share that array
 * and are serialised by a lock (concurrent callers wait); runs on the current device's null
 * stream and returns when the logits are on the host.  Same kernel and results as
 * rkdg_mlp_forward on the same weights. */
#define RKDG_MLP_FORWARD_NOALLOC_MAX 16384
rkdg_status rkdg_mlp_forward_noalloc(const rkdg_mlp_weights* w, const float* x, int32_t n, float* logit);
const char* rkdg_mlp_trainer_error(const rkdg_mlp_trainer* t);

void rkdg_mlp_trainer_destroy(rkdg_mlp_trainer* t);

#ifdef __cplusplus
}
#endif

#endif /* RKDG_H */
