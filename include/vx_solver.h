/*
 * vx_solver.h — C ABI of the MI355X (gfx950) corotational elastic PCG solver.
 *
 * Drop-in boundary: vx_solve_pcg replaces the reference's device solve
 *   pub(crate) fn solve_pcg(rhs: &[Vec3], neighbors: &[[u32; 6]],
 *       diagonal: &[[f32; 9]], blocks: &[[f32; 54]],
 *       inverse_diagonal: &[[f32; 9]], max_iterations: usize, tolerance: f32)
 *       -> Result<(Vec<Vec3>, usize, f32), String>
 * (src/cuda_solver.rs:301-309, called from src/solver.rs:1778-1786).
 * Plain pointers and sizes only; the caller owns every buffer.
 *
 * Every function returns VX_OK (0) or a nonzero status; vx_last_error()
 * returns the calling thread's message for the last nonzero status. The
 * nonzero statuses correspond to the reference's Err branches, so a caller
 * keeps its CPU fallback (src/solver.rs:1792-1808).
 *
 * All entry points serialise on one process-wide device context (device 0,
 * one HIP stream), as the reference serialises on its CUDA context mutex
 * (src/cuda_solver.rs:66, 326-336).
 */
#ifndef VX_SOLVER_H
#define VX_SOLVER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define VX_OK 0
/* Device/runtime initialisation failed; cached for the process lifetime like
 * the reference's OnceLock (src/cuda_solver.rs:326-332). */
#define VX_ERR_INIT 1
/* Problem does not fit: the reference's "at most u32::MAX active REV"
 * (src/cuda_solver.rs:314-315), or device memory exhausted. */
#define VX_ERR_TOO_LARGE 2
/* "produced a non-finite residual norm" (src/cuda_solver.rs:470-472, 484-486). */
#define VX_ERR_NONFINITE_RESIDUAL 3
/* "did not converge" (src/cuda_solver.rs:487-492). */
#define VX_ERR_NOT_CONVERGED 4
/* "produced a non-finite correction" (src/cuda_solver.rs:493-495). */
#define VX_ERR_NONFINITE_CORRECTION 5
/* Malformed arguments (null pointer with n > 0, a neighbour index that is
 * neither EMPTY nor < n, unknown handle). No reference counterpart: the
 * reference would read out of bounds on such input. */
#define VX_ERR_INVALID_ARGUMENT 6
/* A HIP runtime call failed after initialisation. */
#define VX_ERR_DEVICE 7

#define VX_EMPTY 0xFFFFFFFFu

/* Thread-local message of the last nonzero status returned on this thread. */
const char *vx_last_error(void);

/* Initialise the device context (optional; every entry point does it). */
int32_t vx_init(void);

/* Measurement aid, no counterpart in the reference: the median duration
 * (microseconds) of `reps` launches of a STREAM copy kernel reading `bytes`
 * and writing as many (device buffers allocated and freed by the call), the
 * bandwidth a plain copy reaches beside the solver kernels' rooflines. */
int32_t vx_copy_bandwidth(uint64_t bytes, uint32_t reps, double *out_us);

/* Measurement aid, no counterpart in the reference: how the PCG iteration
 * chunks reach the device. VX_LAUNCH_GRAPHS (the default) replays one
 * captured hipGraph per chunk; VX_LAUNCH_DIRECT launches the same kernels with
 * the same arguments in the same order one by one, for tools that trace
 * kernel dispatches. Results are identical bit for bit in both modes. */
#define VX_LAUNCH_GRAPHS 0
#define VX_LAUNCH_DIRECT 1
int32_t vx_set_launch_mode(int32_t mode);

/* ---- drop-in solve ----------------------------------------------------- */
/*
 * vx_solve_pcg — replaces cuda_solver::solve_pcg (src/cuda_solver.rs:301-501).
 *   rhs               n*3   f32, right-hand side (reference Vec3 layout)
 *   neighbors         n*6   u32, per-row neighbour slots, VX_EMPTY if absent
 *   diagonal          n*9   f32, row-major 3x3 diagonal blocks
 *   blocks            n*54  f32, 6 row-major 3x3 off-diagonal blocks per row
 *   inverse_diagonal  n*9   f32, row-major inverse of each diagonal block
 *   n                 active rows
 *   max_iterations    cap (the reference passes config.cuda_iterations)
 *   tolerance         relative residual ||r|| / ||rhs||
 *   out_correction    n*3   f32, caller-owned
 *   out_iterations, out_relative_residual
 * Semantics follow the reference device solve: block-Jacobi PCG with f32
 * vectors, f64 dot products and f32 alpha/beta, the residual checked every
 * 16 iterations and at the cap with `<=`. n == 0 and ||rhs|| <= 1e-12 return
 * VX_OK with a zero correction without touching the device.
 */
int32_t vx_solve_pcg(const float *rhs, const uint32_t *neighbors, const float *diagonal,
                     const float *blocks, const float *inverse_diagonal, uint32_t n,
                     uint32_t max_iterations, float tolerance, float *out_correction,
                     uint32_t *out_iterations, float *out_relative_residual);

/*
 * vx_solve_pcg_f64 — the plain FP64 GPU solve over the same reference
 * layout: inputs widened to f64, block inverses recomputed in f64 from the
 * diagonal, FP64 vectors, scalars and reductions, same convergence checks and
 * statuses as vx_solve_pcg. No reference counterpart (the reference's device
 * solve is f32); it is the drop-in comparison baseline.
 */
int32_t vx_solve_pcg_f64(const float *rhs, const uint32_t *neighbors, const float *diagonal,
                         const float *blocks, const float *inverse_diagonal, uint32_t n,
                         uint32_t max_iterations, float tolerance, float *out_correction,
                         uint32_t *out_iterations, float *out_relative_residual);

/*
 * vx_solve_pcg_noalloc — vx_solve_pcg without any device allocation. No
 * counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, semantics, statuses and results (bit for bit)
 * as vx_solve_pcg. Every device buffer of the solve lives in one statically
 * allocated 64 MiB scratch area per device (held until the process ends),
 * which bounds n: about 90,000 rows; a larger n returns VX_ERR_TOO_LARGE
 * before touching the device. The PCG chunks are launched directly (no graph
 * is instantiated) and the host waits with stream synchronisations (no
 * events are created). Call vx_init() once at set-up: the first call of the
 * library initialises the device. The scratch area is shared by every call
 * on a device and keeps the last call's contents; calls are serialised by
 * the library's context lock (calls that overlapped on one device would
 * race on it).
 */
int32_t vx_solve_pcg_noalloc(const float *rhs, const uint32_t *neighbors, const float *diagonal,
                             const float *blocks, const float *inverse_diagonal, uint32_t n,
                             uint32_t max_iterations, float tolerance, float *out_correction,
                             uint32_t *out_iterations, float *out_relative_residual);

/* ---- device-resident structured scenes (build extensions) --------------- */
/*
 * A box scene of nx*ny*nz REVs kept resident in HBM: SoA state, on-device
 * rotation fit (src/solver.rs:327-417), bond assembly
 * (src/solver.rs:1542-1636), and a mixed-precision PCG (FP32 block-Jacobi
 * PCG with f64 reductions whose residual is replaced by the FP64 true
 * residual every time it has dropped by inner_tolerance). The
 * reference has no counterpart beyond solve_pcg; these entry points carry the
 * same status/error convention.
 *
 * Voxel index v = x + nx*(y + ny*z). Bonds are stored per +axis:
 * neighbor[a*n + v] is v + stride_a or VX_EMPTY.
 */
typedef struct vx_scene_desc {
    uint32_t nx, ny, nz;
    const float *positions;        /* n*3 */
    const float *rotations;        /* n*4 quaternion (x, y, z, w) */
    const float *velocities;       /* n*3 */
    const float *mass;             /* n */
    const uint8_t *kinematic;      /* n */
    const float *young;            /* n, modulus along the fibre */
    const float *young_perp;       /* n */
    const float *shear;            /* n */
    const float *solid;            /* n, solid fraction */
    const float *fibers;           /* n*3, body-frame fibre direction */
    const float *anisotropy;       /* n */
    const float *delta_temp;       /* n, temperature minus stress-free temperature */
    const float *alpha_par;        /* n */
    const float *alpha_perp;       /* n */
    const uint32_t *neighbor;      /* 3*n */
    const float *damage;           /* 3*n, may be NULL (zero) */
    const float *kscale;           /* 3*n, may be NULL (one) */
    const float *formation_offset; /* 3*n*3, may be NULL (zero) */
    const float *plastic_strain;   /* 3*n*3, may be NULL (zero) */
    const float *targets;          /* n*3 world coordinates, NaN = free axis; may be NULL */
    const float *loads;            /* n*3 persistent applied loads; may be NULL */
    /* Fracture inputs. strength: n*12 per-REV strength rows (Material
     * fields, src/material.rs:20-63) in VX_S_* order; required when a step
     * runs the fracture pass, else may be NULL. bond_history: 3*n*6 per +axis
     * bond in VX_H_* order, may be NULL (a new scene: zero). */
    const float *strength;
    const float *bond_history;
} vx_scene_desc;

/* Strength row fields (Material fields, src/material.rs:20-63). S_MODE is
 * VX_BRITTLE, VX_DUCTILE or VX_MIXED (FractureMode, src/material.rs:4-8). */
enum {
    VX_S_TENSION, VX_S_TENSION_PERP, VX_S_COMPRESSION, VX_S_COMPRESSION_PERP, VX_S_SHEAR, VX_S_SHEAR_PERP,
    VX_S_YIELD, VX_S_HARDENING, VX_S_PLASTIC_FAILURE, VX_S_MODE, VX_S_GC, VX_S_GC_PERP, VX_STRENGTH_FIELDS
};
#define VX_BRITTLE 0
#define VX_DUCTILE 1
#define VX_MIXED 2
/* Bond history fields beyond damage and plastic strain (Bond,
 * src/bond.rs:4-16). */
enum {
    VX_H_EQUIVALENT_PLASTIC_STRAIN, VX_H_PLASTIC_WORK, VX_H_FATIGUE_DAMAGE, VX_H_PREVIOUS_STRESS_RATIO,
    VX_H_REVERSAL_STRESS_RATIO, VX_H_LOADING_INCREMENT, VX_HISTORY_FIELDS
};

typedef struct vx_scene vx_scene;

typedef struct vx_solve_params {
    double tolerance;          /* outer FP64 relative true residual ||b - A x|| / ||b|| */
    double inner_tolerance;    /* FP32 residual drop between FP64 residual replacements (3e-3) */
    uint32_t max_inner;        /* inner iteration cap per solve */
    uint32_t max_outer;        /* FP64 residual replacements cap */
    uint32_t warm_start;       /* 1: start from the scene's last correction */
    uint32_t mode;             /* VX_SOLVE_MIXED or VX_SOLVE_F64 */
    uint32_t preconditioner;   /* VX_PRECOND_BLOCK_JACOBI, VX_PRECOND_RED_BLACK or VX_PRECOND_MULTIGRID */
    uint32_t reserved;         /* must be 0 */
} vx_solve_params;

#define VX_SOLVE_MIXED 0 /* FP32 PCG with FP64 residual replacement (reliable updates) */
#define VX_SOLVE_F64 1   /* plain FP64 PCG on the structured layout */

/* Preconditioners of the structured solve. Block-Jacobi is the reference's
 * device preconditioner (src/cuda_solver.rs:366-500, src/cuda/pcg.ptx:104-155).
 * Red-black: CG on the black rows' reduced system S = D_bb - A_br D_rr^-1 A_rb
 * (REVs coloured by the parity of x + y + z) with block-Jacobi, which is CG
 * preconditioned by red-black block symmetric Gauss-Seidel on the full system;
 * the red rows are recovered exactly, and the tolerance applies to the full
 * system's FP64 residual. It has no counterpart in the reference (a stronger
 * device preconditioner than block-Jacobi, cf. the CPU block IC(0) of
 * src/solver.rs:1390-1496). Iterations it reports are reduced-system
 * iterations: about half as many as block-Jacobi's, each three kernel
 * launches against two and about 1.1 times the HBM bytes. */
#define VX_PRECOND_BLOCK_JACOBI 0
#define VX_PRECOND_RED_BLACK 1
/* Aggregation multigrid: one V(1,1) cycle per iteration as the preconditioner
 * of the FP32 inner PCG (2x2x2 aggregates, Galerkin coarse operators with the
 * bond stencil's shape, damped block-Jacobi smoothing, the coarsest level of
 * at most 64 nodes solved by its Cholesky factor), rebuilt after every
 * assembly. One-scene arenas (vx_scene_*) and VX_SOLVE_MIXED only. No
 * counterpart in the reference: about 25x fewer iterations than
 * block-Jacobi, each a dozen small launches and about 2.2 times the bytes. */
#define VX_PRECOND_MULTIGRID 2

typedef struct vx_solve_report {
    uint32_t inner_iterations; /* inner PCG iterations of the solve */
    uint32_t outer_sweeps;     /* FP64 true-residual evaluations */
    uint32_t active_rows;
    uint32_t status;
    double relative_residual;  /* final FP64 true relative residual */
    double rhs_norm;
} vx_solve_report;

int32_t vx_scene_create(const vx_scene_desc *desc, vx_scene **out);
int32_t vx_scene_destroy(vx_scene *scene);
/* Predict (gravity, targets) for one substep of length h; stores the
 * predicted positions used by assembly. */
int32_t vx_scene_predict(vx_scene *scene, double h, const double *gravity3);
/* Rotation fit; writes the largest angle change. */
int32_t vx_scene_fit_rotations(vx_scene *scene, double *out_max_angle);
/* Assemble (M/h^2 + J^T K J) dx = -J^T K C for the current state. */
int32_t vx_scene_assemble(vx_scene *scene, double h);
/* Solve the assembled system; the correction stays on the device. */
int32_t vx_scene_solve(vx_scene *scene, const vx_solve_params *params, vx_solve_report *report);
/* positions += correction on the active rows. */
int32_t vx_scene_apply(vx_scene *scene);
/* Velocities from the solved positions after a substep of length h. */
int32_t vx_scene_finish_substep(vx_scene *scene, double h);
/* Copy state out (any pointer may be NULL). positions/velocities n*3,
 * rotations n*4, correction n*3 (zero on inactive rows). */
int32_t vx_scene_download(vx_scene *scene, double *positions, double *rotations, double *velocities,
                          double *correction);
/* Export the assembled system in the reference packing (src/solver.rs:
 * 1652-1672, 1731-1777) for the active rows: neighbours (active*6), blocks
 * (active*54 row-major), diagonal (active*9), rhs (active*3) in f64, and the
 * voxel index of each active row. Any pointer may be NULL. */
int32_t vx_scene_export_rows(vx_scene *scene, uint32_t *active, uint32_t *neighbors, double *blocks,
                             double *diagonal, double *rhs, uint32_t *out_active_count);
/* The multigrid hierarchy the last VX_PRECOND_MULTIGRID solve or step of the
 * scene built (VX_ERR_INVALID_ARGUMENT before the first): level `level` (1 =
 * the first coarse level) as its node grid (dims[3]; always written when
 * non-NULL, so a call with the other pointers NULL sizes the buffers; 0 0 0
 * past the coarsest level), its diagonal blocks (6 per node, xx xy xz yy yz
 * zz), its couplings to the next node along each axis (27 per node, [axis][3x3
 * row-major], A[I, I+e] = -C) and its free-component masks (bit c: component
 * c free), node-major, in FP32 as the cycle reads them. No counterpart in the
 * reference: for checking the Galerkin products. Any pointer may be NULL. */
int32_t vx_scene_multigrid_level(vx_scene *scene, uint32_t level, uint32_t *dims, float *diagonal, float *couplings,
                                 uint8_t *masks);
/* Mean duration (microseconds) of the FP32 structured bond SpMV kernel in a
 * graph of `reps` back-to-back launches on the library stream (launched one by
 * one under VX_LAUNCH_DIRECT), timed with HIP events; the median of five
 * replays. A microbenchmark: the kernel alone, without the update kernel
 * between launches that a solve has. */
int32_t vx_scene_time_spmv(vx_scene *scene, uint32_t reps, double *out_us);

/* Durations of one kernel's dispatches, from each dispatch's own begin and
 * end timestamps (the figures a kernel-dispatch trace reports). */
typedef struct vx_kernel_stats {
    uint32_t dispatches;
    uint32_t reserved;
    double mean_us, median_us, min_us, max_us, total_us;
} vx_kernel_stats;
typedef struct vx_kernel_times {
    vx_kernel_stats spmv;    /* pcg_bond_spmv: beta, p = z + beta p_old, Ap, p.Ap partials */
    vx_kernel_stats update;  /* pcg_update: alpha, d, r, z = D^-1 r, r.z and r.r partials */
} vx_kernel_times;
/* Measurement aid, no counterpart in the reference: vx_scene_solve with the
 * inner kernels launched one by one (no graph replay), every bond SpMV and
 * update dispatch timestamped. The solve's results are identical bit for bit
 * to vx_scene_solve's. The dispatches include those a chunk issues after the
 * scene's inner solve has ended (as in an untimed solve). */
int32_t vx_scene_solve_timed(vx_scene *scene, const vx_solve_params *params, vx_solve_report *report,
                             vx_kernel_times *times);

/* One fixed step of the solve part of step_with_axis_targets_and_loads
 * (src/solver.rs:475-570, 668-683) on the device: per substep predict, fit
 * rotations, up to corotation_iterations passes of assemble + solve + apply
 * with the reference's early exit, then (with fracture.enabled) the fracture
 * pass, then velocities. Contact, tool and heat are not part of it. */
/* FractureConfig (src/fracture.rs:12-25); enabled = 0 skips the pass. With
 * the pass on, every substep ends with fracture::update (src/fracture.rs:
 * 210-292) of each intact bond of a non-kinematic REV, after the corotation
 * passes and before the velocities (src/solver.rs:627-665); a bond whose
 * damage reaches one leaves the lattice. */
typedef struct vx_fracture_params {
    uint32_t enabled;
    uint32_t reserved;
    double damage_rate;
    double energy_release;
    double fatigue_cycles;
    double fatigue_exponent;
    double fatigue_threshold;
    double plastic_heat_fraction;
} vx_fracture_params;

/* Energies dissipated by the fracture pass (Dissipation, src/fracture.rs:
 * 135-141) and the bonds it broke. */
typedef struct vx_fracture_report {
    uint64_t bonds_broken;
    double plastic_heat;
    double stored_plastic;
    double surface_work;
    double damage_heat;
    double released;
    double seconds;              /* device time of the fracture kernels (HIP events) */
} vx_fracture_report;

typedef struct vx_step_params {
    double fixed_dt;
    double gravity[3];
    uint32_t substeps;
    uint32_t corotation_iterations;
    double corotation_tolerance;
    vx_solve_params solve;
    vx_fracture_params fracture;
} vx_step_params;

typedef struct vx_step_report {
    uint32_t solves;
    uint32_t inner_iterations;   /* summed over solves */
    uint32_t outer_sweeps;       /* summed over solves */
    uint32_t active_rows;
    double last_relative_residual;
    double last_angle;
    double seconds;              /* device wall time of the step */
    vx_fracture_report fracture; /* summed over the step's substeps */
} vx_step_report;

int32_t vx_scene_step(vx_scene *scene, const vx_step_params *params, vx_step_report *report);

/* Per-step inputs of step_with_axis_targets_and_loads (src/solver.rs:
 * 440-482): replace the scene's axis targets (n*3 world coordinates, NaN =
 * free axis; NULL: none) and applied loads (n*3 newtons; NULL: zero) for
 * the following steps. Targets must be finite where not NaN. The warm start
 * is dropped, since the solved axes may change. */
int32_t vx_scene_set_targets_and_loads(vx_scene *scene, const float *targets, const float *loads);

/* Copy the bond state out (any pointer may be NULL): neighbor 3*n (v +
 * stride or VX_EMPTY once broken), damage 3*n, plastic_strain 3*n*3 and
 * history 3*n*6 (VX_H_* order), all per +axis bond [axis][voxel]. */
int32_t vx_scene_download_bonds(vx_scene *scene, uint32_t *neighbor, float *damage, float *plastic_strain,
                                float *history);
/* The bond forces of the scene's last substep, per +axis bond, (3, n, 3)
 * [axis][voxel][component], f64; either pointer may be NULL. For each
 * constraint of the substep's last assembly (every bond except those whose
 * endpoints have no free axis and whose first REV is kinematic), at the
 * solved positions and before the fracture pass updates the bond:
 *   last_force  F (K (u - plastic CELL)) (1 - damage), zero for a kinematic
 *               first REV (the reference's BondWorkspace::last_force,
 *               src/solver.rs:627-640, read by its loading fixtures,
 *               src/experiments.rs:32-49);
 *   lambda      -F^T k (p_j - p_i - rest) h^2 (BondWorkspace::lambda,
 *               src/solver.rs:1824-1827, 1921-1927).
 * Zero for every other bond. Valid after vx_scene_step or
 * vx_scene_finish_substep (a later vx_scene_assemble overwrites them). */
int32_t vx_scene_download_bond_forces(vx_scene *scene, double *last_force, double *lambda);

/* ---- rigid tool (src/tool.rs:26-431) ------------------------------------
 * A rigid voxel group (the reference's Tool) driven into a scene: once set,
 * every vx_scene_step advances it per substep after the prediction, as the
 * reference's step does (src/solver.rs:530-543): microsteps of at most 0.4
 * REV of travel, per tool cell the scene REVs of the spatial bins its
 * oriented box touches, a separating-axis contact test, inelastic normal
 * impulses and Coulomb friction applied in the reference's order, then the
 * removal of the deepest overlap and up to six re-projections
 * (Tool::advance, src/tool.rs:192-431). The contact velocity drifts the REVs
 * by half a substep and enters their final velocities. After the substep's
 * velocities a zero-time advance applies impulses and projection once more
 * and discards its contact velocity (src/solver.rs:701-706). One tool per
 * vx_scene, advanced by vx_scene_step only (not by the stage calls); the
 * neighbourhood it can touch in one substep is limited to 262,144 REVs and
 * each cell to 32 contacts (VX_ERR_TOO_LARGE beyond), a substep to 65,536
 * microsteps (VX_ERR_INVALID_ARGUMENT); a tool thrown further than 256 times
 * its substep's margin fails the step with VX_ERR_DEVICE. */
typedef struct vx_tool_desc {
    uint32_t n_cells;
    uint32_t has_mount;            /* 1: guide mount (Tool::mount_base, src/tool.rs:127-142) */
    const float *cells;            /* n_cells*3, body frame about the centre of mass (m) */
    double mass;                   /* kg */
    double inverse_inertia[3];     /* body-frame principal inverse inertia */
    double position[3];            /* centre of mass, world */
    double rotation[4];            /* x, y, z, w */
    double velocity[3], angular_velocity[3];
    double external_force[3], external_torque[3];  /* persistent, world frame */
    double friction;               /* Coulomb coefficient (SceneConfig fracture.friction) */
    double mount_local_base[3], mount_world_base[3], mount_direction[3], mount_rotation[4];
} vx_tool_desc;
typedef struct vx_tool_state {
    double position[3], rotation[4], velocity[3], angular_velocity[3];
    double friction_heat;          /* J dissipated by tool friction since the tool was set */
    /* microsteps of the last substep's timed advance, contacts of both its
     * advances, candidate REVs of the last advance, status bits */
    uint32_t microsteps, contacts, candidates, status;
} vx_tool_state;
/* Set (or with NULL remove) the scene's tool. */
int32_t vx_scene_set_tool(vx_scene *scene, const vx_tool_desc *tool);
/* The tool's pose and velocities after the last step. */
int32_t vx_scene_tool_state(vx_scene *scene, vx_tool_state *state);

/* Inner PCG iterations per scheduling tick (one hipGraph replay; default 16,
 * even, at most 256). 0 keeps the current value. */
int32_t vx_scene_set_schedule(vx_scene *scene, uint32_t iterations_per_tick);

/* Launch shape of the scene's inner kernels (no reference counterpart; the
 * autotuner's second knob beside the schedule): which workgroup runs which
 * 512-row tile. zcols groups the tile columns of a z-plane so that a tile's
 * z neighbours run a few workgroups before and after it on the same XCD
 * (0: plain tile order; else a power of two up to 64; default 8); deal
 * spreads the tiles over the 8 XCDs in contiguous eighths or round robin,
 * for the SpMV and for the kernels after it; update_threads is the workgroup
 * of the block-Jacobi update kernel per 512-row tile (512: one row per
 * thread, the default; 256: two). Results are identical bit for bit under
 * every shape; only the time changes. */
#define VX_DEAL_AUTO 0        /* SpMV contiguous; the rest round robin above 1,024 tiles (default) */
#define VX_DEAL_CONTIGUOUS 1  /* every inner kernel in contiguous eighths */
#define VX_DEAL_SPLIT 2       /* SpMV contiguous, the kernels after it round robin */
#define VX_DEAL_ROUND_ROBIN 3 /* every inner kernel round robin */
int32_t vx_scene_set_shape(vx_scene *scene, uint32_t zcols, uint32_t deal, uint32_t update_threads);

/* ---- batched and served scenes (build extensions) ----------------------- */
/*
 * A batch is an HBM arena of 512-REV tiles shared by many independent box
 * scenes of uneven size. Every kernel of the solve (predict, fit, assembly,
 * FP64 residual, FP32 inner PCG, apply, velocities) is launched once per
 * scheduler tick for all scenes; each scene advances through its own steps,
 * corotation passes and refinement sweeps with per-scene scalars, early exit
 * and convergence decided on the device. Scenes may be admitted while others
 * run (continuous batching) as long as a contiguous tile range is free, and
 * released when done. The reference has no counterpart.
 */
typedef struct vx_batch vx_batch;

#define VX_SCENE_FREE 0
#define VX_SCENE_RUNNING 1
#define VX_SCENE_DONE 2
#define VX_SCENE_FAILED 3

typedef struct vx_batch_scene_report {
    uint32_t state;              /* VX_SCENE_* */
    uint32_t status;             /* VX_* status of a failed scene */
    uint32_t steps_done;
    uint32_t solves;
    uint64_t inner_iterations;   /* summed over solves */
    uint64_t outer_sweeps;       /* summed over solves */
    uint32_t rows;               /* REVs */
    uint32_t active_rows;
    double last_relative_residual;
    double last_angle;
    vx_fracture_report fracture; /* since admission */
} vx_batch_scene_report;

/* HBM bytes per arena row (for admission budgets). */
uint64_t vx_batch_bytes_per_row(void);
/* capacity_rows is rounded up to whole tiles; xcd_chunk_tiles is the tile
 * interleave across the 8 XCDs (0 = choose from the capacity). */
int32_t vx_batch_create(uint64_t capacity_rows, uint32_t max_scenes, uint32_t xcd_chunk_tiles, vx_batch **out);
int32_t vx_batch_destroy(vx_batch *batch);
/* Schedule of an arena: tiles per XCD chunk of the workgroup-to-tile
 * interleave and inner iterations per tick (even, at most 256); 0 keeps a
 * value. */
int32_t vx_batch_set_schedule(vx_batch *batch, uint32_t xcd_chunk_tiles, uint32_t iterations_per_tick);
/* Largest contiguous free range and total free space, in rows. */
int32_t vx_batch_free_rows(vx_batch *batch, uint64_t *largest, uint64_t *total);
/* Admit a scene that will run `steps` fixed steps with `params`. Returns
 * VX_ERR_TOO_LARGE (nothing admitted) when no contiguous range fits. */
int32_t vx_batch_admit(vx_batch *batch, const vx_scene_desc *desc, const vx_step_params *params, uint32_t steps,
                       uint32_t *out_slot);
/* One scheduler tick for all running scenes; writes how many still run. */
int32_t vx_batch_tick(vx_batch *batch, uint32_t *out_running);
int32_t vx_batch_report(vx_batch *batch, uint32_t slot, vx_batch_scene_report *report);
/* VX_SCENE_* state of every slot (states has max_scenes entries), as known
 * after the last tick; one call instead of a report per slot. */
int32_t vx_batch_states(vx_batch *batch, uint32_t *states);
/* Copy one scene's state out (any pointer may be NULL). */
int32_t vx_batch_download(vx_batch *batch, uint32_t slot, double *positions, double *rotations, double *velocities);
/* Bond state of one scene, as vx_scene_download_bonds. */
int32_t vx_batch_download_bonds(vx_batch *batch, uint32_t slot, uint32_t *neighbor, float *damage,
                                float *plastic_strain, float *history);
/* Bond forces of one scene, as vx_scene_download_bond_forces. */
int32_t vx_batch_download_bond_forces(vx_batch *batch, uint32_t slot, double *last_force, double *lambda);
/* Hold (hold != 0) or resume a running scene: a held scene sits out the
 * ticks (its state and place in the arena are kept; a tick already in flight
 * completes) until it is resumed. The serving loop holds the larger scenes
 * while the smaller ones finish (shortest job first). */
int32_t vx_batch_hold(vx_batch *batch, uint32_t slot, uint32_t hold);
/* Free a finished (or failed) scene's tiles and slot. */
int32_t vx_batch_release(vx_batch *batch, uint32_t slot);

/* ---- contact projection ---------------------------------------------------
 * Drop-in for collision::project (src/collision.rs:33-123): projects voxel
 * pairs apart and reports the contacts, with the reference's results bit for
 * bit (the Gauss-Seidel order over pairs sorted by (i, j) is reproduced).
 *
 *   positions, previous  n*3 f32, updated in place (scene.positions,
 *                        scene.previous; the swept start is previous on entry)
 *   rest_positions       n*3 f32 (scene.rest_positions)
 *   mass                 n f32 (voxel.mass; unread for kinematic voxels)
 *   kinematic            n bytes, nonzero = kinematic (voxel.kinematic)
 *   target_axes          n bytes, bit a set = targets[i][a].is_some(); NULL
 *                        for no targets (an empty targets slice)
 *   bond_neighbor        3*n, [axis][voxel]: scene.bonds[axis][i].neighbor,
 *                        0xFFFFFFFF when empty
 *   contacts, capacity   receive the contacts sorted by (i, j), one per pair
 *
 * One contact is kept per pair, the first found; the reference's
 * sort_unstable_by_key leaves that choice open when a pair's normal changes
 * between passes. VX_ERR_TOO_LARGE when more than `capacity` contacts are
 * found: nothing is written then except report->contacts, so a second call
 * with a larger capacity gives the same result. VX_ERR_INVALID_ARGUMENT for
 * a swept box with no finite bound on some axis (previous and position both
 * non-finite there; the reference would bin such a voxel at cell 0) or one
 * reaching 2^20 cells or more from the origin, or a bond neighbor >= n. A
 * NaN in only one of previous and position is binned by the other, as the
 * reference's f32::min / f32::max do. */
typedef struct vx_contact {
    uint32_t i, j;        /* i < j */
    float normal[3];      /* unit axis normal pointing from i toward j */
    uint32_t new_impact;  /* the pair started with a gap along the normal */
} vx_contact;

typedef struct vx_contact_report {
    uint64_t candidate_pairs; /* broad-phase pairs */
    uint64_t pairs;           /* after dropping intact local continuum pairs */
    uint64_t contacts;        /* contacts found */
    uint32_t passes;          /* projection passes (at most 6) */
    uint32_t rounds;          /* device windows over the pair lists */
    double device_us;         /* device time, inputs resident, no transfers */
    double projection_us;     /* of which the Gauss-Seidel passes */
    uint64_t evaluated;       /* pair evaluations, in windows and relaxation rounds */
    uint32_t redone;          /* window passes after the first run again with more pairs */
    uint32_t relaxed_passes;  /* passes computed by relaxation (the rest by windows) */
    uint32_t relaxation_rounds; /* relaxation rounds, all passes (those left to windows included) */
    uint32_t reserved;
} vx_contact_report;

/* How the passes are computed, for every later call in the process (the
 * results are the same bit for bit either way):
 *   VX_CONTACT_AUTO     relaxation: every pair evaluated on the positions the
 *                       current outcomes of its earlier pairs give, repeated
 *                       until no outcome changes (a fixed point is the
 *                       sequential sweep's result); a pass that does not get
 *                       there within the budget (max_rounds rounds with dirty
 *                       pairs per call, 1536 when 0; 16 evaluations per pair
 *                       and pass; 16 separating pairs per voxel) is run by
 *                       windows from its start, and so are the passes after
 *                       it. The default.
 *   VX_CONTACT_WINDOWS  every pass by windows: one workgroup walks the sorted
 *                       pairs, committing at once the pairs whose predecessors
 *                       cannot change what they read. */
#define VX_CONTACT_AUTO 0
#define VX_CONTACT_WINDOWS 1
int32_t vx_contact_set_method(int32_t method, uint32_t max_rounds);

int32_t vx_contact_project(uint32_t n, float *positions, float *previous, const float *rest_positions,
                           const float *mass, const uint8_t *kinematic, const uint8_t *target_axes,
                           const uint32_t *bond_neighbor, vx_contact *contacts, uint64_t capacity,
                           vx_contact_report *report);

#ifdef __cplusplus
}
#endif
#endif
