/*
 * CPU oracle for the corotational elastic PCG solve of Voxel-sim.
 *
 * TEST INFRASTRUCTURE ONLY. This library is what the HIP path is
 * compared against (and the timed CPU baseline). It is never linked into,
 * loaded by, or called from the product library, and the product path never
 * falls back to it.
 *
 * It restates, in C++ templated on the scalar type, the reference algorithm
 * (paths relative to the reference repository):
 *   src/solver.rs:27-73    basis / tensor_compliance / stiffness_matrix
 *   src/solver.rs:78-95    bond_thermal_rest
 *   src/solver.rs:327-417  fit_rotations
 *   src/solver.rs:1325-1345 make_constraint
 *   src/solver.rs:1347-1537 Cholesky helpers, block IC(0), block-Jacobi
 *   src/solver.rs:1542-1929 solve_coupled: assembly, projection, row packing,
 *                           CPU PCG loop
 *   src/solver.rs:475-570, 668-683 the substep / corotation driver (solve part)
 *   src/cuda_solver.rs:301-501 + src/cuda/pcg.ptx  device PCG semantics
 *
 * Mode VXO_F32 follows the reference's f32 arithmetic order (glam f32 values,
 * f64 dot accumulation, f32 alpha/beta). Mode VXO_F64 is the FP64 oracle:
 * the same formulas evaluated in double.
 *
 * Parity unpinned at bit level: the f32 steps the reference takes through
 * glam 0.27 (Quat::slerp, Quat::from_mat3, Mat3::from_quat, Mat3::inverse;
 * src/solver.rs:101, 365, 383, 398, 1326, 1765) and rayon's f64 summation
 * order cannot be run or read here. They are pinned to tolerance only, by
 * the closed forms, reference tests and invariants checked in tests/.
 *
 * Every array argument is double (or an integer type) regardless of mode; in
 * VXO_F32 mode inputs are rounded to float on entry and results are exact
 * float values widened to double.
 *
 * Matrices are exported row-major: m[r*3 + c].
 */
#ifndef VX_ORACLE_H
#define VX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VXO_F32 0
#define VXO_F64 1

#define VXO_PRECOND_IC0 0
#define VXO_PRECOND_BJ 1

#define VXO_EMPTY 0xFFFFFFFFu

/* Status codes mirror include/vx_solver.h. */
#define VXO_OK 0
#define VXO_ERR_NONFINITE_RESIDUAL 3
#define VXO_ERR_NOT_CONVERGED 4
#define VXO_ERR_NONFINITE_CORRECTION 5

/* Scene state consumed by the solve (reference Scene, src/scene.rs:309-351). */
typedef struct vxo_scene {
    uint32_t n;
    double *positions;            /* n*3 */
    double *rotations;            /* n*4 quaternion (x, y, z, w) */
    double *velocities;           /* n*3, used by vxo_step only */
    const double *mass;           /* n */
    const uint8_t *kinematic;     /* n */
    const double *young;          /* n, modulus along the fibre */
    const double *young_perp;     /* n */
    const double *shear;          /* n */
    const double *solid;          /* n, phase solid fraction */
    const double *fibers;         /* n*3, Microstructure::fibers() (body frame) */
    const double *anisotropy;     /* n */
    const double *delta_temp;     /* n, temperature minus stress-free temperature */
    const double *alpha_par;      /* n, thermal expansion along the fibre */
    const double *alpha_perp;     /* n */
    const uint32_t *neighbor;     /* 3*n, [axis][voxel], VXO_EMPTY if no bond */
    const double *damage;         /* 3*n */
    const double *kscale;         /* 3*n, per-bond stiffness scale (1 = none) */
    const double *formation_offset; /* 3*n*3 */
    const double *plastic_strain;   /* 3*n*3 */
} vxo_scene;

typedef struct vxo_system vxo_system;

/* ---- topology ------------------------------------------------------------ */
/* Lattice bond rule (src/scene.rs:558-570, 1668-1681): for each voxel i and
 * axis a, bond i -> j iff the cell key+e_a is occupied by j and both have the
 * same kinematic flag. keys: n*3 int32 lattice coordinates. Writes
 * neighbor[3*n]. Returns the number of bonds. */
uint32_t vxo_build_topology(uint32_t n, const int32_t *keys, const uint8_t *kinematic,
                            uint32_t *neighbor);

/* ---- constitutive pieces (exposed for known-answer tests) ---------------- */
/* frame = basis(axis, slerp(q_i, q_j, 0.5)); outputs row-major 3x3. */
void vxo_bond_frame(int mode, const double *qi, const double *qj, uint32_t axis, double *frame);
/* stiffness_matrix(scene, i, j, frame) (src/solver.rs:65-73). */
void vxo_stiffness_matrix(int mode, const vxo_scene *s, uint32_t i, uint32_t j,
                          const double *frame, double *k);
/* bond_thermal_rest (src/solver.rs:78-95). */
void vxo_bond_thermal_rest(int mode, const vxo_scene *s, uint32_t i, uint32_t j,
                           const double *frame, double *rest);

/* ---- rotation fit -------------------------------------------------------- */
/* fit_rotations (src/solver.rs:327-417); updates s->rotations and returns the
 * largest quaternion angle change. */
double vxo_fit_rotations(int mode, vxo_scene *s);

/* ---- assembly ------------------------------------------------------------ */
/* solve_coupled assembly (src/solver.rs:1552-1672).
 * predicted: n*3; free_mask: n bytes, bit a set when axis a is free;
 * loads: n*3 applied loads (zero for kinematic voxels). */
vxo_system *vxo_assemble(int mode, const vxo_scene *s, double h, const double *predicted,
                         const uint8_t *free_mask, const double *loads);
void vxo_system_free(vxo_system *sys);
uint32_t vxo_system_active_count(const vxo_system *sys);
uint32_t vxo_system_constraint_count(const vxo_system *sys);
/* Reference device packing (src/solver.rs:1731-1777): per active row the
 * neighbours [u32;6] in incidence order, 6 row-major off-diagonal blocks
 * (54 values, zero-filled), the row-major diagonal and its inverse, and the
 * right-hand side r. active: voxel index of each active row. */
void vxo_system_export_rows(const vxo_system *sys, uint32_t *active, uint32_t *neighbors,
                            double *blocks, double *diagonal, double *inverse, double *rhs);
/* Per voxel: projected diagonal (n*9), masked rhs (n*3), incidence (n*6 edge
 * ids, VXO_EMPTY padded) and incidence lengths (n). */
void vxo_system_export_full(const vxo_system *sys, double *diagonal, double *rhs,
                            uint32_t *incidence, uint8_t *incidence_len);
/* Per constraint: (i, j, axis), frame, material_k, k (row-major), rest. */
void vxo_system_export_constraints(const vxo_system *sys, uint32_t *ija, double *frame,
                                   double *material_k, double *k, double *rest);

/* ---- solves -------------------------------------------------------------- */
/* Reference CPU PCG (src/solver.rs:1833-1911) in the system's mode.
 * threads: OpenMP threads (0 = library default). parallel_cutoff: rows below
 * which loops stay serial (the reference uses 32,768; 0 = always parallel).
 * dx: active*3. seconds: wall time of the PCG loop alone. */
void vxo_solve_cpu(vxo_system *sys, int precond, uint32_t max_iterations, double tolerance,
                   int threads, uint32_t parallel_cutoff, double *dx, uint32_t *iterations,
                   double *residual, double *seconds);

/* Device PCG semantics (src/cuda_solver.rs:301-501, src/cuda/pcg.ptx) over the
 * reference packing. check_every = 16 reproduces the reference. Returns a
 * VXO_* status; dx/iterations/relative are written on VXO_OK. */
int vxo_solve_device_semantics(int mode, uint32_t n, const double *rhs,
                               const uint32_t *neighbors, const double *diagonal,
                               const double *blocks, const double *inverse,
                               uint32_t max_iterations, double tolerance, uint32_t check_every,
                               double *dx, uint32_t *iterations, double *relative);

/* FP64 reference solution: block-Jacobi PCG in double over the reference
 * packing, run until the recomputed relative residual ||b - A x|| / ||b|| is
 * <= tolerance (or max_iterations). Returns the final true relative residual. */
double vxo_solve_exact(uint32_t n, const double *rhs, const uint32_t *neighbors,
                       const double *diagonal, const double *blocks, double tolerance,
                       uint32_t max_iterations, double *x, uint32_t *iterations);

/* FP64 CG on the red-black reduced system S x_b = b_b - A_br D_rr^-1 b_r
 * (black[i] = 1 marks the black rows, odd x + y + z) with block-Jacobi on
 * D_bb, then x_r = D_rr^-1 (b_r - A_rb x_b). Stops before an iteration when
 * ||r|| <= tolerance ||b|| or ||r|| <= inner_tolerance ||r0|| (or after
 * max_iterations). Returns the full system's true relative residual; history
 * (nullable, max_iterations + 1 entries) receives ||r|| before each
 * iteration. */
double vxo_solve_red_black(uint32_t n, const double *rhs, const uint32_t *neighbors, const double *diagonal,
                           const double *blocks, const uint8_t *black, double tolerance, double inner_tolerance,
                           uint32_t max_iterations, double *x, uint32_t *iterations, double *history);

/* y = A x over the reference packing, in the given mode (PTX fma order in
 * VXO_F32). */
void vxo_matvec(int mode, uint32_t n, const uint32_t *neighbors, const double *diagonal,
                const double *blocks, const double *x, double *y);

/* dot_product_single semantics (src/cuda/pcg.ptx:563-637): rows of three f32
 * values (inputs rounded to f32), f64 products, f64 sum. */
double vxo_dot_rows(uint32_t n, const double *a, const double *b);
/* scalar_ratio / dot_reduce_ratio_single (src/cuda/pcg.ptx:437-466, 639-724):
 * num/den cast to the vector type when den > 1e-30, else 0. */
double vxo_guarded_ratio(int mode, double num, double den);

/* ---- fracture (src/fracture.rs:73-133, 210-292; src/solver.rs:106-161, 627-665) */
#define VXO_BRITTLE 0
#define VXO_DUCTILE 1
#define VXO_MIXED 2
/* Per-REV strength (Material, src/material.rs:20-63): VXO_STRENGTH doubles
 * per REV in this order. */
enum {
    VXO_S_TENSION, VXO_S_TENSION_PERP, VXO_S_COMPRESSION, VXO_S_COMPRESSION_PERP, VXO_S_SHEAR,
    VXO_S_SHEAR_PERP, VXO_S_YIELD, VXO_S_HARDENING, VXO_S_PLASTIC_FAILURE, VXO_S_MODE, VXO_S_GC,
    VXO_S_GC_PERP, VXO_STRENGTH
};
/* Limits (src/fracture.rs:63-72), 8 doubles in this order. */
enum { VXO_L_TENSION, VXO_L_COMPRESSION, VXO_L_SHEAR, VXO_L_YIELD, VXO_L_HARDENING, VXO_L_PLASTIC_FAILURE,
       VXO_L_DUCTILITY, VXO_L_GC, VXO_LIMITS };

typedef struct vxo_fracture_config { /* FractureConfig (src/fracture.rs:11-25) */
    double damage_rate, energy_release, fatigue_cycles, fatigue_exponent, fatigue_threshold, plastic_heat_fraction;
} vxo_fracture_config;

typedef struct vxo_bond { /* Bond history (src/bond.rs:4-16) */
    uint32_t neighbor;
    double plastic_strain[3];
    double damage, equivalent_plastic_strain, plastic_work, fatigue_damage;
    double previous_stress_ratio, reversal_stress_ratio, loading_increment;
} vxo_bond;

typedef struct vxo_dissipation { /* Dissipation (src/fracture.rs:135-141), J */
    double plastic_heat, stored_plastic, surface_work, damage_heat, released;
} vxo_dissipation;

/* pair_limits (src/fracture.rs:73-133): a, b are VXO_STRENGTH-value rows. */
void vxo_pair_limits(int mode, const double *a, const double *b, const double alignment[2],
                     const double anisotropy[2], double limits[VXO_LIMITS]);
/* fatigue (src/fracture.rs:190-207). */
void vxo_fatigue(int mode, vxo_bond *bond, double signed_ratio, const vxo_fracture_config *c);
/* update (src/fracture.rs:210-292): u_total in the bond frame (m), k the
 * undamaged local stiffness (row-major 3x3, N/m). */
void vxo_fracture_update(int mode, vxo_bond *bond, const double u_total[3], const double k[9],
                         const double limits[VXO_LIMITS], double dt, const vxo_fracture_config *c,
                         vxo_dissipation *out);

/* Fracture state for vxo_step: the same neighbor / damage / plastic_strain
 * arrays as the scene (written here), the remaining bond history (3*n each)
 * and the strength table (VXO_STRENGTH*n). thermal strength scale is 1 (no
 * temperature model): strength is scaled by min(solid_i, solid_j)^2. */
typedef struct vxo_fracture_state {
    uint32_t *neighbor;
    double *damage, *plastic_strain;
    double *equivalent_plastic_strain, *plastic_work, *fatigue_damage;
    double *previous_stress_ratio, *reversal_stress_ratio, *loading_increment;
    const double *strength;
    vxo_fracture_config config;
    vxo_dissipation dissipated;     /* accumulated */
    uint32_t broken;                /* accumulated */
} vxo_fracture_state;

/* ---- step driver --------------------------------------------------------- */
typedef struct vxo_step_config {
    double fixed_dt;
    uint32_t substeps;
    uint32_t iterations;            /* CPU PCG cap */
    uint32_t corotation_iterations;
    double solver_relative_tolerance;
    double corotation_tolerance;
    double gravity[3];
    int precond;                    /* VXO_PRECOND_* */
    int exact;                      /* 1: solve each pass with vxo_solve_exact at exact_tolerance */
    double exact_tolerance;
} vxo_step_config;

typedef struct vxo_step_report {
    uint32_t pcg_iterations;        /* summed over passes and substeps */
    uint32_t passes;                /* solves performed */
    double solver_residual;         /* last solve */
    double last_angle;              /* last corotation fit angle */
} vxo_step_report;

/* One fixed step without contact, tool, fracture or heat: the solve part of
 * step_with_axis_targets_and_loads (src/solver.rs:446-570, 668-683).
 * targets: n*3 world coordinates, NaN for a free axis (may be NULL).
 * loads: n*3 applied loads (forces + accel*mass) (may be NULL).
 * last_force, lambda: optional 3*n*3 outputs ([axis][voxel][component]) of
 * the last substep, per constraint of its last assembly at the solved
 * positions and before any fracture update: the bond force
 * (src/solver.rs:627-640) and the multiplier (src/solver.rs:1824-1827). */
void vxo_step(int mode, vxo_scene *s, const vxo_step_config *cfg, const double *targets,
              const double *loads, vxo_step_report *report, double *last_force, double *lambda);
/* vxo_step plus, per substep after the corotation passes and before the
 * velocities, the fracture pass over the last assembly's constraints
 * (src/solver.rs:627-665). Bonds that break leave the topology. */
void vxo_step_fracture(int mode, vxo_scene *s, const vxo_step_config *cfg, const double *targets,
                       const double *loads, vxo_fracture_state *fracture, vxo_step_report *report,
                       double *last_force, double *lambda);

/* ---- rigid tool (src/tool.rs:192-431) -------------------------------------
 * The tool's state: cells (n_cells*3, body frame about the centre of mass),
 * mass, principal inverse inertia (body frame), pose (rotation x, y, z, w),
 * velocities, persistent external force and torque, the Coulomb friction
 * coefficient and an optional guide mount (src/tool.rs:17-24). advance updates
 * the pose and velocities in place, adds the friction heat and records the
 * microsteps and contacts of the call. */
typedef struct vxo_tool {
    uint32_t n_cells;
    uint32_t has_mount;
    const double *cells;
    double mass;
    double inverse_inertia[3];
    double position[3];
    double rotation[4];
    double velocity[3];
    double angular_velocity[3];
    double external_force[3];
    double external_torque[3];
    double friction;
    double mount_local_base[3], mount_world_base[3], mount_direction[3], mount_rotation[4];
    double friction_heat;
    uint32_t microsteps, contacts;
} vxo_tool;
/* Tool::advance (src/tool.rs:192-431) over a scene of n voxels: positions
 * (n*3) are read; velocities (n*3) receive the contact impulses; targets
 * (n*3, NaN = free axis, may be NULL) and kinematic flags block axes;
 * contact_velocity (n*3) is the returned per-voxel velocity change. */
void vxo_tool_advance(int mode, vxo_tool *tool, uint32_t n, const double *positions, double *velocities,
                      const double *mass, const uint8_t *kinematic, const double *targets, double dt,
                      const double *gravity, double *contact_velocity);
/* vxo_step_fracture with a tool (may be NULL; fracture may be NULL): per
 * substep after the prediction, the tool advances (src/solver.rs:530-534),
 * positions drift by half a substep of the contact velocity
 * (src/solver.rs:535-543), and the final velocities of free axes include it
 * (src/solver.rs:668-683); then a zero-time advance on the final state
 * (src/solver.rs:701-706), whose contact velocity is discarded. The tool
 * keeps the timed advance's microsteps and the sum of both advances'
 * contacts. */
void vxo_step_tool(int mode, vxo_scene *s, const vxo_step_config *cfg, const double *targets, const double *loads,
                   vxo_tool *tool, vxo_fracture_state *fracture, vxo_step_report *report);

/* ---- contact projection (src/collision.rs:33-123), f32 ------------------
 * vxo_contact_project restates collision::project: positions and previous
 * (n*3 f32) are updated in place; rest (n*3), mass (n), kinematic (n),
 * targets (n, bit a set = axis a held by a target; NULL for none) and
 * bond_neighbor (3*n, [axis][voxel], VXO_EMPTY when absent) are inputs.
 * Contacts are written sorted by (i, j), up to `capacity`; the return value
 * is the full count. stats (may be NULL): candidate pairs, pairs after the
 * local-continuum filter, projection passes, contacts. */
uint64_t vxo_contact_project(uint32_t n, float *positions, float *previous, const float *rest,
                             const float *mass, const uint8_t *kinematic, const uint8_t *targets,
                             const uint32_t *bond_neighbor, uint64_t capacity, uint32_t *out_i,
                             uint32_t *out_j, float *out_normal, uint8_t *out_new_impact, uint64_t *stats);
/* The broad phase alone: sorted candidate pairs (2 per entry) up to
 * `capacity`; returns the count. `previous` is the swept start. */
uint64_t vxo_contact_candidates(uint32_t n, const float *positions, const float *previous,
                                const uint8_t *kinematic, uint32_t *out_pairs, uint64_t capacity);
/* local_continuum_pair (src/collision.rs:520-550): 1 when i and j are joined
 * by a monotone path of intact bonds inside their lattice cell. */
int vxo_local_continuum_pair(uint32_t n, const float *rest, const uint32_t *bond_neighbor, uint32_t i,
                             uint32_t j);

#ifdef __cplusplus
}
#endif
#endif
