/*
 * femtech_amd.h -- C ABI of the MI355X explicit C3D8 engine.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * nonzero code; FemTech's own termination codes are kept (1 unknown material,
 * 3 input error, 12 allocation failure, 19 time step below the failure time
 * step) and the engine adds FT_ERR_* codes >= 50.  ft_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Each entry point names the reference routine it replaces (paths relative to
 * the FemTech source tree).  Arrays follow the reference layout: nodal arrays
 * are 3 doubles per node (x[3 n + d]), connectivity is 8 zero-based node
 * indices per element in C3D8 order (ShapeFunction_C3D8.cpp:22-29), material
 * properties are 9 doubles per part (ReadMaterials.cpp:43-122 order: rho, mu,
 * lambda, k1, k2, g1, tau1, g2, tau2).
 */
#ifndef FEMTECH_AMD_H
#define FEMTECH_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FT_OK 0
#define FT_ERR_MATERIAL 1   /* StressUpdate.cpp:25-26 */
#define FT_ERR_INPUT 3      /* ReadMaterials.cpp:13 and siblings */
#define FT_ERR_ALLOC 12     /* AllocateArrays.cpp:34 */
#define FT_ERR_DT 19        /* StableTimeStep.cpp:35-38 */
#define FT_ERR_HIP 50       /* a HIP runtime call failed */
#define FT_ERR_ARG 51       /* invalid argument (sizes, null pointers, ids) */
#define FT_ERR_STATE 52     /* call out of order (e.g. run before mass) */
#define FT_ERR_NUMERIC 53   /* non-finite force or inverted element seen */

/* Element arithmetic precision (nodal state is always FP64). */
#define FT_FP64 0
#define FT_FP32 1

/* Motion table entry stride and types (boundary conditions as data; the
 * reference drivers encode them in code, e.g. ex9.cpp:187-248). */
#define FT_MOTION_STRIDE 16
#define FT_MOTION_FIXED 0        /* u = v = a = 0                          */
#define FT_MOTION_LINEAR 1       /* u = p1 t, v = p1, a = 0 (ex9.cpp:191)   */
#define FT_MOTION_RAMP_HOLD 2    /* u = p1 min(t/p2, 1), v = p1/p2 (t<p2)   */
#define FT_MOTION_ROT_HALFSINE 3 /* rigid rotation, axis p1..p3, centre
                                    p4..p6, alpha = p7 sin(pi t/p8), t<=p8 */

/* Status bits reported by ft_get_status. */
#define FT_STATUS_NONFINITE 1
#define FT_STATUS_INVERTED 2
#define FT_STATUS_DT_COLLAPSE 4
/* Kept for compatibility, never raised: in the FP32 engine's run loop the
 * step size always comes from FP64 geometry of the minimising element. */
#define FT_STATUS_DT_FP32 16
/* Informational (FP32 engine): in some step the minimum element time step
 * rose by more than 5 % in one step, so elements that had skipped their
 * exact geometry (their lower bound was 5 % above the last minimum) could
 * set it; that step used the best lower bound (smaller than the exact
 * minimum, so the step stays stable). */
#define FT_STATUS_DT_BOUND 32

typedef struct ft_engine ft_engine;

const char *ft_last_error(void);
int ft_version(void);
/* Number of visible HIP devices (0 when none). */
int ft_device_count(int *count);

/* ---- engine lifecycle ---------------------------------------------------
 * Uploads the mesh once; the device then holds the whole mesh as a single
 * partition (replaces ReadInputFile -> PartitionMesh -> AllocateArrays ->
 * ShapeFunctions: ReadInputFile.cpp:26, PartitionMesh.cpp:20,
 * AllocateArrays.cpp:29, ShapeFunctions.cpp:32).  Builds the node-to-element
 * CSR used for the force gather.  material_id per part follows
 * StressUpdate.cpp:7-27 (0 rigid, 1 neo-Hookean, 2 St Venant-Kirchhoff,
 * 3 linear elastic, 4 HGO isotropic, 5 HGO isotropic + Prony), extended with
 * laws the reference does not have: 6 1-term Ogden (props rho, mu, lambda,
 * alpha), 7 1-term Ogden + 2-term Prony (rho, mu, lambda, alpha, -, g1, tau1,
 * g2, tau2), 8 anisotropic HGO with fibre dispersion (rho, mu, lambda, k1,
 * k2, kappa; fibres from ft_set_fibres).  K = lambda + 2 mu / 3 is the bulk
 * modulus of the volumetric term K/4 (J^2 - 1 - 2 ln J) of 4-8.  At most
 * 89,478,485 nodes and as many elements (FT_ERR_ARG beyond). */
int ft_engine_create(int n_nodes, int n_elems, const double *coords,
                     const int *conn, const int *pid, int n_parts,
                     const int *material_id, const double *props,
                     int precision, ft_engine **out);
int ft_engine_destroy(ft_engine *eng);
int ft_engine_info(ft_engine *eng, int *n_nodes, int *n_elems, int *precision);
/* Force-gather layout (no reference counterpart; the reference scatter-adds,
 * GetForce_3D.cpp:39-44).  FT_LAYOUT_CORNER: every node is local corner k of
 * at most one element for each k, so element forces go to per-node corner
 * slots; FT_LAYOUT_NODE_SLOTS: every node has at most 16 elements, and each
 * element writes its corner forces to per-node slots numbered by its rank
 * among the node's elements (element order; ranks 8-15 of the nodes with
 * more than 8 elements in overflow slots); FT_LAYOUT_ELL: element-major
 * forces summed through the node-to-element table.  FT_LAYOUT_CORNER_PAIRS:
 * corner slots where, in the run loop, two consecutive elements of a
 * wavefront that share a face along xi (element i's local nodes 0, 3, 4, 7
 * are element i - 1's nodes 1, 2, 5, 6: a row of a structured block) add
 * the forces of that face's four corners in registers and element i writes
 * the sums, so about half of the corner slots are written; an element
 * without such a partner writes each xi+ corner to the matching xi- slot of
 * its node, directly when no element owns that slot (a row end of a block)
 * or through a small list added after the element pass, so the node sum
 * reads four slots instead of eight.  It applies wherever FT_LAYOUT_CORNER
 * does and reports FT_LAYOUT_CORNER when no element pairs.  Forces asked for outside the run
 * loop (ft_get_force, ft_element_forces, ft_get_force_fields) are written
 * without pairs, bitwise as with FT_LAYOUT_CORNER.  Creation picks the first that applies in the order
 * CORNER_PAIRS, NODE_SLOTS, ELL (FT_LAYOUT_AUTO); ft_engine_set_layout
 * chooses another that applies, before ft_run_init (FT_ERR_ARG when it does
 * not apply, FT_ERR_STATE inside a run).  ell_width: the table's width (the
 * largest node valence, at least 8). */
#define FT_LAYOUT_AUTO 0
#define FT_LAYOUT_CORNER 1
#define FT_LAYOUT_NODE_SLOTS 2
#define FT_LAYOUT_ELL 3
#define FT_LAYOUT_CORNER_PAIRS 4
int ft_engine_layout(ft_engine *eng, int *layout, int *ell_width);
int ft_engine_set_layout(ft_engine *eng, int layout);

/* ---- set-up --------------------------------------------------------------
 * Boundary conditions as data: dof_motion[3 n + d] = -1 for a free DOF or the
 * index of a motion in `motions` (FT_MOTION_STRIDE doubles each).  Marks the
 * DOFs the way the reference's boundary[] array does (AllocateArrays.cpp:38,
 * ex9.cpp:199-243) and sets the per-element "all nodes constrained" flag of
 * StableTimeStep.cpp:12-20. */
int ft_set_boundary(ft_engine *eng, const int *dof_motion, int n_motions,
                    const double *motions);
/* Fibre directions of the anisotropic HGO law (material 8), 3 per element in
 * the reference configuration, normalised here.  Default: (1, 0, 0). */
int ft_set_fibres(ft_engine *eng, const double *dir);
/* Displacement and velocity (3 per node); either may be NULL (zero). */
int ft_set_state(ft_engine *eng, const double *u, const double *v);
/* Prony history of every Gauss point back to zero (the state of a new run,
 * as after ShapeFunctions.cpp:245-252). */
int ft_reset_history(ft_engine *eng);
/* Lumped mass: row-sum of the consistent element mass, gathered per node
 * (AssembleLumpedMass, Mass3D.cpp:127-157).  mass_out: 3 per node or NULL. */
int ft_assemble_lumped_mass(ft_engine *eng, double *mass_out);

/* ---- per-routine entry points (host arrays authoritative) ---------------
 * StableTimeStep (StableTimeStep.cpp:4-40): min over elements not fully
 * constrained of Le / c on the deformed geometry.  u: 3 per node or NULL to
 * use the device displacement.  boundary: 3 ints per node or NULL to use the
 * flags from ft_set_boundary.  Returns FT_ERR_DT below failure_dt. */
int ft_stable_time_step(ft_engine *eng, const double *u, const int *boundary,
                        double failure_dt, double *dt_out);
/* GetForce_3D (GetForce_3D.cpp:5-53): internal force fi (3 per node) at
 * displacement u (NULL = device displacement).  dt is the step size used by
 * the Prony update of material 5 (the reference reads the global dt). */
int ft_get_force(ft_engine *eng, const double *u, double dt, double *fi_out);
/* GetForce_3D with the per-Gauss-point fields the reference's loop leaves in
 * its global arrays (CalculateDeformationGradient, DeterminateF and
 * StressUpdate inside GetForce_3D.cpp:21-34): F (72 per element: 8 Gauss
 * points x 9, column-major), det F (8 per element) and the PK2 stress (48 per
 * element, Voigt), in the caller's element order and the reference's Gauss
 * point order; any output may be NULL.  Rigid elements (material 0) get F
 * and det F and a zero stress.  The Prony history advances as in
 * ft_get_force. */
int ft_get_force_fields(ft_engine *eng, const double *u, double dt, double *fi_out,
                        double *F_out, double *detF_out, double *pk2_out);
/* Per-element forces (24 per element) at u; history is advanced only when
 * update_history != 0.  Exposes InternalForceUpdate.cpp:4-28 summed over the
 * Gauss points of each element. */
int ft_element_forces(ft_engine *eng, const double *u, double dt,
                      int update_history, double *fe_out);

/* ---- resident explicit loop ---------------------------------------------
 * The driver loop of ex9.cpp:79-175 run on the device: initial BCs, dt,
 * GetForce and accelerations, then central-difference steps while
 * time < tmax and steps < max_steps, with dt = dt_factor * StableTimeStep()
 * recomputed every step and CheckEnergy's balance accumulated in FP64.
 * ft_run_init starts a run from the state given by ft_set_state;
 * ft_run_steps advances it (it may be called repeatedly). */
int ft_run_init(ft_engine *eng, double tmax, double dt_factor,
                double failure_dt);
int ft_run_steps(ft_engine *eng, int max_steps, int *steps_done);
/* Time, current dt (the next step's), step count. */
int ft_get_time(ft_engine *eng, double *time, double *dt, int *step);
/* u, v, a, fi (3 per node each; any may be NULL) at the end of the last
 * completed step (CheckEnergy-time values). */
int ft_get_state(ft_engine *eng, double *u, double *v, double *a, double *fi);
/* Output snapshots without stalling the loop (the displacements the
 * examples write every nSteps / nPlotSteps steps, ex9.cpp:156-167 with
 * WriteVTU.cpp:3-263 and the driver's CustomPlot): ft_snapshot_begin copies
 * the current displacement on the device behind the queued steps and moves
 * it to pinned host memory on a second stream while later ft_run_steps calls
 * compute; ft_snapshot_end waits for that copy and gives u (node-major, 3 per
 * node; nullable) with the time and step of instance 0.  One snapshot is in
 * flight at a time (FT_ERR_STATE otherwise). */
int ft_snapshot_begin(ft_engine *eng);
int ft_snapshot_end(ft_engine *eng, double *u, double *time, int *step);
/* The last completed snapshot's displacement in place (pinned host memory,
 * node-major; valid until the next ft_snapshot_begin), for writers that read
 * it without a copy; NULL before the first snapshot or while one is pending. */
const double *ft_snapshot_view(ft_engine *eng);
/* Energy balance (CheckEnergy.cpp:3-85): W_int, W_ext, KE, |KE+W_int-W_ext|
 * after the last completed step (of instance 0). */
int ft_get_energy(ft_engine *eng, double out4[4]);

/* ---- batched independent simulations (no reference counterpart) ---------
 * The mesh given to ft_engine_create may hold several independent meshes
 * ("instances"): instance i owns nodes [node_offsets[i], node_offsets[i+1])
 * and elements [elem_offsets[i], elem_offsets[i+1]) in the caller's
 * numbering, and its elements use only its nodes.  Each instance then has its
 * own stable step, time and step count, energy balance and motion states
 * (at most 64 distinct motions per instance); one launch per kernel advances
 * all of them (blocks never span two instances).  ft_run_init starts every
 * instance; ft_run_steps(n) advances every unfinished instance by up to n
 * steps and reports the largest count.  ft_set_instance_end_times sets a
 * per-instance end time (otherwise the tmax of ft_run_init).
 * ft_get_instance: time, next dt, step, state (0 running, 1 reached its end
 * time, 2 stopped without a stable step: an inverted element or non-finite
 * forces), W_int, W_ext, KE, |balance| of instance i.  A stopped instance
 * does not stop the others, and ft_run_steps on a batch returns FT_OK for
 * it (a single run returns 19 or 53 as described above).  Per-step history
 * and the injury criteria are single-instance features. */
int ft_set_instances(ft_engine *eng, int n_inst, const int *node_offsets,
                     const int *elem_offsets);
int ft_set_instance_end_times(ft_engine *eng, const double *tmax);
int ft_get_instance(ft_engine *eng, int i, double out8[8]);
/* Per-step history (time, dt, energy[4]) of the steps run after this call;
 * capacity 0 disables.  Costs one small reduction kernel per step. */
int ft_set_history(ft_engine *eng, int capacity);
int ft_get_history(ft_engine *eng, double *time, double *dt, double *energy,
                   int *n);
int ft_get_status(ft_engine *eng, int *status);
/* Prony history of material-5 elements (8 Gauss points x 3 tensors x 6
 * Voigt components per element, element order), as FP64. */
int ft_get_visco_history(ft_engine *eng, double *hist);
/* Sets that history (same layout; a host-side StressUpdate of the
 * compatibility layer hands its updates back this way). */
int ft_set_visco_history(ft_engine *eng, const double *hist);

/* Element-averaged Green-Lagrange strain (CalculateStrain, CalculateStrain.cpp:
 * 77-97; 9 column-major values per element) and the principal-strain summary
 * of CalculateMaximumPrincipalStrain (CalculateStrain.cpp:8-75; per element:
 * max principal clamped >= 0, min principal clamped <= 0, max shear), FP64,
 * at displacement u (NULL = device displacement).  Either output may be NULL.
 * Where round-off at a repeated eigenvalue pushes the reference's acos
 * argument outside [-1, 1] (the reference then yields NaN), the argument is
 * clamped, which gives the limit eigenvalues. */
int ft_element_strain(ft_engine *eng, const double *u, double *eavg,
                      double *principal);

/* CalculateAccelerations (CalculateAcclerations.cpp:4-12) on host arrays
 * (3 per node): a = f_net / mass on DOFs with boundary == 0; the other
 * entries of a are returned unchanged.  Needs the lumped mass. */
int ft_accelerations(ft_engine *eng, const double *f_net, const int *boundary,
                     double *a);
/* The sums of CheckEnergy (CheckEnergy.cpp:15-52, one rank) over host arrays
 * (3 per node each): out3[0] = sum dd (fi_prev + fi), out3[1] = sum dd
 * (fe_prev + fe) plus, on DOFs with boundary != 0, dd (fi_prev + fi +
 * m (a + a_prev)), out3[2] = sum m v^2, with dd = u - u_prev.  The caller
 * applies CheckEnergy's factors of 1/2 and accumulates the works. */
int ft_energy_terms(ft_engine *eng, const double *u, const double *u_prev,
                    const double *v, const double *a, const double *a_prev,
                    const double *fi, const double *fi_prev, const double *fe,
                    const double *fe_prev, const int *boundary, double out3[3]);
/* ft_energy_terms on DEVICE arrays, allocating nothing (no counterpart in the
 * reference: for callers that cannot allocate device memory, e.g. code under
 * HIP graph capture or a tight serving loop).  u ... fe_prev, boundary: device
 * pointers, node-major (3 per node), ndof = 3 x the engine's nodes; out3: host
 * pointer, as ft_energy_terms.  The partial sums live in a static 12.3 KB
 * device array, one copy per GPU held until the process ends; any mesh fits
 * (the kernel strides over the DOFs with a fixed grid), so no size is
 * rejected for scratch.  Calls that overlap on one GPU share that array and
 * race.  Needs the lumped mass. */
int ft_energy_terms_noalloc(ft_engine *eng, long ndof, const double *u,
                            const double *u_prev, const double *v, const double *a,
                            const double *a_prev, const double *fi,
                            const double *fi_prev, const double *fe,
                            const double *fe_prev, const int *boundary,
                            double out3[3]);

/* ---- per-Gauss-point routines ----------------------------------------------
 * FemTech's element layer evaluates one element and one Gauss point per call
 * through free functions over global arrays.  ft_point_eval runs n such calls
 * on the device, one thread each, in FP64: `in` holds n input records and
 * `out` n output records of the sizes below (doubles), back to back.  Node
 * data are in C3D8 local order (ShapeFunction_C3D8.cpp:22-29), 3 per node;
 * 3x3 matrices are column-major (CalculateDeformationGradient.cpp:14).
 *   FT_PT_SHAPE     ShapeFunction_C3D8.cpp:4-128.  in 27: chi[3], node
 *                   coordinates[24]; out 33: N[8], dN/dX[24] (node-major),
 *                   det J.
 *   FT_PT_DEFGRAD   CalculateDeformationGradient.cpp:11-25.  in 48: x = X + u
 *                   per node [24], dN/dX[24]; out 9: F.
 *   FT_PT_DETF      DeterminateF.cpp:38-57.  in 9: F; out 1: det F.
 *   FT_PT_INVF      InverseF.cpp (3-D branch).  in 10: F, det F; out 9.
 *   FT_PT_STRESS    StressUpdate.cpp:5-29 and the material routines.  in 20:
 *                   F[9], material id, props[9] (ReadMaterials order), dt;
 *                   out 34: PK2[6] (Voigt 11, 22, 33, 23, 13, 12), Hn_1,
 *                   Hn_2, S0n [3 x 9] (read and updated in place for the
 *                   Prony laws 5 and 7: the caller fills them before the
 *                   call), flag.  FT_ERR_MATERIAL for an unknown id.
 *   FT_PT_BMAT      StrainDisplacementMatrix.cpp:25-77.  in 12: dN_I/dX[3],
 *                   F[9]; out 18: B_I (6 x 3, column-major).
 *   FT_PT_FINT      InternalForceUpdate.cpp:4-28.  in 64: dN/dX[24], F[9],
 *                   PK2[6], w det J, force[24]; out 192: force + w det J
 *                   B^T S [24], B^T S [24], B [6 x 24].
 *   FT_PT_CHARLEN   CalculateCharacteristicLength_C3D8.cpp:3-30.  in 24:
 *                   x = X + u per node; out 1: volume / largest face area.
 *   FT_PT_TIMESTEP  CalculateTimeStep.cpp:7-21.  in 27: x[24], rho, mu,
 *                   lambda; out 1: Le / ce.
 *   FT_PT_INV3      inverse3x3Matrix (math.cpp:8-28).  in 9; out 10:
 *                   inverse[9], determinant. */
#define FT_PT_SHAPE 0
#define FT_PT_DEFGRAD 1
#define FT_PT_DETF 2
#define FT_PT_INVF 3
#define FT_PT_STRESS 4
#define FT_PT_BMAT 5
#define FT_PT_FINT 6
#define FT_PT_CHARLEN 7
#define FT_PT_TIMESTEP 8
#define FT_PT_INV3 9
int ft_point_eval(int op, int n, const double *in, double *out);

/* ---- injury criteria ------------------------------------------------------
 * CalculateInjuryCriterions of FemTech's ex5 (ex5.cpp:1311-1430) with its
 * outputs (ex5.cpp:1040-1240), on the device.  ft_injury_enable selects the
 * included elements (parts with include_part[p] != 0; NULL = all), computes
 * their reference-configuration volumes and resets the criteria; from then on
 * every step of ft_run_steps updates them at the step's displacement, time and
 * step size.  ft_injury_update applies one update at u (NULL = device
 * displacement) with the given time and dt (host-array drivers).
 * ft_injury_results: out[0..15] = max principal strain, its element, time;
 * min principal strain, element, time; max shear, element, time; max PS x SR,
 * element, time; MPS-95 maximum and its time; MPSxSR-95 maximum and its time;
 * out[16..19] = CSDM-15, CSDM-30, MPSR-120, MPSxSR-28 (fractions of the
 * included volume); out[20] = included volume.  flags (per element, may be
 * NULL): bit 0 included, 1 MPS > 0.15 seen, 2 MPS > 0.30 seen, 3 PSR > 120
 * seen, 4 PS x SR > 28 seen, 5 in the MPS-95 list, 6 in the MPSxSR-95 list.
 * The previous principal strain starts at 0 (the reference leaves it unset). */
int ft_injury_enable(ft_engine *eng, const int *include_part);
int ft_injury_disable(ft_engine *eng);
int ft_injury_update(ft_engine *eng, const double *u, double time, double dt);
int ft_injury_results(ft_engine *eng, double out[21], int *flags);

/* ---- launch shapes and captured steps ------------------------------------
 * No counterpart in the reference (its loop is host code, ex9.cpp:107-175).
 * ft_set_graph: ft_run_steps captures `steps` consecutive steps of the loop
 * as one HIP graph (one per starting parity, built on first use) and
 * launches it for every whole chunk of its step count; the rest run as
 * separate launches.  0 (the default) launches every step separately.  Any
 * call that changes the engine's configuration drops the captured graphs;
 * timing (ft_set_timing) uses separate launches.
 * ft_autotune: per material segment, times the run loop's element kernel at
 * every tile size (64, 128, 256 elements per block) and, in FP32, every
 * occupancy variant (1..5 waves per SIMD), `repeats` launches each, plus the
 * default shape (256 elements per block, compiler-chosen occupancy), and
 * keeps the fastest; occupancy variants that spill more than 128 bytes of
 * registers per lane to scratch memory are skipped.  report (may be NULL): 6 doubles per segment = material,
 * elements, waves, tile, best us per launch, default-shape us per launch.
 * The run state is unchanged (the Prony history is not advanced).
 * ft_get_segments: desc (may be NULL) gets 5 ints per segment = material,
 * first and end element (internal, grouped order), waves (0 = default), tile.
 * ft_set_segment_shape sets one segment's shape by hand. */
int ft_set_graph(ft_engine *eng, int steps);
/* ft_set_persistent (no counterpart in the reference): with enable != 0,
 * ft_run_steps runs all of its steps in one cooperative launch, the element
 * phase, the node phase and the history row of every step separated by
 * grid-wide barriers, and ends early once no instance goes on.  Results
 * agree with the per-step launches to floating-point round-off (same bodies,
 * tiles and reduction order; the compiler may fuse multiply-adds differently
 * in the one kernel).  On MI355X it is slower than the per-step launches for
 * every mesh size measured.  The element phase runs the neo-Hookean
 * body when the whole mesh is material 1 and the run-time dispatched body of
 * all laws otherwise; autotuned segment shapes do not apply.  Runs with
 * kernel timing, the surrogate or the injury criteria keep the per-step
 * launches.  A barrier wait is bounded (about 0.3 s); if one runs out the
 * call fails with FT_ERR_HIP and the run is void. */
int ft_set_persistent(ft_engine *eng, int enable);
int ft_autotune(ft_engine *eng, int repeats, double *report, int *n_segments);
int ft_get_segments(ft_engine *eng, int *n_segments, int *desc);
int ft_set_segment_shape(ft_engine *eng, int segment, int waves, int tile);

/* ---- constitutive surrogate network --------------------------------------
 * No counterpart in the reference.  A multilayer perceptron of 16 inputs (12
 * used) -> 4 hidden layers of 128 SiLU units -> 8 outputs (6 used), with BF16
 * weights and activations on the matrix cores and FP32 accumulation; FP32
 * master weights and Adam moments.  Parameters: one FP32 vector of
 * ft_mlp_param_count() values, per layer W [out][in] row-major then b [out].
 * Host arrays: x [n][16] features, t and y [n][8].
 * ft_mlp_forward: y = network(x) (one fused launch, all layers on chip).
 * ft_mlp_gradients: mean squared error over the 6 used outputs on the batch
 * and its gradient (same layout as the parameters).
 * ft_mlp_train: `steps` Adam steps (cosine-decayed rate from lr) on batches
 * drawn on the device from (x, t); final_loss: mean loss of the last 1 %. */
typedef struct ft_mlp ft_mlp;
int ft_mlp_param_count(void);
int ft_mlp_create(unsigned long long seed, ft_mlp **out);
int ft_mlp_destroy(ft_mlp *mlp);
int ft_mlp_get_params(ft_mlp *mlp, float *params);
int ft_mlp_set_params(ft_mlp *mlp, const float *params);
int ft_mlp_forward(ft_mlp *mlp, const float *x, int n, float *y);
int ft_mlp_gradients(ft_mlp *mlp, const float *x, const float *t, int n, double *loss, float *grad);
int ft_mlp_train(ft_mlp *mlp, const float *x, const float *t, int n_data, int steps, int batch,
                 double lr, unsigned long long seed, double *final_loss);

/* ---- surrogate constitutive path (materials 7 and 8) ----------------------
 * No counterpart in the reference.  The network of ft_mlp_* replaces the
 * isochoric part of the PK2 stress of materials 7 (Ogden + Prony) and 8
 * (anisotropic HGO) in the FP32 run loop; the volumetric part and the Prony
 * recurrence stay analytic.  Inputs (12): Cbar - I in the material frame
 * (fibre frame for 8; Voigt 11, 22, 33, 23, 13, 12), J - 1, Ibar1 - 3,
 * Ibar2 - 3, Ibar4 - 1, Ibar6 - 1, law; outputs (6): S_iso in the material
 * frame.  A Gauss point whose inputs leave the training box (per-input
 * range of the training samples) uses the analytic law.
 * ft_surrogate_reserve / ft_surrogate_collect: room for `capacity` training
 * samples; each collect appends the inputs and analytic S_iso (FP64) of
 * every Gauss point of the material 7 / 8 elements at the current device
 * displacement (*total: samples held).
 * ft_surrogate_train: normalises the samples and trains the network on the
 * device (ft_mlp_train); out = final loss, samples, relative L2 error of
 * S_iso on 2^16 samples spread over the set, steps.
 * ft_surrogate_get / ft_surrogate_set: the trained network's parameters
 * (ft_mlp_param_count floats) and normalisation block (80 floats).
 * ft_surrogate_enable: on/off for the run loop (FP32 engine, one instance);
 * every `audit_every`-th step, a fixed share audit_permille / 1000 of the
 * elements also evaluates the analytic law and the differences are tallied.
 * ft_surrogate_stats: out = audited relative L2 error of S_iso, the same
 * relative to the full PK2 stress, fallback share of the Gauss-point
 * evaluations, audited evaluations, surrogate evaluations, fallbacks. */
int ft_surrogate_reserve(ft_engine *eng, long capacity);
int ft_surrogate_collect(ft_engine *eng, long *total);
/* Collected samples [first, first + count): x 16 and t 8 floats each (raw
 * inputs, S_iso in the material frame); either may be NULL.  Sample order
 * per collect: segment by segment, Gauss-point major within a segment. */
int ft_surrogate_samples(ft_engine *eng, long first, long count, float *x, float *t);
int ft_surrogate_train(ft_engine *eng, int steps, int batch, double lr, unsigned long long seed,
                       double out[4]);
int ft_surrogate_get(ft_engine *eng, float *params, float *norm);
int ft_surrogate_set(ft_engine *eng, const float *params, const float *norm);
int ft_surrogate_enable(ft_engine *eng, int enable, int audit_every, int audit_permille);
int ft_surrogate_stats(ft_engine *eng, double out[6]);

/* ---- instrumentation -----------------------------------------------------
 * When enabled (enable = N > 0), the kernel launches of every N-th step of
 * the run loop, and every launch outside it, are bracketed by HIP events on
 * their stream.  Times are summed in milliseconds per kernel class
 * (0 element, 1 nodal/gather, 2 other) together with launch counts. */
int ft_set_timing(ft_engine *eng, int enable);
int ft_get_timing(ft_engine *eng, double ms[3], int launches[3]);
/* Algorithmic bytes moved per launch of the element kernel and of the nodal
 * kernel, as defined in the design notes (compulsory traffic, each array
 * touched once). */
int ft_traffic_model(ft_engine *eng, double bytes[2]);
/* Plain device-to-device streaming copy of `megabytes` (read + write), run
 * `repeats` times; reports GB/s moved.  Used to calibrate the FETCH_SIZE /
 * WRITE_SIZE counters against a kernel of known traffic. */
int ft_copy_probe(double megabytes, int repeats, double *gbs);

/* ---- host-side mesh tools (no GPU needed) -------------------------------
 * Structured hex block nx x ny x nz with spacings h*, nodes numbered
 * x-fastest; interior nodes jittered uniformly by +-jitter*h per coordinate
 * from a 64-bit Mersenne twister seeded with `seed` (draw order node-major,
 * x, y, z).  coords: 3*(nx+1)(ny+1)(nz+1), conn: 8*nx*ny*nz. */
int ft_mesh_block(int nx, int ny, int nz, double hx, double hy, double hz,
                  double jitter, unsigned long long seed, double *coords,
                  int *conn);
/* Ellipsoid with semi-axes a, b, c from an n^3 cube mapped onto the ball
 * (nodes and elements numbered as ft_mesh_block with nx = ny = nz = n; the
 * cube's faces land on the surface).  radius: per element, the normalised
 * radius (0 centre, 1 surface) of the centroid, for region assignment; may
 * be NULL.  coords: 3 (n+1)^3, conn: 8 n^3. */
int ft_mesh_ellipsoid(int n, double a, double b, double c, double *coords,
                      int *conn, double *radius);
/* Abaqus .inp writer readable by ReadAbaqus.cpp:26-277 (one ELSET per part,
 * node ids 1..n in array order, %.17g coordinates). */
int ft_write_inp(const char *path, int n_nodes, int n_elems, const double *coords,
                 const int *conn, const int *pid);
/* Mesh reader for .inp (ReadAbaqus.cpp) and .k (ReadLsDyna.cpp), C3D8 only,
 * with the one-rank node renumbering of PartitionMesh.cpp:485-536 (local
 * node i = i-th smallest global node id).  Two calls: first with NULL arrays
 * to get the sizes, then with arrays of those sizes. */
int ft_read_mesh(const char *path, int *n_nodes, int *n_elems, double *coords,
                 int *conn, int *pid, int *global_node_id);
/* materials.dat reader (ReadMaterials.cpp:8-138): fills material_id[n_parts]
 * and props[9 n_parts]. */
int ft_read_materials(const char *path, int n_parts, int *material_id,
                      double *props);

#ifdef __cplusplus
}
#endif
#endif
