/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of FemTech's explicit central-difference step for
 * C3D8 hexahedra.  It is the checker for the HIP engine: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links, loads or calls anything in this directory.
 *
 * Every routine cites the reference file:line whose semantics it restates
 * (paths relative to the FemTech source tree).  Parity status of each routine
 * is listed in the header comment of oracle.cpp.
 */
#ifndef FT_ORACLE_H
#define FT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Motion table entry: 16 doubles [type, p1 .. p15].
 *   0 FIXED        u = v = a = 0
 *   1 LINEAR       u = p1*t, v = p1, a = 0                  (ex9.cpp:191-243)
 *   2 RAMP_HOLD    u = p1*min(t/p2,1), v = p1/p2 for t<p2 else 0, a = 0
 *   3 ROT_HALFSINE rigid rotation about unit axis p1..p3 through point
 *                  p4..p6, angular acceleration p7*sin(pi t/p8) for t<=p8,
 *                  zero afterwards; u, v, a from the closed-form kinematics.
 */
#define ORACLE_MOTION_STRIDE 16

/* --- element-level routines ------------------------------------------- */
void oracle_gauss_c3d8(double chi[24], double w[8]);
/* xe: 8 nodes x 3 reference coordinates.  dNdX[gp*24 + a*3 + j], detJ[gp]. */
void oracle_shape_c3d8(const double xe[24], double dNdX[192], double detJ[8]);
/* F column-major (F[i + 3j] = F_ij).  hist: 27 doubles (Hn1, Hn2, S0n, each
 * 3x3 column-major) for material 5, else ignored.  Returns 0 or the
 * reference's error code 1 for an unknown material. */
int oracle_stress(int material, const double props[9], const double F[9],
                  double J, double dt, double *hist, double S6[6]);
/* Also materials 6 (1-term Ogden), 7 (Ogden + Prony) and 8 (anisotropic HGO
 * with fibre direction `fibre`, NULL = (1, 0, 0)); see oracle.cpp. */
int oracle_stress_fibre(int material, const double props[9], const double fibre[3],
                        const double F[9], double J, double dt, double *hist,
                        double S6[6]);
double oracle_det3(const double F[9]);
double oracle_hex_volume(const double xe[24]);
double oracle_face_area(const double xe[24], const int idx[4]);
double oracle_char_length(const double xe[24]);
/* Lumped (row-sum) mass of one element, per node (same for the 3 DOFs). */
void oracle_element_mass(const double xe[24], double rho, double m8[8]);

/* --- mesh-level model -------------------------------------------------- */
void *oracle_create(int n_nodes, int n_elems, const double *X, const int *conn,
                    const int *pid, int n_parts, const int *material_id,
                    const double *props /* 9 per part */);
void oracle_destroy(void *m);
void oracle_set_threads(void *m, int threads);
/* dNdX (192 per element) and detJ (8 per element); either may be NULL. */
void oracle_shape_functions(void *m, double *dNdX, double *detJ);
void oracle_lumped_mass(void *m, double *mass /* 3 per node */);
/* One GetForce_3D over the mesh at displacement u (3 per node).  Advances the
 * Prony history of material-5 elements with step size dt.  fi: 3 per node. */
int oracle_get_force(void *m, const double *u, double dt, double *fi);
/* Per-element internal forces (24 per element), no history update unless
 * update_history != 0. */
int oracle_element_forces(void *m, const double *u, double dt,
                          int update_history, double *fe);
/* StableTimeStep at displacement u; boundary: 3 ints per node. */
double oracle_stable_dt(void *m, const double *u, const int *boundary);
/* History (27 doubles per Gauss point, 8 per element) read/write. */
void oracle_get_history(void *m, double *hist);
void oracle_set_history(void *m, const double *hist);

/* Boundary conditions: dof_motion[3*n+d] = -1 (free) or motion index. */
void oracle_set_bc(void *m, const int *dof_motion, int n_motions,
                   const double *motions);
void oracle_set_initial(void *m, const double *u0, const double *v0);
/* Fibre directions (3 per element, normalised) for material 8. */
void oracle_set_fibres(void *m, const double *dir);
/* ex9-style explicit loop (ex9.cpp:79-175).  Runs while time < tmax and
 * steps < max_steps.  Outputs may be NULL.  hist_* have room for
 * max_steps+1 entries (entry 0 = state after the initial GetForce).
 * hist_energy: 4 per step (W_int, W_ext, KE, |KE+W_int-W_ext|).
 * Returns steps taken (>= 0) or -(error code). */
int oracle_run(void *m, int max_steps, double tmax, double dt_factor,
               double failure_dt, double *u_out, double *v_out, double *a_out,
               double *fi_out, double *hist_time, double *hist_dt,
               double *hist_energy, int n_snap, const int *snap_steps,
               double *snap_u);
/* Element-averaged Green-Lagrange strain (9 column-major per element) and
 * principal summary (max >= 0, min <= 0, max shear; 3 per element) at u. */
void oracle_element_strain(void *m, const double *u, double *eavg, double *principal);
/* Injury criteria of ex5 (ex5.cpp:1253-1430, 1040-1240).  After
 * oracle_injury_init, oracle_run updates them after every step.
 * include_part: per part 0/1, NULL = all parts.  out: 21 doubles, flags: per
 * element bits (see oracle.cpp). */
void oracle_injury_init(void *m, const int *include_part);
void oracle_injury_update(void *m, const double *u, double time, double dt);
void oracle_injury_results(void *m, double *out, int *flags);
/* Motion evaluation (used by tests of the kinematics). */
void oracle_motion_eval(const double *motion, const double X[3], double t,
                        double u[3], double v[3], double a[3]);

#ifdef __cplusplus
}
#endif
#endif
