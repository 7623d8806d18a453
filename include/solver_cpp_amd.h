/*
 * solver_cpp_amd.h -- C-ABI of the MI355X-native spectral grid solver and its
 * phenopowerlaw material-point update.
 *
 * Two libraries export exactly these symbols:
 *   solver_cpp_amd/libsolver_cpp_amd.so   HIP/CDNA4 implementation (the product)
 *   oracle/_build/liboracle.so            FP64 CPU restatement (test infrastructure only)
 *
 * Layouts follow the reference: second-order tensors are (3,3) column-major
 * (element (i,j) at i+3j), fields are (3,3,Nx,Ny,Nz) column-major with the
 * tensor components fastest and x fastest among cells, fourth-order tensors
 * are (3,3,3,3) column-major.  Quaternions are (w,x,y,z).
 *
 * Every function that can fail returns 0 on success and a nonzero code
 * otherwise; sc_last_error() returns the message of the last failure on the
 * calling thread.
 */
#ifndef SOLVER_CPP_AMD_H
#define SOLVER_CPP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Configuration                                                              */
/* ------------------------------------------------------------------------- */

/* Elastic + phenopowerlaw constants of one cF phase with N_sl = [12].
 * Replaces the material.yaml phase entry parsed by material.f90:84-205,
 * phase_mechanical_elastic.f90:74-104 and
 * phase_mechanical_plastic_phenopowerlaw.f90:75-290. */
typedef struct sc_material {
  double C_11, C_12, C_44;        /* Pa */
  double xi_0_sl, xi_inf_sl;      /* Pa */
  double h_0_sl_sl;               /* Pa */
  double n_sl, a_sl;              /* stress exponent, hardening exponent */
  double dot_gamma_0_sl;          /* 1/s */
  double h_sl_sl[7];              /* interaction values, cF types 1..7 */
  double atol_xi, atol_gamma;     /* state tolerances (defaults 1.0, 1e-6) */
} sc_material;

/* Phase numerics, defaults of phase_mechanical.f90:280-294. */
typedef struct sc_phase_numerics {
  double r_cutback_min;           /* 1e-3 */
  double r_cutback;               /* 0.25 */
  double r_increase;              /* 1.5 */
  double eps_rel_state;           /* 1e-6 */
  int    N_iter_state_max;        /* 20 */
  int    N_iter_Lp_max;           /* 40 */
  double r_linesearch_Lp;         /* 0.5 */
  double eps_rel_Lp;              /* 1e-6 */
  double eps_abs_Lp;              /* 1e-8 */
} sc_phase_numerics;

/* Grid solver numerics, Config::NumGridMechanical / NumFFT / Numerics
 * (src/grid/config.h:50-78). */
typedef struct sc_grid_numerics {
  int    itmin, itmax;            /* 1, 100 (examples/grid/numerics.yaml: 4, 40) */
  int    update_gamma;            /* 0 */
  double eps_abs_div, eps_rel_div;/* 1e-4, 5e-4 */
  double eps_abs_P, eps_rel_P;    /* 1e3, 1e-3 */
  int    divergence_correction;   /* 0 none, 1 size, 2 size+grid (default 2) */
  int    derivative;              /* 0 continuous, 1 central diff, 2 FWBW */
  int    stag_iter_max;           /* 10 */
  int    max_cut_back;            /* 3 */
  int    ngmres_m;                /* NGMRES history size (PETSc default 30) */
  int    fft_plan_mode;           /* FFT.FFTW_plan_mode (config.cpp:432-441): 0 estimate (heuristic
                                     launch shapes), 1 measure (default: launch shapes timed per grid;
                                     patient and exhaustive map here).  Shapes change no result. */
} sc_grid_numerics;

/* Arithmetic of the spectral stages. */
enum {
  SC_PRECISION_FP64  = 0,  /* everything FP64 (plain form) */
  SC_PRECISION_MIXED = 1   /* FP32 FFT / Gamma / residual fields, FP64 material point,
                              FP64 norms, FP64 finishing sweeps per increment */
};

typedef struct sc_rve_desc {
  int           cells[3];
  double        size[3];          /* m */
  int           n_materials;
  const double* orientations;     /* (4, n_materials), unit quaternions w,x,y,z */
  const int*    material_at;      /* (Nx*Ny*Nz), 0-based material index, x fastest */
  sc_material        material;
  sc_phase_numerics  phase_num;
  sc_grid_numerics   grid_num;
  int           precision;        /* SC_PRECISION_* */
} sc_rve_desc;

/* One load step, Config::LoadStep (src/grid/config.h:80-96).
 * Masks are 1 where the entry was 'x' (not prescribed). */
enum { SC_DEF_F = 0, SC_DEF_DOT_F = 1, SC_DEF_L = 2 };
enum { SC_STRESS_P = 0, SC_STRESS_DOT_P = 1 };
typedef struct sc_loadstep {
  int    deformation_type;        /* SC_DEF_* */
  double deformation[9];          /* (3,3) column-major */
  int    deformation_mask[9];
  int    stress_type;             /* SC_STRESS_* */
  double stress[9];
  int    stress_mask[9];
  double rot_q[4];                /* R, (w,x,y,z); identity = (1,0,0,0) */
  double t;                       /* duration, s */
  int    N;                       /* increments */
  double r;                       /* geometric ratio (1 = constant) */
  int    estimate_rate;           /* 1 */
  int    f_out;                   /* result cadence */
} sc_loadstep;

/* Config::SolutionState (src/grid/config.h:35-40) plus counters. */
typedef struct sc_solution {
  int converged;
  int stag_converged;
  int terminally_ill;
  int required_iterations;        /* total_iter at exit */
  int residual_evaluations;       /* form_residual calls in this solve */
  double err_div, err_BC;         /* last values */
} sc_solution;

/* Run statistics of one RVE context (for measurement). */
typedef struct sc_stats {
  int64_t residual_evaluations;   /* form_residual calls */
  int64_t material_point_updates; /* N * residual evaluations */
  int64_t increments;             /* converged increments */
  int64_t cutbacks;
  int64_t newton_iterations;      /* sum over points and evaluations (if counted) */
  int64_t state_iterations;       /* sum over points and evaluations (if counted) */
  double  seconds_total;          /* wall time inside sc_rve_* calls */
  int64_t fp32_evaluations;       /* residual evaluations with FP32 spectral stages (mixed precision) */
} sc_stats;

/* ------------------------------------------------------------------------- */
/* Library                                                                    */
/* ------------------------------------------------------------------------- */

const char* sc_last_error(void);
/* "hip-gfx950" for the product, "cpu-oracle" for the oracle. */
const char* sc_backend_name(void);
void sc_default_material(sc_material* m);         /* examples/grid/material.yaml:7-24 */
void sc_default_phase_numerics(sc_phase_numerics* p);
void sc_default_grid_numerics(sc_grid_numerics* g);

/* ------------------------------------------------------------------------- */
/* Device-resident RVE context (the spectral basic solver)                    */
/* ------------------------------------------------------------------------- */

typedef struct sc_rve sc_rve;

/* DiscretizationGrid::init + Spectral::init + materialpoint init
 * (discretization_grid.cpp:8-84, spectral.cpp:12-96, materialpoint.f90:68-156). */
int  sc_rve_create(const sc_rve_desc* desc, sc_rve** out);
void sc_rve_destroy(sc_rve* rve);

/* MechSolverBasic::init, mech_solver_basic.cpp:19-108. */
int sc_rve_init(sc_rve* rve);

/* MechSolverBasic::forward, mech_solver_basic.cpp:163-220, followed by
 * f_materialpoint_forward when cut_back == 0 (damask_grid.cpp:19-34). */
int sc_rve_forward(sc_rve* rve, int cut_back, int guess, double dt, double dt_old,
                   double t_remaining, const sc_loadstep* step);

/* MechSolverBasic::calculate_solution, mech_solver_basic.cpp:110-161. */
int sc_rve_solution(sc_rve* rve, sc_solution* out);

/* MechSolverBasic::update_coords, mech_solver_basic.cpp:222-240
 * (base_update_coords, mech_base.cpp:45-154). */
int sc_rve_update_coords(sc_rve* rve);

/* Restart between increments: the state the reference writes and reads
 * for a restart (phase_mechanical.f90:1264-1298 restartWrite/restartRead,
 * homogenization.f90:381-438, the solver's F, F_lastInc, F_aim fields of
 * the grid restart) -- material-point state and its converged copy, F,
 * F_last_inc, F_dot, coordinates, and the solver scalars -- in this
 * library's own binary file (not DADF5/HDF5).  restore needs a context
 * created from the same descriptor; the load-case position is the caller's.
 * Returns nonzero on I/O errors or a checkpoint of another grid. */
int sc_rve_checkpoint(sc_rve* rve, const char* path);
int sc_rve_restore(sc_rve* rve, const char* path);

/* DADF5 job result file of the context (result_init, result.f90:76-116;
 * material_init's cell_to mappings, result.f90:492-790).  material_yaml is
 * the text of the material configuration: it is stored under setup/, and
 * the phase and homogenization names and the output lists
 * (phase/<name>/mechanical/output, .../plastic/output) are read from it;
 * NULL selects Aluminum / SX / [F, P, F_e, F_p, L_p, O] / [xi_sl]
 * (examples/grid/material.yaml).  numerics_yaml (may be NULL) is stored
 * under setup/.  compression = 1 shuffles and deflates datasets of at least
 * 2^18 values, as the reference does where its HDF5 library allows it
 * (HDF5_utilities.f90:1982-2003); every dataset carries Fletcher-32
 * checksums either way. */
typedef struct sc_result_desc {
  const char* material_yaml;
  const char* numerics_yaml;
  int         compression;
} sc_result_desc;
int sc_rve_result_open(sc_rve* rve, const char* path, const sc_result_desc* desc);
/* materialpoint_result (materialpoint.f90:221-234): increment_<inc> with
 * t/s, phase/<phase>/mechanical/<outputs> of the last evaluation (tensors
 * (N,3,3) with [c][i][j] = T_ij, O (N,4) quaternions of transpose(rotational
 * part of F_e), xi_sl / gamma_sl (N,12) with slip-system labels),
 * homogenization/<name>/mechanical, geometry/u_p (N,3) and u_n
 * ((Nx+1)(Ny+1)Nz,3) displacements of the last update_coords.  While a file
 * is open, sc_rve_run_loadcase writes every increment whose index in its
 * load step is a multiple of the step's f_out (damask_grid.cpp:165-168). */
int sc_rve_result_write(sc_rve* rve, int inc, double t);
int sc_rve_result_close(sc_rve* rve);

/* DamaskGrid::loop_over_loadcase, damask_grid.cpp:82-196, with stag_it_max
 * and max_cut_back from the grid numerics.  Per-increment results are
 * written to inc_iterations / inc_evaluations (length >= total increments,
 * may be NULL).  Returns nonzero if an increment could not be solved. */
int sc_rve_run_loadcase(sc_rve* rve, const sc_loadstep* steps, int n_steps,
                        int* inc_iterations, int* inc_evaluations, int max_incs);

/* One residual evaluation r = form_residual(F) on the current solver state
 * (mech_solver_basic.cpp:286-339); F and r are (3,3,N) host arrays.
 * Exposed for parity tests of a single evaluation. */
int sc_rve_form_residual(sc_rve* rve, const double* F, double* r);

/* Replay of the outer nonlinear solver (PETSc SNESNGMRES with the options of
 * mech_solver_basic.cpp:29-31, m = history size): starting from x0 (n
 * values), the k-th function evaluation returns r_rec[k*n .. k*n+n) and the
 * point it was requested at is written to x_eval[k*n ..).  The solve is
 * declared converged once n_evals evaluations were made.  Used to check the
 * solver arithmetic against a recorded trace. */
int sc_ngmres_replay(int64_t n, int m, const double* x0, int n_evals, const double* r_rec, double* x_eval);

/* Field / state access (host copies, (3,3,N) column-major). */
enum {
  SC_FIELD_F = 0,        /* solver deformation gradient (F_PETSc) */
  SC_FIELD_P = 1,        /* first Piola-Kirchhoff stress of the last evaluation */
  SC_FIELD_FP = 2,       /* plastic deformation gradient */
  SC_FIELD_LP = 3,       /* plastic velocity gradient */
  SC_FIELD_S = 4,        /* second Piola-Kirchhoff stress (lattice frame) */
  SC_FIELD_FE = 5,       /* elastic deformation gradient */
  SC_FIELD_F_LAST_INC = 6,
  SC_FIELD_XI = 7,       /* (12,N) slip resistance */
  SC_FIELD_GAMMA = 8,    /* (12,N) accumulated shear */
  SC_FIELD_X_P = 9,      /* (3,N) cell centre coordinates (after update_coords) */
  SC_FIELD_X_N = 10,     /* (3,(Nx+1)(Ny+1)(Nz+1)) node coordinates */
  SC_FIELD_ITERATIONS = 11 /* (2,N) Lp Newton and state iterations of the last evaluation, as doubles */
};
int sc_rve_get_field(sc_rve* rve, int field, double* out, int64_t n_values);
int sc_rve_set_field(sc_rve* rve, int field, const double* in, int64_t n_values);

/* Solver scalars: P_av(9) P_aim(9) F_aim(9) F_aim_dot(9) F_aim_last_inc(9)
 * C_volAvg(81) C_minMaxAvg(81) S(81) C_ref(81), err_div, err_BC. */
typedef struct sc_solver_scalars {
  double P_av[9], P_aim[9], F_aim[9], F_aim_dot[9], F_aim_last_inc[9];
  double C_volAvg[81], C_minMaxAvg[81], S[81], C_ref[81];
  double err_div, err_BC;
  int    total_iter;
} sc_solver_scalars;
int sc_rve_get_scalars(sc_rve* rve, sc_solver_scalars* out);
int sc_rve_get_stats(sc_rve* rve, sc_stats* out);
int sc_rve_reset_stats(sc_rve* rve);

/* Kernel timing (no counterpart in the reference; for measurement).  With
 * profiling enabled every launch of the classes below is bracketed by HIP
 * events on the context's stream, and the material-point kernels' Newton and
 * state iterations are summed (then also reported in sc_stats). */
enum {
  SC_KT_RESPONSE = 0,  /* material-point update (stress integration)       */
  SC_KT_TANGENT = 1,   /* consistent tangent dPdF                           */
  SC_KT_FFT_FWD = 2,   /* forward 3D FFT of the 9 stress planes             */
  SC_KT_FFT_BWD = 3,   /* inverse 3D FFT of the 9 residual planes           */
  SC_KT_GAMMA = 4,     /* fused Gamma convolution + divergence              */
  SC_KT_REDUCE = 5,    /* reductions and N-GMRES vector updates             */
  SC_KT_OTHER = 6,
  SC_KT_COUNT = 7
};
typedef struct sc_kernel_times {
  int64_t launches[SC_KT_COUNT];
  double  ms[SC_KT_COUNT];        /* summed event time */
  int64_t points[SC_KT_COUNT];    /* summed cells (or wavevectors) processed */
} sc_kernel_times;
int sc_rve_profile(sc_rve* rve, int enable);
/* Profiling of the classes in mask only (bit 1 << SC_KT_*; 0 = off): every
 * timed class adds two events per launch on the solver's stream. */
int sc_rve_profile_classes(sc_rve* rve, unsigned mask);
int sc_rve_get_kernel_times(sc_rve* rve, sc_kernel_times* out);

/* ------------------------------------------------------------------------- */
/* Job server: continuous batching of RVE load increments on one GPU          */
/* (no counterpart in the reference: its driver runs one RVE per process).   */
/* ------------------------------------------------------------------------- */

/* Worst-case device bytes of one RVE context (admission footprint). */
int64_t sc_rve_footprint(const sc_rve_desc* desc);

typedef struct sc_server sc_server;
typedef struct sc_request_result {
  int64_t request, job;
  int     status;                 /* 0 solved, 1 load case exhausted, 2 not solved, 3 error */
  int     increment;              /* the job's increment count after this request */
  int     iterations, evaluations;
  double  P_av[9];                /* volume-averaged P after the increment */
  double  t;                      /* load-case time after the increment, s */
  int64_t material_point_updates; /* cells x residual evaluations */
  double  t_submit, t_start, t_end; /* seconds since sc_server_create */
} sc_request_result;
typedef struct sc_server_stats {
  int64_t budget_bytes, used_bytes, peak_bytes;
  int64_t resident_jobs, completed_requests;
  int     max_concurrent;         /* most requests running at once */
  int64_t evictions;              /* idle resident jobs checkpointed to host memory */
  int64_t batches;                /* joint launches of residual evaluations (batched server) */
  int64_t batched_evaluations;    /* residual evaluations computed in them */
  int64_t max_batch;              /* most evaluations in one joint launch */
  double  batch_seconds;          /* wall time of the joint launches (completion included) */
} sc_server_stats;

/* workers: host threads, each driving one admitted job at a time on its own
 * HIP stream; hbm_budget_bytes bounds the sum of resident jobs' footprints. */
int  sc_server_create(int64_t hbm_budget_bytes, int workers, sc_server** out);
/* As sc_server_create, and every residual evaluation of the jobs that are
 * inside an increment is computed in joint launches on one stream: the
 * material-point update of all their cells in one launch, the P sums and the
 * aims in one each, and per group of jobs with the same grid and arithmetic
 * one gather, one batched FFT per pass, one Gamma convolution and one
 * scatter.  A joint launch starts when every running job waits for an
 * evaluation or the oldest request has waited max_wait_ms.  Each job's
 * results are bitwise those of sc_server_create.  For throughput over many
 * RVEs, workers should be about the number of jobs run at once. */
int  sc_server_create_batched(int64_t hbm_budget_bytes, int workers, double max_wait_ms, sc_server** out);
void sc_server_destroy(sc_server* server);  /* finishes running requests */
/* A job: RVE description (copied) and its load case.  keep_resident = 0
 * releases the job's memory as soon as it has no pending request. */
int  sc_server_add_job(sc_server* server, const sc_rve_desc* desc, const sc_loadstep* steps, int n_steps,
                       int keep_resident, int64_t* job);
/* One increment of the job's load case; increments of a job run in order. */
int  sc_server_submit(sc_server* server, int64_t job, int64_t* request);
/* Blocks until done (timeout_s < 0: no limit; returns 4 on timeout). */
int  sc_server_wait(sc_server* server, int64_t request, double timeout_s, sc_request_result* out);
int  sc_server_remove_job(sc_server* server, int64_t job);
/* A field of a resident job with no request running or pending, in the
 * layouts of sc_rve_get_field (SC_FIELD_*). */
int  sc_server_get_field(sc_server* server, int64_t job, int field, double* out, int64_t n_values);
int  sc_server_get_stats(sc_server* server, sc_server_stats* out);

/* ------------------------------------------------------------------------- */
/* Spectral building blocks (parity tests against the reference's goldens).  */
/* All arrays are host arrays; the product runs them on the device.          */
/* ------------------------------------------------------------------------- */

/* Spectral::init xi tables (spectral.cpp:53-91): xi1st/xi2nd as interleaved
 * complex (re,im) of shape (3, Nx/2+1, Nz, Ny). */
int sc_spectral_xi(const int cells[3], const double size[3], int derivative,
                   double* xi1st, double* xi2nd);

/* 3D real-to-complex forward FFT (unnormalised, e^{-i}) of `howmany`
 * interleaved components: in (howmany, Nx, Ny, Nz) real, out (howmany,
 * Nx/2+1, Nz, Ny) complex interleaved (the FFTW_MPI_TRANSPOSED_OUT layout of
 * fft.cpp:48-52). */
int sc_fft_r2c(const int cells[3], int howmany, const double* in, double* out);
/* Inverse (e^{+i}), scaled by wgt as FFT::backward (fft.h:50-60). */
int sc_fft_c2r(const int cells[3], int howmany, const double* in, double wgt, double* out);

/* MechBase::calculate_divergence_rms, mech_base.cpp:302-398. field (3,3,N). */
int sc_divergence_rms(const int cells[3], const double size[3], const double scaled_size[3],
                      const double* xi1st, const double* field, double* rms);

/* MechBase::gamma_convolution with the operator as intended
 * (mech_base.cpp:400-477, Gamma_lmno = (A^-1)_ln conj(-xi_o) xi_m, from the
 * commented-out fill at :188-205).  C_ref is the reference stiffness as
 * stored by update_gamma (C / wgt); DC = field_aim; unnormalised inverse.
 * field (3,3,N) in place. */
int sc_gamma_convolution(const int cells[3], const double* xi1st, const double* C_ref,
                         const double field_aim[9], double* field);

/* MechBase::calculate_masked_compliance, mech_base.cpp:249-300. */
int sc_masked_compliance(const double C[81], const double rot_q[4], const int stress_mask[9],
                         double S_out[81]);

/* MechBase::base_init scaled geometry, mech_base.cpp:11-33. */
int sc_scaled_geom_size(const int cells[3], const double size[3], int divergence_correction,
                        double scaled[3]);

/* MechBase::forward_field (mech_base.cpp:213-247) and calculate_rate
 * (:479-497) on (3,3,N) fields. aim may be NULL. */
int sc_forward_field(int64_t n_cells, double dt, const double* field_last_inc,
                     const double* rate, const double* aim, double* out);
int sc_calculate_rate(int64_t n_cells, int heterogeneous, const double* field0,
                      const double* field, double dt, const double avRate[9], double* out);

/* P_av / C_volAvg / C_minMaxAvg reductions of Spectral::constitutive_response
 * (spectral.cpp:160-217) on given per-point P (3,3,N) and dPdF (3,3,3,3,N). */
int sc_response_reductions(int64_t n_cells, const double* P, const double* dPdF,
                           const double* rot_q /* may be NULL */,
                           double P_av[9], double C_volAvg[81], double C_minMaxAvg[81]);

/* ------------------------------------------------------------------------- */
/* Material-point building blocks                                             */
/* ------------------------------------------------------------------------- */

/* Batched material-point update for n independent points (the inner loop of
 * homogenization_mechanical_response, homogenization.f90:219-266, followed by
 * mechanical_homogenize, homogenization_mechanical.f90:128-151).
 *
 * Inputs per point (column-major 3x3, SoA over points is NOT used here:
 * arrays are (3,3,n) / (12,n) as in the reference):
 *   F, F0 (phase F0), Fp0, Lp0, S0, xi0, gamma0 -- converged values of the last increment
 *   Lp  -- in: Newton initial guess (Lp of the previous evaluation); out: Lp
 * Outputs: P, Fp, S, Fe, xi, gamma, dPdF (may be NULL), broken (n, may be NULL),
 *          iterations (2n: Newton total, FPI total; may be NULL).
 * Orientation quaternions are not needed: Fp carries the orientation. */
int sc_mp_update(int64_t n, double dt, const sc_material* m, const sc_phase_numerics* num,
                 const double* F, const double* F0, const double* Fp0, const double* Lp0,
                 const double* S0, const double* xi0, const double* gamma0,
                 double* Lp, double* P, double* Fp, double* S, double* Fe,
                 double* xi, double* gamma, double* dPdF, int* broken, int* iterations);

/* Material-point update without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The update of sc_mp_update, enqueued on the caller's stream (NULL: the null
 * stream) on device arrays the caller owns, in structure-of-arrays layout:
 * a (3,3) field is 9 planes of n doubles (plane q = i+3j holds component
 * (i,j) of every point), a slip field 12 planes of n doubles, dPdF 81 planes
 * (plane i+3j+9k+27l), iterations 2 planes of n ints (Newton, state
 * iterations) and broken n ints.  The substep scratch, the pivoting pass's
 * point list and the terminally-ill flag live in one fixed device array of
 * the library sized for SC_MP_NOALLOC_MAX_POINTS points (about 16 MB, held
 * on each GPU the library runs on until the process ends); larger n is
 * refused with an error.  The call allocates nothing.  Calls that overlap on
 * one GPU share that array and race: serialise them. */
#define SC_MP_NOALLOC_MAX_POINTS 32768
typedef struct sc_mp_device_planes {
  const double *F, *F0, *Fp0, *Lp0, *S0, *xi0, *gamma0;  /* inputs, as in sc_mp_update */
  double* Lp;                                            /* in: Newton guess; out: Lp */
  double *P, *Fp, *S, *Fe, *xi, *gamma;                  /* outputs */
  double* dPdF;                                          /* may be NULL */
  int* broken;                                           /* may be NULL */
  int* iterations;                                       /* may be NULL */
} sc_mp_device_planes;
int sc_mp_update_noalloc(int64_t n, double dt, const sc_material* m, const sc_phase_numerics* num,
                         const sc_mp_device_planes* planes, void* stream);

/* Orientation matrix of a unit quaternion, qu2om with P = -1
 * (rotations.f90:423-447); out (3,3) column-major. */
void sc_qu2om(const double q[4], double om[9]);
/* math_rotationalPart (math.f90:1117-1156), used by crystallite_orientations. */
void sc_rotational_part(const double F[9], double R[9]);
/* Quaternion (w >= 0) of an orientation matrix, om2qu (rotations.f90),
 * inverse of sc_qu2om; crystallite_orientations (phase.f90:590) stores the
 * orientation of transpose(rotational part of F_e) this way. */
void sc_om2qu(const double om[9], double q[4]);

/* HDF5 writer of the DADF5 files (the storage of HDF5_utilities.f90:
 * 1453-1500, 1934-2037: chunked datasets with Fletcher-32, shuffle + deflate
 * for large ones).  A host-only utility beneath sc_rve_result_*; names are
 * "group/.../dataset" paths, parent groups are created, dims in C order. */
typedef struct sc_dadf5 sc_dadf5;
int sc_dadf5_create(const char* path, int compression, sc_dadf5** out);
int sc_dadf5_dataset_f64(sc_dadf5* f, const char* name, int rank, const int64_t* dims, const double* data);
int sc_dadf5_dataset_str(sc_dadf5* f, const char* name, const char* text);
int sc_dadf5_attr_str(sc_dadf5* f, const char* object, const char* name, const char* value);
int sc_dadf5_attr_f64(sc_dadf5* f, const char* object, const char* name, double value);
int sc_dadf5_close(sc_dadf5* f);  /* writes the metadata and closes */

/* ------------------------------------------------------------------------- */
/* Drop-in boundary: the Fortran bind(C) exports the reference's C++ grid     */
/* solver calls (src/grid/spectral.h:16-23, damask_grid.h:16-20,              */
/* main.cpp:8-14, discretization_grid.h:9-22, utilities_fortran.h:8-15).      */
/* Arrays behind the fetched pointers are host mirrors owned by the library;  */
/* in the product the material-point state stays resident on the device.     */
/* ------------------------------------------------------------------------- */

/* materialpoint.f90:68-117.  material/load/grid/numerics are paths
 * (material.yaml, load.yaml, geometry .vti, numerics.yaml); the lengths are
 * passed by value as in the reference. */
void f_materialpoint_initBase(const char* material, int n1, const char* load, int n2,
                              const char* grid, int n3, const char* numerics, int n4);
/* materialpoint.f90:138-156 (phase, homogenization and state allocation). */
void f_materialpoint_initDamask(void);
/* materialpoint.f90:208-215: state0 <- state, Fp0 <- Fp, ... */
void f_materialpoint_forward(void);
/* materialpoint.f90:221-234: writes increment_<inc> of the DADF5 job file
 * created by f_materialpoint_initDamask (see sc_boundary_set_result_file). */
void f_materialpoint_result(int* inc, double* t);

/* homogenization.f90:491-510 */
void f_homogenization_fetch_tensor_pointers(double** F0, double** F, double** P,
                                            double** dPdF, void** terminally_ill);
/* homogenization.f90:219-266, 1-based inclusive cell range. */
void f_homogenization_mechanical_response(double* dt, int* cell_start, int* cell_end);
/* homogenization.f90:272 (no thermal field is active: no-op). */
void f_homogenization_thermal_response(double* dt, int* cell_start, int* cell_end);
/* homogenization.f90:299-324 (orientations + P, dPdF homogenization). */
void f_homogenization_mechanical_response2(double* dt, int* execIP, int* execElem);

/* discretization.f90:79-100, :145 */
void f_discretization_init(int* material_at, int* n_cells, double* ip0, int* n_ip,
                           double* node0, int* n_node, int* shared_nodes_begin);
void f_discretization_fetch_ip_node_coord_pointers(double** ip, double** node);

/* VTI.f90:146-202 and :473-497 (bind(C) names of the reference's VTI
 * reader, called by discretization_grid.h:64-76): cells, size = spacing *
 * cells and origin of a VTK image-data file held in memory, and an integer
 * cell dataset returned +1 (1-based material entries).  Errors leave
 * cells = -1 / dataset_size = -1 and set sc_last_error(). */
void f_VTI_readCellsSizeOrigin(const char* content, int length, int* cells, double* size, double* origin);
void f_VTI_readDataset_int(const char* content, int length, const char* label, int label_length,
                           int* dataset, int* dataset_size);

/* Boundary state for tests: resident constitutive fields of the cells
 * (SC_FIELD_FP, _LP, _S, _FE, _P, _XI, _GAMMA; (3,3,N) or (12,N)), the number
 * of f_materialpoint_result calls, and a reset of the process-global state.
 * No counterpart in the reference. */
int  sc_boundary_get_state(int field, double* out, int64_t n_values);
int  sc_boundary_result_count(void);
void sc_boundary_reset(void);
/* DADF5 job file of the boundary.  Default (and after a reset): the
 * reference's job name stem(grid)_stem(load)_stem(material)[_stem(numerics)]
 * + ".hdf5" in the working directory (CLI.f90:311-333, result.f90:87).
 * NULL or "" writes no file; compression as in sc_result_desc.  Takes
 * effect at the next f_materialpoint_initDamask.  No counterpart in the
 * reference (its job name comes from the command line). */
int  sc_boundary_set_result_file(const char* path, int compression);

/* ------------------------------------------------------------------------- */
/* Job inputs (host)                                                          */
/* ------------------------------------------------------------------------- */

/* Geometry file: cells, size, origin and the 'material' cell dataset as
 * 0-based material indices (material_at may be NULL to query the cells). */
int sc_read_vti(const char* path, int cells[3], double size[3], double origin[3], int* material_at,
                int64_t capacity);
/* material.yaml with one cF phenopowerlaw phase and one constituent per
 * material entry (material.f90:84-205): constants and the (4, n) entry
 * orientations (orientations may be NULL to query n). */
int sc_read_material(const char* path, sc_material* m, double* orientations, int capacity, int* n_materials);
/* numerics.yaml (config.cpp:362-460, phase_mechanical.f90:280-294): updates
 * the given structs (initialise them with the defaults first). */
int sc_read_numerics(const char* path, sc_grid_numerics* g, sc_phase_numerics* p);
/* Load-case YAML (config.cpp:240-360): loadsteps in file order (steps may be
 * NULL to query the count).  t, N and r are read as integers, as there. */
int sc_read_load(const char* path, sc_loadstep* steps, int capacity, int* n_steps);

/* math.f90:599 (n x n inverse, err = 1 on failure), :819, :849, :428 */
void f_math_invert(double* invA, int* err, double* A, int* n);
void f_math_3333to99(double* m3333, double* m99);
void f_math_99to3333(double* m99, double* m3333);
void f_math_mul3333xx33(double* A, double* B, double* res);
/* rotations.f90:331-360: passive unless *active == 1. */
void f_rotate_tensor2(double* q, double* T, double* out, int* active);
void f_rotate_tensor4(double* q, double* T, double* out, int* active);

#ifdef __cplusplus
}
#endif
#endif /* SOLVER_CPP_AMD_H */
