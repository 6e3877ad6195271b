/*
 * idp_amd.h - C ABI of the MI355X (gfx950) Newton-with-barrier inner loop.
 *
 * Plain C types only: pointers, sizes, status codes.  Host arrays are
 * borrowed for the duration of a call; device state lives in an opaque
 * context (one HIP stream each, not thread-safe; use one context per thread).
 *
 * Layouts (same as the reference's flattened Storage objects):
 *   positions, velocities, gradients, directions: 3*nV doubles, vertex-major
 *   (x0,y0,z0,x1,...), i.e. sol/rhs of Library/FEM/Shell/IMPLICIT_EULER.h
 *   tets: 4*nT ints; IB: 9*nT doubles, row-major 3x3 inverse rest basis
 *   constraints: 4 ints each, the encoding of Library/FEM/IPC.h:19-740
 *     (-p-1,t0,t1,t2) PT   (-p-1,e0,e1,-m) PE x m   (-p-1,q,-1,-m) PP x m
 *     (a0,a1,b0,b1) EE   (a0,a1,-b0-1,b1) / (p,e0,e1,-o-1) / (p,q,-a-1,-b-1)
 *     mollified EE / PE / PP
 *   local Hessians: 144 doubles per constraint (nv*3 square, row-major,
 *     rows ordered like the reference's triplet emission) plus 4 vertex ids.
 */
#ifndef IDP_AMD_H
#define IDP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference exits or prints; see each entry point) */
#define IDP_OK 0
#define IDP_E_INVALID (-1)      /* bad argument / no mesh */
#define IDP_E_HIP (-2)          /* HIP runtime error */
#define IDP_E_NONPOS_DIST (-3)  /* dist2 <= 0 in a barrier term: IPC.h:773-776 exit(-1) */
#define IDP_E_NOT_SPD (-4)      /* linear solve breakdown: DIRECT_SOLVER.h:42-44 returns false */
#define IDP_E_CCD_ZERO (-5)     /* zero time of impact: IPC.h:2014-2031 exit(-1) */
#define IDP_E_CAPACITY (-6)     /* output array too small; *n holds the size needed */
#define IDP_E_INVALID_START (-7)/* degenerate tet (det F <= 0) at the start of a step */
#define IDP_E_NO_GPU (-8)       /* no HIP device */

typedef struct idp_ctx idp_ctx;

typedef struct idp_solver_opts {
    double pcg_rtol;  /* relative residual of the Newton-direction solve (default 1e-8) */
    int pcg_maxit;    /* default 10000 */
    int max_newton;   /* 0 = until the reference stopping rule holds */
    int verbose;      /* > 0: print a progress line to stderr every `verbose` Newton iterations */
    int precond;      /* 0: cluster block Jacobi, clusters matched over elastic and contact
                         couplings (default); 1: 3x3 block Jacobi; 2: clusters matched over
                         contact couplings only */
    int pcg_cold_start; /* 1: every PCG solve starts from zero; 0 (default): from the second Newton
                           iteration of a step on, from the best multiple of the previous direction */
    int accd_budget;    /* batched engine: ACCD steps a pair may take per Newton iteration before it is
                           handed over to the next one (its instance waits, the others go on); 0: the
                           default (65536), < 0: never.  Results do not depend on it. */
    int pcg_handover;   /* batched engine: 0 (default): once half of the instances' PCG solves have
                           stopped, the others run as many iterations again and then continue in the
                           next Newton iteration while the rest go on; < 0: never.  Results do not
                           depend on it. */
    int accd_max_steps; /* additive-CCD steps one pair may take in one step-size call before it is
                           taken as a hit at the time of impact it reached (a smaller, still
                           conservative step).  0: the default, IDP_ACCD_MAX_STEPS_DEFAULT; < 0: no
                           cap, as in the reference's loop (CCD.h:308-327, 367-392).  Capped pairs
                           are counted in idp_step_stats.accd_capped. */
    int coarse_verts;   /* one-mesh steps and Newton directions: a coarse level in the PCG
                           preconditioner (rigid-body modes of voxel aggregates of about this many
                           free vertices, added to the cluster block Jacobi); 0: the default,
                           IDP_COARSE_VERTS_DEFAULT; < 0: none.  Changes the PCG path within its
                           stated relative residual only. */
} idp_solver_opts;

#define IDP_ACCD_MAX_STEPS_DEFAULT (1 << 20)
#define IDP_COARSE_VERTS_DEFAULT 144

typedef struct idp_step_stats {
    int newton_iters;
    int line_search_trials;
    int constraint_sets;
    int gd_iters;
    long pcg_iters;
    long n_constraints;
    long ccd_candidates;
    double last_alpha, first_ccd_alpha, last_res, last_relres, min_dist2, E;
    double kappa;
    /* stage wall time in ms (stream-synchronous timings) */
    double ms_total, ms_constraint_set, ms_ccd, ms_energy, ms_gradient, ms_hessian, ms_solve;
    /* moving DBC: the CCD-limited fraction of the move taken at the start of
       the step (1: all of it), and the penalty stiffness at the end (0: off) */
    double dbc_alpha, dbc_stiff;
    /* Newton directions taken from a PCG solve that reached pcg_maxit before
       its residual test held (the reference's direct solve has no such case;
       the direction is used as it stands) */
    int pcg_unconverged;
    /* ACCD pairs stopped by the step cap (accd_max_steps) and taken as hits at
       the time of impact they had reached */
    long accd_capped;
    /* lagged friction: constraints in the friction set of the last basis
       update, and the basis updates after the step's first
       (IMPLICIT_EULER.h:757-848) */
    long n_friction;
    int friction_updates;
    /* ACCD pairs ended with no hit by the certificate that their distance
       stays above the stopping gap up to their bound (no steps taken past it) */
    long accd_certified;
} idp_step_stats;

/* Kernel-family profile: device time from HIP events recorded on the
 * context's stream around each launch, plus the algorithmic work of those
 * launches (bytes for HBM-bound families, flops for compute-bound ones). */
enum {
    IDP_K_PCG = 0,          /* whole block-Jacobi PCG solve, launch gaps included (bytes) */
    IDP_K_BARRIER_HESS = 1, /* barrier Hessian + PSD projection (flops) */
    IDP_K_TET_HESS = 2,     /* tet projected 12x12 (flops) */
    IDP_K_ASSEMBLY = 3,     /* elastic + contact BSR value assembly (bytes) */
    IDP_K_CONSTRAINT_SET = 4, /* whole contact-set stage (no work count) */
    IDP_K_CCD = 5,          /* ACCD kernel over candidate pairs (pair count) */
    IDP_K_ENERGY = 6,       /* tet + inertia + barrier energy kernels */
    IDP_K_GRADIENT = 7,     /* tet + barrier gradient and vertex gather */
    IDP_K_PCG_SPMV = 8,     /* PCG SpMV kernel, sampled on one iteration in 8 (bytes) */
    IDP_K_PCG_UPDATE = 9,   /* PCG vector-update kernel, sampled likewise (bytes) */
    IDP_K_FRICTION = 10,    /* friction Hessian kernel (bytes) */
    IDP_K_FRICTION_BASIS = 11, /* friction basis update: flags, scan, records (bytes) */
    IDP_K_COUNT = 16
};
typedef struct idp_profile {
    double ms[IDP_K_COUNT];
    long launches[IDP_K_COUNT];
    double work[IDP_K_COUNT];
    long barrier_pairs[8];   /* Hessian evaluations per constraint kind */
    long barrier_clamped[8]; /* ... of which needed the PSD clamp */
    double ccd_steps;        /* additive-CCD iterations over all pairs */
    double ccd_max_steps;    /* iterations of the longest pair */
} idp_profile;

/* ---- context ---------------------------------------------------------- */
int idp_ctx_create(int device, idp_ctx** out);
int idp_ctx_destroy(idp_ctx* ctx);
const char* idp_last_error(const idp_ctx* ctx);
int idp_device_count(int* n);
/* Device memory held by the library's buffers in this process, over every
 * context and GPU (no counterpart in the reference): bytes now and the most
 * held at once.  Buffers grow to their working sizes on first use and are
 * kept; a serving loop measures a slot's footprint from the difference
 * before and after a context's first solve. */
int idp_device_memory(long long* in_use, long long* peak);
/* Profiling (no counterpart in the reference, which only has wall-clock
 * TIMER_FLAG scopes): on = 1 resets the counters and records every kernel
 * family (an event pair around each launch, constraint-kind counters in the
 * barrier Hessian); on = 2 records only the PCG SpMV and update kernels on
 * one iteration in eight, which leaves the launch stream as it runs
 * unprofiled; on = 0 stops. */
int idp_profile_enable(idp_ctx* ctx, int on);
int idp_profile_read(idp_ctx* ctx, idp_profile* out);
/* Launch configuration of the PCG kernels (no counterpart in the reference;
 * set by the autotuner, idp_amd/autotune.py).  spmv_shape 0..7 indexes
 * {lanes per block row G} x {scalars in wavefront 0, in a fifth wavefront}
 * with G = 4, 8, 16, 32 for 0..3 and again for 4..7; upd_waves 1..8 is the
 * number of 64-row update blocks per workgroup.  upd_waves and the
 * wavefront placement never change results; G changes the order of each
 * block row's sum, so one G serves every context of a process. */
int idp_set_launch_config(idp_ctx* ctx, int spmv_shape, int upd_waves);
int idp_get_launch_config(idp_ctx* ctx, int* spmv_shape, int* upd_waves);
/* Lanes per order-9 constraint (point-triangle, edge-edge, mollified) in the
 * projection stage of the barrier Hessian: 3 (a lane triple holding K in
 * registers, one eigenvector component each) or 8, 16, 32, 64 (a team running
 * a parallel Jacobi in LDS); default 3.  No counterpart in the reference;
 * set by the autotuner.  Results agree to rounding across the choices. */
int idp_set_barrier_team(idp_ctx* ctx, int lanes);
int idp_get_barrier_team(idp_ctx* ctx, int* lanes);

/* ---- host-side set-up helpers (no device work) ------------------------- */
/* Compute_Vol_And_Inv_Basis (Library/FEM/DEFORMATION_GRADIENT.h:39-76),
 * FIXED_COROTATED::All_Append_FEM Lame parameters (Library/Physics/FIXED_COROTATED.h:56-80),
 * Compute_Mass_And_Init_Velocity_NoAlloc lumped mass (Library/FEM/ELEM_TO_NODE.h:11-47). */
int idp_fem_init(int nV, const double* X, int nT, const int* tets, double density, double E, double nu,
                 double* vol, double* IB, double* lambda, double* mu, double* mass);
/* Find_Surface_TriMesh<T,false> + Find_Surface_Primitives_And_Compute_Area
 * (Library/Utils/MESHIO.h:486-545, 769-835).  bnode >= nV ints, btri >= 12*nT,
 * bedge >= 12*nT. */
int idp_surface(int nV, const double* X, int nT, const int* tets, int* nb, int* ne, int* nf, int* bnode,
                int* bedge, int* btri);
/* Initialize_Discrete_Shell<double,3,false,false> (Library/FEM/Shell/DISCRETE_SHELL.h:219-354),
 * bound as FEM.DiscreteShell.Initialize_Shell_Hinge: keeps the triangles with
 * a nonzero first-fundamental-form determinant, in order (kept: 3 nE ints,
 * nE_kept their count); mass (nV, accumulated into: zero it for a pure
 * shell) += |cross(X2-X1, X3-X1)| / 2 * thickness * density / 3 per kept
 * triangle in order; vol = area * thickness, lambda = E nu / (1 - nu^2),
 * mu = E / (2 (1 + nu)) per kept triangle. */
int idp_shell_init(int nV, const double* X, int nE, const int* tris, double density, double E, double nu,
                   double thickness, int* nE_kept, int* kept, double* mass, double* vol, double* lambda, double* mu);
/* The boundary primitives of a mesh with tets and shell triangles as the
 * shell step lists them (Library/FEM/Shell/IMPLICIT_EULER.h:233-241): the
 * tets' surface triangles (MESHIO.h:486-545) followed by the shell triangles
 * in order, then MESHIO.h:769-835 over that list.  nT may be 0.
 * bnode >= nV, bedge >= 2 (3 nE + 12 nT), btri >= 3 (nE + 4 nT) ints. */
int idp_shell_surface(int nV, const double* X, int nT, const int* tets, int nE, const int* tris, int* nb, int* ne,
                      int* nf, int* bnode, int* bedge, int* btri);
/* Hinge stencils and rest data of a shell (Compute_Discrete_Shell_Inv_Basis,
 * Library/FEM/Shell/DISCRETE_SHELL.h:132-215): for every edge of the
 * triangles (first-seen orientation, sorted: MESHIO.h:728-766) with a
 * triangle on both sides, the stencil (v, e0, e1, w) and (rest dihedral
 * angle, |e|, (|n1| + |n2|) / (6 |e|)); tri_IB (nullable, 3 per triangle): the
 * rest first fundamental forms.  hinges >= 4 * 3 nE / 2 ints, hinge_rest
 * >= 3 * 3 nE / 2 doubles. */
int idp_shell_hinges(int nV, const double* X, int nE, const int* tris, int* nH, int* hinges, double* hinge_rest,
                     double* tri_IB);
/* Initialize_EIPC<double,false> as bound to Initialize_OIPC
 * (Library/FEM/Shell/DISCRETE_SHELL.h:555-577, 1105). kappa3 = {kappa, kappa_max, 0}. */
int idp_initialize_oipc(double dhat, int nV, const double* mass, double stiff_mult, double* kappa3,
                        double* dHat2);

/* ---- device mesh and state -------------------------------------------- */
/* Uploads topology, rest positions (nodeAttr.x0: mollifier thresholds),
 * tetAttr.IB, tetElasticityAttr vol/lambda/mu and nodeAttr.m, extracts the
 * surface primitives and precomputes the static elastic BSR pattern. */
int idp_set_mesh(idp_ctx* ctx, int nV, const double* X_rest, int nT, const int* tets, const double* IB,
                 const double* vol, const double* lambda, const double* mu, const double* mass);
/* A mesh with shell triangles (Elem of the shell step), with or without
 * tets (nT may be 0; then tets..mu may be null): tri_vol is the vol of the
 * shell's elasticityAttr (area * thickness, from idp_shell_init).  The
 * surface primitives are those of idp_shell_surface.  One instance.  Shell
 * triangles carry the flow variant's -h/2 x^T L x term
 * (Library/FEM/Shell/INC_POTENTIAL.h:53-73, 191-211, 323-339) and, in
 * non-flow steps, the membrane and hinge bending of idp_set_shell_elasticity
 * (IDP_E_INVALID until it is called).  hinges (4 per hinge, from
 * idp_shell_hinges; nH may be 0) couple the four vertices of each stencil in
 * the static BSR pattern. */
int idp_set_mesh_shell(idp_ctx* ctx, int nV, const double* X_rest, const double* mass, int nT, const int* tets,
                       const double* IB, const double* vol, const double* lambda, const double* mu, int nE,
                       const int* tris, const double* tri_vol, int nH, const int* hinges);
/* The non-flow shell terms (Advance_One_Step_IE_Hinge on a shell): the
 * neo-Hookean membrane (Library/FEM/Shell/MEMBRANE.h:8-312, useNH) per
 * triangle from its rest first fundamental form tri_IB (3 per triangle:
 * |E01|^2, E01.E02, |E02|^2) and Lame parameters, and the discrete hinge
 * bending (Library/FEM/Shell/BENDING.h:9-80, 150-214, 440-497) per hinge of
 * idp_set_mesh_shell from hinge_rest (3 per hinge: rest angle, |e|, h) with
 * stiffness bend_k (bendingStiffMult E t^3 / (24 (1 - nu^2)),
 * DISCRETE_SHELL.h:333-339).  Elements whose vertices are all Dirichlet
 * nodes contribute nothing, as there. */
int idp_set_shell_elasticity(idp_ctx* ctx, const double* tri_IB, const double* tri_lambda, const double* tri_mu,
                             const double* hinge_rest, double bend_k);
/* Update_Normal_Flow_Neumann (Library/FEM/Shell/DISCRETE_SHELL.h:359-395) on
 * the device: b (3 nV, host) = magnitude * m_v * n_v / |n_v| with n_v the sum
 * of cross(X2-X1, X3-X1) over the shell triangles of v in order, at X (host,
 * 3 nV; null: the context's current x).  Equal bit for bit to the
 * sequential loop. */
int idp_normal_flow_force(idp_ctx* ctx, const double* X, double magnitude, double* b);
/* Steps given no body force (b = null) recompute the normal-flow force from
 * their start positions on the device first, as the shell driver calls
 * Update_Normal_Flow_Neumann before each step (FEMDiscreteShellBase.py:269-273);
 * magnitude 0 turns it off (the default). */
int idp_set_normal_flow(idp_ctx* ctx, double magnitude);
/* Lagged IPC friction of this context's steps: the mu, epsv2, fricIterAmt,
 * compNodeRange and muComp arguments of Advance_One_Step_IE_Discrete_Shell
 * (Library/FEM/Shell/IMPLICIT_EULER.h:166-169; the friction terms are
 * Library/FEM/FRICTION.h:16-660).  At the step's start, and again each time
 * the Newton loop meets its tolerance while fewer than fric_iters bases have
 * been made (fric_iters <= 0: no limit), the friction set is rebuilt from
 * the barrier set: the non-mollified constraints with their closest points,
 * tangent bases and normal forces; the potential mu sum N f0(|T^T dx|) with
 * epsv h as the static bound then enters the energy, gradient and Hessian.
 * ncomp > 0: comp_range (ncomp ascending vertex bounds: vertex v is in the
 * first component c with v < comp_range[c]) and mu_comp (ncomp^2,
 * mu_comp[c0 + c1 ncomp]) scale each normal force and mu is taken as 1
 * (Compute_Friction_Coef, FRICTION.h:119-169).  mu 0 and ncomp 0 turn
 * friction off (the default).  Steps of one mesh (idp_advance_one_step);
 * idp_advance_batch and serving refuse a context with friction on. */
int idp_set_friction(idp_ctx* ctx, double mu, double epsv2, int fric_iters, int ncomp, const int* comp_range,
                     const double* mu_comp);
/* DBC vertex list (VECTOR_STORAGE DBC of the reference, fixed targets = current x). */
int idp_set_dbc(idp_ctx* ctx, int n, const int* idx);
/* Moving Dirichlet nodes: node idx[i] is driven to targets[3 i .. 3 i + 2]
 * in every step (IMPLICIT_EULER.h:304-401, 438-474, 701-755: the CCD-limited
 * move, then an augmented-Lagrangian penalty when the move is cut short).
 * Replaces the set of idp_set_dbc; nodes may not repeat.  The reference keeps
 * these as the rows (node, x, y, z) of its DBC storage. */
int idp_set_dbc_targets(idp_ctx* ctx, int n, const int* idx, const double* targets);
int idp_set_state(idp_ctx* ctx, const double* X, const double* v);
int idp_get_state(idp_ctx* ctx, double* X, double* v);
/* Load_Velocity_X0 (Library/Utils/MESHIO.h:987-1005): v = (X_target - x) / h
 * on the device, from the context's current positions x.  X_target (3 nV,
 * host) goes through a page-locked staging buffer held by the context; the
 * call returns once it is copied there (the copy and the kernel run on the
 * context's stream). */
int idp_load_velocity(idp_ctx* ctx, const double* X_target, double h);
/* device-to-device save/restore of (x, v, kappa) for repeated solves */
int idp_snapshot_state(idp_ctx* ctx);
int idp_restore_state(idp_ctx* ctx);
int idp_surface_sizes(idp_ctx* ctx, int* nb, int* ne, int* nf);
/* The boundary primitives idp_set_mesh extracted on the device, in the order
 * of Find_Surface_TriMesh<T,false> + Find_Surface_Primitives_And_Compute_Area
 * (Library/Utils/MESHIO.h:486-545, 769-835): nodes (nb), edges (2 ne),
 * triangles (3 nf); sizes from idp_surface_sizes; any pointer may be null. */
int idp_get_surface(idp_ctx* ctx, int* bnode, int* bedge, int* btri);

/* ---- the hot path ------------------------------------------------------ */
/* Advance_One_Step_IE_Discrete_Shell<double,3,false,false,flow> for tet
 * meshes (Library/FEM/Shell/IMPLICIT_EULER.h:152-891), bound as
 * FEM.DiscreteShell.Advance_One_Step_IE_Hinge (flow = 0) and
 * Advance_One_Step_IE_Flow (flow = 1) at DISCRETE_SHELL.h:1110, 1113:
 * projected Newton with the IPC barrier, ACCD-filtered backtracking line
 * search, kappa doubling and the max-of-last-3 residual stop.
 *   b: body force, 3*nV doubles (the reference's `b`), or null for zero; the
 *      step targets x~ = x + h v + h^2 M^-1 b with the lumped mass
 *      (IMPLICIT_EULER.h:196-217).
 *   flow != 0: v is zeroed before x~ is formed, the tet elasticity term is
 *      left out (INC_POTENTIAL.h:53-73: a tet mesh has no shell flow term),
 *      and the loop also ends after any iteration that leaves no contact
 *      (IMPLICIT_EULER.h:850-854).
 * Mutates the context's x, v and kappa3[0] in place (like X, nodeAttr.v and
 * kappaVec).  Returns the Newton iteration count (>= 0) or a negative status:
 * IDP_E_CCD_ZERO when a CCD step size comes out as 0 (the reference exits,
 * IPC.h:2014-2031), IDP_E_NONPOS_DIST, IDP_E_INVALID_START. */
int idp_advance_one_step(idp_ctx* ctx, double h, double newton_tol, int with_collision, double dHat2,
                         double* kappa3, double thickness, const double* b, int flow, const idp_solver_opts* opts,
                         idp_step_stats* stats);

/* ---- stage entry points (parity tests and callers with their own loop) --- */
/* Compute_Constraint_Set (Library/FEM/IPC.h:19-740) at X (host, 3*nV);
 * *n receives the constraint count; fetch with idp_get_constraints. */
int idp_constraint_set(idp_ctx* ctx, const double* X, double dHat2, double thickness, long* n);
int idp_get_constraints(idp_ctx* ctx, int* cs, long cap);
/* Compute_Barrier (IPC.h:742-941): E = sum of barrier terms (not +=). */
int idp_barrier_energy(idp_ctx* ctx, const double* X, long n, const int* cs, double dHat2, double kappa,
                       double thickness, double* E);
/* Allocation-free variant of idp_barrier_energy.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * X_dev (3*nV doubles) and cs_dev (4*n ints) are DEVICE pointers; *E is the
 * same sum, bit for bit, as idp_barrier_energy's on the same input.  The
 * call makes no device allocation: its partial sums, parameters and block
 * table live in one fixed scratch array per GPU (about 1 MB, held until the
 * process ends), so n is limited to IDP_NOALLOC_MAX_CONSTRAINTS
 * (IDP_E_CAPACITY beyond it).  Calls that overlap on one GPU, from any
 * context, share that array and race.  Single-mesh contexts only. */
#define IDP_NOALLOC_MAX_CONSTRAINTS (1L << 24)
int idp_barrier_energy_noalloc(idp_ctx* ctx, const double* X_dev, long n, const int* cs_dev, double dHat2,
                               double kappa, double thickness, double* E);
/* Compute_Barrier_Gradient (IPC.h:943-1256): g (3*nV) is overwritten. */
int idp_barrier_gradient(idp_ctx* ctx, const double* X, long n, const int* cs, double dHat2, double kappa,
                         double thickness, double* g);
/* Compute_Barrier_Hessian (IPC.h:1258-1731) per constraint, PSD-projected
 * when project != 0 (makePD, Library/Math/UTILS.h:9-27). */
int idp_barrier_hessian_local(idp_ctx* ctx, const double* X, long n, const int* cs, double dHat2, double kappa,
                              double thickness, int project, double* H, int* vid);
/* Friction stage entry points (FRICTION.h), on the context's friction
 * coefficients (idp_set_friction).  idp_friction_basis: Compute_Friction_Basis
 * (and Compute_Friction_Coef) at X (host, 3 nV; null: the context's x) on the
 * n barrier constraints cs; Xn (host, 3 nV) is kept as the start positions of
 * the later calls; *nf = friction constraints.  idp_get_friction: their int4
 * entries, closest-point parameters (2 each), tangent bases (6 each: column
 * 0, column 1) and normal forces (cap >= nf).  The energy (mu sum N f0, with
 * epsv h the static bound: h is the step), the gradient (3 nV, summed per
 * vertex) and the local Hessians (144 per constraint: the dense (3 nv)^2
 * row-major matrix, zero padded) are evaluated at X for that set.  After
 * idp_friction_basis, idp_newton_direction includes the friction terms of
 * that set with its Xn. */
int idp_friction_basis(idp_ctx* ctx, const double* X, const double* Xn, long n, const int* cs, double dHat2,
                       double kappa, double thickness, long* nf);
int idp_get_friction(idp_ctx* ctx, int* fcs, double* cp, double* basis, double* nforce, long cap);
int idp_friction_energy(idp_ctx* ctx, const double* X, double h, double* E);
int idp_friction_gradient(idp_ctx* ctx, const double* X, double h, double* g);
int idp_friction_hessian_local(idp_ctx* ctx, const double* X, double h, int project, double* H);
/* Compute_Min_Dist2 (IPC.h:2246-2388) */
int idp_min_dist2(idp_ctx* ctx, const double* X, long n, const int* cs, double thickness, double* dist2,
                  double* min_dist2);
/* Compute_Intersection_Free_StepSize (IPC.h:1879-2244); IDP_E_CCD_ZERO when
 * the step size comes out as 0 (*alpha_out is still written).  The ACCD step
 * cap is the context's (idp_set_accd_max_steps, or the last step's
 * idp_solver_opts.accd_max_steps). */
int idp_ccd_stepsize(idp_ctx* ctx, const double* X, const double* p, double thickness, double alpha_in,
                     double* alpha_out, long* n_candidates);
/* ACCD step cap of the context's stage calls: n > 0 steps, 0 the default
 * (IDP_ACCD_MAX_STEPS_DEFAULT), < 0 none (the reference's loop). */
int idp_set_accd_max_steps(idp_ctx* ctx, int n);
/* No-hit certificate of long ACCD pairs (on by default): a pair expected to
 * take more than 1024 steps tries it after 32, any other at 1024, and each again
 * at four times the steps it has taken (at least 256),
 * is ended with no hit when a Bernstein bound shows that its distance stays
 * above the stopping gap up to its bound (the outcome its remaining steps
 * would reach; pairs that hit are always stepped).  No counterpart in the
 * reference, whose loop steps every pair; on = 0 steps them all here too. */
int idp_set_accd_certificate(idp_ctx* ctx, int on);
/* ACCD pairs stopped by the step cap and ended by the certificate in the
 * context's last CCD call (diagnostics; no counterpart in the reference) */
int idp_ccd_last_stats(idp_ctx* ctx, long* capped, long* certified);
/* Neo-Hookean tet terms with weight w = h^2 (INC_POTENTIAL.h:98-109, 228-229, 355-363) */
int idp_elastic_energy(idp_ctx* ctx, const double* X, double w, double* E, int* valid);
int idp_elastic_gradient(idp_ctx* ctx, const double* X, double w, double* g);
int idp_elastic_hessian_local(idp_ctx* ctx, const double* X, double w, int project, double* H);
/* One Newton direction at X: assemble the projected Hessian (elastic + mass +
 * barrier for the constraint set at X) into BSR, apply the DBC projection and
 * solve H sol = -g with block-Jacobi PCG.  xtilde is the inertia target. */
int idp_newton_direction(idp_ctx* ctx, const double* X, const double* xtilde, double h, double dHat2,
                         double kappa, double thickness, const idp_solver_opts* opts, double* g, double* sol,
                         double* relres, int* iters);
/* idp_newton_direction with the flow variant's terms (Advance_One_Step_IE_Flow,
 * Library/FEM/Shell/INC_POTENTIAL.h:53-73, 191-211, 323-339): no tet
 * elasticity, the shell triangles' -h/2 x^T L x term, inertia and barrier. */
int idp_newton_direction_flow(idp_ctx* ctx, const double* X, const double* xtilde, double h, double dHat2,
                              double kappa, double thickness, const idp_solver_opts* opts, double* g, double* sol,
                              double* relres, int* iters);
/* BSR of the last assembled system: sizes, then arrays (rowptr nV+1, col nnzb, val 9*nnzb) */
int idp_get_bsr(idp_ctx* ctx, long* nnzb, int* rowptr, int* col, double* val, long cap);

/* ---- batched independent solves ---------------------------------------- */
/* No counterpart in the reference, which advances one mesh per call
 * (Advance_One_Step_IE_Discrete_Shell, IMPLICIT_EULER.h:152-891).  B meshes
 * ("instances") are held in one context: arrays are concatenated over
 * instances (nV[i] vertices, nT[i] tets each; tets in instance-local vertex
 * ids).  Every stage runs over all instances at once, and each instance
 * follows its own Newton / line-search / kappa sequence, with every reduction
 * taken over the instance alone, so an instance's result is the same whether
 * it is solved alone or with others.  idp_set_state / idp_get_state /
 * idp_set_dbc address the concatenated vertex arrays. */
int idp_set_mesh_batch(idp_ctx* ctx, int B, const int* nV, const int* nT, const double* X_rest, const int* tets,
                       const double* IB, const double* vol, const double* lambda, const double* mu,
                       const double* mass);
int idp_batch_size(idp_ctx* ctx, int* B);
/* One implicit-Euler step of every instance: newton_tol[B], dHat2[B],
 * kappa3[3 B] (in/out), newton_iters[B] (out), stats[B] (out, may be null). */
int idp_advance_batch(idp_ctx* ctx, double h, const double* newton_tol, int with_collision, const double* dHat2,
                      double* kappa3, double thickness, const idp_solver_opts* opts, int* newton_iters,
                      idp_step_stats* stats);

/* ---- continuous batching (serving) -------------------------------------- */
/* No counterpart in the reference.  The instances of a batch mesh are slots;
 * a request is started in a free slot while other slots are mid-solve, and
 * every idp_serve_iterate call advances all busy slots by one Newton
 * iteration.  configure: time step and solver options for all slots. */
int idp_serve_configure(idp_ctx* ctx, double h, int with_collision, double thickness, const idp_solver_opts* opts);
/* start n requests in the given free slots: X and v concatenated over the
 * listed slots (3 nV[slot] each), newton_tol[n], dHat2[n], kappa3[3 n] */
int idp_serve_start(idp_ctx* ctx, int n, const int* slots, const double* X, const double* v,
                    const double* newton_tol, const double* dHat2, const double* kappa3);
/* one Newton iteration of every busy slot; finished[B] (may be null) flags
 * the slots whose solve ended in this call; *n_active = busy slots left.  A
 * slot whose CCD pair ran past the step budget (idp_solver_opts.accd_budget)
 * continues that pair in the next call before its line search. */
int idp_serve_iterate(idp_ctx* ctx, int* n_active, int* finished);
/* result of a slot (positions, velocities, kappa3, stats; any may be null) */
int idp_serve_result(idp_ctx* ctx, int slot, double* X, double* v, double* kappa3, idp_step_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
