// Oracle (test infrastructure only): C entry points of the CPU restatement of
// IDP's Newton-with-barrier path.  Loaded by tests/, __graft_entry__.smoke()
// and the cpu_baseline leg of bench.py as the checker; the product never
// links it.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_E_NONPOS_DIST = -1, ORC_E_DEGENERATE = -2, ORC_E_INVALID_START = -3 };

typedef struct orc_stats {
    int n_newton;
    int line_search_trials;
    int n_constraint_sets;
    long linear_iters;
    long n_constraints;
    double last_alpha, first_ccd_alpha, last_res, last_relres, min_dist2, E;
    double t_constraint_set, t_ccd, t_solve, t_hessian, t_gradient, t_energy;
    double dbc_alpha, dbc_stiff; /* moving DBC: CCD-limited move at the start, penalty stiffness at the end */
    long n_friction;             /* friction constraints at the last basis update */
    int friction_updates;        /* lagged basis updates after the first */
} orc_stats;

int orc_surface(int nV, const double* X, int nT, const int* tets, int* nb, int* ne, int* nf, int* bnode,
                int* bedge, int* btri);
long orc_constraint_set(int nV, const double* X, const double* X0, const unsigned char* dbc, int nb,
                        const int* bnode, int ne, const int* bedge, int nf, const int* btri, double dHat2,
                        double thickness, int use_hash, int* out, long cap);
int orc_barrier_energy(const double* X, const double* X0, long n, const int* cs, double dHat2, double kappa,
                       double thickness, double* E_out, double* per_pair);
int orc_barrier_gradient(const double* X, const double* X0, long n, const int* cs, double dHat2, double kappa,
                         double thickness, int nV, double* g);
int orc_barrier_hessian_local(const double* X, const double* X0, long n, const int* cs, double dHat2,
                              double kappa, double thickness, int projectSPD, double* H_out, int* vid_out);
int orc_min_dist2(const double* X, long n, const int* cs, double thickness, double* dist2, double* minDist2);
int orc_ccd_stepsize(int nV, const double* X, const double* p, const unsigned char* dbc, int nb, const int* bnode,
                     int ne, const int* bedge, int nf, const int* btri, double thickness, double alpha_in,
                     double* alpha_out, long* n_cand);
int orc_geom_probe(int kind, const double* x, int* type, double* d2);
int orc_fem_init(int nV, const double* X, int nT, const int* tets, double rho, double E, double nu, double* vol,
                 double* IB, double* lam, double* mu, double* mass);
int orc_elastic_energy(const double* X, int nT, const int* tets, const double* IB, const double* vol,
                       const double* lam, const double* mu, double w, double* E);
void orc_elastic_gradient(const double* X, int nT, const int* tets, const double* IB, const double* vol,
                          const double* lam, const double* mu, double w, double* g);
void orc_elastic_hessian_local(const double* X, int nT, const int* tets, const double* IB, const double* vol,
                               const double* lam, const double* mu, double w, int projectSPD, double* H);
long orc_newton_system(int nV, const double* X, const double* Xtilde, const double* X0, const double* mass, int nT,
                       const int* tets, const double* IB, const double* vol, const double* lam, const double* mu,
                       int nb, const int* bnode, int ne, const int* bedge, int nf, const int* btri, int nDBC,
                       const int* dbc_idx, double h, double dHat2, double kappa, double thickness, int* rowptr,
                       int* col, double* val, double* g, long cap);
int orc_advance_one_step(int nV, double* X, const double* X0, double* vel, const double* mass, int nT,
                         const int* tets, const double* IB, const double* vol, const double* lam,
                         const double* mu, int nb, const int* bnode, int ne, const int* bedge, int nf,
                         const int* btri, int nDBC, const int* dbc_idx, const double* dbc_target, double h,
                         double NewtonTol, int withCollision, double dHat2, double* kappa, double thickness,
                         double lin_rtol, int max_newton, const double* body_force, int flow, orc_stats* st);
/* ACCD step cap per pair and call for orc_ccd_stepsize and the Newton loop:
 * <= 0 none (the reference's loop, CCD.h:308-327, 367-392; the default); a
 * pair that reaches it is taken as a hit at the time of impact it reached. */
void orc_set_accd_max_steps(int n);

#ifdef __cplusplus
}
#endif
