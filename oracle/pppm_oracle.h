/*
 * pppm_oracle.h -- CPU restatement of the reference PPPM vortex-velocity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in the
 * product path (the HIP library) links, loads or calls it.
 *
 * Arithmetic follows the reference: FP32 where the reference computes in FP32
 * (hash keys, origin-shifted positions, near-field coordinate differences),
 * FP64 everywhere else.  Each function cites the reference file:line it
 * restates (paths relative to the reference repository root).
 */
#ifndef PPPM_ORACLE_H
#define PPPM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_KERNEL_EXPCUBIC = 0, ORC_KERNEL_GAUSSIAN = 1 };
enum { ORC_MG_VCYCLE = 0, ORC_MG_REF_FMG = 1 };
enum { ORC_PROLONG_CONSTANT = 0, ORC_PROLONG_TRILINEAR = 1 };

typedef struct {
    uint32_t g;          /* cells per axis (cubic grid, cubic hash)           */
    float    hf;         /* hash cell size as the hash stores it (float)      */
    float    originf[3]; /* hash origin (float3)                              */
    double   origin[3];  /* origin used by P2M / M2P / near field (double4)   */
    int      K;          /* near-field range in cells                          */
    int      kernel;     /* ORC_KERNEL_*                                       */
    double   core_a;     /* smoothing core radius (absolute units)             */
    int      mg_schedule;/* ORC_MG_*                                           */
    int      nu1, nu2;   /* smoothing iterations (one = red + black)          */
    double   mg_omega;   /* SOR factor (1 = Gauss-Seidel)                       */
    double   mg_tol;     /* relative residual target (VCYCLE)                   */
    int      mg_max_cycles;
    int      mg_prolong; /* ORC_PROLONG_* */
} orc_params;

typedef struct {
    int    mg_cycles[3];
    double mg_relres[3];
    uint32_t strays;     /* particles whose deposit cell differs from hash cell */
    double t_hash, t_p2m, t_mg, t_far, t_eval; /* seconds */
} orc_diag;

/* ---- particle_hash --------------------------------------------------- */
uint32_t orc_cell_key(float px, float py, float pz, uint32_t g, float hf,
                      float ox, float oy, float oz);
void orc_hash_keys(const float *pos4, uint32_t n, uint32_t g, float hf,
                   const float originf[3], uint32_t *key, uint32_t *idx);
void orc_sort_pairs(uint32_t *key, uint32_t *idx, uint32_t n);
void orc_cell_tables(const uint32_t *key_sorted, uint32_t n, uint32_t ncells,
                     uint32_t *start, uint32_t *end);
void orc_reorder(const uint32_t *idx, const void *in, uint32_t stride,
                 void *out, uint32_t precision, uint32_t n);
void orc_deorder(const uint32_t *idx, const void *in, uint32_t stride,
                 void *out, uint32_t precision, uint32_t n);

/* ---- P2M / RHS ------------------------------------------------------- */
void orc_p2m(const uint32_t *start, const uint32_t *end, const float *pos4_sorted,
             const double *mass_sorted, double cell_h, double *grid,
             uint32_t g, uint32_t n, const double origin[3]);
void orc_rhs(double *grid, double h_sqr, double G, uint64_t ncells);

/* ---- multigrid ------------------------------------------------------- */
int  orc_mg_levels(int nx, int ny, int nz, int dims_out[][3], int max_levels);
void orc_mg_smooth(double *x, const double *b, int nx, int ny, int nz,
                   int iters, double omega);
void orc_mg_residual(const double *x, const double *b, double *r,
                     int nx, int ny, int nz);
void orc_mg_restrict(double *bc, const double *rf, int cx, int cy, int cz);
void orc_mg_prolong(double *xf, const double *xc, int fx, int fy, int fz);
void orc_mg_prolong_tri(double *xf, const double *xc, int fx, int fy, int fz);
void orc_mg_coarse_exact(double *x, const double *b, int nx, int ny, int nz);
double orc_mg_beta(int level);
double orc_mg_norm2(const double *v, uint64_t n);
double orc_mg_relres(const double *x, const double *b, int nx, int ny, int nz);
int  orc_mg_solve(double *x, const double *b, int nx, int ny, int nz,
                  int schedule, int nu1, int nu2, double omega, double tol,
                  int max_cycles, int prolong, double *relres_out);

/* ---- curl, far field, M2P, near field ------------------------------- */
void orc_curl(const double *Fx, const double *Fy, const double *Fz,
              double *u, double *v, double *w, int nx, int ny, int nz, double h);
void orc_far_dir_table(int K, double h, double *dir /* (2K+1)^3 x 3 */);
void orc_far_field(const double *wx, const double *wy, const double *wz,
                   const double *Ux, const double *Uy, const double *Uz,
                   double *fx, double *fy, double *fz, double h, int g, int K);
void orc_m2p(const float *pos4_sorted, const double *fx, const double *fy,
             const double *fz, double *pu, double *pv, double *pw,
             double cell_h, int g, uint32_t m, const double origin[3]);
void orc_near_field(const uint32_t *start, const uint32_t *end,
                    const float *evala4, const float *evalb4, int segment,
                    const float *vort4, const double *wx, const double *wy,
                    const double *wz, double *u, double *v, double *w,
                    double cell_h, int g, int K, uint32_t m, uint32_t n,
                    const double origin[3], int kernel, double core_a);
double orc_kernel_coeff(double r, int kernel, double a);

/* ---- whole path ------------------------------------------------------ */
int orc_pppm_eval(const orc_params *p, uint32_t n, const float *vort_pos4,
                  const double *wx, const double *wy, const double *wz,
                  uint32_t m, const float *eval_pos4, const float *other_end4,
                  double *ux, double *uy, double *uz, orc_diag *diag);
void orc_direct_sum(uint32_t n, const float *vort_pos4, const double *wx,
                    const double *wy, const double *wz, uint32_t m,
                    const float *eval_pos4, int kernel, double core_a,
                    double *ux, double *uy, double *uz);
void orc_reference_domain(const float *pos4, uint32_t n, uint32_t g,
                          double origin_out[3], double *L_out);
/* ---- time step of the vortex-segment method -------------------------- */
void orc_segments_from_vortices(const float *cent4, float *enda4, float *endb4, const double *wx,
                                const double *wy, const double *wz, double *kappa, double h, uint32_t n);
void orc_move_points(float *pos4, double dt, const double *u4, uint32_t n);
void orc_segment_centres(const float *posc4, const float *posb4, float *posa4, uint32_t n);
void orc_vortices_from_segments(const float *posa4, const float *posb4, const double *wx, const double *wy,
                                const double *wz, double *dwx, double *dwy, double *dwz, const double *kappa,
                                uint32_t n);
void orc_gaussian_smooth(const uint32_t *start, const uint32_t *end, const float *pos4,
                         const float *pos4_sorted, double nu, double *ox, double *oy, double *oz,
                         const double *ix, const double *iy, const double *iz, double h, int g, uint32_t n,
                         const double origin[3]);

/* ---- baroclinic source (addBaroclinicForce) ------------------------- */
void orc_density_domain(const float *pos4, uint32_t n, double cell_target, uint32_t max_dim, double origin[3],
                        double *h_out, uint32_t *dim_out);
void orc_assign_density(const uint32_t *start, const uint32_t *end, const float *pos4_sorted, const float *rho,
                        double cell_h, uint32_t dim, uint32_t n, const double origin[3], float *grid);
void orc_grad_density(const float *phi, uint32_t dim, double cell_h, float *grad);
void orc_baroclinic(const float *vpos4, double *wx, double *wy, double *wz, const float *grad, double dt_beta,
                    uint32_t n, uint32_t dim, double inv_h, const double origin[3], const double g[3]);

/* ---- potential-flow gradient path (PotentialFieldSolver) ------------- */
void orc_pot_gradient(const double *phi, double *dphi, double h, double direction, int n);
void orc_pot_dirichlet(double *rhs, const double center[3], double monopole, const double corner[3], double h, int n);
void orc_pot_far_field(const double *rho, double *phi, double *gforce, double *far, double h, double G,
                       double direction, int n, int K);
void orc_pot_interp(const float *pos4, const double *far, double *pf, double h, int n, uint32_t m,
                    const double origin[3]);
void orc_pot_near(const uint32_t *start, const uint32_t *end, const float *pe4, const float *pm4, const double *mass,
                  double *pf, double h, double G, int n, int K, uint32_t M, uint32_t N, const double origin[3]);
void orc_pot_outside(const float *pos4, const float lo[3], const float hi[3], double *pf, uint32_t M);

void orc_pot_scalar_far(const double *rho, double *phi, double h, double G, int n, int K);
void orc_pot_interp_scalar(const float *pos4, const double *far, double *pf, double h, int n, uint32_t m,
                           const double origin[3]);
void orc_pot_near_scalar(const uint32_t *start, const uint32_t *end, const float *pe4, const float *pm4,
                         const double *mass, double *pf, double h, double G, int n, int K, uint32_t M, uint32_t N,
                         const double origin[3]);

int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
