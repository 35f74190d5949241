/*
 * pppm_oracle.c -- CPU restatement of the reference PPPM vortex-velocity path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py.  The HIP product library never links it.
 *
 * Build: gcc -O2 -fopenmp -fPIC -shared -ffp-contract=off (see oracle/Makefile).
 * FP contraction is off so that results are reproducible run to run.
 *
 * Deliberate, documented departures from the reference (each one is a bug or a
 * race in the reference; effects are stated where the function is):
 *   - coarse-grid solve: exact separable sine-transform solve instead of the
 *     reference's racy 100-iteration in-place relaxation
 *     (multigrid/multigrid/MultiGrid_kernel.cu:422-498);
 *   - near-field small-radius branch: -expm1 form instead of the Taylor branch
 *     whose 1/6 and 1/24 are integer zero (GPU_kernel.cu:2399-2401);
 *   - empty cells get cellEnd = 0xFFFFFFFF (the reference never clears cellEnd,
 *     particle_hash/particle_hash/SpaceHashFunctional.cu:82);
 *   - prolongation clamps the coarse index for odd fine dimensions (the
 *     reference reads past the coarse array, MultiGrid_kernel.cu:929-930).
 */
#include "pppm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_FOUR_PI 12.5663706144          /* GPU_kernel.cu:14, 2395 */
#define REF_DIAG_PSI 0.250580605913965     /* GPU_kernel.cu:1989 */
#define EMPTY 0xFFFFFFFFu

static double wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ===================================================================== */
/* particle_hash                                                          */
/* ===================================================================== */

/* float -> int as the GPU converts it (round toward zero, saturating, NaN -> 0) */
static int f2i_sat(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int)f;
}

/* particle_hash/particle_hash/particle_hash_kernel.cu:32-48 (calcGridPos,
 * calcGridHash): axis = floor((p - o) / h) in FP32 with a correctly rounded
 * division; int % uint wraps negative or too-large cells; __umul24 products are
 * exact for g <= 4096. */
uint32_t orc_cell_key(float px, float py, float pz, uint32_t g, float hf,
                      float ox, float oy, float oz)
{
    volatile float dx = px - ox, dy = py - oy, dz = pz - oz;
    volatile float qx = dx / hf, qy = dy / hf, qz = dz / hf;
    int ix = f2i_sat(floorf(qx)), iy = f2i_sat(floorf(qy)), iz = f2i_sat(floorf(qz));
    uint32_t ux = (uint32_t)ix % g, uy = (uint32_t)iy % g, uz = (uint32_t)iz % g;
    return (uz * g + uy) * g + ux;
}

/* particle_hash_kernel.cu:51-70 (calcHashD): key per particle, index = i */
void orc_hash_keys(const float *pos4, uint32_t n, uint32_t g, float hf,
                   const float originf[3], uint32_t *key, uint32_t *idx)
{
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < (int64_t)n; i++) {
        key[i] = orc_cell_key(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], g, hf,
                              originf[0], originf[1], originf[2]);
        idx[i] = (uint32_t)i;
    }
}

/* SpaceHashFunctional.cu:97-102 (sortParticles).  The reference calls
 * thrust::sort_by_key; the build's contract is the STABLE order by
 * (key, original index), realised here as an LSD radix sort (stable). */
void orc_sort_pairs(uint32_t *key, uint32_t *idx, uint32_t n)
{
    uint32_t *k2 = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *v2 = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *ks = key, *vs = idx, *kd = k2, *vd = v2;
    int shift;
    for (shift = 0; shift < 32; shift += 11) {
        uint32_t cnt[2048];
        uint32_t i, sum = 0;
        memset(cnt, 0, sizeof(cnt));
        for (i = 0; i < n; i++) cnt[(ks[i] >> shift) & 2047u]++;
        for (i = 0; i < 2048; i++) { uint32_t c = cnt[i]; cnt[i] = sum; sum += c; }
        for (i = 0; i < n; i++) {
            uint32_t d = (ks[i] >> shift) & 2047u;
            uint32_t o = cnt[d]++;
            kd[o] = ks[i];
            vd[o] = vs[i];
        }
        { uint32_t *t = ks; ks = kd; kd = t; t = vs; vs = vd; vd = t; }
    }
    if (ks != key) {
        memcpy(key, ks, sizeof(uint32_t) * n);
        memcpy(idx, vs, sizeof(uint32_t) * n);
    }
    free(k2);
    free(v2);
}

/* SpaceHashFunctional.cu:71-94 + particle_hash_kernel.cu:74-126
 * (FindCellStart / FindCellStartD): start[k] = first sorted index with key k,
 * end[k] = one past the last; empty cells hold 0xFFFFFFFF in both. */
void orc_cell_tables(const uint32_t *key_sorted, uint32_t n, uint32_t ncells,
                     uint32_t *start, uint32_t *end)
{
    uint32_t i;
    for (i = 0; i < ncells; i++) { start[i] = EMPTY; end[i] = EMPTY; }
    for (i = 0; i < n; i++) {
        uint32_t k = key_sorted[i];
        if (i == 0 || k != key_sorted[i - 1]) {
            if (k < ncells) start[k] = i;
            if (i > 0 && key_sorted[i - 1] < ncells) end[key_sorted[i - 1]] = i;
        }
        if (i == n - 1 && k < ncells) end[k] = i + 1;
    }
}

/* particle_hash_kernel.cu:128-154 (applyReorder): out[i*s+c] = in[idx[i]*s+c],
 * precision 1 = float, 2 = double */
void orc_reorder(const uint32_t *idx, const void *in, uint32_t stride,
                 void *out, uint32_t precision, uint32_t n)
{
    uint32_t i, c;
    for (i = 0; i < n; i++)
        for (c = 0; c < stride; c++) {
            if (precision == 1)
                ((float *)out)[(size_t)i * stride + c] = ((const float *)in)[(size_t)idx[i] * stride + c];
            else
                ((double *)out)[(size_t)i * stride + c] = ((const double *)in)[(size_t)idx[i] * stride + c];
        }
}

/* particle_hash_kernel.cu:158-184 (removeReorder): out[idx[i]*s+c] = in[i*s+c] */
void orc_deorder(const uint32_t *idx, const void *in, uint32_t stride,
                 void *out, uint32_t precision, uint32_t n)
{
    uint32_t i, c;
    for (i = 0; i < n; i++)
        for (c = 0; c < stride; c++) {
            if (precision == 1)
                ((float *)out)[(size_t)idx[i] * stride + c] = ((const float *)in)[(size_t)i * stride + c];
            else
                ((double *)out)[(size_t)idx[i] * stride + c] = ((const double *)in)[(size_t)i * stride + c];
        }
}

/* ===================================================================== */
/* P2M                                                                    */
/* ===================================================================== */

/* GPU_kernel.cu:31-40 (shape_function_mass): box weight on [-1/2, 1/2) */
static double box_w(double r) { return (r < 0.5 && r >= -0.5) ? 1.0 : 0.0; }

/* GPU_kernel.cu:218-271 (ParticleToMesh_kernel): each cell gathers over the 5^3
 * hash cells around it; the particle position is shifted by the double origin
 * and stored back to float (ppos.x -= origin.x), then compared with the cell
 * centre in double.  Result times h^-3. */
void orc_p2m(const uint32_t *start, const uint32_t *end, const float *pos4_sorted,
             const double *mass_sorted, double cell_h, double *grid,
             uint32_t g, uint32_t n, const double origin[3])
{
    const double invh = 1.0 / cell_h;
    const int G = (int)g;
    int64_t c;
#pragma omp parallel for schedule(static)
    for (c = 0; c < (int64_t)G * G * G; c++) {
        int i = (int)(c % G), j = (int)((c / G) % G), k = (int)(c / ((int64_t)G * G));
        double sum = 0.0;
        float cx = (float)((double)i + 0.5), cy = (float)((double)j + 0.5), cz = (float)((double)k + 0.5);
        int kk, jj, ii;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    uint32_t cell, p;
                    if (kk < 0 || kk >= G || jj < 0 || jj >= G || ii < 0 || ii >= G) continue;
                    cell = (uint32_t)(((int64_t)kk * G + jj) * G + ii);
                    for (p = start[cell]; p < end[cell]; p++) {
                        float px, py, pz;
                        double w;
                        if (p >= n) continue;
                        px = (float)((double)pos4_sorted[4 * (size_t)p] - origin[0]);
                        py = (float)((double)pos4_sorted[4 * (size_t)p + 1] - origin[1]);
                        pz = (float)((double)pos4_sorted[4 * (size_t)p + 2] - origin[2]);
                        w = box_w((double)px * invh - (double)cx) * box_w((double)py * invh - (double)cy) *
                            box_w((double)pz * invh - (double)cz);
                        if (w != 0.0) sum += w * mass_sorted[p];
                    }
                }
        grid[c] = invh * invh * invh * sum;
    }
}

/* GPU_kernel.cu:524-543 (compute_rhs_kernel): rhs = v * h^2 * G */
void orc_rhs(double *grid, double h_sqr, double G, uint64_t ncells)
{
    uint64_t i;
    for (i = 0; i < ncells; i++) grid[i] = grid[i] * h_sqr * G;
}

/* ===================================================================== */
/* multigrid (cell-centred 7-point Poisson, zero ghost values)            */
/* ===================================================================== */

#define IDX(i, j, k) (((size_t)(k) * ny + (j)) * nx + (i))

/* multigrid/multigrid/MultiGrid.h:115-127 (m_ComputeLevels) and 93-98 */
int orc_mg_levels(int nx, int ny, int nz, int dims_out[][3], int max_levels)
{
    int L = 0;
    int x = nx, y = ny, z = nz;
    dims_out[0][0] = x; dims_out[0][1] = y; dims_out[0][2] = z;
    while ((long)x * y * z > 512 && L + 1 < max_levels) {
        x = x / 2; if (x == 0) x = 1;
        y = y / 2; if (y == 0) y = 1;
        z = z / 2; if (z == 0) z = 1;
        L++;
        dims_out[L][0] = x; dims_out[L][1] = y; dims_out[L][2] = z;
    }
    return L + 1;
}

/* MultiGrid_kernel.cu:722-820 (red/black_sweep_GS_periodic_kerneld_type2) and
 * Multigrid.cu:202-239: one iteration = the (i+j+k)-odd colour, then the even
 * colour; x = (1-w) x + w (sum_nb - b) / 6 with out-of-grid neighbours 0.
 * beta != 0 (coarse levels of the tolerance-driven V-cycle only): an
 * out-of-grid neighbour is beta * x (ghost on the same wall as the finest
 * level), so the diagonal becomes 6 - beta * (number of ghosts). */
static void mg_smooth_beta(double *x, const double *b, int nx, int ny, int nz,
                           int iters, double omega, double beta)
{
    const double invdiag = 1.0 / 6.0;
    const double one_m_w = 1.0 - omega;
    int it, color;
    for (it = 0; it < iters; it++)
        for (color = 1; color >= 0; color--) {
            int64_t kj;
#pragma omp parallel for schedule(static)
            for (kj = 0; kj < (int64_t)nz * ny; kj++) {
                int k = (int)(kj / ny), j = (int)(kj % ny);
                int i = ((j + k + color) & 1);
                for (; i < nx; i += 2) {
                    size_t t = IDX(i, j, k);
                    int nout = (i == 0) + (i == nx - 1) + (j == 0) + (j == ny - 1) + (k == 0) + (k == nz - 1);
                    double lv = (i == 0) ? 0.0 : x[t - 1];
                    double rv = (i == nx - 1) ? 0.0 : x[t + 1];
                    double uv = (j == ny - 1) ? 0.0 : x[t + nx];
                    double dv = (j == 0) ? 0.0 : x[t - nx];
                    double fv = (k == 0) ? 0.0 : x[t - (size_t)nx * ny];
                    double bv = (k == nz - 1) ? 0.0 : x[t + (size_t)nx * ny];
                    double r = lv + rv;
                    r += uv + dv;
                    r += fv + bv;
                    r -= b[t];
                    if (beta == 0.0 || nout == 0)
                        x[t] = one_m_w * x[t] + omega * r * invdiag;
                    else
                        x[t] = one_m_w * x[t] + omega * r / (6.0 - beta * nout);
                }
            }
        }
}

void orc_mg_smooth(double *x, const double *b, int nx, int ny, int nz,
                   int iters, double omega)
{
    mg_smooth_beta(x, b, nx, ny, nz, iters, omega, 0.0);
}

/* MultiGrid_kernel.cu:821-871 (compute_residual_periodic_kernel2):
 * r = b - (sum_nb - 6 x); beta as in mg_smooth_beta. */
static void mg_residual_beta(const double *x, const double *b, double *r,
                             int nx, int ny, int nz, double beta)
{
    int64_t t0;
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)nx * ny * nz; t0++) {
        int i = (int)(t0 % nx), j = (int)((t0 / nx) % ny), k = (int)(t0 / ((int64_t)nx * ny));
        size_t t = (size_t)t0;
        int nout = (i == 0) + (i == nx - 1) + (j == 0) + (j == ny - 1) + (k == 0) + (k == nz - 1);
        double lv = (i == 0) ? 0.0 : x[t - 1];
        double rv = (i == nx - 1) ? 0.0 : x[t + 1];
        double uv = (j == ny - 1) ? 0.0 : x[t + nx];
        double dv = (j == 0) ? 0.0 : x[t - nx];
        double fv = (k == 0) ? 0.0 : x[t - (size_t)nx * ny];
        double bv = (k == nz - 1) ? 0.0 : x[t + (size_t)nx * ny];
        double s2 = lv + rv;
        s2 += uv + dv;
        s2 += fv + bv;
        s2 -= 6 * x[t];
        if (beta != 0.0 && nout) s2 += beta * nout * x[t];
        r[t] = b[t] - s2;
    }
}

void orc_mg_residual(const double *x, const double *b, double *r,
                     int nx, int ny, int nz)
{
    mg_residual_beta(x, b, r, nx, ny, nz, 0.0);
}

/* Ghost factor of level l of the tolerance-driven V-cycle: the finest level
 * (l = 0) has ghost value 0, i.e. its wall sits half a fine cell outside the
 * last face; a level with spacing H = 2^l h keeps that wall by extrapolating
 * linearly through its last cell centre, ghost = beta_l x0 with
 * beta_l = -(1 - 2^-l) / (1 + 2^-l). */
double orc_mg_beta(int level)
{
    double s = ldexp(1.0, -level);
    return -(1.0 - s) / (1.0 + s);
}

/* MultiGrid_kernel.cu:873-910 (restriction_periodic_kernel2):
 * b_c = 0.5 * sum of the 8 children (fine dims are 2x the coarse dims) */
void orc_mg_restrict(double *bc, const double *rf, int cx, int cy, int cz)
{
    const int nx = 2 * cx, ny = 2 * cy; /* fine strides as the reference uses them */
    int64_t t0;
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)cx * cy * cz; t0++) {
        int i = (int)(t0 % cx), j = (int)((t0 / cx) % cy), k = (int)(t0 / ((int64_t)cx * cy));
        size_t a = IDX(2 * i, 2 * j, 2 * k);
        size_t ay = a + nx, az = a + (size_t)nx * ny, ayz = az + nx;
        double r = 0;
        r += rf[a];
        r += rf[a + 1];
        r += rf[ay];
        r += rf[ay + 1];
        r += rf[az];
        r += rf[az + 1];
        r += rf[ayz];
        r += rf[ayz + 1];
        bc[t0] = r * 0.5;
    }
}

/* Restriction with general (possibly odd) fine dims: children beyond the
 * coarse range (the last odd fine cell) are dropped, as in the reference. */
static void mg_restrict_dims(double *bc, const double *rf, int cx, int cy, int cz,
                             int fx, int fy)
{
    const int nx = fx, ny = fy;
    int64_t t0;
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)cx * cy * cz; t0++) {
        int i = (int)(t0 % cx), j = (int)((t0 / cx) % cy), k = (int)(t0 / ((int64_t)cx * cy));
        size_t a = IDX(2 * i, 2 * j, 2 * k);
        size_t ay = a + nx, az = a + (size_t)nx * ny, ayz = az + nx;
        double r = 0;
        r += rf[a];
        r += rf[a + 1];
        r += rf[ay];
        r += rf[ay + 1];
        r += rf[az];
        r += rf[az + 1];
        r += rf[ayz];
        r += rf[ayz + 1];
        bc[t0] = r * 0.5;
    }
}

/* MultiGrid_kernel.cu:912-934 (prolongation_periodic_kernel2):
 * x_f += x_c[i/2, j/2, k/2] (coarse index clamped for odd fine dims) */
void orc_mg_prolong(double *xf, const double *xc, int fx, int fy, int fz)
{
    const int cx = fx / 2 > 0 ? fx / 2 : 1, cy = fy / 2 > 0 ? fy / 2 : 1, cz = fz / 2 > 0 ? fz / 2 : 1;
    int64_t t0;
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)fx * fy * fz; t0++) {
        int i = (int)(t0 % fx), j = (int)((t0 / fx) % fy), k = (int)(t0 / ((int64_t)fx * fy));
        int ic = i / 2, jc = j / 2, kc = k / 2;
        if (ic > cx - 1) ic = cx - 1;
        if (jc > cy - 1) jc = cy - 1;
        if (kc > cz - 1) kc = cz - 1;
        xf[t0] += xc[((size_t)kc * cy + jc) * cx + ic];
    }
}

/* Trilinear cell-centred prolongation (no reference counterpart; the
 * reference pair above does not give a convergent V-cycle beyond 64^3):
 * x_f += sum over the 2x2x2 nearest coarse cells with 1-D weights 3/4 (parent)
 * and 1/4 (neighbour on the fine cell's side); coarse cells outside the grid
 * are ghost zeros. */
static void mg_prolong_tri_beta(double *xf, const double *xc, int fx, int fy, int fz, double betac)
{
    const int cx = fx / 2 > 0 ? fx / 2 : 1, cy = fy / 2 > 0 ? fy / 2 : 1, cz = fz / 2 > 0 ? fz / 2 : 1;
    int64_t t0;
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)fx * fy * fz; t0++) {
        int i = (int)(t0 % fx), j = (int)((t0 / fx) % fy), k = (int)(t0 / ((int64_t)fx * fy));
        int ci[2], cj[2], ck[2];
        double wi[2] = {0.75, 0.25}, wj[2] = {0.75, 0.25}, wk[2] = {0.75, 0.25};
        double s = 0;
        int a, b2, c2;
        ci[0] = i / 2; ci[1] = (i & 1) ? ci[0] + 1 : ci[0] - 1;
        cj[0] = j / 2; cj[1] = (j & 1) ? cj[0] + 1 : cj[0] - 1;
        ck[0] = k / 2; ck[1] = (k & 1) ? ck[0] + 1 : ck[0] - 1;
        if (ci[0] > cx - 1) { ci[0] = cx - 1; }
        if (cj[0] > cy - 1) { cj[0] = cy - 1; }
        if (ck[0] > cz - 1) { ck[0] = cz - 1; }
        if (fx == 1) { wi[0] = 1.0; wi[1] = 0.0; }
        if (fy == 1) { wj[0] = 1.0; wj[1] = 0.0; }
        if (fz == 1) { wk[0] = 1.0; wk[1] = 0.0; }
        /* a coarse neighbour outside the grid is the ghost betac * (its
         * in-grid mirror, the parent along that axis) */
        for (c2 = 0; c2 < 2; c2++) {
            double sy = 0, fk = 1.0;
            int kk = ck[c2];
            if (wk[c2] == 0.0) continue;
            if (kk < 0 || kk > cz - 1) { kk = ck[0]; fk = betac; }
            for (b2 = 0; b2 < 2; b2++) {
                double sx = 0, fj = 1.0;
                int jj = cj[b2];
                if (wj[b2] == 0.0) continue;
                if (jj < 0 || jj > cy - 1) { jj = cj[0]; fj = betac; }
                for (a = 0; a < 2; a++) {
                    double fi = 1.0;
                    int ii = ci[a];
                    if (wi[a] == 0.0) continue;
                    if (ii < 0 || ii > cx - 1) { ii = ci[0]; fi = betac; }
                    sx += wi[a] * fi * xc[((size_t)kk * cy + jj) * cx + ii];
                }
                sy += wj[b2] * fj * sx;
            }
            s += wk[c2] * fk * sy;
        }
        xf[t0] += s;
    }
}

void orc_mg_prolong_tri(double *xf, const double *xc, int fx, int fy, int fz)
{
    mg_prolong_tri_beta(xf, xc, fx, fy, fz, 0.0);
}

/* Orthonormal sine basis of the 1-D zero-ghost second difference
 * tridiag(1,-2,1) of size n: S[k][i] = sqrt(2/(n+1)) sin(pi (k+1)(i+1)/(n+1)),
 * eigenvalue -4 sin^2(pi (k+1) / (2 (n+1))). */
static void dst_basis(int n, double *S, double *lam)
{
    int k, i;
    const double pi = 3.14159265358979323846;
    for (k = 0; k < n; k++) {
        double s = sin(pi * (k + 1) / (2.0 * (n + 1)));
        lam[k] = -4.0 * s * s;
        for (i = 0; i < n; i++)
            S[k * n + i] = sqrt(2.0 / (n + 1)) * sin(pi * (double)(k + 1) * (i + 1) / (n + 1));
    }
}

static void apply_axis(const double *S, int n, double *v, int nx, int ny, int nz, int axis)
{
    double tmp[512];
    int a, b2, k, i;
    int na = axis == 0 ? ny : nx, nb = axis == 2 ? ny : nz;
    for (b2 = 0; b2 < nb; b2++)
        for (a = 0; a < na; a++) {
            size_t base, step;
            if (axis == 0) { base = IDX(0, a, b2); step = 1; }
            else if (axis == 1) { base = IDX(a, 0, b2); step = (size_t)nx; }
            else { base = IDX(a, b2, 0); step = (size_t)nx * ny; }
            for (k = 0; k < n; k++) {
                double s = 0;
                for (i = 0; i < n; i++) s += S[k * n + i] * v[base + i * step];
                tmp[k] = s;
            }
            for (k = 0; k < n; k++) v[base + k * step] = tmp[k];
        }
}

/* Eigen-decomposition of the symmetric n x n matrix A (destroyed) by cyclic
 * Jacobi rotations: on return E[k*n+i] is component i of eigenvector k and
 * lam[k] its eigenvalue. */
static void eig_sym_jacobi(int n, double *A, double *E, double *lam)
{
    int i, j, p, q, sweep;
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) E[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (p = 0; p < n; p++)
            for (q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-300) break;
        for (p = 0; p < n; p++)
            for (q = p + 1; q < n; q++) {
                double apq = A[p * n + q], app, aqq, theta, t, c, sn;
                if (apq == 0.0) continue;
                app = A[p * n + p];
                aqq = A[q * n + q];
                theta = (aqq - app) / (2.0 * apq);
                t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                c = 1.0 / sqrt(t * t + 1.0);
                sn = t * c;
                for (i = 0; i < n; i++) { /* A = J^T A J */
                    double aip = A[i * n + p], aiq = A[i * n + q];
                    A[i * n + p] = c * aip - sn * aiq;
                    A[i * n + q] = sn * aip + c * aiq;
                }
                for (i = 0; i < n; i++) {
                    double api = A[p * n + i], aqi = A[q * n + i];
                    A[p * n + i] = c * api - sn * aqi;
                    A[q * n + i] = sn * api + c * aqi;
                }
                for (i = 0; i < n; i++) { /* rows of E are eigenvectors */
                    double ep = E[p * n + i], eq = E[q * n + i];
                    E[p * n + i] = c * ep - sn * eq;
                    E[q * n + i] = sn * ep + c * eq;
                }
            }
    }
    for (i = 0; i < n; i++) lam[i] = A[i * n + i];
}

/* Eigenbasis of the 1-D second difference of size n whose out-of-grid
 * neighbours are beta * (boundary value): tridiag(1,-2,1) with the two end
 * diagonals -2 + beta (a single cell: -2 + 2 beta).  beta = 0 uses the closed
 * sine basis. */
static void coarse_basis(int n, double beta, double *E, double *lam)
{
    double *A;
    int i;
    if (beta == 0.0) { dst_basis(n, E, lam); return; }
    A = (double *)calloc((size_t)n * n, sizeof(double));
    for (i = 0; i < n; i++) {
        A[i * n + i] = -2.0;
        if (i > 0) A[i * n + i - 1] = 1.0;
        if (i < n - 1) A[i * n + i + 1] = 1.0;
    }
    A[0] += beta;
    A[(n - 1) * n + (n - 1)] += beta;
    eig_sym_jacobi(n, A, E, lam);
    free(A);
}

static void transpose_sq(const double *S, double *T, int n)
{
    int i, j;
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) T[j * n + i] = S[i * n + j];
}

/* Coarsest-level solve.  The reference (MultiGrid_kernel.cu:422-498) runs 100
 * racy in-place relaxations; the build solves the <=512-cell system exactly in
 * the separable eigenbasis: x = E^T (Lambda^-1 (E b)). */
static void mg_coarse_exact_beta(double *x, const double *b, int nx, int ny, int nz, double beta)
{
    int dims[3], a, i, j, k;
    double *E[3], *ET[3], *lam[3], *v;
    dims[0] = nx; dims[1] = ny; dims[2] = nz;
    for (a = 0; a < 3; a++) {
        int n = dims[a];
        E[a] = (double *)malloc(sizeof(double) * n * n);
        ET[a] = (double *)malloc(sizeof(double) * n * n);
        lam[a] = (double *)malloc(sizeof(double) * n);
        coarse_basis(n, beta, E[a], lam[a]);
        transpose_sq(E[a], ET[a], n);
    }
    v = (double *)malloc(sizeof(double) * nx * ny * nz);
    memcpy(v, b, sizeof(double) * nx * ny * nz);
    apply_axis(E[0], nx, v, nx, ny, nz, 0);
    apply_axis(E[1], ny, v, nx, ny, nz, 1);
    apply_axis(E[2], nz, v, nx, ny, nz, 2);
    for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) v[IDX(i, j, k)] /= (lam[0][i] + lam[1][j] + lam[2][k]);
    apply_axis(ET[0], nx, v, nx, ny, nz, 0);
    apply_axis(ET[1], ny, v, nx, ny, nz, 1);
    apply_axis(ET[2], nz, v, nx, ny, nz, 2);
    memcpy(x, v, sizeof(double) * nx * ny * nz);
    free(v);
    for (a = 0; a < 3; a++) { free(E[a]); free(ET[a]); free(lam[a]); }
}

void orc_mg_coarse_exact(double *x, const double *b, int nx, int ny, int nz)
{
    mg_coarse_exact_beta(x, b, nx, ny, nz, 0.0);
}

double orc_mg_norm2(const double *v, uint64_t n)
{
    double s = 0;
    int64_t i;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (i = 0; i < (int64_t)n; i++) s += v[i] * v[i];
    return sqrt(s);
}

double orc_mg_relres(const double *x, const double *b, int nx, int ny, int nz)
{
    size_t n = (size_t)nx * ny * nz;
    double *r = (double *)malloc(sizeof(double) * n);
    double nr, nb;
    orc_mg_residual(x, b, r, nx, ny, nz);
    nr = orc_mg_norm2(r, n);
    nb = orc_mg_norm2(b, n);
    free(r);
    return nb > 0 ? nr / nb : nr;
}

typedef struct {
    int L;
    int d[16][3];
    double *x[16], *b[16], *r[16];
} mg_h;

/* MultiGrid.cpp:128-201 (m_Vcycle): zero the levels below, smooth/residual/
 * restrict down, coarse solve, prolong/smooth up. */
static void mg_vcycle(mg_h *m, int l0, int nu1, int nu2, double omega, int prolong)
{
    /* the tolerance-driven cycle (trilinear prolongation) keeps every level's
     * wall where the finest level has it; the reference pair uses zero ghosts */
    const int consistent = (prolong == ORC_PROLONG_TRILINEAR);
    int l;
    for (l = l0 + 1; l < m->L; l++) {
        size_t n = (size_t)m->d[l][0] * m->d[l][1] * m->d[l][2];
        memset(m->x[l], 0, n * sizeof(double));
        memset(m->b[l], 0, n * sizeof(double));
    }
    for (l = l0; l < m->L - 1; l++) {
        double beta = consistent ? orc_mg_beta(l) : 0.0;
        mg_smooth_beta(m->x[l], m->b[l], m->d[l][0], m->d[l][1], m->d[l][2], nu1, omega, beta);
        mg_residual_beta(m->x[l], m->b[l], m->r[l], m->d[l][0], m->d[l][1], m->d[l][2], beta);
        mg_restrict_dims(m->b[l + 1], m->r[l], m->d[l + 1][0], m->d[l + 1][1], m->d[l + 1][2],
                         m->d[l][0], m->d[l][1]);
    }
    mg_coarse_exact_beta(m->x[m->L - 1], m->b[m->L - 1], m->d[m->L - 1][0], m->d[m->L - 1][1],
                         m->d[m->L - 1][2], consistent ? orc_mg_beta(m->L - 1) : 0.0);
    for (l = m->L - 2; l >= l0; l--) {
        double beta = consistent ? orc_mg_beta(l) : 0.0;
        if (consistent)
            mg_prolong_tri_beta(m->x[l], m->x[l + 1], m->d[l][0], m->d[l][1], m->d[l][2], orc_mg_beta(l + 1));
        else
            orc_mg_prolong(m->x[l], m->x[l + 1], m->d[l][0], m->d[l][1], m->d[l][2]);
        mg_smooth_beta(m->x[l], m->b[l], m->d[l][0], m->d[l][1], m->d[l][2], nu2, omega, beta);
    }
}

/* Whole solve.  ORC_MG_REF_FMG restates MultiGrid.cpp:202-250 (one full
 * multigrid pass plus one V-cycle, no tolerance test); ORC_MG_VCYCLE repeats
 * V(nu1,nu2) cycles until ||b - A x|| / ||b|| <= tol.  Returns cycles run. */
int orc_mg_solve(double *x, const double *b, int nx, int ny, int nz,
                 int schedule, int nu1, int nu2, double omega, double tol,
                 int max_cycles, int prolong, double *relres_out)
{
    mg_h m;
    int l, cycles = 0;
    m.L = orc_mg_levels(nx, ny, nz, m.d, 16);
    m.x[0] = x;
    m.b[0] = (double *)b;
    for (l = 0; l < m.L; l++) {
        size_t n = (size_t)m.d[l][0] * m.d[l][1] * m.d[l][2];
        m.r[l] = (double *)calloc(n, sizeof(double));
        if (l > 0) {
            m.x[l] = (double *)calloc(n, sizeof(double));
            m.b[l] = (double *)calloc(n, sizeof(double));
        }
    }
    if (schedule == ORC_MG_REF_FMG) {
        for (l = 0; l < m.L - 1; l++) {
            orc_mg_residual(m.x[l], m.b[l], m.r[l], m.d[l][0], m.d[l][1], m.d[l][2]);
            mg_restrict_dims(m.b[l + 1], m.r[l], m.d[l + 1][0], m.d[l + 1][1], m.d[l + 1][2],
                             m.d[l][0], m.d[l][1]);
        }
        orc_mg_coarse_exact(m.x[m.L - 1], m.b[m.L - 1], m.d[m.L - 1][0], m.d[m.L - 1][1],
                            m.d[m.L - 1][2]);
        for (l = m.L - 2; l >= 0; l--) {
            if (prolong == ORC_PROLONG_TRILINEAR)
                orc_mg_prolong_tri(m.x[l], m.x[l + 1], m.d[l][0], m.d[l][1], m.d[l][2]);
            else
                orc_mg_prolong(m.x[l], m.x[l + 1], m.d[l][0], m.d[l][1], m.d[l][2]);
            mg_vcycle(&m, l, nu1, nu2, omega, prolong);
        }
        mg_vcycle(&m, 0, nu1, nu2, omega, prolong);
        cycles = 1;
    } else {
        double rel = orc_mg_relres(x, b, nx, ny, nz);
        while (cycles < max_cycles && rel > tol) {
            mg_vcycle(&m, 0, nu1, nu2, omega, prolong);
            cycles++;
            rel = orc_mg_relres(x, b, nx, ny, nz);
        }
    }
    if (relres_out) *relres_out = orc_mg_relres(x, b, nx, ny, nz);
    for (l = 0; l < m.L; l++) {
        free(m.r[l]);
        if (l > 0) { free(m.x[l]); free(m.b[l]); }
    }
    return cycles;
}

/* ===================================================================== */
/* curl, far-field buffer                                                 */
/* ===================================================================== */

/* GPU_kernel.cu:23-30 (sample3D): clamp-to-edge read */
static double sample3(const double *f, int x, int y, int z, int nx, int ny, int nz)
{
    int i = x < 0 ? 0 : (x > nx - 1 ? nx - 1 : x);
    int j = y < 0 ? 0 : (y > ny - 1 ? ny - 1 : y);
    int k = z < 0 ? 0 : (z > nz - 1 ? nz - 1 : z);
    return f[IDX(i, j, k)];
}

/* GPU_kernel.cu:1750-1789 (compute_curl_kernel): central differences with
 * clamped indices; u = dyFz - dzFy, v = dzFx - dxFz, w = dxFy - dyFx */
void orc_curl(const double *Fx, const double *Fy, const double *Fz,
              double *u, double *v, double *w, int nx, int ny, int nz, double h)
{
    const double inv_h = 1.0 / h;
    int64_t t0;
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)nx * ny * nz; t0++) {
        int i = (int)(t0 % nx), j = (int)((t0 / nx) % ny), k = (int)(t0 / ((int64_t)nx * ny));
        double Fz_up = sample3(Fz, i, j + 1, k, nx, ny, nz), Fz_down = sample3(Fz, i, j - 1, k, nx, ny, nz);
        double Fz_left = sample3(Fz, i - 1, j, k, nx, ny, nz), Fz_right = sample3(Fz, i + 1, j, k, nx, ny, nz);
        double Fy_front = sample3(Fy, i, j, k - 1, nx, ny, nz), Fy_back = sample3(Fy, i, j, k + 1, nx, ny, nz);
        double Fy_left = sample3(Fy, i - 1, j, k, nx, ny, nz), Fy_right = sample3(Fy, i + 1, j, k, nx, ny, nz);
        double Fx_front = sample3(Fx, i, j, k - 1, nx, ny, nz), Fx_back = sample3(Fx, i, j, k + 1, nx, ny, nz);
        double Fx_up = sample3(Fx, i, j + 1, k, nx, ny, nz), Fx_down = sample3(Fx, i, j - 1, k, nx, ny, nz);
        u[t0] = 0.5 * inv_h * (Fz_up - Fz_down + Fy_front - Fy_back);
        v[t0] = 0.5 * inv_h * (Fx_back - Fx_front + Fz_left - Fz_right);
        w[t0] = 0.5 * inv_h * (Fy_right - Fy_left + Fx_down - Fx_up);
    }
}

/* GPU_kernel.cu:10-17 (cpu_compute_r): 4 pi |r1 - r2|, or 1e24 if coincident */
static double cpu_compute_r(double ax, double ay, double az, double bx, double by, double bz)
{
    double r_sqr = (ax - bx) * (ax - bx) + (ay - by) * (ay - by) + (az - bz) * (az - bz);
    if (r_sqr > 1e-12) return REF_FOUR_PI * (sqrt(r_sqr));
    return 1e+24;
}

/* GPU_kernel.cu:2646-2677: dir[D] = central difference (step h) of
 * 1/(4 pi |X - X'|) with respect to X, X = centre of cell 0, X' = centre of
 * cell D, for D in [-K,K]^3; dir[0] = 0.  Layout: ((dz+K)(2K+1)+(dy+K))(2K+1)+(dx+K). */
void orc_far_dir_table(int K, double h, double *dir)
{
    const int W = 2 * K + 1;
    int kk, jj, ii;
    for (kk = -K; kk <= K; kk++)
        for (jj = -K; jj <= K; jj++)
            for (ii = -K; ii <= K; ii++) {
                double *d = dir + 3 * (((kk + K) * W + (jj + K)) * W + (ii + K));
                double Xp_x = ((double)ii + 0.5) * h, Xp_y = ((double)jj + 0.5) * h, Xp_z = ((double)kk + 0.5) * h;
                double X_x = 0.5 * h, X_y = 0.5 * h, X_z = 0.5 * h;
                if (ii == 0 && jj == 0 && kk == 0) { d[0] = d[1] = d[2] = 0.0; continue; }
                d[0] = 0.5 * ((1.0 / cpu_compute_r(X_x + h, X_y, X_z, Xp_x, Xp_y, Xp_z)) -
                              (1.0 / cpu_compute_r(X_x - h, X_y, X_z, Xp_x, Xp_y, Xp_z))) / h;
                d[1] = 0.5 * ((1.0 / cpu_compute_r(X_x, X_y + h, X_z, Xp_x, Xp_y, Xp_z)) -
                              (1.0 / cpu_compute_r(X_x, X_y - h, X_z, Xp_x, Xp_y, Xp_z))) / h;
                d[2] = 0.5 * ((1.0 / cpu_compute_r(X_x, X_y, X_z + h, Xp_x, Xp_y, Xp_z)) -
                              (1.0 / cpu_compute_r(X_x, X_y, X_z - h, Xp_x, Xp_y, Xp_z))) / h;
            }
}

/* GPU_kernel.cu:2634-2734 (BiotSavartComputeFarField) with its Step1 kernel
 * (2487-2564), compute_diag_Psi_kernel (1976-2001), getCurl and
 * subtract_diagonal_grid_kernel (2566-2630):
 *   far[c] = U[c] - sum_{D != 0, |D|<=K, c+D in grid} dir[D] x (w[c+D] h^3)
 *                 - curl(0.250580605913965 h^2 w)[c] */
void orc_far_field(const double *wx, const double *wy, const double *wz,
                   const double *Ux, const double *Uy, const double *Uz,
                   double *fx, double *fy, double *fz, double h, int g, int K)
{
    const int W = 2 * K + 1, nx = g, ny = g;
    const size_t G = (size_t)g * g * g;
    double *dir = (double *)malloc(sizeof(double) * 3 * W * W * W);
    double *px = (double *)malloc(sizeof(double) * G), *py = (double *)malloc(sizeof(double) * G),
           *pz = (double *)malloc(sizeof(double) * G);
    double *gu = (double *)malloc(sizeof(double) * G), *gv = (double *)malloc(sizeof(double) * G),
           *gw = (double *)malloc(sizeof(double) * G);
    int64_t t0;
    orc_far_dir_table(K, h, dir);
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)G; t0++) {
        int i = (int)(t0 % g), j = (int)((t0 / g) % g), k = (int)(t0 / ((int64_t)g * g));
        double vx = 0, vy = 0, vz = 0;
        int kk, jj, ii;
        for (kk = k - K; kk <= k + K; kk++)
            for (jj = j - K; jj <= j + K; jj++)
                for (ii = i - K; ii <= i + K; ii++) {
                    double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;
                    if (kk >= 0 && kk < g && jj >= 0 && jj < g && ii >= 0 && ii < g) {
                        size_t r = IDX(ii, jj, kk);
                        double v = h * h * h;
                        ox = wx[r] * v; oy = wy[r] * v; oz = wz[r] * v;
                        if (!(kk == k && jj == j && ii == i)) {
                            const double *d = dir + 3 * (((kk - k + K) * W + (jj - j + K)) * W + (ii - i + K));
                            dx = d[0]; dy = d[1]; dz = d[2];
                        }
                    }
                    vx += dy * oz - dz * oy;
                    vy += dz * ox - dx * oz;
                    vz += dx * oy - dy * ox;
                }
        fx[t0] = Ux[t0] - vx;
        fy[t0] = Uy[t0] - vy;
        fz[t0] = Uz[t0] - vz;
    }
    {
        const double h2 = h * h;
#pragma omp parallel for schedule(static)
        for (t0 = 0; t0 < (int64_t)G; t0++) {
            px[t0] = REF_DIAG_PSI * wx[t0] * h2;
            py[t0] = REF_DIAG_PSI * wy[t0] * h2;
            pz[t0] = REF_DIAG_PSI * wz[t0] * h2;
        }
    }
    orc_curl(px, py, pz, gu, gv, gw, g, g, g, h);
#pragma omp parallel for schedule(static)
    for (t0 = 0; t0 < (int64_t)G; t0++) {
        fx[t0] -= gu[t0];
        fy[t0] -= gv[t0];
        fz[t0] -= gw[t0];
    }
    free(dir); free(px); free(py); free(pz); free(gu); free(gv); free(gw);
}

/* ===================================================================== */
/* M2P, near field                                                        */
/* ===================================================================== */

/* GPU_kernel.cu:51-84 (poly_shape, interpolate_weights): hat function */
static double hat(double r)
{
    double rr = fabs(r);
    return rr < 1 ? 1 - rr : 0;
}

/* GPU_kernel.cu:2886-2964 (BiotSavartInterpolate_far_field_kernel2):
 * trilinear (hat) interpolation over the in-grid cells of the 5^3 block around
 * the point's cell; points whose own cell is outside the grid get nothing. */
void orc_m2p(const float *pos4_sorted, const double *fx, const double *fy,
             const double *fz, double *pu, double *pv, double *pw,
             double cell_h, int g, uint32_t m, const double origin[3])
{
    const double invh = 1.0 / cell_h;
    int64_t t;
#pragma omp parallel for schedule(static)
    for (t = 0; t < (int64_t)m; t++) {
        float qx = (float)((double)pos4_sorted[4 * t] - origin[0]);
        float qy = (float)((double)pos4_sorted[4 * t + 1] - origin[1]);
        float qz = (float)((double)pos4_sorted[4 * t + 2] - origin[2]);
        int i = (int)floor(qx * invh), j = (int)floor(qy * invh), k = (int)floor(qz * invh);
        double sx = 0, sy = 0, sz = 0;
        int kk, jj, ii;
        pu[t] = pv[t] = pw[t] = 0.0;
        if (!(i >= 0 && i < g && j >= 0 && j < g && k >= 0 && k < g)) continue;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    size_t c;
                    double w;
                    float cx = (float)((double)ii + 0.5), cy = (float)((double)jj + 0.5), cz = (float)((double)kk + 0.5);
                    if (!(kk >= 0 && kk < g && jj >= 0 && jj < g && ii >= 0 && ii < g)) continue;
                    c = ((size_t)kk * g + jj) * g + ii;
                    w = hat((double)cx - qx * invh) * hat((double)cy - qy * invh) * hat((double)cz - qz * invh);
                    if (w == 0.0) continue;
                    sx += fx[c] * w;
                    sy += fy[c] * w;
                    sz += fz[c] * w;
                }
        pu[t] = sx;
        pv[t] = sy;
        pw[t] = sz;
    }
}

/* Smoothing-kernel coefficient q(r/a) / (4 pi r^3).
 * EXPCUBIC restates GPU_kernel.cu:2389-2402: q = 1 - exp(-(r/a)^3), computed
 * as -expm1 so that the small-radius branch needs no Taylor series.
 * GAUSSIAN: q = erf(rho/sqrt2) - sqrt(2/pi) rho exp(-rho^2/2) (series below
 * rho = 0.5 to avoid cancellation). */
double orc_kernel_coeff(double r, int kernel, double a)
{
    double rho = r / a, q;
    if (kernel == ORC_KERNEL_GAUSSIAN) {
        if (rho < 0.5) {
            /* sqrt(2/pi) * sum_n (-1)^n rho^(2n+3) / (2^n n! (2n+3)) */
            double r2 = rho * rho, term = rho * r2, s = 0;
            int nn;
            double fact = 1.0;
            for (nn = 0; nn < 8; nn++) {
                if (nn > 0) fact *= -(2.0 * nn) / 1.0; /* (-2)^n n! */
                s += term / (fact * (2 * nn + 3));
                term *= r2;
            }
            q = 0.79788456080286535588 * s;
        } else {
            q = erf(rho * 0.70710678118654752440) - 0.79788456080286535588 * rho * exp(-0.5 * rho * rho);
        }
    } else {
        q = -expm1(-rho * rho * rho);
    }
    return q / (REF_FOUR_PI * r * r * r);
}

/* GPU_kernel.cu:2262-2436 (BiotSavartPPCorrScaleMN_kernel): for the hash cells
 * within +-(K+1) of the evaluation point's cell, add the regularised
 * Biot-Savart term of every vortex whose float coordinate difference is below
 * (K+1/2) h on all three axes; skip r < 1e-12, and with segment == 1 skip
 * vortices within 1e-12 of the other end.  u += M2P value (already in u). */
void orc_near_field(const uint32_t *start, const uint32_t *end,
                    const float *evala4, const float *evalb4, int segment,
                    const float *vort4, const double *wx, const double *wy,
                    const double *wz, double *u, double *v, double *w,
                    double cell_h, int g, int K, uint32_t m, uint32_t n,
                    const double origin[3], int kernel, double core_a)
{
    const double invh = 1.0 / cell_h;
    const double correct_radius = ((double)K + 0.5) / invh;
    int64_t t;
#pragma omp parallel for schedule(dynamic, 64)
    for (t = 0; t < (int64_t)m; t++) {
        float ax = (float)((double)evala4[4 * t] - origin[0]);
        float ay = (float)((double)evala4[4 * t + 1] - origin[1]);
        float az = (float)((double)evala4[4 * t + 2] - origin[2]);
        float bx = 0, by = 0, bz = 0;
        int i = (int)floor(ax * invh), j = (int)floor(ay * invh), k = (int)floor(az * invh);
        double vx = 0, vy = 0, vz = 0;
        int kk, jj, ii;
        if (segment == 1) {
            bx = (float)((double)evalb4[4 * t] - origin[0]);
            by = (float)((double)evalb4[4 * t + 1] - origin[1]);
            bz = (float)((double)evalb4[4 * t + 2] - origin[2]);
        }
        for (kk = k - K - 1; kk <= k + K + 1; kk++)
            for (jj = j - K - 1; jj <= j + K + 1; jj++)
                for (ii = i - K - 1; ii <= i + K + 1; ii++) {
                    uint32_t cell, p;
                    if (!(kk >= 0 && kk < g && jj >= 0 && jj < g && ii >= 0 && ii < g)) continue;
                    cell = (uint32_t)(((int64_t)kk * g + jj) * g + ii);
                    for (p = start[cell]; p < end[cell]; p++) {
                        float sx, sy, sz;
                        double dx, dy, dz, r1, coeff;
                        if (p >= n) continue;
                        sx = (float)((double)vort4[4 * (size_t)p] - origin[0]);
                        sy = (float)((double)vort4[4 * (size_t)p + 1] - origin[1]);
                        sz = (float)((double)vort4[4 * (size_t)p + 2] - origin[2]);
                        if (!((double)fabsf(sx - ax) < correct_radius && (double)fabsf(sy - ay) < correct_radius &&
                              (double)fabsf(sz - az) < correct_radius))
                            continue;
                        dx = (double)(float)(ax - sx);
                        dy = (double)(float)(ay - sy);
                        dz = (double)(float)(az - sz);
                        r1 = sqrt(dx * dx + dy * dy + dz * dz);
                        coeff = (r1 < 1e-12) ? 0.0 : orc_kernel_coeff(r1, kernel, core_a);
                        if (segment == 1) {
                            double ex = (double)(float)(bx - sx), ey = (double)(float)(by - sy),
                                   ez = (double)(float)(bz - sz);
                            if (sqrt(ex * ex + ey * ey + ez * ez) < 1e-12) coeff = 0.0;
                        }
                        {
                            double ox = wx[p], oy = wy[p], oz = wz[p];
                            vx += (oy * dz - oz * dy) * coeff;
                            vy += (oz * dx - ox * dz) * coeff;
                            vz += (ox * dy - oy * dx) * coeff;
                        }
                    }
                }
        u[t] = u[t] + vx;
        v[t] = v[t] + vy;
        w[t] = w[t] + vz;
    }
}

/* ===================================================================== */
/* whole path                                                             */
/* ===================================================================== */

/* BiotSavartSolverImp.cpp:378-409 (setDomain): bounding box of the float
 * positions in double, L = 3 max extent, origin = centre - L/2. */
void orc_reference_domain(const float *pos4, uint32_t n, uint32_t g,
                          double origin_out[3], double *L_out)
{
    double maxx, minx, maxy, miny, maxz, minz, L;
    uint32_t i;
    (void)g;
    maxx = minx = pos4[0];
    maxy = miny = pos4[1];
    maxz = minz = pos4[2];
    for (i = 1; i < n; i++) {
        if (pos4[4 * i] < minx) minx = pos4[4 * i]; else if (pos4[4 * i] > maxx) maxx = pos4[4 * i];
        if (pos4[4 * i + 1] < miny) miny = pos4[4 * i + 1]; else if (pos4[4 * i + 1] > maxy) maxy = pos4[4 * i + 1];
        if (pos4[4 * i + 2] < minz) minz = pos4[4 * i + 2]; else if (pos4[4 * i + 2] > maxz) maxz = pos4[4 * i + 2];
    }
    L = maxx - minx;
    if (maxy - miny > L) L = maxy - miny;
    if (maxz - minz > L) L = maxz - minz;
    L *= 3;
    origin_out[0] = (minx + 0.5 * (maxx - minx)) - 0.5 * L;
    origin_out[1] = (miny + 0.5 * (maxy - miny)) - 0.5 * L;
    origin_out[2] = (minz + 0.5 * (maxz - minz)) - 0.5 * L;
    *L_out = L;
}

/* BiotSavartSolver::computeFarFieldBuffer (BiotSavartSolverImp.cpp:755-816,
 * m_ParticleToMesh 517-602, m_SolvePoisson 603-620, m_ComputeCurl 621-639)
 * followed by evaluateVelocity (410-516): hash + reorder the vortices, P2M,
 * three Poisson solves, curl, far-field buffer, hash + reorder the evaluation
 * points, M2P, near field, unsort. */
int orc_pppm_eval(const orc_params *p, uint32_t n, const float *vort_pos4,
                  const double *wx, const double *wy, const double *wz,
                  uint32_t m, const float *eval_pos4, const float *other_end4,
                  double *ux, double *uy, double *uz, orc_diag *diag)
{
    const uint32_t g = p->g;
    const size_t G = (size_t)g * g * g;
    const double cell_h = (double)p->hf;
    uint32_t *key = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *start = (uint32_t *)malloc(sizeof(uint32_t) * G), *end = (uint32_t *)malloc(sizeof(uint32_t) * G);
    float *spos = (float *)malloc(sizeof(float) * 4 * (n ? n : 1));
    double *sw[3], *wg[3], *psi[3], *U[3], *far[3];
    const double *win[3];
    uint32_t *ekey, *eidx, *estart, *eend;
    float *epos, *eoth = NULL;
    double *us[3];
    double t0, t1;
    int c;
    win[0] = wx; win[1] = wy; win[2] = wz;
    if (diag) memset(diag, 0, sizeof(*diag));

    t0 = wtime();
    orc_hash_keys(vort_pos4, n, g, p->hf, p->originf, key, idx);
    orc_sort_pairs(key, idx, n);
    orc_cell_tables(key, n, (uint32_t)G, start, end);
    orc_reorder(idx, vort_pos4, 4, spos, 1, n);
    for (c = 0; c < 3; c++) {
        sw[c] = (double *)malloc(sizeof(double) * (n ? n : 1));
        orc_reorder(idx, win[c], 1, sw[c], 2, n);
    }
    t1 = wtime();
    if (diag) diag->t_hash = t1 - t0;

    t0 = t1;
    for (c = 0; c < 3; c++) {
        wg[c] = (double *)malloc(sizeof(double) * G);
        orc_p2m(start, end, spos, sw[c], cell_h, wg[c], g, n, p->origin);
    }
    t1 = wtime();
    if (diag) diag->t_p2m = t1 - t0;

    t0 = t1;
    for (c = 0; c < 3; c++) {
        double *rhs = (double *)malloc(sizeof(double) * G);
        double rel = 0;
        int cyc;
        memcpy(rhs, wg[c], sizeof(double) * G);
        orc_rhs(rhs, cell_h * cell_h, -1.0, G);
        psi[c] = (double *)calloc(G, sizeof(double));
        cyc = orc_mg_solve(psi[c], rhs, (int)g, (int)g, (int)g, p->mg_schedule, p->nu1, p->nu2,
                           p->mg_omega, p->mg_tol, p->mg_max_cycles, p->mg_prolong, &rel);
        if (diag) { diag->mg_cycles[c] = cyc; diag->mg_relres[c] = rel; }
        free(rhs);
    }
    t1 = wtime();
    if (diag) diag->t_mg = t1 - t0;

    t0 = t1;
    for (c = 0; c < 3; c++) {
        U[c] = (double *)malloc(sizeof(double) * G);
        far[c] = (double *)malloc(sizeof(double) * G);
    }
    orc_curl(psi[0], psi[1], psi[2], U[0], U[1], U[2], (int)g, (int)g, (int)g, cell_h);
    if (p->K >= 0) {
        orc_far_field(wg[0], wg[1], wg[2], U[0], U[1], U[2], far[0], far[1], far[2], cell_h, (int)g, p->K);
    } else { /* vortex-in-cell (BiotSavartSolverImp.cpp:487-497): interpolate grid_U */
        for (c = 0; c < 3; c++) memcpy(far[c], U[c], sizeof(double) * G);
    }
    t1 = wtime();
    if (diag) diag->t_far = t1 - t0;

    t0 = t1;
    ekey = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    eidx = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    estart = (uint32_t *)malloc(sizeof(uint32_t) * G);
    eend = (uint32_t *)malloc(sizeof(uint32_t) * G);
    epos = (float *)malloc(sizeof(float) * 4 * (m ? m : 1));
    orc_hash_keys(eval_pos4, m, g, p->hf, p->originf, ekey, eidx);
    orc_sort_pairs(ekey, eidx, m);
    orc_cell_tables(ekey, m, (uint32_t)G, estart, eend);
    orc_reorder(eidx, eval_pos4, 4, epos, 1, m);
    if (other_end4) {
        eoth = (float *)malloc(sizeof(float) * 4 * (m ? m : 1));
        orc_reorder(eidx, other_end4, 4, eoth, 1, m);
    }
    for (c = 0; c < 3; c++) us[c] = (double *)calloc(m ? m : 1, sizeof(double));
    orc_m2p(epos, far[0], far[1], far[2], us[0], us[1], us[2], cell_h, (int)g, m, p->origin);
    orc_near_field(start, end, epos, eoth, other_end4 ? 1 : 0, spos, sw[0], sw[1], sw[2], us[0], us[1],
                   us[2], cell_h, (int)g, p->K, m, n, p->origin, p->kernel, p->core_a);
    orc_deorder(eidx, us[0], 1, ux, 2, m);
    orc_deorder(eidx, us[1], 1, uy, 2, m);
    orc_deorder(eidx, us[2], 1, uz, 2, m);
    t1 = wtime();
    if (diag) diag->t_eval = t1 - t0;

    for (c = 0; c < 3; c++) { free(sw[c]); free(wg[c]); free(psi[c]); free(U[c]); free(far[c]); free(us[c]); }
    free(key); free(idx); free(start); free(end); free(spos);
    free(ekey); free(eidx); free(estart); free(eend); free(epos);
    if (eoth) free(eoth);
    return 0;
}

/* FP64 direct O(N M) regularised Biot-Savart sum with the same kernel and the
 * same r < 1e-12 exclusion, no cutoff: u_i = sum_j q(r/a)/(4 pi r^3) w_j x (x_i - x_j). */
void orc_direct_sum(uint32_t n, const float *vort_pos4, const double *wx,
                    const double *wy, const double *wz, uint32_t m,
                    const float *eval_pos4, int kernel, double core_a,
                    double *ux, double *uy, double *uz)
{
    int64_t t;
#pragma omp parallel for schedule(dynamic, 16)
    for (t = 0; t < (int64_t)m; t++) {
        double x = eval_pos4[4 * t], y = eval_pos4[4 * t + 1], z = eval_pos4[4 * t + 2];
        double vx = 0, vy = 0, vz = 0;
        uint32_t j;
        for (j = 0; j < n; j++) {
            double dx = x - (double)vort_pos4[4 * (size_t)j], dy = y - (double)vort_pos4[4 * (size_t)j + 1],
                   dz = z - (double)vort_pos4[4 * (size_t)j + 2];
            double r = sqrt(dx * dx + dy * dy + dz * dz), coeff;
            if (r < 1e-12) continue;
            coeff = orc_kernel_coeff(r, kernel, core_a);
            vx += (wy[j] * dz - wz[j] * dy) * coeff;
            vy += (wz[j] * dx - wx[j] * dz) * coeff;
            vz += (wx[j] * dy - wy[j] * dx) * coeff;
        }
        ux[t] = vx;
        uy[t] = vy;
        uz[t] = vz;
    }
}

/* ---- time step of the vortex-segment method ---------------------------
 * VortexSimulationSystemImp.cpp:90-125 (solveTimeStep) and the kernels it
 * calls.  Positions are float4 arrays (x, y, z, w); only the fields the
 * reference writes are written. */

/* GPU_kernel.cu:3182-3235 compute_posb_kernel / ComputePosb: segment ends
 * enda = c - s/2, endb = c + s/2 with s = w when |w| < h (kappa = 1), else
 * s = w h / |w| (kappa = |w| / h). */
void orc_segments_from_vortices(const float *cent4, float *enda4, float *endb4, const double *wx,
                                const double *wy, const double *wz, double *kappa, double h, uint32_t n)
{
    uint32_t i;
    for (i = 0; i < n; i++) {
        double s[3] = {wx[i], wy[i], wz[i]};
        double l = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        int c;
        kappa[i] = l < h ? 1.0 : l / h;
        for (c = 0; c < 3; c++) {
            double half = l < h ? 0.5 * s[c] : 0.5 * s[c] / l * h;
            double a = (double)cent4[4 * (size_t)i + c];
            endb4[4 * (size_t)i + c] = (float)(a + half);
            enda4[4 * (size_t)i + c] = (float)(a - half);
        }
    }
}

/* GPU_kernel.cu:3118-3140 move_particle_kernel: pos += u dt in FP64, stored
 * as float with w = 0.  u is double4 (x, y, z, w). */
void orc_move_points(float *pos4, double dt, const double *u4, uint32_t n)
{
    uint32_t i;
    for (i = 0; i < n; i++) {
        int c;
        for (c = 0; c < 3; c++)
            pos4[4 * (size_t)i + c] = (float)((double)pos4[4 * (size_t)i + c] + u4[4 * (size_t)i + c] * dt);
        pos4[4 * (size_t)i + 3] = 0.0f;
    }
}

/* GPU_kernel.cu:3237-3254 compute_posa_kernel: the centre is 0.5 (b + c)
 * with the sum taken in float. */
void orc_segment_centres(const float *posc4, const float *posb4, float *posa4, uint32_t n)
{
    uint32_t i;
    for (i = 0; i < n; i++) {
        int c;
        for (c = 0; c < 3; c++) {
            float sum = posb4[4 * (size_t)i + c] + posc4[4 * (size_t)i + c];
            posa4[4 * (size_t)i + c] = (float)(0.5 * (double)sum);
        }
    }
}

/* GPU_kernel.cu:3143-3180 compute_vort_kernel: dw = kappa (b - a) - w with
 * b - a taken in float. */
void orc_vortices_from_segments(const float *posa4, const float *posb4, const double *wx, const double *wy,
                                const double *wz, double *dwx, double *dwy, double *dwz, const double *kappa,
                                uint32_t n)
{
    uint32_t i;
    for (i = 0; i < n; i++) {
        double sx = (double)(posb4[4 * (size_t)i] - posa4[4 * (size_t)i]);
        double sy = (double)(posb4[4 * (size_t)i + 1] - posa4[4 * (size_t)i + 1]);
        double sz = (double)(posb4[4 * (size_t)i + 2] - posa4[4 * (size_t)i + 2]);
        dwx[i] = kappa[i] * sx - wx[i];
        dwy[i] = kappa[i] * sy - wy[i];
        dwz[i] = kappa[i] * sz - wz[i];
    }
}

/* GPU_kernel.cu:3327-3438 particle_gaussian_kernel / ParticleGaussianSmooth:
 * o[t] = sum_p e^{-d^2 / (2 nu^2)} i[p] / sum_p e^{-d^2 / (2 nu^2)} over the
 * sorted particles p of the hash cells within +-2 of t's cell (total weight
 * below 1e-12 divides by 1e24).  pos4 in caller order, pos4_sorted and i_*
 * in hash order; d is taken between origin-shifted float coordinates. */
void orc_gaussian_smooth(const uint32_t *start, const uint32_t *end, const float *pos4,
                         const float *pos4_sorted, double nu, double *ox, double *oy, double *oz,
                         const double *ix, const double *iy, const double *iz, double h, int g, uint32_t n,
                         const double origin[3])
{
    const double invh = 1.0 / h;
    int64_t t;
#pragma omp parallel for schedule(dynamic, 64)
    for (t = 0; t < (int64_t)n; t++) {
        float ax = (float)((double)pos4[4 * t] - origin[0]);
        float ay = (float)((double)pos4[4 * t + 1] - origin[1]);
        float az = (float)((double)pos4[4 * t + 2] - origin[2]);
        int i = (int)floor(ax * invh), j = (int)floor(ay * invh), k = (int)floor(az * invh);
        double sx = 0, sy = 0, sz = 0, tw = 0;
        int kk, jj, ii;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    uint32_t cell, p;
                    if (!(kk >= 0 && kk < g && jj >= 0 && jj < g && ii >= 0 && ii < g)) continue;
                    cell = (uint32_t)(((int64_t)kk * g + jj) * g + ii);
                    for (p = start[cell]; p < end[cell]; p++) {
                        float bx, by, bz;
                        double dx, dy, dz, d2, wgt;
                        if (p >= n) continue;
                        bx = (float)((double)pos4_sorted[4 * (size_t)p] - origin[0]);
                        by = (float)((double)pos4_sorted[4 * (size_t)p + 1] - origin[1]);
                        bz = (float)((double)pos4_sorted[4 * (size_t)p + 2] - origin[2]);
                        dx = (double)(float)(bx - ax);
                        dy = (double)(float)(by - ay);
                        dz = (double)(float)(bz - az);
                        d2 = dx * dx + dy * dy + dz * dz;
                        wgt = exp(-d2 / (2 * nu * nu));
                        tw += wgt;
                        sx += wgt * ix[p];
                        sy += wgt * iy[p];
                        sz += wgt * iz[p];
                    }
                }
        if (tw < 1e-12) tw = 1e+24;
        ox[t] = sx / tw;
        oy[t] = sy / tw;
        oz[t] = sz / tw;
    }
}

/* ---- baroclinic source (addBaroclinicForce) ---------------------------
 * VortexSimulationSystemImp.cpp:320-367 (setDensityDomain), 369-465, and
 * GPU_kernel.cu:3655-3898. */

/* setDensityDomain + the grid rule of addBaroclinicForce: L = 1.2 x the
 * largest extent, origin = box centre - L/2, dim = min(floor(L/cell) + 1,
 * max_dim), h = L / dim. */
void orc_density_domain(const float *pos4, uint32_t n, double cell_target, uint32_t max_dim, double origin[3],
                        double *h_out, uint32_t *dim_out)
{
    double mn[3], mx[3], L;
    uint32_t i, dim;
    int c;
    for (c = 0; c < 3; c++) mn[c] = mx[c] = pos4[c];
    for (i = 1; i < n; i++)
        for (c = 0; c < 3; c++) {
            double v = pos4[4 * (size_t)i + c];
            if (v < mn[c]) mn[c] = v; else if (v > mx[c]) mx[c] = v;
        }
    L = mx[0] - mn[0];
    if (mx[1] - mn[1] > L) L = mx[1] - mn[1];
    if (mx[2] - mn[2] > L) L = mx[2] - mn[2];
    L = 1.2 * L;
    for (c = 0; c < 3; c++) origin[c] = (mn[c] + 0.5 * (mx[c] - mn[c])) - 0.5 * L;
    dim = (uint32_t)floor(L / cell_target) + 1;
    if (dim > max_dim) dim = max_dim;
    *dim_out = dim;
    *h_out = L / (double)dim;
}

static double orc_poly_shape2(double r)
{
    double rr = fabs(r), r1 = 1 - rr, r2 = r1 * r1;
    return rr < 1 ? r2 * r2 : 0.0;
}

static double orc_poly_shape(double r)
{
    double rr = fabs(r);
    return rr < 1 ? 1 - rr : 0.0;
}

/* GPU_kernel.cu:3655-3725 assign_density_kernel over the +-2 hash cells.  An
 * empty cell (start = EMPTY) takes the `end < start` branch: the reference
 * never clears its end table, so the branch adds the weight of the empty
 * cell's centre with value 0. */
void orc_assign_density(const uint32_t *start, const uint32_t *end, const float *pos4_sorted, const float *rho,
                        double cell_h, uint32_t dim, uint32_t n, const double origin[3], float *grid)
{
    const double invh = 1.0 / cell_h;
    const int d = (int)dim;
    int64_t c;
#pragma omp parallel for schedule(static)
    for (c = 0; c < (int64_t)dim * dim * dim; c++) {
        int i = (int)(c % dim), j = (int)((c / dim) % dim), k = (int)(c / ((int64_t)dim * dim));
        double gx = (double)i + 0.5, gy = (double)j + 0.5, gz = (double)k + 0.5;
        double sum = 0, tw = 0;
        int kk, jj, ii;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    uint32_t q, p;
                    if (!(kk >= 0 && kk < d && jj >= 0 && jj < d && ii >= 0 && ii < d)) continue;
                    q = ((uint32_t)kk * dim + (uint32_t)jj) * dim + (uint32_t)ii;
                    if (start[q] != EMPTY && end[q] >= start[q]) {
                        for (p = start[q]; p < end[q]; p++) {
                            float px, py, pz;
                            double dx, dy, dz, w;
                            if (p >= n) continue;
                            px = (float)((double)pos4_sorted[4 * (size_t)p] - origin[0]);
                            py = (float)((double)pos4_sorted[4 * (size_t)p + 1] - origin[1]);
                            pz = (float)((double)pos4_sorted[4 * (size_t)p + 2] - origin[2]);
                            dx = gx - px * invh;
                            dy = gy - py * invh;
                            dz = gz - pz * invh;
                            w = orc_poly_shape2(dx * dx + dy * dy + dz * dz);
                            sum += w * (double)rho[p];
                            tw += w;
                        }
                    } else {
                        double dx = gx - ((double)ii + 0.5), dy = gy - ((double)jj + 0.5), dz = gz - ((double)kk + 0.5);
                        double w = orc_poly_shape2(dx * dx + dy * dy + dz * dz);
                        sum += 0.0;
                        tw += w;
                    }
                }
        if (tw < 1e-12) tw = 1e+24;
        grid[c] = (float)(sum / tw);
    }
}

/* GPU_kernel.cu:3752-3791 grad_density_kernel: periodic central differences. */
void orc_grad_density(const float *phi, uint32_t dim, double cell_h, float *grad)
{
    const double invh = 1.0 / cell_h;
    uint64_t c, G = (uint64_t)dim * dim * dim;
    for (c = 0; c < G; c++) {
        uint32_t i = (uint32_t)(c % dim), j = (uint32_t)((c / dim) % dim), k = (uint32_t)(c / ((uint64_t)dim * dim));
        uint32_t im = (i + dim - 1) % dim, ip = (i + 1) % dim, jm = (j + dim - 1) % dim, jp = (j + 1) % dim;
        uint32_t km = (k + dim - 1) % dim, kp = (k + 1) % dim;
#define PHI(a, b, e) ((double)phi[((uint64_t)(e) * dim + (b)) * dim + (a)])
        grad[3 * c] = (float)((PHI(ip, j, k) - PHI(im, j, k)) * 0.5 * invh);
        grad[3 * c + 1] = (float)((PHI(i, jp, k) - PHI(i, jm, k)) * 0.5 * invh);
        grad[3 * c + 2] = (float)((PHI(i, j, kp) - PHI(i, j, km)) * 0.5 * invh);
#undef PHI
    }
}

/* GPU_kernel.cu:3810-3880 baroclinic_kernel over the +-2 grid cells:
 * w += dt_beta * 0.1 * (sum_cells grad hat-weight) x g. */
void orc_baroclinic(const float *vpos4, double *wx, double *wy, double *wz, const float *grad, double dt_beta,
                    uint32_t n, uint32_t dim, double inv_h, const double origin[3], const double g[3])
{
    const int d = (int)dim;
    int64_t t;
#pragma omp parallel for schedule(static)
    for (t = 0; t < (int64_t)n; t++) {
        float px = (float)((double)vpos4[4 * t] - origin[0]);
        float py = (float)((double)vpos4[4 * t + 1] - origin[1]);
        float pz = (float)((double)vpos4[4 * t + 2] - origin[2]);
        double qx = px * inv_h, qy = py * inv_h, qz = pz * inv_h;
        int i = (int)floor(qx), j = (int)floor(qy), k = (int)floor(qz);
        double sx = 0, sy = 0, sz = 0, bx, by, bz;
        int kk, jj, ii;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    uint64_t q;
                    double w;
                    if (!(kk >= 0 && kk < d && jj >= 0 && jj < d && ii >= 0 && ii < d)) continue;
                    q = ((uint64_t)kk * dim + (uint64_t)jj) * dim + (uint64_t)ii;
                    w = orc_poly_shape((double)ii + 0.5 - qx) * orc_poly_shape((double)jj + 0.5 - qy) *
                        orc_poly_shape((double)kk + 0.5 - qz);
                    sx += (double)grad[3 * q] * w;
                    sy += (double)grad[3 * q + 1] * w;
                    sz += (double)grad[3 * q + 2] * w;
                }
        bx = sy * g[2] - sz * g[1];
        by = sz * g[0] - sx * g[2];
        bz = sx * g[1] - sy * g[0];
        wx[t] += dt_beta * 0.1 * bx;
        wy[t] += dt_beta * 0.1 * by;
        wz[t] += dt_beta * 0.1 * bz;
    }
}

/* ---- potential-flow gradient path (PotentialFieldSolver) ----------------
 * PotentialFieldSolver.cpp:441-739 and GPU_kernel.cu:560-610, 750-912,
 * 964-1200, 1520-1594, 1678-1718, 2977-3117.  Vectors are (x, y, z) triples. */

/* compute_gradient_kernel (GPU_kernel.cu:560-587): periodic central differences. */
void orc_pot_gradient(const double *phi, double *dphi, double h, double direction, int n)
{
    const double invh = 1.0 / h;
    int64_t c;
    for (c = 0; c < (int64_t)n * n * n; c++) {
        int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
        int im = (i + n - 1) % n, ip = (i + 1) % n, jm = (j + n - 1) % n, jp = (j + 1) % n;
        int km = (k + n - 1) % n, kp = (k + 1) % n;
#define PH(a, b, e) phi[((int64_t)(e) * n + (b)) * n + (a)]
        dphi[3 * c] = direction * (PH(ip, j, k) - PH(im, j, k)) * 0.5 * invh;
        dphi[3 * c + 1] = direction * (PH(i, jp, k) - PH(i, jm, k)) * 0.5 * invh;
        dphi[3 * c + 2] = direction * (PH(i, j, kp) - PH(i, j, km)) * 0.5 * invh;
#undef PH
    }
}

static double orc_four_pi_r(double x, double y, double z, const double c[3])
{
    double r2 = (x - c[0]) * (x - c[0]) + (y - c[1]) * (y - c[1]) + (z - c[2]) * (z - c[2]);
    return r2 > 1e-12 ? 12.5663706144 * sqrt(r2) : 1e+12;
}

/* apply_Dirichlet_kernel (GPU_kernel.cu:2977-3101): rhs -= M / (4 pi r) at the
 * ghost centre of each boundary face, faces in the order x-, x+, y-, y+, z-, z+. */
void orc_pot_dirichlet(double *rhs, const double center[3], double monopole, const double corner[3], double h, int n)
{
    int64_t c;
    for (c = 0; c < (int64_t)n * n * n; c++) {
        int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
        double g[6][3] = {{i - 1.0 + 0.5, j + 0.5, k + 0.5}, {i + 1.0 + 0.5, j + 0.5, k + 0.5},
                          {i + 0.5, j - 1.0 + 0.5, k + 0.5}, {i + 0.5, j + 1.0 + 0.5, k + 0.5},
                          {i + 0.5, j + 0.5, k - 1.0 + 0.5}, {i + 0.5, j + 0.5, k + 1.0 + 0.5}};
        int on[6] = {i == 0, i == n - 1, j == 0, j == n - 1, k == 0, k == n - 1};
        int f;
        for (f = 0; f < 6; f++)
            if (on[f]) {
                double x = g[f][0] * h + corner[0], y = g[f][1] * h + corner[1], z = g[f][2] * h + corner[2];
                rhs[c] -= monopole / orc_four_pi_r(x, y, z, center);
            }
    }
}

/* PotentialComputeFarField (GPU_kernel.cu:1098-1200): far = gforce - sum_{D != 0}
 * dir[D] (G direction) rho[c+D] h^3; then phi = 0.250580605913965 rho h^2 G,
 * gforce = its gradient (direction), far -= gforce. */
void orc_pot_far_field(const double *rho, double *phi, double *gforce, double *far, double h, double G,
                       double direction, int n, int K)
{
    int W = 2 * K + 1;
    double *dir = (double *)malloc(sizeof(double) * 3 * W * W * W);
    int64_t c, G3 = (int64_t)n * n * n;
    orc_far_dir_table(K, h, dir);
#pragma omp parallel for schedule(static)
    for (c = 0; c < G3; c++) {
        int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
        double sx = 0, sy = 0, sz = 0;
        int ii, jj, kk;
        for (kk = k - K; kk <= k + K; kk++)
            for (jj = j - K; jj <= j + K; jj++)
                for (ii = i - K; ii <= i + K; ii++) {
                    const double *d;
                    double m;
                    if (!(kk >= 0 && kk < n && jj >= 0 && jj < n && ii >= 0 && ii < n)) continue;
                    if (kk == k && jj == j && ii == i) continue;
                    m = (G * direction) * rho[((int64_t)kk * n + jj) * n + ii] * (h * h * h);
                    d = dir + 3 * (((kk - k + K) * W + (jj - j + K)) * W + (ii - i + K));
                    sx += d[0] * m;
                    sy += d[1] * m;
                    sz += d[2] * m;
                }
        far[3 * c] = gforce[3 * c] - sx;
        far[3 * c + 1] = gforce[3 * c + 1] - sy;
        far[3 * c + 2] = gforce[3 * c + 2] - sz;
    }
    for (c = 0; c < G3; c++) phi[c] = 0.250580605913965 * rho[c] * (h * h) * G;
    orc_pot_gradient(phi, gforce, h, direction, n);
    for (c = 0; c < 3 * G3; c++) far[c] -= gforce[c];
    free(dir);
}

/* PotentialInterpolate_far_field_kernel2 (GPU_kernel.cu:1520-1576) */
void orc_pot_interp(const float *pos4, const double *far, double *pf, double h, int n, uint32_t m,
                    const double origin[3])
{
    const double invh = 1.0 / h;
    int64_t t;
#pragma omp parallel for schedule(static)
    for (t = 0; t < (int64_t)m; t++) {
        float px = (float)((double)pos4[4 * t] - origin[0]);
        float py = (float)((double)pos4[4 * t + 1] - origin[1]);
        float pz = (float)((double)pos4[4 * t + 2] - origin[2]);
        double fi = floor(px * invh), fj = floor(py * invh), fk = floor(pz * invh);
        double qx = px * invh, qy = py * invh, qz = pz * invh, sx = 0, sy = 0, sz = 0;
        int i, j, k, ii, jj, kk;
        if (!(fi >= 0 && fi < n && fj >= 0 && fj < n && fk >= 0 && fk < n)) continue;
        i = (int)fi;
        j = (int)fj;
        k = (int)fk;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    int64_t q;
                    double w;
                    if (!(kk >= 0 && kk < n && jj >= 0 && jj < n && ii >= 0 && ii < n)) continue;
                    q = ((int64_t)kk * n + jj) * n + ii;
                    w = orc_poly_shape((double)(float)(ii + 0.5) - qx) * orc_poly_shape((double)(float)(jj + 0.5) - qy) *
                        orc_poly_shape((double)(float)(kk + 0.5) - qz);
                    sx += far[3 * q] * w;
                    sy += far[3 * q + 1] * w;
                    sz += far[3 * q + 2] * w;
                }
        pf[3 * t] = sx;
        pf[3 * t + 1] = sy;
        pf[3 * t + 2] = sz;
    }
}

/* Potential_PPCorrMN_kernel (GPU_kernel.cu:750-870), G already times direction */
void orc_pot_near(const uint32_t *start, const uint32_t *end, const float *pe4, const float *pm4, const double *mass,
                  double *pf, double h, double G, int n, int K, uint32_t M, uint32_t N, const double origin[3])
{
    const double invh = 1.0 / h, R = ((double)K + 0.5) / invh, a = 0.05;
    int64_t t;
#pragma omp parallel for schedule(dynamic, 64)
    for (t = 0; t < (int64_t)M; t++) {
        float ax = (float)((double)pe4[4 * t] - origin[0]);
        float ay = (float)((double)pe4[4 * t + 1] - origin[1]);
        float az = (float)((double)pe4[4 * t + 2] - origin[2]);
        int i = (int)floor(ax * invh), j = (int)floor(ay * invh), k = (int)floor(az * invh);
        double sx = 0, sy = 0, sz = 0;
        int ii, jj, kk;
        for (kk = k - K - 1; kk <= k + K + 1; kk++)
            for (jj = j - K - 1; jj <= j + K + 1; jj++)
                for (ii = i - K - 1; ii <= i + K + 1; ii++) {
                    uint32_t q, p;
                    if (!(kk >= 0 && kk < n && jj >= 0 && jj < n && ii >= 0 && ii < n)) continue;
                    q = (uint32_t)(((int64_t)kk * n + jj) * n + ii);
                    for (p = start[q]; p < end[q]; p++) {
                        float bx, by, bz;
                        double dx, dy, dz, r1, r3;
                        if (p >= N) continue;
                        bx = (float)((double)pm4[4 * (size_t)p] - origin[0]);
                        by = (float)((double)pm4[4 * (size_t)p + 1] - origin[1]);
                        bz = (float)((double)pm4[4 * (size_t)p + 2] - origin[2]);
                        if (!((double)fabsf(bx - ax) < R && (double)fabsf(by - ay) < R && (double)fabsf(bz - az) < R))
                            continue;
                        dx = (double)(float)(bx - ax);
                        dy = (double)(float)(by - ay);
                        dz = (double)(float)(bz - az);
                        r1 = sqrt(dx * dx + dy * dy + dz * dz);
                        r3 = r1 > a ? 12.5663706144 * r1 * r1 * r1 : (r1 < 1e-12 ? 1e+14 : 12.5663706144 * a * a * a);
                        sx += G * dx / r3 * mass[p];
                        sy += G * dy / r3 * mass[p];
                        sz += G * dz / r3 * mass[p];
                    }
                }
        pf[3 * t] = pf[3 * t] + sx;
        pf[3 * t + 1] = pf[3 * t + 1] + sy;
        pf[3 * t + 2] = pf[3 * t + 2] + sz;
    }
}

/* Potentialparticle_outside_kernel (GPU_kernel.cu:1678-1709): zero outside the box */
void orc_pot_outside(const float *pos4, const float lo[3], const float hi[3], double *pf, uint32_t M)
{
    uint32_t t;
    for (t = 0; t < M; t++) {
        double x = pos4[4 * (size_t)t], y = pos4[4 * (size_t)t + 1], z = pos4[4 * (size_t)t + 2];
        if (x < lo[0] || x > hi[0] || y < lo[1] || y > hi[1] || z < lo[2] || z > hi[2])
            pf[3 * (size_t)t] = pf[3 * (size_t)t + 1] = pf[3 * (size_t)t + 2] = 0.0;
    }
}

/* PotentialComputeFarFieldScalarStep1_kernel (GPU_kernel.cu:1201-1267), G
 * already times direction */
void orc_pot_scalar_far(const double *rho, double *phi, double h, double G, int n, int K)
{
    int64_t c;
#pragma omp parallel for schedule(static)
    for (c = 0; c < (int64_t)n * n * n; c++) {
        int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
        double sum = 0.0;
        int ii, jj, kk;
        for (kk = k - K; kk <= k + K; kk++)
            for (jj = j - K; jj <= j + K; jj++)
                for (ii = i - K; ii <= i + K; ii++) {
                    double m, w;
                    if (!(kk >= 0 && kk < n && jj >= 0 && jj < n && ii >= 0 && ii < n)) continue;
                    m = G * rho[((int64_t)kk * n + jj) * n + ii] * (h * h);
                    if (ii == i && jj == j && kk == k) {
                        w = 0.25;
                    } else {
                        double x = (double)ii - (double)i, y = (double)jj - (double)j, z = (double)kk - (double)k;
                        w = 1.0 / (12.5663706144 * sqrt(x * x + y * y + z * z));
                    }
                    sum += w * m;
                }
        phi[c] = phi[c] - sum;
    }
}

/* PotentialInterpolate_far_fieldScalar_kernel2 (GPU_kernel.cu:1595-1640) */
void orc_pot_interp_scalar(const float *pos4, const double *far, double *pf, double h, int n, uint32_t m,
                           const double origin[3])
{
    const double invh = 1.0 / h;
    int64_t t;
    for (t = 0; t < (int64_t)m; t++) {
        float px = (float)((double)pos4[4 * t] - origin[0]);
        float py = (float)((double)pos4[4 * t + 1] - origin[1]);
        float pz = (float)((double)pos4[4 * t + 2] - origin[2]);
        double fi = floor(px * invh), fj = floor(py * invh), fk = floor(pz * invh);
        double qx = px * invh, qy = py * invh, qz = pz * invh, sum = 0;
        int i, j, k, ii, jj, kk;
        if (!(fi >= 0 && fi < n && fj >= 0 && fj < n && fk >= 0 && fk < n)) continue;
        i = (int)fi;
        j = (int)fj;
        k = (int)fk;
        for (kk = k - 2; kk <= k + 2; kk++)
            for (jj = j - 2; jj <= j + 2; jj++)
                for (ii = i - 2; ii <= i + 2; ii++) {
                    if (!(kk >= 0 && kk < n && jj >= 0 && jj < n && ii >= 0 && ii < n)) continue;
                    sum += far[((int64_t)kk * n + jj) * n + ii] *
                           (orc_poly_shape((double)(float)(ii + 0.5) - qx) * orc_poly_shape((double)(float)(jj + 0.5) - qy) *
                            orc_poly_shape((double)(float)(kk + 0.5) - qz));
                }
        pf[t] = sum;
    }
}

/* Potential_PPCorrMNScalar_kernel (GPU_kernel.cu:1301-1419), G already times
 * direction; the visited cells reach K + 1 on the side of the point's half */
void orc_pot_near_scalar(const uint32_t *start, const uint32_t *end, const float *pe4, const float *pm4,
                         const double *mass, double *pf, double h, double G, int n, int K, uint32_t M, uint32_t N,
                         const double origin[3])
{
    const double invh = 1.0 / h, R = ((double)K + 0.5) / invh, a = 0.125;
    int64_t t;
#pragma omp parallel for schedule(dynamic, 64)
    for (t = 0; t < (int64_t)M; t++) {
        float ax = (float)((double)pe4[4 * t] - origin[0]);
        float ay = (float)((double)pe4[4 * t + 1] - origin[1]);
        float az = (float)((double)pe4[4 * t + 2] - origin[2]);
        int i = (int)floor(ax * invh), j = (int)floor(ay * invh), k = (int)floor(az * invh);
        double px = ax * invh, py = ay * invh, pz = az * invh, s = 0;
        int i0 = px > i + 0.5 ? i - K : i - K - 1, i1 = px > i + 0.5 ? i + K + 1 : i + K;
        int j0 = py > j + 0.5 ? j - K : j - K - 1, j1 = py > j + 0.5 ? j + K + 1 : j + K;
        int k0 = pz > k + 0.5 ? k - K : k - K - 1, k1 = pz > k + 0.5 ? k + K + 1 : k + K;
        int ii, jj, kk;
        for (kk = k0; kk <= k1; kk++)
            for (jj = j0; jj <= j1; jj++)
                for (ii = i0; ii <= i1; ii++) {
                    uint32_t q, p;
                    if (!(kk >= 0 && kk < n && jj >= 0 && jj < n && ii >= 0 && ii < n)) continue;
                    q = (uint32_t)(((int64_t)kk * n + jj) * n + ii);
                    for (p = start[q]; p < end[q]; p++) {
                        float bx, by, bz;
                        double dx, dy, dz, r1, r2;
                        if (p >= N) continue;
                        bx = (float)((double)pm4[4 * (size_t)p] - origin[0]);
                        by = (float)((double)pm4[4 * (size_t)p + 1] - origin[1]);
                        bz = (float)((double)pm4[4 * (size_t)p + 2] - origin[2]);
                        if (!((double)fabsf(bx - ax) < R && (double)fabsf(by - ay) < R && (double)fabsf(bz - az) < R))
                            continue;
                        dx = (double)(float)(bx - ax);
                        dy = (double)(float)(by - ay);
                        dz = (double)(float)(bz - az);
                        r1 = sqrt(dx * dx + dy * dy + dz * dz);
                        r2 = r1 > a ? 12.5663706144 * r1 : (r1 < 1e-12 ? 1e+14 : 12.5663706144 * a);
                        s += G / r2 * mass[p];
                    }
                }
        pf[t] = pf[t] + s;
    }
}
