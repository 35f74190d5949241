/*
 * pppm_particle_mesh.h -- the reference's particle-mesh free functions of the
 * Biot-Savart path, same names and arguments (paths relative to the
 * reference repository: hybridFastSummation/hybridFastSummation/ParticleMesh.h,
 * kernels in GPU_kernel.cu).  FP64, device pointers, enqueued on the null
 * stream like the reference.  Arrays "sorted" are in the hash order of
 * start/end (ParticleManager / calcHash + sortParticles + reorder).
 *
 * C++ callers only (the arguments are HIP vector types).
 */
#ifndef PPPM_PARTICLE_MESH_H
#define PPPM_PARTICLE_MESH_H
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" {

/* ParticleMesh.h:9-18, GPU_kernel.cu:218-297: gridval[c] = h^-3 * sum of mass
 * over the sorted particles within +-2 hash cells of c whose box weight about
 * c's centre is one. */
void ParticleToMesh(uint32_t *start, uint32_t *end, float4 *pos, double *mass, double cell_h, double *gridval,
                    uint3 griddim, uint3 hashdim, uint32_t num_particle, double4 origin);

/* ParticleMesh.h:40, GPU_kernel.cu:524-543: gridval *= h_sqr * G. */
void ComputeRHS(double *gridval, double h_sqr, double G, uint32_t num_cell);

/* ParticleMesh.h:139-148, GPU_kernel.cu:1750-1805: (u, v, w) = curl F with
 * central differences, indices clamped to the grid. */
void getCurl(double *Fx, double *Fy, double *Fz, double *u, double *v, double *w, uint32_t dimx, uint32_t dimy,
             uint32_t dimz, double h);

/* ParticleMesh.h:150-157, GPU_kernel.cu:1809-1870: trilinear (hat-weighted)
 * interpolation of mesh_value at each particle over the in-grid cells. */
void MeshToParticle(float4 *pos, double *mesh_value, double *particle_value, double h, uint3 gdim, uint3 hdim,
                    uint32_t num_particles, double4 origin);

/* ParticleMesh.h:194-216, GPU_kernel.cu:2262-2485: near-field correction of
 * the sorted evaluation points against the sorted vortices, added to u, v, w
 * (in/out; u_pm, v_pm, w_pm are unused, as in the reference).  Core radius
 * 0.08, box |d|inf < (K + 1/2) h over hash cells +-(K+1); is_segment = 1
 * also skips vortices within 1e-12 of evalposb.  K < 0: nothing. */
void BiotSavartPPCorrScaleMN(uint32_t *start, uint32_t *end, float4 *evalposa, float4 *evalposb, uint32_t is_segment,
                             float4 *vortpos, double *vort_x, double *vort_y, double *vort_z, double *u, double *v,
                             double *w, double *u_pm, double *v_pm, double *w_pm, double cellh, uint3 gdim,
                             uint3 hdim, int K, uint32_t M, uint32_t N, double4 origin);

/* ParticleMesh.h:281-286, GPU_kernel.cu:2487-2734: far-field buffer
 * u = g_u - sum_{D != 0, |D|inf <= K} dir[D] x vrho[c+D] h^3 - curl(0.250580605913965 h^2 vrho),
 * with g_u the grid velocity on entry.  As in the reference, psi_* and g_*
 * are overwritten (psi = 0.2505806 h^2 vrho, g = its curl).  K < 0: nothing. */
void BiotSavartComputeFarField(uint32_t *start, uint32_t *end, float4 *pos, double *vrho_x, double *vrho_y,
                               double *vrho_z, double *psi_x, double *psi_y, double *psi_z, double *g_u, double *g_v,
                               double *g_w, double *u, double *v, double *w, double cell_h, uint3 gdim, uint3 hdim,
                               int K, uint32_t num_particles, double4 origin);

/* ParticleMesh.h:288-294, GPU_kernel.cu:2886-2976: trilinear interpolation
 * of the far-field buffer at each point whose cell is inside the grid
 * (other points' outputs are left untouched). */
void BiotSavartInterpolateFarField(float4 *pos, double *far_u, double *far_v, double *far_w, double *p_u, double *p_v,
                                   double *p_w, double cell_h, uint32_t dimx, uint32_t dimy, uint32_t dimz,
                                   uint32_t num_particles, double4 origin);
/* Time step (VortexSimulationSystemImp.cpp:90-125), ParticleMesh.h:254-309:
 * ComputePosb GPU_kernel.cu:3182-3235, move_particle 3118-3140 (w = 0),
 * ComputePosa 3237-3254, ComputeVortex 3143-3180, VectorAdd 3442-3460,
 * AddURotation 3589-3606.  Same arithmetic as the pppm_* entry points of the
 * same names in pppm.h. */
void ComputePosb(float4 *posa, float4 *posc, float4 *posb, double *vortx, double *vorty, double *vortz, double *k,
                 double h, uint32_t num_particles);
void move_particle(float4 *pos, double dt, double4 *U, uint32_t num_particles);
void ComputePosa(float4 *posc, float4 *posb, float4 *posa, uint32_t num_particles);
void ComputeVortex(float4 *posa, float4 *posb, double *vortx, double *vorty, double *vortz, double *dvortx,
                   double *dvorty, double *dvortz, double *k, uint32_t num_particles);
void VectorAdd(double *A, double *B, double *C, double a, double b, uint32_t num);
void AddURotation(double4 *U, double *u, double *v, double *w, uint32_t num);

/* ParticleMesh.h:264-279, GPU_kernel.cu:3327-3438: o[t] = Gaussian-weighted
 * mean (weight e^{-d^2/(2 nu^2)}) of the sorted i over the sorted particles
 * of the hash cells within +-2 of pos[t]'s cell; dt is unused, as in the
 * reference. */
void ParticleGaussianSmooth(uint32_t *start, uint32_t *end, float4 *pos, float4 *pos_sort, double nu, double dt,
                            double *o_vortx, double *o_vorty, double *o_vortz, double *i_vortx, double *i_vorty,
                            double *i_vortz, double h, uint3 gdim, uint32_t N, double4 origin);
/* Baroclinic source, ParticleMesh.h (AssignDensity, GradDensity,
 * ParticleGetBaroclinic), GPU_kernel.cu:3655-3898: grid density from the
 * sorted smoke over start/end tables (cubic grid, hash grid = density grid),
 * its periodic central differences, and w += dt_beta * 0.1 * (grad rho x
 * (0, 10, 0)) at each vortex.  vort_size, hashdim, dimy and dimz are unused
 * (cubic grids), as in the reference. */
void AssignDensity(uint32_t *start, uint32_t *end, float4 *pos, float *rho, float *vort_size, double cell_h,
                   float *gridval, uint3 griddim, uint3 hashdim, uint32_t num_particle, double4 origin);
void GradDensity(float *phi, float3 *dphi, double cell_h, uint32_t dimx, uint32_t dimy, uint32_t dimz);
void ParticleGetBaroclinic(float4 *vortex_pos, double *vort_x, double *vort_y, double *vort_z, float3 *grad_rho,
                           double dt_beta, uint32_t num_vort, uint32_t dimx, uint32_t dimy, uint32_t dimz,
                           double inv_h, double4 origin);
/* Potential-flow gradient path (PotentialFieldSolver.cpp:441-739
 * computeFarFieldBuffer, evaluateGradient), ParticleMesh.h:43-120, 245;
 * kernels GPU_kernel.cu:560-610, 750-912, 964-987, 988-1200, 1520-1594,
 * 1678-1718, 2977-3117.  P2M, RHS and the FMG solves are ParticleToMesh,
 * ComputeRHS and MultiGridSolver3DPeriod_double above.
 *  ComputeGradient: dphi = direction * periodic central differences of phi.
 *  applyDirichlet: RHS -= M / (4 pi r) at each boundary face's ghost centre.
 *  ComputeDiagPhi: phi = 0.250580605913965 gridval h^2 G.
 *  PotentialComputeFarField: pforce = gforce - sum_{D != 0} dir[D] G direction
 *    rho h^3 - grad(diagonal phi); phi and gforce are overwritten.
 *  PotentialInterpolateFarField: hat-weighted far field at in-grid points.
 *  Potential_PPCorrMN: pforce += G direction d m / (4 pi max(r, 0.05)^3) over
 *    masses in the box |d|inf < (K + 1/2) h.
 *  PotentialComputeGradForOutParticle: zero gradient outside [bbmin, bbmax]
 *    (the reference computes the monopole there and stores zero). */
void ComputeGradient(double *phi, double3 *dphi, double cell_h, double direction, uint32_t dimx, uint32_t dimy,
                     uint32_t dimz);
void applyDirichlet(double *RHS, double4 center, double monopole, double4 grid_coner, double h, uint32_t dimx,
                    uint32_t dimy, uint32_t dimz);
void ComputeDiagPhi(double *gridval, double *phi, double h_sqr, double G, uint32_t num_cell);
void PotentialComputeFarField(uint32_t *start, uint32_t *end, float4 *pos, double *mass, double *gridval, double *phi,
                              double3 *gforce, double3 *pforce, double cell_h, double G, double direction,
                              uint3 gdim, uint3 hdim, int K, uint32_t num_particles, double4 origin);
void PotentialInterpolateFarField(float4 *pos, double3 *far_force, double3 *p_force, double cell_h, uint32_t dimx,
                                  uint32_t dimy, uint32_t dimz, uint32_t num_particles, double4 origin);
void Potential_PPCorrMN(uint32_t *start, uint32_t *end, float4 *pos_eval, float4 *pos_mass, double *mass,
                        double3 *pforce, double invh, double h, double G, double direction, uint3 gdim, uint3 hdim,
                        uint2 gstride, uint2 hstride, int K, uint32_t M, uint32_t N, double4 origin);
void PotentialComputeGradForOutParticle(float4 *pos, double total_mass, double3 mass_center, float3 bbmin,
                                        float3 bbmax, double G, double direction, double3 *p_force, uint32_t M);
/* The scalar potential (PotentialFieldSolver.cpp:695-739,
 * computeFarFieldPotentialBuffer and evaluateScalarPotential; kernels
 * GPU_kernel.cu:1201-1300, 1301-1462, 1595-1676):
 *  PotentialComputeScalarFarField: phi -= sum_{|D|inf <= K} w(D) G direction
 *    rho h^2, w = 1/4 at D = 0, else 1 / (4 pi |D|) in cells.
 *  PotentialInterpolateFarFieldScalar: hat-weighted phi at in-grid points.
 *  Potential_PPCorrMNScalar: p_phi += G direction m / (4 pi max(r, 0.125)). */
void PotentialComputeScalarFarField(double *mass, double *gridval, double *phi, double cell_h, double G,
                                    double direction, uint3 gdim, uint3 hdim, int K);
void PotentialInterpolateFarFieldScalar(float4 *pos, double *far_force, double *p_force, double cell_h,
                                        uint32_t dimx, uint32_t dimy, uint32_t dimz, uint32_t num_particles,
                                        double4 origin);
void Potential_PPCorrMNScalar(uint32_t *start, uint32_t *end, float4 *pos_eval, float4 *pos_mass, double *mass,
                              double *p_phi, double invh, double h, double G, double direction, uint3 gdim,
                              uint3 hdim, uint2 gstride, uint2 hstride, int K, uint32_t M, uint32_t N,
                              double4 origin);
}

#endif
