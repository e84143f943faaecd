/*
 * libh3d host-side model builder: the inputs the reference's q3d() builds
 * before it calls main() (grid pitch, Z metrics, modelA voxelisation,
 * material properties, heat source, boundary-condition set), restated as
 * native code so that a caller without the Julia front end can produce
 * bit-identical inputs.  All arrays use the layout of include/h3d.h.
 *
 * Reference interfaces replaced (paths are into the kenoogl/H2 tree):
 *   dx = round(1.2e-3/NX, digits=8)     src/heat3ds.jl:197-200   -> h3d_model_pitch
 *   Zcoordinate.genZ!/Zcase2!           src/Zcoord.jl:88-166     -> h3d_model_genz
 *   modelA.fillID!                      src/modelA.jl:545-556    -> h3d_model_fill_id
 *   modelA.setProperties!               src/modelA.jl:499-542    -> h3d_model_properties
 *   HeatSrc!                            src/heat3ds.jl:55-64     -> h3d_model_heat_src
 *   set_mode3_bc_parameters             src/heat3ds.jl:27-46     -> h3d_model_mode3_bc
 * and one input the reference does not have:
 *   h3d_model_power_map: a seeded synthetic per-cell power map.
 */
#ifndef H3D_MODEL_H
#define H3D_MODEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* round(1.2e-3 / n, digits=8) */
int h3d_model_pitch(int n, double *d);

/* Z faces/centres/widths for nk = 30*r cells.  r == 1 is Zcase2! exactly;
 * r > 1 splits each Zcase2! cell into r equal cells.  Z has nk+3 entries,
 * ZC and dZ nk+2 (genZ! layout).  Returns H3D_ERR_ARG (-1) otherwise. */
int h3d_model_genz(int nk, double *Z, double *ZC, double *dZ);

/* fillID!: voxelise the modelA IC stack into material ids 1..7 over the
 * whole (MX,MY,MZ) array (halo included).  ID must be zero on entry.
 * nthreads <= 0 uses the hardware concurrency. */
int h3d_model_fill_id(uint8_t *ID, int MX, int MY, int MZ, double dx, double dy, const double *Z,
                      int nthreads);

/* setProperties!: λ, ρ, cp from the ids (n = MX*MY*MZ). */
int h3d_model_properties(double *lam, double *rho, double *cp, const uint8_t *ID, size_t n);

/* HeatSrc!: hsrc = q_src on interior cells whose id is 7 (power grid);
 * the reference uses q_src = Q_src = 1.6e11 W/m^3. */
int h3d_model_heat_src(double *hsrc, const uint8_t *ID, int MX, int MY, int MZ, double q_src);

/* set_mode3_bc_parameters: x±, y±, z+ convective (h = 5 W/m²K, θ∞ = 300 K),
 * z- isothermal 300 K; face order x-,x+,y-,y+,z-,z+. */
int h3d_model_mode3_bc(int type[6], double temperature[6], double heat_flux[6], double htc[6],
                       double ambient[6]);

/* Seeded synthetic power map on the power-grid cells (id 7): a log-normal
 * (mu = 0, sigma = 0.5) weight per cell times (1 + 4*sum of 16 Gaussian
 * hotspots of standard deviation 40 um whose centres are uniform over
 * [0.1, 1.1] mm squared), scaled so that sum(hsrc * dx*dy*dZ[k]) equals
 * total_watts.  Uses a counter-based splitmix64 stream, so the map depends
 * only on (seed, grid).  Writes hsrc on interior id-7 cells and 0 elsewhere. */
int h3d_model_power_map(double *hsrc, const uint8_t *ID, int MX, int MY, int MZ, double dx, double dy,
                        const double *dZ, uint64_t seed, double total_watts);

#ifdef __cplusplus
}
#endif
#endif
