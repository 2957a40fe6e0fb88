/*
 * prg_synth.h - seeded synthetic oriented point clouds for tests and
 * benchmarks.  No counterpart in the reference, which only reads files
 * (PointStream.inl); these produce the benchmark workloads in memory with a
 * fully specified RNG (splitmix64 + Box-Muller), so the oracle and the GPU
 * see identical bytes.  Outputs are float xyz[3n] and unit normals nxyz[3n].
 */
#ifndef PRG_SYNTH_H
#define PRG_SYNTH_H
#include <stdint.h>
#include "prg.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Area-uniform points on the sphere of radius 0.4 centred at (0.5,0.5,0.5):
 * g ~ N(0, I3), p = c + 0.4 g/|g|, n = g/|g| (exact radial normals). */
prg_status prg_synth_sphere(int64_t n, uint64_t seed, float* xyz, float* nxyz);

/* Torus R=0.30, r=0.12 centred in the unit cube, area-uniform by rejection
 * (accept with probability (R + r cos phi)/(R + r)), Gaussian position noise
 * sigma=0.002, analytic normals tilted by a 0.05-rad Gaussian angle. */
prg_status prg_synth_torus(int64_t n, uint64_t seed, float* xyz, float* nxyz);

/* Union of `shapes` seeded superquadrics (exponents U[0.3,2.5], semi-axes
 * U[0.1,0.3], random rotations, centres U[0.3,0.7]^3), area-proportional
 * samples on the parts of each surface outside the other shapes, position
 * noise `sigma`, analytic normals. */
prg_status prg_synth_superquadrics(int64_t n, int32_t shapes, float sigma, uint64_t seed, float* xyz, float* nxyz);

#ifdef __cplusplus
}
#endif
#endif
