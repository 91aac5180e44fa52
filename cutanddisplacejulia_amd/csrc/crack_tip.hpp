// Crack-tip (free) edges of a triangle mesh and the stress intensities K_I..K_III at them, estimated from the
// displacement discontinuity of the tip triangles (GetCrackTipElements3D.jl, StressIntensity3D.jl).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdj {

// Geometry columns per edge block (n x CT_GEOM, column-major): FeLe, FeMd xyz, FeEv xyz, FeM2Ev xyz, FeM2ELe,
// IntAng, Area.  K columns per edge block (n x CT_K): K1, K2, K3, StrainEnergy.  Edge blocks are P1P2, P1P3,
// P2P3 (third corner P3, P2, P1); flags are n x 3 bytes in the same order.
constexpr int CT_GEOM = 13;
constexpr int CT_K = 4;

// Hash-table slots for 3n edges (a power of two, at least twice the edge count).
uint64_t crack_tip_slots(int64_t n);

// Free-edge flags (1 = free) and the edge geometry for all three edge blocks.  slots: crack_tip_slots(n) ints,
// matched: 3n ints (both scratch).
hipError_t crack_tip_elements(const double* mid, const double* P1, const double* P2, const double* P3,
                              const double* nrm, int64_t n, int* slots, int* matched, uint8_t* flags, double* geom,
                              hipStream_t st);

// K1, K2, K3 and StrainEnergy of every free edge (NaN elsewhere) from the flags and geometry above.
hipError_t stress_intensity(const double* Dn, const double* Dss, const double* Dds, int64_t n, double G, double nu,
                            const double* nrm, const uint8_t* flags, const double* geom, double* K, hipStream_t st);

}  // namespace cdj
