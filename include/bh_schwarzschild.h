/*
 * bh_schwarzschild.h -- C++ entry point of the Schwarzschild lookup table,
 * with the signature and ownership the reference's tests link against
 * (tests/bh_sanity_test.cu:12-15; schwarzschild_blackhole.cu:142).
 */
#ifndef BH_SCHWARZSCHILD_API_H
#define BH_SCHWARZSCHILD_API_H

#include "bh_types.h"

#ifdef __cplusplus
/* New table with caller-owned new[] arrays, or nullptr (message on stderr). */
GeodesicLookupTable* precomputeGeodesics(float mass, int num_rays);
#endif

#endif /* BH_SCHWARZSCHILD_API_H */
