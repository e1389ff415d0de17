/*
 * bh_types.h -- ray state and result types of the drop-in interface.
 *
 * Member names and order match the reference's include/bh_types.h:52-135
 * (SchwarzschildMetric), :163-251 (GeodesicState, 8 floats),
 * :281-357 (GeodesicLookupTable, caller-owned new[] arrays), so the
 * structures are layout-compatible with code built against them.
 * GeodesicStateD is in bh_kerr.h, as in the reference.
 */
#ifndef BH_TYPES_H
#define BH_TYPES_H

#include "bh_common.h"

struct SchwarzschildMetric {
    float mass;
    float Rs;

    BH_HD explicit SchwarzschildMetric(float m) : mass(m), Rs(BH_SCHWARZSCHILD_MULTIPLIER * m) {}

    BH_HD inline float g_tt(float r) const { return -(1.0f - Rs / r); }
    BH_HD inline float g_rr(float r) const { return 1.0f / fmaxf(BH_EPS, 1.0f - Rs / r); }
    BH_HD inline float g_theta_theta(float r) const { return r * r; }
    BH_HD inline float g_phi_phi(float r, float theta) const {
        float s = sinf(theta);
        return r * r * s * s;
    }
};

/* Photon state: position (t, r, theta, phi) and covariant momentum. */
struct GeodesicState {
    float t;
    float r;
    float theta;
    float phi;
    float pt;
    float pr;
    float ptheta;
    float pphi;

    BH_HD GeodesicState() : t(0), r(0), theta(0), phi(0), pt(0), pr(0), ptheta(0), pphi(0) {}
};

/* Result of precomputeGeodesics(); every array has num_entries elements and
 * is owned by the caller (delete[] each, then delete the table). */
struct GeodesicLookupTable {
    float* impact_parameters;
    float* deflection_angles;
    float* bending_angles;
    float* closest_approaches;
    int* ray_status;
    int num_entries;
    float mass;
};

#endif /* BH_TYPES_H */
