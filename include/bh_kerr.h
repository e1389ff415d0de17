/*
 * bh_kerr.h -- FP64 ray state of the drop-in interface.
 *
 * The reference's include/bh_kerr.h defines GeodesicStateD (:553-556) after
 * including bh_common.h and bh_types.h (:48-49); sources written against it
 * include this header for the double-precision state and get the rest of
 * the drop-in types with it.  Member names and order match, so the
 * structure is layout-compatible (8 doubles, 64 bytes).  The reference's
 * per-ray device functions in that header are replaced by the kernels of
 * libnullpath.so (include/nullpath.h).
 */
#ifndef BH_KERR_H
#define BH_KERR_H

#include "bh_common.h"
#include "bh_types.h"

/* FP64 photon state: position (t, r, theta, phi) and covariant momentum. */
struct GeodesicStateD {
    double t, r, theta, phi;
    double pt, pr, ptheta, pphi;
};

#endif /* BH_KERR_H */
