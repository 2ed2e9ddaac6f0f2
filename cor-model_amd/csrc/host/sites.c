/*
 * sites.c -- station geometry helpers for batch callers.
 *
 * The reference takes a geocentric station (radius in Re, geocentric latitude) and an
 * arrival direction in GEO spherical angles (src/io.c:7-11).  Site lists given in
 * geodetic coordinates (WGS84) and directions given as zenith/azimuth in the local
 * frame are converted here in FP64; nothing here has a counterpart in the reference.
 */
#include <math.h>
#include "cor_host.h"

#define WGS84_A_KM 6378.137
#define WGS84_F (1.0 / 298.257223563)
#define RE_KM 6371.2

int cor_geodetic_to_geocentric(double lat_deg, double h_km, double *r0_re, double *lat_c_deg) {
    if (!r0_re || !lat_c_deg || !isfinite(lat_deg) || !isfinite(h_km) || fabs(lat_deg) > 90.) return COR_E_INVALID;
    const double e2 = WGS84_F * (2. - WGS84_F);
    const double p = lat_deg * (3.14159265358979323846 / 180.);
    const double sp = sin(p), cp = cos(p);
    const double n = WGS84_A_KM / sqrt(1. - e2 * sp * sp);
    const double x = (n + h_km) * cp;
    const double z = (n * (1. - e2) + h_km) * sp;
    *r0_re = sqrt(x * x + z * z) / RE_KM;
    *lat_c_deg = atan2(z, x) * (180. / 3.14159265358979323846);
    return COR_OK;
}

/* Arrival direction from zenith angle and azimuth (clockwise from geographic north)
 * in the local frame at geocentric (lat, lon): the unit vector pointing up the
 * arrival direction, returned as GEO latitude/longitude as the reference's the1/fi1
 * expect.  Zenith 0 gives the station's own (lat, lon). */
int cor_direction_to_geo(double lat_deg, double lon_deg, double zenith_deg, double azimuth_deg,
                         double *the1_deg, double *fi1_deg) {
    if (!the1_deg || !fi1_deg) return COR_E_INVALID;
    const double d2r = 3.14159265358979323846 / 180.;
    const double la = lat_deg * d2r, lo = lon_deg * d2r, ze = zenith_deg * d2r, az = azimuth_deg * d2r;
    const double up[3] = {cos(la) * cos(lo), cos(la) * sin(lo), sin(la)};
    const double east[3] = {-sin(lo), cos(lo), 0.};
    const double north[3] = {-sin(la) * cos(lo), -sin(la) * sin(lo), cos(la)};
    const double cz = cos(ze), sz = sin(ze), ca = cos(az), sa = sin(az);
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = cz * up[k] + sz * (ca * north[k] + sa * east[k]);
    const double rho = sqrt(v[0] * v[0] + v[1] * v[1]);
    *the1_deg = atan2(v[2], rho) / d2r;
    double f = atan2(v[1], v[0]) / d2r;
    if (f < 0.) f += 360.;
    *fi1_deg = f;
    return COR_OK;
}

/* Penumbra refinement: the scan that re-traces [Rl - del, Ru] of a finished scan at
 * del / factor, clamped to the coarse scan's range (a reference run of the same
 * station with that rig, rk and del traces the same rigidities). */
int cor_refine_scan(const cor_scan *coarse, const cor_cutoff *cut, int32_t factor, cor_scan *fine) {
    if (!coarse || !cut || !fine || factor < 1 || !(coarse->del > 0.) || !isfinite(cut->rl) ||
        !isfinite(cut->ru))
        return COR_E_INVALID;
    double lo = cut->rl - coarse->del;
    if (lo < coarse->rig) lo = coarse->rig;
    double hi = cut->ru;
    if (hi > coarse->rk) hi = coarse->rk;
    if (hi < lo) hi = lo;
    *fine = *coarse;
    fine->rig = lo;
    fine->rk = hi;
    fine->del = coarse->del / factor;
    return COR_OK;
}
