/*
 * cor_host.h -- internal declarations of the host C library (the public API is
 * include/cor.h and include/cor_gpu.h).
 */
#ifndef COR_HOST_H
#define COR_HOST_H

#include "cor.h"
#include "cor_gpu.h"

#define COR_IGRF14_NROWS 54   /* lib/internal/coefs/igrf_14_coefs.h:90 */
#define COR_TO_DEG 57.295779513   /* lib/internal/internal.h:6 */

extern const double cor_igrf14_rows[COR_IGRF14_NROWS][INTERNAL_COEFS_SIZE];

/* CHAOS-7.16 (lib/chaos/coefs/chaos_7_16.h:384-388: nm_max 186): the triangle m <= n
 * of G and H, row after row (chaos716_table.c) */
#define COR_CHAOS716_NM 186
#define COR_CHAOS716_NTRI (COR_CHAOS716_NM * (COR_CHAOS716_NM + 1) / 2)
extern const double cor_chaos716_tri[2][COR_CHAOS716_NTRI];

/* Fill a cor_model_set from a NULL-terminated Model list (src/main.c:268-299 order). */
int cor_models_to_set(const Model **models, cor_model_set *set);

#endif /* COR_HOST_H */
