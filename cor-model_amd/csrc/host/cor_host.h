/*
 * cor_host.h -- internal declarations of the host C library (the public API is
 * include/cor.h and include/cor_gpu.h).
 */
#ifndef COR_HOST_H
#define COR_HOST_H

#include "cor.h"
#include "cor_gpu.h"

#define COR_TO_DEG 57.295779513   /* lib/internal/internal.h:6 */

/* CHAOS-7.16 (lib/chaos/coefs/chaos_7_16.h:384-388: nm_max 186): the triangle m <= n
 * of G and H, row after row (chaos716_table.c) */
#define COR_CHAOS716_NM 186
#define COR_CHAOS716_NTRI (COR_CHAOS716_NM * (COR_CHAOS716_NM + 1) / 2)
extern const double cor_chaos716_tri[2][COR_CHAOS716_NTRI];

/* Fill a cor_model_set from a NULL-terminated Model list (src/main.c:268-299 order). */
int cor_models_to_set(const Model **models, cor_model_set *set);

/* Set the message cor_gpu_last_hip_error returns (host-side failures). */
void cor_set_error(const char *msg);

/* The per-scan text of trajectory_simulation_cutoff (src/trajectory.c:117-121,
 * 156-160): one line per allowed rigidity in ascending order, then the summary. */
void cor_write_cutoff_lines(FILE *out, const cor_scan *scan, const cor_traj *traj, int64_t n,
                            const cor_cutoff *cut);

#endif /* COR_HOST_H */
