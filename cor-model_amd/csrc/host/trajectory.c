/*
 * trajectory.c -- trajectory_simulation_cutoff of the reference API
 * (src/trajectory.c:6-161): the whole rigidity scan is traced by the GPU tracer in
 * FP64 and the result is printed exactly as the reference prints it (one line per
 * allowed rigidity in ascending order, then the cutoff summary).
 */
#include <stdlib.h>
#include "cor_host.h"

void cor_write_cutoff_lines(FILE *out, const cor_scan *scan, const cor_traj *traj, int64_t n,
                            const cor_cutoff *cut) {
    for (int64_t i = 0; i < n; i++) {
        const cor_traj *t = &traj[i];
        if (t->status != COR_ALLOWED) continue;
        const double rig = scan->rig + (int)i * scan->del;   /* src/trajectory.c:14 */
        fprintf(out, "%10.6f%15.10f%12.6f%10.3f%10.3f%10.3f%10.3f%12.6f%16.2f\n", rig, t->vc, t->r, t->td,
                t->fei, t->ast, t->asf, t->time, t->length_m / 1000.);
    }
    fprintf(out, "  CUTOFF with rigidities P(S),P(C),P(M) are:\n%12.5f%12.5f%12.5f\n\n", cut->rl, cut->ru, cut->rc);
}

void trajectory_simulation_cutoff(InfileHeader *header, const Model **models, ModelsParams *models_params,
                                  FILE *outfile, uint64_t step_limit, double tu_angle,
                                  double particle_mass, uint32_t parallel) {
    (void)parallel;
    cor_model_set set;
    const cor_scan scan = {header->r0, header->the0, header->fi0, header->the1, header->fi1,
                           header->rig, header->del, header->rk, header->zn, header->nk1, 0};
    const cor_trace_opts opts = {step_limit, tu_angle, particle_mass, COR_PREC_FP64, 0, 0, 0, 0, 0, 0, 0};
    int rc = cor_models_to_set(models, &set);
    const int64_t n = cor_scan_count(&scan);
    if (rc == COR_OK && n < 0) rc = COR_E_INVALID;
    cor_traj *traj = NULL;
    cor_cutoff cut;
    if (rc == COR_OK) {
        traj = (cor_traj *)calloc((size_t)(n > 0 ? n : 1), sizeof(cor_traj));
        rc = traj ? cor_gpu_trace(&set, models_params, &scan, 1, &opts, traj, &cut) : COR_E_NOMEM;
    }
    if (rc != COR_OK) {
        fprintf(stderr, "Error: GPU tracer failed: %s (%s)\n", cor_gpu_error_string(rc), cor_gpu_last_hip_error());
        free(traj);
        return;
    }
    cor_write_cutoff_lines(outfile, &scan, traj, n, &cut);
    free(traj);
}

/* src/trajectory.c:163-273 (the trajectory mode): the tracks of the header's scan,
 * traced on the GPU in FP64, one line per accepted step that does not end the
 * trajectory, to tragsm (GSM position [m]) and trasph (radius [Re], latitude,
 * longitude [deg]), each followed by speed, flight time and path length, in the
 * reference's formats. */
void trajectory_simulation(InfileHeader *header, const Model **models, ModelsParams *models_params, FILE *tragsm,
                           FILE *trasph, uint64_t step_limit, double tu_angle, double particle_mass) {
    cor_model_set set;
    const cor_scan scan = {header->r0, header->the0, header->fi0, header->the1, header->fi1,
                           header->rig, header->del, header->rk, header->zn, header->nk1, 0};
    const cor_trace_opts opts = {step_limit, tu_angle, particle_mass, COR_PREC_FP64, 0, 0, 0, 0, 0, 0, 0};
    int rc = cor_models_to_set(models, &set);
    const int64_t n = cor_scan_count(&scan);
    if (rc == COR_OK && n < 0) rc = COR_E_INVALID;
    int64_t total = 0;
    int64_t *counts = NULL;
    cor_track_point *pts = NULL;
    if (rc == COR_OK) {
        counts = (int64_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t));
        rc = counts ? cor_gpu_trace_tracks(&set, models_params, &scan, 1, &opts, counts, &total, NULL, 0)
                    : COR_E_NOMEM;
    }
    if (rc == COR_OK && total > 0) {
        pts = (cor_track_point *)malloc(sizeof(cor_track_point) * (size_t)total);
        rc = pts ? cor_gpu_trace_tracks(&set, models_params, &scan, 1, &opts, counts, &total, pts, total)
                 : COR_E_NOMEM;
    }
    if (rc != COR_OK) {
        fprintf(stderr, "Error: GPU tracer failed: %s (%s)\n", cor_gpu_error_string(rc), cor_gpu_last_hip_error());
    } else {
        for (int64_t k = 0; k < total; k++) {
            const cor_track_point *t = &pts[k];
            fprintf(tragsm, "%6d%17.5f%17.5f%17.5f%14.2f%15.10f%18.2f\n", (int)t->step, t->gsm[0], t->gsm[1],
                    t->gsm[2], t->v, t->time, t->length_m);
            fprintf(trasph, "%6d%17.5f%17.5f%17.5f%14.2f%15.10f%18.2f\n", (int)t->step, t->r, t->td, t->fei, t->v,
                    t->time, t->length_m);
        }
    }
    free(pts);
    free(counts);
}

/* src/trajectory.c:275-356: one RK step with the frozen field, on the GPU (the
 * tracer's FP64 step, so the same bits as a step inside the tracer). */
void calculate_position_and_speed(double h, double d, Vector geo_field, Vector *car_vel, Vector *geo_pos) {
    const double f[3] = {geo_field.x, geo_field.y, geo_field.z};
    double v[3] = {car_vel->x, car_vel->y, car_vel->z};
    double x[3] = {geo_pos->x, geo_pos->y, geo_pos->z};
    const int rc = cor_gpu_rk_step(&h, &d, f, v, x, 1);
    if (rc != COR_OK) {
        fprintf(stderr, "Error: GPU step failed: %s (%s)\n", cor_gpu_error_string(rc), cor_gpu_last_hip_error());
        return;
    }
    car_vel->x = v[0];
    car_vel->y = v[1];
    car_vel->z = v[2];
    geo_pos->x = x[0];
    geo_pos->y = x[1];
    geo_pos->z = x[2];
}
