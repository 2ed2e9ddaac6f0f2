/*
 * io.c -- the reference's input parsing and output headers (src/io.c:4-127).  The
 * formats are the reference's, so files and outputs are interchangeable.
 */
#include <string.h>
#include "cor_host.h"

/* src/io.c:4-20 */
int read_cutoff_infile_header(FILE *input_file, InfileHeader *h) {
    const int items = fscanf(input_file, "%lf %lf %lf\n%lf %lf %lf\n%lf %lf\n%d %d %d %d %d %d %d\n%d %d %d %lf",
                             &h->rig, &h->zn, &h->rk, &h->r0, &h->the0, &h->fi0, &h->the1, &h->fi1,
                             &h->iy, &h->mes, &h->ide, &h->id, &h->ih, &h->min, &h->is, &h->nk1,
                             &h->iopt, &h->ist, &h->del);
    return items < 0 ? -1 : 0;
}

/* src/io.c:22-50 */
int write_cutoff_outfile_header(FILE *out, InfileHeader *h, const Model **models, uint64_t steps_count) {
    if (fprintf(out, "\n\n\n                ASYMPTOTIC COORDINATES\n   calculated by model(s): ") < 0) return -1;
    const char *pad = "";
    for (int i = 0; models[i]; i++) {
        if (fprintf(out, "%s%s", pad, models[i]->name) < 0) return -1;
        pad = ", ";
    }
    const int res = fprintf(out,
                            "\n Station with geo.latitude:%9.3f  & longitude:%9.3f & radius : %8.5f\n"
                            " Direction of trajectory with latitude: %9.3f & longitude: %9.3f\n"
                            " Date: %4d %2d %2d  time: %2d hod %2d min %2d sec\n"
                            " Starting rigidity : %8.4f GV Epsilon=%7.4f\n Limit of total number of steps : %lu \n\n"
                            " rig : v : rad : eth : efi : ath : afi : time : length\n",
                            h->the0, h->fi0, h->r0, h->the1, h->fi1, h->iy, h->mes, h->ide, h->ih, h->min,
                            h->is, h->rig, h->del, (unsigned long)steps_count);
    return res < 0 ? -1 : 0;
}

/* src/io.c:52-67 */
int read_field_infile_header(FILE *input_file, InfileHeader *h) {
    const int items = fscanf(input_file, "%lf %lf %lf\n%lf %lf %lf %lf\n%d %d %d %d %d %d %d\n", &h->r0,
                             &h->the0, &h->fi0, &h->lat_step, &h->lon_step, &h->the1, &h->fi1, &h->iy,
                             &h->mes, &h->ide, &h->id, &h->ih, &h->min, &h->is);
    return items < 0 ? -1 : 0;
}

/* src/io.c:69-96 */
int write_field_outfile_header(FILE *out, InfileHeader *h, const Model **models) {
    if (fprintf(out, "\n\n\n                GEOMAGNETIC FIELD\n Geomagnetic Model(s): ") < 0) return -1;
    const char *pad = "";
    for (int i = 0; models[i]; i++) {
        if (fprintf(out, "%s%s", pad, models[i]->name) < 0) return -1;
        pad = ", ";
    }
    const int items = fprintf(out,
                              "\n lat step: %.3f, lon step: %.3f\n"
                              " Date: %4d %2d %2d  time: %2d h %2d min %2d sec\n\n"
                              " lat : lon : B [nT] : BXGEO [nT] : BYGEO [nT] : BZGEO [nT]\n",
                              h->lat_step, h->lon_step, h->iy, h->mes, h->ide, h->ih, h->min, h->is);
    return items < 0 ? -1 : 0;
}

/* src/io.c:98-113: Dst first, then Pdyn, By, Bz, W1..W6 */
int read_ts05_params(FILE *input_file, ModelsParams *p) {
    const int items = fscanf(input_file, "%lf %lf %lf %lf\n%lf %lf %lf %lf %lf %lf\n", &p->PARMOD[1],
                             &p->PARMOD[0], &p->PARMOD[2], &p->PARMOD[3], &p->PARMOD[4], &p->PARMOD[5],
                             &p->PARMOD[6], &p->PARMOD[7], &p->PARMOD[8], &p->PARMOD[9]);
    return items != 10 ? -1 : 0;
}

/* src/io.c:115-127: Dst first, then Pdyn, By, Bz */
int read_t96_params(FILE *input_file, ModelsParams *p) {
    const int items = fscanf(input_file, "%lf %lf %lf %lf\n", &p->PARMOD[1], &p->PARMOD[0], &p->PARMOD[2],
                             &p->PARMOD[3]);
    return items != 4 ? -1 : 0;
}
