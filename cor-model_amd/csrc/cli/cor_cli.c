/*
 * cor_cli.c -- the `cor` command line on the GPU library.
 *
 * Same modes, flags, defaults, messages and output files as the reference's command
 * line (src/main.c:12-315 with the options of src/args.c:11-73): cutoff, field and
 * trajectory modes; -p, -o, -s, -u, -m, -t, -i, -e, -c; one INFILE.  The text it
 * writes is the library's, which is the reference's format byte for byte.
 *
 * Two additions, neither of which changes a single-file run:
 *   * several INFILEs in cutoff mode are traced together in one GPU batch (files that
 *     share the epoch and model inputs share one launch), each written to its own
 *     output: <INFILE>.out, or DIR/<basename>.out with -o DIR (an existing directory);
 *   * --precision mixed traces with the mixed-precision tracer (FP32 field, FP64
 *     state) instead of the FP64 path the reference API runs.
 */
#include <argp.h>
#include <errno.h>
#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "cor.h"
#include "cor_gpu.h"
#include "cor_host.h"

const char *argp_program_version = "cor-model 3.0 (MI355X tracer)";

enum { OPT_PRECISION = 1000 };

typedef struct {
    char **inputs;
    int n_inputs;
    char *output;
    char *internal, *external, *crustal;
    char *steps, *tu, *mode, *particle;
    int parallel;
    int precision;
} Cli;

static const struct argp_option cli_options[] = {
    {"par", 'p', 0, 0, "Parallel computation (the GPU always traces in parallel)", 0},
    {"output", 'o', "FILE", 0,
     "Write to FILE instead of standard output (required in trajectory mode; a directory "
     "with several cutoff INFILEs)", 0},
    {"steps-count", 's', "COUNT", 0, "Step limit of one trajectory (default: 25000)", 0},
    {"tu-angle", 'u', "RAD", 0, "Largest velocity turn of one step (default: 0.01 rad)", 0},
    {"mode", 'm', "MODE", 0, "cutoff, field or trajectory (default: cutoff)", 0},
    {"particle-type", 't', "TYPE", 0, "proton or electron (default: proton)", 0},
    {"precision", OPT_PRECISION, "PREC", 0, "cutoff mode: fp64 (default) or mixed", 0},
    {0, 0, 0, 0, "Field models:", 0},
    {"internal-field", 'i', "MODEL", 0, "igrf9 .. igrf14, hist, cals10k.2 (default: igrf14 when no model is named)", 0},
    {"external-field", 'e', "MODEL", 0, "t96 or ts05 (default: none)", 0},
    {"crustal-field", 'c', "MODEL", 0, "chaos7.16 (default: none)", 0},
    {0}};

static error_t cli_parse(int key, char *arg, struct argp_state *state) {
    Cli *c = state->input;
    switch (key) {
        case 'p': c->parallel = 1; break;
        case 'o': c->output = arg; break;
        case 's': c->steps = arg; break;
        case 'u': c->tu = arg; break;
        case 'm': c->mode = arg; break;
        case 't': c->particle = arg; break;
        case 'i': c->internal = arg; break;
        case 'e': c->external = arg; break;
        case 'c': c->crustal = arg; break;
        case OPT_PRECISION:
            if (strcmp(arg, "fp64") == 0) c->precision = COR_PREC_FP64;
            else if (strcmp(arg, "mixed") == 0) c->precision = COR_PREC_MIXED;
            else argp_error(state, "unknown precision '%s'", arg);
            break;
        case ARGP_KEY_ARG: c->inputs[c->n_inputs++] = arg; break;
        case ARGP_KEY_END:
            if (c->n_inputs < 1) argp_usage(state);
            break;
        default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

/* src/main.c:12-31: step limit, turn angle and particle mass */
static int trace_options(const Cli *c, uint64_t *steps, double *tu, double *mass) {
    *steps = 25000;
    if (c->steps && string_to_uint64(c->steps, steps) == -1) {
        fprintf(stderr, "Error: Invalid steps count '%s'\n", c->steps);
        return 1;
    }
    *tu = .01;
    if (c->tu && string_to_double(c->tu, tu) == -1) {
        fprintf(stderr, "Error: Invalid TU angle '%s'\n", c->tu);
        return 1;
    }
    *mass = strcmp(c->particle, "electron") == 0 ? DEF_HM0_E : DEF_HM0_P;
    return 0;
}

/* src/main.c:51-66: every model's setup reads its lines of the input */
static int setup_models(FILE *in, const InfileHeader *h, const Model **models, ModelsParams *p) {
    for (int i = 0; models[i]; i++) {
        if (!models[i]->setup_fun) continue;
        if (models[i]->setup_fun(in, h, p, models[i]->setup_data) != 0) {
            fprintf(stderr, "Error: Invalid input file\n");
            return 1;
        }
    }
    return 0;
}

typedef struct {
    const char *path;
    InfileHeader h;
    ModelsParams p;
    cor_scan scan;
    int64_t n, first;
    int group;
} CutoffInput;

/* src/main.c:33-49: header, its rigidity step, the models' lines */
static int read_cutoff_input(const char *path, const Model **models, CutoffInput *ci) {
    FILE *in = fopen(path, "r");
    if (!in) {
        perror("Error: Cannot open input file for reading");
        return 1;
    }
    memset(&ci->h, 0, sizeof ci->h);
    memset(&ci->p, 0, sizeof ci->p);
    if (read_cutoff_infile_header(in, &ci->h) != 0) {
        fprintf(stderr, "Error: Invalid input file\n");
        fclose(in);
        return 1;
    }
    if (ci->h.del <= 0.) {
        fprintf(stderr, "Error: Rigidity step size must be bigger than 0\n");
        fclose(in);
        return 1;
    }
    if (setup_models(in, &ci->h, models, &ci->p)) {
        fclose(in);
        return 1;
    }
    if (fclose(in) != 0) {
        perror("Error: Cannot close input file");
        return 1;
    }
    ci->path = path;
    const InfileHeader *h = &ci->h;
    const cor_scan s = {h->r0, h->the0, h->fi0, h->the1, h->fi1, h->rig, h->del, h->rk, h->zn, h->nk1, 0};
    ci->scan = s;
    ci->n = cor_scan_count(&ci->scan);
    return ci->n < 0 ? (fprintf(stderr, "Error: Invalid input file\n"), 1) : 0;
}

static FILE *open_output(const Cli *c, const char *input, int batch) {
    if (!batch) {
        if (!c->output) return stdout;
        FILE *f = fopen(c->output, "w");
        if (!f) perror("Error: Cannot open output file for writing");
        return f;
    }
    char name[4096];
    char base_buf[4096];
    snprintf(base_buf, sizeof base_buf, "%s", input);
    if (c->output)
        snprintf(name, sizeof name, "%s/%s.out", c->output, basename(base_buf));
    else
        snprintf(name, sizeof name, "%s.out", input);
    FILE *f = fopen(name, "w");
    if (!f) perror("Error: Cannot open output file for writing");
    return f;
}

/* src/main.c:12-93 for one or many input files: the scans of files with identical
 * epoch and model inputs are traced in one GPU launch, each file's text written as
 * trajectory_simulation_cutoff writes it. */
static int cutoff(const Cli *c, const Model **models) {
    uint64_t steps;
    double tu, mass;
    if (trace_options(c, &steps, &tu, &mass)) return 1;
    const int batch = c->n_inputs > 1;
    if (batch && c->output) {
        struct stat st;
        if (stat(c->output, &st) != 0 || !S_ISDIR(st.st_mode)) {
            fprintf(stderr, "Error: with several input files -o must name an existing directory\n");
            return 1;
        }
    }
    CutoffInput *in = calloc((size_t)c->n_inputs, sizeof *in);
    if (!in) return 1;
    int rc = 0;
    for (int k = 0; k < c->n_inputs && !rc; k++) rc = read_cutoff_input(c->inputs[k], models, &in[k]);
    if (rc) {
        free(in);
        return 1;
    }
    cor_model_set set;
    if (cor_models_to_set(models, &set) != COR_OK) {
        fprintf(stderr, "Error: %s\n", cor_gpu_last_hip_error());
        free(in);
        return 1;
    }
    /* groups: identical ModelsParams bytes (epoch tables, rotation, external inputs) */
    int n_groups = 0;
    for (int k = 0; k < c->n_inputs; k++) {
        in[k].group = -1;
        for (int j = 0; j < k; j++)
            if (memcmp(&in[j].p, &in[k].p, sizeof(ModelsParams)) == 0) {
                in[k].group = in[j].group;
                break;
            }
        if (in[k].group < 0) in[k].group = n_groups++;
    }
    const cor_trace_opts opts = {steps, tu, mass, c->precision, 1, 0, 0, 0, 0, 0, 0};
    for (int g = 0; g < n_groups && !rc; g++) {
        int ns = 0;
        int64_t total = 0;
        for (int k = 0; k < c->n_inputs; k++)
            if (in[k].group == g) {
                in[k].first = total;
                total += in[k].n;
                ns++;
            }
        cor_scan *scans = calloc((size_t)ns, sizeof *scans);
        cor_cutoff *cut = calloc((size_t)ns, sizeof *cut);
        cor_traj *traj = calloc((size_t)(total > 0 ? total : 1), sizeof *traj);
        const ModelsParams *params = NULL;
        int s = 0;
        for (int k = 0; k < c->n_inputs; k++)
            if (in[k].group == g) {
                scans[s++] = in[k].scan;
                params = &in[k].p;
            }
        int e = (scans && cut && traj) ? cor_gpu_trace(&set, params, scans, ns, &opts, traj, cut) : COR_E_NOMEM;
        if (e != COR_OK) {
            fprintf(stderr, "Error: GPU tracer failed: %s (%s)\n", cor_gpu_error_string(e), cor_gpu_last_hip_error());
            rc = 1;
        }
        s = 0;
        for (int k = 0; k < c->n_inputs && !rc; k++) {
            if (in[k].group != g) continue;
            FILE *out = open_output(c, in[k].path, batch);
            if (!out) {
                rc = 1;
                break;
            }
            if (write_cutoff_outfile_header(out, &in[k].h, models, steps) != 0) {
                fprintf(stderr, "Error: Cannot write output file header\n");
                rc = 1;
            } else {
                cor_write_cutoff_lines(out, &in[k].scan, traj + in[k].first, in[k].n, &cut[s]);
            }
            s++;
            if (out != stdout && fclose(out) != 0) {
                perror("Error: Cannot close output file");
                rc = 1;
            }
        }
        free(scans);
        free(cut);
        free(traj);
    }
    free(in);
    return rc;
}

/* src/main.c:95-141 */
static int field(const Cli *c, const Model **models) {
    FILE *in = fopen(c->inputs[0], "r");
    if (!in) {
        perror("Error: Cannot open input file for reading");
        return 1;
    }
    InfileHeader h;
    ModelsParams p;
    memset(&h, 0, sizeof h);
    memset(&p, 0, sizeof p);
    if (read_field_infile_header(in, &h) != 0) {
        fprintf(stderr, "Error: Invalid input file\n");
        fclose(in);
        return 1;
    }
    if (setup_models(in, &h, models, &p)) {
        fclose(in);
        return 1;
    }
    if (fclose(in) != 0) {
        perror("Error: Cannot close input file");
        return 1;
    }
    FILE *out = open_output(c, c->inputs[0], 0);
    if (!out) return 1;
    if (write_field_outfile_header(out, &h, models) != 0) {
        fprintf(stderr, "Error: Cannot write output file header\n");
        if (out != stdout) fclose(out);
        return 1;
    }
    field_simulation(&h, &p, models, out, (uint32_t)c->parallel);
    if (out != stdout && fclose(out) != 0) {
        perror("Error: Cannot close output file");
        return 1;
    }
    return 0;
}

/* src/main.c:143-255: tracks to tragsm_<OUTPUT> and trasph_<OUTPUT> */
static int trajectory(const Cli *c, const Model **models) {
    if (!c->output) {
        fprintf(stderr, "Error: Output file name is required for trajectory calculation mode\n");
        return 1;
    }
    uint64_t steps;
    double tu, mass;
    if (trace_options(c, &steps, &tu, &mass)) return 1;
    FILE *in = fopen(c->inputs[0], "r");
    if (!in) {
        perror("Error: Cannot open input file for reading");
        return 1;
    }
    InfileHeader h;
    ModelsParams p;
    memset(&h, 0, sizeof h);
    memset(&p, 0, sizeof p);
    if (read_cutoff_infile_header(in, &h) != 0) {
        fprintf(stderr, "Error: Invalid input file\n");
        fclose(in);
        return 1;
    }
    if (h.rig != h.rk) {
        fprintf(stderr, "Error: Invalid input file - starting and ending rigidity must be the same for trajectory "
                        "calculation mode\n");
        fclose(in);
        return 1;
    }
    if (setup_models(in, &h, models, &p)) {
        fclose(in);
        return 1;
    }
    if (fclose(in) != 0) {
        perror("Error: Cannot close input file");
        return 1;
    }
    const size_t len = strlen(c->output) + 8;
    char *gsm_name = malloc(len), *sph_name = malloc(len);
    if (!gsm_name || !sph_name) return 1;
    snprintf(gsm_name, len, "tragsm_%s", c->output);
    snprintf(sph_name, len, "trasph_%s", c->output);
    FILE *gsm = fopen(gsm_name, "w");
    FILE *sph = gsm ? fopen(sph_name, "w") : NULL;
    free(gsm_name);
    free(sph_name);
    if (!gsm || !sph) {
        if (gsm) fclose(gsm);
        perror("Error: Cannot open trajectory output file for writing");
        return 1;
    }
    trajectory_simulation(&h, models, &p, gsm, sph, steps, tu, mass);
    const int e1 = fclose(gsm), e2 = fclose(sph);
    if (e1 != 0 || e2 != 0) {
        perror("Error: Cannot close trajectory output file");
        return 1;
    }
    return 0;
}

int main(int argc, char **argv) {
    Cli c;
    memset(&c, 0, sizeof c);
    c.inputs = calloc((size_t)(argc > 1 ? argc : 1), sizeof(char *));
    if (!c.inputs) return 1;
    c.mode = "cutoff";
    c.particle = "proton";
    c.precision = COR_PREC_FP64;
    struct argp parser = {cli_options, cli_parse, "INFILE...",
                          "Cutoff rigidities, asymptotic directions, field maps and particle tracks in the "
                          "geomagnetic field, traced on an MI355X GPU",
                          0, 0, 0};
    argp_parse(&parser, argc, argv, 0, NULL, &c);
    if (!c.internal && !c.external && !c.crustal) c.internal = "igrf14";

    /* src/main.c:266-299: the model list, internal, external, crustal */
    const Model *models[4] = {NULL};
    int k = 0;
    const struct {
        const Model **reg;
        const char *name, *what;
    } pick[3] = {{internal_models, c.internal, "internal"},
                 {external_models, c.external, "external"},
                 {crustal_models, c.crustal, "crustal"}};
    for (int j = 0; j < 3; j++) {
        if (!pick[j].name) continue;
        const Model *m = get_model(pick[j].reg, pick[j].name);
        if (!m) {
            fprintf(stderr, "Error: Invalid %s geomagnetic field model: '%s'\n", pick[j].what, pick[j].name);
            return 1;
        }
        models[k++] = m;
    }
    if (c.n_inputs > 1 && strcmp(c.mode, "cutoff") != 0) {
        fprintf(stderr, "Error: several input files are taken in cutoff mode only\n");
        return 1;
    }
    int rc;
    if (strcmp(c.mode, "field") == 0) {
        rc = field(&c, models);
    } else if (strcmp(c.mode, "cutoff") == 0) {
        rc = cutoff(&c, models);
    } else if (strcmp(c.mode, "trajectory") == 0) {
        rc = trajectory(&c, models);
    } else {
        fprintf(stderr, "Error: Invalid calculation mode\n");
        rc = 1;
    }
    free(c.inputs);
    return rc;
}
