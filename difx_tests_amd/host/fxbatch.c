/*
 * fxbatch.c -- uneven subbands through the native API, one hipGraph launch
 * per subint (fx_batch_process_ex).
 *
 * Usage: fxbatch [config-file] [--binary] [--launch] [--tune]
 *
 * Config: the grammar of main.c:14-129 (NBIT, COMPLEX, NPOL, LO, BANDWIDTH,
 * NCHAN, NUMFFTS, NANT + antenna lines) extended with
 *   SUBBAND nchan bandwidth lo numffts antennas
 *            one line per subband; antennas is a comma-separated list of
 *            indices into the NANT list, or ALL
 *   NSUBINT n             subints to process (default 1)
 *   PREC F32|F16|F8|F32V  X-engine precision (default F32)
 * An antenna's FILE may hold one %d, replaced by the subband index.  A file
 * is a continuous sample stream of its subband; subint s reads numffts + 1
 * FFTs of samples (allocDataHost.c:21) starting at FFT s*numffts, with
 * global FFT indices s*numffts + k and buffer origin s*numffts.
 * Outputs, normalised by the valid-FFT counts: vis_sb<i>.out (the reference's
 * vis.out format) or, with --binary, vis_sb<i>.bin (DIFXFXV1), and
 * runtime.fxbatch.log.  --launch runs plain launches instead of graphs;
 * --tune autotunes every subband first (FX_TUNE_CACHE keeps the result).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

#include "difxfx.h"

#define MAXSB 64

typedef struct {
    int nchan, numffts, nant;
    double bw, lo;
    int *ants;   /* indices into the antenna list */
} Subband;

typedef struct {
    int nbit, npol, iscomplex, nant, nsubint, prec, nsb;
    double lo, bw;
    int nchan, numffts;
    char **names, **files;
    double *delays;       /* [nant][3] */
    double *fileoffsets;  /* [nant]    */
    Subband sb[MAXSB];
} Config;

static void die(const char *msg, const char *arg)
{
    fprintf(stderr, "Error: %s%s%s\n", msg, arg ? " " : "", arg ? arg : "");
    exit(1);
}

/* File name of subband `sb` from an antenna's FILE: the first "%d" is replaced
 * by the subband index; any other '%' is rejected (the name is never used as a
 * format string).  Exits when the result does not fit `cap` bytes. */
static void subband_file(char *out, size_t cap, const char *pattern, int sb)
{
    const char *d = strstr(pattern, "%d");
    for (const char *q = pattern; *q; q++)
        if (*q == '%' && q != d) die("FILE may hold no '%' other than one %d:", pattern);
    int n;
    if (d)
        n = snprintf(out, cap, "%.*s%d%s", (int)(d - pattern), pattern, sb, d + 2);
    else
        n = snprintf(out, cap, "%s", pattern);
    if (n < 0 || (size_t)n >= cap) die("file name too long:", pattern);
}

static void check(int rc, const char *what)
{
    if (rc) {
        fprintf(stderr, "Error: %s: %s\n", what, fx_last_error());
        exit(1);
    }
}

static char *next_tok(void)
{
    return strtok(NULL, " \n");
}

static void parse_subband(Config *c, char *tok)
{
    if (c->nsb >= MAXSB) die("too many subbands", NULL);
    Subband *s = &c->sb[c->nsb];
    char *t[5];
    t[0] = tok;
    for (int i = 1; i < 5; i++)
        if (!(t[i] = next_tok())) die("Insufficient data for SUBBAND", NULL);
    s->nchan = atoi(t[0]);
    s->bw = atof(t[1]);
    s->lo = atof(t[2]);
    s->numffts = atoi(t[3]);
    if (s->nchan < 1 || !(s->bw > 0) || s->numffts < 1) die("bad SUBBAND line", NULL);
    s->ants = NULL;
    s->nant = -1;                       /* resolved after NANT is known */
    if (strcasecmp(t[4], "ALL") != 0) {
        int n = 1;
        for (const char *q = t[4]; *q; q++) n += *q == ',';
        s->ants = malloc(sizeof(int) * n);
        s->nant = 0;
        for (char *q = strtok_r(t[4], ",", &t[0]); q; q = strtok_r(NULL, ",", &t[0])) s->ants[s->nant++] = atoi(q);
    }
    c->nsb++;
}

static void parse_config(const char *path, Config *c)
{
    FILE *f = fopen(path, "r");
    if (!f) die("Error Opening File", path);
    char line[4096];
    int antToRead = 0, iant = 0;
    memset(c, 0, sizeof(*c));
    c->npol = 2;
    c->nbit = 2;
    c->nsubint = 1;
    c->prec = FX_PREC_F32;
    while (fgets(line, sizeof(line), f)) {
        char *tok = strtok(line, " \n");
        if (!tok) die("Could not parse an empty line", NULL);
        if (antToRead) {
            c->names[iant] = strdup(tok);
            char *file = next_tok();
            if (!file) die("Insufficient data for an antenna", NULL);
            c->files[iant] = strdup(file);
            for (int i = 0; i < 4; i++) {
                char *v = next_tok();
                if (!v) die("Insufficient data for an antenna", NULL);
                if (i < 3) c->delays[3 * iant + i] = atof(v);
                else c->fileoffsets[iant] = atof(v);
            }
            iant++;
            antToRead--;
            continue;
        }
        char *val = next_tok();
        if (strcasecmp(tok, "SUBBAND") == 0) {
            if (!val) die("Insufficient data for SUBBAND", NULL);
            parse_subband(c, val);
            continue;
        }
        if (!val) die("Insufficient data for", tok);
        if (strcasecmp(tok, "COMPLEX") == 0) c->iscomplex = atoi(val);
        else if (strcasecmp(tok, "NBIT") == 0) c->nbit = atoi(val);
        else if (strcasecmp(tok, "NPOL") == 0) c->npol = atoi(val);
        else if (strcasecmp(tok, "NCHAN") == 0) c->nchan = atoi(val);
        else if (strcasecmp(tok, "LO") == 0) c->lo = atof(val);
        else if (strcasecmp(tok, "BANDWIDTH") == 0) c->bw = atof(val);
        else if (strcasecmp(tok, "NUMFFTS") == 0) c->numffts = atoi(val);
        else if (strcasecmp(tok, "NSUBINT") == 0) c->nsubint = atoi(val);
        else if (strcasecmp(tok, "PREC") == 0) {
            if (strcasecmp(val, "F32") == 0) c->prec = FX_PREC_F32;
            else if (strcasecmp(val, "F16") == 0) c->prec = FX_PREC_F16;
            else if (strcasecmp(val, "F8") == 0) c->prec = FX_PREC_F8;
            else if (strcasecmp(val, "F32V") == 0) c->prec = FX_PREC_F32V;
            else die("unknown PREC", val);
        } else if (strcasecmp(tok, "NANT") == 0) {
            c->nant = atoi(val);
            if (c->nant < 1) die("NANT must be positive", NULL);
            antToRead = c->nant;
            c->names = calloc(c->nant, sizeof(char *));
            c->files = calloc(c->nant, sizeof(char *));
            c->delays = calloc(3 * (size_t)c->nant, sizeof(double));
            c->fileoffsets = calloc(c->nant, sizeof(double));
        } else {
            fprintf(stderr, "Error: Unknown keyword \"%s\"\n", tok);
        }
    }
    fclose(f);
    if (antToRead) die("Insufficient antenna lines", NULL);
    if (!(c->nbit == 2 && !c->iscomplex) && !(c->nbit == 8 && c->iscomplex))
        die("supported inputs are NBIT 2 real and NBIT 8 complex", NULL);
    if (c->npol != 2) die("NPOL must be 2", NULL);
    if (c->nsubint < 1) die("NSUBINT must be positive", NULL);
    if (c->nsb == 0) {                   /* single band from the plain keywords */
        c->sb[0].nchan = c->nchan;
        c->sb[0].bw = c->bw;
        c->sb[0].lo = c->lo;
        c->sb[0].numffts = c->numffts;
        c->sb[0].nant = -1;
        c->nsb = 1;
    }
    for (int i = 0; i < c->nsb; i++) {
        Subband *s = &c->sb[i];
        if (s->nant < 0) {
            s->nant = c->nant;
            s->ants = malloc(sizeof(int) * c->nant);
            for (int j = 0; j < c->nant; j++) s->ants[j] = j;
        }
        for (int j = 0; j < s->nant; j++)
            if (s->ants[j] < 0 || s->ants[j] >= c->nant) die("SUBBAND antenna index out of range", NULL);
    }
}

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int main(int argc, char **argv)
{
    const char *cfgpath = "./test.conf";
    int binary = 0, mode = FX_BATCH_GRAPH, tune = 0;
    for (int i = 1; i < argc; i++) {
        if (strcmp(argv[i], "--binary") == 0) binary = 1;
        else if (strcmp(argv[i], "--launch") == 0) mode = FX_BATCH_LAUNCH;
        else if (strcmp(argv[i], "--tune") == 0) tune = 1;
        else cfgpath = argv[i];
    }
    Config c;
    parse_config(cfgpath, &c);
    const int nsb = c.nsb;
    fx_context *ctx[MAXSB];
    uint8_t **bufs[MAXSB];
    FILE **fps[MAXSB];
    long long bytes[MAXSB], step[MAXSB];
    int *iters[MAXSB], nf[MAXSB], origins[MAXSB];
    for (int i = 0; i < nsb; i++) {
        Subband *s = &c.sb[i];
        const int F = c.iscomplex ? s->nchan : 2 * s->nchan;
        /* bytes of one FFT of samples, and of a subint buffer (numffts + 1 FFTs) */
        step[i] = c.iscomplex ? 4LL * F : F / 2;
        bytes[i] = step[i] * (s->numffts + 1);
        fx_params prm = {s->nant, s->nchan, s->lo, s->bw, c.prec, s->numffts, bytes[i], c.iscomplex};
        check(fx_create(&prm, &ctx[i]), "fx_create");
        double *d = malloc(sizeof(double) * 3 * s->nant), *o = malloc(sizeof(double) * s->nant);
        bufs[i] = malloc(sizeof(uint8_t *) * s->nant);
        fps[i] = malloc(sizeof(FILE *) * s->nant);
        for (int j = 0; j < s->nant; j++) {
            const int a = s->ants[j];
            memcpy(d + 3 * j, c.delays + 3 * a, 3 * sizeof(double));
            o[j] = c.fileoffsets[a];
            char name[4096];
            subband_file(name, sizeof(name), c.files[a], i);
            if (!(fps[i][j] = fopen(name, "rb"))) die("cannot open input file", name);
            bufs[i][j] = malloc((size_t)bytes[i]);
        }
        check(fx_set_model(ctx[i], d, o), "fx_set_model");
        free(d);
        free(o);
        check(fx_reset(ctx[i]), "fx_reset");
        nf[i] = s->numffts;
        iters[i] = malloc(sizeof(int) * s->numffts);
    }
    fx_batch *batch;
    check(fx_batch_create(ctx, nsb, &batch), "fx_batch_create");
    FILE *log = fopen("runtime.fxbatch.log", "w");
    double total = 0.0;
    for (int sub = 0; sub < c.nsubint; sub++) {
        for (int i = 0; i < nsb; i++) {
            for (int j = 0; j < c.sb[i].nant; j++) {
                memset(bufs[i][j], 0, (size_t)bytes[i]);
                if (fseek(fps[i][j], (long)(step[i] * nf[i] * (long long)sub), SEEK_SET) == 0) {
                    size_t got = fread(bufs[i][j], 1, (size_t)bytes[i], fps[i][j]);
                    (void)got;              /* a short final read leaves zeros, as a short file does */
                }
            }
            check(fx_upload_data(ctx[i], (const uint8_t *const *)bufs[i], bytes[i]), "fx_upload_data");
            for (int k = 0; k < nf[i]; k++) iters[i][k] = sub * nf[i] + k;
            origins[i] = sub * nf[i];
        }
        if (tune && sub == 0)
            for (int i = 0; i < nsb; i++) {
                double rep[5];
                check(fx_autotune(ctx[i], nf[i], 3, NULL, rep), "fx_autotune");
                if (log) fprintf(log, "subband %d tuned: plan %d tile %d\n", i, (int)rep[0], (int)rep[1]);
            }
        const double t0 = now_ms();
        check(fx_batch_process_ex(batch, nf, (const int *const *)iters, origins, 1, mode), "fx_batch_process_ex");
        check(fx_batch_sync(batch), "fx_batch_sync");
        const double t1 = now_ms();
        total += t1 - t0;
        if (log) fprintf(log, "subint %d: %.4f ms\n", sub, t1 - t0);
    }
    if (log) {
        fprintf(log, "total: %.4f ms for %d subints x %d subbands (%s)\n", total, c.nsubint, nsb,
                mode == FX_BATCH_GRAPH ? "one graph launch per subint" : "plain launches");
        fclose(log);
    }
    for (int i = 0; i < nsb; i++) {
        char name[64];
        snprintf(name, sizeof(name), binary ? "vis_sb%d.bin" : "vis_sb%d.out", i);
        check(fx_save_vis(ctx[i], name, binary ? FX_VIS_BINARY : FX_VIS_TEXT, FX_VIS_NORMALIZE), "fx_save_vis");
    }
    fx_batch_destroy(batch);
    for (int i = 0; i < nsb; i++) {
        for (int j = 0; j < c.sb[i].nant; j++) {
            fclose(fps[i][j]);
            free(bufs[i][j]);
        }
        fx_destroy(ctx[i]);
    }
    return 0;
}
