/*
 * fxmain.c -- host program driving the kernel.h entry points, with the
 * behaviour of the reference's main.c (main.c:14-129 config grammar,
 * allocDataHost.c:8-22 buffer size, openFiles.c:7-45 input, main.c:323-460
 * call sequence), linked against libdifxfx.so instead of the IPP kernel.
 *
 * Usage: fxmain [config-file] [split]
 *   config-file  default ./test.conf (main.c:324)
 *   split        time splits of the subint (main.c:371 uses 1)
 *
 * Differences from main.c: antenna count is not capped at 20 (main.c:10),
 * names and file names are not truncated by fixed buffers, and a missing
 * input file is a fatal error instead of a later fread on NULL.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

#include "kernel.h"

typedef struct {
    int nbit, npol, iscomplex, nchan, nant, numffts;
    double lo, bw;
    char **names, **files;
    double **delays;      /* [nant][3] */
    double *fileoffsets;  /* [nant]    */
} Config;

static void need(const char *tok, const char *what)
{
    if (!tok) {
        fprintf(stderr, "Error: Insufficient data for %s\n", what);
        exit(1);
    }
}

/* main.c:14-129 */
static void parse_config(const char *path, Config *c)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "Error Opening File %s\n", path);
        exit(1);
    }
    char line[1024];
    int antToRead = 0, iant = 0;
    memset(c, 0, sizeof(*c));
    c->npol = 2;
    c->iscomplex = 0;
    c->nbit = 2;
    while (fgets(line, sizeof(line), f)) {
        char *tok = strtok(line, " \n");
        if (!tok) {
            fprintf(stderr, "Error: Could not parse \"%s\"\n", line);
            exit(1);
        }
        if (antToRead) {
            c->names[iant] = strdup(tok);
            tok = strtok(NULL, " \n");
            if (!tok) {
                fprintf(stderr, "Error: Insufficient data for antenna %d\n", iant);
                exit(1);
            }
            c->files[iant] = strdup(tok);
            tok = strtok(NULL, " \n");
            for (int i = 0; i < 3; i++) {
                if (!tok) {
                    fprintf(stderr, "Error: Insufficient data for antenna %d\n", iant);
                    exit(1);
                }
                sscanf(tok, "%lf", &c->delays[iant][i]);
                tok = strtok(NULL, " \n");
            }
            if (!tok) {
                fprintf(stderr, "Error: Insufficient data for antenna %d\n", iant);
                exit(1);
            }
            sscanf(tok, "%lf", &c->fileoffsets[iant]);
            iant++;
            antToRead--;
            continue;
        }
        char *val = strtok(NULL, " \n");
        if (strcasecmp(tok, "COMPLEX") == 0) {
            need(val, "COMPLEX");
            sscanf(val, "%d", &c->iscomplex);
        } else if (strcasecmp(tok, "NBIT") == 0) {
            need(val, "NBIT");
            sscanf(val, "%d", &c->nbit);
        } else if (strcasecmp(tok, "NPOL") == 0) {
            need(val, "NPOL");
            sscanf(val, "%d", &c->npol);
        } else if (strcasecmp(tok, "NCHAN") == 0) {
            need(val, "NCHAN");
            sscanf(val, "%d", &c->nchan);
        } else if (strcasecmp(tok, "LO") == 0) {
            need(val, "LO");
            sscanf(val, "%lf", &c->lo);
        } else if (strcasecmp(tok, "BANDWIDTH") == 0) {
            need(val, "BANDWIDTH");
            sscanf(val, "%lf", &c->bw);
        } else if (strcasecmp(tok, "NUMFFTS") == 0) {
            need(val, "NUMFFTS");
            sscanf(val, "%d", &c->numffts);
        } else if (strcasecmp(tok, "NANT") == 0) {
            need(val, "NANT");
            sscanf(val, "%d", &c->nant);
            if (c->nant < 1) {
                fprintf(stderr, "Error: NANT must be positive\n");
                exit(1);
            }
            c->names = calloc((size_t)c->nant, sizeof(char *));
            c->files = calloc((size_t)c->nant, sizeof(char *));
            c->delays = calloc((size_t)c->nant, sizeof(double *));
            c->fileoffsets = calloc((size_t)c->nant, sizeof(double));
            for (int i = 0; i < c->nant; i++) c->delays[i] = calloc(3, sizeof(double));
            antToRead = c->nant;
            iant = 0;
        } else {
            fprintf(stderr, "Error: Unknown keyword \"%s\"\n", tok);
        }
    }
    fclose(f);
    if (antToRead) {
        fprintf(stderr, "Error: Insufficient data for antenna %d\n", iant);
        exit(1);
    }
}

/* main.c:249-287 */
static void print_config(const Config *c)
{
    printf("nbit: %d\n", c->nbit);
    printf("nPol: %d\n", c->npol);
    printf("iscomplex: %s\n", c->iscomplex ? "true" : "false");
    printf("nchan: %d\n", c->nchan);
    printf("nant: %d\n", c->nant);
    printf("lo: %lf\n", c->lo);
    printf("bandwidth: %lf\n", c->bw);
    printf("numffts: %d\n", c->numffts);
    printf("Antenna names:\n");
    for (int i = 0; i < c->nant; i++) printf("%s\n", c->names[i]);
    printf("Antenna files:\n");
    for (int i = 0; i < c->nant; i++) printf("%s\n", c->files[i]);
    printf("Delays:\n");
    for (int i = 0; i < c->nant; i++) {
        printf("Antenna %d:\n", i);
        for (int j = 0; j < 3; j++) printf("%e ", c->delays[i][j]);
        printf("\n");
    }
    printf("Antenna file offsets:\n");
    for (int i = 0; i < c->nant; i++) printf("%e\n", c->fileoffsets[i]);
}

int main(int argc, char **argv)
{
    const char *path = argc > 1 ? argv[1] : "./test.conf";
    const int split_size = argc > 2 ? atoi(argv[2]) : 1;
    Config cfg;
    parse_config(path, &cfg);
    print_config(&cfg);
    if (cfg.iscomplex || cfg.npol != 2) {
        fprintf(stderr, "Error: only real (COMPLEX 0) dual-polarisation data are supported\n");
        return 1;
    }
    if (split_size < 1 || cfg.numffts % split_size) {
        fprintf(stderr, "Error: NUMFFTS must be a multiple of the split\n");
        return 1;
    }

    /* allocDataHost.c:8-22 */
    const int cfactor = cfg.iscomplex ? 1 : 2;
    const long long subintbytes = (long long)cfg.nchan * cfactor * (cfg.numffts + 1) * cfg.nbit / 8 * cfg.npol;

    /* openFiles.c:7-45 */
    printf("subintbytes : %lld \n", subintbytes);
    u8 **inputdata = malloc(sizeof(u8 *) * (size_t)cfg.nant);
    for (int a = 0; a < cfg.nant; a++) {
        FILE *fa = fopen(cfg.files[a], "rb");
        if (!fa) {
            fprintf(stderr, "Problem with file %s - does it exist?\n", cfg.files[a]);
            return 1;
        }
        inputdata[a] = malloc((size_t)subintbytes);
        if (!inputdata[a]) {
            fprintf(stderr, "Unable to allocate %lld bytes. Quitting\n", subintbytes);
            return 1;
        }
        size_t got = fread(inputdata[a], 1, (size_t)subintbytes, fa);
        if ((long long)got != subintbytes) {
            fprintf(stderr, "Error: Problem reading data (Read %zu bytes)\n", got);
            return 1;
        }
        fclose(fa);
    }

    struct timespec *starttime = malloc(sizeof(struct timespec));
    char *starttimestring = malloc(64);
    int fftchannels, stridesize, substridesize, fractionalLoFreq, nbaselines;
    double sampletime;
    const int nbl = cfg.nant * (cfg.nant - 1) / 2;
    cf32 ***unpacked = malloc((size_t)split_size * cfg.nant * sizeof(cf32 **));
    cf32 ***channelised = malloc((size_t)split_size * cfg.nant * sizeof(cf32 **));
    cf32 ***conjchannels = malloc((size_t)split_size * cfg.nant * sizeof(cf32 **));
    cf32 ***visibilities = malloc((size_t)split_size * (nbl > 0 ? nbl : 1) * sizeof(cf32 **));
    int *baselineCount = malloc((size_t)split_size * (nbl > 0 ? nbl : 1) * sizeof(int));

    init_FxKernel(&cfg.nant, &cfg.nchan, &cfg.nbit, &cfg.lo, &cfg.bw, starttime, starttimestring, &fftchannels,
                  &sampletime, &stridesize, &substridesize, &fractionalLoFreq, unpacked, channelised, conjchannels,
                  visibilities, &nbaselines, baselineCount, split_size);

    long long diff_ms;
    int *baselineCount_to_norm = malloc((size_t)split_size * (nbl > 0 ? nbl : 1) * sizeof(int));
    int *iter = malloc(sizeof(int) * (size_t)cfg.numffts);
    for (int i = 0; i < cfg.numffts; i++) iter[i] = i;
    cf32 ***vis_to_norm = malloc((size_t)split_size * (nbl > 0 ? nbl : 1) * sizeof(cf32 **));
    int numffts_split = cfg.numffts / split_size;

    /* main.c:401-408 */
    for (int i = 0; i < split_size; i++) {
        processAntennasAndBaselinePara(&cfg.nant, &numffts_split, &fftchannels, cfg.fileoffsets, &sampletime,
                                       inputdata, &unpacked[i * cfg.nant], &channelised[i * cfg.nant],
                                       &conjchannels[i * cfg.nant], cfg.delays, &cfg.nbit, &substridesize,
                                       &stridesize, &fractionalLoFreq, &cfg.lo, &cfg.nchan, &visibilities[i * nbl],
                                       &baselineCount[i * nbl], &iter[numffts_split * i], &cfg.bw,
                                       &vis_to_norm[i * nbl], &baselineCount_to_norm[i * nbl], split_size);
    }

    cf32 ***vis_for_split = malloc((size_t)(nbl > 0 ? nbl : 1) * sizeof(cf32 **));
    int *baselineCount_for_split = malloc((size_t)(nbl > 0 ? nbl : 1) * sizeof(int));
    merge(split_size, nbaselines, visibilities, &cfg.nant, &cfg.nchan, baselineCount, vis_for_split,
          baselineCount_for_split);

    /* main.c:435-441 normalises the merged slab in split chunks; one call
       over all baselines is the same operation */
    cf32 ***vis_out = malloc((size_t)(nbl > 0 ? nbl : 1) * sizeof(cf32 **));
    processNormalize(&nbaselines, baselineCount_for_split, vis_for_split, &cfg.nchan, vis_out);

    endTiming(starttime, &diff_ms);
    saveVisibilities(&nbaselines, &cfg.nchan, vis_out, &cfg.bw);
    saveLog(&diff_ms, starttimestring);
    return 0;
}
