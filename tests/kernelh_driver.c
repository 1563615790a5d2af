/*
 * kernelh_driver.c -- calls the per-(antenna, FFT) kernel.h functions of
 * libdifxfx.so the way a C caller of the reference would, on an input case
 * written by tests/test_gpu_host.py, and dumps every result for the test to
 * compare with the oracle.
 *
 * Input (little endian): int32 nant, nchan, nbytes, fftindex, numffts;
 *   double lo, bw; double delays[nant][3]; double fileoffsets[nant];
 *   uint8 data[nant][nbytes]
 * Output, in order:
 *   getStationDelay(antenna 0, fftindex): double mean, a, b
 *   stationDelayAndOffset(fftindex, numffts): int32 offset, valid; double fd, a, b
 *   processAntennas (last FFT numffts-1): int32 antValid[nant];
 *     cf32 unpacked[nant][2][F], channelised[nant][2][F], conj[nant][2][nchan]
 *   processBaselinePara on those: cf32 vis[nbl][4][nchan], int32 count[nbl]
 *   merge + processNormalize of that slab: cf32 vis[nbl][4][nchan]
 *   processAntennasAndBaselinePara over FFTs 0..numffts-1 (split 1) on fresh
 *   buffers: cf32 unpacked[nant][2][F], channelised[nant][2][F],
 *   conj[nant][2][nchan] (each antenna's last valid FFT), vis[nbl][4][nchan],
 *   int32 count[nbl]
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "kernel.h"

static void rd(void *p, size_t n, FILE *f)
{
    if (fread(p, 1, n, f) != n) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: kernelh_driver in.bin out.bin\n");
        return 2;
    }
    FILE *fi = fopen(argv[1], "rb");
    FILE *fo = fopen(argv[2], "wb");
    if (!fi || !fo) return 2;
    int32_t hdr[5];
    rd(hdr, sizeof(hdr), fi);
    int nant = hdr[0], nchan = hdr[1], nbytes = hdr[2], fftindex = hdr[3], numffts = hdr[4];
    double lo, bw;
    rd(&lo, 8, fi);
    rd(&bw, 8, fi);
    double **delays = malloc(sizeof(double *) * nant);
    double *fileoff = malloc(sizeof(double) * nant);
    for (int j = 0; j < nant; j++) {
        delays[j] = malloc(3 * sizeof(double));
        rd(delays[j], 3 * sizeof(double), fi);
    }
    rd(fileoff, sizeof(double) * nant, fi);
    u8 **data = malloc(sizeof(u8 *) * nant);
    for (int j = 0; j < nant; j++) {
        data[j] = malloc(nbytes);
        rd(data[j], nbytes, fi);
    }
    fclose(fi);

    int nbit = 2, fftchannels, stridesize, substridesize, fractionalLoFreq, nbaselines;
    double sampletime;
    struct timespec t0;
    char tstr[64];
    const int nbl = nant * (nant - 1) / 2;
    cf32 ***unpacked = malloc(sizeof(cf32 **) * nant);
    cf32 ***channelised = malloc(sizeof(cf32 **) * nant);
    cf32 ***conjc = malloc(sizeof(cf32 **) * nant);
    cf32 ***vis = malloc(sizeof(cf32 **) * (nbl ? nbl : 1));
    int *count = malloc(sizeof(int) * (nbl ? nbl : 1));
    init_FxKernel(&nant, &nchan, &nbit, &lo, &bw, &t0, tstr, &fftchannels, &sampletime, &stridesize, &substridesize,
                  &fractionalLoFreq, unpacked, channelised, conjc, vis, &nbaselines, count, 1);
    const int F = fftchannels;

    double m, a, b;
    getStationDelay(0, fftindex, &m, &a, &b, delays);
    fwrite(&m, 8, 1, fo);
    fwrite(&a, 8, 1, fo);
    fwrite(&b, 8, 1, fo);

    int *antValid = calloc(nant, sizeof(int));
    int offset;
    double fd, da, db;
    int it = fftindex;
    stationDelayAndOffset(delays, &it, fileoff, antValid, &offset, &fd, &da, &db, &sampletime, &fftchannels, numffts);
    int32_t ov[2] = {offset, antValid[0]};
    fwrite(ov, 4, 2, fo);
    fwrite(&fd, 8, 1, fo);
    fwrite(&da, 8, 1, fo);
    fwrite(&db, 8, 1, fo);

    FxKernel kernel = {0};
    processAntennas(&kernel, &nant, &numffts, &fftchannels, fileoff, &sampletime, antValid, data, unpacked,
                    channelised, conjc, delays, &nbit, &substridesize, &stridesize, &fractionalLoFreq, &lo, &nchan);
    for (int j = 0; j < nant; j++) fwrite(&antValid[j], 4, 1, fo);
    for (int j = 0; j < nant; j++)
        for (int p = 0; p < 2; p++) fwrite(unpacked[j][p], sizeof(cf32), F, fo);
    for (int j = 0; j < nant; j++)
        for (int p = 0; p < 2; p++) fwrite(channelised[j][p], sizeof(cf32), F, fo);
    for (int j = 0; j < nant; j++)
        for (int p = 0; p < 2; p++) fwrite(conjc[j][p], sizeof(cf32), nchan, fo);

    processBaselinePara(&nant, antValid, channelised, conjc, vis, &nchan, count);
    for (int i = 0; i < nbl; i++)
        for (int p = 0; p < 4; p++) fwrite(vis[i][p], sizeof(cf32), nchan, fo);
    fwrite(count, sizeof(int), nbl, fo);

    cf32 ***vout = malloc(sizeof(cf32 **) * (nbl ? nbl : 1));
    int *cout = malloc(sizeof(int) * (nbl ? nbl : 1));
    cf32 ***vnorm = malloc(sizeof(cf32 **) * (nbl ? nbl : 1));
    merge(1, nbaselines, vis, &nant, &nchan, count, vout, cout);
    processNormalize(&nbaselines, cout, vout, &nchan, vnorm);
    for (int i = 0; i < nbl; i++)
        for (int p = 0; p < 4; p++) fwrite(vnorm[i][p], sizeof(cf32), nchan, fo);

    /* the hot entry on fresh buffers (init_FxKernel zeroes them) */
    cf32 ***unp2 = malloc(sizeof(cf32 **) * nant);
    cf32 ***chan2 = malloc(sizeof(cf32 **) * nant);
    cf32 ***conj2 = malloc(sizeof(cf32 **) * nant);
    cf32 ***vis2 = malloc(sizeof(cf32 **) * (nbl ? nbl : 1));
    cf32 ***vis2out = malloc(sizeof(cf32 **) * (nbl ? nbl : 1));
    int *count2 = malloc(sizeof(int) * (nbl ? nbl : 1));
    int *count2out = malloc(sizeof(int) * (nbl ? nbl : 1));
    init_FxKernel(&nant, &nchan, &nbit, &lo, &bw, &t0, tstr, &fftchannels, &sampletime, &stridesize, &substridesize,
                  &fractionalLoFreq, unp2, chan2, conj2, vis2, &nbaselines, count2, 1);
    int *iter = malloc(sizeof(int) * numffts);
    for (int k = 0; k < numffts; k++) iter[k] = k;
    double *fo2 = malloc(sizeof(double) * nant);
    for (int j = 0; j < nant; j++) fo2[j] = fileoff[j];
    processAntennasAndBaselinePara(&nant, &numffts, &fftchannels, fo2, &sampletime, data, unp2, chan2, conj2, delays,
                                   &nbit, &substridesize, &stridesize, &fractionalLoFreq, &lo, &nchan, vis2, count2,
                                   iter, &bw, vis2out, count2out, 1);
    for (int j = 0; j < nant; j++)
        for (int p = 0; p < 2; p++) fwrite(unp2[j][p], sizeof(cf32), F, fo);
    for (int j = 0; j < nant; j++)
        for (int p = 0; p < 2; p++) fwrite(chan2[j][p], sizeof(cf32), F, fo);
    for (int j = 0; j < nant; j++)
        for (int p = 0; p < 2; p++) fwrite(conj2[j][p], sizeof(cf32), nchan, fo);
    for (int i = 0; i < nbl; i++)
        for (int p = 0; p < 4; p++) fwrite(vis2out[i][p], sizeof(cf32), nchan, fo);
    fwrite(count2out, sizeof(int), nbl, fo);
    fclose(fo);
    return 0;
}
