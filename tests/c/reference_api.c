/*
 * A C caller of the reference library's own entry points, compiled against
 * include/resynth_mi355x.h alone and linked to libresynth_mi355x.so: the way a program
 * built for the reference (src/testSynth.c:60-212, src/resynthesizer/resynthesizer.c:
 * 505-513) would use this library as a drop-in.  It uses only reference names and
 * signatures: imageSynth, imageSynth2, engine, setDefaultParams,
 * prepareImageFormatIndices, countPixelelsPerPixelForFormat, new_pixmap, free_map.
 *
 *   reference_api known
 *       the known answers of src/testSynth.c:163-212, through imageSynth
 *   reference_api synth  FMT W H IMAGE.raw MASK.raw OUT.raw
 *   reference_api synth2 FMT W H IMAGE.raw MASK.raw MASK2.raw OUT.raw
 *       imageSynth / imageSynth2 with default parameters on unpadded rows
 *   reference_api engine TW TH TARGET.raw CW CH CORPUS.raw COLORS MAPS TILEH TILEV TYPE
 *                        MAPWEIGHT SENS PATCH PROBES OUT.raw
 *       engine() on engine-layout pixmaps ([mask, colours, maps] per pixel)
 *
 * Prints one line per check and exits 0 when every check passed.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "resynth_mi355x.h"

static int progressCalls = 0, lastPercent = -1;

static void progress(int percent, void *context)
{
  (void)context;
  progressCalls++;
  lastPercent = percent;
}

static int failures = 0;

static void expect_bytes(const char *what, const unsigned char *got, const unsigned char *want, size_t n)
{
  int ok = memcmp(got, want, n) == 0;
  printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) {
    failures++;
    printf("  got ");
    for (size_t i = 0; i < n; i++) printf("%02x ", got[i]);
    printf("\n  want");
    for (size_t i = 0; i < n; i++) printf(" %02x", want[i]);
    printf("\n");
  }
}

static void expect_int(const char *what, int got, int want)
{
  printf("%s %s (%d)\n", got == want ? "PASS" : "FAIL", what, got);
  if (got != want) failures++;
}

/* src/testSynth.c:86-212: padded rows, each case's expected bytes of the synthesized row */
static void known_answers(void)
{
  TImageSynthParameters P;
  int cancel = 0;
  setDefaultParams(&P);

  unsigned char image[42] = {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0,
                             0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0,
                             0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 8, 0, 0};
  unsigned char mask[12] = {0, 0, 0, 0, 0, 0xFF, 0, 0, 0, 0, 0, 0};
  ImageBuffer img = {image, 3, 3, 14}, msk = {mask, 3, 3, 4};
  expect_int("imageSynth RGBA 3x3 returns 0", imageSynth(&img, &msk, T_RGBA, &P, progress, NULL, &cancel), 0);
  const unsigned char row1[12] = {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0};
  expect_bytes("centre pixel synthesized, alpha kept", image + 14, row1, 12);
  expect_bytes("pad bytes untouched", image + 12, (const unsigned char[]){0, 0}, 2);

  unsigned char image2[14] = {128, 128, 128, 0xFF, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0};
  unsigned char mask2[4] = {0, 0xFF, 0, 0};
  ImageBuffer img2 = {image2, 3, 1, 14}, msk2 = {mask2, 3, 1, 4};
  expect_int("imageSynth mixed transparency", imageSynth(&img2, &msk2, T_RGBA, &P, progress, NULL, &cancel), 0);
  expect_bytes("80 80 80 ff  80 80 80 01  00 00 00 00", image2,
               (const unsigned char[]){0x80, 0x80, 0x80, 0xff, 0x80, 0x80, 0x80, 1, 0, 0, 0, 0}, 12);

  unsigned char rgb[22] = {128, 128, 128, 1, 1, 1, 2, 2, 2, 5, 5, 64, 64, 64, 4, 4, 4, 3, 3, 3, 7, 7};
  unsigned char mask3[8] = {0, 0, 0, 0, 0, 0xFF, 0, 0};
  ImageBuffer imgRGB = {rgb, 3, 2, 11}, msk3 = {mask3, 3, 2, 4};
  expect_int("imageSynth RGB", imageSynth(&imgRGB, &msk3, T_RGB, &P, progress, NULL, &cancel), 0);
  expect_bytes("RGB row 1", rgb + 11, (const unsigned char[]){0x40, 0x40, 0x40, 1, 1, 1, 3, 3, 3}, 9);

  unsigned char grayA[8] = {128, 0xFF, 64, 1, 1, 0, 0, 0};
  ImageBuffer imgGA = {grayA, 3, 1, 8};
  expect_int("imageSynth GrayA", imageSynth(&imgGA, &msk2, T_GrayA, &P, progress, NULL, &cancel), 0);
  expect_bytes("80 ff  80 01  01 00", grayA, (const unsigned char[]){0x80, 0xff, 0x80, 1, 1, 0}, 6);

  unsigned char gray[5] = {128, 64, 1, 0, 0};
  ImageBuffer imgG = {gray, 3, 1, 5};
  expect_int("imageSynth Gray, NULL parameters", imageSynth(&imgG, &msk2, T_Gray, NULL, progress, NULL, &cancel), 0);
  expect_bytes("80  01  01", gray, (const unsigned char[]){0x80, 1, 1}, 3);

  /* error conditions (src/testSynth.c:195-212) */
  unsigned char image3[14] = {128, 128, 128, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0};
  ImageBuffer img3 = {image3, 3, 1, 14};
  expect_int("all transparent: empty corpus", imageSynth(&img3, &msk2, T_RGBA, &P, progress, NULL, &cancel),
             IMAGE_SYNTH_ERROR_EMPTY_CORPUS);
  P.patchSize = 65;
  expect_int("patchSize out of range", imageSynth(&imgG, &msk2, T_Gray, &P, progress, NULL, &cancel),
             IMAGE_SYNTH_ERROR_PATCH_SIZE_EXCEEDED);
  P.patchSize = 10;
  ImageBuffer bad = {mask3, 1, 1, 1};
  expect_int("mask not the image's size", imageSynth(&imgG, &bad, T_Gray, &P, progress, NULL, &cancel),
             IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH);
  unsigned char empty[4] = {0, 0, 0, 0};
  ImageBuffer mskE = {empty, 3, 1, 4};
  expect_int("empty target", imageSynth(&imgG, &mskE, T_Gray, &P, progress, NULL, &cancel),
             IMAGE_SYNTH_ERROR_EMPTY_TARGET);
  expect_int("invalid format", imageSynth(&imgG, &msk2, (TImageFormat)666, &P, progress, NULL, &cancel),
             IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT);
  P.patchSize = 30;
  P.matchContextType = 9;
  expect_int("match context type out of range", imageSynth(&imgG, &msk2, T_Gray, &P, progress, NULL, &cancel),
             IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE);
  expect_int("progress was reported", progressCalls > 0, 1);

  /* cancel before the first progress point: success, image left as it was */
  unsigned char keep[5] = {128, 64, 1, 0, 0};
  memcpy(gray, keep, 5);
  cancel = 1;
  setDefaultParams(&P);
  expect_int("cancelled call returns 0", imageSynth(&imgG, &msk2, T_Gray, &P, progress, NULL, &cancel), 0);
  expect_bytes("cancelled call leaves the image", gray, keep, 5);
}

static unsigned char *read_file(const char *path, size_t n)
{
  FILE *f = fopen(path, "rb");
  unsigned char *p = (unsigned char *)malloc(n ? n : 1);
  if (!f || !p || fread(p, 1, n, f) != n) {
    fprintf(stderr, "cannot read %zu bytes of %s\n", n, path);
    exit(2);
  }
  fclose(f);
  return p;
}

static void write_file(const char *path, const unsigned char *p, size_t n)
{
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, 1, n, f) != n) {
    fprintf(stderr, "cannot write %s\n", path);
    exit(2);
  }
  fclose(f);
}

int main(int argc, char **argv)
{
  if (argc >= 2 && strcmp(argv[1], "known") == 0) {
    known_answers();
  } else if (argc == 8 && strcmp(argv[1], "synth") == 0) {
    TImageFormat fmt = (TImageFormat)atoi(argv[2]);
    unsigned w = (unsigned)atoi(argv[3]), h = (unsigned)atoi(argv[4]);
    unsigned bpp = countPixelelsPerPixelForFormat(fmt);
    unsigned char *image = read_file(argv[5], (size_t)w * h * bpp), *mask = read_file(argv[6], (size_t)w * h);
    ImageBuffer img = {image, w, h, (size_t)w * bpp}, msk = {mask, w, h, w};
    int cancel = 0;
    expect_int("imageSynth", imageSynth(&img, &msk, fmt, NULL, progress, NULL, &cancel), 0);
    write_file(argv[7], image, (size_t)w * h * bpp);
    free(image);
    free(mask);
  } else if (argc == 9 && strcmp(argv[1], "synth2") == 0) {
    TImageFormat fmt = (TImageFormat)atoi(argv[2]);
    unsigned w = (unsigned)atoi(argv[3]), h = (unsigned)atoi(argv[4]);
    unsigned bpp = countPixelelsPerPixelForFormat(fmt);
    unsigned char *image = read_file(argv[5], (size_t)w * h * bpp), *mask = read_file(argv[6], (size_t)w * h),
                  *mask2 = read_file(argv[7], (size_t)w * h);
    ImageBuffer img = {image, w, h, (size_t)w * bpp}, msk = {mask, w, h, w}, msk2 = {mask2, w, h, w};
    int cancel = 0;
    expect_int("imageSynth2", imageSynth2(&img, &msk, &msk2, fmt, NULL, progress, NULL, &cancel), 0);
    write_file(argv[8], image, (size_t)w * h * bpp);
    free(image);
    free(mask);
    free(mask2);
  } else if (argc == 18 && strcmp(argv[1], "engine") == 0) {
    unsigned tw = (unsigned)atoi(argv[2]), th = (unsigned)atoi(argv[3]);
    unsigned cw = (unsigned)atoi(argv[5]), ch = (unsigned)atoi(argv[6]);
    unsigned colors = (unsigned)atoi(argv[8]), maps = (unsigned)atoi(argv[9]);
    TImageSynthParameters P;
    setDefaultParams(&P);
    P.isMakeSeamlesslyTileableHorizontally = atoi(argv[10]);
    P.isMakeSeamlesslyTileableVertically = atoi(argv[11]);
    P.matchContextType = atoi(argv[12]);
    P.mapWeight = atof(argv[13]);
    P.sensitivityToOutliers = atof(argv[14]);
    P.patchSize = (unsigned)atoi(argv[15]);
    P.maxProbeCount = (unsigned)atoi(argv[16]);
    TFormatIndices ix;
    prepareImageFormatIndices(&ix, colors, maps, 0, 0, maps > 0);
    const unsigned depth = ix.total_bpp;
    Map target, corpus;
    new_pixmap(&target, tw, th, depth);
    new_pixmap(&corpus, cw, ch, depth);
    unsigned char *t = read_file(argv[4], (size_t)tw * th * depth), *c = read_file(argv[7], (size_t)cw * ch * depth);
    memcpy(target.data->data, t, (size_t)tw * th * depth);
    memcpy(corpus.data->data, c, (size_t)cw * ch * depth);
    int cancel = 0;
    expect_int("engine", engine(P, &ix, &target, &corpus, progress, NULL, &cancel), 0);
    write_file(argv[17], (const unsigned char *)target.data->data, (size_t)tw * th * depth);
    free_map(&target);
    free_map(&corpus);
    free(t);
    free(c);
  } else {
    fprintf(stderr, "usage: see the header of tests/c/reference_api.c\n");
    return 2;
  }
  printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
