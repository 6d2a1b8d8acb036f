/* rct_read_bench.c -- TEST INFRASTRUCTURE (the CPU baseline of .rct ingest).
 *
 * Times the reference runner's input loop, rabbitRunner.c:266-283 without its
 * iterate calls: ctFileReaderOpenFile, then the N records read with
 * ctFileReaderReadImage (ctFileReader.c:47-61: one fread of the 12-double
 * matrix, one of the image) into a buffer of b projections, b at a time.
 * Links the reference's own ctFileReader.o, built from its sources by this
 * directory's Makefile.  Prints one JSON line: images, bytes, seconds.
 *
 * usage: rct_read_bench <file.rct> <b>      (b must divide the image count)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "ctFileReader.h"

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: %s <file.rct> <b>\n", argv[0]);
    return 2;
  }
  RabbitCtFileType f;
  ctFileReaderOpenFile(argv[1], &f);
  const int b      = atoi(argv[2]);
  const int n      = (int)f.header.numberOfImages;
  const size_t pix = (size_t)f.header.imageDimension[0] * f.header.imageDimension[1];
  if (b < 1 || n % b) {
    fprintf(stderr, "b = %d does not divide %d images\n", b, n);
    return 2;
  }
  double **mat = malloc(sizeof(double *) * b);
  float **img  = malloc(sizeof(float *) * b);
  for (int i = 0; i < b; i++) {
    mat[i] = malloc(12 * sizeof(double));
    img[i] = aligned_alloc(64, (pix * sizeof(float) + 63) / 64 * 64);
    if (!mat[i] || !img[i])
      return 3;
  }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int left = n; left > 0; left -= b)
    for (int i = 0; i < b; i++)
      ctFileReaderReadImage(&f, mat[i], img[i]);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  ctFileReaderClose(&f);
  const double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  const double bytes = (double)n * (12.0 * sizeof(double) + (double)pix * sizeof(float));
  printf("{\"images\": %d, \"bytes\": %.0f, \"seconds\": %.6f}\n", n, bytes, s);
  return 0;
}
