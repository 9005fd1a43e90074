/*
 * mrcread -- drop-in for NUDIM's experimental-data converter
 * (examples/exp/mrcread.c:3-90 with mrcfile.c:5-210): an MRC tilt stack to
 * the ./files/prj-<j>.txt projections that `reconstruction` reads.
 *
 * What the reference does, and this program does the same:
 *   - reads the 1024-byte header in the machine's byte order and requires
 *     "MAP" at byte 208 (mrc_read_head, mrcfile.c:5-17); prints the header
 *     lines of mrcread.c:19-28, "Image type" and "Space Group" on one line;
 *   - takes section k (z) as nx*ny values at byte 1024 + k*nx*ny*size, for
 *     modes 0 (unsigned byte), 1 (signed short) and 2 (float), converted to
 *     float (mrc_read_slice axis 'z', mrcfile.c:182-207); the extended
 *     header is not skipped (the reference reads from byte 1024 whatever
 *     `next` says);
 *   - drops section 0; for k = 1..nz-1 prints "<max> <min> " of the section
 *     (extremes started at -100000 / 100000, mrcread.c:45), moves value
 *     (j, i) to column (i-61) mod nx and, for the first 61 columns, to row
 *     (j-1) mod ny (mrcread.c:49-63), and writes max - value + min (contrast
 *     inverted, in double) as "%f\n" to ./files/prj-<nz-1-k>.txt
 *     (mrcread.c:67-80).
 *
 * The reference has no command line (input BBa.ali).  Here the optional
 * first argument names the input; the default is BBa.ali.
 *
 * Differences, all where the reference's behaviour is undefined: a file that
 * is not MRC, a mode other than 0, 1, 2, a stack shorter than its header
 * says and a missing ./files directory are reported with exit code 1 (the
 * reference goes on with unset or unread data); nx below 61 wraps the
 * column with a non-negative modulus.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "textio.h"

#define HEADER_BYTES 1024
#define SHIFT_X 61 /* columns moved to the end of the row (mrcread.c:49) */
#define SHIFT_Y 1  /* rows those columns move down by */

static int32_t geti(const unsigned char *h, int off) {
  int32_t v;
  memcpy(&v, h + off, 4);
  return v;
}
static float getf(const unsigned char *h, int off) {
  float v;
  memcpy(&v, h + off, 4);
  return v;
}
static int16_t gets16(const unsigned char *h, int off) {
  int16_t v;
  memcpy(&v, h + off, 2);
  return v;
}

int main(int argc, char *argv[]) {
  const char *path = argc > 1 ? argv[1] : "BBa.ali";
  FILE *fin = fopen(path, "rb");
  if (!fin) {
    fprintf(stderr, "cannot open %s\n", path);
    return 1;
  }
  unsigned char h[HEADER_BYTES];
  if (fread(h, 1, HEADER_BYTES, fin) != HEADER_BYTES || h[208] != 'M' || h[209] != 'A' || h[210] != 'P') {
    printf("Error with function 'mrc_read_head()'! Warning: Not MRC format! \n");
    fclose(fin);
    return 1;
  }
  const int nx = geti(h, 0), ny = geti(h, 4), nz = geti(h, 8), mode = geti(h, 12);

  printf("Data Dimension: %d %d %d\n", nx, ny, nz);
  printf("Data type (Mode): %d\n", mode);
  printf("Nstart: %d %d %d\n", geti(h, 16), geti(h, 20), geti(h, 24));
  printf("Grid size: %d %d %d\n", geti(h, 28), geti(h, 32), geti(h, 36));
  printf("Length: %f %f %f\n", getf(h, 40), getf(h, 44), getf(h, 48));
  printf("Angles: %f %f %f\n", getf(h, 52), getf(h, 56), getf(h, 60));
  printf("Map column: %d %d %d\n", geti(h, 64), geti(h, 68), geti(h, 72));
  printf("Density: %f %f %f\n", getf(h, 76), getf(h, 80), getf(h, 84));
  printf("Image type: %d", gets16(h, 88));
  printf("Space Group: %d\n", gets16(h, 90));
  fflush(stdout);

  const int psize = mode == 0 ? 1 : mode == 1 ? 2 : mode == 2 ? 4 : 0;
  if (!psize) {
    fprintf(stderr, "MRC mode %d is not read (modes 0, 1 and 2 are)\n", mode);
    fclose(fin);
    return 1;
  }
  if (nx <= 0 || ny <= 0 || nz <= 0) {
    fprintf(stderr, "invalid MRC dimensions %d %d %d\n", nx, ny, nz);
    fclose(fin);
    return 1;
  }
  const size_t plane = (size_t)nx * ny;
  /* the whole stack in one read; section 0 is read but not written */
  unsigned char *raw = (unsigned char *)malloc(plane * nz * psize);
  float *sec = (float *)malloc(sizeof(float) * plane);
  double *out = (double *)malloc(sizeof(double) * plane);
  size_t *dest = (size_t *)malloc(sizeof(size_t) * plane);
  if (!raw || !sec || !out || !dest) {
    fprintf(stderr, "out of memory\n");
    return 1;
  }
  if (fread(raw, psize, plane * nz, fin) != plane * nz) {
    fprintf(stderr, "%s holds fewer than the %d x %d x %d values its header gives\n", path, nx, ny, nz);
    fclose(fin);
    return 1;
  }
  fclose(fin);

  /* destination of value (j, i): one permutation shared by every section */
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int col = ((i - SHIFT_X) % nx + nx) % nx;
      const int row = i < SHIFT_X ? ((j - SHIFT_Y) % ny + ny) % ny : j;
      dest[(size_t)j * nx + i] = (size_t)row * nx + col;
    }

  const int threads = textio_threads();
  for (int k = 1; k < nz; k++) {
    const unsigned char *src = raw + (size_t)k * plane * psize;
    for (size_t t = 0; t < plane; t++) {
      if (mode == 2) memcpy(&sec[t], src + 4 * t, 4);
      else if (mode == 1) {
        int16_t s;
        memcpy(&s, src + 2 * t, 2);
        sec[t] = (float)s;
      } else sec[t] = (float)src[t];
    }
    double mx = -100000, mn = 100000;
    for (size_t t = 0; t < plane; t++) {
      if (sec[t] > mx) mx = sec[t];
      if (sec[t] < mn) mn = sec[t];
    }
    for (size_t t = 0; t < plane; t++) out[dest[t]] = mx - (double)sec[t] + mn;
    printf("%f %f \n", mx, mn);
    char name[64];
    snprintf(name, sizeof(name), "./files/prj-%d.txt", nz - 1 - k);
    if (textio_write_doubles(name, out, (long)plane, "%f\n", threads) != 0) {
      fprintf(stderr, "cannot write %s\n", name);
      return 1;
    }
  }
  free(raw);
  free(sec);
  free(out);
  free(dest);
  return 0;
}
