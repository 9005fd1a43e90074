/* Test driver for nudim_amd/host/textio.c (built by tests/test_textio.py).
 *   textio_check write <in.bin> <out.txt> <fmt> <threads>
 *   textio_check read <in.txt> <out.bin> <count> <threads>   (prints the count read) */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "textio.h"

int main(int argc, char **argv) {
  if (argc != 6) return 2;
  int threads = atoi(argv[5]);
  if (strcmp(argv[1], "write") == 0) {
    FILE *fp = fopen(argv[2], "rb");
    if (!fp) return 2;
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp) / 8;
    fseek(fp, 0, SEEK_SET);
    double *v = malloc(8 * (n + 1));
    if (fread(v, 8, n, fp) != (size_t)n) return 2;
    fclose(fp);
    return textio_write_doubles(argv[3], v, n, argv[4], threads) == 0 ? 0 : 1;
  }
  long count = atol(argv[4]);
  double *v = calloc(count + 1, 8);
  long got = textio_read_doubles(argv[2], v, count, threads);
  printf("%ld\n", got);
  FILE *fp = fopen(argv[3], "wb");
  if (!fp) return 2;
  if (got > 0) fwrite(v, 8, got, fp);
  fclose(fp);
  return 0;
}
