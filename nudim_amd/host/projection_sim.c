/*
 * projection_sim -- drop-in for NUDIM's tilt-series simulator
 * (src/projection_sim.c:71-617), method 2D1D on an MI355X through
 * libnudim_hip.so (nudim_project_2d1d: FFT along z, per-plane 2D NFFT at the
 * polar nodes, inverse 2D FFT per tilt, FP64).
 *
 * Same command line (:96-202), messages and exit codes (:204-245), input
 * (a text density map read as data[z][y][x], :257-266) and output
 * (./files/prj-<j>.txt, one "%f\n" value per line, projection[nz][np],
 * :396-403).  --noise --snr adds Gaussian noise of standard deviation
 * sigma(slice)/snr to each tilt (add_noise, :45-67), drawn with the
 * Box-Muller polar method from MT19937 seeded as GSL's default generator
 * (seed 0 -> 4357, one stream over the tilts in order).
 *
 * Differences, all where the reference's behaviour is undefined or absent:
 * the 3D and DFT methods are not built (message, exit code 1); nx must equal
 * ny and nx, nz, np be even; --min above 90 is rejected with the --max/--min
 * message (:229 tests --max twice); an unknown --method, a short density
 * file and a missing ./files directory are reported with exit code 1.
 */
#include <getopt.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nudim.h"
#include "textio.h"

static void show_help(const char *s) {
  printf("\nSyntax: %s [options]\n\n", s);
  printf(
      "Purpose: Simulate serials of tomograms from 3D density arranged in dimensions nx*ny*nz.\n"
      "         To be noted, rotation axes is along direction Z, perpendicular to paper. X is horizontal while Y is vertical.\n"
      "         And rotation angle is between X direction and incident rays.\n"
      "         Thereby, the size of each projection is np*nz.\n"
      "\n"
      "-h, --help                    Display this help message.\n"
      "\n"
      "-i, --infile=<file>           File from which to get the 3D density.\n"
      "    --nx=<nx>                 X dimension.\n"
      "    --ny=<ny>                 Y dimension.\n"
      "    --nz=<nz>                 Z dimension.\n"
      "-p, --pad=<np>                Dimension of projection in XY plane.\n"
      "    --max=<maximum gama>.     Maximum angle of tomograms for simulation.\n"
      "    --min=<minimum gama>.     Minimum angle of tomograms for simulation.\n"
      "-t, --interval=<delta angle>. Angular sampling interval.\n"
      "    --noise.                  Add noise to tomograms.\n"
      "    --snr=<value>.            Signal to noise level.\n"
      "-m  --method=<2D1D/3D/DFT>.   Use 2D NFFT + 1D FFT (default) or 3D NFFT (slow) or Discrete Fourier transform.\n"
      "\n");
}

/* MT19937 (Matsumoto and Nishimura 1998) with GSL's seeding */
typedef struct {
  uint32_t mt[624];
  int i;
} mt_t;

static void mt_seed(mt_t *r, unsigned long s) {
  if (s == 0) s = 4357;
  r->mt[0] = (uint32_t)(s & 0xffffffffUL);
  for (int i = 1; i < 624; i++) r->mt[i] = (uint32_t)(1812433253UL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + i);
  r->i = 624;
}

static uint32_t mt_next(mt_t *r) {
  if (r->i >= 624) {
    for (int k = 0; k < 624; k++) {
      uint32_t y = (r->mt[k] & 0x80000000u) | (r->mt[(k + 1) % 624] & 0x7fffffffu);
      r->mt[k] = r->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    r->i = 0;
  }
  uint32_t y = r->mt[r->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* uniform on (0, 1) */
static double uniform_pos(mt_t *r) {
  double x;
  do x = mt_next(r) / 4294967296.0;
  while (x == 0);
  return x;
}

/* normal deviate of standard deviation sigma, polar method (Knuth, TAOCP
 * vol. 2, 3.4.1 algorithm P) */
static double gaussian(mt_t *r, double sigma) {
  double x, y, r2;
  do {
    x = -1 + 2 * uniform_pos(r);
    y = -1 + 2 * uniform_pos(r);
    r2 = x * x + y * y;
  } while (r2 > 1.0 || r2 == 0);
  return sigma * y * sqrt(-2.0 * log(r2) / r2);
}

int main(int argc, char *argv[]) {
  int c;
  char *infile = NULL, *method = NULL, *rval;
  double interval = -1, gam_max = 100, gam_min = 100, snr = -1;
  int nx = -1, ny = -1, nz = -1, np = -1, config_noise = 0;

  const struct option longopts[] = {
      {"help", 0, NULL, 'h'}, {"infile", 1, NULL, 'i'}, {"pad", 1, NULL, 'p'},   {"interval", 1, NULL, 't'},
      {"method", 1, NULL, 'm'}, {"nx", 1, NULL, 2},     {"ny", 1, NULL, 3},      {"nz", 1, NULL, 4},
      {"max", 1, NULL, 5},      {"min", 1, NULL, 6},    {"snr", 1, NULL, 7},     {"noise", 0, NULL, 8},
      {0, 0, NULL, 0}};

  while ((c = getopt_long(argc, argv, "hi:p:t:m:", longopts, NULL)) != -1) {
    switch (c) {
      case 'h': show_help(argv[0]); return 0;
      case 'i': infile = strdup(optarg); break;
      case 'p':
        np = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid X dimension!\n"); return 1; }
        break;
      case 't':
        interval = strtod(optarg, &rval);
        if (*rval != '\0') { fprintf(stderr, "Invalid angular sampling interval!\n"); return 1; }
        break;
      case 'm': method = strdup(optarg); break;
      case 2:
        nx = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid X dimension!\n"); return 1; }
        break;
      case 3:
        ny = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid Y dimension!\n"); return 1; }
        break;
      case 4:
        nz = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid Z dimension!\n"); return 1; }
        break;
      case 5:
        gam_max = strtod(optarg, &rval);
        if (*rval != '\0') { fprintf(stderr, "Invalid maximum collecting angle!\n"); return 1; }
        break;
      case 6:
        gam_min = strtod(optarg, &rval);
        if (*rval != '\0') { fprintf(stderr, "Invalid minimum collecting angle!\n"); return 1; }
        break;
      case 7:
        snr = strtod(optarg, &rval);
        if (*rval != '\0') { fprintf(stderr, "Invalid signal to noise ratio!\n"); return 1; }
        break;
      case 8: config_noise = 1; break;
      default: return 1;
    }
  }

  if (infile == NULL) {
    fprintf(stderr, "You need to provide 3D density with --infile.\n");
    return 1;
  }
  if (nx < 0 || ny < 0 || nz < 0) {
    fprintf(stderr, "You need to provide valid dimensions with --nx, --ny, --nz.\n");
    return 1;
  }
  if (np < 0) {
    fprintf(stderr, "You need to provide valid projection dimension in XY plane with --pad.\n");
    return 1;
  }
  if (interval < 0) {
    fprintf(stderr, "You need to provide valid angular sampling interval with --interval.\n");
    return 1;
  }
  if (gam_max < -90 || gam_max > 90 || gam_min < -90 || gam_min > 90) {
    fprintf(stderr, "You need to provide valid scope of tomogram angles (-90,90] with --max, --min.\n");
    return 1;
  }
  if (method == NULL) {
    fprintf(stderr, "You didn't specify simulating method, so I'm  using"
                    " the 2D1D method, which is fastest.\n");
  } else if (strcmp(method, "2D1D") != 0) {
    if (strcmp(method, "3D") == 0 || strcmp(method, "DFT") == 0)
      fprintf(stderr, "Method %s is not built here: use the 2D1D method.\n", method);
    else
      fprintf(stderr, "Unknown simulating method %s (2D1D, 3D or DFT).\n", method);
    return 1;
  }
  if (config_noise && snr < 0) {
    fprintf(stderr, "You must provide valid signal to noise ratio with --snr or delete --noise.\n");
    return 1;
  }
  if (nx != ny || (nx & 1) || (nz & 1) || (np & 1) || nx == 0 || nz == 0 || np == 0) {
    fprintf(stderr, "This build simulates square even x-y planes (nx == ny) with even nz and np.\n");
    return 1;
  }

  /* the density, data[z][y][x] */
  const long V = (long)nx * ny * nz;
  double *data = (double *)malloc(sizeof(double) * V);
  long got = textio_read_doubles(infile, data, V, textio_threads());
  if (got < 0) {
    fprintf(stderr, "Error : Data file %s doesn't exist!\n ", infile);
    return 1;
  }
  if (got < V) {
    fprintf(stderr, "Error : Data file %s holds %ld of the %ld values!\n", infile, got, V);
    return 1;
  }
  free(infile);

  const int nslices = (int)((gam_max - gam_min) / interval + 0.5) + 1;
  if (nslices < 1) {
    fprintf(stderr, "You need to provide valid scope of tomogram angles (-90,90] with --max, --min.\n");
    return 1;
  }
  const long S = (long)np * nz;
  double *proj = (double *)malloc(sizeof(double) * S * nslices);
  if (nudim_project_2d1d(nx, ny, nz, np, gam_min, gam_max, interval, data, proj) != NUDIM_OK) {
    fprintf(stderr, "projection_sim: %s\n", nudim_last_error());
    return 1;
  }
  free(data);

  mt_t rng;
  if (config_noise) mt_seed(&rng, 0);
  const double scale = (double)np * nz;
  for (int j = 0; j < nslices; j++) {
    double *p = proj + (size_t)j * S;
    if (config_noise) {
      /* add_noise on the unscaled slice (:45-67) */
      double sum = 0, sum2 = 0;
      for (long i = 0; i < S; i++) {
        double v = p[i] * scale;
        sum += v;
        sum2 += v * v;
      }
      double sigma = sqrt(sum2 / S - pow(sum / S, 2));
      for (long i = 0; i < S; i++) p[i] = (p[i] * scale + gaussian(&rng, sigma / snr)) / scale;
    }
    char filename[1024];
    snprintf(filename, 1023, "./files/prj-%i.txt", j);
    if (textio_write_doubles(filename, p, S, "%f\n", textio_threads()) != 0) {
      fprintf(stderr, "Error : cannot write %s (does ./files exist?)\n", filename);
      return 1;
    }
  }
  free(proj);
  return 0;
}
