/*
 * recon_cli -- the command-line drivers of NUDIM's reconstructions on an
 * MI355X through libnudim_hip.so: `reconstruction` (src/reconstruction.c:50-733)
 * and `reconstruction_parallel` (src/reconstruction-parallel.c:100-890, its
 * MPI master/worker split becomes one GPU job).
 *
 * Same command lines (src/reconstruction.c:91-243, reconstruction-parallel.c:
 * 108-249), validation messages and exit codes (:245-302 / :250-310), same
 * files: ./files/<prefix>-<j>.txt in (:341-358), recovered-0.txt (%le, full
 * grid, :531-540) and recovered.txt (%f, central nx x nx crop, :707-719),
 * the --print line (:603-604).  The parallel program differs as its source
 * does: --height without --center (the band is centred at np/2, :813), TV
 * over the whole volume (:766), DIFFUSION with lambda 0.4 (:764), the
 * gradient-magnitude dumps grad_before.txt / grad_after.txt at the first
 * pass (:770-791, show_grad_mag, denoise.c:381-402) instead of the density
 * dumps at pass 100 (reconstruction.c:637-658), the -parallel file names
 * (:613, :843) and the closing "Total time used" line (:883).
 *
 * Device arithmetic: FP64 (the plain path, reference NFFT sizes) unless the
 * environment sets NUDIM_PRECISION=fp32 (the FP32 batched form).
 * NUDIM_DEVICE selects the HIP device.
 *
 * Differences, all where the reference's behaviour is undefined: a missing or
 * short projection file and an unknown --method are reported with a message
 * and exit code 1 (the reference dereferences a NULL FILE* or reads an
 * uninitialised enum); --random uses MT19937 seeded as GSL's default
 * generator (seed 0 -> 4357), in the serial program's fill order (the
 * parallel program's order depends on its MPI rank count).
 *
 * Text files are read and written by several host threads (textio.c); the
 * values parsed and the bytes written are those of the reference's
 * fscanf / fprintf calls.
 */
#include <getopt.h>
#include <math.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nudim.h"
#include "recon_cli.h"
#include "textio.h"

static void show_help(const char *s, int parallel) {
  printf("\nSyntax: %s [options]\n\n", s);
  printf(
      "Purpose: Reconstruct 3D object from series of equally-angled tomograms.\n"
      "         To be noted, rotation axes is along direction Z, perpendicular to paper. X is horizontal while Y is vertical.\n"
      "         And rotation angle is between X direction and incident rays.\n"
      "         Thereby, the size of each projection is nx*nz , and nx is first ordered, namely projection[nz][nx].\n"
      "\n"
      "-h, --help                    Display this help message.\n"
      "\n"
      "    --prefix=<prefix name>.   Prefix name of series of projections.\n"
      "    --nx=<nx>                 X dimension size of projection.\n"
      "    --nz=<nz>                 Z dimension size of projection.\n"
      "-p, --pad=<np>                Padded size along radial direction, which must be at least larger than nx to achieve oversampling.\n"
      "    --max=<maximum gama>.     Maximum angle of tomograms.\n"
      "    --min=<minimum gama>.     Minimum angle of tomograms.\n"
      "-t, --interval=<delta angle>. Angular sampling interval.\n"
      "    --padangle=<pad interval>.Filling unknown region every padangle, which should be divided by the above interval.\n"
      "    --weight.                 Add weight to fourier slices to compensate for uneven sampling density among different frequencies.\n"
      "    --denoise.                Filtering 3D density at each iteration.\n"
      "    --method=<DIFFUSION/TV>.  Use anisotropic diffusion or total variation to denoise 3D density.\n"
      "    --nhio=<nhio>.            Number of HIO iterations.\n"
      "    --ner=<ner>.              Number of ER iterations.\n"
      "%s"
      "    --random.                 Padding unknown projections with randoms or zeros.\n"
      "    --print.                  Print out R factor with each cycle of iteration.\n"
      "\n", parallel ? "    --height.                 Sample height for constraint.\n"
                      : "    --height=<nheight>.       Sample height for constraint (in pixel).\n"
                        "    --center=<ncenter>.       Sample center for constraint (in pixel).\n");
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int fail_lib(const char *what) {
  fprintf(stderr, "%s: %s\n", what, nudim_last_error());
  return 1;
}

/* grad_mag of denoise.c:381-402 on img [nz][ny][nx]: the density on the
 * border, the backward-difference gradient magnitude inside */
static void grad_mag(const double *img, int nz, int ny, int nx, double *out) {
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nx; k++) {
        size_t o = ((size_t)i * ny + j) * nx + k;
        if (i < 1 || i >= nz - 1 || j < 1 || j >= ny - 1 || k < 1 || k >= nx - 1) {
          out[o] = img[o];
          continue;
        }
        out[o] = sqrt(pow(img[o] - img[o - (size_t)ny * nx], 2) + pow(img[o] - img[o - nx], 2) +
                      pow(img[o] - img[o - 1], 2));
      }
}

int nudim_recon_main(int argc, char *argv[], int parallel) {
  struct timespec t_start;
  clock_gettime(CLOCK_MONOTONIC, &t_start);
  int c;
  char *prefix = NULL, *method = NULL, *rval;
  double interval = -1, padangle = -1, gam_max = 100, gam_min = 100;
  int nx = -1, nz = -1, nheight = -1, ncenter = -1, np = -1, nhio = -1, ner = -1;
  int config_denoise = 0, config_random = 0, config_weight = 0, config_print = 0, config_height = 0;
  int denoise_m = NUDIM_TV;

  const struct option longopts[] = {
      {"help", 0, NULL, 'h'},    {"pad", 1, NULL, 'p'},      {"interval", 1, NULL, 't'},
      {"prefix", 1, NULL, 2},    {"nx", 1, NULL, 3},         {"nz", 1, NULL, 4},
      {"max", 1, NULL, 5},       {"min", 1, NULL, 6},        {"method", 1, NULL, 7},
      {"nhio", 1, NULL, 8},      {"ner", 1, NULL, 9},        {"denoise", 0, NULL, 10},
      {"random", 0, NULL, 11},   {"print", 0, NULL, 12},     {"padangle", 1, NULL, 13},
      {"height", 1, NULL, 14},   {"weight", 0, NULL, 16},    {"center", 1, NULL, 15},
      {0, 0, NULL, 0}};
  /* the parallel program has no --center (reconstruction-parallel.c:108-127) */
  struct option opts[sizeof(longopts) / sizeof(longopts[0])];
  memcpy(opts, longopts, sizeof(longopts));
  if (parallel) opts[17] = opts[18];

  while ((c = getopt_long(argc, argv, "hp:t:", opts, NULL)) != -1) {
    switch (c) {
      case 'h': show_help(argv[0], parallel); return 0;
      case 'p':
        np = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid X dimension!\n"); return 1; }
        break;
      case 't':
        interval = strtod(optarg, &rval);
        if (*rval != '\0') { fprintf(stderr, "Invalid angular sampling interval!\n"); return 1; }
        break;
      case 2: prefix = strdup(optarg); break;
      case 3:
        nx = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid X dimension!\n"); return 1; }
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
      case 7: method = strdup(optarg); break;
      case 8:
        nhio = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid number of HIO!\n"); return 1; }
        break;
      case 9:
        ner = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid number of ER!\n"); return 1; }
        break;
      case 10: config_denoise = 1; break;
      case 11: config_random = 1; break;
      case 12: config_print = 1; break;
      case 13:
        padangle = strtod(optarg, &rval);
        if (*rval != '\0') { fprintf(stderr, "Invalid padded angular interval!\n"); return 1; }
        break;
      case 14:
        config_height = 1;
        nheight = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid sample height!\n"); return 1; }
        break;
      case 15:
        config_height = 1;
        ncenter = strtol(optarg, &rval, 10);
        if (*rval != '\0') { fprintf(stderr, "Invalid sample center!\n"); return 1; }
        break;
      case 16: config_weight = 1; break;
      default: return 1;
    }
  }

  if (prefix == NULL) {
    fprintf(stderr, "You need to provide prefix name of tomogram files with --prefix.\n");
    return 1;
  }
  if (nx < 0 || nz < 0) {
    fprintf(stderr, "You need to provide valid projection size with --nx, --nz.\n");
    return 1;
  }
  if (np < 0) {
    fprintf(stderr, "You need to provide valid projection dimension in XY plane with --pad.\n");
    return 1;
  }
  if (nhio < 0) {
    fprintf(stderr, "You need to provide valid nhio with --nhio.\n");
    return 1;
  }
  if (ner < 0) {
    fprintf(stderr, "You need to provide valid ner with --ner.\n");
    return 1;
  }
  if (interval < 0) {
    fprintf(stderr, "You need to provide valid angular sampling interval with --interval.\n");
    return 1;
  }
  if (padangle < 0 || interval / padangle - (int)(interval / padangle + 0.1) > 1e-5) {
    fprintf(stderr, "You need to provide valid padded angular sampling interval with --padangle.\n");
    return 1;
  }
  if (gam_max < -90 || gam_max > 90 || gam_min < -90 || gam_max > 90) {
    fprintf(stderr, "You need to provide valid scope of tomogram angles (-90,90] with --max, --min.\n");
    return 1;
  }
  if (config_denoise && method == NULL) {
    fprintf(stderr, "You must provide valid denoising method or delete --denoise.\n");
    return 1;
  }
  if (parallel && config_height && nheight <= 0) {
    fprintf(stderr, "You must provide valid sample height --height.\n");
    return 1;
  }
  if (!parallel && config_height && (nheight <= 0 || ncenter <= 0)) {
    fprintf(stderr, "You must provide valid both sample height and center with --height and --center.\n");
    return 1;
  }
  if (parallel && config_height) ncenter = np / 2; /* band (np-h)/2 .. (np+h)/2, :813 */
  if (method != NULL) {
    if (strcmp(method, "DIFFUSION") == 0) {
      denoise_m = NUDIM_DIFFUSION;
    } else if (strcmp(method, "TV") == 0) {
      denoise_m = NUDIM_TV;
    } else if (config_denoise) {
      fprintf(stderr, "Unknown denoising method %s (use DIFFUSION or TV).\n", method);
      return 1;
    }
    free(method);
  }

  nudim_params prm;
  nudim_default_params(&prm);
  prm.nx = nx; prm.nz = nz; prm.np = np;
  prm.gam_min = gam_min; prm.gam_max = gam_max; prm.interval = interval; prm.padangle = padangle;
  prm.weight = config_weight; prm.random = config_random; prm.seed = 0;
  prm.config_height = config_height; prm.nheight = nheight; prm.ncenter = ncenter;
  prm.denoise = config_denoise; prm.method = denoise_m;
  prm.nhio = nhio; prm.ner = ner;
  if (parallel) {
    prm.tv_full = 1;
    prm.diffusion_lambda = 0.4f;
  }
  const char *prec = getenv("NUDIM_PRECISION");
  prm.precision = (prec && (strcmp(prec, "fp32") == 0 || strcmp(prec, "FP32") == 0)) ? NUDIM_FP32 : NUDIM_FP64_PLAIN;
  /* without density compensation the polar normal equations are badly
     scaled (node density ~ 1/r) and the FP32 CGNR ends ~5e-2 from the FP64
     result after its ten steps; the FP64 path is accurate either way */
  if (prm.precision == NUDIM_FP32 && !config_weight)
    fprintf(stderr, "warning: NUDIM_PRECISION=fp32 without --weight is accurate to about 5e-2 only; "
                    "add --weight or use the FP64 path\n");
  const char *dev = getenv("NUDIM_DEVICE");
  prm.device = dev ? atoi(dev) : 0;

  /* phase times on stderr with NUDIM_CLI_TIMES set (stdout as the reference's) */
  double tp[7];
  tp[0] = now_s();
  nudim_ctx *ctx = NULL;
  if (nudim_create(&prm, &ctx) != 0) return fail_lib("nudim_create");
  tp[1] = now_s();
  int nmeas = 0, M = 0, nang = 0;
  nudim_sizes(ctx, &nmeas, &M, &nang);

  /* read in the tilt series, src/reconstruction.c:333-358 */
  size_t per = (size_t)nx * nz;
  double *proj = (double *)malloc(sizeof(double) * per * nmeas);
  if (!proj) { fprintf(stderr, "out of memory\n"); return 1; }
  const int io_threads = textio_threads();
  for (int j = 0; j < nmeas; j++) {
    char filename[1024];
    snprintf(filename, 1023, "./files/%s-%i.txt", prefix, j);
    long got = textio_read_doubles(filename, proj + (size_t)j * per, (long)per, io_threads);
    if (got < 0) {
      fprintf(stderr, "Error : Projection file %s doesn't exist!\n", filename);
      return 1;
    }
    if ((size_t)got < per) {
      fprintf(stderr, "Error : Projection file %s holds fewer than %zu values!\n", filename, per);
      return 1;
    }
  }
  free(prefix);
  tp[2] = now_s();

  if (nudim_load_projections(ctx, proj) != 0) return fail_lib("nudim_load_projections");
  free(proj);
  if (nudim_initial(ctx) != 0) return fail_lib("nudim_initial");
  tp[3] = now_s();

  size_t V = (size_t)np * np * nz;
  double *img = (double *)malloc(sizeof(double) * V);
  if (!img) { fprintf(stderr, "out of memory\n"); return 1; }
  if (nudim_get_volume(ctx, img) != 0) return fail_lib("nudim_get_volume");
  const char *f0name = parallel ? "recovered-0-parallel.txt" : "recovered-0.txt";
  if (textio_write_doubles(f0name, img, (long)V, "%le\n", io_threads) != 0) {
    fprintf(stderr, "cannot write %s\n", f0name);
    return 1;
  }
  tp[4] = now_s();

  /* density before / after the denoiser at pass 100 (serial, :637-658), or
   * its gradient magnitude at pass 0 (parallel, :770-791) */
  const int dump_pass = parallel ? 0 : 100;
  double *before = NULL, *after = NULL;
  if (config_denoise && nhio + ner > dump_pass) {
    before = (double *)malloc(sizeof(double) * V);
    after = (double *)malloc(sizeof(double) * V);
    if (!before || !after) { fprintf(stderr, "out of memory\n"); return 1; }
    nudim_set_dump(ctx, dump_pass, before, after);
  }

  for (int nn = 0; nn < nhio + ner; nn++) {
    double r = 0;
    if (nudim_iterate(ctx, 1, &r) != 0) return fail_lib("nudim_iterate");
    if (config_print) printf("Iteration: %d Measured Residual: %f\n", nn, r);
    if (nn == dump_pass && before) {
      if (parallel) {
        double *g = (double *)malloc(sizeof(double) * V);
        if (!g) { fprintf(stderr, "out of memory\n"); return 1; }
        grad_mag(before, nz, np, np, g);
        memcpy(before, g, sizeof(double) * V);
        grad_mag(after, nz, np, np, g);
        memcpy(after, g, sizeof(double) * V);
        free(g);
      }
      if (textio_write_doubles("grad_before.txt", before, (long)V, "%f\n", io_threads) != 0 ||
          textio_write_doubles("grad_after.txt", after, (long)V, "%f\n", io_threads) != 0) {
        fprintf(stderr, "cannot write grad_before.txt / grad_after.txt\n");
        return 1;
      }
    }
  }

  tp[5] = now_s();
  if (nudim_get_volume(ctx, img) != 0) return fail_lib("nudim_get_volume");
  /* central nx x nx crop of every plane, in the reference's loop order */
  size_t nc = 0;
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++) {
        if (j >= (np + nx) / 2 || j < (np - nx) / 2 || k >= (np + nx) / 2 || k < (np - nx) / 2) continue;
        img[nc++] = img[((size_t)i * np + j) * np + k];
      }
  const char *fname = parallel ? "recovered-parallel.txt" : "recovered.txt";
  if (textio_write_doubles(fname, img, (long)nc, "%f\n", io_threads) != 0) {
    fprintf(stderr, "cannot write %s\n", fname);
    return 1;
  }
  tp[6] = now_s();
  if (getenv("NUDIM_CLI_TIMES"))
    fprintf(stderr,
            "nudim-cli-times create %.3f read %.3f load+initial %.3f write0 %.3f iterate %.3f write %.3f "
            "io-threads %d\n",
            tp[1] - tp[0], tp[2] - tp[1], tp[3] - tp[2], tp[4] - tp[3], tp[5] - tp[4], tp[6] - tp[5], io_threads);

  free(img);
  free(before);
  free(after);
  nudim_destroy(ctx);
  if (parallel) {
    struct timespec t_end;
    clock_gettime(CLOCK_MONOTONIC, &t_end);
    printf("Total time used : %f seconds.\n",
           (double)(t_end.tv_sec - t_start.tv_sec) + 1e-9 * (double)(t_end.tv_nsec - t_start.tv_nsec));
  }
  return 0;
}
