/* reconstruction_parallel -- drop-in for NUDIM's `reconstruction_parallel`
 * (src/reconstruction-parallel.c) on one GPU: no MPI ranks, same flags,
 * numerics and files as that program. */
#include "recon_cli.h"

int main(int argc, char *argv[]) { return nudim_recon_main(argc, argv, 1); }
