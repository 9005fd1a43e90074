/* recon_cli -- shared driver of the reconstruction programs (see recon_cli.c). */
#ifndef NUDIM_RECON_CLI_H
#define NUDIM_RECON_CLI_H

/* parallel = 0: src/reconstruction.c; 1: src/reconstruction-parallel.c */
int nudim_recon_main(int argc, char *argv[], int parallel);

#endif
