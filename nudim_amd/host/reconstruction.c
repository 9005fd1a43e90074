/* reconstruction -- drop-in for NUDIM's serial `reconstruction` (src/reconstruction.c). */
#include "recon_cli.h"

int main(int argc, char *argv[]) { return nudim_recon_main(argc, argv, 0); }
