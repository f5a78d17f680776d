/*
 * tab2bin -- emissivity tables to the binary layout, with the reference
 * CLI's arguments (reference src/tab2bin.c):
 *   tab2bin <ctl> [KEY VALUE ...]
 * writes TBLBASE_<nu>_<emitter>.bin for every emitter and channel of <ctl>
 * from the binary table if present, else the ASCII one.
 */
#include <stdio.h>
#include <stdlib.h>

#include "jurassic_amd.h"

int main(int argc, char *argv[]) {
  static ctl_t ctl;
  if (argc < 2) {
    printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__, "Give parameters: <ctl>");
    return EXIT_FAILURE;
  }
  read_ctl(argc, argv, &ctl);
  jur_tab2bin(&ctl);
  return EXIT_SUCCESS;
}
