/*
 * kernel -- command-line Jacobian on the GPU library, with the reference
 * CLI's arguments (reference src/kernel.c):
 *   kernel <ctl> <obs> <atm> <aero> <kernel> [KEY VALUE ...]
 * <aero> is "-" for clear air.  The matrix is written (measurement space
 * along rows, state space along columns) only with WRITE_MATRIX = 1, as in
 * the reference.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "jurassic_amd.h"

static void fail(const char *msg) {
  printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__, msg);
  exit(EXIT_FAILURE);
}

int main(int argc, char *argv[]) {
  static atm_t atm;
  static ctl_t ctl;
  static obs_t obs;
  static aero_t aero;
  if (argc < 6) fail("Give parameters: <ctl> <obs> <atm> <aero> <kernel>");
  read_ctl(argc, argv, &ctl);
  read_obs(NULL, argv[2], &ctl, &obs);
  read_atm(NULL, argv[3], &ctl, &atm);
  if (strcmp(argv[4], "-") != 0 && ctl.sca_n > 0) {
    read_aero(NULL, argv[4], &ctl, &aero);
    get_opt_prop(&ctl, &aero);
  } else if (strcmp(argv[4], "-") == 0 && ctl.sca_n > 0) {
    fail("Please give aerosol file name or set SCA_N=0 for clear air simulation!");
  }
  const size_t n = jur_state_size(&ctl, &atm, &aero);
  const size_t m = jur_meas_size(&ctl, &obs);
  if (n <= 0) fail("No state vector elements!");
  if (m <= 0) fail("No measurement vector elements!");
  double *k = calloc(m * n, sizeof(double));
  if (!k) fail("Out of memory!");
  jur_gsl_matrix mat = {m, n, n, k, NULL, 0};
  kernel(&ctl, &atm, &obs, &aero, &mat);
  write_matrix(NULL, argv[5], &ctl, &mat, &atm, &aero, &obs, "y", "x", "r");
  free(k);
  return EXIT_SUCCESS;
}
