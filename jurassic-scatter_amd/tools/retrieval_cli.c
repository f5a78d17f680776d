/*
 * retrieval -- command-line optimal-estimation retrieval on the GPU library,
 * with the reference CLI's arguments (reference src/retrieval.c:5-83):
 *   retrieval <ctl> <dirlist> [KEY VALUE ...]
 * For every directory named in <dirlist>: a priori atm_apr.tab and
 * aero_apr.tab, measurements obs_meas.tab; the outputs of optimal_estimation
 * go into the same directory.  With RESMAX > 0 the retrieval is repeated
 * without the measurements whose residual |1 - y_i / y_meas| reaches RESMAX
 * percent, while any remain.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "jurassic_amd.h"

static void fail(const char *msg) {
  printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__, msg);
  exit(EXIT_FAILURE);
}

int main(int argc, char *argv[]) {
  static ret_t ret;
  static ctl_t ctl;
  static atm_t atm_i, atm_apr;
  static obs_t obs_i, obs_meas;
  static aero_t aero_i, aero_apr;
  if (argc < 3) fail("Give parameters: <ctl> <dirlist>");
  read_ctl(argc, argv, &ctl);
  read_ret(argc, argv, &ctl, &ret);
  FILE *dirlist = fopen(argv[2], "r");
  if (!dirlist) fail("Cannot open directory list!");
  char fmt[32];
  snprintf(fmt, sizeof fmt, "%%%ds", LEN - 1);
  while (fscanf(dirlist, fmt, ret.dir) != EOF) {
    printf("\nRetrieve in directory %s...\n\n", ret.dir);
    read_atm(ret.dir, "atm_apr.tab", &ctl, &atm_apr);
    read_aero(ret.dir, "aero_apr.tab", &ctl, &aero_apr);
    read_obs(ret.dir, "obs_meas.tab", &ctl, &obs_meas);
    for (;;) {
      optimal_estimation(&ret, &ctl, &obs_meas, &obs_i, &atm_apr, &atm_i, &aero_apr, &aero_i);
      /* screen the radiance residuals */
      int nbad = 0;
      for (int id = 0; id < ctl.nd; id++)
        for (int ir = 0; ir < obs_meas.nr; ir++)
          if (ret.resmax > 0 && isfinite(obs_i.rad[id][ir]) && isfinite(obs_meas.rad[id][ir]) &&
              fabs(1 - obs_i.rad[id][ir] / obs_meas.rad[id][ir]) >= ret.resmax / 100) {
            obs_meas.rad[id][ir] = obs_i.rad[id][ir] = NAN;
            nbad++;
          }
      if (nbad > 0 && jur_meas_size(&ctl, &obs_meas) > 0)
        printf("\nFound %d bad measurements. Redo retrieval...\n\n", nbad);
      else
        break;
    }
  }
  fclose(dirlist);
  printf("\nRetrieval done...\n");
  return EXIT_SUCCESS;
}
