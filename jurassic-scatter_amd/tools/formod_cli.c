/*
 * formod -- command-line forward model on the GPU library, with the
 * reference CLI's arguments and keys (reference src/formod.c:24-191):
 *   formod <ctl> <obs> <atm> <rad> [KEY VALUE ...]
 * keys TASK (c: per-emitter contributions), DIRLIST, AEROFILE.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "jurassic_amd.h"

static void call_formod(ctl_t *ctl, const char *wrkdir, const char *obsfile, const char *atmfile,
                        const char *radfile, const char *task, const char *aerofile) {
  static atm_t atm, atm2;
  static obs_t obs;
  static aero_t aero;
  char filename[LEN];
  read_obs(wrkdir, obsfile, ctl, &obs);
  read_atm(wrkdir, atmfile, ctl, &atm);
  if (aerofile[0] != '-' && ctl->sca_n > 0) {
    read_aero(wrkdir, aerofile, ctl, &aero);
    get_opt_prop(ctl, &aero);
  } else if (aerofile[0] == '-' && ctl->sca_n > 0) {
    printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__,
           "Please give aerosol file name or set SCA_N=0 for clear air simulation!");
    exit(EXIT_FAILURE);
  }
  formod(ctl, &atm, &obs, &aero);
  write_obs(wrkdir, radfile, ctl, &obs);
  if (task[0] == 'c' || task[0] == 'C') {
    ctl->ctm_n2 = 0;
    ctl->ctm_o2 = 0;
    for (int ig = 0; ig < ctl->ng; ig++) {
      atm2 = atm;
      for (int iw = 0; iw < ctl->nw; iw++)
        for (int ip = 0; ip < atm2.np; ip++) atm2.k[iw][ip] = 0;
      for (int ig2 = 0; ig2 < ctl->ng; ig2++)
        if (ig2 != ig)
          for (int ip = 0; ip < atm2.np; ip++) atm2.q[ig2][ip] = 0;
      formod(ctl, &atm2, &obs, &aero);
      snprintf(filename, LEN, "%s.%s", radfile, ctl->emitter[ig]);
      write_obs(wrkdir, filename, ctl, &obs);
    }
    atm2 = atm;
    for (int ig = 0; ig < ctl->ng; ig++)
      for (int ip = 0; ip < atm2.np; ip++) atm2.q[ig][ip] = 0;
    formod(ctl, &atm2, &obs, &aero);
    snprintf(filename, LEN, "%s.EXTINCT", radfile);
    write_obs(wrkdir, filename, ctl, &obs);
  }
}

int main(int argc, char *argv[]) {
  static ctl_t ctl;
  static char dirlist[LEN], wrkdir[LEN], task[LEN], aerofile[LEN];
  if (argc < 5) {
    printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__,
           "Give parameters: <ctl> <obs> <atm> <rad>");
    return EXIT_FAILURE;
  }
  read_ctl(argc, argv, &ctl);
  scan_ctl(argc, argv, "TASK", -1, "-", task);
  scan_ctl(argc, argv, "DIRLIST", -1, "-", dirlist);
  scan_ctl(argc, argv, "AEROFILE", -1, "-", aerofile);
  if (dirlist[0] == '-') {
    call_formod(&ctl, NULL, argv[2], argv[3], argv[4], task, aerofile);
  } else {
    FILE *in = fopen(dirlist, "r");
    if (!in) {
      printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__, "Cannot open directory list!");
      return EXIT_FAILURE;
    }
    while (fscanf(in, "%4999s", wrkdir) != EOF) {
      printf("\nWorking directory: %s\n", wrkdir);
      call_formod(&ctl, wrkdir, argv[2], argv[3], argv[4], task, aerofile);
    }
    fclose(in);
  }
  return EXIT_SUCCESS;
}
