/*
 * raytrace -- command-line LOS dump on the GPU library, with the reference
 * CLI's arguments and output (reference src/raytrace.c):
 *   raytrace <ctl> <obs> <atm> [AEROFILE <aero>] [KEY VALUE ...]
 * writes los.<ray> in the working directory for every ray of <obs>: time,
 * altitude, longitude, latitude, pressure, temperature, volume mixing
 * ratios, window extinctions, column densities, the aerosol beta_e / beta_s
 * / beta_a seen at each point and the segment length.
 */
#include <stdio.h>
#include <stdlib.h>

#include "jurassic_amd.h"

static void fail(const char *msg) {
  printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__, msg);
  exit(EXIT_FAILURE);
}

/* The aerosol coefficient of layer il scaled by the point's aerosol factor;
   points outside every layer (il < 0) see none. */
static double aero_at(double (*beta)[NDMAX], int il, int id, double af) {
  return il >= 0 ? beta[il][id] * af : 0.0;
}

int main(int argc, char *argv[]) {
  static atm_t atm;
  static ctl_t ctl;
  static los_t los;
  static obs_t obs;
  static aero_t aero;
  char filename[LEN], aerofile[LEN];
  if (argc < 4) fail("Give parameters: <ctl> <obs> <atm>");
  read_ctl(argc, argv, &ctl);
  scan_ctl(argc, argv, "AEROFILE", -1, "-", aerofile);
  read_obs(NULL, argv[2], &ctl, &obs);
  read_atm(NULL, argv[3], &ctl, &atm);
  if (aerofile[0] != '-' && ctl.sca_n > 0) {
    read_aero(NULL, aerofile, &ctl, &aero);
    get_opt_prop(&ctl, &aero);
  } else if (aerofile[0] == '-' && ctl.sca_n > 0) {
    fail("Please give aerosol file name or set SCA_N=0 for clear air simulation!");
  }
  for (int ir = 0; ir < obs.nr; ir++) {
    raytrace(&ctl, &atm, &obs, &aero, &los, ir);
    snprintf(filename, sizeof filename, "los.%d", ir);
    FILE *out = fopen(filename, "w");
    if (!out) fail("Cannot create los.tab!");
    fprintf(out,
            "# $1 = time (seconds since 2000-01-01T00:00Z)\n"
            "# $2 = LOS point altitude [km]\n"
            "# $3 = LOS point longitude [deg]\n"
            "# $4 = LOS point latitude [deg]\n"
            "# $5 = LOS point pressure [hPa]\n"
            "# $6 = LOS point temperature [K]\n");
    for (int ig = 0; ig < ctl.ng; ig++)
      fprintf(out, "# $%d = LOS point %s volume mixing ratio \n", 7 + ig, ctl.emitter[ig]);
    for (int iw = 0; iw < ctl.nw; iw++)
      fprintf(out, "# $%d = LOS point window %d extinction [1/km]\n", 7 + ctl.ng + iw, iw);
    for (int ig = 0; ig < ctl.ng; ig++)
      fprintf(out, "# $%d = LOS point %s column density [molec/cm^2] \n", 7 + ctl.ng + ctl.nw + ig, ctl.emitter[ig]);
    const int b = 7 + ctl.ng + ctl.nw + ctl.ng;
    for (int id = 0; id < ctl.nd; id++)
      fprintf(out, "# $%d = LOS point beta_e(%.4f cm-1) [km-1] \n", b + id, ctl.nu[id]);
    for (int id = 0; id < ctl.nd; id++)
      fprintf(out, "# $%d = LOS point beta_s(%.4f cm-1) [km-1] \n", b + ctl.nd + id, ctl.nu[id]);
    for (int id = 0; id < ctl.nd; id++)
      fprintf(out, "# $%d = LOS point beta_a(%.4f cm-1) [km-1] \n", b + 2 * ctl.nd + id, ctl.nu[id]);
    fprintf(out, "# $%d = LOS segement length [km] \n", b + 3 * ctl.nd);
    fprintf(out, "\n");
    for (int ip = 0; ip < los.np; ip++) {
      fprintf(out, "%.2f %g %g %g %g %g", obs.time[ir], los.z[ip], los.lon[ip], los.lat[ip], los.p[ip], los.t[ip]);
      for (int ig = 0; ig < ctl.ng; ig++) fprintf(out, " %g", los.q[ip][ig]);
      for (int iw = 0; iw < ctl.nw; iw++) fprintf(out, " %g", los.k[ip][iw]);
      for (int ig = 0; ig < ctl.ng; ig++) fprintf(out, " %g", los.u[ip][ig]);
      for (int id = 0; id < ctl.nd; id++) fprintf(out, " %g", aero_at(aero.beta_e, los.aeroi[ip], id, los.aerofac[ip]));
      for (int id = 0; id < ctl.nd; id++) fprintf(out, " %g", aero_at(aero.beta_s, los.aeroi[ip], id, los.aerofac[ip]));
      for (int id = 0; id < ctl.nd; id++) fprintf(out, " %g", aero_at(aero.beta_a, los.aeroi[ip], id, los.aerofac[ip]));
      fprintf(out, " %g", los.ds[ip]);
      fprintf(out, "\n");
    }
    fclose(out);
    printf("Wrote output to %s \n", filename);
  }
  return EXIT_SUCCESS;
}
