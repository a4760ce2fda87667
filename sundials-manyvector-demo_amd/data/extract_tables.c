/* Extract the 50 primordial-chemistry rate tables used by the network from the
 * HDF5 file shipped with the reference (tests/primordial_ode/primordial_tables.h5)
 * into one raw little-endian float64 blob, in the fixed order listed below.
 *
 * The order is the order in which the reference reads them:
 *   rate tables    raja_primordial_network.cpp:472-512  (k01..k22, no k20)
 *   cooling tables raja_primordial_network.cpp:~600-700 (25 tables)
 *   gamma tables   raja_primordial_network.cpp:860-866  (4 tables)
 * Each table has 1024 entries (TSIZE, raja_primordial_network.cpp:43).
 *
 * Build and run (this container only; the GPU box never needs HDF5):
 *   sh make_tables.sh
 */
#include <stdio.h>
#include <stdlib.h>
#include "hdf5.h"
#include "hdf5_hl.h"

#define TSIZE 1024

static const char *names[] = {
  "k01", "k02", "k03", "k04", "k05", "k06", "k07", "k08", "k09", "k10", "k11",
  "k12", "k13", "k14", "k15", "k16", "k17", "k18", "k19", "k21", "k22",
  "brem_brem", "ceHeI_ceHeI", "ceHeII_ceHeII", "ceHI_ceHI", "cie_cooling_cieco",
  "ciHeI_ciHeI", "ciHeII_ciHeII", "ciHeIS_ciHeIS", "ciHI_ciHI", "compton_comp_",
  "gloverabel08_gael", "gloverabel08_gaH2", "gloverabel08_gaHe",
  "gloverabel08_gaHI", "gloverabel08_gaHp", "gloverabel08_h2lte",
  "h2formation_h2mcool", "h2formation_h2mheat", "h2formation_ncrd1",
  "h2formation_ncrd2", "h2formation_ncrn", "reHeII1_reHeII1",
  "reHeII2_reHeII2", "reHeIII_reHeIII", "reHII_reHII",
  "gammaH2_1", "dgammaH2_1_dT", "gammaH2_2", "dgammaH2_2_dT"
};

int main(int argc, char **argv)
{
  if (argc != 3) {
    fprintf(stderr, "usage: %s <primordial_tables.h5> <out.f64>\n", argv[0]);
    return 2;
  }
  const int ntab = (int)(sizeof(names) / sizeof(names[0]));
  hid_t fid = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fid < 0) { fprintf(stderr, "cannot open %s\n", argv[1]); return 1; }
  FILE *out = fopen(argv[2], "wb");
  if (!out) { fprintf(stderr, "cannot write %s\n", argv[2]); return 1; }
  double buf[TSIZE];
  for (int t = 0; t < ntab; t++) {
    char path[64];
    snprintf(path, sizeof(path), "/%s", names[t]);
    hsize_t dims[1] = {0};
    if (H5LTget_dataset_info(fid, path, dims, NULL, NULL) < 0 || dims[0] != TSIZE) {
      fprintf(stderr, "dataset %s missing or not %d long\n", path, TSIZE);
      return 1;
    }
    if (H5LTread_dataset_double(fid, path, buf) < 0) {
      fprintf(stderr, "read of %s failed\n", path);
      return 1;
    }
    if (fwrite(buf, sizeof(double), TSIZE, out) != TSIZE) return 1;
    printf("%s\n", names[t]);
  }
  fclose(out);
  H5Fclose(fid);
  return 0;
}
