/* C ABI of the CPU oracle -- TEST INFRASTRUCTURE ONLY (see orc_common.hpp). */
#ifndef ORACLE_H
#define ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* single-rank grid description; bc[] = {xl, xr, yl, yr, zl, zr} codes of euler3D.hpp:90-93 */
typedef struct {
  long nx, ny, nz;
  int nchem;
  int bc[6];
  double gamma, dx, dy, dz;
} orc_grid;

void orc_set_threads(int n);
int orc_max_threads(void);
int orc_face_flux(const double *stencil, int dir, int nchem, double gamma, double *fface);
int orc_fe(const orc_grid *g, const double *const *w, double *const *wdot);
double orc_max_wave_speed(const orc_grid *g, const double *const *w);
double orc_stability_dt(const orc_grid *g, const double *const *w, double cfl);
void orc_face_flux_opcount(int nchem, unsigned long long *out);

#ifdef __cplusplus
}
#endif
#endif
