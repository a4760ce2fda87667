// Solution output / restart input (serial HDF5) and the restart parameter file, in the
// reference's format (io.cpp:645-1162), for one rank. Host code: device fields are staged
// through host buffers with synchronous copies.
#include <hdf5.h>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mvd_io.h"

namespace {

std::string in_dir(const char *dir, const char *name) {
  if (!dir || !*dir) return name;
  std::string s(dir);
  if (s.back() != '/') s += '/';
  return s + name;
}

bool grid_ok(const mvd_params *p) {
  return p && p->nx > 0 && p->ny > 0 && p->nz > 0 && p->nchem >= 0;
}

// code -> CGS factors, in the reference's order of operations (euler3D.hpp:389-391)
void unit_factors(const mvd_params *p, double (&u)[5]) {
  const double M = p->mass_units, L = p->length_units, T = p->time_units;
  const double dens = M / L / L / L, mom = M / L / L / T, ener = M / L / T / T;
  u[0] = dens;
  u[1] = u[2] = u[3] = mom;
  u[4] = ener;
}

bool to_host(const double *src, double *dst, size_t n, bool dev) {
  if (!dev) {
    for (size_t i = 0; i < n; i++) dst[i] = src[i];
    return true;
  }
  return hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
}

bool from_host(const double *src, double *dst, size_t n, bool dev) {
  if (!dev) {
    for (size_t i = 0; i < n; i++) dst[i] = src[i];
    return true;
  }
  return hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
}

const char *const FLUID_NAMES[5] = {"Density", "x-Momentum", "y-Momentum", "z-Momentum",
                                    "TotalEnergy"};

std::string chem_name(int v) {
  char b[16];
  std::snprintf(b, sizeof b, "Chemical-%03hu", (unsigned short)v);
  return b;
}

// one [nz][ny][nx] double dataset written from a host array
bool write_field(hid_t file, hid_t space, const char *name, const double *data) {
  const hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT,
                              H5P_DEFAULT);
  if (ds < 0) return false;
  const herr_t e = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  return e >= 0;
}

bool write_scalar(hid_t file, const char *name, hid_t type, const void *v) {
  const hid_t sp = H5Screate(H5S_SCALAR);
  const hid_t ds = H5Dcreate2(file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  bool ok = ds >= 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  return ok;
}

// 0 = read, MVD_IO_MISMATCH = shape differs from [nz][ny][nx], MVD_IO_ERROR otherwise
int read_field(hid_t file, const char *name, const mvd_params *p, double *data) {
  const hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
  if (ds < 0) return MVD_IO_ERROR;
  const hid_t sp = H5Dget_space(ds);
  hsize_t d[3] = {0, 0, 0};
  int rc = 0;
  if (H5Sget_simple_extent_ndims(sp) != 3) rc = MVD_IO_MISMATCH;
  else if (H5Sget_simple_extent_dims(sp, d, nullptr) < 0) rc = MVD_IO_ERROR;
  else if (d[0] != (hsize_t)p->nz || d[1] != (hsize_t)p->ny || d[2] != (hsize_t)p->nx)
    rc = MVD_IO_MISMATCH;
  if (rc == 0 && H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    rc = MVD_IO_ERROR;
  H5Sclose(sp);
  H5Dclose(ds);
  return rc;
}

}  // namespace

extern "C" int mvd_write_parameters(const char *dir, const mvd_params *p,
                                    const mvd_run_options *o, double tcur, double hcur,
                                    int iout) {
  if (!grid_ok(p) || !o) return MVD_BAD_ARGUMENT;
  FILE *f = std::fopen(in_dir(dir, "restart_parameters.txt").c_str(), "w");
  if (!f) return MVD_IO_ERROR;
  // the keys and their order are those of the reference's file (io.cpp:656-708), which its
  // input parser reads back (io.cpp:231-297)
  std::fprintf(f, "# Euler3D restart file\n");
  std::fprintf(f, "xl = %.16e\n", p->xl);
  std::fprintf(f, "xr = %.16e\n", p->xr);
  std::fprintf(f, "yl = %.16e\n", p->yl);
  std::fprintf(f, "yr = %.16e\n", p->yr);
  std::fprintf(f, "zl = %.16e\n", p->zl);
  std::fprintf(f, "zr = %.16e\n", p->zr);
  std::fprintf(f, "t0 = %.16e\n", tcur);
  std::fprintf(f, "tf = %.16e\n", o->tf);
  std::fprintf(f, "gamma = %.16e\n", p->gamma);
  std::fprintf(f, "MassUnits = %.16e\n", p->mass_units);
  std::fprintf(f, "LengthUnits = %.16e\n", p->length_units);
  std::fprintf(f, "TimeUnits = %.16e\n", p->time_units);
  std::fprintf(f, "nx = %li\n", p->nx);
  std::fprintf(f, "ny = %li\n", p->ny);
  std::fprintf(f, "nz = %li\n", p->nz);
  static const char *const bcname[6] = {"xlbc", "xrbc", "ylbc", "yrbc", "zlbc", "zrbc"};
  for (int b = 0; b < 6; b++) std::fprintf(f, "%s = %i\n", bcname[b], p->bc[b]);
  std::fprintf(f, "cfl = %.16e\n", p->cfl);
  std::fprintf(f, "nout = %i\n", o->nout - iout);
  std::fprintf(f, "showstats = %i\n", o->showstats);
  std::fprintf(f, "order = %i\n", o->order);
  std::fprintf(f, "dense_order = %i\n", o->dense_order);
  std::fprintf(f, "etable = %i\n", o->etable);
  std::fprintf(f, "itable = %i\n", o->itable);
  std::fprintf(f, "mtable = %i\n", o->mtable);
  std::fprintf(f, "adapt_method = %i\n", o->adapt_method);
  std::fprintf(f, "maxnef = %i\n", o->maxnef);
  std::fprintf(f, "mxhnil = %i\n", o->mxhnil);
  std::fprintf(f, "mxsteps = %i\n", o->mxsteps);
  std::fprintf(f, "safety = %.16e\n", o->safety);
  std::fprintf(f, "bias = %.16e\n", o->bias);
  std::fprintf(f, "growth = %.16e\n", o->growth);
  std::fprintf(f, "pq = %i\n", o->pq);
  std::fprintf(f, "k1 = %.16e\n", o->k1);
  std::fprintf(f, "k2 = %.16e\n", o->k2);
  std::fprintf(f, "k3 = %.16e\n", o->k3);
  std::fprintf(f, "etamx1 = %.16e\n", o->etamx1);
  std::fprintf(f, "etamxf = %.16e\n", o->etamxf);
  std::fprintf(f, "h0 = %.16e\n", hcur);
  std::fprintf(f, "hmin = %.16e\n", o->hmin);
  std::fprintf(f, "hmax = %.16e\n", o->hmax);
  std::fprintf(f, "fixedstep = %i\n", o->fixedstep);
  std::fprintf(f, "htrans = %.16e\n", o->htrans);
  std::fprintf(f, "rtol = %.16e\n", o->rtol);
  std::fprintf(f, "atol = %.16e\n", o->atol);
  std::fprintf(f, "fusedkernels = %i\n", o->fusedkernels);
  std::fprintf(f, "localreduce = %i\n", o->localreduce);
  std::fprintf(f, "restart = %i\n", iout);
  return std::fclose(f) == 0 ? 0 : MVD_IO_ERROR;
}

extern "C" int mvd_output_solution(const char *dir, const mvd_params *p,
                                   const mvd_run_options *o, double tcur, double hcur,
                                   int iout, const double *const *w, int on_device) {
  if (!grid_ok(p) || !w || iout < 0) return MVD_BAD_ARGUMENT;
  for (int q = 0; q < 5; q++)
    if (!w[q]) return MVD_BAD_ARGUMENT;
  if (p->nchem > 0 && !w[5]) return MVD_BAD_ARGUMENT;
  if (o) {
    const int rc = mvd_write_parameters(dir, p, o, tcur, hcur, iout);
    if (rc) return rc;
  }
  const size_t N = (size_t)p->nx * p->ny * p->nz;
  char fname[32];
  std::snprintf(fname, sizeof fname, "output-%07i.hdf5", iout);
  // file layout settings of the reference (io.cpp:773-776)
  const hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
  H5Pset_sieve_buf_size(fapl, 262144);
  H5Pset_alignment(fapl, 524288, 262144);
  const hid_t file = H5Fcreate(in_dir(dir, fname).c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
  H5Pclose(fapl);
  if (file < 0) return MVD_IO_ERROR;
  int rc = 0;
  const int nchem = p->nchem;
  const double domain[6] = {p->zl, p->zr, p->yl, p->yr, p->xl, p->xr};
  if (!write_scalar(file, "time", H5T_NATIVE_DOUBLE, &tcur) ||
      !write_scalar(file, "nchem", H5T_NATIVE_INT, &nchem))
    rc = MVD_IO_ERROR;
  if (rc == 0) {
    const hsize_t dd[2] = {3, 2};
    const hid_t sp = H5Screate_simple(2, dd, nullptr);
    const hid_t ds =
        H5Dcreate2(file, "domain", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, domain) < 0)
      rc = MVD_IO_ERROR;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
  }
  const hsize_t dims[3] = {(hsize_t)p->nz, (hsize_t)p->ny, (hsize_t)p->nx};
  const hid_t space = H5Screate_simple(3, dims, nullptr);
  std::vector<double> buf(N);
  double u[5];
  unit_factors(p, u);
  for (int q = 0; q < 5 && rc == 0; q++) {
    if (!to_host(w[q], buf.data(), N, on_device)) { rc = MVD_DEVICE_ERROR; break; }
    for (size_t i = 0; i < N; i++) buf[i] = u[q] * buf[i];     // N_VScale(Units, w_q)
    if (!write_field(file, space, FLUID_NAMES[q], buf.data())) rc = MVD_IO_ERROR;
  }
  if (rc == 0 && nchem > 0) {
    std::vector<double> all(N * nchem);
    if (!to_host(w[5], all.data(), N * nchem, on_device)) rc = MVD_DEVICE_ERROR;
    for (int v = 0; v < nchem && rc == 0; v++) {
      for (size_t i = 0; i < N; i++) buf[i] = all[i * nchem + v];
      if (!write_field(file, space, chem_name(v).c_str(), buf.data())) rc = MVD_IO_ERROR;
    }
  }
  H5Sclose(space);
  if (H5Fclose(file) < 0 && rc == 0) rc = MVD_IO_ERROR;
  return rc;
}

extern "C" int mvd_read_restart(const char *dir, const mvd_params *p, int restart, double *t,
                                double *const *w, int on_device) {
  if (!grid_ok(p) || !w || !t || restart < 0) return MVD_BAD_ARGUMENT;
  for (int q = 0; q < 5; q++)
    if (!w[q]) return MVD_BAD_ARGUMENT;
  if (p->nchem > 0 && !w[5]) return MVD_BAD_ARGUMENT;
  char fname[32];
  std::snprintf(fname, sizeof fname, "output-%07i.hdf5", restart);
  const hid_t file = H5Fopen(in_dir(dir, fname).c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return MVD_IO_ERROR;
  int rc = 0;
  double tf = 0.0;
  // "time" (its type must be the working precision), then nchem, then the domain
  {
    const hid_t ds = H5Dopen2(file, "time", H5P_DEFAULT);
    if (ds < 0) {
      rc = MVD_IO_ERROR;
    } else {
      const hid_t ty = H5Dget_type(ds);
      if (H5Tequal(ty, H5T_NATIVE_DOUBLE) <= 0) rc = MVD_IO_MISMATCH;
      else if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &tf) < 0)
        rc = MVD_IO_ERROR;
      H5Tclose(ty);
      H5Dclose(ds);
    }
  }
  if (rc == 0) {
    int nchem = -1;
    const hid_t ds = H5Dopen2(file, "nchem", H5P_DEFAULT);
    if (ds < 0 || H5Dread(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &nchem) < 0)
      rc = MVD_IO_ERROR;
    else if (nchem != p->nchem)
      rc = MVD_IO_MISMATCH;
    if (ds >= 0) H5Dclose(ds);
  }
  if (rc == 0) {
    const hid_t ds = H5Dopen2(file, "domain", H5P_DEFAULT);
    if (ds < 0) {
      rc = MVD_IO_ERROR;
    } else {
      const hid_t sp = H5Dget_space(ds);
      hsize_t d[2] = {0, 0};
      double dom[6];
      if (H5Sget_simple_extent_ndims(sp) != 2 || H5Sget_simple_extent_dims(sp, d, nullptr) < 0 ||
          d[0] != 3 || d[1] != 2) {
        rc = MVD_IO_MISMATCH;
      } else if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, dom) < 0) {
        rc = MVD_IO_ERROR;
      } else {
        const double want[6] = {p->zl, p->zr, p->yl, p->yr, p->xl, p->xr};
        for (int a = 0; a < 6; a++)
          if (std::fabs(dom[a] - want[a]) > 1e-14) rc = MVD_IO_MISMATCH;
      }
      H5Sclose(sp);
      H5Dclose(ds);
    }
  }
  const size_t N = (size_t)p->nx * p->ny * p->nz;
  std::vector<double> buf(N);
  double u[5];
  unit_factors(p, u);
  for (int q = 0; q < 5 && rc == 0; q++) {
    rc = read_field(file, FLUID_NAMES[q], p, buf.data());
    if (rc) break;
    const double s = 1.0 / u[q];
    for (size_t i = 0; i < N; i++) buf[i] = s * buf[i];        // N_VScale(1/Units, w_q)
    if (!from_host(buf.data(), w[q], N, on_device)) rc = MVD_DEVICE_ERROR;
  }
  if (rc == 0 && p->nchem > 0) {
    const int nchem = p->nchem;
    std::vector<double> all(N * nchem);
    for (int v = 0; v < nchem && rc == 0; v++) {
      rc = read_field(file, chem_name(v).c_str(), p, buf.data());
      for (size_t i = 0; i < N && rc == 0; i++) all[i * nchem + v] = buf[i];
    }
    if (rc == 0 && !from_host(all.data(), w[5], N * nchem, on_device)) rc = MVD_DEVICE_ERROR;
  }
  if (H5Fclose(file) < 0 && rc == 0) rc = MVD_IO_ERROR;
  if (rc == 0) *t = tf;
  return rc;
}
