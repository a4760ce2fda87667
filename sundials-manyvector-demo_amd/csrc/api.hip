// C-ABI layer (include/mvd_rhs.h): context lifetime, argument checks, launches.
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <cmath>
#include "mvd_internal.hpp"
#include "mvd_chem.hpp"
#include "../../include/mvd_rhs.h"

struct mvd_context {
  mvd::Grid g;
  mvd_params p;
  hipStream_t stream = nullptr;
  mvd::FeScratch fe;
  unsigned *host_flag = nullptr;      // pinned
  double energy_units = 1.0;
  mvd::ChemState chem;
};

namespace {

bool dev_ok(hipError_t e) { return e == hipSuccess; }

int check_ptrs(const mvd_context *ctx, const double *const *w) {
  if (!w) return MVD_BAD_ARGUMENT;
  for (int v = 0; v < 5; v++)
    if (!w[v]) return MVD_BAD_ARGUMENT;
  if (ctx->g.nchem > 0 && !w[5]) return MVD_BAD_ARGUMENT;
  return MVD_SUCCESS;
}

mvd::Fields fields_of(const mvd_context *ctx, const double *const *w) {
  mvd::Fields f;
  for (int v = 0; v < 5; v++) f.w[v] = w[v];
  f.chem = ctx->g.nchem > 0 ? w[5] : nullptr;
  return f;
}

mvd::FieldsOut fields_out(const mvd_context *ctx, double *const *w) {
  mvd::FieldsOut f;
  for (int v = 0; v < 5; v++) f.w[v] = w[v];
  f.chem = ctx->g.nchem > 0 ? w[5] : nullptr;
  return f;
}

int fe_common(mvd_context *ctx, const double *const *w, double *const *wdot, bool plain,
              bool wait) {
  if (!ctx) return MVD_BAD_ARGUMENT;
  int rc = check_ptrs(ctx, w);
  if (rc) return rc;
  rc = check_ptrs(ctx, (const double *const *)wdot);
  if (rc) return rc;
  rc = mvd::fe_launch(ctx->g, fields_of(ctx, w), fields_out(ctx, wdot), ctx->fe, ctx->stream,
                      plain);
  if (rc == -2) return MVD_UNSUPPORTED;
  if (rc) return MVD_DEVICE_ERROR;
  if (!wait) return MVD_SUCCESS;
  if (!dev_ok(hipMemcpyAsync(ctx->host_flag, ctx->fe.flag, sizeof(unsigned),
                             hipMemcpyDeviceToHost, ctx->stream)))
    return MVD_DEVICE_ERROR;
  if (!dev_ok(hipStreamSynchronize(ctx->stream))) return MVD_DEVICE_ERROR;
  ctx->fe.timer.collect();
  return (*ctx->host_flag != 0u) ? MVD_ILLEGAL_STATE : MVD_SUCCESS;
}

}  // namespace

extern "C" {

int mvd_create(const mvd_params *p, void *stream, mvd_context **out) {
  if (!p || !out) return MVD_BAD_ARGUMENT;
  *out = nullptr;
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return MVD_BAD_ARGUMENT;
  if (p->nx > (1L << 20) || p->ny > (1L << 20) || p->nz > (1L << 20)) return MVD_BAD_ARGUMENT;
  if (p->nchem < 0 || !mvd::fe_nchem_supported(p->nchem)) return MVD_UNSUPPORTED;
  for (int d = 0; d < 3; d++) {
    const int lo = p->bc[2 * d], hi = p->bc[2 * d + 1];
    if (lo < 0 || lo > 3 || hi < 0 || hi > 3) return MVD_BAD_ARGUMENT;
    // SetupDecomp: a periodic face must be paired with a periodic face (euler3D.hpp:443-457)
    if ((lo == MVD_BC_PERIODIC) != (hi == MVD_BC_PERIODIC)) return MVD_BAD_ARGUMENT;
  }
  if (!(p->xr > p->xl) || !(p->yr > p->yl) || !(p->zr > p->zl)) return MVD_BAD_ARGUMENT;
  mvd_context *c = new mvd_context;
  c->p = *p;
  c->g.nx = (int)p->nx; c->g.ny = (int)p->ny; c->g.nz = (int)p->nz;
  c->g.nchem = p->nchem;
  for (int i = 0; i < 6; i++) c->g.bc[i] = p->bc[i];
  c->g.gamma = p->gamma;
  c->g.dx = (p->xr - p->xl) / p->nx;   // SetupDecomp, euler3D.hpp:412-414
  c->g.dy = (p->yr - p->yl) / p->ny;
  c->g.dz = (p->zr - p->zl) / p->nz;
  c->stream = (hipStream_t)stream;
  // derived energy unit (UpdateUnits, euler3D.hpp:391)
  const double mu = p->mass_units > 0 ? p->mass_units : 1.0;
  const double lu = p->length_units > 0 ? p->length_units : 1.0;
  const double tu = p->time_units > 0 ? p->time_units : 1.0;
  c->energy_units = mu / lu / tu / tu;
  const int nv = 5 + p->nchem;
  const long nx = p->nx, ny = p->ny, nz = p->nz;
  const size_t fx = (size_t)nv * (nx + 1) * ny * nz, fy = (size_t)nv * nx * (ny + 1) * nz,
               fz = (size_t)nv * nx * ny * (nz + 1);
  bool ok = dev_ok(hipMalloc(&c->fe.Fx, fx * sizeof(double))) &&
            dev_ok(hipMalloc(&c->fe.Fy, fy * sizeof(double))) &&
            dev_ok(hipMalloc(&c->fe.Fz, fz * sizeof(double))) &&
            dev_ok(hipMalloc(&c->fe.flag, sizeof(unsigned) * 4)) &&
            dev_ok(hipMalloc(&c->fe.red, sizeof(double) * 4096)) &&
            dev_ok(hipHostMalloc(&c->host_flag, sizeof(unsigned) * 4, hipHostMallocDefault));
  c->fe.red_len = 4096;
  if (!ok) {
    mvd_destroy(c);
    return MVD_DEVICE_ERROR;
  }
  if (!dev_ok(hipMemsetAsync(c->fe.flag, 0, sizeof(unsigned) * 4, c->stream)) ||
      !dev_ok(hipStreamSynchronize(c->stream))) {
    mvd_destroy(c);
    return MVD_DEVICE_ERROR;
  }
  *out = c;
  return MVD_SUCCESS;
}

void mvd_destroy(mvd_context *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  else (void)hipDeviceSynchronize();
  (void)hipFree(c->fe.Fx);
  (void)hipFree(c->fe.Fy);
  (void)hipFree(c->fe.Fz);
  (void)hipFree(c->fe.flag);
  (void)hipFree(c->fe.red);
  if (c->fe.timer.on)
    for (int i = 0; i < 2 * mvd::PROF_NSLOT; i++) (void)hipEventDestroy(c->fe.timer.ev[i]);
  if (c->host_flag) (void)hipHostFree(c->host_flag);
  mvd::chem_free(c->chem);
  delete c;
}

int mvd_set_stream(mvd_context *ctx, void *stream) {
  if (!ctx) return MVD_BAD_ARGUMENT;
  ctx->stream = (hipStream_t)stream;
  return MVD_SUCCESS;
}

int mvd_fe(mvd_context *ctx, double, const double *const *w, double *const *wdot) {
  return fe_common(ctx, w, wdot, false, true);
}

int mvd_fe_plain(mvd_context *ctx, double, const double *const *w, double *const *wdot) {
  return fe_common(ctx, w, wdot, true, true);
}

int mvd_fe_async(mvd_context *ctx, double, const double *const *w, double *const *wdot) {
  return fe_common(ctx, w, wdot, false, false);
}

int mvd_last_state_flags(mvd_context *ctx) {
  if (!ctx) return MVD_BAD_ARGUMENT;
  if (!dev_ok(hipMemcpyAsync(ctx->host_flag, ctx->fe.flag, sizeof(unsigned),
                             hipMemcpyDeviceToHost, ctx->stream)) ||
      !dev_ok(hipStreamSynchronize(ctx->stream)))
    return MVD_DEVICE_ERROR;
  ctx->fe.timer.collect();
  return (int)*ctx->host_flag;
}

int mvd_profile(mvd_context *ctx, int on) {
  if (!ctx) return MVD_BAD_ARGUMENT;
  mvd::KernelTimer &t = ctx->fe.timer;
  if (on && !t.on) {
    for (int i = 0; i < 2 * mvd::PROF_NSLOT; i++)
      if (!dev_ok(hipEventCreate(&t.ev[i]))) return MVD_DEVICE_ERROR;
  }
  if (!on && t.on) {
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 2 * mvd::PROF_NSLOT; i++) (void)hipEventDestroy(t.ev[i]);
  }
  t.on = on != 0;
  for (int s = 0; s < mvd::PROF_NSLOT; s++) { t.ms[s] = 0.0; t.count[s] = 0; t.pending[s] = false; }
  return MVD_SUCCESS;
}

int mvd_profile_read(mvd_context *ctx, double *ms, long *count, int nslot) {
  if (!ctx || !ms || !count || nslot <= 0) return MVD_BAD_ARGUMENT;
  mvd::KernelTimer &t = ctx->fe.timer;
  if (!dev_ok(hipStreamSynchronize(ctx->stream))) return MVD_DEVICE_ERROR;
  t.collect();
  for (int s = 0; s < nslot; s++) {
    ms[s] = s < mvd::PROF_NSLOT ? t.ms[s] : 0.0;
    count[s] = s < mvd::PROF_NSLOT ? t.count[s] : 0;
  }
  return mvd::PROF_NSLOT;
}

int mvd_max_wave_speed(mvd_context *ctx, const double *const *w, double *alpha) {
  if (!ctx || !alpha) return MVD_BAD_ARGUMENT;
  int rc = check_ptrs(ctx, w);
  if (rc) return rc;
  if (mvd::max_wave_speed_launch(ctx->g, fields_of(ctx, w), ctx->fe, ctx->stream, alpha))
    return MVD_DEVICE_ERROR;
  return MVD_SUCCESS;
}

int mvd_stability(mvd_context *ctx, const double *const *w, double, double *dt_stab) {
  if (!dt_stab) return MVD_BAD_ARGUMENT;
  double alpha = 0.0;
  int rc = mvd_max_wave_speed(ctx, w, &alpha);
  if (rc) return rc;
  *dt_stab = ctx->p.cfl * std::fmin(std::fmin(ctx->g.dx, ctx->g.dy), ctx->g.dz) / alpha;
  return MVD_SUCCESS;
}

int mvd_fslow(mvd_context *ctx, double t, double *const *w, double *const *wdot) {
  if (!ctx) return MVD_BAD_ARGUMENT;
  if (ctx->g.nchem <= 0) return MVD_UNSUPPORTED;
  if (!ctx->chem.ready) return MVD_NOT_READY;
  int rc = check_ptrs(ctx, (const double *const *)w);
  if (rc) return rc;
  rc = check_ptrs(ctx, (const double *const *)wdot);
  if (rc) return rc;
  if (ctx->chem.n != (long)ctx->g.nx * ctx->g.ny * ctx->g.nz || ctx->g.nchem != mvd::NSP)
    return MVD_BAD_ARGUMENT;
  if (mvd::fslow_prepare_launch(ctx->g, w, ctx->chem.scale, 1.0 / ctx->energy_units, ctx->stream))
    return MVD_DEVICE_ERROR;
  rc = fe_common(ctx, (const double *const *)w, wdot, false, false);
  if (rc) return rc;
  if (mvd::fslow_finish_launch(ctx->g, wdot, ctx->stream)) return MVD_DEVICE_ERROR;
  const long nch = ctx->chem.n * mvd::NSP;
  if (mvd::chem_scale_launch(nch, w[5], ctx->chem.scale, 0, ctx->stream)) return MVD_DEVICE_ERROR;
  if (!dev_ok(hipMemcpyAsync(ctx->host_flag, ctx->fe.flag, sizeof(unsigned),
                             hipMemcpyDeviceToHost, ctx->stream)) ||
      !dev_ok(hipStreamSynchronize(ctx->stream)))
    return MVD_DEVICE_ERROR;
  return (*ctx->host_flag != 0u) ? MVD_ILLEGAL_STATE : MVD_SUCCESS;
}

// ---- chemistry ----

int mvd_chem_setup(mvd_context *ctx, const double *tables_host, long ncell, double current_z) {
  if (!ctx || !tables_host || ncell <= 0) return MVD_BAD_ARGUMENT;
  return mvd::chem_setup(ctx->chem, tables_host, ncell, current_z, ctx->stream);
}

int mvd_chem_prepare(mvd_context *ctx, double *chem) {
  if (!ctx || !chem) return MVD_BAD_ARGUMENT;
  if (!ctx->chem.ready) return MVD_NOT_READY;
  return mvd::chem_prepare(ctx->chem, chem, ctx->stream);
}

int mvd_chem_apply_scaling(mvd_context *ctx, double *chem) {
  if (!ctx || !chem) return MVD_BAD_ARGUMENT;
  if (!ctx->chem.ready) return MVD_NOT_READY;
  if (mvd::chem_scale_launch(ctx->chem.n * mvd::NSP, chem, ctx->chem.scale, 1, ctx->stream))
    return MVD_DEVICE_ERROR;
  return dev_ok(hipStreamSynchronize(ctx->stream)) ? MVD_SUCCESS : MVD_DEVICE_ERROR;
}

int mvd_chem_unapply_scaling(mvd_context *ctx, double *chem) {
  if (!ctx || !chem) return MVD_BAD_ARGUMENT;
  if (!ctx->chem.ready) return MVD_NOT_READY;
  if (mvd::chem_scale_launch(ctx->chem.n * mvd::NSP, chem, ctx->chem.scale, 0, ctx->stream))
    return MVD_DEVICE_ERROR;
  return dev_ok(hipStreamSynchronize(ctx->stream)) ? MVD_SUCCESS : MVD_DEVICE_ERROR;
}

int mvd_chem_rhs(mvd_context *ctx, double, const double *y, double *ydot) {
  if (!ctx || !y || !ydot) return MVD_BAD_ARGUMENT;
  if (!ctx->chem.ready) return MVD_NOT_READY;
  return mvd::chem_rhs(ctx->chem, y, ydot, ctx->stream, true);
}

int mvd_chem_jac(mvd_context *ctx, double, const double *y, double *J) {
  if (!ctx || !y || !J) return MVD_BAD_ARGUMENT;
  if (!ctx->chem.ready) return MVD_NOT_READY;
  return mvd::chem_jac(ctx->chem, y, J, ctx->stream, true);
}

double *mvd_chem_state(mvd_context *ctx, int which) {
  if (!ctx || !ctx->chem.ready) return nullptr;
  return mvd::chem_state_ptr(ctx->chem, which);
}

}  // extern "C"
