// ekpnp_io.cpp -- the reference's file outputs, restart input and run-time
// diagnostics, per channel, on top of the public C ABI (include/ekpnp.h).
//
// Formats follow LBM.cu byte for byte where they are text:
//   data_end.dat     save_data_end    LBM.cu:1272-1312 (9 columns, %10.6f)
//   restart input    read_data        LBM.cu:1350-1391
//   data.dat         save_data_tecplot LBM.cu:1100-1269 (Tecplot POINT zone)
//   charge_data      save_data_dmd    LBM.cu:1314-1348 (one c+ per line)
//   current          current          LBM.cu:1394-1417
//   umax.dat         record_umax      LBM.cu:1419-1451
// The file writers work on host copies of the channel's fields; the
// wall-row extrapolations they apply change the copies only, never the
// device state.  current and record_umax reduce on the device and copy two
// numbers back.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ekpnp.h"
#include "ekpnp_internal.h"

namespace {

struct Fields {
  int nx = 0, ny = 0;
  ekpnp_channel_params p{};
  std::vector<double> rho, ux, uy, cp, cn, psi, ex, ey;
  size_t at(int x, int y) const { return (size_t)y * nx + x; }
};

int fetch(ekpnp_ctx* ctx, int ch, Fields& f) {
  int rc = ekpnp_get_params(ctx, ch, &f.p);
  if (rc) return rc;
  f.nx = f.p.nx;
  f.ny = f.p.ny;
  const size_t nn = (size_t)f.nx * f.ny;
  std::vector<double>* dst[8] = {&f.rho, &f.ux, &f.uy, &f.cp, &f.cn, &f.psi, &f.ex, &f.ey};
  const int ids[8] = {EKPNP_FIELD_RHO, EKPNP_FIELD_UX, EKPNP_FIELD_UY, EKPNP_FIELD_CPOS,
                      EKPNP_FIELD_CNEG, EKPNP_FIELD_PSI, EKPNP_FIELD_EX, EKPNP_FIELD_EY};
  for (int k = 0; k < 8; ++k) {
    dst[k]->resize(nn);
    rc = ekpnp_get_field(ctx, ch, ids[k], dst[k]->data());
    if (rc) return rc;
  }
  return EKPNP_OK;
}

// v(x, 0) = 2 v(x, 1) - v(x, 2)
void extrap_bottom(std::vector<double>& v, const Fields& f) {
  for (int x = 0; x < f.nx; ++x) v[f.at(x, 0)] = 2.0 * v[f.at(x, 1)] - v[f.at(x, 2)];
}
// v(x, ny-1) = 2 v(x, ny-2) - v(x, ny-3)
void extrap_top(std::vector<double>& v, const Fields& f) {
  const int n = f.ny;
  for (int x = 0; x < f.nx; ++x) v[f.at(x, n - 1)] = 2.0 * v[f.at(x, n - 2)] - v[f.at(x, n - 3)];
}

}  // namespace

extern "C" {

int ekpnp_save_data_end(ekpnp_ctx* ctx, int channel, double time, const char* path) {
  if (!ctx || !path) return EKPNP_ERR_INVALID;
  Fields f;
  int rc = fetch(ctx, channel, f);
  if (rc) return rc;
  FILE* fp = std::fopen(path, "w");
  if (!fp) return EKPNP_ERR_IO;
  for (int y = 0; y < f.ny; ++y)
    for (int x = 0; x < f.nx; ++x) {
      const size_t i = f.at(x, y);
      std::fprintf(fp, "%10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f\n", time, f.ux[i], f.uy[i],
                   f.rho[i], f.cp[i], f.psi[i], f.ex[i], f.ey[i], f.cn[i]);
    }
  return std::fclose(fp) == 0 ? EKPNP_OK : EKPNP_ERR_IO;
}

int ekpnp_read_data(ekpnp_ctx* ctx, int channel, const char* path, double* time_out) {
  if (!ctx || !path) return EKPNP_ERR_INVALID;
  ekpnp_channel_params p;
  int rc = ekpnp_get_params(ctx, channel, &p);
  if (rc) return rc;
  const size_t nn = (size_t)p.nx * p.ny;
  std::vector<double> ux(nn), uy(nn), rho(nn), cp(nn), psi(nn), ex(nn), ey(nn), cn(nn);
  FILE* fp = std::fopen(path, "r");
  if (!fp) return EKPNP_ERR_IO;
  double t = 0.0;
  for (size_t i = 0; i < nn; ++i) {
    if (std::fscanf(fp, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &t, &ux[i], &uy[i], &rho[i], &cp[i], &psi[i], &ex[i],
                    &ey[i], &cn[i]) != 9) {
      std::fclose(fp);
      return EKPNP_ERR_IO;
    }
  }
  std::fclose(fp);
  // E is not stored: the step takes E = -grad(psi) of the file's psi
  rc = ekpnp_set_fields(ctx, channel, rho.data(), ux.data(), uy.data(), cp.data(), cn.data(), psi.data());
  if (rc) return rc;
  if (time_out) *time_out = t;
  return EKPNP_OK;
}

int ekpnp_save_tecplot(ekpnp_ctx* ctx, int channel, double time, const char* path, int first) {
  if (!ctx || !path) return EKPNP_ERR_INVALID;
  Fields f;
  int rc = fetch(ctx, channel, f);
  if (rc) return rc;
  const int NX = f.nx, NY = f.ny;
  const double dx = f.p.dx, dy = f.p.dy;
  // constants of LBM.cu:1119-1121
  const double Xs = 1000.0, Re = 170.07 / 100.0, extu = 0.25;
  const size_t nn = (size_t)NX * NY;
  std::vector<double> rhodx(nn, 0.0), rhody(nn, 0.0), vor(nn, 0.0), vord(nn, 0.0), curl(nn, 0.0);
  auto& c = f.cp;
  auto xm = [&](int x) { return x == 0 ? NX - 1 : x - 1; };
  auto xp = [&](int x) { return x == NX - 1 ? 0 : x + 1; };
  // gradients of c+ and vorticity on rows 1..NY-2, periodic in x (LBM.cu:1145-1180)
  for (int y = 1; y < NY - 1; ++y)
    for (int x = 0; x < NX; ++x) {
      rhodx[f.at(x, y)] = (c[f.at(xp(x), y)] - c[f.at(xm(x), y)]) / dx / 2;
      rhody[f.at(x, y)] = (c[f.at(x, y + 1)] - c[f.at(x, y - 1)]) / dy / 2;
      vor[f.at(x, y)] = (f.uy[f.at(xp(x), y)] - f.uy[f.at(xm(x), y)]) / dx / 2 -
                        (f.ux[f.at(x, y + 1)] - f.ux[f.at(x, y - 1)]) / dy / 2;
    }
  // Laplacian of the vorticity: rows 2..NY-5 (and 1..NY-3 at x = 0, NX-1);
  // the rows these leave out are filled by copies below (LBM.cu:1182-1196)
  auto lap = [&](int x, int y) {
    return (vor[f.at(xp(x), y)] + vor[f.at(x, y + 1)] + vor[f.at(xm(x), y)] + vor[f.at(x, y - 1)] -
            4 * vor[f.at(x, y)]) / (dy * dy);
  };
  for (int y = 2; y < NY - 4; ++y)
    for (int x = 1; x < NX - 1; ++x) vord[f.at(x, y)] = lap(x, y);
  for (int y = 2; y < NY - 2; ++y) {
    vord[f.at(0, y)] = lap(0, y);
    vord[f.at(NX - 1, y)] = lap(NX - 1, y);
  }
  // wall rows (LBM.cu:1202-1232)
  for (auto* v : {&f.rho, &f.cp, &f.cn, &f.ux, &f.uy, &rhodx, &rhody, &vor}) {
    extrap_bottom(*v, f);
    extrap_top(*v, f);
  }
  for (int x = 0; x < NX; ++x) {
    vord[f.at(x, 1)] = vord[f.at(x, 2)];
    vord[f.at(x, 0)] = vord[f.at(x, 1)];
    vord[f.at(x, NY - 4)] = vord[f.at(x, NY - 5)];
    vord[f.at(x, NY - 3)] = vord[f.at(x, NY - 4)];
    vord[f.at(x, NY - 2)] = vord[f.at(x, NY - 3)];
    vord[f.at(x, NY - 1)] = vord[f.at(x, NY - 2)];
  }
  for (size_t i = 0; i < nn; ++i) {
    curl[i] = (rhodx[i] * f.ey[i] - rhody[i] * f.ex[i]) / 1e5;
    vor[i] /= extu;
    vord[i] /= extu;
  }
  FILE* fp = std::fopen(path, first ? "w" : "a");
  if (!fp) return EKPNP_ERR_IO;
  if (first)
    std::fprintf(fp, "%s\n",
                 "VARIABLES=\"x\",\"y\",\"u\",\"v\",\"p\",\"Positive charge\",\"phi\",\"Ex\",\"Ey\",\"vorticity\","
                 "\"curlForce\",\"vorticity_diff\",\"XcurlForce\",\"Negative charge\"");
  std::fprintf(fp, "\n");
  std::fprintf(fp, "ZONE T=\"t=%g\", F=POINT, I = %d, J = %d\n", time, NX, NY);
  for (int y = 0; y < NY; ++y)
    for (int x = 0; x < NX; ++x) {
      const size_t i = f.at(x, y);
      std::fprintf(fp,
                   "%g %g %10.12f %10.12f %10.12f %10.12f %10.12f %10.12f %10.12f %10.12f %10.12f %10.12f %10.12f "
                   "%10.12f\n",
                   dx * x, dy * y, f.ux[i], f.uy[i], f.rho[i], f.cp[i], f.psi[i], f.ex[i], f.ey[i], vor[i], curl[i],
                   1 / Re * vord[i], Xs * curl[i], f.cn[i]);
    }
  return std::fclose(fp) == 0 ? EKPNP_OK : EKPNP_ERR_IO;
}

int ekpnp_save_charge_data(ekpnp_ctx* ctx, int channel, const char* path) {
  if (!ctx || !path) return EKPNP_ERR_INVALID;
  Fields f;
  int rc = fetch(ctx, channel, f);
  if (rc) return rc;
  extrap_top(f.cp, f);
  FILE* fp = std::fopen(path, "a");
  if (!fp) return EKPNP_ERR_IO;
  for (int y = 0; y < f.ny; ++y)
    for (int x = 0; x < f.nx; ++x) std::fprintf(fp, "%10.6f \n", f.cp[f.at(x, y)]);
  return std::fclose(fp) == 0 ? EKPNP_OK : EKPNP_ERR_IO;
}

int ekpnp_current(ekpnp_ctx* ctx, int channel, double* out) {
  if (!ctx || !out) return EKPNP_ERR_INVALID;
  ekpnp_channel_params p;
  int rc = ekpnp_get_params(ctx, channel, &p);
  if (rc) return rc;
  double sum = 0.0;
  rc = ekpnp::channel_diagnostics(ctx, channel, nullptr, &sum);
  if (rc) return rc;
  *out = sum * p.mob_pos * p.dy;
  return EKPNP_OK;
}

int ekpnp_record_umax(ekpnp_ctx* ctx, int channel, double time, const char* path, double* umax_out) {
  if (!ctx) return EKPNP_ERR_INVALID;
  // the reference's umax is the largest u_y (starting from 0), top row
  // extrapolated
  double umax = 0.0;
  int rc = ekpnp::channel_diagnostics(ctx, channel, &umax, nullptr);
  if (rc) return rc;
  if (umax_out) *umax_out = umax;
  if (path) {
    FILE* fp = std::fopen(path, "a");
    if (!fp) return EKPNP_ERR_IO;
    std::fprintf(fp, "%10.6f %10.6f\n", time, umax);
    if (std::fclose(fp) != 0) return EKPNP_ERR_IO;
  }
  return EKPNP_OK;
}

}  // extern "C"
