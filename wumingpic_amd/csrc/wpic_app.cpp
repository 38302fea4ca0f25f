// wpic_app.cpp -- the reference's problem drivers on one MI355X: the 2D
// Weibel app (2d/proj/weibel/app.f90), the 2D reconnection app
// (2d/proj/reconnection/app.f90) and the 3D Weibel app
// (3d/proj/weibel/app.f90), with their JSON config, derived
// parameters, initial state or restart, output cadence (particles, tracers,
// moments + energy history, restart at the elapsed-time limit and at the
// end), driven through the C-ABIs of wpic.h and wpic_io.h.
//
// The config is the reference's own ("config" and "parameter" sections).
// The problem is the one whose parameters the config holds: "t_ani" for the
// Weibel app, "nbg"/"ncs" for the reconnection app, "udrift" for the
// counter-streaming electron beams of the headline benchmark (an extension:
// the reference's Weibel app loads an anisotropic Maxwellian instead),
// "u_inject" for the shock app (2d/proj/shock/app.f90), its box growing from
// n_x_ini cells by relocate every intvl_expand steps.
// Optional extension keys in "parameter": "seed" (the counter-based draws
// replacing the reference's OS-seeded ones), "e1" (the X-point perturbation
// the reconnection app hard-codes as 0.12).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wpic.h"
#include "../../include/wpic3d.h"
#include "../../include/wpic_app.h"
#include "../../include/wpic_io.h"
#include "wpic_json.h"

namespace wpic {
int set_error(int code, const std::string& msg);
}

namespace {

using wpic::set_error;
using namespace wpic::json;

const double PI = 3.141592653589793;  // 4*atan(1)

struct Config {
  JVal root;
  std::string path;
  int verbose = 0;
  std::string datadir;
  double max_elapsed = 0;
  int max_it = 0, intvl_ptcl = 0, intvl_mom = 0, intvl_orb = 0, intvl_expand = 0;
  std::string restart_file;
};

bool get_num(const JVal* o, const char* k, double& v) {
  const JVal* x = o ? o->get(k) : nullptr;
  if (!x) return false;
  if (x->k == JVal::INT) v = (double)x->i;
  else if (x->k == JVal::REAL) v = x->d;
  else return false;
  return true;
}
bool get_int(const JVal* o, const char* k, int& v) {
  double d;
  if (!get_num(o, k, d)) return false;
  v = (int)d;
  return true;
}

// Fortran '(f8.2)' and '(e12.5)' edit descriptors of energy_history's line
std::string fmt_f(double v, int w, int d) {
  char b[64];
  std::snprintf(b, sizeof b, "%*.*f", w, d, v);
  std::string s = b;
  if ((int)s.size() > w) return std::string(w, '*');
  return s;
}
std::string fmt_e(double v, int w, int d) {
  if (!std::isfinite(v)) {
    std::string s = std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity");
    return std::string(w > (int)s.size() ? w - (int)s.size() : 0, ' ') + s;
  }
  char b[64];
  std::snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(v));
  // d.ddd e+XX  ->  0.dddd E+XX (exponent + 1)
  std::string digits(1, b[0]);
  const char* p = b + 1;
  if (*p == '.') p++;
  while (*p && *p != 'e') digits += *p++;
  int e10 = *p == 'e' ? std::atoi(p + 1) : 0;
  const int fexp = v == 0.0 ? 0 : e10 + 1;
  char eb[16];
  if (std::abs(fexp) <= 99)
    std::snprintf(eb, sizeof eb, "E%c%02d", fexp < 0 ? '-' : '+', std::abs(fexp));
  else
    std::snprintf(eb, sizeof eb, "%c%03d", fexp < 0 ? '-' : '+', std::abs(fexp));
  std::string s = std::string(std::signbit(v) ? "-" : "") + "0." + digits + eb;
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - (int)s.size(), ' ') + s;
}

uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int fail(const char* what) {
  std::fprintf(stderr, "%s\n", what);
  return -1;
}

// '(i7.7, "_restart")'
std::string restart_name(int it) {
  char name[32];
  if (it < 0 || it > 9999999)
    std::snprintf(name, sizeof name, "*******_restart");
  else
    std::snprintf(name, sizeof name, "%07d_restart", it);
  return name;
}

// save_restart's config update (app.f90:550-580): "restart_file" set in the
// "config" section, printed back where the config was read
void rewrite_config(Config& cfg, const std::string& name) {
  JVal* c = nullptr;
  for (auto& kv : cfg.root.obj)
    if (kv.first == "config") c = &kv.second;
  if (c) {
    bool found = false;
    for (auto& kv : c->obj)
      if (kv.first == "restart_file") {
        kv.second = JVal::string(name);
        found = true;
      }
    if (!found) c->add("restart_file", JVal::string(name));
    std::string txt;
    print(cfg.root, 0, txt);
    txt += "\n";
    FILE* f = std::fopen(cfg.path.c_str(), "wb");
    if (f) {
      std::fwrite(txt.data(), 1, txt.size(), f);
      std::fclose(f);
    }
  }
  cfg.restart_file = name;
}

struct App {
  Config cfg;
  wpic_params2d P{};
  wpic_ctx* ctx = nullptr;
  wpic_paraio* io = nullptr;
  int np = 0;  // the reference's particle slots per row (its up array)
  int nx = 0, ny = 0, n0 = 0;
  double wpe = 0, wpi = 0, wge = 0, wgi = 0, vte = 0, vti = 0;
  std::vector<double> up, uf, mom;
  std::vector<int> np2, cumcnt;

  ~App() {
    if (io) wpic_paraio_finalize(io);
    if (ctx) wpic_destroy(ctx);
  }

  int chk(int rc) {
    if (rc) std::fprintf(stderr, "%s\n", wpic_last_error());
    return rc;
  }

  void alloc_host() {
    up.assign((size_t)P.ndim * np * ny * P.nsp, 0.0);
    uf.assign((size_t)6 * (nx + 4) * (ny + 4), 0.0);
    np2.assign((size_t)ny * P.nsp, 0);
    cumcnt.assign((size_t)(nx + 1) * ny * P.nsp, 0);
  }

  // energy_history (2d/proj/weibel/app.f90:479-545): time, kinetic, E, B, total
  int energy_history(int it) {
    double e[3];
    if (chk(wpic_energy2d(ctx, e))) return -1;
    const std::string fn = cfg.datadir + "energy.dat";
    FILE* f = std::fopen(fn.c_str(), it == 0 ? "w" : "a");
    if (!f) return fail(("cannot open " + fn).c_str());
    const double total = e[0] + e[1] + e[2];
    std::fprintf(f, "%s %s %s %s %s\n", fmt_f(it * P.delt, 8, 2).c_str(), fmt_e(e[0], 12, 5).c_str(),
                 fmt_e(e[1], 12, 5).c_str(), fmt_e(e[2], 12, 5).c_str(), fmt_e(total, 12, 5).c_str());
    std::fclose(f);
    return 0;
  }

  int pull_particles() {
    if (chk(wpic_download_up2d(ctx, up.data(), np2.data(), cumcnt.data(), np))) return -1;
    return chk(wpic_download_fields2d(ctx, uf.data()));
  }

  // save_restart (app.f90:550-580): the restart pair, and the config with
  // "restart_file" set, printed back where it was read
  int save_restart(int it) {
    const std::string name = restart_name(it);
    if (pull_particles()) return -1;
    int nxe = P.nxge;  // the shock's active edge (io__output(up, uf, np2, nxs, nxe, ...))
    if (chk(wpic_get_nxe2d(ctx, &nxe))) return -1;
    if (chk(wpic_paraio_output(io, up.data(), uf.data(), np2.data(), P.nxgs, nxe, it, name.c_str()))) return -1;
    rewrite_config(cfg, name);
    return 0;
  }

  int run(int mode) {
    const JVal* par = cfg.root.get("parameter");
    int num_process = 1;
    get_int(par, "num_process", num_process);
    if (!get_int(par, "n_x", nx) || !get_int(par, "n_y", ny)) return fail("Error: n_x and n_y are required");
    double mass_ratio;
    if (!get_num(par, "mass_ratio", mass_ratio)) return fail("Error: mass_ratio is required");
    double seedd = 20261202;
    get_num(par, "seed", seedd);
    const unsigned long long seed = (unsigned long long)seedd;
    const bool weibel = par && par->get("t_ani"), beams = par && par->get("udrift"),
               recon = par && par->get("nbg"), shock = par && par->get("u_inject");
    if ((int)weibel + (int)beams + (int)recon + (int)shock != 1)
      return fail("Error: the parameter section must describe one problem (t_ani, udrift, nbg/ncs or u_inject)");

    P.ndim = 6;
    P.nsp = 2;
    P.nxgs = 2;
    P.nxge = P.nxgs + nx - 1;
    P.nygs = 2;
    P.nyge = P.nygs + ny - 1;
    P.delx = 1.0;
    P.c = 1.0;
    P.gfac = 0.501;
    P.mode = mode;
    P.r[0] = mass_ratio;
    P.r[1] = 1.0;
    double b0 = 0, x0 = 0, y0 = 0, lcs = 0, rtemp = 0, t_ani = 0, udrift = 0, e1 = 0.12;
    double u0 = 0, v0 = 0, theta = 0, phi = 0, l_damp = 0, pflux = 0;
    int nbg = 0, ncs = 0, n_x_ini = 0;
    if (shock) {
      // 2d/proj/shock/app.f90:240-345; the box spans nxs..nxe = nxgs + n_x_ini
      // and grows by relocate every intvl_expand steps
      double sigma_e = 0, omega_pe = 0, u_inject = 0;
      get_int(par, "n_ppc", n0);
      if (!get_int(par, "n_x_ini", n_x_ini)) return fail("Error: n_x_ini is required");
      if (n_x_ini < 3 || n_x_ini > nx - 1) return fail("Error: n_x_ini must lie in [3, n_x - 1]");
      get_num(par, "u_inject", u_inject);
      get_num(par, "sigma_e", sigma_e);
      get_num(par, "omega_pe", omega_pe);
      get_num(par, "v_the", vte);
      get_num(par, "v_thi", vti);
      get_num(par, "theta_bn", theta);
      get_num(par, "phi_bn", phi);
      get_num(par, "l_damp_ini", l_damp);
      theta = theta * PI / 180;
      phi = phi * PI / 180;
      np = n0 * nx * 5;
      P.bc = WPIC_BC_SHOCK;
      P.delt = 1.0 * P.delx / P.c;  // cfl = 1
      get_num(par, "delt", P.delt);
      u0 = -std::fabs(u_inject);
      const double gam0 = std::sqrt(1 + (u0 * u0) / (P.c * P.c));
      v0 = u0 / gam0;
      wpe = omega_pe;
      wge = omega_pe * std::sqrt(sigma_e);
      wpi = wpe / std::sqrt(mass_ratio);
      wgi = wge / mass_ratio;
      P.q[0] = +std::sqrt(gam0 * P.r[0] / (4 * PI * n0 / (P.delx * P.delx))) * wpi;
      P.q[1] = -std::sqrt(gam0 * P.r[1] / (4 * PI * n0 / (P.delx * P.delx))) * wpe;
      b0 = P.r[0] * P.c / P.q[0] * wgi * gam0;
      if ((long long)n0 * (nx - 2) > np) return fail("Error: Too large number of particles");
      // inject: n0 |v0| dt dx (nyge - nygs + 1) particles per step (app.f90:735)
      pflux = n0 * std::fabs(v0) * P.delt * P.delx * (P.nyge - P.nygs + 1);
      P.capacity = (long long)n0 * (nx - 1) * ny + (long long)(cfg.max_it + 1) * ((long long)pflux + 1);
    } else if (weibel || beams) {
      // 2d/proj/weibel/app.f90:255-318
      double sigma_e = 0, omega_pe = 0, v_the = 0, v_thi = 0;
      get_int(par, "n_ppc", n0);
      get_num(par, "sigma_e", sigma_e);
      get_num(par, "omega_pe", omega_pe);
      get_num(par, "v_the", v_the);
      get_num(par, "v_thi", v_thi);
      get_num(par, "t_ani", t_ani);
      get_num(par, "udrift", udrift);
      np = n0 * nx * 5;
      P.bc = WPIC_BC_PERIODIC;
      P.delt = 1.0 * P.delx / P.c;  // cfl = 1
      get_num(par, "delt", P.delt);  // extension: the benchmark's dt = 0.5
      wpe = omega_pe;
      wge = omega_pe * std::sqrt(sigma_e);
      wpi = wpe / std::sqrt(mass_ratio);
      wgi = wge / mass_ratio;
      vte = v_the;
      vti = v_thi;
      P.q[0] = +std::sqrt(P.r[0] / (4 * PI * n0 / (P.delx * P.delx))) * wpi;
      P.q[1] = -std::sqrt(P.r[1] / (4 * PI * n0 / (P.delx * P.delx))) * wpe;
      b0 = P.r[0] * P.c / P.q[0] * wgi;
      if ((long long)n0 * nx > np) return fail("Error: Too large number of particles");
      P.capacity = (long long)n0 * nx * ny;
    } else {
      // 2d/proj/reconnection/app.f90:281-305
      double alpha = 0;
      get_num(par, "alpha", alpha);
      get_num(par, "rtemp", rtemp);
      get_num(par, "lcs", lcs);
      get_int(par, "nbg", nbg);
      get_int(par, "ncs", ncs);
      get_num(par, "e1", e1);
      n0 = nbg + ncs;
      np = n0 * nx;
      P.bc = WPIC_BC_WALLX;
      P.delt = 0.5 * P.delx / P.c;  // cfl = 0.5
      const double ldb = P.delx;
      vte = std::sqrt(rtemp) * P.c / (std::sqrt(1 + rtemp) * alpha);
      vti = vte * std::sqrt(P.r[1] / P.r[0]) / std::sqrt(rtemp);
      wpe = vte / ldb / std::sqrt(2.0);
      wpi = wpe * std::sqrt(P.r[1] / P.r[0]);
      wge = wpe / alpha;
      wgi = wge / mass_ratio;
      P.q[0] = +std::sqrt(P.r[0] / (4 * PI * n0 / (P.delx * P.delx))) * wpi;
      P.q[1] = -std::sqrt(P.r[1] / (4 * PI * n0 / (P.delx * P.delx))) * wpe;
      b0 = P.r[0] * P.c / P.q[0] * wgi;
      x0 = 0.5 * (P.nxge + P.nxgs) * P.delx;
      y0 = 0.5 * (P.nyge - P.nygs) * P.delx;
      lcs = lcs * P.c / wpi;
      // counts are conserved; room for the initial load
      const long long row = (long long)(nbg * (P.nxge - P.nxgs) + ncs * 2 * lcs);
      P.capacity = row * ny;
    }
    if (num_process != 1)
      std::fprintf(stderr, "note: num_process = %d in the config; this driver runs one rank on one GPU\n",
                   num_process);

    // restart: the counts come from the file
    int it0 = 0;
    alloc_host();
    if (chk(wpic_paraio_init2d(&io, P.ndim, np, P.nsp, P.nxgs, P.nxge, P.nygs, P.nyge, P.nygs, P.nyge, 1, 0,
                               P.delx, P.delt, P.c, P.q, P.r, cfg.datadir.c_str())))
      return -1;
    if (!cfg.restart_file.empty()) {
      int nxs, nxe;
      if (chk(wpic_paraio_input(io, up.data(), uf.data(), np2.data(), &nxs, &nxe, &it0, cfg.restart_file.c_str())))
        return -1;
      if (shock && (nxs != P.nxgs || nxe < P.nxgs + 3 || nxe > P.nxge))
        return fail("Error: the restart file's nxs, nxe lie outside the box");
      long long most = 0;
      for (int s = 0; s < P.nsp; s++) {
        long long t = 0;
        for (int r = 0; r < ny; r++) t += np2[(size_t)s * ny + r];
        most = std::max(most, t);
      }
      P.capacity = std::max(P.capacity, most);
      if (chk(wpic_create2d(&ctx, &P))) return -1;
      if (shock && chk(wpic_set_nxe2d(ctx, nxe))) return -1;
      if (chk(wpic_upload_gp2d(ctx, up.data(), np2.data(), np)) || chk(wpic_upload_fields2d(ctx, uf.data())))
        return -1;
      // io__input into gp, then sort__bucket(up, gp, ...) (app.f90:346-348)
      if (chk(wpic_sort_bucket2d(ctx))) return -1;
    } else {
      if (chk(wpic_create2d(&ctx, &P))) return -1;
      if (chk(shock ? wpic_paraio_param_shock(io, n0, wpe, wpi, wge, wgi, vti, vte, u0, "init_param")
                    : wpic_paraio_param(io, n0, wpe, wpi, wge, wgi, vti, vte, "init_param")))
        return -1;
      int rc;
      if (weibel)
        rc = wpic_ic_weibel2d(ctx, seed, n0, vti, vte, t_ani, b0);
      else if (beams)
        rc = wpic_ic_beams2d(ctx, seed, n0, udrift, vte, vti, b0);
      else if (shock)
        rc = wpic_set_nxe2d(ctx, P.nxgs + n_x_ini)
                 ? -1
                 : wpic_ic_shock2d(ctx, seed, n0, v0, vti, vte, l_damp, b0, theta, phi);
      else
        rc = wpic_ic_harris2d(ctx, seed, nbg, ncs, lcs, x0, y0, b0, e1, rtemp, vti, vte);
      if (chk(rc)) return -1;
      if (!shock && energy_history(0)) return -1;  // the shock app keeps no energy history
    }
    if (shock && chk(wpic_set_injection2d(ctx, u0))) return -1;
    std::vector<int> nl(ny);

    mom.assign((size_t)7 * (nx + 2) * (ny + 2) * P.nsp, 0.0);
    const double t0 = now_s();
    for (int it = it0 + 1; it <= cfg.max_it; it++) {
      if (chk(wpic_step2d(ctx, 1))) return -1;
      if (shock) {
        // inject (app.f90:731-760): int(pflux) particles plus one with the
        // fractional probability, split evenly over the rows with the
        // remainder given to randomly chosen rows
        uint64_t h = splitmix(seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)it));
        int nginj = (int)pflux;
        if ((double)(h >> 11) * 0x1.0p-53 < pflux - (int)pflux) nginj++;
        std::vector<int> idx(ny);
        for (int r = 0; r < ny; r++) {
          nl[r] = nginj / ny;
          idx[r] = r;
        }
        for (int r = ny - 1; r > 0; r--) {
          h = splitmix(h);
          std::swap(idx[r], idx[(int)(h % (uint64_t)(r + 1))]);
        }
        for (int r = 0; r < nginj % ny; r++) nl[idx[r]]++;
        if (chk(wpic_inject2d(ctx, seed, it, nl.data(), v0, vti, vte, l_damp, b0, theta, phi))) return -1;
        // expand box (app.f90:122-125)
        if (cfg.intvl_expand > 0 && it % cfg.intvl_expand == 0 &&
            chk(wpic_relocate2d(ctx, seed, it, n0, v0, vti, vte, l_damp, b0, theta, phi)))
          return -1;
      }
      if (cfg.intvl_ptcl > 0 && it % cfg.intvl_ptcl == 0) {
        if (pull_particles() || chk(wpic_paraio_ptcl(io, up.data(), uf.data(), np2.data(), it))) return -1;
      }
      if (cfg.intvl_orb > 0 && it % cfg.intvl_orb == 0) {
        if (pull_particles() || chk(wpic_paraio_orb(io, up.data(), uf.data(), np2.data(), it))) return -1;
      }
      if (cfg.intvl_mom > 0 && it % cfg.intvl_mom == 0) {
        if (chk(wpic_moments2d(ctx)) || chk(wpic_download_mom2d(ctx, mom.data())) ||
            chk(wpic_download_fields2d(ctx, uf.data())) || chk(wpic_paraio_mom(io, mom.data(), uf.data(), it)))
          return -1;
        if (!shock && energy_history(it)) return -1;
      }
      const double etime = now_s() - t0;
      if (etime >= cfg.max_elapsed) {
        if (save_restart(it)) return -1;
        std::fprintf(stderr, "*** Elapsed time limit exceeded \n");
        std::fprintf(stderr, "*** A snapshot %s has been saved\n", cfg.restart_file.c_str());
        return 0;
      }
      if (cfg.verbose >= 1)
        std::printf("*** Time step: %7d completed in %s sec.\n", it, fmt_e(etime, 10, 2).c_str());
    }
    // the Weibel app saves the final state, numbered max_it + 1
    // (weibel/app.f90:150-153); the reconnection app only finalizes
    // (reconnection/app.f90:146-148)
    return recon ? 0 : save_restart(cfg.max_it + 1);  // the shock app saves it too (app.f90:168-171)
  }
};

// The 3D Weibel app (3d/proj/weibel/app.f90:87-160 loop, :266-380 init):
// cfl 1, np = 3 n_ppc nx, energy.dat with one kinetic column per species
// ('(f10.2, 5(1x, e12.5))', :509-577), the final state saved as restart
// max_it + 1.  "udrift" selects the counter-streaming beams of the 3D Weibel
// workload instead of the lattice state.
struct App3 {
  Config cfg;
  wpic_params3d P{};
  wpic_ctx3d* ctx = nullptr;
  wpic_paraio* io = nullptr;
  int np = 0, nx = 0, ny = 0, nz = 0, n0 = 0;
  double wpe = 0, wpi = 0, wge = 0, wgi = 0, vte = 0, vti = 0;
  std::vector<double> up, uf, mom;
  std::vector<int> np2, cumcnt;

  ~App3() {
    if (io) wpic_paraio_finalize(io);
    if (ctx) wpic_destroy3d(ctx);
  }
  int chk(int rc) {
    if (rc) std::fprintf(stderr, "%s\n", wpic_last_error());
    return rc;
  }
  int energy_history(int it) {
    double e[4];
    if (chk(wpic_energy_species3d(ctx, e))) return -1;
    const std::string fn = cfg.datadir + "energy.dat";
    FILE* f = std::fopen(fn.c_str(), it == 0 ? "w" : "a");
    if (!f) return fail(("cannot open " + fn).c_str());
    const double total = ((e[0] + e[1]) + e[2]) + e[3];
    std::fprintf(f, "%s %s %s %s %s %s\n", fmt_f(it * P.delt, 10, 2).c_str(), fmt_e(e[0], 12, 5).c_str(),
                 fmt_e(e[1], 12, 5).c_str(), fmt_e(e[2], 12, 5).c_str(), fmt_e(e[3], 12, 5).c_str(),
                 fmt_e(total, 12, 5).c_str());
    std::fclose(f);
    return 0;
  }
  int pull_particles() {
    if (up.empty()) {
      up.assign((size_t)P.ndim * np * ny * nz * P.nsp, 0.0);
      np2.assign((size_t)ny * nz * P.nsp, 0);
      cumcnt.assign((size_t)(nx + 1) * ny * nz * P.nsp, 0);
    }
    if (chk(wpic_download_up3d(ctx, up.data(), np2.data(), cumcnt.data(), np))) return -1;
    return chk(wpic_download_fields3d(ctx, uf.data()));
  }
  int save_restart(int it) {
    const std::string name = restart_name(it);
    if (pull_particles()) return -1;
    if (chk(wpic_paraio_output(io, up.data(), uf.data(), np2.data(), P.nxgs, P.nxge, it, name.c_str()))) return -1;
    rewrite_config(cfg, name);
    return 0;
  }
  int run(int mode) {
    const JVal* par = cfg.root.get("parameter");
    int num_process = 1;
    get_int(par, "num_process", num_process);
    if (!get_int(par, "n_x", nx) || !get_int(par, "n_y", ny) || !get_int(par, "n_z", nz))
      return fail("Error: n_x, n_y and n_z are required");
    double mass_ratio;
    if (!get_num(par, "mass_ratio", mass_ratio)) return fail("Error: mass_ratio is required");
    double seedd = 20261202;
    get_num(par, "seed", seedd);
    const unsigned long long seed = (unsigned long long)seedd;
    const bool weibel = par && par->get("t_ani"), beams = par && par->get("udrift");
    if ((int)weibel + (int)beams != 1)
      return fail("Error: the parameter section must describe one problem (t_ani or udrift)");
    double sigma_e = 0, omega_pe = 0, t_ani = 0, udrift = 0;
    get_int(par, "n_ppc", n0);
    get_num(par, "sigma_e", sigma_e);
    get_num(par, "omega_pe", omega_pe);
    get_num(par, "v_the", vte);
    get_num(par, "v_thi", vti);
    get_num(par, "t_ani", t_ani);
    get_num(par, "udrift", udrift);
    P.ndim = 7;
    P.nsp = 2;
    P.nxgs = 2;
    P.nxge = P.nxgs + nx - 1;
    P.nygs = 2;
    P.nyge = P.nygs + ny - 1;
    P.nzgs = 2;
    P.nzge = P.nzgs + nz - 1;
    P.delx = 1.0;
    P.c = 1.0;
    P.gfac = 0.501;
    P.mode = mode;
    P.delt = 1.0 * P.delx / P.c;  // cfl = 1
    get_num(par, "delt", P.delt);  // extension: the benchmark's dt = 0.5
    np = n0 * nx * 3;
    const double pi = PI;
    wpe = omega_pe;
    wge = omega_pe * std::sqrt(sigma_e);
    wpi = wpe / std::sqrt(mass_ratio);
    wgi = wge / mass_ratio;
    P.r[0] = mass_ratio;
    P.r[1] = 1.0;
    P.q[0] = +std::sqrt(P.r[0] / (4.0 * pi * n0)) * wpi;
    P.q[1] = -std::sqrt(P.r[1] / (4.0 * pi * n0)) * wpe;
    const double b0 = P.r[0] * P.c / P.q[0] * wgi;
    if ((long long)n0 * nx > np || n0 <= 0) return fail("Error: Too large number of particles");
    P.capacity = (long long)n0 * nx * ny * nz;
    if (num_process != 1)
      std::fprintf(stderr, "note: num_process = %d in the config; this driver runs one rank on one GPU\n",
                   num_process);
    uf.assign((size_t)6 * (nx + 4) * (ny + 4) * (nz + 4), 0.0);
    int it0 = 0;
    if (chk(wpic_paraio_init3d(&io, P.ndim, np, P.nsp, P.nxgs, P.nxge, P.nygs, P.nyge, P.nzgs, P.nzge, P.nygs,
                               P.nyge, P.nzgs, P.nzge, 1, 0, P.delx, P.delt, P.c, P.q, P.r, cfg.datadir.c_str())))
      return -1;
    if (!cfg.restart_file.empty()) {
      up.assign((size_t)P.ndim * np * ny * nz * P.nsp, 0.0);
      np2.assign((size_t)ny * nz * P.nsp, 0);
      cumcnt.assign((size_t)(nx + 1) * ny * nz * P.nsp, 0);
      int nxs, nxe;
      if (chk(wpic_paraio_input(io, up.data(), uf.data(), np2.data(), &nxs, &nxe, &it0, cfg.restart_file.c_str())))
        return -1;
      if (chk(wpic_create3d(&ctx, &P))) return -1;
      if (chk(wpic_upload_gp3d(ctx, up.data(), np2.data(), np)) || chk(wpic_upload_fields3d(ctx, uf.data())))
        return -1;
      if (chk(wpic_sort_bucket3d(ctx))) return -1;
    } else {
      if (chk(wpic_create3d(&ctx, &P))) return -1;
      if (chk(wpic_paraio_param(io, n0, wpe, wpi, wge, wgi, vti, vte, "init_param"))) return -1;
      const int rc = weibel ? wpic_ic_weibel3d(ctx, seed, n0, vti, vte, t_ani, b0)
                            : wpic_ic_beams3d(ctx, seed, n0, udrift, vte, vti, b0);
      if (chk(rc)) return -1;
      if (energy_history(0)) return -1;
    }
    mom.assign((size_t)7 * (nx + 2) * (ny + 2) * (nz + 2) * P.nsp, 0.0);
    const double t0 = now_s();
    for (int it = it0 + 1; it <= cfg.max_it; it++) {
      if (chk(wpic_step3d(ctx, 1))) return -1;
      if (cfg.intvl_ptcl > 0 && it % cfg.intvl_ptcl == 0) {
        if (pull_particles() || chk(wpic_paraio_ptcl(io, up.data(), uf.data(), np2.data(), it))) return -1;
      }
      if (cfg.intvl_orb > 0 && it % cfg.intvl_orb == 0) {
        if (pull_particles() || chk(wpic_paraio_orb(io, up.data(), uf.data(), np2.data(), it))) return -1;
      }
      if (cfg.intvl_mom > 0 && it % cfg.intvl_mom == 0) {
        if (chk(wpic_moments3d(ctx)) || chk(wpic_download_mom3d(ctx, mom.data())) ||
            chk(wpic_download_fields3d(ctx, uf.data())) || chk(wpic_paraio_mom(io, mom.data(), uf.data(), it)))
          return -1;
        if (energy_history(it)) return -1;
      }
      const double etime = now_s() - t0;
      if (etime >= cfg.max_elapsed) {
        if (save_restart(it)) return -1;
        std::fprintf(stderr, "*** Elapsed time limit exceeded \n");
        std::fprintf(stderr, "*** A snapshot %s has been saved\n", cfg.restart_file.c_str());
        return 0;
      }
      if (cfg.verbose >= 1)
        std::printf("*** Time step: %7d completed in %s sec.\n", it, fmt_e(etime, 10, 2).c_str());
    }
    return save_restart(cfg.max_it + 1);
  }
};

// load_config (weibel/app.f90:160-260): the "config" section
int load_config(const char* config, Config& cfg) {
  const std::string path = config && *config ? config : "config.json";
  std::string txt;
  if (!read_file(path, txt)) {
    std::fprintf(stderr, "Error: %s does not exists\n", path.c_str());
    return set_error(WPIC_EINVAL, "Error: " + path + " does not exists");
  }
  Parser ps{txt.data(), txt.data() + txt.size()};
  cfg.root = ps.val();
  if (!ps.ok || cfg.root.k != JVal::OBJ) {
    std::fprintf(stderr, "Error: failed to load %s\n", path.c_str());
    return set_error(WPIC_EINVAL, "Error: failed to load " + path);
  }
  cfg.path = path;
  const JVal* c = cfg.root.get("config");
  get_int(c, "verbose", cfg.verbose);
  const JVal* dd = c ? c->get("datadir") : nullptr;
  cfg.datadir = (dd && dd->k == JVal::STR ? dd->s : std::string(".")) + "/";
  get_num(c, "max_elapsed", cfg.max_elapsed);
  get_int(c, "max_it", cfg.max_it);
  get_int(c, "intvl_ptcl", cfg.intvl_ptcl);
  get_int(c, "intvl_mom", cfg.intvl_mom);
  get_int(c, "intvl_orb", cfg.intvl_orb);
  get_int(c, "intvl_expand", cfg.intvl_expand);
  const JVal* rf = c ? c->get("restart_file") : nullptr;
  if (rf && rf->k == JVal::STR && !rf->s.empty()) cfg.restart_file = rf->s;
  return 0;
}

}  // namespace

extern "C" int wpic_app2d_main(const char* config, int mode) {
  App app;
  const int lc = load_config(config, app.cfg);
  if (lc) return lc;
  const int rc = app.run(mode);
  return rc ? (rc < 0 ? WPIC_EINVAL : rc) : 0;
}

extern "C" int wpic_app3d_main(const char* config, int mode) {
  App3 app;
  const int lc = load_config(config, app.cfg);
  if (lc) return lc;
  const int rc = app.run(mode);
  return rc ? (rc < 0 ? WPIC_EINVAL : rc) : 0;
}
