// Command-line driver over the C ABI, the counterpart of the reference's
// main() loop (smaiso-wcomp.cpp:139-2498) for one GPU.
//
//   smaiso_run nu ep_arg [options]
//
// nu and ep_arg keep the reference's argv meaning (S:249: nu = argv[1];
// S:242-243: epsilon = -0.001*argv[2]).  Options:
//   --n NX NY NZ      grid (default 64 64 64)
//   --dt DT           time step (default 0.001, S:262)
//   --rho0 R          isotropic density rho_0 (default 0.01, S:251)
//   --rho-m R         stabilisation density (default rho_s/2, S:255)
//   --divv-switch K   relaxation steps (default 50, S:214)
//   --steps K         run K steps (default: until L1 < the stop threshold, S:1081)
//   --lim-l1 X        stop threshold of L1 when --steps is absent (default
//                     1e-6, the reference's limL1, S:1029)
//   --ic reference|interface   IC: the reference's layers (S:571-590) or the
//                     tanh interface of smaiso_make_interface_ic (default)
//   --load DIR [R]    the reference's load mode (argv[4] = 1, S:727-748):
//                     psi from DIR/psi0.dat .. psi<R-1>.dat (R = 1 by
//                     default), rho = 0 (S:1053), relaxation skipped (S:1312)
//   --restart FILE    start from a full-state checkpoint (smaiso_save_checkpoint)
//   --checkpoint FILE write a full-state checkpoint at every save and at the end
//   --mixed           FP32-mixed precision
//   --out DIR         the reference's output files in DIR (S:800-919): every
//                     50 steps L1.dat, mass.dat, energy.dat (S:2246-2262);
//                     every --save-every steps (default 500 = stepL1*stepSave,
//                     S:206-207) psi0.dat, psiMid.dat, surfPsi.dat, vx/vy/vz,
//                     vsolx/vsoly/vsolz, divv, rho, amp, energyMid (S:2273-2446)
//   --save-every K    save cadence in steps (a multiple of 50)
// Values are written with the reference's ofstream formatting (6 significant
// digits), one per line, in its array orders.
#include "smaiso.h"
#include "run_args.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

static void usage() {
  std::fprintf(stderr,
               "usage: smaiso_run nu ep_arg [--n NX NY NZ] [--dt DT] [--rho0 R] [--rho-m R]\n"
               "                  [--divv-switch K] [--steps K] [--lim-l1 X] [--ic reference|interface]\n"
               "                  [--load DIR [R]] [--restart FILE] [--checkpoint FILE]\n"
               "                  [--mixed] [--out DIR] [--save-every K]\n");
}

static bool check(int rc, const char* what, smaiso_handle* h) {
  if (rc == SMAISO_OK) return true;
  std::fprintf(stderr, "%s: %s\n", what, smaiso_last_error(h));
  return false;
}

// truncate (create) an output file, as S:805-917 does for each
static void create(const std::string& path) { std::ofstream f(path); }

static void append(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios_base::app);
  for (double x : v) f << x << "\n";
}

int main(int argc, char** argv) {
  RunArgs args;
  std::string err;
  if (parse_run_args(argc, argv, &args, &err) != 0) {
    std::fprintf(stderr, "smaiso_run: %s\n", err.c_str());
    usage();
    return 2;
  }
  smaiso_params& p = args.p;
  const long long steps = args.steps;
  const double lim_l1 = args.lim_l1;
  const int save_every = args.save_every, load_ranks = args.load_ranks;
  const std::string &ic = args.ic, &out = args.out, &load_dir = args.load_dir, &restart = args.restart,
                    &ckpt = args.ckpt;
  const int step_l1 = 50;  // S:206
  smaiso_handle* h = nullptr;
  if (!check(smaiso_create(&p, &h), "smaiso_create", nullptr)) return 1;

  const size_t n3 = (size_t)p.nx * p.ny * p.nz;
  const double amp = std::sqrt((3 * p.beta + std::sqrt(9 * p.beta * p.beta + 40 * p.epsilon * p.gamma)) / 5);
  const double kp = (1 - p.rho0) / amp;  // S:252
  const double rho_s = kp * amp + p.rho0;  // S:253
  std::vector<double> psi(n3, 0.0), rho(n3, 0.0);
  if (!restart.empty()) {
    if (!check(smaiso_load_checkpoint(h, restart.c_str()), "smaiso_load_checkpoint", h)) return 1;
  } else {
    if (!load_dir.empty()) {
      std::vector<std::string> names;
      for (int r = 0; r < load_ranks; ++r) names.push_back(load_dir + "/psi" + std::to_string(r) + ".dat");
      std::vector<const char*> cp;
      for (auto& n : names) cp.push_back(n.c_str());
      if (!check(smaiso_read_psi_text(cp.data(), (int32_t)cp.size(), (int64_t)n3, psi.data()),
                 "smaiso_read_psi_text", nullptr))
        return 1;
    } else if (ic == "reference") {
      // S:571-590: Amp cos(q0 k dz) for Nz/2-8 < k < Nz/2+8, rho = 0 (S:1053)
      for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j)
          for (int k = 0; k < p.nz; ++k)
            if (k > p.nz / 2 - 8 && k < p.nz / 2 + 8)
              psi[((size_t)i * p.ny + j) * p.nz + k] = amp * std::cos(p.q0 * k * p.dz);
    } else {
      smaiso_ic c{1.0, -1.0, 2.0, p.rho0, 1.0, 0.0, 1e-3, 1};
      smaiso_make_interface_ic(&p, &c, nullptr, psi.data(), rho.data());
    }
    if (!check(smaiso_set_state(h, psi.data(), rho.data()), "smaiso_set_state", h)) return 1;
  }

  const char* mids[] = {"psiMid.dat", "energyMid.dat", "vx.dat", "vy.dat", "vz.dat", "vsolx.dat",
                        "vsoly.dat", "vsolz.dat", "sx.dat", "sy.dat", "sz.dat", "divv.dat",
                        "rho.dat", "amp.dat", "surfPsi.dat", "L1.dat", "mass.dat", "energy.dat"};
  if (!out.empty()) {
    std::ofstream info(out + "/info");  // S:805-830 (the keys that apply here)
    info << "Nx: " << p.nx << "\n" << "Ny: " << p.ny << "\n" << "Nz: " << p.nz << "\n"
         << "q0 (z direction): " << p.q0 << "\n" << "epsilon: " << p.epsilon << "\n"
         << "beta: " << p.beta << "\n" << "gamma: " << p.gamma << "\n"
         << "Viscosity (nu): " << p.nu << "\n" << "2nd Viscosity (lambda): " << p.lambda << "\n"
         << "Density penalty (zeta): " << p.zeta << "\n" << "dx: " << p.dx << "\n"
         << "dt: " << p.dt << "\n" << "stepL1: " << step_l1 << "\n"
         << "stepSave: " << save_every / step_l1 << "\n" << "divv switch: " << p.divv_switch << "\n"
         << "Initial amp: " << amp << "\n" << "kp: " << kp << "\n" << "rho_0: " << p.rho0 << "\n";
    for (const char* m : mids) create(out + "/" + m);
  }
  std::vector<double> surf((size_t)p.nx * p.ny, 0.0), slice((size_t)p.nx * p.nz), full;

  auto save = [&]() -> bool {
    if (!out.empty()) {
      // surface tracker (S:2278-2318), then psi<rank>.dat (S:2334-2345)
      if (!check(smaiso_surface_height(h, rho_s - 0.15, surf.data()), "smaiso_surface_height", h)) return false;
      full.resize(n3);
      if (!check(smaiso_get_field(h, SMAISO_FIELD_PSI, full.data()), "smaiso_get_field", h)) return false;
      {
        std::ofstream f(out + "/psi0.dat");
        for (double x : full) f << x << "\n";
      }
      const struct { int field; const char* name; } mid[] = {
          {SMAISO_FIELD_PSI, "psiMid.dat"}, {-1, "surfPsi.dat"}, {SMAISO_FIELD_VX, "vx.dat"},
          {SMAISO_FIELD_VY, "vy.dat"}, {SMAISO_FIELD_VZ, "vz.dat"}, {SMAISO_FIELD_VSOLX, "vsolx.dat"},
          {SMAISO_FIELD_VSOLY, "vsoly.dat"}, {SMAISO_FIELD_VSOLZ, "vsolz.dat"},
          {SMAISO_FIELD_DIVV, "divv.dat"}, {SMAISO_FIELD_RHO, "rho.dat"}, {SMAISO_FIELD_AMP, "amp.dat"}};
      for (const auto& m : mid) {
        if (m.field < 0) { append(out + "/" + m.name, surf); continue; }
        if (!check(smaiso_get_slice(h, m.field, p.ny / 2, slice.data()), "smaiso_get_slice", h)) return false;
        append(out + "/" + m.name, slice);
      }
      // energyMid.dat: energy_local of S:2227-2230 (the free-energy density
      // without the density weighting), from the psi and L psi slices
      std::vector<double> lps(slice.size());
      if (!check(smaiso_get_slice(h, SMAISO_FIELD_PSI, p.ny / 2, slice.data()), "smaiso_get_slice", h) ||
          !check(smaiso_get_slice(h, SMAISO_FIELD_LPSI, p.ny / 2, lps.data()), "smaiso_get_slice", h))
        return false;
      for (size_t q = 0; q < slice.size(); ++q) {
        const double ps = slice[q];
        slice[q] = -0.5 * p.epsilon * std::pow(ps, 2) + 0.5 * std::pow(lps[q], 2) -
                   0.25 * p.beta * std::pow(ps, 4) + (p.gamma / 6) * std::pow(ps, 6);
      }
      append(out + "/energyMid.dat", slice);
    }
    if (!ckpt.empty() && !check(smaiso_save_checkpoint(h, ckpt.c_str()), "smaiso_save_checkpoint", h))
      return false;
    return true;
  };

  double l1 = 1.0, mass = 0, energy = 0;
  long long done = 0;
  auto t0 = std::chrono::steady_clock::now();
  while (steps < 0 ? (l1 > lim_l1) : (done < steps)) {
    int chunk = step_l1;
    if (steps >= 0 && steps - done < chunk) chunk = (int)(steps - done);
    if (!check(smaiso_step(h, chunk), "smaiso_step", h)) return 1;
    done += chunk;
    int rc = smaiso_diagnostics(h, &l1, &mass, &energy);
    std::printf("step %lld L1 %.6g mass %.10g energy %.10g\n", (long long)smaiso_step_count(h), l1, mass, energy);
    std::fflush(stdout);
    if (!out.empty()) {
      append(out + "/L1.dat", {l1});
      append(out + "/mass.dat", {mass});
      append(out + "/energy.dat", {energy});
    }
    if (rc == SMAISO_E_NONFINITE) { std::fprintf(stderr, "state blew up\n"); return 3; }
    if (chunk == step_l1 && done % save_every == 0 && !save()) return 1;
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("%lld steps in %.3f s (%.2f steps/s)\n", done, sec, done / sec);
  if (!ckpt.empty() && !check(smaiso_save_checkpoint(h, ckpt.c_str()), "smaiso_save_checkpoint", h)) return 1;
  smaiso_destroy(h);
  return 0;
}
