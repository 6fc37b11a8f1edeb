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
//   --steps K         run K steps (default: until L1 < 1e-6, S:1081)
//   --ic reference|interface   IC: the reference's layers (S:571-590) or the
//                     tanh interface of smaiso_make_interface_ic (default)
//   --mixed           FP32-mixed precision
//   --out DIR         append L1.dat, mass.dat, energy.dat every 50 steps
//                     (S:2246-2262), one value per line
#include "smaiso.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static void usage() {
  std::fprintf(stderr,
               "usage: smaiso_run nu ep_arg [--n NX NY NZ] [--dt DT] [--rho0 R] [--rho-m R]\n"
               "                  [--divv-switch K] [--steps K] [--ic reference|interface]\n"
               "                  [--mixed] [--out DIR]\n");
}

int main(int argc, char** argv) {
  if (argc < 3) { usage(); return 2; }
  smaiso_params p;
  smaiso_params_default(&p);
  p.nx = p.ny = p.nz = 64;
  p.nu = std::atof(argv[1]);
  p.epsilon = -0.001 * std::atof(argv[2]);
  long long steps = -1;
  std::string ic = "interface", out;
  for (int a = 3; a < argc; ++a) {
    std::string s = argv[a];
    auto need = [&](int k) { if (a + k >= argc) { usage(); std::exit(2); } };
    if (s == "--n") { need(3); p.nx = std::atoi(argv[++a]); p.ny = std::atoi(argv[++a]); p.nz = std::atoi(argv[++a]); }
    else if (s == "--dt") { need(1); p.dt = std::atof(argv[++a]); }
    else if (s == "--rho0") { need(1); p.rho0 = std::atof(argv[++a]); }
    else if (s == "--rho-m") { need(1); p.rho_m = std::atof(argv[++a]); }
    else if (s == "--divv-switch") { need(1); p.divv_switch = std::atoi(argv[++a]); }
    else if (s == "--steps") { need(1); steps = std::atoll(argv[++a]); }
    else if (s == "--ic") { need(1); ic = argv[++a]; }
    else if (s == "--mixed") { p.precision = SMAISO_FP32_MIXED; }
    else if (s == "--out") { need(1); out = argv[++a]; }
    else { usage(); return 2; }
  }
  smaiso_handle* h = nullptr;
  if (smaiso_create(&p, &h) != SMAISO_OK) {
    std::fprintf(stderr, "smaiso_create: %s\n", smaiso_last_error(nullptr));
    return 1;
  }
  const size_t n3 = (size_t)p.nx * p.ny * p.nz;
  std::vector<double> psi(n3, 0.0), rho(n3, 0.0);
  if (ic == "reference") {
    // S:571-590: Amp cos(q0 k dz) for Nz/2-8 < k < Nz/2+8, rho = 0 (S:1053)
    const double amp = std::sqrt((3 * p.beta + std::sqrt(9 * p.beta * p.beta + 40 * p.epsilon * p.gamma)) / 5);
    for (int i = 0; i < p.nx; ++i)
      for (int j = 0; j < p.ny; ++j)
        for (int k = 0; k < p.nz; ++k)
          if (k > p.nz / 2 - 8 && k < p.nz / 2 + 8)
            psi[((size_t)i * p.ny + j) * p.nz + k] = amp * std::cos(p.q0 * k * p.dz);
  } else {
    smaiso_ic c{1.0, -1.0, 2.0, p.rho0, 1.0, 0.0, 1e-3, 1};
    smaiso_make_interface_ic(&p, &c, nullptr, psi.data(), rho.data());
  }
  if (smaiso_set_state(h, psi.data(), rho.data()) != SMAISO_OK) {
    std::fprintf(stderr, "smaiso_set_state: %s\n", smaiso_last_error(h));
    return 1;
  }
  const int step_l1 = 50;  // S:206
  const double lim_l1 = 1e-6;  // S:1029
  double l1 = 1.0, mass = 0, energy = 0;
  long long done = 0;
  auto t0 = std::chrono::steady_clock::now();
  while (steps < 0 ? (l1 > lim_l1) : (done < steps)) {
    int chunk = step_l1;
    if (steps >= 0 && steps - done < chunk) chunk = (int)(steps - done);
    if (smaiso_step(h, chunk) != SMAISO_OK) {
      std::fprintf(stderr, "smaiso_step: %s\n", smaiso_last_error(h));
      return 1;
    }
    done += chunk;
    int rc = smaiso_diagnostics(h, &l1, &mass, &energy);
    std::printf("step %lld L1 %.6g mass %.10g energy %.10g\n", done, l1, mass, energy);
    std::fflush(stdout);
    if (!out.empty()) {
      const char* names[3] = {"L1.dat", "mass.dat", "energy.dat"};
      const double vals[3] = {l1, mass, energy};
      for (int q = 0; q < 3; ++q) {
        FILE* f = std::fopen((out + "/" + names[q]).c_str(), "a");
        if (f) { std::fprintf(f, "%g\n", vals[q]); std::fclose(f); }
      }
    }
    if (rc == SMAISO_E_NONFINITE) { std::fprintf(stderr, "state blew up\n"); return 3; }
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("%lld steps in %.3f s (%.2f steps/s)\n", done, sec, done / sec);
  smaiso_destroy(h);
  return 0;
}
