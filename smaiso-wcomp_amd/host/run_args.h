// Command-line arguments of smaiso_run (see smaiso_run.cpp for their
// meaning); header-only so that the host sanitizer checker parses argv with
// the same code as the driver.
#pragma once

#include "smaiso.h"

#include <cmath>
#include <cstdlib>
#include <string>

struct RunArgs {
  smaiso_params p;
  long long steps = -1;  // -1: until L1 falls below lim_l1
  double lim_l1 = 1e-6;  // S:1029
  int save_every = 500;  // stepL1 * stepSave, S:206-207
  int load_ranks = 1;
  std::string ic = "interface", out, load_dir, restart, ckpt;
};

// 0 on success; 2 with the reason in *err for a usage error.  Numbers must
// parse whole (no trailing characters) and lie in range.
inline int parse_run_args(int argc, char** argv, RunArgs* a, std::string* err) {
  auto num = [&](const char* s, double* v) {
    char* end = nullptr;
    *v = std::strtod(s, &end);
    return end != s && *end == '\0' && std::isfinite(*v);
  };
  auto inum = [&](const char* s, long long lo, long long hi, long long* v) {
    char* end = nullptr;
    *v = std::strtoll(s, &end, 10);
    return end != s && *end == '\0' && *v >= lo && *v <= hi;
  };
  if (argc < 3) { *err = "need nu and ep_arg"; return 2; }
  smaiso_params_default(&a->p);
  a->p.nx = a->p.ny = a->p.nz = 64;
  double nu = 0, ep = 0;
  if (!num(argv[1], &nu) || !num(argv[2], &ep)) { *err = "nu and ep_arg must be numbers"; return 2; }
  a->p.nu = nu;                  // S:249
  a->p.epsilon = -0.001 * ep;    // S:242-243
  for (int i = 3; i < argc; ++i) {
    const std::string s = argv[i];
    auto need = [&](int k) {
      if (i + k >= argc) { *err = s + " needs " + std::to_string(k) + " value(s)"; return false; }
      return true;
    };
    double d = 0;
    long long v = 0;
    if (s == "--n") {
      if (!need(3)) return 2;
      long long n[3];
      for (int q = 0; q < 3; ++q)
        if (!inum(argv[++i], 1, 1 << 20, &n[q])) { *err = "--n takes three positive extents"; return 2; }
      a->p.nx = (int)n[0]; a->p.ny = (int)n[1]; a->p.nz = (int)n[2];
    } else if (s == "--dt" || s == "--rho0" || s == "--rho-m" || s == "--lim-l1") {
      if (!need(1)) return 2;
      if (!num(argv[++i], &d)) { *err = s + " takes a number"; return 2; }
      if (s == "--dt") a->p.dt = d;
      else if (s == "--rho0") a->p.rho0 = d;
      else if (s == "--rho-m") a->p.rho_m = d;
      else if (d > 0) a->lim_l1 = d;
      else { *err = "--lim-l1 must be a positive number"; return 2; }
    } else if (s == "--divv-switch") {
      if (!need(1) || !inum(argv[++i], 0, 1 << 30, &v)) { *err = "--divv-switch takes a count >= 0"; return 2; }
      a->p.divv_switch = (int)v;
    } else if (s == "--steps") {
      if (!need(1) || !inum(argv[++i], 0, 1LL << 40, &v)) { *err = "--steps takes a count >= 0"; return 2; }
      a->steps = v;
    } else if (s == "--save-every") {
      if (!need(1) || !inum(argv[++i], 1, 1 << 30, &v)) { *err = "--save-every takes a count >= 1"; return 2; }
      a->save_every = (int)v;
    } else if (s == "--ic") {
      if (!need(1)) return 2;
      a->ic = argv[++i];
      if (a->ic != "reference" && a->ic != "interface") { *err = "--ic is reference or interface"; return 2; }
    } else if (s == "--load") {
      if (!need(1)) return 2;
      a->load_dir = argv[++i];
      if (i + 1 < argc && argv[i + 1][0] != '-') {
        if (!inum(argv[++i], 1, 1 << 16, &v)) { *err = "--load DIR R takes a rank count >= 1"; return 2; }
        a->load_ranks = (int)v;
      }
    } else if (s == "--restart") {
      if (!need(1)) return 2;
      a->restart = argv[++i];
    } else if (s == "--checkpoint") {
      if (!need(1)) return 2;
      a->ckpt = argv[++i];
    } else if (s == "--mixed") {
      a->p.precision = SMAISO_FP32_MIXED;
    } else if (s == "--out") {
      if (!need(1)) return 2;
      a->out = argv[++i];
    } else {
      *err = "unknown option " + s;
      return 2;
    }
  }
  if (a->save_every % 50 != 0) { *err = "--save-every must be a positive multiple of 50"; return 2; }
  if (!a->load_dir.empty()) a->p.load_mode = 1;
  return 0;
}
