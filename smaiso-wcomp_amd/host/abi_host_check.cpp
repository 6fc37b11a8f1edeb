// Host-side checks of the C ABI's host-only code under the address and
// undefined-behaviour sanitizers (built by `make asan`, run by
// tests/test_host_sanitizers.py): parameter validation, the interface IC and
// its vector-Jacobian product, the reader of the reference's psi<rank>.dat
// restart files (S:727-752), checkpoint header parsing, and the driver's
// argument parser, each fed valid and malformed inputs.  Exits 0 when every
// check holds; a sanitizer report aborts the run.
#include "host_abi.h"
#include "run_args.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

using namespace smaiso;

static int g_failed = 0;
#define CHECK(cond)                                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond);    \
      ++g_failed;                                                                \
    }                                                                            \
  } while (0)

static std::string tmpdir;

static std::string write_file(const std::string& name, const void* data, size_t n) {
  const std::string path = tmpdir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (n) std::fwrite(data, 1, n, f);
  std::fclose(f);
  return path;
}
static std::string write_text(const std::string& name, const std::string& text) {
  return write_file(name, text.data(), text.size());
}

static smaiso_params small_params(int nx, int ny, int nz) {
  smaiso_params p;
  smaiso_params_default(&p);
  p.nx = nx; p.ny = ny; p.nz = nz;
  return p;
}

static void check_validate() {
  std::string why;
  smaiso_params p = small_params(8, 8, 8);
  CHECK(validate_params(&p, &why) == SMAISO_OK);
  CHECK(validate_params(nullptr, &why) == SMAISO_E_INVALID);
  smaiso_params q = p;
  q.nz = 1;  // N = 1
  CHECK(validate_params(&q, &why) == SMAISO_E_INVALID && why.find(">= 2") != std::string::npos);
  q = p; q.dt = NAN;
  CHECK(validate_params(&q, &why) == SMAISO_E_INVALID);
  q = p; q.nu = -1;
  CHECK(validate_params(&q, &why) == SMAISO_E_INVALID);
  q = p; q.precision = 7;
  CHECK(validate_params(&q, &why) == SMAISO_E_INVALID);
  q = p; q.beta = 0; q.epsilon = 1; q.gamma = -1;  // 9 beta^2 + 40 eps gamma < 0
  CHECK(validate_params(&q, &why) == SMAISO_E_INVALID);
  q = p; q.nx = q.ny = q.nz = 2000;  // 8e9 points
  CHECK(validate_params(&q, &why) == SMAISO_E_INVALID);
}

static void check_ic() {
  smaiso_ic ic{1.0, -1.0, 2.0, 0.9, 1.0, 0.5, 1e-3, 7};
  // N = 1 along an axis is a valid IC grid (only the stepper needs N >= 2)
  for (int shape = 0; shape < 3; ++shape) {
    const int nx = shape == 0 ? 1 : 3, ny = shape == 1 ? 1 : 2, nz = shape == 2 ? 1 : 5;
    smaiso_params p = small_params(nx, ny, nz);
    const size_t n3 = (size_t)nx * ny * nz;
    std::vector<double> psi(n3), rho(n3), h((size_t)nx * ny, 3.0);
    CHECK(smaiso_make_interface_ic(&p, &ic, nullptr, psi.data(), rho.data()) == SMAISO_OK);
    CHECK(smaiso_make_interface_ic(&p, &ic, h.data(), psi.data(), rho.data()) == SMAISO_OK);
    for (size_t i = 0; i < n3; ++i) CHECK(std::isfinite(psi[i]) && rho[i] >= 0.9 - 1e-12 && rho[i] <= 1.0 + 1e-12);
    // dL/dh against central differences of L = sum(w psi + u rho)
    std::vector<double> w(n3), u(n3), dh((size_t)nx * ny);
    for (size_t i = 0; i < n3; ++i) { w[i] = std::sin(0.3 * i + 1); u[i] = std::cos(0.7 * i); }
    CHECK(smaiso_interface_ic_vjp(&p, &ic, h.data(), w.data(), u.data(), dh.data()) == SMAISO_OK);
    smaiso_ic quiet = ic;
    for (size_t c = 0; c < h.size(); ++c) {
      double lp = 0, lm = 0;
      const double e = 1e-6;
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        std::vector<double> hh = h;
        hh[c] += sgn * e;
        CHECK(smaiso_make_interface_ic(&p, &quiet, hh.data(), psi.data(), rho.data()) == SMAISO_OK);
        double l = 0;
        for (size_t i = 0; i < n3; ++i) l += w[i] * psi[i] + u[i] * rho[i];
        (sgn > 0 ? lp : lm) = l;
      }
      const double fd = (lp - lm) / (2 * e);
      CHECK(std::fabs(fd - dh[c]) <= 1e-6 * (1 + std::fabs(fd)));
    }
  }
  smaiso_params p = small_params(3, 2, 5);
  std::vector<double> a(30), b(30), c(6);
  CHECK(smaiso_make_interface_ic(&p, &ic, nullptr, nullptr, b.data()) == SMAISO_E_INVALID);
  CHECK(smaiso_interface_ic_vjp(&p, &ic, nullptr, a.data(), nullptr, c.data()) == SMAISO_E_INVALID);
  smaiso_ic flat = ic;
  flat.width = 0;
  CHECK(smaiso_make_interface_ic(&p, &flat, nullptr, a.data(), b.data()) == SMAISO_E_INVALID);
  smaiso_params bad = small_params(0, 2, 5);
  CHECK(smaiso_make_interface_ic(&bad, &ic, nullptr, a.data(), b.data()) == SMAISO_E_INVALID);
  CHECK(std::strlen(smaiso::abi_error()) > 0);
}

static void check_psi_text() {
  const std::string r0 = write_text("psi0.dat", "1\n2.5\n-3e-2\n");
  const std::string r1 = write_text("psi1.dat", "4 5\n6\n");
  const std::string trunc = write_text("psi_trunc.dat", "7\n8");
  const std::string junk = write_text("psi_junk.dat", "1 2 abc 4\n");
  const std::string empty = write_text("psi_empty.dat", "");
  std::vector<double> psi(6, 0.0);
  const char* two[] = {r0.c_str(), r1.c_str()};
  CHECK(smaiso_read_psi_text(two, 2, 6, psi.data()) == SMAISO_OK);
  CHECK(psi[0] == 1 && psi[2] == -3e-2 && psi[5] == 6);
  // fewer values than needed, in one file and across files
  const char* t[] = {trunc.c_str()};
  CHECK(smaiso_read_psi_text(t, 1, 6, psi.data()) == SMAISO_E_INVALID);
  CHECK(std::strstr(smaiso::abi_error(), "hold 2 values") != nullptr);
  const char* j[] = {junk.c_str()};
  CHECK(smaiso_read_psi_text(j, 1, 4, psi.data()) == SMAISO_E_INVALID);
  const char* e[] = {empty.c_str(), r0.c_str()};
  CHECK(smaiso_read_psi_text(e, 2, 3, psi.data()) == SMAISO_OK && psi[1] == 2.5);
  const std::string missing = tmpdir + "/nope.dat";
  const char* m[] = {missing.c_str()};
  CHECK(smaiso_read_psi_text(m, 1, 1, psi.data()) == SMAISO_E_INVALID);
  const char* nul[] = {nullptr};
  CHECK(smaiso_read_psi_text(nul, 1, 1, psi.data()) == SMAISO_E_INVALID);
  CHECK(smaiso_read_psi_text(two, 0, 1, psi.data()) == SMAISO_E_INVALID);
  CHECK(smaiso_read_psi_text(two, 2, -1, psi.data()) == SMAISO_E_INVALID);
  CHECK(smaiso_read_psi_text(two, 2, 0, psi.data()) == SMAISO_OK);
}

static void check_checkpoint_header() {
  CkptHeader good;
  ckpt_header_init(&good, 16, 8, 32, SMAISO_FP64, 8, 120);
  std::string why;
  CkptHeader out;
  auto read = [&](const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    const int rc = ckpt_header_read(f, 16, 8, 32, SMAISO_FP64, 8, &out, &why);
    std::fclose(f);
    return rc;
  };
  CHECK(read(write_file("good.ck", &good, sizeof good)) == SMAISO_OK && out.nloop == 120);
  CHECK(read(write_file("empty.ck", nullptr, 0)) == SMAISO_E_INVALID);
  CHECK(read(write_file("half.ck", &good, sizeof good / 2)) == SMAISO_E_INVALID);
  CkptHeader h = good;
  h.magic[0] = 'X';
  CHECK(read(write_file("magic.ck", &h, sizeof h)) == SMAISO_E_INVALID && why == "not a smaiso checkpoint");
  h = good; h.version = 9;
  CHECK(read(write_file("version.ck", &h, sizeof h)) == SMAISO_E_INVALID && why.find("version 9") != std::string::npos);
  h = good; h.nloop = -5;
  CHECK(read(write_file("nloop.ck", &h, sizeof h)) == SMAISO_E_INVALID && why.find("negative") != std::string::npos);
  h = good; h.nz = 31;
  CHECK(read(write_file("grid.ck", &h, sizeof h)) == SMAISO_E_INVALID && why.find("does not match") != std::string::npos);
  h = good; h.precision = SMAISO_FP32_MIXED; h.elem = 4;
  CHECK(read(write_file("prec.ck", &h, sizeof h)) == SMAISO_E_INVALID);
  const std::string text = write_text("text.ck", "psi0.dat is text, not a checkpoint\n");
  CHECK(read(text) == SMAISO_E_INVALID);
}

static int parse(std::vector<std::string> words, RunArgs* a, std::string* err) {
  std::vector<char*> argv;
  for (auto& w : words) argv.push_back(&w[0]);
  argv.push_back(nullptr);
  return parse_run_args((int)words.size(), argv.data(), a, err);
}

static void check_args() {
  RunArgs a;
  std::string err;
  CHECK(parse({"smaiso_run", "1", "5", "--n", "16", "8", "32", "--dt", "0.01", "--steps", "100", "--lim-l1", "1e-4",
               "--load", "dir", "3", "--mixed", "--save-every", "100"}, &a, &err) == 0);
  CHECK(a.p.nx == 16 && a.p.ny == 8 && a.p.nz == 32 && a.p.dt == 0.01 && a.steps == 100 && a.lim_l1 == 1e-4);
  CHECK(a.load_ranks == 3 && a.p.load_mode == 1 && a.p.precision == SMAISO_FP32_MIXED && a.save_every == 100);
  CHECK(std::fabs(a.p.epsilon + 0.005) < 1e-15 && a.p.nu == 1);
  const std::vector<std::vector<std::string>> bad = {
      {"smaiso_run"},
      {"smaiso_run", "1"},
      {"smaiso_run", "x", "5"},
      {"smaiso_run", "1", "5", "--n", "16", "8"},
      {"smaiso_run", "1", "5", "--n", "16", "8", "0"},
      {"smaiso_run", "1", "5", "--n", "16x", "8", "8"},
      {"smaiso_run", "1", "5", "--dt"},
      {"smaiso_run", "1", "5", "--dt", "fast"},
      {"smaiso_run", "1", "5", "--steps", "-3"},
      {"smaiso_run", "1", "5", "--lim-l1", "0"},
      {"smaiso_run", "1", "5", "--lim-l1", "nan"},
      {"smaiso_run", "1", "5", "--save-every", "70"},
      {"smaiso_run", "1", "5", "--ic", "bogus"},
      {"smaiso_run", "1", "5", "--load", "dir", "0"},
      {"smaiso_run", "1", "5", "--divv-switch", "-1"},
      {"smaiso_run", "1", "5", "--bogus"},
  };
  for (const auto& w : bad) {
    RunArgs b;
    std::string e;
    CHECK(parse(w, &b, &e) == 2 && !e.empty());
  }
}

int main() {
  char tmpl[] = "/tmp/smaiso_host_check_XXXXXX";
  if (!mkdtemp(tmpl)) return 2;
  tmpdir = tmpl;
  check_validate();
  check_ic();
  check_psi_text();
  check_checkpoint_header();
  check_args();
  std::string cmd = "rm -rf " + tmpdir;
  if (std::system(cmd.c_str()) != 0) std::fprintf(stderr, "could not remove %s\n", tmpdir.c_str());
  if (g_failed) {
    std::fprintf(stderr, "%d checks failed\n", g_failed);
    return 1;
  }
  std::printf("host ABI checks passed\n");
  return 0;
}
