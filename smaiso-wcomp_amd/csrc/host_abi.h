// Host-only parts of the C ABI (no HIP): the handle-less error text,
// parameter validation, the checkpoint header, and (in host_abi.cpp) the
// exported functions that touch no device: smaiso_params_default,
// smaiso_make_interface_ic, smaiso_interface_ic_vjp, smaiso_read_psi_text.
// Plain C++, so that it also builds with the host sanitizers.
#pragma once

#include "smaiso.h"

#include <cstdint>
#include <cstdio>
#include <string>

namespace smaiso {

// Record msg as the text smaiso_last_error(NULL) returns (per thread) and
// return code.
int abi_fail(int code, const std::string& msg);
// The calling thread's handle-less error text.
const char* abi_error();

// SMAISO_OK, or SMAISO_E_INVALID with the reason in *why.
int validate_params(const smaiso_params* p, std::string* why);

// Checkpoint file: this header, then psiq, psiq_old, rho, f_rho (FP64) and
// psi, N^l of the last step (work precision), each nx*ny*nz values.
struct CkptHeader {
  char magic[8];
  int32_t version, nx, ny, nz, precision, elem;
  int64_t nloop;
};
constexpr int32_t kCkptVersion = 1;
void ckpt_header_init(CkptHeader* h, int nx, int ny, int nz, int precision, int elem, int64_t nloop);
// Reads and checks the header at the start of f against the grid,
// precision and element size a handle expects: SMAISO_OK, or
// SMAISO_E_INVALID with the reason in *why (not a checkpoint, another
// version, a negative step count, another grid or precision).
int ckpt_header_read(std::FILE* f, int nx, int ny, int nz, int precision, int elem, CkptHeader* out,
                     std::string* why);

}  // namespace smaiso
