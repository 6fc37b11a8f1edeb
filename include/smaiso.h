/*
 * smaiso.h - C ABI of the MI355X pseudo-spectral stepper for the weakly
 * compressible smectic-A / isotropic phase-field model (order parameter psi,
 * density rho, quasi-static velocity v).
 *
 * The reference (smaiso-wcomp.cpp, cited as S:) has no library API: its hot
 * path is the body of main()'s loop `while (L1 > limL1)` (S:1081-2198), and
 * its only operator interface is FFTW's r2r plan/execute pair (S:499-545,
 * S:2483-2492).  Each entry point below names the reference code it stands
 * in for.
 *
 * Conventions
 *  - Plain C types only.  Status: 0 = ok, < 0 = error (SMAISO_E_*).  The text
 *    of the last error is returned by smaiso_last_error().
 *  - Host arrays use the reference layout: index (i*Ny + j)*Nz + k with k
 *    fastest (S:943), FP64.
 *  - The caller owns host buffers; the library owns device memory.  One handle
 *    owns one HIP stream; a handle is not thread-safe, distinct handles are
 *    independent.
 */
#ifndef SMAISO_H
#define SMAISO_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMAISO_ABI_VERSION 1

/* status codes */
#define SMAISO_OK 0
#define SMAISO_E_INVALID (-1)   /* bad argument or parameter */
#define SMAISO_E_HIP (-2)       /* HIP runtime error */
#define SMAISO_E_NOMEM (-3)     /* device allocation failed */
#define SMAISO_E_STATE (-4)     /* call out of order (e.g. step before set_state) */
#define SMAISO_E_NONFINITE (-5) /* diagnostics found NaN/Inf in the state */

/* precision of the stepper */
#define SMAISO_FP64 0 /* every field, transform and kernel in FP64 */
/* FP32 storage of the work fields (psi, grad psi, grad rho, amp, L psi, mu,
 * forces, velocities, Nl, ...); FP64 state (psiq, psiq_old), rho and f_rho;
 * grad rho formed from the FP64 rho (its y and x passes in FP32 arithmetic
 * under SMAISO_STAGE_GRAD_RHO, its z component in FP64 by the density pass);
 * FP64 arithmetic in the forces and density kernels, the psi update, the
 * implicit denominators and the compensated reductions; the transform that
 * forms psi from psiq keeps FP64 intermediates.  Transform passes and the
 * velocity solve compute in FP64 (FP64 LDS tiles), except those of the
 * stages selected by smaiso_set_fp32_stages, which compute in FP32. */
#define SMAISO_FP32_MIXED 1

/* FFTW r2r kind numbering used by smaiso_r2r3d (S:499-545) */
#define SMAISO_REDFT10 0  /* DCT-II  */
#define SMAISO_REDFT01 1  /* DCT-III */
#define SMAISO_RODFT10 2  /* DST-II  */
#define SMAISO_RODFT01 3  /* DST-III */

/* fields readable with smaiso_get_field (real space unless noted) */
#define SMAISO_FIELD_PSI 0
#define SMAISO_FIELD_RHO 1
#define SMAISO_FIELD_AMP 2        /* Gaussian-smoothed amplitude, S:1254-1289 */
#define SMAISO_FIELD_VX 3         /* total velocity vsol + virr, S:2125-2131 */
#define SMAISO_FIELD_VY 4
#define SMAISO_FIELD_VZ 5
#define SMAISO_FIELD_VSOLX 6      /* solenoidal part, S:2094-2104 */
#define SMAISO_FIELD_VSOLY 7
#define SMAISO_FIELD_VSOLZ 8
#define SMAISO_FIELD_VIRRX 9      /* irrotational part, S:2106-2116 */
#define SMAISO_FIELD_VIRRY 10
#define SMAISO_FIELD_VIRRZ 11
#define SMAISO_FIELD_DIVV 12      /* div v, S:2085-2088 */
#define SMAISO_FIELD_LPSI 13      /* alpha (q0^2 + lap) psi at the step start, S:1293-1308 */
#define SMAISO_FIELD_MU 14        /* chemical potential, S:1411-1413 */
#define SMAISO_FIELD_ENERGY 15    /* free-energy density of the diagnostics, S:2227-2233 */
#define SMAISO_FIELD_PSIQ 16      /* normalised DCT coefficients of psi (spectral) */
#define SMAISO_FIELD_PSIQ_OLD 17  /* previous psiq (spectral) */
#define SMAISO_FIELD_NLQ 18       /* DCT of the nonlinear term (spectral, unnormalised) */
#define SMAISO_FIELD_F_RHO 19     /* AB2 history of the density, S:2136 */
#define SMAISO_FIELD_DIVF 20      /* div f, S:2090-2092 */
#define SMAISO_FIELD_GPSIX 21     /* grad psi, S:1226-1236 */
#define SMAISO_FIELD_GPSIY 22
#define SMAISO_FIELD_GPSIZ 23
#define SMAISO_FIELD_GRHOX 24     /* grad rho, S:1240-1250 */
#define SMAISO_FIELD_GRHOY 25
#define SMAISO_FIELD_GRHOZ 26
#define SMAISO_FIELD_LAPRHOL 27   /* lap(rho L psi), S:1312-1380 */
#define SMAISO_FIELD_PSI_OLD 28   /* psi at the start of the last step */
#define SMAISO_FIELD_COUNT 29

/*
 * Parameters (S:206-283).  smaiso_params_default() fills the reference's
 * constants for a 512^3 grid; the caller then edits what it needs.
 * Derived quantities follow the reference: Amp = sqrt((3 beta +
 * sqrt(9 beta^2 + 40 epsilon gamma))/5) (S:250), kp = (1 - rho0)/Amp (S:252),
 * rho_s = kp Amp + rho0 (S:253), scale = 1/(8 Nx Ny Nz) (S:278).
 */
typedef struct smaiso_params {
  int32_t nx, ny, nz;        /* grid (S:218) */
  double dx, dy, dz;         /* spacing, 2 pi/(q0 Nw) in the reference (S:267-272) */
  double dt;                 /* time step (S:262) */
  double alpha, beta, gamma; /* phase-field coefficients (S:239-241) */
  double epsilon;            /* -0.001*ep_arg (S:242-243) */
  double q0;                 /* layer wavenumber (S:244) */
  double nu;                 /* viscosity, argv[1] (S:249) */
  double lambda;             /* second viscosity (S:256) */
  double zeta;               /* density coupling (S:257) */
  double rho0;               /* isotropic density rho_0 (S:251) */
  double rho_m;              /* stabilisation density in C1, C2 (S:255: rho_s/2); <= 0 means rho_s/2 */
  int32_t divv_switch;       /* relaxation steps with rho slaved to amp (S:214) */
  int32_t load_mode;         /* 1 = restart semantics: skip relaxation (S:1312) */
  int32_t precision;         /* SMAISO_FP64 or SMAISO_FP32_MIXED */
  int32_t slab_n0;           /* reference slab height for quirk Q3 (S:1932); 0 = Nx */
} smaiso_params;

/* Interface initial condition (builder-defined, see smaiso_make_interface_ic). */
typedef struct smaiso_ic {
  double amplitude;   /* A in psi = A cos(q0 z) * profile */
  double h0;          /* mean interface height; < 0 means Lz/2 */
  double width;       /* tanh width w */
  double rho_iso;     /* density above the interface */
  double rho_sm;      /* density below the interface */
  double undulation;  /* u in h = h0 + u cos(2 pi x/Lx) cos(2 pi y/Ly) */
  double noise;       /* uniform noise amplitude on psi */
  uint64_t seed;      /* splitmix64 seed */
} smaiso_ic;

typedef struct smaiso_handle smaiso_handle;

/* Library ABI version (SMAISO_ABI_VERSION). No reference counterpart. */
int smaiso_abi_version(void);

/* Fill the reference's constants (S:206-283, argv defaults nu=1, ep_arg=5,
 * 512^3 grid). */
void smaiso_params_default(smaiso_params* p);

/* Replaces the reference's setup S:293-545 and the operator tables S:936-963:
 * validates p, allocates device fields on the current HIP device and builds
 * the transform tables.  Returns SMAISO_E_INVALID for N < 2, dt <= 0, nu <= 0,
 * lambda + 2 nu <= 0, non-finite parameters or an unknown precision. */
int smaiso_create(const smaiso_params* p, smaiso_handle** out);

/* Replaces fftw_destroy_plan/MPI::Finalize (S:2483-2496). NULL is a no-op. */
int smaiso_destroy(smaiso_handle* h);

/* Text of the last error on h (or of the last failed smaiso_create when h is
 * NULL).  Never NULL. */
const char* smaiso_last_error(const smaiso_handle* h);

/* Replaces the IC / load block and the pre-loop (S:556-796, S:965-1010):
 * psiq = scale*DCT(psi), psi = iDCT(psiq), psiq_old = psiq, f_rho = 0 and the
 * step counter reset (quirk Q5).  psi, rho: host FP64 arrays of Nx*Ny*Nz. */
int smaiso_set_state(smaiso_handle* h, const double* psi, const double* rho);

/* As smaiso_set_state with device FP64 arrays (already in HBM). */
int smaiso_set_state_device(smaiso_handle* h, const void* psi_dev, const void* rho_dev);

/* Replaces nsteps iterations of the loop body S:1081-2198.  Synchronous. */
int smaiso_step(smaiso_handle* h, int32_t nsteps);

/* Enqueue nsteps on the handle's stream without waiting. */
int smaiso_step_async(smaiso_handle* h, int32_t nsteps);

/* Wait for the handle's stream. */
int smaiso_synchronize(smaiso_handle* h);

/* Steps completed since the last set_state (the reference's nLoop). */
int64_t smaiso_step_count(const smaiso_handle* h);

/* Stages of the SMAISO_FP32_MIXED step that compute in FP32 rather than
 * FP64: the transform passes of a stage (FP32 LDS tiles and tables), or the
 * velocity solve's arithmetic.  The transform that forms psi from psiq, the
 * forces and density kernels, the psi update and the reductions always
 * compute in FP64.  FP32 arithmetic in a transform rounds every mode of a
 * line relative to the line's norm, so it is kept to the stages where that
 * rounding does not reach psi. */
#define SMAISO_STAGE_GRAD_PSI (1u << 0) /* grad psi, one pass per axis (S:1104-1236) */
#define SMAISO_STAGE_AMP (1u << 1)      /* Gaussian amplitude filter's factors (S:1254-1289);
                                         * d_z psi, formed in the same z pass, stays FP64 */
#define SMAISO_STAGE_LPSI (1u << 2)     /* L psi (S:1293-1308) */
#define SMAISO_STAGE_LAPRHOL (1u << 3)  /* lap(rho L psi) (S:1312-1342) */
#define SMAISO_STAGE_FORCES (1u << 4)   /* y and x passes of fgrad, erho, S_a (S:1427-1457) */
#define SMAISO_STAGE_DIVV (1u << 5)     /* div v's transform (S:2042-2088); its z pass, shared
                                         * with v's in the fused velocity pass, is FP32 only
                                         * with SMAISO_STAGE_VELOCITY too */
#define SMAISO_STAGE_DIVF (1u << 6)     /* div f's y and x passes (S:2042-2092); the z term is
                                         * formed in FP64 by the density pass */
#define SMAISO_STAGE_VELOCITY (1u << 7) /* v's inverse transforms (S:2094-2117) */
#define SMAISO_STAGE_NL (1u << 8)       /* y and x passes of Nl's transform (S:2166-2168) */
#define SMAISO_STAGE_VELSOLVE (1u << 9) /* the velocity solve's arithmetic (S:1479-1974) */
#define SMAISO_STAGE_GRAD_RHO (1u << 10) /* grad rho's y and x passes from the FP64 rho, at
                                          * length 256 (S:1104-1250) */
#define SMAISO_STAGES_ALL 0x7ffu

/* The FP32-arithmetic stages of a SMAISO_FP32_MIXED handle (SMAISO_STAGE_*
 * bits; the default is smaiso_fp32_stages_default()).  Ignored by FP64
 * handles.  No reference counterpart. */
int smaiso_set_fp32_stages(smaiso_handle* h, uint32_t mask);
uint32_t smaiso_fp32_stages_default(void);

/* enable != 0: smaiso_step/smaiso_step_async replay pairs of steady steps
 * (past the relaxation branch, S:1312-1380) from one captured hipGraph of two
 * steps instead of launching them one by one (off by default; ignored while
 * timing is enabled).  Results are identical to eager steps.  No reference
 * counterpart: it replaces S:1081-2198 run twice. */
int smaiso_set_graph(smaiso_handle* h, int32_t enable);

/* Copy a field (SMAISO_FIELD_*) of the last completed step to host FP64 out
 * (Nx*Ny*Nz values, reference layout).  Replaces reading the reference's
 * arrays (psi_local, rho_local, ...) and its slice writers (S:2323-2446).
 * The state fields (PSI, RHO, PSIQ, PSIQ_OLD, F_RHO) are readable once the
 * state is set; every other field is produced by a step and returns
 * SMAISO_E_STATE until a step has completed since the last set_state or
 * checkpoint load (NLQ is also restored by a checkpoint load). */
int smaiso_get_field(smaiso_handle* h, int32_t field, double* out);

/* As smaiso_get_field into a device FP64 buffer. */
int smaiso_get_field_device(smaiso_handle* h, int32_t field, void* out_dev);

/* Diagnostics of the last completed step, S:2202-2263: L1 = sum|psi-psi_old|/
 * (dt sum|psi|), mass = sum(rho) dx dy dz, free energy (S:2227-2233).  FP64
 * compensated sums on the device.  SMAISO_E_STATE until a step has completed
 * since the last set_state or checkpoint load.  Returns SMAISO_E_NONFINITE
 * (values still written) when any of them is NaN or Inf (the reference's loop
 * silently exits in that case, S:1081). */
int smaiso_diagnostics(smaiso_handle* h, double* l1, double* mass, double* energy);

/* Standalone 3-D r2r transform with FFTW semantics (unnormalised), the
 * equivalent of fftw_mpi_plan_r2r_3d + fftw_execute (S:499-545): kinds[a]
 * is the kind on axis a (0 = x, slowest).  in/out: host FP64 arrays of
 * nx*ny*nz; dtype SMAISO_FP64 stores and computes in FP64,
 * SMAISO_FP32_MIXED rounds the input to FP32, stores every pass's output in
 * FP32 and computes each pass in FP64.  Uses the current HIP device and the
 * null stream; synchronous. */
int smaiso_r2r3d(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz,
                 const double* in, double* out, int32_t dtype);

/* As smaiso_r2r3d on device buffers of the given dtype (double or float;
 * FP64 arithmetic either way), enqueued on `stream` (a hipStream_t, NULL =
 * null stream), asynchronous. */
int smaiso_r2r3d_device(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz,
                        const void* in_dev, void* out_dev, int32_t dtype, void* stream);

/* No counterpart in the reference: smaiso_r2r3d for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and results
 * (bit-identical) as smaiso_r2r3d, but the device staging of the data and
 * the transform tables live in a fixed device array of about 24 MiB that
 * each GPU the library runs on holds until the process ends; the call makes
 * no hipMalloc/hipFree.  Limits: at most SMAISO_NOALLOC_MAX_POINTS points
 * and SMAISO_NOALLOC_MAX_LEN per axis, else SMAISO_E_INVALID.  Every call on
 * one GPU uses that GPU's one array: calls from different threads would race
 * on it, so the library serialises them (each call holds a process-wide lock
 * until its result is on the host).  Uses the null stream; synchronous. */
#define SMAISO_NOALLOC_MAX_POINTS 2097152 /* 128^3 */
#define SMAISO_NOALLOC_MAX_LEN 1024
int smaiso_r2r3d_noalloc(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz,
                         const double* in, double* out, int32_t dtype);

/* No reference counterpart: the HBM rate (GB/s, bytes read plus bytes
 * written per second) a plain copy kernel reaches on the current device,
 * copying `bytes` (a multiple of 16, at least 1 MiB) `reps` times between
 * two buffers it allocates and frees; the measured reference beside the
 * 8 TB/s specification.  Synchronous. */
int smaiso_hbm_copy_rate(int64_t bytes, int32_t reps, double* gbs);

/* Builder-defined initial condition (the reference's own IC is S:571-590):
 *   profile = 1/2 (1 - tanh((z - h(x,y))/width)),  z = k dz, x = i dx, y = j dy
 *   h(x,y)  = h0 + undulation cos(2 pi x/Lx) cos(2 pi y/Ly)  (or height[i*ny+j])
 *   psi = amplitude cos(q0 z) profile + noise (2u - 1)
 *   rho = rho_iso + (rho_sm - rho_iso) profile
 * u = U[0,1) from splitmix64 (state seed + (n+1)*0x9E3779B97F4A7C15, top 53
 * bits) in row-major order.  height may be NULL.  Host only. */
int smaiso_make_interface_ic(const smaiso_params* p, const smaiso_ic* ic, const double* height,
                             double* psi, double* rho);

/* Per-kernel-class timing with HIP events on the handle's stream.
 * enable != 0 starts collecting (and clears totals). */
int smaiso_timing_enable(smaiso_handle* h, int32_t enable);

/* Kernel classes for smaiso_timing_get. */
#define SMAISO_KC_R2R_Z 0       /* transform pass along z (plain or with a fused load) */
#define SMAISO_KC_R2R_Y 1       /* transform pass along y */
#define SMAISO_KC_R2R_X 2       /* transform pass along x (also the fused mid operators) */
#define SMAISO_KC_SPECTRAL 3    /* per-mode kernels */
#define SMAISO_KC_POINTWISE 4   /* real-space kernels (also the forces fused into a z pass) */
#define SMAISO_KC_REDUCE 5      /* diagnostics */
#define SMAISO_KC_R2R_AXIS 6    /* one-axis operator passes (forward, operator, inverse along one axis) */
#define SMAISO_KC_COUNT 7

/* Total milliseconds, launch count and algorithmic bytes moved by kernel
 * class kc since timing was enabled. */
int smaiso_timing_get(smaiso_handle* h, int32_t kc, double* total_ms, int64_t* launches,
                      double* bytes);

/*
 * Adjoint (no reference counterpart): reverse-mode gradients of
 *   L = sum over the grid of (psi(T) - target)^2
 * after nsteps steps from smaiso_set_state(psi0, rho0), with respect to psi0,
 * rho0 (host FP64 arrays of Nx*Ny*Nz), zeta and nu.  The forward trajectory
 * is stored every checkpoint_every steps and each segment is recomputed from
 * its checkpoint before its steps are transposed (hand-written adjoints of
 * S:1081-2198 with transposed DCT/DST passes), so device memory grows with
 * nsteps/checkpoint_every + checkpoint_every stored states.  The adjoint
 * fields are FP64; with SMAISO_FP32_MIXED the trajectory it differentiates is
 * the mixed-precision one.  Needs the steady branch for every step
 * (divv_switch = 0 or load_mode = 1; SMAISO_E_INVALID otherwise).  dpsi0 and
 * drho0 may be NULL.  The handle's state is consumed: call smaiso_set_state
 * before stepping it again.
 */
int smaiso_adjoint(smaiso_handle* h, const double* psi0, const double* rho0, const double* target,
                   int32_t nsteps, int32_t checkpoint_every, double* loss, double* dpsi0,
                   double* drho0, double* dzeta, double* dnu);

/* enable != 0 (the default): while smaiso_adjoint recomputes a segment it
 * keeps each step's intermediates (20 work fields and two FP64 fields per
 * step, checkpoint_every steps), so every step is recomputed once; it does
 * so only when they fit in free device memory with 1 GiB to spare.
 * enable == 0, or no room: each step is recomputed again from its stored
 * state before its transpose (less memory, one more forward step per step).
 * No reference counterpart. */
int smaiso_set_adjoint_store(smaiso_handle* h, int32_t enable);

/* dL/dh(x, y) of the interface initial condition (smaiso_make_interface_ic)
 * from dL/dpsi and dL/drho of its output: sum over z of (amplitude cos(q0 z)
 * dpsi + (rho_sm - rho_iso) drho) * d profile/dh, d profile/dh =
 * sech^2((z - h)/width)/(2 width).  height may be NULL (the analytic h0 +
 * undulation map).  dheight: Nx*Ny.  Host only. */
int smaiso_interface_ic_vjp(const smaiso_params* p, const smaiso_ic* ic, const double* height,
                            const double* dpsi, const double* drho, double* dheight);

/*
 * Autotuner (no reference counterpart).  For each axis whose length runs on
 * the LDS FFT path, times a forward + inverse plain pass of the handle's
 * batch geometry with each of SMAISO_TILE_VARIANTS tile shapes (lines per
 * tile, threads per workgroup; variant 0 is the built-in default) and keeps
 * the fastest for that handle's later plain passes.  FP64 is tuned always
 * (the density transforms), FP32 for SMAISO_FP32_MIXED handles.  ms
 * (optional): 2 x 3 x SMAISO_TILE_VARIANTS milliseconds [dtype FP64, FP32]
 * [axis x, y, z][variant], -1 where not timed; chosen (optional): 2 x 3
 * variants.  Results change only by rounding (the FFT radix plan depends on
 * the tile).  Allocates scratch and synchronises; a captured step graph is
 * recaptured afterwards.
 */
#define SMAISO_TILE_VARIANTS 4
int smaiso_autotune(smaiso_handle* h, float* ms, int32_t* chosen);

/* smaiso_r2r3d_device with tile variant `variant` forced on every axis
 * (SMAISO_E_INVALID if the variant is not built for a length); synchronous.
 * For validating the autotuner's variants. */
int smaiso_r2r3d_tiles(const int32_t kinds[3], int32_t nx, int32_t ny, int32_t nz, const void* in_dev,
                       void* out_dev, int32_t dtype, int32_t variant, void* stream);

/* Data of the reference's output writers (S:2273-2446).  smaiso_get_slice:
 * the x-z slice at y index j of a field (saveMid, S:94-129, with j = Ny/2):
 * out[i*Nz + k] = field[(i*Ny + j)*Nz + k], Nx*Nz values.  smaiso_surface_height:
 * the surface tracker (S:2278-2318): surf[i*Ny + j] = the largest k with
 * rho > threshold (the reference uses rho_s - 0.15 with rho_s = kp Amp + rho0
 * = 1); a column with no such k keeps the value surf held on entry (the
 * reference's stale entry).  Both synchronous. */
int smaiso_get_slice(smaiso_handle* h, int32_t field, int32_t j, double* out);
int smaiso_surface_height(smaiso_handle* h, double threshold, double* surf);

/* Full-state checkpoint of a single handle (fixes the reference's psi-only
 * restart, S:727-796 and S:2334-2345): psiq, psiq_old, rho, f_rho, psi, the
 * last step's N^l and nLoop, binary.  Loading needs the same grid and
 * precision; stepping on after a load is bit-identical to never stopping.
 * A file that is not a checkpoint, or one of another grid or precision,
 * returns SMAISO_E_INVALID and leaves the handle as it was; a truncated
 * one returns SMAISO_E_INVALID and leaves the handle without a state. */
int smaiso_save_checkpoint(smaiso_handle* h, const char* path);
int smaiso_load_checkpoint(smaiso_handle* h, const char* path);

/* Reader of the reference's restart files psi<rank>.dat (S:727-748: one value
 * per line, each rank's x-slab in rank order): reads whitespace-separated
 * values from paths[0], paths[1], ... until `count` (= Nx*Ny*Nz) values are in
 * psi.  SMAISO_E_INVALID if a file is missing or the files hold fewer values.
 * Host only. */
int smaiso_read_psi_text(const char* const* paths, int32_t npaths, int64_t count, double* psi);

/*
 * Ensembles: `count` independent simulations (each one the reference's whole
 * program, S:1081-2263, with its own parameters and grid) stepped together on
 * one device.  Members that share grid shape, spacing, dt, alpha, beta,
 * gamma, epsilon, q0, the effective rho_m, divv_switch, load_mode, precision
 * and slab_n0 form one group and run as one batch (every kernel launch covers
 * all members of the group); nu, lambda, zeta and rho0 may differ inside a
 * group.  Groups run on their own streams, and steady steps of all groups
 * replay from one captured hipGraph (smaiso_ensemble_set_graph, on by
 * default).  Every member's results equal those of a single handle with the
 * same parameters and state.  No reference counterpart (the reference runs one
 * simulation per MPI job).
 */
typedef struct smaiso_ensemble smaiso_ensemble;

int smaiso_ensemble_create(const smaiso_params* members, int32_t count, smaiso_ensemble** out);
int smaiso_ensemble_destroy(smaiso_ensemble* e);
/* Text of the last error on e (or of the last failed smaiso_ensemble_create when e is NULL). */
const char* smaiso_ensemble_last_error(const smaiso_ensemble* e);
/* Number of batches the members were grouped into. */
int smaiso_ensemble_groups(const smaiso_ensemble* e);
/* As smaiso_set_state for one member; every member must be set before the
 * first step.  Setting a member restarts the step count (nLoop, with quirk
 * Q5's first-step history) of the batch it belongs to.  If that batch has
 * stepped since it was last set, its other members count as unset again and
 * smaiso_ensemble_step returns SMAISO_E_STATE until each has been set. */
int smaiso_ensemble_set_state(smaiso_ensemble* e, int32_t member, const double* psi,
                              const double* rho);
int smaiso_ensemble_set_graph(smaiso_ensemble* e, int32_t enable);
/* nsteps of every member; synchronous. */
int smaiso_ensemble_step(smaiso_ensemble* e, int32_t nsteps);
/* Steps since the member's batch was last set (-1 for a bad member). */
int64_t smaiso_ensemble_step_count(const smaiso_ensemble* e, int32_t member);
int smaiso_ensemble_get_field(smaiso_ensemble* e, int32_t member, int32_t field, double* out);
int smaiso_ensemble_diagnostics(smaiso_ensemble* e, int32_t member, double* l1, double* mass,
                                double* energy);
/* smaiso_autotune for every batch of the ensemble. */
int smaiso_ensemble_autotune(smaiso_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* SMAISO_H */
