/* bns_nurates_amd.h -- C ABI of the MI355X gray M1 neutrino-opacity library.
 *
 * Drop-in boundary for the gray (energy-integrated) M1 opacity path of
 * bns_nurates, and for its spectral (energy-dependent) opacities.  Every
 * entry point is plain C: pointers, sizes and ints; no HIP or torch types.
 * Device pointers are HIP device addresses; streams are passed as `void*` (a
 * hipStream_t).  Errors are returned as status codes, never by abort();
 * bnsamd_last_error() gives the message of the context's last failure.
 *
 * Data layout (SoA, one double per point per field, field f of point i at
 * base[f * ld + i], ld >= npts):
 *   point block  21 fields  BNSAMD_PT_*   (MyEOSParams + M1Quantities)
 *   distribution 28 fields  BNSAMD_DI_*   (NuDistributionParams), optional
 *   output       20 fields  BNSAMD_OUT_*  (M1Opacities)
 * Units are the reference's: nb [nm^-3], T and chemical potentials [MeV]
 * (rest masses included), n [nm^-3], J [MeV nm^-3]; outputs eta_0
 * [nm^-3 s^-1], eta [MeV nm^-3 s^-1], kappa [nm^-1].
 */
#ifndef BNS_NURATES_AMD_H_
#define BNS_NURATES_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
  BNSAMD_OK = 0,
  BNSAMD_ERR_INVALID = 1,       /* bad argument: null pointer, size, order */
  BNSAMD_ERR_HIP = 2,           /* HIP runtime failure */
  BNSAMD_ERR_NO_DEVICE = 3,     /* no usable HIP device */
  BNSAMD_ERR_NO_QUADRATURE = 4, /* bnsamd_set_quadrature not called */
  BNSAMD_ERR_NOMEM = 5          /* device allocation failed */
};

/* ---- field indices --------------------------------------------------------
 * point block: MyEOSParams fields read by the path (bns_nurates.hpp:289-306)
 * then M1Quantities n, J, chi per species (bns_nurates.hpp:409-416); species
 * order nue, anue, nux, anux (bns_nurates.hpp:57-60). */
enum {
  BNSAMD_PT_NB = 0, BNSAMD_PT_TEMP = 1, BNSAMD_PT_YP = 2, BNSAMD_PT_YN = 3,
  BNSAMD_PT_MU_P = 4, BNSAMD_PT_MU_N = 5, BNSAMD_PT_MU_E = 6, BNSAMD_PT_DU = 7,
  BNSAMD_PT_DM_EFF = 8, BNSAMD_PT_N = 9, BNSAMD_PT_J = 13, BNSAMD_PT_CHI = 17,
  BNSAMD_PT_NFIELDS = 21
};
/* distribution block: NuDistributionParams (bns_nurates.hpp:387-401) */
enum {
  BNSAMD_DI_W_T = 0, BNSAMD_DI_TEMP_T = 4, BNSAMD_DI_ETA_T = 8, BNSAMD_DI_W_F = 12,
  BNSAMD_DI_TEMP_F = 16, BNSAMD_DI_C_F = 20, BNSAMD_DI_BETA_F = 24,
  BNSAMD_DI_NFIELDS = 28
};
/* output block: M1Opacities (bns_nurates.hpp:463-474) */
enum {
  BNSAMD_OUT_ETA_0 = 0, BNSAMD_OUT_KAPPA_0_A = 4, BNSAMD_OUT_ETA = 8,
  BNSAMD_OUT_KAPPA_A = 12, BNSAMD_OUT_KAPPA_S = 16, BNSAMD_OUT_NFIELDS = 20
};

/* spectral output block: SpectralOpacities (bns_nurates.hpp:517-523), one
 * field per (quantity, species), each field [nnu][ld]: value of energy e at
 * point i at base[(f * nnu + e) * ld + i] */
enum {
  BNSAMD_SP_J = 0, BNSAMD_SP_KAPPA = 4, BNSAMD_SP_J_S = 8, BNSAMD_SP_KAPPA_S = 12,
  BNSAMD_SP_NFIELDS = 16
};

/* OpacityFlags (bns_nurates.hpp:418-426), same field order and types */
typedef struct bnsamd_flags {
  int use_abs_em, use_pair, use_brem, use_inelastic_scatt, use_iso;
} bnsamd_flags;

/* OpacityParams (bns_nurates.hpp:338-352) with every bool as int;
 * brem_implementation: 0 BREM_HR98, 1 BREM_BRT06, 2 BREM_GP19 (over the
 * context's bnsamd_set_gp19_table; without one it adds zero, as the
 * reference's table-free configuration does, kernel_brem_GP19.hpp:483) */
typedef struct bnsamd_params {
  int use_dU, use_dm_eff, use_WM_ab, use_WM_sc, use_decay, brem_implementation,
      use_NN_medium_corr, neglect_blocking;
} bnsamd_params;

/* precision of a gray opacity evaluation */
enum { BNSAMD_FP64 = 0, BNSAMD_MIXED = 1 };

typedef struct bnsamd_context bnsamd_context;

/* ---- context --------------------------------------------------------------
 * One context per device; it owns a stream (used until bnsamd_set_stream
 * selects another; NULL selects HIP's default stream), the resident
 * quadrature tables and the staging buffers of the host entry point.  Calls
 * on one context are serialised by an internal lock; results never depend on
 * how points are batched.  bnsamd_set_stream waits for the work already
 * queued on the previous stream before switching, so that buffers the
 * library reuses (staging, fallback workspaces) are never in use on two
 * streams at once. */
int bnsamd_create(int device, bnsamd_context** out);
void bnsamd_destroy(bnsamd_context* ctx);
const char* bnsamd_last_error(const bnsamd_context* ctx);
int bnsamd_set_stream(bnsamd_context* ctx, void* hip_stream);
void* bnsamd_get_stream(bnsamd_context* ctx);
int bnsamd_synchronize(bnsamd_context* ctx);
int bnsamd_version(void);

/* GaussLegendre on [x1, x2] (integration.hpp:28-87): the same Newton
 * iteration and stopping rule, host code, any n >= 1. */
int bnsamd_gauss_legendre(int n, double x1, double x2, double* points, double* weights);

/* Builds (or re-selects from the context's cache) the resident tables for a
 * 1D quadrature (n1 nodes) and a 2D quadrature (n2 nodes) on [0, 1], as
 * ComputeM1Opacities receives quad_1d and quad_2d
 * (m1_opacities.hpp:1306-1313).  1 <= n1, n2 <= 32; the reference caps 2n at
 * BS_N_MAX = 20 (bns_nurates.hpp:68). */
int bnsamd_set_quadrature(bnsamd_context* ctx, int n1, const double* x1, const double* w1,
                          int n2, const double* x2, const double* w2);

/* The GP19 bremsstrahlung table (BremKernelAbsGP19, kernel_brem_GP19.hpp:
 * 35-480, whose GP19Table.hpp data the reference tree does not ship): the
 * response function S(nb, Ye, T, w) on the grid of axes nb [fm^-3] (with its
 * natural logarithms lognb, as the table states them), Ye, T [MeV] and the
 * pair energy w [MeV] (with logw), data[((i n_ye + j) n_T + k) n_w + l]; each
 * axis strictly increasing with at least 2 values.  Copied to the device;
 * brem_implementation = BREM_GP19 then uses it in bnsamd_m1_opacities and
 * bnsamd_spectral_opacities (the 2 and 4 MeV and 48 and 50 MeV temperatures
 * of the extrapolations must lie on the T axis, as in the reference's
 * table).  The electron fraction is the point block's yp (there is no
 * separate ye field; GP19 is the only reader of ye on the path).  n_nb = 0
 * clears the table: BREM_GP19 then adds zero, as the reference's table-free
 * configuration does (kernel_brem_GP19.hpp:483-493). */
int bnsamd_set_gp19_table(bnsamd_context* ctx, int n_nb, const double* nb, const double* lognb,
                          int n_ye, const double* ye, int n_T, const double* T, int n_w,
                          const double* w, const double* logw, const double* data);

/* Gray M1 opacities of npts points, device buffers, asynchronous on the
 * context stream.  Replaces ComputeM1Opacities (stim_abs = 1) and
 * ComputeM1OpacitiesNotStimulated (stim_abs = 0), m1_opacities.hpp:1297-1313,
 * called once per point by the reference's drivers (tests/tests.hpp:756).
 * d_distr may be NULL: the distribution parameters are then reconstructed on
 * the device from (n, J, chi) as CalculateDistrParamsFromM1 does
 * (distribution.hpp:323-333, the callers' step at tests/tests.hpp:731).
 * precision: BNSAMD_FP64, or BNSAMD_MIXED (the distribution values TotalNuF
 * and the HR98 channel functions in FP32; Fermi integrals, the pair, BRT06
 * and NES/NPS kernels, detailed balance, integrands and sums in FP64). */
int bnsamd_m1_opacities(bnsamd_context* ctx, int64_t npts, int64_t ld, const double* d_pts,
                        const double* d_distr, const bnsamd_flags* flags,
                        const bnsamd_params* params, int stim_abs, int precision,
                        double* d_out);

/* Same, from and to host memory (synchronous; stages through the context's
 * device buffers).  ld of the host blocks is npts. */
int bnsamd_m1_opacities_host(bnsamd_context* ctx, int64_t npts, const double* h_pts,
                             const double* h_distr, const bnsamd_flags* flags,
                             const bnsamd_params* params, int stim_abs, int precision,
                             double* h_out);

/* Same as bnsamd_m1_opacities_host, staged through device memory that the
 * library's code object holds (a static array sized for
 * BNSAMD_STATIC_MAX_POINTS points, 9 MB per device, held until the process
 * ends): the call allocates nothing.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Rejects npts above
 * BNSAMD_STATIC_MAX_POINTS with BNSAMD_ERR_INVALID.  Every call on one
 * device shares that device's block (calls on one context are serialised;
 * calls through two contexts on the same device that overlap race).
 * Synchronous; ld of the host blocks is npts. */
#define BNSAMD_STATIC_MAX_POINTS 16384
int bnsamd_m1_opacities_host_static(bnsamd_context* ctx, int64_t npts, const double* h_pts,
                                    const double* h_distr, const bnsamd_flags* flags,
                                    const bnsamd_params* params, int stim_abs, int precision,
                                    double* h_out);

/* Spectral opacities of npts points at nnu neutrino energies each, device
 * buffers, asynchronous on the context stream.  Replaces
 * ComputeSpectralOpacitiesNotStimulatedAbs (stim_abs = 0) and
 * ComputeSpectralOpacitiesStimulatedAbs (stim_abs = 1),
 * m1_opacities.hpp:1505-1660, called once per (energy, point) by the
 * reference's drivers (tests/tests.hpp:1174-1176, mwe.cpp:232,312).  The
 * integrals over the partner energy use the context's 1D quadrature
 * (n1 of bnsamd_set_quadrature, the reference's quad_1d).  Energies [MeV]:
 * d_nu[e] for every point when nu_ld == 0, else d_nu[e * nu_ld + i]
 * (nu_ld >= npts).  d_out: BNSAMD_SP_* fields, 16 * nnu rows of ld doubles.
 * d_distr as for bnsamd_m1_opacities (NULL: reconstructed from n, J, chi). */
int bnsamd_spectral_opacities(bnsamd_context* ctx, int64_t npts, int64_t ld, const double* d_pts,
                              const double* d_distr, int64_t nnu, const double* d_nu,
                              int64_t nu_ld, const bnsamd_flags* flags,
                              const bnsamd_params* params, int stim_abs, double* d_out);

/* Same, with the output as an array of SpectralOpacities structs
 * (bns_nurates.hpp:517-523), one per (point, energy): d_out[npts][nnu][16],
 * value of field f (BNSAMD_SP_*) of energy e at point i at
 * d_out[(i * nnu + e) * 16 + f] -- what the reference's drivers hold when they
 * keep the struct that ComputeSpectralOpacities*Abs returns per (energy,
 * point) (tests/tests.hpp:1174-1176).  Each point's values are one contiguous
 * run, so every cache line of the output is written whole by one wavefront. */
int bnsamd_spectral_opacities_aos(bnsamd_context* ctx, int64_t npts, int64_t ld,
                                  const double* d_pts, const double* d_distr, int64_t nnu,
                                  const double* d_nu, int64_t nu_ld, const bnsamd_flags* flags,
                                  const bnsamd_params* params, int stim_abs, double* d_out);

/* Same as bnsamd_spectral_opacities, from and to host memory (synchronous;
 * ld = npts).  h_nu holds nnu
 * energies shared by all points (per_point_energies = 0) or nnu * npts
 * (per_point_energies = 1, energy e of point i at h_nu[e * npts + i]). */
int bnsamd_spectral_opacities_host(bnsamd_context* ctx, int64_t npts, const double* h_pts,
                                   const double* h_distr, int64_t nnu, const double* h_nu,
                                   int per_point_energies, const bnsamd_flags* flags,
                                   const bnsamd_params* params, int stim_abs, double* h_out);

/* CalculateDistrParamsFromM1 (distribution.hpp:323-333) on the device:
 * d_distr[28][ld] from the n, J, chi fields of d_pts. */
int bnsamd_distr_from_m1(bnsamd_context* ctx, int64_t npts, int64_t ld, const double* d_pts,
                         double* d_distr);

/* NuEquilibriumParams + ComputeM1DensitiesEq (distribution.hpp:253-290,
 * 447-473) on the device: writes the equilibrium distribution into d_distr
 * and the equilibrium n, J into the point block d_pts (chi untouched). */
int bnsamd_equilibrium(bnsamd_context* ctx, int64_t npts, int64_t ld, double* d_pts,
                       double* d_distr);

/* ComputeM1DensitiesEq (distribution.hpp:447-473) on the device: the n, J
 * fields of d_pts from the eta_t fields of d_distr and the temperature. */
int bnsamd_m1_densities_eq(bnsamd_context* ctx, int64_t npts, int64_t ld, const double* d_distr,
                           double* d_pts);

/* Host-memory forms of the three set-up steps above (synchronous, staged
 * through the context's device buffers; ld = npts).  They stand in for the
 * per-point host calls CalculateDistrParamsFromM1 (distribution.hpp:324),
 * NuEquilibriumParams (:254) + ComputeM1DensitiesEq (:448) and
 * ComputeM1DensitiesEq alone, which the reference's drivers make before
 * ComputeM1Opacities (tests/tests.hpp:731, mwe.cpp:192-200,307). */
int bnsamd_distr_from_m1_host(bnsamd_context* ctx, int64_t npts, const double* h_pts,
                              double* h_distr);
int bnsamd_equilibrium_host(bnsamd_context* ctx, int64_t npts, double* h_pts, double* h_distr);
int bnsamd_m1_densities_eq_host(bnsamd_context* ctx, int64_t npts, const double* h_distr,
                                double* h_pts);

/* Complete Fermi-Dirac integral F_k, k = 0..5 (FDI_0..FDI_p5,
 * functions.hpp:1551-2837), elementwise on device arrays. */
int bnsamd_fdi(bnsamd_context* ctx, int k, int64_t n, const double* d_y, double* d_out);

/* ---- serving --------------------------------------------------------------
 * A request server for many concurrent callers (meshblocks of uneven sizes):
 * requests are queued (a large one as pieces of chunk_points points),
 * admitted against a fixed device-memory budget, served smallest request
 * first (a piece waiting beyond starvation_ms goes first), batched with other
 * pending pieces of the same quadrature order, precision, flags and
 * parameters, and run on several streams so that their work overlaps.  A
 * request's results do not depend on its batch or pieces.  No
 * counterpart in the reference, whose callers invoke ComputeM1Opacities per
 * point inside their own loops (tests/tests.hpp:272-274,756); each request is
 * the batched equivalent of such a loop, with the distribution reconstructed
 * from (n, J, chi) as CalculateDistrParamsFromM1 does
 * (distribution.hpp:323-333) and the same quadrature for quad_1d and quad_2d,
 * as all of the reference's callers pass. */
typedef struct bnsamd_server bnsamd_server;

typedef struct bnsamd_server_config {
  int device;               /* HIP device */
  size_t hbm_budget_bytes;  /* device arena for in-flight batches: 328 B per point */
  int streams;              /* batches in flight at once, 1..16 */
  int64_t max_batch_points; /* upper limit of chunk_points */
  int64_t chunk_points;     /* points per launch: a larger request runs as pieces of this
                               many points, interleaved with other requests (0: 32768) */
  double starvation_ms;     /* a piece queued longer than this is served first (0: 200) */
} bnsamd_server_config;

typedef struct bnsamd_server_stats {
  int64_t requests_submitted, requests_completed, points_completed, batches;
  int64_t max_inflight_bytes, max_concurrent_batches, arena_bytes;
} bnsamd_server_stats;

int bnsamd_server_create(const bnsamd_server_config* cfg, bnsamd_server** out);
/* Finishes every submitted request, then releases the server.  No call may
 * be waiting on the server when it is destroyed. */
void bnsamd_server_destroy(bnsamd_server* s);
const char* bnsamd_server_last_error(const bnsamd_server* s);
/* Queues npts points (host SoA block h_pts[21][npts]); the outputs
 * h_out[20][npts] are written when the request completes.  Both buffers must
 * stay valid until bnsamd_server_wait returns for this request; pinned
 * buffers (bnsamd_host_alloc) let the copies run asynchronously.  Rejects a
 * request whose piece is larger than the server's HBM budget. */
int bnsamd_server_submit(bnsamd_server* s, int64_t npts, const double* h_pts, int nquad,
                         const bnsamd_flags* flags, const bnsamd_params* params, int stim_abs,
                         int precision, double* h_out, int64_t* request_id);
int bnsamd_server_poll(bnsamd_server* s, int64_t request_id, int* done);
/* Blocks until the request has completed; returns its status and, if
 * latency_us is not NULL, the time from submission to completion.  The id is
 * released afterwards. */
int bnsamd_server_wait(bnsamd_server* s, int64_t request_id, double* latency_us);
int bnsamd_server_get_stats(bnsamd_server* s, bnsamd_server_stats* out);

/* Pinned (page-locked) host memory for request buffers. */
int bnsamd_host_alloc(size_t bytes, void** ptr);
void bnsamd_host_free(void* ptr);

/* ---- learned form: BF16-MFMA MLP ensemble ------------------------------------
 * No counterpart in the reference.  A small network per ensemble member maps
 * 15 features of a point (log nb, log T, Ye and, for nue, anue, nux, log n,
 * log J, chi; then the degeneracies mu_e / T and (mu_n - mu_p) / T, and
 * 1 / T) to 15 standardised logarithms of the outputs (eta_0, eta,
 * kappa_0_a, kappa_a, kappa_s of nue, anue, nux); it is trained on the GPU by
 * the library's own kernels on data the FP64 path generates, and evaluated by
 * a fused kernel whose points fall back to the FP64 path when the members
 * disagree or the inputs leave the training box. */
typedef struct bnsamd_mlp bnsamd_mlp;

/* feature standardisation x = (f - feat_mean) / feat_std, the training box
 * (raw features), and the output model: log output o =
 * lin[0][o] + sum_i lin[1 + i][o] x_i + out_mean[o] + out_std[o] y_o, where
 * y is the network's output (the network learns the residual of a linear
 * model in the standardised features) */
typedef struct bnsamd_mlp_norm {
  float feat_mean[15], feat_std[15];
  float out_mean[15], out_std[15];
  float box_lo[15], box_hi[15];
  float lin[16][15];
  float res_lo[15], res_hi[15]; /* range of the standardised residual in training */
  float err_a[15], err_k[15];   /* acceptance coefficients: set by bnsamd_mlp_set_acceptance,
                                   left unchanged by bnsamd_mlp_set_norm */
} bnsamd_mlp_norm;

int bnsamd_mlp_create(int device, int members, int max_batch, bnsamd_mlp** out);
void bnsamd_mlp_destroy(bnsamd_mlp* m);
const char* bnsamd_mlp_last_error(const bnsamd_mlp* m);
/* floats per member in the parameter blob (W_0, b_0, ..., W_4, b_4) */
int bnsamd_mlp_param_count(void);
int bnsamd_mlp_set_params(bnsamd_mlp* m, int member, const float* h_params);
int bnsamd_mlp_get_params(bnsamd_mlp* m, int member, float* h_params);
int bnsamd_mlp_set_norm(bnsamd_mlp* m, const bnsamd_mlp_norm* norm);
/* loss and gradient (host blob) of one member on `rows` rows of normalised
 * features d_x [rows][15] and standardised targets d_y [rows][15] */
int bnsamd_mlp_gradients(bnsamd_mlp* m, int member, int rows, const float* d_x, const float* d_y,
                         float* h_grad, float* h_loss);
/* Adam training of one member: `epochs` passes over n rows in the order
 * d_perm, batches of `batch` rows, cosine learning rate lr0 -> lr1; the mean
 * loss of each epoch goes to h_epoch_loss[epochs] (may be NULL) */
int bnsamd_mlp_train(bnsamd_mlp* m, int member, int64_t n, const float* d_x, const float* d_y,
                     const int32_t* d_perm, int epochs, int batch, float lr0, float lr1,
                     float* h_epoch_loss);
/* the same training for every member at once, member e on its own stream and
 * workspace (created on first use), reading its rows in the order
 * d_perm[e * n .. e * n + n); the members' steps are interleaved so that
 * their small matrix products share the device.  Member e's result is the
 * one bnsamd_mlp_train gives with d_perm + e * n; epoch losses go to
 * h_epoch_loss[members][epochs] (may be NULL) */
int bnsamd_mlp_train_ensemble(bnsamd_mlp* m, int64_t n, const float* d_x, const float* d_y,
                              const int32_t* d_perm, int epochs, int batch, float lr0, float lr1,
                              float* h_epoch_loss);
/* standardised outputs d_y [rows][15] of one member (training-path kernels) */
int bnsamd_mlp_predict(bnsamd_mlp* m, int member, int rows, const float* d_x, float* d_y);
/* The configuration the ensemble's training targets were made with: every
 * input of the FP64 path that is not a feature.  A point block evaluated by
 * the learned form must agree with it: the same flags, parameters and
 * stimulated-absorption switch, the context's 1D and 2D quadratures of order
 * nquad, and per point dU and dm_eff equal to the values here, yn = 1 - yp,
 * and the anux moments (n, J, chi) equal to the nux ones (the features carry
 * nux only).  beta_guard: points with |mu_np| < beta_guard * T are left to
 * the FP64 path, mu_np = mu_n - mu_p - dQ - dU the chemical-potential
 * difference of the beta rates (formulas_beta_iso.hpp:161-180): the
 * reference's EtaNNAbs switches branch where mu_np or the sign of
 * (n_n - n_p) / mu_np changes (formulas_beta_iso.hpp:79-96), and the rates
 * jump there. */
typedef struct bnsamd_mlp_domain {
  bnsamd_flags flags;
  bnsamd_params params;
  int stim_abs;
  int nquad;
  double dU, dm_eff;
  double beta_guard;
} bnsamd_mlp_domain;
int bnsamd_mlp_set_domain(bnsamd_mlp* m, const bnsamd_mlp_domain* domain);

/* Acceptance rule of the learned form: the predicted error of log output o
 * (natural-log units) is err_a[o] + err_k[o] * s_o, s_o the standard
 * deviation of the members' log predictions.  Until this is called,
 * err_a = 0 and err_k = 1 (the predicted error is the spread itself).  The
 * coefficients come from a calibration on held-out points with FP64 targets
 * (bns_nurates_amd/surrogate.py calibrate). */
int bnsamd_mlp_set_acceptance(bnsamd_mlp* m, const float* err_a, const float* err_k);

/* Acceptance judge of the learned form: a network of the members' shape,
 * member 0 of `judge`, whose outputs (through judge's own output model,
 * bnsamd_mlp_norm) are the predicted log margins
 * log(|pred_o - ref_o| / (tol |ref_o| + 16 D_o)) of the ensemble's 15
 * outputs at one tolerance (bns_nurates_amd/surrogate.py train_judge).
 * When set, a point is also recomputed by the FP64 path when its largest
 * predicted margin exceeds threshold (outputs the FP64 path sets to zero
 * below its density thresholds excepted; a residual outside the judge's
 * training range counts as +inf).  The judge's parameters and normalisation
 * block are copied: `judge` may be changed or destroyed afterwards.  Its
 * feature standardisation and training box must equal the ensemble's
 * (BNSAMD_ERR_INVALID otherwise).  judge == NULL removes the judge.  A call
 * rejected by these checks changes nothing; one that fails while copying
 * leaves a judge that sends every point to the FP64 path.  No counterpart
 * in the reference. */
int bnsamd_mlp_set_judge(bnsamd_mlp* m, bnsamd_mlp* judge, double threshold);

/* The ensemble's raw predictions on npts points (calibration and
 * diagnostics): d_logmean[15][ld] the mean log output (nue, anue, nux x
 * eta_0, eta, kappa_0_a, kappa_a, kappa_s), d_spread[15][ld] its spread s_o
 * (-1 for an output the FP64 path sets to zero below its density threshold,
 * +inf when the predicted residual leaves its training range), d_outside[ld]
 * 1 where the point is outside the training box or the domain.  Asynchronous
 * on the context's stream. */
int bnsamd_mlp_infer(bnsamd_context* ctx, bnsamd_mlp* m, int64_t npts, int64_t ld,
                     const double* d_pts, double* d_logmean, float* d_spread, int32_t* d_outside);

/* The learned form of bnsamd_m1_opacities (same point and output blocks, on
 * the context's stream): every member of the ensemble evaluates each point.
 * A point is recomputed by the FP64 path when its features leave the
 * training box, it leaves the training domain (bnsamd_mlp_domain), a
 * predicted residual leaves the range seen in training, the predicted
 * error of any output (bnsamd_mlp_set_acceptance) exceeds error_bound
 * (natural-log units: log1p(tol) for a relative tolerance tol), or a judge
 * is set and rejects it (bnsamd_mlp_set_judge); the other
 * points take the ensemble mean.  When flags, params, stim_abs or the
 * context's quadrature orders differ from the training domain, every point
 * is recomputed.  Fails with BNSAMD_ERR_INVALID before
 * bnsamd_mlp_set_domain.  The number of recomputed points goes to
 * *n_fallback.  Synchronous: the outputs are complete on return. */
int bnsamd_surrogate_m1_opacities(bnsamd_context* ctx, bnsamd_mlp* m, int64_t npts, int64_t ld,
                                  const double* d_pts, float error_bound,
                                  const bnsamd_flags* flags, const bnsamd_params* params,
                                  int stim_abs, double* d_out, int64_t* n_fallback);

#ifdef __cplusplus
}
#endif

#endif /* BNS_NURATES_AMD_H_ */
