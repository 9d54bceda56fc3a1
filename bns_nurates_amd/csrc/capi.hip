// capi.hip -- the extern "C" boundary (include/bns_nurates_amd.h): context,
// resident quadrature tables, launches and host staging.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bns_nurates_amd.h"
#include "context.h"
#include "gray_kernel.h"
#include "gray_math.h"
#include "gray_tables.h"
#include "spectral_kernel.h"

namespace {

using bns::TableLayout;

struct Tables {
  TableLayout L;
  double* dev = nullptr;
};

// Staging block of bnsamd_m1_opacities_host_static: point, distribution and
// output blocks of up to BNSAMD_STATIC_MAX_POINTS points, in device memory
// that the code object itself holds (no allocation at any call).  Host code
// reaches it only through the address hipGetSymbolAddress gives per device.
constexpr size_t kScratchDoubles = size_t(BNSAMD_STATIC_MAX_POINTS) *
                                   (BNSAMD_PT_NFIELDS + BNSAMD_DI_NFIELDS + BNSAMD_OUT_NFIELDS);
static __device__ __align__(256) double scratch[kScratchDoubles];

// device address of `scratch` per device, set once per device at context
// creation with that device current
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

hipError_t scratch_address(int device, double** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

}  // namespace

struct bnsamd_context {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  // resident tables keyed by (n1, n2, node and weight bits)
  std::map<std::vector<double>, Tables> cache;
  const Tables* current = nullptr;
  // staging buffers of the host entry point
  double* stage = nullptr;
  size_t stage_doubles = 0;
  // this device's static staging block (bnsamd_m1_opacities_host_static)
  double* scratch = nullptr;
  // GP19 bremsstrahlung table (bnsamd_set_gp19_table): one device block, the
  // bns::Gp19Table that describes it first, then axes and data; gp19 points
  // to that struct (nullptr: no table)
  double* gp19_dev = nullptr;
  const bns::Gp19Table* gp19 = nullptr;
};

namespace {

int fail(bnsamd_context* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(bnsamd_context* c, hipError_t e, const char* what) {
  return fail(c, BNSAMD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// grows the context's staging buffer of the host entry points to n doubles
int ensure_stage(bnsamd_context* c, size_t n, const char* who) {
  if (n <= c->stage_doubles) return BNSAMD_OK;
  if (c->stage) (void)hipFree(c->stage);
  c->stage = nullptr;
  c->stage_doubles = 0;
  if (hipMalloc(&c->stage, n * sizeof(double)) != hipSuccess)
    return fail(c, BNSAMD_ERR_NOMEM, std::string(who) + ": hipMalloc");
  c->stage_doubles = n;
  return BNSAMD_OK;
}

int upload(bnsamd_context* c, double* d, const double* h, size_t n) {
  return hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess ? BNSAMD_OK : BNSAMD_ERR_HIP;
}

// copies n doubles back and waits for the stream
int download(bnsamd_context* c, double* h, const double* d, size_t n, const char* who) {
  hipError_t e = hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? BNSAMD_OK : hip_fail(c, e, who);
}


// ---- small elementwise kernels ------------------------------------------------
template <int K>
__global__ void k_fdi(int64_t n, const double* __restrict__ y, double* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bns::fdi<K>(y[i]);
}

// NuEquilibriumParams + ComputeM1DensitiesEq (distribution.hpp:253-290,
// 447-473), one point per thread
__global__ void k_equilibrium(int64_t n, int64_t ld, double* __restrict__ pts,
                              double* __restrict__ distr) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double T = pts[BNSAMD_PT_TEMP * ld + i];
  double eta_nue =
      (pts[BNSAMD_PT_MU_E * ld + i] - pts[BNSAMD_PT_MU_N * ld + i] + pts[BNSAMD_PT_MU_P * ld + i]) / T;
  double eta[4] = {eta_nue, -eta_nue, 0., 0.};
  double npre = bns::c::FourPiHc3 * (T * T * T), jpre = npre * T;
  for (int s = 0; s < 4; ++s) {
    distr[(BNSAMD_DI_W_T + s) * ld + i] = 1.;
    distr[(BNSAMD_DI_TEMP_T + s) * ld + i] = T;
    distr[(BNSAMD_DI_ETA_T + s) * ld + i] = eta[s];
    distr[(BNSAMD_DI_W_F + s) * ld + i] = 0.;
    distr[(BNSAMD_DI_TEMP_F + s) * ld + i] = T;
    distr[(BNSAMD_DI_C_F + s) * ld + i] = 0.6;
    distr[(BNSAMD_DI_BETA_F + s) * ld + i] = 1.;
    double F[6];
    bns::fdi_bundle<2, 3>(eta[s], F);
    pts[(BNSAMD_PT_N + s) * ld + i] = npre * F[2];
    pts[(BNSAMD_PT_J + s) * ld + i] = jpre * F[3];
  }
}

// ComputeM1DensitiesEq (distribution.hpp:447-473) for a given distribution:
// n, J from eta_t and T, one point per thread
__global__ void k_m1_densities_eq(int64_t n, int64_t ld, const double* __restrict__ distr,
                                  double* __restrict__ pts) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double T = pts[BNSAMD_PT_TEMP * ld + i];
  double npre = bns::c::FourPiHc3 * (T * T * T), jpre = npre * T;
  for (int s = 0; s < 4; ++s) {
    double F[6];
    bns::fdi_bundle<2, 3>(distr[(BNSAMD_DI_ETA_T + s) * ld + i], F);
    pts[(BNSAMD_PT_N + s) * ld + i] = npre * F[2];
    pts[(BNSAMD_PT_J + s) * ld + i] = jpre * F[3];
  }
}

// CalculateDistrParamsFromM1 (distribution.hpp:51-243), one point per thread
__global__ void k_distr_from_m1(int64_t n, int64_t ld, const double* __restrict__ pts,
                                double* __restrict__ distr, double gamma36) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int s = 0; s < 4; ++s) {
    double n0 = pts[(BNSAMD_PT_N + s) * ld + i], J0 = pts[(BNSAMD_PT_J + s) * ld + i];
    double chi = pts[(BNSAMD_PT_CHI + s) * ld + i];
    double nn = n0 > 1e-100 ? n0 : 1e-100;
    double JJ = J0 > 1e-100 ? J0 : 1e-100;
    double r = nn / JJ;
    double y = nn * (r * r * r) / bns::c::FourPiHc3;
    double et;
    if (y < 0.005) {
      double num = y * (y * (y * (y * (y * (y + 0.19926987701997) + 38865.0149220478) +
                   14364.6331099737) + 5750.1878570758) + 1120.71610972194) + 1.60356108438235e-8;
      double den = y * (y * (y * (y * (y * (y + 1.0) + 38840.0743174942) - 99.9009680656931) -
                   171.874844843596) + 75.7101579899442) + 83.0160130941424;
      et = log(num / den);
    } else if (y <= 0.7) {
      double num = y * (y * (y * (y * (y * (41.3836568203438 * y + 32.5515666786612) -
                   157.774993512235) + 66.5726772426253) + 14.4883415579211) +
                   0.315360380575709) + 0.000660414331285249;
      double den = y * (y * (y * (y * (y * (y + 1.8888797407042) - 5.35488690539183) +
                   1.94673781342617) + 0.483128792035557) + 0.0113386564109086) +
                   2.64160073447322e-5;
      et = num / den - 30.;
    } else if (y > 0.7 && y < 0.7901234567745267) {
      double num = y * (y * (y * (y * (y * (3852.81416018959 * y - 5316.18895799799) +
                   1102.91561586553) + 1.54082262710661e-6) + 1732.89925128741) -
                   1769.59868329086) + 586.406885304906;
      double den = y * (y * (y * (y * (y * (y + 255.936658313629) + 9.42360945627147e-5) -
                   81.2467063138386) - 180.100197053091) - 89.0343496217014) +
                   143.849128123195;
      et = exp(num / den);
    } else {
      et = 20.;
    }
    et = et < 20. ? et : 20.;
    double F[6];
    bns::fdi_bundle<2, 3>(et, F);
    double Tnu = JJ / (nn * 3.6);
    distr[(BNSAMD_DI_W_T + s) * ld + i] = 1.5 * (1. - chi);
    distr[(BNSAMD_DI_ETA_T + s) * ld + i] = et;
    distr[(BNSAMD_DI_TEMP_T + s) * ld + i] = F[2] * JJ / (F[3] * nn);
    distr[(BNSAMD_DI_W_F + s) * ld + i] = 0.5 * (bns::mul_rn(3., chi) - 1.);
    distr[(BNSAMD_DI_C_F + s) * ld + i] = 0.6;
    distr[(BNSAMD_DI_TEMP_F + s) * ld + i] = Tnu;
    distr[(BNSAMD_DI_BETA_F + s) * ld + i] = nn / (bns::c::FourPiHc3 * gamma36 * pow(Tnu, 3.6));
  }
}

int check_ctx(bnsamd_context* c) { return c ? BNSAMD_OK : BNSAMD_ERR_INVALID; }

int run_gray(bnsamd_context* c, int64_t npts, int64_t ld, const double* d_pts,
             const double* d_distr, const bnsamd_flags* flags, const bnsamd_params* params,
             int stim_abs, int precision, double* d_out) {
  if (npts < 0 || ld < npts || !d_pts || !d_out || !flags || !params)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities: invalid argument");
  if (stim_abs != 0 && stim_abs != 1)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities: stim_abs must be 0 or 1");
  if (params->brem_implementation < 0 || params->brem_implementation > 2)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities: unknown brem_implementation");
  if (!c->current) return fail(c, BNSAMD_ERR_NO_QUADRATURE, "bnsamd_set_quadrature not called");
  if (npts > int64_t(0x7fffffff))
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities: npts above 2^31 - 1 per call");
  if (npts == 0) return BNSAMD_OK;
  bns::GrayKernelArgs a{};
  a.npts = npts;
  a.ld = ld;
  a.pts = d_pts;
  a.distr = d_distr;
  a.out = d_out;
  a.tab = c->current->dev;
  a.L = c->current->L;
  a.flags = bns::PointFlags{flags->use_abs_em, flags->use_pair, flags->use_brem,
                            flags->use_inelastic_scatt, flags->use_iso};
  a.params = bns::PointParams{params->use_dU, params->use_dm_eff, params->use_WM_ab,
                              params->use_WM_sc, params->use_decay,
                              params->brem_implementation, params->use_NN_medium_corr,
                              params->neglect_blocking};
  a.stim = stim_abs;
  a.gamma36 = bns::gamma_stirling_36();
  a.gp19 = c->gp19;
  hipError_t e;
  if (precision == BNSAMD_FP64) {
    e = bns::launch_gray_m1_f64(a, c->stream);
  } else if (precision == BNSAMD_MIXED) {
    e = bns::launch_gray_m1_mixed(a, c->stream);
  } else {
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities: unknown precision");
  }
  if (e != hipSuccess) return hip_fail(c, e, "gray M1 kernel launch");
  return BNSAMD_OK;
}

int run_spectral(bnsamd_context* c, int64_t npts, int64_t ld, const double* d_pts,
                 const double* d_distr, int64_t nnu, const double* d_nu, int64_t nu_ld,
                 const bnsamd_flags* flags, const bnsamd_params* params, int stim_abs,
                 double* d_out, bool aos = false) {
  if (npts < 0 || ld < npts || nnu < 0 || !d_pts || !d_nu || !d_out || !flags || !params ||
      (nu_ld != 0 && nu_ld < npts))
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_spectral_opacities: invalid argument");
  if (aos && npts > INT64_MAX / BNSAMD_SP_NFIELDS / (nnu > 0 ? nnu : 1))
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_spectral_opacities_aos: output too large");
  if (stim_abs != 0 && stim_abs != 1)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_spectral_opacities: stim_abs must be 0 or 1");
  if (params->brem_implementation < 0 || params->brem_implementation > 2)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_spectral_opacities: unknown brem_implementation");
  if (!c->current) return fail(c, BNSAMD_ERR_NO_QUADRATURE, "bnsamd_set_quadrature not called");
  if (npts == 0 || nnu == 0) return BNSAMD_OK;
  const TableLayout& L = c->current->L;
  bns::SpectralKernelArgs a{};
  a.npts = npts;
  a.ld = ld;
  a.pts = d_pts;
  a.distr = d_distr;
  a.nnu = nnu;
  a.nu = d_nu;
  a.nu_ld = nu_ld;
  a.out = d_out;
  a.aos = aos ? 1 : 0;
  a.q1 = c->current->dev + L.q1;
  a.n = L.n1;
  a.ech = bns::spectral_energies_per_chunk(L.n1);
  a.nchunk = (nnu + a.ech - 1) / a.ech;
  a.cpw = bns::spectral_chunks_per_group(npts, a.nchunk);
  if (npts > int64_t(0x7fffffff) / ((a.nchunk + a.cpw - 1) / a.cpw))
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_spectral_opacities: too many workgroups");
  a.flags = bns::PointFlags{flags->use_abs_em, flags->use_pair, flags->use_brem,
                            flags->use_inelastic_scatt, flags->use_iso};
  a.params = bns::PointParams{params->use_dU, params->use_dm_eff, params->use_WM_ab,
                              params->use_WM_sc, params->use_decay,
                              params->brem_implementation, params->use_NN_medium_corr,
                              params->neglect_blocking};
  a.stim = stim_abs;
  a.gamma36 = bns::gamma_stirling_36();
  a.gp19 = c->gp19;
  hipError_t e = bns::launch_spectral_f64(a, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "spectral kernel launch");
  return BNSAMD_OK;
}

}  // namespace

extern "C" {

int bnsamd_version(void) { return 1; }

int bnsamd_create(int device, bnsamd_context** out) {
  if (!out) return BNSAMD_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BNSAMD_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return BNSAMD_ERR_INVALID;
  auto* c = new bnsamd_context();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = scratch_address(device, &c->scratch);  // with the device current
  if (e != hipSuccess) {
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return BNSAMD_ERR_HIP;
  }
  c->stream = c->own_stream;
  *out = c;
  return BNSAMD_OK;
}

void bnsamd_destroy(bnsamd_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->cache) (void)hipFree(kv.second.dev);
  if (c->stage) (void)hipFree(c->stage);
  if (c->gp19_dev) (void)hipFree(c->gp19_dev);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

const char* bnsamd_last_error(const bnsamd_context* c) { return c ? c->err.c_str() : ""; }

int bnsamd_set_stream(bnsamd_context* c, void* s) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  hipStream_t ns = static_cast<hipStream_t>(s);  // NULL: HIP's default (null) stream
  // buffers the library reuses in stream order (the host entry's staging
  // buffer, the learned form's fallback workspace) must not be in use on two
  // streams at once: finish what the previous stream still has queued
  if (ns != c->stream) {
    (void)hipSetDevice(c->device);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "bnsamd_set_stream: previous stream");
  }
  c->stream = ns;
  return BNSAMD_OK;
}

void* bnsamd_get_stream(bnsamd_context* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int bnsamd_synchronize(bnsamd_context* c) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(c->device);
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? BNSAMD_OK : hip_fail(c, e, "hipStreamSynchronize");
}

int bnsamd_gauss_legendre(int n, double x1, double x2, double* pts, double* w) {
  if (n < 1 || !pts || !w) return BNSAMD_ERR_INVALID;
  // Newton iteration on P_n from cos(pi (i + 3/4) / (n + 1/2)), stop at
  // |dz| <= 1e-10, weights from the derivative of the previous iterate
  // (integration.hpp:28-87)
  const double pi = bns::c::Pi;
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 0; i < m; ++i) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), z1, pp, p1, p2, p3;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 0; j < n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j + 1.0) * z * p2 - j * p3) / (j + 1);
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > 1.0e-10);
    pts[i] = xm - xl * z;
    pts[n - 1 - i] = xm + xl * z;
    w[i] = 2.0 * xl / ((1.0 - z * z) * (pp * pp));
    w[n - 1 - i] = w[i];
  }
  return BNSAMD_OK;
}

int bnsamd_set_quadrature(bnsamd_context* c, int n1, const double* x1, const double* w1, int n2,
                          const double* x2, const double* w2) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  if (n1 < 1 || n2 < 1 || n1 > bns::kMaxQuad || n2 > bns::kMaxQuad || !x1 || !w1 || !x2 || !w2)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_set_quadrature: orders must be in [1, 32]");
  for (int i = 0; i < n2; ++i)
    if (!(x2[i] > 0. && x2[i] < 1.))
      return fail(c, BNSAMD_ERR_INVALID, "bnsamd_set_quadrature: nodes must lie in (0, 1)");
  std::vector<double> key;
  key.push_back(n1);
  key.push_back(n2);
  key.insert(key.end(), x1, x1 + n1);
  key.insert(key.end(), w1, w1 + n1);
  key.insert(key.end(), x2, x2 + n2);
  key.insert(key.end(), w2, w2 + n2);
  auto it = c->cache.find(key);
  if (it != c->cache.end()) {
    c->current = &it->second;
    return BNSAMD_OK;
  }
  (void)hipSetDevice(c->device);
  TableLayout L = bns::gray_table_layout(n1, n2);
  std::vector<double> host(L.total);
  bns::build_gray_tables(n1, x1, w1, n2, x2, w2, host.data());
  Tables t;
  t.L = L;
  hipError_t e = hipMalloc(&t.dev, sizeof(double) * L.total);
  if (e != hipSuccess) return fail(c, BNSAMD_ERR_NOMEM, "bnsamd_set_quadrature: hipMalloc");
  e = hipMemcpyAsync(t.dev, host.data(), sizeof(double) * L.total, hipMemcpyHostToDevice,
                     c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)hipFree(t.dev);
    return hip_fail(c, e, "bnsamd_set_quadrature: upload");
  }
  auto ins = c->cache.emplace(std::move(key), t);
  c->current = &ins.first->second;
  return BNSAMD_OK;
}

int bnsamd_set_gp19_table(bnsamd_context* c, int n_nb, const double* nb, const double* lognb,
                          int n_ye, const double* ye, int n_T, const double* T, int n_w,
                          const double* w, const double* logw, const double* data) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  // the previous table may be read by work still queued on the stream
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_set_gp19_table: stream");
  if (c->gp19_dev) (void)hipFree(c->gp19_dev);
  c->gp19_dev = nullptr;
  c->gp19 = nullptr;
  if (n_nb == 0) return BNSAMD_OK;  // cleared: GP19 adds zero again
  auto increasing = [](int n, const double* a) {
    if (!a || n < 2) return false;
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(a[i]) || (i && !(a[i] > a[i - 1]))) return false;
    return true;
  };
  if (!increasing(n_nb, nb) || !increasing(n_nb, lognb) || !increasing(n_ye, ye) ||
      !increasing(n_T, T) || !increasing(n_w, w) || !increasing(n_w, logw) || !data)
    return fail(c, BNSAMD_ERR_INVALID,
                "bnsamd_set_gp19_table: axes need >= 2 finite, strictly increasing values");
  const size_t nd = size_t(n_nb) * n_ye * n_T * n_w;
  constexpr size_t head = (sizeof(bns::Gp19Table) + sizeof(double) - 1) / sizeof(double);
  const size_t total = head + 2 * size_t(n_nb) + n_ye + n_T + 2 * size_t(n_w) + nd;
  std::vector<double> host(total);
  double* h = host.data();
  const double* parts[7] = {nb, lognb, ye, T, w, logw, data};
  const size_t sizes[7] = {size_t(n_nb), size_t(n_nb), size_t(n_ye), size_t(n_T), size_t(n_w),
                           size_t(n_w), nd};
  size_t off[7];
  size_t o = head;
  for (int k = 0; k < 7; o += sizes[k], ++k) {
    off[k] = o;
    std::memcpy(h + o, parts[k], sizes[k] * sizeof(double));
  }
  double* d = nullptr;
  if (hipMalloc(&d, total * sizeof(double)) != hipSuccess)
    return fail(c, BNSAMD_ERR_NOMEM, "bnsamd_set_gp19_table: hipMalloc");
  const bns::Gp19Table tab{n_nb, n_ye, n_T, n_w, d + off[0], d + off[1], d + off[2], d + off[3],
                           d + off[4], d + off[5], d + off[6]};
  std::memcpy(h, &tab, sizeof(tab));
  e = hipMemcpy(d, h, total * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(c, e, "bnsamd_set_gp19_table: upload");
  }
  c->gp19_dev = d;
  c->gp19 = reinterpret_cast<const bns::Gp19Table*>(d);
  return BNSAMD_OK;
}

int bnsamd_m1_opacities(bnsamd_context* c, int64_t npts, int64_t ld, const double* d_pts,
                        const double* d_distr, const bnsamd_flags* flags,
                        const bnsamd_params* params, int stim_abs, int precision,
                        double* d_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  return run_gray(c, npts, ld, d_pts, d_distr, flags, params, stim_abs, precision, d_out);
}

int bnsamd_m1_opacities_host(bnsamd_context* c, int64_t npts, const double* h_pts,
                             const double* h_distr, const bnsamd_flags* flags,
                             const bnsamd_params* params, int stim_abs, int precision,
                             double* h_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  if (npts < 0 || !h_pts || !h_out)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities_host: invalid argument");
  if (npts == 0) return BNSAMD_OK;
  (void)hipSetDevice(c->device);
  size_t need = size_t(npts) * (BNSAMD_PT_NFIELDS + BNSAMD_OUT_NFIELDS +
                                (h_distr ? BNSAMD_DI_NFIELDS : 0));
  if (int rc = ensure_stage(c, need, "bnsamd_m1_opacities_host")) return rc;
  double* dp = c->stage;
  double* dout = dp + size_t(npts) * BNSAMD_PT_NFIELDS;
  double* dd = h_distr ? dout + size_t(npts) * BNSAMD_OUT_NFIELDS : nullptr;
  hipError_t e = hipMemcpyAsync(dp, h_pts, sizeof(double) * npts * BNSAMD_PT_NFIELDS,
                                hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && dd)
    e = hipMemcpyAsync(dd, h_distr, sizeof(double) * npts * BNSAMD_DI_NFIELDS,
                       hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_m1_opacities_host: upload");
  int rc = run_gray(c, npts, npts, dp, dd, flags, params, stim_abs, precision, dout);
  if (rc != BNSAMD_OK) return rc;
  e = hipMemcpyAsync(h_out, dout, sizeof(double) * npts * BNSAMD_OUT_NFIELDS,
                     hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_m1_opacities_host: download");
  return BNSAMD_OK;
}

int bnsamd_m1_opacities_host_static(bnsamd_context* c, int64_t npts, const double* h_pts,
                                    const double* h_distr, const bnsamd_flags* flags,
                                    const bnsamd_params* params, int stim_abs, int precision,
                                    double* h_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  if (npts < 0 || !h_pts || !h_out)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_opacities_host_static: invalid argument");
  if (npts > BNSAMD_STATIC_MAX_POINTS)
    return fail(c, BNSAMD_ERR_INVALID,
                "bnsamd_m1_opacities_host_static: npts above BNSAMD_STATIC_MAX_POINTS (" +
                    std::to_string(BNSAMD_STATIC_MAX_POINTS) + ")");
  if (npts == 0) return BNSAMD_OK;
  (void)hipSetDevice(c->device);
  // point block, output block, distribution block at npts leading dimension
  double* dp = c->scratch;
  double* dout = dp + size_t(npts) * BNSAMD_PT_NFIELDS;
  double* dd = h_distr ? dout + size_t(npts) * BNSAMD_OUT_NFIELDS : nullptr;
  hipError_t e = hipMemcpyAsync(dp, h_pts, sizeof(double) * npts * BNSAMD_PT_NFIELDS,
                                hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && dd)
    e = hipMemcpyAsync(dd, h_distr, sizeof(double) * npts * BNSAMD_DI_NFIELDS,
                       hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_m1_opacities_host_static: upload");
  int rc = run_gray(c, npts, npts, dp, dd, flags, params, stim_abs, precision, dout);
  if (rc != BNSAMD_OK) return rc;
  e = hipMemcpyAsync(h_out, dout, sizeof(double) * npts * BNSAMD_OUT_NFIELDS,
                     hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_m1_opacities_host_static: download");
  return BNSAMD_OK;
}

int bnsamd_spectral_opacities(bnsamd_context* c, int64_t npts, int64_t ld, const double* d_pts,
                              const double* d_distr, int64_t nnu, const double* d_nu,
                              int64_t nu_ld, const bnsamd_flags* flags,
                              const bnsamd_params* params, int stim_abs, double* d_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  return run_spectral(c, npts, ld, d_pts, d_distr, nnu, d_nu, nu_ld, flags, params, stim_abs,
                      d_out);
}

int bnsamd_spectral_opacities_aos(bnsamd_context* c, int64_t npts, int64_t ld,
                                  const double* d_pts, const double* d_distr, int64_t nnu,
                                  const double* d_nu, int64_t nu_ld, const bnsamd_flags* flags,
                                  const bnsamd_params* params, int stim_abs, double* d_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  return run_spectral(c, npts, ld, d_pts, d_distr, nnu, d_nu, nu_ld, flags, params, stim_abs,
                      d_out, true);
}

int bnsamd_spectral_opacities_host(bnsamd_context* c, int64_t npts, const double* h_pts,
                                   const double* h_distr, int64_t nnu, const double* h_nu,
                                   int per_point_energies, const bnsamd_flags* flags,
                                   const bnsamd_params* params, int stim_abs, double* h_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  if (npts < 0 || nnu < 0 || !h_pts || !h_nu || !h_out ||
      (per_point_energies != 0 && per_point_energies != 1))
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_spectral_opacities_host: invalid argument");
  if (npts == 0 || nnu == 0) return BNSAMD_OK;
  (void)hipSetDevice(c->device);
  const size_t nnu_vals = size_t(nnu) * (per_point_energies ? size_t(npts) : 1);
  const size_t nout = size_t(npts) * size_t(nnu) * BNSAMD_SP_NFIELDS;
  size_t need = size_t(npts) * (BNSAMD_PT_NFIELDS + (h_distr ? BNSAMD_DI_NFIELDS : 0)) +
                nnu_vals + nout;
  if (int rc = ensure_stage(c, need, "bnsamd_spectral_opacities_host")) return rc;
  double* dp = c->stage;
  double* dout = dp + size_t(npts) * BNSAMD_PT_NFIELDS;
  double* dnu = dout + nout;
  double* dd = h_distr ? dnu + nnu_vals : nullptr;
  hipError_t e = hipMemcpyAsync(dp, h_pts, sizeof(double) * npts * BNSAMD_PT_NFIELDS,
                                hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dnu, h_nu, sizeof(double) * nnu_vals, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && dd)
    e = hipMemcpyAsync(dd, h_distr, sizeof(double) * npts * BNSAMD_DI_NFIELDS,
                       hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_spectral_opacities_host: upload");
  int rc = run_spectral(c, npts, npts, dp, dd, nnu, dnu, per_point_energies ? npts : 0, flags,
                        params, stim_abs, dout);
  if (rc != BNSAMD_OK) return rc;
  e = hipMemcpyAsync(h_out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "bnsamd_spectral_opacities_host: download");
  return BNSAMD_OK;
}

int bnsamd_distr_from_m1(bnsamd_context* c, int64_t npts, int64_t ld, const double* d_pts,
                         double* d_distr) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  if (npts < 0 || ld < npts || !d_pts || !d_distr)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_distr_from_m1: invalid argument");
  if (npts == 0) return BNSAMD_OK;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  unsigned blocks = unsigned((npts + 255) / 256);
  hipLaunchKernelGGL(k_distr_from_m1, dim3(blocks), dim3(256), 0, c->stream, npts, ld, d_pts,
                     d_distr, bns::gamma_stirling_36());
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BNSAMD_OK : hip_fail(c, e, "k_distr_from_m1");
}

int bnsamd_equilibrium(bnsamd_context* c, int64_t npts, int64_t ld, double* d_pts,
                       double* d_distr) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  if (npts < 0 || ld < npts || !d_pts || !d_distr)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_equilibrium: invalid argument");
  if (npts == 0) return BNSAMD_OK;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  unsigned blocks = unsigned((npts + 255) / 256);
  hipLaunchKernelGGL(k_equilibrium, dim3(blocks), dim3(256), 0, c->stream, npts, ld, d_pts,
                     d_distr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BNSAMD_OK : hip_fail(c, e, "k_equilibrium");
}

int bnsamd_m1_densities_eq(bnsamd_context* c, int64_t npts, int64_t ld, const double* d_distr,
                           double* d_pts) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  if (npts < 0 || ld < npts || !d_pts || !d_distr)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_m1_densities_eq: invalid argument");
  if (npts == 0) return BNSAMD_OK;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  unsigned blocks = unsigned((npts + 255) / 256);
  hipLaunchKernelGGL(k_m1_densities_eq, dim3(blocks), dim3(256), 0, c->stream, npts, ld,
                     d_distr, d_pts);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BNSAMD_OK : hip_fail(c, e, "k_m1_densities_eq");
}

// ---- host-memory forms of the distribution set-up (synchronous) -------------
// op 0: distr_from_m1 (pts -> distr), 1: equilibrium (pts -> distr, pts n/J),
// 2: m1_densities_eq (distr -> pts n/J)
static int distr_host(bnsamd_context* c, int op, int64_t npts, double* h_pts, double* h_distr,
                      const char* who) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  if (npts < 0 || !h_pts || !h_distr) return fail(c, BNSAMD_ERR_INVALID, std::string(who) +
                                                  ": invalid argument");
  if (npts == 0) return BNSAMD_OK;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  const size_t np = size_t(npts), P = np * BNSAMD_PT_NFIELDS, D = np * BNSAMD_DI_NFIELDS;
  if (int rc = ensure_stage(c, P + D, who)) return rc;
  double* dp = c->stage;
  double* dd = dp + P;
  if (upload(c, dp, h_pts, P) || (op == 2 && upload(c, dd, h_distr, D)))
    return fail(c, BNSAMD_ERR_HIP, std::string(who) + ": upload");
  unsigned blocks = unsigned((npts + 255) / 256);
  if (op == 0)
    hipLaunchKernelGGL(k_distr_from_m1, dim3(blocks), dim3(256), 0, c->stream, npts, npts, dp,
                       dd, bns::gamma_stirling_36());
  else if (op == 1)
    hipLaunchKernelGGL(k_equilibrium, dim3(blocks), dim3(256), 0, c->stream, npts, npts, dp, dd);
  else
    hipLaunchKernelGGL(k_m1_densities_eq, dim3(blocks), dim3(256), 0, c->stream, npts, npts, dd,
                       dp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, who);
  if (op != 2)
    if (int rc = download(c, h_distr, dd, D, who)) return rc;
  if (op != 0)
    if (int rc = download(c, h_pts, dp, P, who)) return rc;
  return BNSAMD_OK;
}

int bnsamd_distr_from_m1_host(bnsamd_context* c, int64_t npts, const double* h_pts,
                              double* h_distr) {
  return distr_host(c, 0, npts, const_cast<double*>(h_pts), h_distr, "bnsamd_distr_from_m1_host");
}

int bnsamd_equilibrium_host(bnsamd_context* c, int64_t npts, double* h_pts, double* h_distr) {
  return distr_host(c, 1, npts, h_pts, h_distr, "bnsamd_equilibrium_host");
}

int bnsamd_m1_densities_eq_host(bnsamd_context* c, int64_t npts, const double* h_distr,
                                double* h_pts) {
  return distr_host(c, 2, npts, h_pts, const_cast<double*>(h_distr),
                    "bnsamd_m1_densities_eq_host");
}

int bnsamd_fdi(bnsamd_context* c, int k, int64_t n, const double* d_y, double* d_out) {
  if (check_ctx(c)) return BNSAMD_ERR_INVALID;
  if (k < 0 || k > 5 || n < 0 || !d_y || !d_out)
    return fail(c, BNSAMD_ERR_INVALID, "bnsamd_fdi: k must be in 0..5");
  if (n == 0) return BNSAMD_OK;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  dim3 grid(unsigned((n + 255) / 256)), block(256);
  switch (k) {
    case 0: hipLaunchKernelGGL(k_fdi<0>, grid, block, 0, c->stream, n, d_y, d_out); break;
    case 1: hipLaunchKernelGGL(k_fdi<1>, grid, block, 0, c->stream, n, d_y, d_out); break;
    case 2: hipLaunchKernelGGL(k_fdi<2>, grid, block, 0, c->stream, n, d_y, d_out); break;
    case 3: hipLaunchKernelGGL(k_fdi<3>, grid, block, 0, c->stream, n, d_y, d_out); break;
    case 4: hipLaunchKernelGGL(k_fdi<4>, grid, block, 0, c->stream, n, d_y, d_out); break;
    default: hipLaunchKernelGGL(k_fdi<5>, grid, block, 0, c->stream, n, d_y, d_out); break;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BNSAMD_OK : hip_fail(c, e, "k_fdi");
}

}  // extern "C"

bool bns::context_quadrature_orders(bnsamd_context* c, int* n1, int* n2) {
  if (!c) return false;
  std::lock_guard<std::mutex> g(c->mu);
  if (!c->current) return false;
  *n1 = c->current->L.n1;
  *n2 = c->current->L.n2;
  return true;
}
