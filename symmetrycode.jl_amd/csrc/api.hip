// C ABI of libsymcode_hip.so (declared in include/symcode.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include "../../include/symcode.h"
#include "ops.hpp"
#include "pmlp_link.hpp"
#include "cnn_link.hpp"

struct sc_ctx : sc::Ctx {};

namespace {
thread_local std::string g_create_err;

// Scratch of sc_energy_noalloc: one copy per GPU, part of this code object
// (no allocation), kept until the process ends.  Host code never uses the
// array's name as an address: its device address on each device comes from
// hipGetSymbolAddress, taken once per device (with that device current) by
// the first sc_create there.  Calls that overlap on one GPU share the array
// and race.
static __device__ __align__(256) double scratch[sc::NOALLOC_DOUBLES];
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

void init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw sc::Error(1, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  SC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_dev[device] = static_cast<double*>(p);
}

template <class F>
int guard(sc_ctx* c, F&& f) {
  if (!c) return 1;
  try {
    SC_HIP(hipSetDevice(c->device));
    f();
    if (c->sync_calls) SC_HIP(hipStreamSynchronize(c->stream));
    return 0;
  } catch (sc::Error& e) {
    c->err = e.what();
    return e.code;
  } catch (std::exception& e) {
    c->err = e.what();
    return 9;
  }
}

// run f with a typed zero of the context precision
template <class F>
void by_prec(sc::Ctx& c, F&& f) {
  if (c.prec == 64) f(double(0));
  else f(float(0));
}

void need(bool ok, const char* msg) {
  if (!ok) throw sc::Error(1, msg);
}
}  // namespace

#define TYPED(c, ...)                                    \
  by_prec(*(c), [&](auto z_) {                           \
    using T = decltype(z_);                              \
    using CT = sc::C<T>;                                 \
    (void)sizeof(CT);                                    \
    __VA_ARGS__;                                         \
  })

// elements between the spectral row pitch nhp and n / 2 + 1 rounded up to
// 128 bytes
#ifndef SC_NHP_PAD
#define SC_NHP_PAD 0
#endif

extern "C" {

int sc_create(int D, int n, double l, int precision, int device, sc_ctx** out) {
  if (!out) return 1;
  *out = nullptr;
  if (D != 2 && D != 3) { g_create_err = "D must be 2 or 3"; return 1; }
  if (n <= 0 || n % 2 != 0) { g_create_err = "Only even number of grid points supported."; return 1; }
  if (precision != 32 && precision != 64) { g_create_err = "precision must be 32 or 64"; return 1; }
  if (!(l > 0)) { g_create_err = "l must be positive"; return 1; }
  sc_ctx* c = new sc_ctx();
  c->D = D; c->n = n; c->l = l; c->prec = precision; c->device = device;
  c->nh = n / 2 + 1;
  const int align = 128 / (precision == 64 ? 16 : 8);
  c->nhp = (c->nh + align - 1) / align * align + SC_NHP_PAD;
  c->nz = D == 3 ? n : 1;
  c->spec_elems = long(c->nhp) * n * c->nz;
  c->phys_elems = long(n) * n * c->nz;
  try {
    SC_HIP(hipSetDevice(device));
    init_scratch(device);
    SC_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
    SC_HIP(hipStreamCreate(&c->stream));
    c->own_stream = true;
    bool ok = precision == 64 ? sc::Ops<double>::supported(n) : sc::Ops<float>::supported(n);
    if (ok) {
      if (precision == 64) sc::Ops<double>::init_tables(*c);
      else sc::Ops<float>::init_tables(*c);
    }
  } catch (std::exception& e) {
    g_create_err = e.what();
    delete c;
    return 3;
  }
  *out = c;
  return 0;
}

int sc_destroy(sc_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->tw_n, c->tw_m, c->dscr, c->vmax, c->forceA, c->dtdev};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (void* b : {c->bufA, c->bufB, c->bufT, c->scr, c->ustart, c->du, c->acc, c->ust}) c->dfree(b);
  for (void* b : c->ws) c->dfree(b);
  for (void* b : c->adjw) c->dfree(b);
  if (c->bandadj) (void)hipFree(c->bandadj);
  if (c->dtlog) (void)hipFree(c->dtlog);
  if (c->hbuf) (void)hipHostFree(c->hbuf);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char* sc_last_error(const sc_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int sc_info(const sc_ctx* c, int64_t* info) {
  if (!c || !info) return 1;
  info[0] = c->D; info[1] = c->n; info[2] = c->nh; info[3] = c->nhp; info[4] = c->nz;
  info[5] = c->spec_elems; info[6] = c->phys_elems; info[7] = c->prec;
  return 0;
}

int sc_set_stream(sc_ctx* c, void* s) {
  return guard(c, [&] {
    SC_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream) SC_HIP(hipStreamDestroy(c->stream));
    c->own_stream = false;
    c->stream = (hipStream_t)s;
  });
}
int sc_set_sync(sc_ctx* c, int s) { if (!c) return 1; c->sync_calls = s != 0; return 0; }
int sc_sync(sc_ctx* c) { return guard(c, [&] { SC_HIP(hipStreamSynchronize(c->stream)); }); }

int sc_profile(sc_ctx* c, int enable) {
  return guard(c, [&] {
    SC_HIP(hipStreamSynchronize(c->stream));
    c->prof.clear();
    c->ev_used = 0;
    c->prof_on = enable != 0;
  });
}
int sc_profile_report(sc_ctx* c, char* buf, int buflen) {
  return guard(c, [&] {
    need(buf && buflen > 0, "buffer");
    SC_HIP(hipStreamSynchronize(c->stream));
    std::vector<std::string> names;
    std::vector<double> tot;
    std::vector<long> cnt;
    for (auto& r : c->prof) {
      float ms = 0;
      SC_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      size_t i = 0;
      while (i < names.size() && names[i] != r.name) ++i;
      if (i == names.size()) { names.push_back(r.name); tot.push_back(0); cnt.push_back(0); }
      tot[i] += ms;
      cnt[i] += 1;
    }
    std::string out;
    for (size_t i = 0; i < names.size(); ++i)
      out += names[i] + " " + std::to_string(cnt[i]) + " " + std::to_string(tot[i]) + "\n";
    std::strncpy(buf, out.c_str(), size_t(buflen) - 1);
    buf[buflen - 1] = 0;
  });
}

// Host <-> device copies go through the context's pinned staging buffer in
// chunks of at most 32 MB: the runtime is never handed pageable memory, so
// every transfer is a DMA ordered on the context's stream after the work
// queued before it (a torch allocation's zero fill, a kernel), whatever path
// the runtime takes for pageable copies.  rows x width bytes, pitches in bytes.
static void staged_2d(sc_ctx* c, char* dst, size_t dpitch, const char* src, size_t spitch, size_t width,
                      size_t rows, bool h2d) {
  const size_t chunk = size_t(32) << 20;
  const size_t R = std::max<size_t>(1, chunk / width);
  char* pin = reinterpret_cast<char*>(c->hostbuf(long((std::min(R, rows) * width + 7) / 8)));
  for (size_t r0 = 0; r0 < rows; r0 += R) {
    const size_t nr = std::min(R, rows - r0);
    if (h2d) {
      for (size_t r = 0; r < nr; ++r) std::memcpy(pin + r * width, src + (r0 + r) * spitch, width);
      SC_HIP(hipMemcpy2DAsync(dst + r0 * dpitch, dpitch, pin, width, width, nr, hipMemcpyHostToDevice, c->stream));
      SC_HIP(hipStreamSynchronize(c->stream));
    } else {
      SC_HIP(hipMemcpy2DAsync(pin, width, src + r0 * spitch, spitch, width, nr, hipMemcpyDeviceToHost, c->stream));
      SC_HIP(hipStreamSynchronize(c->stream));
      for (size_t r = 0; r < nr; ++r) std::memcpy(dst + (r0 + r) * dpitch, pin + r * width, width);
    }
  }
}

int sc_upload_spectral(sc_ctx* c, void* dev, const void* host, int ncomp) {
  return guard(c, [&] {
    need(dev && host && ncomp > 0, "null pointer or ncomp");
    const size_t e = c->celem();
    const size_t rows = size_t(c->n) * c->nz;
    for (int k = 0; k < ncomp; ++k) {
      char* d = (char*)dev + size_t(k) * c->spec_elems * e;
      const char* h = (const char*)host + size_t(k) * c->nh * rows * e;
      staged_2d(c, d, c->nhp * e, h, c->nh * e, c->nh * e, rows, true);
    }
  });
}
int sc_download_spectral(sc_ctx* c, void* host, const void* dev, int ncomp) {
  return guard(c, [&] {
    need(dev && host && ncomp > 0, "null pointer or ncomp");
    const size_t e = c->celem();
    const size_t rows = size_t(c->n) * c->nz;
    for (int k = 0; k < ncomp; ++k) {
      const char* d = (const char*)dev + size_t(k) * c->spec_elems * e;
      char* h = (char*)host + size_t(k) * c->nh * rows * e;
      staged_2d(c, h, c->nh * e, d, c->nhp * e, c->nh * e, rows, false);
    }
  });
}
int sc_upload_physical(sc_ctx* c, void* dev, const void* host, int ncomp) {
  return guard(c, [&] {
    need(dev && host && ncomp > 0, "null pointer or ncomp");
    const size_t row = size_t(c->n) * c->relem(), rows = size_t(ncomp) * c->phys_elems / c->n;
    staged_2d(c, (char*)dev, row, (const char*)host, row, row, rows, true);
  });
}
int sc_download_physical(sc_ctx* c, void* host, const void* dev, int ncomp) {
  return guard(c, [&] {
    need(dev && host && ncomp > 0, "null pointer or ncomp");
    const size_t row = size_t(c->n) * c->relem(), rows = size_t(ncomp) * c->phys_elems / c->n;
    staged_2d(c, (char*)host, row, (const char*)dev, row, row, rows, false);
  });
}

static void need_fft(sc_ctx* c) {
  if (!c->tw_n) throw sc::Error(2, "no FFT plan for n = " + std::to_string(c->n));
}
static double nD(const sc_ctx* c) { return std::pow(double(c->n), c->D); }

int sc_to_phys(sc_ctx* c, void* phys, const void* spec, int ncomp) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::c2r(*c, (T*)phys, (const CT*)spec, ncomp, 1.0, false)); });
}
int sc_to_spec(sc_ctx* c, void* spec, const void* phys, int ncomp) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::r2c(*c, (CT*)spec, (const T*)phys, ncomp, 1.0 / nD(c))); });
}
int sc_irfft(sc_ctx* c, void* phys, const void* spec, int ncomp) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::c2r(*c, (T*)phys, (const CT*)spec, ncomp, 1.0 / nD(c), false)); });
}
int sc_rfft(sc_ctx* c, void* spec, const void* phys, int ncomp) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::r2c(*c, (CT*)spec, (const T*)phys, ncomp, 1.0)); });
}

int sc_twothirds(sc_ctx* c, void* u, int ncomp) { return guard(c, [&] { TYPED(c, sc::Ops<T>::twothirds(*c, (CT*)u, ncomp)); }); }
int sc_project(sc_ctx* c, void* u) { return guard(c, [&] { TYPED(c, sc::Ops<T>::project(*c, (CT*)u)); }); }
int sc_symmetrize(sc_ctx* c, void* u, int ncomp) { return guard(c, [&] { TYPED(c, sc::Ops<T>::symmetrize(*c, (CT*)u, ncomp)); }); }
int sc_viscosity(sc_ctx* c, void* s, const void* u, double nu) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::viscosity(*c, (CT*)s, (const CT*)u, nu)); });
}
int sc_tensordivergence(sc_ctx* c, void* du, const void* s) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::tensordivergence(*c, (CT*)du, (const CT*)s)); });
}
int sc_vectordivergence(sc_ctx* c, void* d, const void* u) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::vectordivergence(*c, (CT*)d, (const CT*)u)); });
}
int sc_vectorgradient(sc_ctx* c, void* G, const void* u) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::vectorgradient(*c, (CT*)G, (const CT*)u)); });
}
int sc_strainrate(sc_ctx* c, void* S, const void* u) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::strainrate(*c, (CT*)S, (const CT*)u)); });
}
int sc_gaussianfilter(sc_ctx* c, void* u, int ncomp, double delta) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::gaussian(*c, (CT*)u, ncomp, delta)); });
}
int sc_cutoff(sc_ctx* les, sc_ctx* dns, void* ub, const void* u, int ncomp) {
  return guard(les, [&] {
    need(dns && dns->D == les->D && dns->prec == les->prec && dns->n >= les->n, "incompatible contexts");
    TYPED(les, sc::Ops<T>::cutoff(*les, *dns, (CT*)ub, (const CT*)u, ncomp));
  });
}
int sc_inverse_cutoff(sc_ctx* les, sc_ctx* dns, void* u, const void* ub, int ncomp) {
  return guard(les, [&] {
    need(dns && dns->D == les->D && dns->prec == les->prec && dns->n >= les->n, "incompatible contexts");
    TYPED(les, sc::Ops<T>::inverse_cutoff(*les, *dns, (CT*)u, (const CT*)ub, ncomp));
  });
}
int sc_axpy(sc_ctx* c, void* y, const void* x, double a, int ncomp) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::axpy(*c, (CT*)y, (const CT*)x, a, ncomp)); });
}
int sc_make_tracefree_spectral(sc_ctx* c, void* tau) { return guard(c, [&] { TYPED(c, sc::Ops<T>::spec_tracefree(*c, (CT*)tau)); }); }
int sc_make_tracefree_physical(sc_ctx* c, void* tau) { return guard(c, [&] { TYPED(c, sc::Ops<T>::phys_tracefree(*c, (T*)tau)); }); }

int sc_nonlinearity(sc_ctx* c, void* s, const void* u) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::nonlinearity(*c, (CT*)s, (const CT*)u, nullptr)); });
}
int sc_stress(sc_ctx* c, void* s, const void* u, double nu) {
  return guard(c, [&] {
    need_fft(c);
    TYPED(c, sc::Ops<T>::nonlinearity(*c, (CT*)s, (const CT*)u, nullptr); sc::Ops<T>::viscosity(*c, (CT*)s, (const CT*)u, nu));
  });
}
int sc_convectiondiffusion(sc_ctx* c, void* du, const void* u, double nu) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::rhs(*c, (CT*)du, (const CT*)u, nu)); });
}
int sc_wray3(sc_ctx* c, void* u, double dt, double nu) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::wray3(*c, (CT*)u, dt, nu)); });
}
int sc_rk4(sc_ctx* c, void* u, double dt, double nu) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::rk4(*c, (CT*)u, dt, nu)); });
}
int sc_rk4_cfl(sc_ctx* c, void* u, double cfl, double nu, double dtmax, double* dt) {
  return guard(c, [&] {
    need_fft(c);
    need(cfl > 0 && dtmax > 0, "cfl and dtmax must be positive");
    double d = 0;
    TYPED(c, d = sc::Ops<T>::rk4_dev(*c, (CT*)u, dtmax, nu, cfl));
    if (dt) *dt = d;
  });
}
int sc_rk4_steps(sc_ctx* c, void* u, int nsteps, double cfl, double nu, double dtmax, double* dts) {
  return guard(c, [&] {
    need_fft(c);
    need(nsteps >= 0 && dtmax > 0, "nsteps must be >= 0 and dtmax positive");
    TYPED(c, sc::Ops<T>::rk4_steps(*c, (CT*)u, nsteps, cfl, nu, dtmax, dts));
  });
}
int sc_wray3_cfl(sc_ctx* c, void* u, double cfl, double nu, double dtmax, double* dt) {
  return guard(c, [&] {
    need_fft(c);
    need(cfl > 0 && dtmax > 0, "cfl and dtmax must be positive");
    double d = 0;
    TYPED(c, d = sc::Ops<T>::wray3_dev(*c, (CT*)u, dtmax, nu, cfl, -1, 0, 0));
    if (dt) *dt = d;
  });
}
int sc_les_wray3_cfl(sc_ctx* c, void* u, double cfl, double nu, double dtmax, int closure, double p1, double p2,
                     double* dt) {
  return guard(c, [&] {
    need_fft(c);
    need(cfl > 0 && dtmax > 0, "cfl and dtmax must be positive");
    need(closure >= 0 && closure <= 7, "closure kind");
    double d = 0;
    TYPED(c, d = sc::Ops<T>::wray3_dev(*c, (CT*)u, dtmax, nu, cfl, closure, p1, p2));
    if (dt) *dt = d;
  });
}
int sc_les_rk4(sc_ctx* c, void* u, double dt, double nu, int closure, double p1, double p2) {
  return guard(c, [&] {
    need_fft(c);
    need(closure >= 0 && closure <= 7, "closure kind");
    TYPED(c, sc::Ops<T>::rk4_dev(*c, (CT*)u, dt, nu, 0.0, closure, p1, p2));
  });
}
int sc_les_rk4_cfl(sc_ctx* c, void* u, double cfl, double nu, double dtmax, int closure, double p1, double p2,
                   double* dt) {
  return guard(c, [&] {
    need_fft(c);
    need(cfl > 0 && dtmax > 0, "cfl and dtmax must be positive");
    need(closure >= 0 && closure <= 7, "closure kind");
    double d = 0;
    TYPED(c, d = sc::Ops<T>::rk4_dev(*c, (CT*)u, dtmax, nu, cfl, closure, p1, p2));
    if (dt) *dt = d;
  });
}
int sc_propose_timestep(sc_ctx* c, const void* u, double nu, double* dt) {
  return guard(c, [&] {
    need_fft(c);
    need(dt != nullptr, "null dt");
    double vmax = 0;
    TYPED(c, vmax = sc::Ops<T>::max_velocity(*c, (const CT*)u));
    const double h = c->l / c->n;
    const double dconv = h / vmax;
    const double ddiff = h * h / c->D / 2 / nu;
    *dt = std::fmin(dconv, ddiff);
  });
}
int sc_set_band_forcing(sc_ctx* c, double injection, int kband) {
  if (!c) return 1;
  if (injection <= 0) { c->force_kind = 0; return 0; }
  if (kband < 1 || kband >= c->n / 2) { c->err = "kband out of range"; return 1; }
  c->force_kind = 1;
  c->injection = injection;
  c->kband = kband;
  return 0;
}

int sc_energy(sc_ctx* c, const void* u, int ncomp, double* e) {
  return guard(c, [&] { TYPED(c, *e = sc::Ops<T>::sum_modes(*c, (const CT*)u, ncomp, 0, 0.0) / 2); });
}
int sc_energy_noalloc(sc_ctx* c, const void* u, int ncomp, double* e) {
  return guard(c, [&] {
    need(u && e && ncomp > 0, "null pointer or ncomp");
    double* scr = g_scratch_dev[c->device];
    need(scr != nullptr, "no scratch on this device");
    double s = 0;
    TYPED(c, sc::Ops<T>::sum_modes_noalloc(*c, (const CT*)u, ncomp, 0, 0.0, scr, &s));
    *e = s / 2;
  });
}
int sc_spectraldot(sc_ctx* c, const void* u, const void* v, int ncomp, double* re, double* im) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::dot_modes(*c, (const CT*)u, (const CT*)v, ncomp, re, im)); });
}
int sc_norm2(sc_ctx* c, const void* u, int ncomp, double* s) {
  return guard(c, [&] { TYPED(c, *s = sc::Ops<T>::sum_modes(*c, (const CT*)u, ncomp, 3, 0.0)); });
}
int sc_dissipation(sc_ctx* c, const void* u, double nu, double* eps) {
  return guard(c, [&] { TYPED(c, *eps = sc::Ops<T>::sum_modes(*c, (const CT*)u, c->D, 1, nu)); });
}
int sc_filter_reynolds(sc_ctx* c, const void* ubar, double visc, double delta, double* re) {
  return guard(c, [&] {
    need(visc > 0, "visc must be positive");
    double a2 = 0;
    TYPED(c, a2 = sc::Ops<T>::sum_modes(*c, (const CT*)ubar, c->D, 4, 1.0));
    *re = delta * delta * std::sqrt(a2) / visc;
  });
}
int sc_palinstrophy(sc_ctx* c, const void* u, double nu, double* p) {
  return guard(c, [&] { TYPED(c, *p = sc::Ops<T>::sum_modes(*c, (const CT*)u, c->D, 2, nu)); });
}
int sc_spectrum(sc_ctx* c, const void* u, int ncomp, int nshell, double* s) {
  return guard(c, [&] {
    need(nshell > 0 && s, "nshell / output");
    std::vector<double> h(nshell + 1);
    TYPED(c, sc::Ops<T>::shell_energy(*c, (const CT*)u, ncomp, nshell + 1, h.data()));
    for (int i = 0; i < nshell; ++i) s[i] = h[i + 1];
  });
}
int sc_shell_energies(sc_ctx* c, const void* u, int ncomp, int nshell, double* out) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::shell_energy(*c, (const CT*)u, ncomp, nshell, out)); });
}
int sc_shell_scale(sc_ctx* c, void* u, int ncomp, int nshell, const double* f) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::shell_scale(*c, (CT*)u, ncomp, nshell, f)); });
}
int sc_maintain_shell_energy(sc_ctx* c, void* u, int nshell, const double* eref) {
  return guard(c, [&] {
    need(nshell > 0 && eref, "nshell / eref");
    std::vector<double> e(nshell + 1), f(nshell + 1, 1.0);
    TYPED(c, sc::Ops<T>::shell_energy(*c, (const CT*)u, c->D, nshell + 1, e.data()));
    for (int i = 1; i <= nshell; ++i) f[i] = std::sqrt(eref[i - 1] / e[i]);
    TYPED(c, sc::Ops<T>::shell_scale(*c, (CT*)u, c->D, nshell + 1, f.data()));
  });
}

int sc_randn(sc_ctx* c, void* u, int ncomp, uint64_t seed) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::randn(*c, (CT*)u, ncomp, seed)); });
}
int sc_randomfield(sc_ctx* c, void* u, uint64_t seed, int nshell, const double* E0) {
  return guard(c, [&] {
    need(nshell > 0 && E0, "nshell / E0");
    TYPED(c, {
      CT* uu = (CT*)u;
      sc::Ops<T>::randn(*c, uu, c->D, seed);
      sc::Ops<T>::symmetrize(*c, uu, c->D);
      sc::Ops<T>::project(*c, uu);
      sc::Ops<T>::twothirds(*c, uu, c->D);
      std::vector<double> e(nshell), f(nshell);
      sc::Ops<T>::shell_energy(*c, uu, c->D, nshell, e.data());
      for (int k = 0; k < nshell; ++k) f[k] = e[k] > 0 ? std::sqrt(E0[k] / e[k]) : 0.0;
      sc::Ops<T>::shell_scale(*c, uu, c->D, nshell, f.data());
      sc::Ops<T>::twothirds(*c, uu, c->D);
    });
  });
}
int sc_taylorgreen(sc_ctx* c, void* u, double V0, int doproject) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::taylorgreen(*c, (CT*)u, V0, doproject)); });
}

int sc_gradient_phys(sc_ctx* c, void* GG, const void* u) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::gradient_phys(*c, (T*)GG, (const CT*)u)); });
}
int sc_closure_clark(sc_ctx* c, void* tau, const void* GG, double delta) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::clark(*c, (T*)tau, (const T*)GG, delta)); });
}
int sc_closure_smagorinsky(sc_ctx* c, void* tau, const void* GG, double cs, double delta) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::smagorinsky(*c, (T*)tau, (const T*)GG, cs, delta)); });
}
int sc_sfs_dissipation(sc_ctx* c, void* eps, const void* tau, const void* GG) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::sfs_dissipation(*c, (T*)eps, (const T*)tau, (const T*)GG)); });
}
int sc_shell_transfer(sc_ctx* c, const void* u, const void* tau, int nshell, double* out) {
  return guard(c, [&] {
    need(nshell > 0 && out, "nshell / output");
    std::vector<double> h(nshell + 1);
    TYPED(c, sc::Ops<T>::shell_transfer(*c, (const CT*)u, (const T*)tau, nshell + 1, h.data()));
    for (int i = 0; i < nshell; ++i) out[i] = h[i + 1];
  });
}
int sc_sample_mean(sc_ctx* c, const void* x, long n, double* mean) {
  return guard(c, [&] {
    need(mean != nullptr && n > 0, "output / samples");
    TYPED(c, *mean = sc::Ops<T>::sample_mean(*c, (const T*)x, n));
  });
}
int sc_moments(sc_ctx* c, const void* x, long n, double* out) {
  return guard(c, [&] {
    need(out != nullptr, "output");
    TYPED(c, sc::Ops<T>::sample_moments(*c, (const T*)x, n, out));
  });
}
int sc_apriori_metrics(sc_ctx* c, const void* model, const void* ref, long n, double* out) {
  return guard(c, [&] {
    need(out != nullptr, "output");
    TYPED(c, sc::Ops<T>::apriori(*c, (const T*)model, (const T*)ref, n, out));
  });
}
int sc_closure_verstappen(sc_ctx* c, void* tau, const void* GG, double cq, double delta) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::verstappen(*c, (T*)tau, (const T*)GG, cq, delta)); });
}
int sc_closure_bardina(sc_ctx* c, void* tau, const void* u, double s, double delta) {
  return guard(c, [&] { TYPED(c, sc::Ops<T>::bardina(*c, (T*)tau, (const CT*)u, s, delta)); });
}
int sc_closure_dynsmag(sc_ctx* c, void* tau, const void* u, double delta, double* coeff) {
  return guard(c, [&] {
    need_fft(c);
    double cc = 0;
    TYPED(c, cc = sc::Ops<T>::dynsmag(*c, (T*)tau, (const CT*)u, delta));
    if (coeff) *coeff = cc;
  });
}
int sc_les_rhs(sc_ctx* c, void* du, const void* u, double nu, int closure, double p1, double p2) {
  return guard(c, [&] {
    need_fft(c);
    need(closure >= 0 && closure <= 7, "closure kind");
    TYPED(c, sc::Ops<T>::les_rhs(*c, (CT*)du, (const CT*)u, nu, closure, p1, p2));
  });
}
int sc_les_wray3(sc_ctx* c, void* u, double dt, double nu, int closure, double p1, double p2) {
  return guard(c, [&] {
    need_fft(c);
    need(closure >= 0 && closure <= 7, "closure kind");
    TYPED(c, sc::Ops<T>::les_wray3(*c, (CT*)u, dt, nu, closure, p1, p2));
  });
}

int sc_rhs_vjp(sc_ctx* c, void* out, const void* u, const void* w, double nu) {
  return guard(c, [&] { need_fft(c); TYPED(c, sc::Ops<T>::rhs_vjp(*c, (CT*)out, (const CT*)u, (const CT*)w, nu)); });
}
int sc_rk4_vjp(sc_ctx* c, void* out, const void* u0, const void* ubar, double dt, double nu) {
  return guard(c, [&] {
    need_fft(c);
    need(out != u0 && out != ubar, "out must not alias u0 or ubar");
    TYPED(c, sc::Ops<T>::rk4_vjp(*c, (CT*)out, (const CT*)u0, (const CT*)ubar, dt, nu));
  });
}
int sc_wray3_vjp(sc_ctx* c, void* out, const void* u0, const void* ubar, double dt, double nu) {
  return guard(c, [&] {
    need_fft(c);
    need(out != u0 && out != ubar, "out must not alias u0 or ubar");
    TYPED(c, sc::Ops<T>::wray3_vjp(*c, (CT*)out, (const CT*)u0, (const CT*)ubar, dt, nu));
  });
}

int sc_set_cnn_closure(sc_ctx* c, sc_cnn* net, double s_in, double s_out, double extra) {
  return guard(c, [&] {
    if (!net) {
      c->cnn = nullptr;
      return;
    }
    need(c->D == 3, "the learned closure is three-dimensional");
    need(cnnlink::grid_n(net) == c->n, "network grid size differs from the context's");
    need(cnnlink::out_channels(net) == 6, "the network must output the 6 packed stress components");
    need(cnnlink::in_channels(net) == 3 || cnnlink::in_channels(net) == 4,
         "the network input is the 3 velocity components (+ 1 constant channel)");
    need(s_in > 0 && s_out != 0, "scales: s_in > 0, s_out != 0");
    c->cnn = net;
    c->cnn_sin = s_in;
    c->cnn_sout = s_out;
    c->cnn_extra = extra;
  });
}
int sc_set_pmlp_closure(sc_ctx* c, sc_pmlp* net, double delta, double visc, double redelta_mu,
                        double redelta_sigma) {
  return guard(c, [&] {
    if (net) {
      need(pmlplink::dim(net) == c->D, "the network's dimension differs from the grid's");
      need(delta > 0, "delta > 0");
      if (pmlplink::uses_redelta(net))
        need(visc > 0 && redelta_sigma > 0, "a Re_Delta input needs visc > 0 and redelta_sigma > 0");
    }
    c->pmlp = net;
    c->pmlp_sym = 0;
    c->pmlp_delta = delta;
    c->pmlp_visc = visc;
    c->pmlp_mu = redelta_mu;
    c->pmlp_sigma = redelta_sigma;
  });
}
int sc_set_pmlp_symmetrize(sc_ctx* c, int on) {
  return guard(c, [&] {
    need(!on || (c->pmlp && pmlplink::arch(c->pmlp) == 0), "the group average needs an attached mlp network");
    c->pmlp_sym = on ? 1 : 0;
  });
}
int sc_closure_pmlp(sc_ctx* c, void* tau, const void* u) {
  return guard(c, [&] {
    need(c->pmlp != nullptr, "no network attached (sc_set_pmlp_closure)");
    need_fft(c);
    const int ng = c->D * c->D;
    TYPED(c, {
      T* GG = reinterpret_cast<T*>(c->work(1, size_t(ng) * c->phys_elems * sizeof(T)));
      sc::Ops<T>::gradient_phys(*c, GG, (const CT*)u);
      pmlplink::closure_tau<T>(c->pmlp, GG, c->phys_elems, c->pmlp_delta, c->pmlp_visc, c->pmlp_mu,
                               c->pmlp_sigma, c->pmlp_sym, (T*)tau, c->stream);
    });
  });
}
int sc_closure_cnn(sc_ctx* c, void* tau, const void* u) {
  return guard(c, [&] {
    need_fft(c);
    TYPED(c, sc::Ops<T>::cnn_tau(*c, (T*)tau, (const CT*)u));
  });
}
int sc_les_rhs_vjp(sc_ctx* c, void* out, const void* u, const void* w, double nu, int closure) {
  return guard(c, [&] {
    need_fft(c);
    need(closure == 0 || closure == 4, "the transpose exists for closure kinds 0 and 4");
    TYPED(c, sc::Ops<T>::rhs_vjp(*c, (CT*)out, (const CT*)u, (const CT*)w, nu, closure));
  });
}
int sc_les_rk4_vjp(sc_ctx* c, void* out, const void* u0, const void* ubar, double dt, double nu, int closure) {
  return guard(c, [&] {
    need_fft(c);
    need(out != u0 && out != ubar, "out must not alias u0 or ubar");
    need(closure == 0 || closure == 4, "the transpose exists for closure kinds 0 and 4");
    TYPED(c, sc::Ops<T>::rk4_vjp(*c, (CT*)out, (const CT*)u0, (const CT*)ubar, dt, nu, closure));
  });
}
int sc_les_wray3_vjp(sc_ctx* c, void* out, const void* u0, const void* ubar, double dt, double nu, int closure) {
  return guard(c, [&] {
    need_fft(c);
    need(out != u0 && out != ubar, "out must not alias u0 or ubar");
    need(closure == 0 || closure == 4, "the transpose exists for closure kinds 0 and 4");
    TYPED(c, sc::Ops<T>::wray3_vjp(*c, (CT*)out, (const CT*)u0, (const CT*)ubar, dt, nu, closure));
  });
}
int sc_lincomb(sc_ctx* c, void* dst, const void* x, double a, const void* y, double b, int ncomp) {
  return guard(c, [&] {
    need(ncomp >= 1 && ncomp <= 9, "ncomp");
    TYPED(c, sc::Ops<T>::lincomb(*c, (CT*)dst, (const CT*)x, a, (const CT*)y, b, ncomp));
  });
}

}  // extern "C"
