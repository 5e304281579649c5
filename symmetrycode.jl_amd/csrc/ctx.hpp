// Library context: one grid (D, n, l), one precision, one device, one stream.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include <vector>
#include <stdexcept>
#include "solver_kernels.hpp"

#ifndef SC_STAGGER
#define SC_STAGGER 1
#endif
// choose the tiled buffer and the RK4 state buffers among candidate
// allocations by the rates of the passes that write them
#ifndef SC_TUNE_TILED
#define SC_TUNE_TILED 1
#endif
// place the pipeline buffers of a large grid in scattered pieces of device
// memory, SC_SCATTER_PIECE_MB each (Ctx::scattered_allocs)
#ifndef SC_SCATTER
#define SC_SCATTER 1
#endif
#ifndef SC_SCATTER_PIECE_MB
#define SC_SCATTER_PIECE_MB 256
#endif

struct sc_cnn;   // learned closure (cnn.hip)
struct sc_pmlp;  // pointwise MLP closure (pmlp.hip)

namespace sc {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SC_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::sc::Error(3, std::string(#call) + ": " + hipGetErrorString(e_));             \
  } while (0)

struct Ctx {
  int D = 3, n = 0, nh = 0, nhp = 0, nz = 0, prec = 64, device = 0;
  int ncu = 256;  // compute units of the device (persistent grids)
  double l = 0;
  long spec_elems = 0, phys_elems = 0;  // per component
  // per component of the buffers that hold either a spectral field or the
  // z-pencil-tiled state of the multi-step path (>= spec_elems)
  long state_elems = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool sync_calls = true;
  void* tw_n = nullptr;  // forward twiddles exp(-2 pi i k / n), k < n
  void* tw_m = nullptr;  // forward twiddles exp(-2 pi i k / (n/2)), k < n/2
  // lazily allocated scratch (device layout)
  void* bufA = nullptr;   // D spectral comps: inverse pipeline
  void* bufB = nullptr;   // 6 spectral comps: forward pipeline / stress
  void* bufT = nullptr;   // 6 slots, z-pencil tiled: forward y-pass output
  bool tiled_tuned = false;  // bufT chosen among candidate allocations (choose_placement)
  bool inv_tuned = false;    // bufA likewise
  bool state_tuned = false;  // ust, acc chosen among candidate allocations (tune_state)
  bool scattered = false;    // the pipeline buffers were placed by scattered_allocs
  void* scr = nullptr;    // general spectral scratch
  int scr_comp = 0;
  void* ustart = nullptr; // RK storage (D comps each)
  void* du = nullptr;
  void* acc = nullptr;
  void* ust = nullptr;
  double* dscr = nullptr; long dscr_n = 0;
  unsigned long long* vmax = nullptr;
  double* forceA = nullptr;
  double* dtdev = nullptr;  // time step of the current step (device)
  double* dtlog = nullptr;  // step sizes of a multi-step call (device)
  long dtlog_n = 0;
  void* ws[16] = {};       // named workspaces (closures, LES, adjoint RHS)
  size_t ws_bytes[16] = {};
  void* adjw[8] = {};      // adjoint step workspaces
  size_t adjw_bytes[8] = {};
  double* bandadj = nullptr;
  double* hbuf = nullptr;  // pinned host scratch
  long hbuf_n = 0;
  // learned closure of les_rhs (closure kind 4): tau = cnn_sout * net(v / cnn_sin, cnn_extra)
  ::sc_cnn* cnn = nullptr;
  double cnn_sin = 1.0, cnn_sout = 1.0, cnn_extra = 0.0;
  int cnn_slot = 0;  // activation slot the closure's forward writes (the adjoint keeps stages in 1..3)
  // pointwise MLP closure (closure kind 7): tau = Delta^2 |G|^2 mlp(G / |G|, z)
  ::sc_pmlp* pmlp = nullptr;
  double pmlp_delta = 0.0, pmlp_visc = 1.0, pmlp_mu = 0.0, pmlp_sigma = 1.0;
  int pmlp_sym = 0;  // group-average the attached mlp (symmetrize_pointwise)
  // forcing applied inside the RHS: 0 none, 1 linear band
  int force_kind = 0;
  double injection = 0.1;
  int kband = 2;
  std::string err;
  // per-launch HIP-event timing (sc_profile)
  struct ProfRec { const char* name; hipEvent_t a, b; };
  bool prof_on = false;
  std::vector<ProfRec> prof;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      SC_HIP(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  void prof_begin(const char* name) {
    if (!prof_on) return;
    ProfRec r{name, next_event(), next_event()};
    SC_HIP(hipEventRecord(r.a, stream));
    prof.push_back(r);
  }
  void prof_end() {
    if (!prof_on) return;
    SC_HIP(hipEventRecord(prof.back().b, stream));
  }

  // Large buffers start at staggered offsets inside their allocations
  // (hipMalloc returns them all at the same large alignment, so element o of
  // every buffer would map to the same HBM channel); dmalloc / dfree
  std::vector<std::pair<void*, void*>> blocks;  // (buffer, allocation)
  int nalloc = 0;
  void dmalloc(void** p, size_t bytes) {
    const size_t off = SC_STAGGER ? size_t(nalloc % 16) * (size_t(260) << 10) : 0;
    void* base = nullptr;
    SC_HIP(hipMalloc(&base, bytes + off));
    ++nalloc;
    *p = static_cast<char*>(base) + off;
    blocks.push_back({*p, base});
  }
  // Run allocs(), which allocates about `need` bytes, while every other
  // piece (SC_SCATTER_PIECE_MB) of a ballast spread over the device's free
  // memory is held,
  // so that those allocations are made of separate pieces.  The streaming
  // passes of a large grid run slower when their buffers sit in long
  // physically contiguous ranges, which is what a device whose memory has not
  // been allocated before hands out (measured at 810^3 FP64: the forward
  // y-pass at 11.3-11.4 instead of 10.1-10.4 ms per launch, as with
  // hipDeviceMallocContiguous); processes that start on memory earlier
  // processes have used get separated pieces and the fast rates.  With
  // 256 MiB pieces the first process on a fresh device ran every pass at its
  // fast rate (1 GiB pieces left the fused z-stage 3 % slow).  The
  // ballast B is bounded so that allocs() still finds need + 8 GiB free
  // (free memory during allocs(): F - B / 2).  Everything held is freed
  // before returning.
  template <class F> void scattered_allocs(size_t need, F&& allocs) {
    std::vector<void*> keep;
    const size_t piece = size_t(SC_SCATTER_PIECE_MB) << 20, margin = size_t(8) << 30;
    size_t freeb = 0, totalb = 0;
    SC_HIP(hipMemGetInfo(&freeb, &totalb));
    size_t ballast = 0;
    if (freeb > need + margin) ballast = std::min(freeb - margin, 2 * (freeb - need - margin));
    try {
      for (size_t held = 0; held + piece <= ballast; held += piece) {
        void* p = nullptr;
        if (hipMalloc(&p, piece) != hipSuccess) {
          (void)hipGetLastError();
          break;
        }
        keep.push_back(p);
      }
      for (size_t i = 1; i < keep.size(); i += 2) {
        SC_HIP(hipFree(keep[i]));
        keep[i] = nullptr;
      }
      allocs();
    } catch (...) {
      for (void* p : keep)
        if (p) (void)hipFree(p);
      throw;
    }
    for (void* p : keep)
      if (p) SC_HIP(hipFree(p));
  }
  void dfree(void* p) {
    if (!p) return;
    for (size_t i = 0; i < blocks.size(); ++i)
      if (blocks[i].first == p) {
        (void)hipFree(blocks[i].second);
        blocks.erase(blocks.begin() + long(i));
        return;
      }
  }

  int tdim() const { return D == 3 ? 6 : 3; }
  size_t celem() const { return prec == 64 ? 16 : 8; }
  size_t relem() const { return prec == 64 ? 8 : 4; }
  SpecGeom geom() const {
    SpecGeom g;
    g.D = D; g.n = n; g.nh = nh; g.nhp = nhp; g.nz = nz;
    g.s_comp = spec_elems; g.total = long(nhp) * n * nz;
    g.kf = 3.14159265358979323846 / l * 2;
    return g;
  }
  long state_or_spec() const { return state_elems > spec_elems ? state_elems : spec_elems; }
  // spectral buffers; sized for the tiled state as well, since the state
  // buffers (du, ust, acc) hold either layout
  void* ensure_spec(void*& p, int ncomp) {
    if (!p) {
      dmalloc(&p, size_t(ncomp) * state_or_spec() * celem());
      SC_HIP(hipMemsetAsync(p, 0, size_t(ncomp) * state_or_spec() * celem(), stream));
    }
    return p;
  }
  void* scratch(int ncomp) {
    if (scr_comp < ncomp) {
      dfree(scr);
      scr = nullptr;
      dmalloc(&scr, size_t(ncomp) * spec_elems * celem());
      SC_HIP(hipMemsetAsync(scr, 0, size_t(ncomp) * spec_elems * celem(), stream));
      scr_comp = ncomp;
    }
    return scr;
  }
  void* work(int slot, size_t bytes) {
    if (ws_bytes[slot] < bytes) {
      dfree(ws[slot]);
      ws[slot] = nullptr;
      dmalloc(&ws[slot], bytes);
      SC_HIP(hipMemsetAsync(ws[slot], 0, bytes, stream));
      ws_bytes[slot] = bytes;
    }
    return ws[slot];
  }
  void* adj(int slot, size_t bytes) {
    if (adjw_bytes[slot] < bytes) {
      dfree(adjw[slot]);
      adjw[slot] = nullptr;
      dmalloc(&adjw[slot], bytes);
      SC_HIP(hipMemsetAsync(adjw[slot], 0, bytes, stream));
      adjw_bytes[slot] = bytes;
    }
    return adjw[slot];
  }
  double* dscratch(long n_) {
    if (dscr_n < n_) {
      if (dscr) SC_HIP(hipFree(dscr));
      SC_HIP(hipMalloc(&dscr, n_ * sizeof(double)));
      dscr_n = n_;
    }
    return dscr;
  }
  double* hostbuf(long n_) {
    if (hbuf_n < n_) {
      if (hbuf) SC_HIP(hipHostFree(hbuf));
      SC_HIP(hipHostMalloc(&hbuf, n_ * sizeof(double)));
      hbuf_n = n_;
    }
    return hbuf;
  }
};

}  // namespace sc
