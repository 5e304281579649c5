// Interface between the solver and the learned closure (cnn.hip).  Every call
// launches on the network's current stream and does not synchronise; errors
// throw std::runtime_error.  Buffers are channels-last over one cube of side n.
#pragma once
#include <hip/hip_runtime.h>

struct sc_cnn;

namespace cnnlink {

struct Bufs {
  void* x;    // bf16 [n^3][cin_p]  network input
  float* y;   // fp32 [n^3][cout_p] network output
  void* dy;   // bf16 [n^3][cout_p] output cotangent
  float* dx;  // fp32 [n^3][cin_p]  input cotangent
};

int grid_n(const sc_cnn* c);
int in_channels(const sc_cnn* c);
int out_channels(const sc_cnn* c);
Bufs buffers(sc_cnn* c);                               // allocated on first call
hipStream_t set_stream(sc_cnn* c, hipStream_t s);     // returns the previous stream
void forward(sc_cnn* c, const void* x, float* y, int batch);
// accumulates the parameter gradients; dx may be null
void backward(sc_cnn* c, const void* x, const void* dy, float* dx, int batch);
// Activation slots: one-cube evaluations whose activations stay in sample
// slot s of the network's batch buffers until that slot is evaluated again,
// so a backward can follow without recomputing the forward.  slots(c) =
// min(4, batch_max); slot s's input buffer is slot_x(c, s) (slot 0's is
// buffers(c).x).
int slots(sc_cnn* c);
void* slot_x(sc_cnn* c, int s);
void forward_slot(sc_cnn* c, int s, float* y);
void backward_slot(sc_cnn* c, int s, const void* dy, float* dx);
// x channels 0..2 = scale * v_m (component stride `stride`), channel 3 = extra when cin > 3
template <typename T> void pack_velocity(sc_cnn* c, const T* v, long stride, double scale, double extra, void* x);
// tau_c = scale * y[:, c], c < 6, component-major
template <typename T> void unpack_tau(sc_cnn* c, const float* y, double scale, T* tau, long stride);
// dy[:, c] = scale * (symmetric cotangent of tau_c from the 9 gradient components g9)
template <typename T> void pack_tau_cotangent(sc_cnn* c, const T* g9, long stride, double scale, void* dy);
// a_m += scale * dx[:, m], m < 3
template <typename T> void add_velocity_cotangent(sc_cnn* c, const float* dx, double scale, T* a, long stride);

}  // namespace cnnlink
