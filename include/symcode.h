/*
 * symcode.h -- C ABI of libsymcode_hip.so, an MI355X (gfx950) implementation
 * of the pseudo-spectral incompressible Navier-Stokes hot path of
 * agdestein/SymmetryCode.jl.
 *
 * Every entry point returns an int status (0 = ok, 1 = invalid argument,
 * 2 = unsupported size, 3 = HIP error, 9 = internal); sc_last_error() gives
 * the message.  No exception crosses the ABI.
 *
 * Fields are device pointers in the device layout of the context:
 *   spectral scalar : complex T [nz][n][nhp]  (x fastest; nhp >= n/2+1 pads
 *                     each x-row to a 128-byte multiple, pad entries unused)
 *   physical scalar : T [nz][n][n]            (identical to the reference)
 * with nz = n in 3D and 1 in 2D and T = double (precision 64) or float (32).
 * A vector / tensor field is ncomp scalars back to back (component stride =
 * spec_elems or phys_elems from sc_info), in the reference's NamedTuple
 * order: vectors x,y(,z); symmetric tensors xx,yy,zz,xy,yz,zx (2D xx,yy,xy);
 * velocity gradients xx,yx,zx,xy,yy,zy,xz,yz,zz (G.ij = d u_i / d x_j).
 * sc_upload_spectral / sc_download_spectral convert from / to the reference's
 * host layout ((n/2+1, n, n) column-major complex).
 *
 * Calls are synchronous on return by default (the reference's apply!
 * synchronises after every kernel); sc_set_sync(ctx, 0) makes them
 * stream-ordered.  One context per host thread; not reentrant.
 *
 * Reference interfaces are cited as path:line in agdestein/SymmetryCode.jl.
 */
#ifndef SYMCODE_H
#define SYMCODE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sc_ctx sc_ctx;

/* Grid{D}(; l, n) (src/solver.jl:131-151) + getcache/getplan
 * (src/solver.jl:956-965, 432).  precision: 64 or 32. */
int sc_create(int D, int n, double l, int precision, int device, sc_ctx** out);
int sc_destroy(sc_ctx* ctx);
/* message of the last failed call on ctx (or of sc_create when ctx is NULL) */
const char* sc_last_error(const sc_ctx* ctx);
/* info[0..7] = D, n, nh, nhp, nz, spec_elems, phys_elems, precision */
int sc_info(const sc_ctx* ctx, int64_t* info);
/* launch on an external hipStream_t from now on (NULL = the null stream);
 * the stream created by sc_create is released */
int sc_set_stream(sc_ctx* ctx, void* hip_stream);
int sc_set_sync(sc_ctx* ctx, int sync);
int sc_sync(sc_ctx* ctx);

/* per-launch timing with HIP events on the context's stream (no reference
 * counterpart; replaces the reference's wall clocks around its loops).
 * sc_profile(ctx, 1) clears and starts recording; sc_profile_report writes
 * one line "kernel count total_ms" per kernel into buf. */
int sc_profile(sc_ctx* ctx, int enable);
int sc_profile_report(sc_ctx* ctx, char* buf, int buflen);

/* layout conversion, host reference layout <-> device layout */
int sc_upload_spectral(sc_ctx* ctx, void* dev, const void* host, int ncomp);
int sc_download_spectral(sc_ctx* ctx, void* host, const void* dev, int ncomp);
int sc_upload_physical(sc_ctx* ctx, void* dev, const void* host, int ncomp);
int sc_download_physical(sc_ctx* ctx, void* host, const void* dev, int ncomp);

/* to_phys! / to_spec! (src/solver.jl:210-215): unnormalised C2R, and R2C
 * divided by n^D.  sc_irfft / sc_rfft are the plan's ldiv! (normalised C2R)
 * and mul! (unnormalised R2C) used by dealias_phys! (src/solver.jl:218-223).
 * The spectral input is left intact (the reference's C2R destroys it). */
int sc_to_phys(sc_ctx* ctx, void* phys, const void* spec, int ncomp);
int sc_to_spec(sc_ctx* ctx, void* spec, const void* phys, int ncomp);
int sc_irfft(sc_ctx* ctx, void* phys, const void* spec, int ncomp);
int sc_rfft(sc_ctx* ctx, void* spec, const void* phys, int ncomp);

/* spectral kernels of src/solver.jl and src/filtering.jl */
int sc_twothirds(sc_ctx* ctx, void* u, int ncomp);                        /* :371-389 */
int sc_project(sc_ctx* ctx, void* u);                                     /* :350-368 */
int sc_symmetrize(sc_ctx* ctx, void* u, int ncomp);                       /* :398-430 */
int sc_viscosity(sc_ctx* ctx, void* sigma, const void* u, double visc);   /* :458-476 */
int sc_tensordivergence(sc_ctx* ctx, void* du, const void* sigma);        /* :509-521 */
int sc_vectordivergence(sc_ctx* ctx, void* div, const void* u);           /* :484-493 */
int sc_vectorgradient(sc_ctx* ctx, void* G, const void* u);               /* :711-733 */
int sc_strainrate(sc_ctx* ctx, void* S, const void* u);                   /* :776-794 */
int sc_gaussianfilter(sc_ctx* ctx, void* u, int ncomp, double delta);     /* filtering.jl:44-67 */
/* cutoff! (src/filtering.jl:28-34): ubar on the coarse context les,
 * u on the fine context dns (same D and precision) */
int sc_cutoff(sc_ctx* les, sc_ctx* dns, void* ubar, const void* u, int ncomp);
int sc_inverse_cutoff(sc_ctx* les, sc_ctx* dns, void* u, const void* ubar, int ncomp); /* :36-42 */
/* y .+= a .* x over ncomp spectral components (the axpy! of wray3!, :1002) */
int sc_axpy(sc_ctx* ctx, void* y, const void* x, double a, int ncomp);
/* make_tracefree! (src/solver.jl:305-317) on a spectral / physical tensor */
int sc_make_tracefree_spectral(sc_ctx* ctx, void* tau);
int sc_make_tracefree_physical(sc_ctx* ctx, void* tau);

/* the RHS: nonlinearity! (:434-456), stress! (:478-482) and the DNS RHS
 * convectiondiffusion!(du, u, grid, cache; visc) (:949-953) */
int sc_nonlinearity(sc_ctx* ctx, void* sigma, const void* u);
int sc_stress(sc_ctx* ctx, void* sigma, const void* u, double visc);
int sc_convectiondiffusion(sc_ctx* ctx, void* du, const void* u, double visc);
/* wray3!(convectiondiffusion!, u, dt, grid, cache; visc) (:968-1012) */
int sc_wray3(sc_ctx* ctx, void* u, double dt, double visc);
/* classic RK4 with projected stages (no reference counterpart; configs use it).
 * Buffer placement: on a 3D context whose buffers exceed 1 GB, the first
 * RHS and the first RK4 step time a pass into extra allocations of the
 * buffers it writes (the inverse pipeline's buffer: up to three more; the
 * tiled y-pass buffer: up to five more; the RK4 stage state and accumulator:
 * up to three more of each), keep the fastest and free the rest: a one-time
 * peak of up to 5 x 6 C (first RHS) or 3 x 2 x 3 C
 * (first RK4 step) more device memory, C = one spectral field, bounded to
 * leave 8 GB free; results do not depend on the choice. */
int sc_rk4(sc_ctx* ctx, void* u, double dt, double visc);
/* one step with dt = min(cfl * propose_timestep(u), dtmax), the velocity
 * maximum taken from the first stage's own C2R (the caller's
 * `Δt = min(cfl * propose_timestep(...), tstop - t)`, src/data.jl:61-62,
 * src/les.jl:209-210); *dt receives the step taken */
int sc_rk4_cfl(sc_ctx* ctx, void* u, double cfl, double visc, double dtmax, double* dt);
/* nsteps steps of sc_rk4_cfl (cfl > 0: dt = min(cfl * propose_timestep(u),
 * dtmax) per step) or of sc_rk4 (cfl <= 0: dt = dtmax) in one call, the
 * reference's march loop without the per-step host work (src/data.jl:57-66).
 * 3D: the Runge-Kutta state stays in z-pencil tiles between the passes
 * (converted from and back to the device layout of u once per call) and the
 * stage update, viscous term, forcing and projection are fused into the last
 * forward pass.  dts (nullable) receives the nsteps step sizes. */
int sc_rk4_steps(sc_ctx* ctx, void* u, int nsteps, double cfl, double visc, double dtmax, double* dts);
int sc_wray3_cfl(sc_ctx* ctx, void* u, double cfl, double visc, double dtmax, double* dt);
int sc_les_wray3_cfl(sc_ctx* ctx, void* u, double cfl, double visc, double dtmax, int closure, double p1,
                     double p2, double* dt);
/* propose_timestep(u, grid, visc, cache) (:931-947) */
int sc_propose_timestep(sc_ctx* ctx, const void* u, double visc, double* dt);
/* linear band forcing f = A u on 0 < |k| <= kband with A = injection /
 * sum_band |u|^2 (nearest reference: forced_rhs!, src/data.jl:12-23); added
 * inside sc_convectiondiffusion / sc_wray3 / sc_rk4.  injection <= 0 disables. */
int sc_set_band_forcing(sc_ctx* ctx, double injection, int kband);

/* FP64 reductions */
int sc_energy(sc_ctx* ctx, const void* u, int ncomp, double* e);          /* energy, :643-646 */
/* sc_energy without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device or pinned memory (code under HIP graph
 * capture, serving loops, harnesses that forbid it).  Its partial sums live in
 * a fixed scratch array of 532 KB per GPU that the library holds until the
 * process ends; fields of at most 1.7e10 entries per component (error 1
 * beyond).  The sum runs in chunks, so it equals sc_energy to rounding, not
 * bit for bit.  Calls that overlap on one GPU share the scratch and race. */
int sc_energy_noalloc(sc_ctx* ctx, const void* u, int ncomp, double* e);
int sc_spectraldot(sc_ctx* ctx, const void* u, const void* v, int ncomp,
                   double* re, double* im);                                /* :708-709 */
int sc_dissipation(sc_ctx* ctx, const void* u, double visc, double* eps);  /* get_dissipation!, :796-811 */
/* sum over the stored half spectrum of |u|^2, no RFFT weights: the squared
 * Julia norm of the rfft arrays, as e_post = norm(u - u_ref) / norm(u_ref)
 * uses it (src/les.jl:97) */
int sc_norm2(sc_ctx* ctx, const void* u, int ncomp, double* s);
int sc_palinstrophy(sc_ctx* ctx, const void* u, double visc, double* p);   /* get_palinstrophy!, :813-827 */
/* filter_reynolds(ubar, g, visc, Delta) (:768-774): Re_Delta = Delta^2
 * sqrt(sum_ij getenergy(2/3 (i k_j ubar_i))) / visc, the gradient truncated
 * to the 2/3 box whatever ubar holds */
int sc_filter_reynolds(sc_ctx* ctx, const void* ubar, double visc, double delta, double* re);
/* spectrum (:664-699): s[i-1] = shell i energy, i = 1..nshell */
int sc_spectrum(sc_ctx* ctx, const void* u, int ncomp, int nshell, double* s);
/* energy of integer shells 0..nshell-1 (getshells, :1080-1131) and their rescaling */
int sc_shell_energies(sc_ctx* ctx, const void* u, int ncomp, int nshell, double* out);
int sc_shell_scale(sc_ctx* ctx, void* u, int ncomp, int nshell, const double* factor);
/* maintain_shell_energy! (:1146-1152) for shells 1..nshell with reference
 * energies eref[0..nshell-1] from energy_shells (:1135-1141) */
int sc_maintain_shell_energy(sc_ctx* ctx, void* u, int nshell, const double* eref);

/* initial conditions */
int sc_randn(sc_ctx* ctx, void* u, int ncomp, uint64_t seed);
/* randomfield (:576-641) with target shell energies E0[k], k = 0..nshell-1
 * (E0[k] = totalenergy * profile(k) / sum profile) */
int sc_randomfield(sc_ctx* ctx, void* u, uint64_t seed, int nshell, const double* E0);
int sc_taylorgreen(sc_ctx* ctx, void* u, double V0, int doproject);       /* :537-570 */

/* closures (src/closures.jl) on physical fields; tau packed, GG nonsym */
int sc_gradient_phys(sc_ctx* ctx, void* GG, const void* u);               /* les.jl:11-16 */
int sc_closure_clark(sc_ctx* ctx, void* tau, const void* GG, double delta);            /* :37-42,64 */
int sc_closure_smagorinsky(sc_ctx* ctx, void* tau, const void* GG, double cs, double delta); /* :44-51 */
/* create_dynamic_smagorinsky model (:188-271); coeff receives c */
int sc_closure_dynsmag(sc_ctx* ctx, void* tau, const void* u, double delta, double* coeff);
/* create_verstappen (:53-62,68-71), the Q-R model, 3D only */
int sc_closure_verstappen(sc_ctx* ctx, void* tau, const void* GG, double c, double delta);
/* create_bardina (:73-115): trace-free scale similarity with test filter s * Delta */
int sc_closure_bardina(sc_ctx* ctx, void* tau, const void* u, double s, double delta);
/* LES statistics (solve_les's reduce!, src/les.jl:94-139; compute_sfs_stats,
 * src/analysis.jl:41-137), on device fields of the context's precision:
 * eps = -tau:S pointwise (contract_dissipation, src/solver.jl:746-759, with
 * S = sym(GG)); */
int sc_sfs_dissipation(sc_ctx* ctx, void* eps, const void* tau, const void* GG);
/* the shell-resolved transfer sum over shell i of w Re(conj(u) . clo),
 * clo = tensordivergence(2/3 to_spec(tau)), out[i-1] for shells i = 1..nshell
 * (the shells of spectrum) */
int sc_shell_transfer(sc_ctx* ctx, const void* u, const void* tau, int nshell, double* out);
/* FP64 mean of n device samples (the SFS drain -mean(tau:S), src/les.jl:121) */
int sc_sample_mean(sc_ctx* ctx, const void* x, long n, double* mean);
/* moments of n device samples (src/analysis.jl:9-18) and the backscatter
 * fraction: out = mean, median, std (n - 1), skewness, fraction < 0 */
int sc_moments(sc_ctx* ctx, const void* x, long n, double* out);
/* relerr = |b - a| / |a| and crosscor (mean-removed) of n stacked entries,
 * b the model, a the reference (src/analysis.jl:116-121) */
int sc_apriori_metrics(sc_ctx* ctx, const void* model, const void* ref, long n, double* out);
/* les!(du, u, grid, cache; model, visc) (src/les.jl:4-33) with a built-in
 * closure: 0 none, 1 Clark(p1 = Delta), 2 Smagorinsky(p1 = C_S, p2 = Delta),
 * 3 dynamic Smagorinsky(p1 = Delta), 4 the learned closure of
 * sc_set_cnn_closure, 5 Verstappen(p1 = C, p2 = Delta), 6 Bardina(p1 = s,
 * p2 = Delta), 7 the pointwise MLP of sc_set_pmlp_closure; the band forcing
 * is added when enabled */
int sc_les_rhs(sc_ctx* ctx, void* du, const void* u, double visc, int closure, double p1, double p2);
int sc_les_wray3(sc_ctx* ctx, void* u, double dt, double visc, int closure, double p1, double p2);
/* solve_les! with RK4 in place of wray3 (src/les.jl:188-247; RK4 has no
 * reference counterpart), fixed dt or CFL-limited */
int sc_les_rk4(sc_ctx* ctx, void* u, double dt, double visc, int closure, double p1, double p2);
int sc_les_rk4_cfl(sc_ctx* ctx, void* u, double cfl, double visc, double dtmax, int closure, double p1, double p2,
                   double* dt);

/* adjoint (no reference counterpart: the reference trains a-priori only).
 * Inner product: the RFFT-weighted one of spectraldot (src/solver.jl:708).
 * sc_rhs_vjp: out = J(u)^T w for convectiondiffusion! (+ band forcing).
 * sc_rk4_vjp / sc_wray3_vjp: out = (d step / d u0)^T ubar for one step of
 * fixed dt, recomputing the stage values from the checkpoint u0. */
int sc_rhs_vjp(sc_ctx* ctx, void* out, const void* u, const void* w, double visc);
int sc_rk4_vjp(sc_ctx* ctx, void* out, const void* u0, const void* ubar, double dt, double visc);
int sc_wray3_vjp(sc_ctx* ctx, void* out, const void* u0, const void* ubar, double dt, double visc);
/* the same for les! with closure 0 (none) or 4 (the attached learned closure);
 * with closure 4 the network is re-run on each stage value and its parameter
 * gradients accumulate (sc_cnn_zero_grad clears them) */
int sc_les_rhs_vjp(sc_ctx* ctx, void* out, const void* u, const void* w, double visc, int closure);
int sc_les_rk4_vjp(sc_ctx* ctx, void* out, const void* u0, const void* ubar, double dt, double visc, int closure);
int sc_les_wray3_vjp(sc_ctx* ctx, void* out, const void* u0, const void* ubar, double dt, double visc, int closure);
/* dst = a x + b y over ncomp spectral components (dst may alias x or y) */
int sc_lincomb(sc_ctx* ctx, void* dst, const void* x, double a, const void* y, double b, int ncomp);

/* ------------------------------------------------------------------------
 * Learned closure: 3x3x3 circular-padding CNN (cin -> 64 -> 64 -> 64 -> cout,
 * tanh-GELU after the hidden layers, no bias on the last), BF16 MFMA with FP32
 * accumulation and FP32 master weights, trained with Adam.  No reference
 * counterpart: the reference's learned closures are pointwise (1x1x1) nets
 * (PointwiseConv / mlp, src/nets.jl:37-81,247-298), which this net reduces to
 * when every non-centre tap is zero.  n in {64, 128}, cin <= 8, cout <= 16;
 * tensors are channels-last: input bf16 [batch][n^3][cin_p] (cin_p = 8),
 * output fp32 [batch][n^3][cout_p] (cout_p = 16).
 */
typedef struct sc_cnn sc_cnn;
int sc_cnn_create(int n, int batch_max, int cin, int width, int cout, uint64_t seed, sc_cnn** out);
int sc_cnn_destroy(sc_cnn* net);
const char* sc_cnn_last_error(const sc_cnn* net);
/* info[0..7] = n, batch_max, cin, width, cout, cin_p, cout_p, nparams */
int sc_cnn_info(const sc_cnn* net, int64_t* info);
int sc_cnn_set_stream(sc_cnn* net, void* hip_stream);
/* FP32 master parameters, host copies: conv weights [27][cout_p][cin_p] per
 * layer (tap t = (dz+1)*9 + (dy+1)*3 + (dx+1)), then the 3 hidden biases */
int sc_cnn_get_params(sc_cnn* net, float* host);
int sc_cnn_set_params(sc_cnn* net, const float* host);
int sc_cnn_get_grads(sc_cnn* net, float* host);
int sc_cnn_zero_grad(sc_cnn* net);
/* ncomp physical fields (component stride cstride, batch stride bstride
 * elements, fp64 != 0 for double) -> bf16 input; channel ncomp = extra when
 * has_extra (the constant Re_Delta feature) */
int sc_cnn_pack_input(sc_cnn* net, const void* fields, int fp64, int ncomp, long cstride, long bstride, int batch,
                      double scale, int has_extra, double extra, void* x);
/* cout physical fields -> fp32 target [batch][n^3][cout] (times scale) */
int sc_cnn_pack_target(sc_cnn* net, const void* fields, int fp64, long cstride, long bstride, int batch,
                       double scale, void* target);
int sc_cnn_forward(sc_cnn* net, const void* x, void* y, int batch);
/* MSE on the first cout channels; dy = d loss / d y as bf16 */
int sc_cnn_mse(sc_cnn* net, const void* y, const void* target, void* dy, int batch, double* loss);
/* loss of the first cout channels, dy = d loss / d y as bf16: kind 0 the MSE,
 * kind 1 the reference's relative loss sum|y - t|^2 / (sum|t|^2 + eps(Float32))
 * over the batch (create_loss, src/training.jl:3-9) */
int sc_cnn_loss(sc_cnn* net, const void* y, const void* target, void* dy, int batch, int kind, double* loss);
/* accumulate parameter gradients of the last forward; dx (nullable) = d loss / d x (fp32) */
int sc_cnn_backward(sc_cnn* net, const void* x, const void* dy, void* dx, int batch);
int sc_cnn_adam(sc_cnn* net, double lr, double beta1, double beta2, double eps);
/* OptimiserChain(ClipNorm(clip), AdamW(lr, (beta1, beta2), weight_decay, eps))
 * of the reference's train (src/training.jl:50; Optimisers.jl 0.4 rules): the
 * gradient of each parameter array (4 conv weights, 3 biases) is scaled by
 * min(1, clip / |g|) (clip <= 0: no clipping), then
 * w -= lr (mhat / (sqrt(vhat) + eps) + weight_decay w).  The accumulated
 * gradient is left as it is. */
int sc_cnn_adamw(sc_cnn* net, double lr, double beta1, double beta2, double eps, double weight_decay, double clip);
/* trace removal of the output (on != 0; cout must be 6): channels xx, yy, zz
 * lose their mean, as the last layer of the reference's nets does
 * (src/nets.jl:275-293); forward, backward and the closure follow it */
int sc_cnn_set_tracefree(sc_cnn* net, int on);
/* per-launch HIP-event timing: enable (clears the record), then a report of
 * lines "name count total_ms executed_flop" (flop = 2 * voxels * 27 * cin_p * cout_p) */
int sc_cnn_profile(sc_cnn* net, int enable);
int sc_cnn_profile_report(sc_cnn* net, char* buf, int buflen);
/* attach net as closure kind 4 of sc_les_*: tau = s_out * net(C2R(2/3 u) / s_in,
 * extra) on the channels (xx, yy, zz, xy, yz, zx); the net must have the
 * context's n, cin 3 or 4 (channel 3 = extra), cout 6, and outlive the
 * attachment.  net = NULL detaches.  (The reference's closure hook is the
 * `model(u, grad)` of les!, src/les.jl:4-33, built by fullchain,
 * src/nets.jl:311-334.) */
int sc_set_cnn_closure(sc_ctx* ctx, sc_cnn* net, double s_in, double s_out, double extra);
/* the attached learned closure's tau (packed, physical) for the spectral u */
int sc_closure_cnn(sc_ctx* ctx, void* tau, const void* u);

/* Pointwise MLP closure: the reference's `mlp` net (src/nets.jl:247-298)
 * behind `fullchain` (src/nets.jl:311-334).  D in {2, 3}; input channels
 * D^2 (+1 with use_redelta, the standardised log Re_Delta), hidden widths
 * c1, c2, c3 with GELU, output D == 3 ? 6 : 3 with the `symm` trace removal.
 * FP32 parameters in Lux's order [W1 (c_in, c_out) column-major, b1, W2, b2,
 * W3, b3, W4]; voxel matrices channels-last [nvox][c] FP32; nvox_max bounds
 * every call (activations for that many voxels are kept for backward). */
typedef struct sc_pmlp sc_pmlp;
/* A network handle owns its parameters, optimiser state, activations and
 * inference buffers: calls on one handle are serialised by the caller (its
 * training calls run on its own stream, closure evaluations on the solver's),
 * and one handle attached to several contexts must not be evaluated by them
 * concurrently. */
int sc_pmlp_create(int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max, sc_pmlp** out);
/* The tensor-basis network: tbnn_net (src/nets.jl:535-556) behind tbnn
 * (:558-583).  Inputs are the tb_kernel! invariants of the normalised
 * gradient (5 in 3D, 2 in 2D; +1 with use_redelta), outputs the coefficients
 * of its deviatoric basis (7 in 3D, 2 in 2D), contracted with the basis to the
 * packed stress.  Input rows carry the basis tensors after the network
 * inputs (row width info[9]); every other sc_pmlp_* call takes either kind. */
int sc_tbnn_create(int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max, sc_pmlp** out);
/* The group-equivariant network: equivariant_net (src/nets.jl:102-238), |G|
 * = 48 (2D: 8) copies of each hidden channel whose weights are synthesised
 * from compact trainables by get_weight_synthesis (src/symmetry.jl:156-197),
 * the sink's D^2 outputs made symmetric and trace-free.  The trainable vector
 * (get/set_params, gradients, AdamW) is the compact one in Lux's order: lift
 * (weight (D^2, c1), bias[, weight_re (1, c1)]), mid_1 and mid_2 (weight
 * (|G|, c_out, c_in), bias), sink (weight (D^2, c3)); fullchain inputs and
 * closure 7 as for the mlp. */
int sc_equi_create(int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max, sc_pmlp** out);
int sc_pmlp_destroy(sc_pmlp* net);
const char* sc_pmlp_last_error(const sc_pmlp* net);
/* info[0..11] = D, network inputs, c1, c2, c3, stress components, nparams
 * (trainable), nvox_max, kind (0 mlp, 1 TBNN, 2 equivariant), input row
 * width, |G| copies per hidden channel (1 unless equivariant), network
 * parameters */
int sc_pmlp_info(const sc_pmlp* net, int64_t* info);
int sc_pmlp_set_stream(sc_pmlp* net, void* hip_stream);
int sc_pmlp_get_params(sc_pmlp* net, float* host);
int sc_pmlp_set_params(sc_pmlp* net, const float* host);
int sc_pmlp_get_grads(sc_pmlp* net, float* host);
int sc_pmlp_zero_grad(sc_pmlp* net);
/* create_dataloader's pair (src/data.jl:705-741) from the physical velocity
 * gradient GG (D^2 comps, nonsym order, component stride np) and, when tau
 * (packed, physical) is given, the target: x = G / (|G| + eps) (+ z),
 * y = tau / (Delta^2 |G|^2 + eps), z the standardised log Re_Delta of the
 * snapshot; fp64 != 0 for double fields, eps = eps of the field type;
 * x [np][row width], y [np][stress components] FP32 (the training precision).
 * For the TBNN, x rows are [invariants, z, basis (component fastest)]
 * (create_dataloader_tbnn, src/data.jl:759-805). */
int sc_pmlp_features(sc_pmlp* net, const void* GG, const void* tau, int fp64, long np, double delta, double z,
                     void* x, void* y);
/* y = symm(mlp(x)) (TBNN: the basis contraction of its coefficients);
 * keeps the activations for sc_pmlp_backward */
int sc_pmlp_forward(sc_pmlp* net, const void* x, void* y, long nvox);
/* accumulate parameter gradients of the last forward for dy = dL/dy */
int sc_pmlp_backward(sc_pmlp* net, const void* x, const void* dy, long nvox);
/* kind 0 MSE, kind 1 the relative loss of create_loss (src/training.jl:3-9) */
int sc_pmlp_loss(sc_pmlp* net, const void* y, const void* target, void* dy, long nvox, int kind, double* loss);
/* per-array ClipNorm(clip) (clip <= 0: none) then AdamW (src/training.jl:50) */
int sc_pmlp_adamw(sc_pmlp* net, double lr, double beta1, double beta2, double eps, double weight_decay, double clip);
/* attach net as closure kind 7 of sc_les_*: fullchain's model
 * (src/nets.jl:311-334), tau = Delta^2 |G|^2 net(G / (|G| + eps(T)) [, z])
 * with G = C2R(2/3 grad u); with a Re_Delta input, z = (log Re_Delta -
 * redelta_mu) / redelta_sigma of the live field, Re_Delta = Delta^2
 * sqrt(mean |G|^2) / visc.  A TBNN handle gives tbnn's model instead.
 * Inference runs in FP64 (the reference widens the trained parameters with
 * f64) on FP64 or FP32 fields; net = NULL detaches. */
int sc_set_pmlp_closure(sc_ctx* ctx, sc_pmlp* net, double delta, double visc, double redelta_mu,
                        double redelta_sigma);
/* symmetrize_pointwise (src/nets.jl:350-389) for the attached mlp: closure 7
 * averages R_e^T m(R_e G R_e^T) R_e over the cube's 48 (2D: 8) symmetries.
 * Attaching a network switches it off. */
int sc_set_pmlp_symmetrize(sc_ctx* ctx, int on);
/* the attached pointwise closure's tau (packed, physical) for the spectral u */
int sc_closure_pmlp(sc_ctx* ctx, void* tau, const void* u);

#ifdef __cplusplus
}
#endif
#endif
