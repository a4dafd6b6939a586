/*
 * augur.h -- C ABI of libaugur_hip.so, the MI355X (gfx950) kernels behind the
 * particle-mesh (PM) gravity step and its scatter-field CNN surrogate.
 *
 * Conventions (every entry point):
 *   - Plain pointers to device memory, sizes in elements, an explicit HIP
 *     stream passed as void* (hipStream_t).  All work is stream-ordered; no
 *     call synchronises the host, allocates device memory or frees it.
 *   - Scratch memory is caller-provided: augur_*_workspace_bytes() gives the
 *     size, the caller passes (ws, ws_bytes).  Calls are re-entrant and safe
 *     under stream capture (no hipMemset nodes; buffers are cleared by
 *     kernels).
 *   - Return 0 (AUGUR_OK) or a negative status; augur_last_error() returns a
 *     thread-local description of the last failure.
 *   - Positions are FP32 (B, N, D) row-major, D in {2, 3}, component 0 = x.
 *     Grids are (B, C, G^D) with x fastest: cell ((b*G + iz)*G + iy)*G + ix.
 *   - "off" is added to positions before dividing by the cell size h:
 *     u = (pos + off)/h - 0.5 (the reference's (pos + extent/2)/h - 0.5 with
 *     off = extent/2; off = 0 for a periodic box [0, G*h)).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * Faraday-dot-py/Augur).
 */
#ifndef AUGUR_H
#define AUGUR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AUGUR_OK 0
#define AUGUR_EINVAL (-1)       /* bad argument or shape */
#define AUGUR_EUNSUPPORTED (-2) /* valid but not supported (e.g. non power-of-two FFT) */
#define AUGUR_EHIP (-3)         /* HIP runtime error */
#define AUGUR_EWORKSPACE (-4)   /* workspace missing or too small */

const char* augur_last_error(void);
int augur_abi_version(void);

/* ---------------------------------------------------------------- a1 keys */

/* Pair-term cell keys: c = clamp(floor((p + off)/cell), 0, C-1) per axis,
 * key = b*C^D + lin(c) (x fastest); masked (mask == 0) -> B*C^D.  mask may be
 * NULL (all valid).  counts (B*C^D + 1 entries, may be NULL) receives the
 * occupancy.  Replaces ScatterField._cell_keys, scripts/scatter_field.py:302-311. */
int augur_cell_keys(const float* pos, const float* mask, int64_t B, int64_t N, int D, float off, float cell,
                    int64_t C, int64_t* keys, int64_t* counts, void* stream);

/* CIC base corner i0 = floor(u) and weight w = u - i0 per axis, FP32 with the
 * reference's rounding order.  i0, w: (n, D).
 * Replaces ScatterField._corners, scripts/scatter_field.py:179-182. */
int augur_cic_base(const float* pos, int64_t n, int D, float off, float h, int32_t* i0, float* w, void* stream);

/* -------------------------------------------------------------- a3 sort */

/* Stable sort by cell: perm orders indices by (key, index), bit-exact with
 * torch.argsort(key, stable=True); starts (nkeys + 1 entries, may be NULL)
 * receives cumsum(counts) - counts with starts[nkeys] = n.  Keys in
 * [0, nkeys), nkeys < 2^31 (out-of-range keys are clamped).  Replaces scripts/scatter_field.py:331-333. */
size_t augur_sort_workspace_bytes(int64_t n, int64_t nkeys);
int augur_sort_by_cell(const int64_t* keys, int64_t n, int64_t nkeys, int64_t* perm, int64_t* starts, void* ws,
                       size_t ws_bytes, void* stream);

/* ----------------------------------------------------------- a2 deposit */

/* Deterministic CIC deposit.  Channels: [m] (vel == NULL) or
 * [m, m*v_0, .., m*v_{D-1}], each times `scale` (the reference's
 * in_scale/h^D).  mass: (B, N) or NULL with mass_value for all; mask (B, N)
 * or NULL.  periodic != 0 wraps corners mod G, else drops corners outside
 * [0, G).  Output (B, C, G^D).  Every contribution is accumulated as an exact
 * fixed-point integer, so the grid is bitwise independent of particle order,
 * thread count and launch shape.
 * Replaces ScatterField.scatter (scripts/scatter_field.py:237-250) and
 * cic_density (model/tiered/energy.py:9-34). */
size_t augur_cic_deposit_workspace_bytes(int64_t B, int D, int64_t G, int C);
int augur_cic_deposit(const float* pos, const float* vel, const float* mass, float mass_value, const float* mask,
                      int64_t B, int64_t N, int D, int64_t G, float off, float h, int periodic, float scale, float* out,
                      void* ws, size_t ws_bytes, void* stream);
/* The same fixed-point deposit for a ragged batch: scene b holds particles
 * [starts[b], starts[b+1]) of pos (total = starts[B] particles, starts an
 * int64 device array of B+1 entries), one launch for all scenes; out is
 * (B, C, G^D) as above, workspace as augur_cic_deposit_workspace_bytes(B, D,
 * G, C).  Replaces the padded, masked batch of ScatterField.scatter
 * (scripts/scatter_field.py:237-250 with mask) when the scenes' particle
 * counts differ (the training batches of scripts/train_scatter_field.py:
 * 158-255). */
int augur_cic_deposit_ragged(const float* pos, const float* vel, const float* mass, float mass_value,
                             const int64_t* starts, int64_t B, int64_t total, int D, int64_t G, float off, float h,
                             int periodic, float scale, float* out, void* ws, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------ a6 gather */

/* CIC gather of a (B, C, G^D) field at (B, N, D) positions -> (B, N, C).
 * Replaces ScatterField.gather, scripts/scatter_field.py:252-261. */
int augur_cic_gather(const float* field, const float* pos, int64_t B, int64_t N, int D, int64_t G, int C, float off,
                     float h, int periodic, float* out, void* stream);

/* The TSC ("quad") deposit and gather: the 3-point quadratic-spline stencil of
 * ScatterField._quad_axis (scripts/scatter_field.py:184-188, used by
 * _corner_iter / _corners4 when deposit="quad", :190-235) in place of CIC's
 * 2 points: i0 = round(u) half to even, d = u - i0, axis weights
 * 0.5 (0.5 - d)^2, 0.75 - d^2, 0.5 (0.5 + d)^2 at i0 - 1, i0, i0 + 1.
 * Arguments, workspace (augur_cic_deposit_workspace_bytes), channels, the
 * fixed-point order-independent sums and the open/periodic edges are those
 * of augur_cic_deposit / augur_cic_gather. */
int augur_tsc_deposit(const float* pos, const float* vel, const float* mass, float mass_value, const float* mask,
                      int64_t B, int64_t N, int D, int64_t G, float off, float h, int periodic, float scale, float* out,
                      void* ws, size_t ws_bytes, void* stream);
int augur_tsc_gather(const float* field, const float* pos, int64_t B, int64_t N, int D, int64_t G, int C, float off,
                     float h, int periodic, float* out, void* stream);

/* ---------------------------------------------------------------- a4 FFT */

/* Batched real-to-complex FFT over the last ndim (2 or 3) axes; shape =
 * [ny, nx] or [nz, ny, nx], powers of two in [4, 512].  The input is
 * (batch, in_shape) and is zero-extended to `shape` (in_shape may be NULL =
 * shape), which gives the reference's rfft2(x, s=(2G, 2G)) padding.  Output
 * (batch, .., nx/2 + 1) complex64 interleaved, unnormalised (torch "backward").
 * Replaces torch.fft.rfft2 at scripts/scatter_field.py:273,284 and
 * scripts/pm_baseline_force_error.py:74, and rfftn for the 3D solve. */
size_t augur_fft_workspace_bytes(int ndim, const int64_t* shape, int64_t batch);
int augur_fft_r2c(const float* in, const int64_t* in_shape, float* out, int ndim, const int64_t* shape,
                  int64_t batch, void* ws, size_t ws_bytes, void* stream);

/* Inverse of augur_fft_r2c times `scale` (pass 1/prod(shape) for torch's
 * irfftn).  The real output is cropped to out_shape (NULL = shape), which
 * gives the reference's irfft2(...)[..., :G, :G].  The complex input is
 * consumed (used as scratch).
 * Replaces torch.fft.irfft2 at scripts/scatter_field.py:285 and
 * scripts/pm_baseline_force_error.py:74. */
int augur_fft_c2r(float* in, float* out, const int64_t* out_shape, int ndim, const int64_t* shape, int64_t batch,
                  float scale, void* ws, size_t ws_bytes, void* stream);

/* Pointwise complex product out = a * b (b broadcast over `batch`), n complex
 * elements per batch entry, times `scale`.  The spectral product of
 * scripts/pm_baseline_force_error.py:74 and scripts/scatter_field.py:285. */
int augur_complex_mul(const float* a, const float* b, float* out, int64_t n, int64_t batch, float scale, void* stream);

/* ------------------------------------------- a4 + a5 Poisson and gradient */

/* Periodic Poisson solve with spectral gradient on a (B, G, G, G/2+1) rfftn
 * spectrum of rho: phi_hat = -4 pi Gconst rho_hat / |k|^2 (phi_hat(0) = 0),
 * acc_hat_d = -i k_d phi_hat with the Nyquist plane of axis d zeroed,
 * k = 2 pi n / L.  acc_hat: (B, 3, G, G, G/2+1) [a_x, a_y, a_z]; phi_hat may be
 * NULL.  Outputs are multiplied by `scale` (fold the inverse-FFT 1/G^3 here).
 * New for the 3D periodic solve; replaces the kernel convolution + neg_grad of
 * scripts/pm_baseline_force_error.py:71-75. */
int augur_poisson_green_grad(const float* rho_hat, int64_t B, int64_t G, float L, float Gconst, float scale,
                             float* acc_hat, float* phi_hat, void* stream);

/* Softened open-boundary kernel K = -1/sqrt(dx^2 + dy^2 + eps^2) at
 * min-image distances d = min(i, 2G - i)*h on the (2G)^2 domain.
 * Replaces scripts/pm_baseline_force_error.py:71-73. */
int augur_open_kernel_2d(int64_t G, float h, float eps, float* out, void* stream);

/* Zero-padded central-difference -grad of (B, G, G) -> (B, 2, G, G)
 * [-dphi/dx, -dphi/dy].  Replaces ScatterField.neg_grad,
 * scripts/scatter_field.py:551-555. */
int augur_neg_grad_fd(const float* phi, int64_t B, int64_t G, float h, float* out, void* stream);

/* ------------------------------------------------------- a9 / a10 oracle */

/* FP64 softened direct sum, G = 1: for each target t,
 * a_t = sum_{j != idx_t} m_j (x_j - x_t) (|x_j - x_t|^2 + eps^2)^-1.5.
 * Sources: FP64 (pos64) or FP32 (pos32, promoted exactly); exactly one is
 * non-NULL, shape (n, D).  mass (n) or NULL (all 1).  targets: (t) indices
 * into the sources, or NULL for all.  out: (t, D) FP64.  Sources are split
 * into slices whose partial sums are added in a fixed order (run-to-run
 * identical).  Replaces exact_acc, scripts/scatter_scaling_ops.py:56-71. */
size_t augur_direct_sum_workspace_bytes(int64_t n, int64_t t, int D);
int augur_direct_sum_f64(const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                         const int64_t* targets, int64_t t, double eps, double* out, void* ws, size_t ws_bytes,
                         void* stream);

/* augur_direct_sum_f64 with its partial sums in a static device array of the
 * library (16 MiB per GPU, held until the process ends) instead of a caller
 * workspace: no allocation of any kind.  No counterpart in the reference; it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Refuses (AUGUR_EINVAL)
 * sizes whose partials exceed the array (above 699050 targets in 3D); every
 * call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * race. */
int augur_direct_sum_f64_noalloc(const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                                 const int64_t* targets, int64_t t, double eps, double* out, void* stream);

/* FP64 softened potential sums phi_t = sum_{j != idx_t} m_j (|x_j - x_t|^2 +
 * eps^2)^-1/2, out (t) FP64, same arguments and workspace as augur_direct_sum_f64.  The
 * pair-potential part of exact_energy (scripts/scatter_scaling_ops.py:90-103)
 * and gravity_sim.energy (scripts/gravity_sim.py:14-18) is -0.5 sum m_t phi_t. */
int augur_direct_potential_f64(const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                               const int64_t* targets, int64_t t, double eps, double* out, void* ws, size_t ws_bytes,
                               void* stream);

/* FP64 Ewald sum for the periodic cube [0, L)^3 with the mean density removed
 * (the problem the periodic PM solve approximates), Plummer softening eps on
 * the nearest image only.  Real space: minimum image only, erfc split with
 * alpha = alpha_l / L (the dropped images are below erfc(alpha_l / 2));
 * reciprocal: |n_k|_inf <= kmax.  Sources FP32 (n, 3), mass (n) FP64 or NULL
 * (= mass_value).  out: (t, 3) FP64.  New (the reference is isolated). */
size_t augur_ewald_workspace_bytes(int64_t n, int64_t t, int kmax);
int augur_ewald_f64(const float* pos32, const double* mass, double mass_value, int64_t n, const int64_t* targets,
                    int64_t t, double L, double eps, double alpha_l, int kmax, double* out, void* ws, size_t ws_bytes,
                    void* stream);

/* Deterministic FP64 reductions over (n, 3) FP32 velocities and optional
 * (n) FP32 potentials: out[0..2] = momentum sum m v, out[3] = kinetic
 * energy 0.5 sum m v^2, out[4] = 0.5 sum m phi (0 if phi == NULL),
 * out[5] = sum m.  Replaces the FP64 momentum / energy sums of
 * scripts/scatter_bh_1b.py:150-186 and scatter_field.momentum (:768-769). */
size_t augur_reduce_workspace_bytes(int64_t n);
int augur_reduce_energy_momentum_f64(const float* vel, const float* mass, float mass_value, const float* phi,
                                     int64_t n, double* out, void* ws, size_t ws_bytes, void* stream);

/* ----------------------------------------------- periodic PM stepper (a7) */

/* The plain PM step on a periodic cube [0, L)^3, G^3 cells (G a power of two
 * in [32, 512]), particles held sorted by CIC base cell in caller-owned state:
 *   pos, vel, acc: (n, 3) FP32; ids: (n) int32 original index; mass (n) or
 *   NULL (= mass_value).  The workspace also carries the stepper's device
 *   bookkeeping (cell offsets of the last sort, drift bound): use one
 *   workspace per particle state.
 * augur_pm_prepare(): stable sort of the state by CIC base cell in tile-major
 *   order (32x16x16-cell tiles x-fastest, cells x-fastest inside a tile;
 *   physical reorder of pos, vel, acc, mass, ids), the cell offsets and the
 *   deposit / gather work plans.
 * augur_pm_accel(): acc = PM force at pos in any order (index sort by cell,
 *   tiled deposit -> FFT Poisson + spectral gradient -> CIC gather); uses
 *   the workspace as scratch (it resets the stepper bookkeeping).
 * augur_pm_kdk(): nsteps of KDK leapfrog with the fused gather + kick + drift:
 *   v += dt/2 a; x += dt v (wrapped); a = PM(x); v += dt/2 a.  The state is
 *   re-sorted on the device (pos, vel, mass, ids permuted together) whenever
 *   the accumulated drift since the last sort could exceed one cell; acc
 *   holds the acceleration at the final positions on return.
 * Together these replace ScatterField._step (verlet; scripts/scatter_field.py:
 * 677-688), pm_acc (scripts/pm_baseline_force_error.py:68-76) and
 * gravity_sim.rollout's KDK (scripts/gravity_sim.py:31-44) for 3D periodic PM. */
size_t augur_pm_workspace_bytes(int64_t n, int64_t G);
int augur_pm_prepare(float* pos, float* vel, float* acc, float* mass, int32_t* ids, int64_t n, int64_t G, float L,
                     void* ws, size_t ws_bytes, void* stream);
int augur_pm_accel(const float* pos, float* acc, const float* mass, float mass_value, int64_t n, int64_t G, float L,
                   float Gconst, void* ws, size_t ws_bytes, void* stream);
/* Periodic CIC mass density rho (G^3, FP32) of particles in any order by the
 * same sort-by-cell tiled deposit (scratch workspace as augur_pm_accel).
 * Replaces cic_density (model/tiered/energy.py:9-34) for the periodic cube. */
int augur_pm_density(const float* pos, const float* mass, float mass_value, int64_t n, int64_t G, float L, float* rho,
                     void* ws, size_t ws_bytes, void* stream);
/* The PM grids of particles in any order: rho (G^3 FP32 mass density) and
 * force (3 x G^3 FP32 planes a_x, a_y, a_z), as computed inside
 * augur_pm_accel (scratch workspace as there).  Replaces the grid half of
 * ScatterField.force (scripts/scatter_field.py:620-666: scatter + kernel_acc
 * before the gather) for 3D periodic PM; the inputs of the surrogate. */
int augur_pm_fields(const float* pos, const float* mass, float mass_value, int64_t n, int64_t G, float L, float Gconst,
                    float* rho, float* force, void* ws, size_t ws_bytes, void* stream);
int augur_pm_kdk(float* pos, float* vel, float* acc, float* mass, int32_t* ids, float mass_value, int64_t n, int64_t G,
                 float L, float Gconst, float dt, int nsteps, void* ws, size_t ws_bytes, void* stream);
/* augur_pm_kdk with per-step FP64 diagnostics fused into the gather:
 * diag[4 s .. 4 s + 3] = [sum m vx, sum m vy, sum m vz, sum m |v|^2 / 2] at
 * the synchronised velocity after step s (the closing half kick), summed in
 * FP64 in a fixed order (deterministic).  The stepping is identical to
 * augur_pm_kdk, bit for bit.  Replaces the per-tick FP64 momentum and
 * kinetic-energy sums of the large-N loop (scripts/scatter_bh_1b.py:150-186,
 * diagnostics(), run each tick by Sim; also scatter_scaling_ops.py:90-103
 * for the kinetic part), which take a separate pass over the state. */
int augur_pm_kdk_diag(float* pos, float* vel, float* acc, float* mass, int32_t* ids, float mass_value, int64_t n,
                      int64_t G, float L, float Gconst, float dt, int nsteps, double* diag, void* ws, size_t ws_bytes,
                      void* stream);
/* One KDK step of augur_pm_kdk in two halves, so a caller can change the
 * force field between them (the surrogate's correction):
 * augur_pm_step_begin(): with first != 0 the opening half kick and drift from
 *   acc; then the gated re-sort, the deposit and the FFT solve, leaving the
 *   density and the force planes in the workspace at the offsets
 *   augur_pm_grid_offsets() reports (rho: G^3 FP32; force: 3 x G^3 FP32).
 * augur_pm_step_end(): the fused gather of the workspace force planes with
 *   the closing half kick and, unless last != 0, the next opening half kick
 *   and drift (acc is written on the last step).
 * Replace the force evaluation and update of ScatterField._step (verlet,
 * scripts/scatter_field.py:677-688) when the field comes from _grid_force
 * (:620-666) with a network applied to the PM grids. */
int augur_pm_step_begin(float* pos, float* vel, float* acc, float* mass, int32_t* ids, float mass_value, int64_t n,
                        int64_t G, float L, float Gconst, float dt, int first, void* ws, size_t ws_bytes,
                        void* stream);
int augur_pm_step_end(float* pos, float* vel, float* acc, int64_t n, int64_t G, float L, float dt, int last, void* ws,
                      size_t ws_bytes, void* stream);
int augur_pm_grid_offsets(int64_t n, int64_t G, size_t* rho_offset, size_t* force_offset);

/* The plain-PM fallback of a learned force step, decided on the device (no
 * host synchronisation, so a k-step call can be captured in a HIP graph).
 * gate: 4 uint32 words [fallback (sticky within a call), revert (this step),
 * step counter, first step that fell back (~0u: none)].
 * augur_sur_gate(phase, ...):
 *   0  reset the gate (the start of a call);
 *   1  scratch[3 p ..] = CIC gather of the force planes (3 x G^3 FP32, the PM
 *      field of augur_pm_step_begin) at pos[probe[p]], p < P; clear revert;
 *   2  gather again (after the learned field was applied); the residual
 *      sqrt(sum |a - ref|^2 / sum |ref|^2) (FP64 sums, per workgroup and
 *      then in workgroup order) goes to log[step]; above thr[step]
 *      (thr[nlog - 1] past the end) it sets fallback and revert and records
 *      the step; the step counter advances.
 * scratch: augur_sur_gate_scratch_bytes(P) bytes (16-byte aligned).
 * augur_pm_force_replay(): while gate[1] (pass &gate[1]) is set, re-makes
 * the force planes from the PM spectra the step left in the workspace (the
 * last FFT pass of augur_pm_step_begin again: the same bits); otherwise the
 * launch exits at once.  Replaces the north star's "plain PM fallback when
 * a residual check against it exceeds a threshold"; the reference has no
 * such check (its nearest is the device-side cell_max_occ overflow test of
 * the pair term, scripts/scatter_field.py:324-334). */
size_t augur_sur_gate_scratch_bytes(int64_t P);
int augur_sur_gate(int phase, const float* force, const float* pos, const int64_t* probe, int64_t P, int64_t G,
                   float L, void* scratch, uint32_t* gate, const float* thr, float* log, int nlog, void* stream);
int augur_pm_force_replay(int64_t n, int64_t G, float L, const uint32_t* gate, void* ws, size_t ws_bytes,
                          void* stream);

/* One launch over a ragged batch of B scenes (B <= 1024): scene b holds
 * particles [starts[b], starts[b+1]) of pos/vel/acc (FP32 (total, 3), starts
 * an int64 device array of B+1 nondecreasing offsets, starts[B] = total), and
 * its force field is force[b], three FP32 planes a_x, a_y, a_z of G^3 cells
 * ((B, 3, G, G, G), the batched PM solve's layout).  Periodic box [0, L)^3,
 * G a power of two.  The KDK step splits around the force evaluation:
 *   phase 0 (GATHER):     acc = F(x)
 *   phase 1 (OPEN):       v += dt/2 acc; x += dt v            (force unused)
 *   phase 2 (CLOSE):      a = F(x); v += dt/2 a; acc = a
 *   phase 3 (CLOSE_OPEN): a = F(x); v += 2 x dt/2 a; x += dt v (acc not stored)
 * so k steps are OPEN, then (k - 1) x [fields; CLOSE_OPEN], then [fields;
 * CLOSE].  Per particle the arithmetic is the single-scene stepper's (each
 * kick and drift one FMA), so batching does not change a scene's bits.  No
 * host synchronisation (capturable).  Replaces the padded, masked (B, N, D)
 * batch stepping of ScatterField._step / rollout (scripts/scatter_field.py:
 * 677-688, 745) for scenes of uneven particle counts. */
int augur_persistent_step(float* pos, float* vel, float* acc, const int64_t* starts, int64_t B, int64_t total,
                          const float* force, int64_t G, float L, float dt, int phase, void* stream);

/* augur_persistent_step that also records the largest |v| component of its
 * drifts (phases OPEN and CLOSE_OPEN) into drift_word, the drift word of a
 * planned ragged deposit's workspace (augur_ragged_plan_drift_word), with
 * device atomics: the bound from which that deposit decides its re-sort.
 * Same bits as augur_persistent_step.  Part of the ragged batch stepping
 * that replaces ScatterField._step / rollout (scripts/scatter_field.py:
 * 677-688, 745), as augur_persistent_step. */
int augur_persistent_step_tracked(float* pos, float* vel, float* acc, const int64_t* starts, int64_t B,
                                  int64_t total, const float* force, int64_t G, float L, float dt, int phase,
                                  uint32_t* drift_word, void* stream);

/* The PM force planes of B density grids in one pass sequence: rho (B, G, G,
 * G) FP32 -> force (B, 3, G, G, G) FP32 (a_x, a_y, a_z planes per grid), the
 * fused FFT passes of augur_pm_step_begin with the batch as a second grid
 * dimension, so every grid gets the bits augur_pm_solve_grid gives it alone.
 * Periodic box [0, L)^3, G a power of two in [32, 512], 1 <= B <= 65535.
 * New for the 3D periodic solve (as augur_poisson_green_grad, which it fuses
 * with the transforms); replaces the kernel FFT convolution + neg_grad of
 * scripts/pm_baseline_force_error.py:71-75 for a batch of scenes. */
/* The periodic 3D fixed-point CIC deposit of a ragged batch (as
 * augur_cic_deposit_ragged with D = 3, periodic, off = 0, no velocity
 * channels: the same units, so the same bits) on the PM stepper's tiled LDS
 * deposit instead of global atomics: the batch is sorted into scene-major
 * tile-major order (keys scene G^3 + tile-major cell; the sort permutes an
 * index array, not the particles), each scene's 32x16x16-cell tiles sum their
 * grown regions in LDS and write their cells, dense tiles merge through int64
 * images.  out: (B, G, G, G) FP32.  G a power of two in [32, 512], B G^3 <
 * 2^32, total < 2^31; h the cell size, scale the deposit scale (1/h^3 for a
 * density).  Replaces the index_add_ deposit of scripts/scatter_field.py:
 * 237-250 for a batch of 3D periodic scenes. */
size_t augur_cic_deposit_ragged_tiled_workspace_bytes(int64_t B, int64_t G, int64_t total);
int augur_cic_deposit_ragged_tiled(const float* pos, const float* mass, float mass_value, const int64_t* starts,
                                   int64_t B, int64_t total, int64_t G, float h, float scale, float* out, void* ws,
                                   size_t ws_bytes, void* stream);

/* augur_cic_deposit_ragged_tiled with its sort and plans kept in the
 * caller's workspace between calls (the same size) and redone only when the
 * particles may have moved a cell: force_sort != 0 sorts unconditionally
 * (the first call, or after the caller changed the positions); otherwise
 * the drift since the last sort is accumulated on the device from the drift
 * word (drift_dt times the largest |v| recorded by
 * augur_persistent_step_tracked since the previous call) and the batch is
 * re-sorted when it would pass one cell.  Same bits as the unplanned
 * deposit; no host synchronisation.  Replaces, as augur_cic_deposit_ragged_
 * tiled, the deposit of scripts/scatter_field.py:237-250; the drift-gated
 * re-sort is the stepper's (the reference sorts every call,
 * scripts/scatter_field.py:331-333). */
size_t augur_cic_deposit_ragged_planned_workspace_bytes(int64_t B, int64_t G, int64_t total);
int augur_cic_deposit_ragged_planned(const float* pos, const float* mass, float mass_value, const int64_t* starts,
                                     int64_t B, int64_t total, int64_t G, float h, float scale, float drift_dt,
                                     int force_sort, float* out, void* ws, size_t ws_bytes, void* stream);
/* The drift word of a planned ragged deposit's workspace (a device pointer
 * for augur_persistent_step_tracked), NULL if the workspace is too small. */
uint32_t* augur_ragged_plan_drift_word(void* ws, size_t ws_bytes);

size_t augur_pm_solve_grids_workspace_bytes(int64_t B, int64_t G);
int augur_pm_solve_grids(const float* rho, float* force, int64_t B, int64_t G, float L, float Gconst, void* ws,
                         size_t ws_bytes, void* stream);

/* ------------------------------------- large-N driver with compact state */

/* Particle state in compact encodings, 3 values per particle: positions
 * penc 0 FP32, 1 FIX32 (x = q L / 2^32, uint32), 2 FIX16 (x = q L / 2^16,
 * uint16) -- fixed point wraps the periodic box [0, L) by integer overflow --
 * and velocities venc 0 FP32, 1 FP16, 2 BF16 (round to nearest even).
 * augur_lp_encode / augur_lp_decode convert n particles from / to FP32 (either
 * array may be NULL).  The state is used in the merged KDK form M = v +
 * (dt/2) a of the reference's billion-body loop (scripts/scatter_bh_1b.py:
 * 1-9,107-146, with its fp16/bf16 state, :32,102-103):
 * augur_lp_drift_deposit(): particles [lo, hi) of n_total, x += dt M (dt = 0:
 *   no drift) stored back in the encoding, then the fixed-point CIC deposit
 *   of the PM stepper (mass mass_value each, scale 1/h^3) added into img
 *   (G^3 int64 units, kept zero between steps), so the density is bitwise the
 *   same for any chunking and order;
 * augur_lp_finish(): rho (G^3 FP32) from img, img cleared;
 * augur_pm_solve_grid(): the PM force planes of a density grid (the fused FFT
 *   passes of augur_pm_step_begin; workspace augur_pm_workspace_bytes(1, G));
 * augur_grid_mean_acc_f64(): out[0..2] = sum_c rho_c a_c, out[3] = sum_c rho_c
 *   (FP64, fixed order), the grid-weighted mean acceleration of
 *   scatter_bh_1b.py:119-120 (workspace augur_grid_mean_workspace_bytes());
 * augur_lp_kick(): particles [lo, hi): M += kick (a(x) - out[0..2]/out[3])
 *   with a the CIC gather of the planes (abar NULL: no mean removed). */
int augur_lp_encode(void* pos, int penc, void* vel, int venc, int64_t n, float L, const float* pin, const float* vin,
                    void* stream);
int augur_lp_decode(const void* pos, int penc, const void* vel, int venc, int64_t n, float L, float* pout,
                    float* vout, void* stream);
int augur_lp_drift_deposit(void* pos, int penc, const void* vel, int venc, int64_t lo, int64_t hi, int64_t n_total,
                           float mass_value, float dt, int64_t G, float L, unsigned long long* img, void* stream);
int augur_lp_finish(unsigned long long* img, int64_t G, int64_t n_total, float mass_value, float L, float* rho,
                    void* stream);
int augur_pm_solve_grid(const float* rho, float* force, int64_t G, float L, float Gconst, void* ws, size_t ws_bytes,
                        void* stream);
size_t augur_grid_mean_workspace_bytes(void);
int augur_grid_mean_acc_f64(const float* rho, const float* force, int64_t G, double* out, void* ws, size_t ws_bytes,
                            void* stream);
int augur_lp_kick(const void* pos, int penc, void* vel, int venc, int64_t lo, int64_t hi, float kick,
                  const float* force, const double* abar, int64_t G, float L, void* stream);

/* ------------------------------------------ a11 / a12 surrogate network */

/* BF16 MFMA 3x3x3 periodic convolution, FP32 accumulation, on a G^3 grid
 * (G a power of two in [32, 1024]).  x: BF16 channels-last [B][G^3][cin],
 * cin in {16, 32} (zero-pad narrower inputs), or 4 for an input layer in
 * mode 1 (the network's first convolution); w: BF16 [27][32][cin] with
 * tap = (dz*3 + dy)*3 + dx for offsets d - 1 and rows co >= cout zero;
 * bias: FP32 [32].  Modes (epilogue):
 *   0: y_out FP32 channels-first [B][cout][G^3] = conv + bias;
 *   1: z_out BF16 [B][G^3][32] = conv + bias (may be NULL), h_out = gelu(z);
 *   2: as 1 with a residual: h_out = x + gelu(z) (cin = cout = 32);
 *   3: y_out FP32 channels-last [B][G^3][cout] = conv (+ addend, same layout),
 *      no bias (the data gradient, with augur_conv3_flip_weights());
 *   4: as 3 with cin = cout = 32, and also h_out BF16 [B][G^3][32] =
 *      y_out * gelu'(z_out), z_out read here (the previous layer's saved
 *      pre-activation): the data gradient fused with the next GELU backward
 *      (augur_gelu_bwd's arithmetic), so y_out is not read back; y_out may
 *      be NULL when only dz is wanted (the last layer of the chain).
 * gelu is the tanh form (torch's approximate='tanh').  Replaces the
 * nn.Conv2d + F.gelu layers of LocalNet (scripts/scatter_field.py:15-26),
 * here in 3D with periodic padding.  Runs the library's default tile. */
int augur_conv3_bf16_fwd(const uint16_t* x, int cin, const uint16_t* w, const float* bias, int64_t B, int64_t G,
                         int mode, int cout, uint16_t* z_out, uint16_t* h_out, float* y_out, const float* addend,
                         void* stream);

/* The same convolution on an explicit tile: a kernel family (variant) and the
 * number of z planes each workgroup marches through (zc, a power of two in
 * [2, G]; the default is 16).  Families:
 *   AUGUR_CONV_TILE_ZM     32x32x16 MFMAs, every wave all 32 output channels;
 *   AUGUR_CONV_TILE_SPLIT  eight waves, each half the output channels
 *                          (cin 32 with 32 outputs: modes 1, 2, and 3 with cout 32);
 *   AUGUR_CONV_TILE_NARROW 16x16x32 MFMAs over at most 16 outputs (cin 32, modes 0, 3);
 *   AUGUR_CONV_TILE_IN4    the 4-channel input layer (cin 4, mode 1);
 *   AUGUR_CONV_TILE_DEFAULT the library's choice (IN4, else NARROW, else SPLIT, else ZM).
 * A family computes each voxel the same way for every zc, so zc changes the
 * time and not the bits.  augur_conv3_tile_variants() returns the bit mask
 * (1 << family) of the families that can run a layer (0: unsupported layer).
 * No counterpart in the reference (cuDNN picks its own algorithms); the
 * autotuner in augur_amd/autotune.py times the candidates. */
#define AUGUR_CONV_TILE_DEFAULT 0
#define AUGUR_CONV_TILE_ZM 1
#define AUGUR_CONV_TILE_SPLIT 2
#define AUGUR_CONV_TILE_NARROW 3
#define AUGUR_CONV_TILE_IN4 4  /* cin 4, mode 1, 32 outputs: 4 taps x 4 channels per MFMA K-step */
int augur_conv3_tile_variants(int cin, int mode, int cout);
int augur_conv3_bf16_fwd_tile(const uint16_t* x, int cin, const uint16_t* w, const float* bias, int64_t B, int64_t G,
                              int mode, int cout, uint16_t* z_out, uint16_t* h_out, float* y_out, const float* addend,
                              int variant, int zc, void* stream);

/* The output layer written straight into a force field: field FP32
 * channels-first [B][cout][G^3] becomes alpha (conv + bias) + field
 * (accumulate != 0, fmaf per voxel) or alpha (conv + bias) (accumulate = 0),
 * cin 32, cout <= 16, cout G^3 * 4 bytes (one scene's field) below 2 GiB, on
 * the NARROW family with zc planes per workgroup (0: the default).  While
 * skip (may be NULL) points at a non-zero word the launch
 * does nothing.  Replaces the separate `force += scale * net(x)` of a
 * correction net (the grid half of ScatterField._grid_force,
 * scripts/scatter_field.py:657-662, where the net's field is added to the
 * kernel acceleration) with no extra pass over the field. */
int augur_conv3_bf16_fwd_into(const uint16_t* x, int cin, const uint16_t* w, const float* bias, int64_t B,
                              int64_t G, int cout, float alpha, int accumulate, float* field, const uint32_t* skip,
                              int zc, void* stream);

/* y += alpha x over n FP32 values (y, x 16-byte aligned); the launch does
 * nothing while skip (may be NULL) points at a non-zero word.  A stored
 * correction field added to later steps' PM force planes (correction reuse
 * every k steps); no counterpart in the reference, whose net runs every
 * force evaluation (scripts/scatter_field.py:620-666). */
int augur_field_add(float* y, const float* x, float alpha, int64_t n, const uint32_t* skip, void* stream);

/* Weight gradient dw[27][32][cin] = sum_v dz[v][co] x[v + o_tap][ci] and
 * db[32] = sum_v dz[v][co] (FP32), dz BF16 channels-last [B][G^3][32],
 * cin 4 (the input layer on its 4 real channels), 16 or 32.  nwg persistent
 * workgroups each write a partial (cin 4: one per wave) that is summed in a
 * fixed order (run-to-run identical).  Replaces autograd's conv weight
 * gradient (scripts/train_scatter_field.py:225 loss.backward()). */
size_t augur_conv3_wgrad_workspace_bytes(int cin, int64_t B, int64_t G, int nwg);
int augur_conv3_bf16_wgrad(const uint16_t* x, int cin, const uint16_t* dz, int64_t B, int64_t G, float* dw, float* db,
                           int nwg, void* ws, size_t ws_bytes, void* stream);

/* Data-gradient weights W'[t'][a][c] = W[26 - t'][c][a] (c < cout_fwd,
 * a < cin_fwd, else 0), W: [27][32][cin_fwd] -> W': [27][32][cinp]; the data
 * gradient of a layer is then augur_conv3_bf16_fwd(dz, cinp, W', mode 3). */
int augur_conv3_flip_weights(const uint16_t* w, int cin_fwd, int cout_fwd, int cinp, uint16_t* wt, void* stream);

/* dz = dh * gelu'(z): dh FP32, z BF16, dz BF16, n elements. */
int augur_gelu_bwd(const float* dh, const uint16_t* z, int64_t n, uint16_t* dz, void* stream);

/* FP32 channels-first [B][C][G^3] * scale -> BF16 channels-last [B][G^3][CP],
 * channels C..CP-1 zero. */
int augur_pack_cl_bf16(const float* in, int C, int64_t B, int64_t G, int CP, float scale, uint16_t* out,
                       void* stream);
/* The correction net's input from the PM grids in one pass: BF16
 * channels-last [B][G^3][CP] with channel 0 = rho * rho_mul + rho_add (the
 * scaled density contrast), channels 1..3 = force * force_mul (force: B x 3
 * planes of G^3), the rest zero; CP 4 or a multiple of 8.  The input stage of
 * _grid_force (scripts/scatter_field.py:620-666: the density grid and the
 * field handed to the net), for the 3D correction net. */
int augur_pack_features_bf16(const float* rho, const float* force, int64_t B, int64_t G, float rho_mul, float rho_add,
                             float force_mul, int CP, uint16_t* out, void* stream);

/* Adam step on FP32 master weights (bias-corrected, step >= 1), gradient
 * times gscale; p_bf16 (may be NULL) receives the BF16 copy.  Replaces
 * torch.optim.Adam in scripts/train_scatter_field.py:158-255. */
int augur_adam_update(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                      float eps, int step, float gscale, uint16_t* p_bf16, void* stream);

/* The training step's gradient-norm clip and non-finite skip, decided on the
 * device (no host read, so a captured step stays capturable):
 * augur_grad_sumsq() adds sum g^2 (FP64) of one gradient tensor to acc[0];
 * augur_grad_clip_decide() takes norm = sqrt(acc[0]), skips the step when
 * the norm or *loss (may be NULL) is not finite, sets *gscale = min(1,
 * max_norm / (norm + 1e-6)) (max_norm <= 0: 1), advances the step counter
 * opt[0] on a step taken or the skip counter opt[1] on a skip, sets opt[2]
 * to 1 on a skip (0 otherwise), keeps the norm in acc[1] and clears acc[0];
 * augur_adam_update_gated() is augur_adam_update with the step from opt[0],
 * the gradient times *gscale, and no write at all while opt[2] != 0.
 * Replaces clip_grad_norm_(1.0) and the isfinite skip of
 * scripts/train_scatter_field.py:225-235. */
int augur_grad_sumsq(const float* g, int64_t n, double* acc, void* stream);
int augur_grad_clip_decide(double* acc, const float* loss, float max_norm, float* gscale, int* opt, void* stream);
int augur_adam_update_gated(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                            float eps, const int* opt, const float* gscale, uint16_t* p_bf16, void* stream);

/* ------------------------------------- 2D reference scatter-field step */

/* FP32 3x3 convolution with zero padding, bias and an epilogue, on
 * channels-first grids: out[b][co][y][x] = act(bias[co] + sum w[co][ci][dy][dx]
 * in[b][ci][y + dy - 1][x + dx - 1]) with w in nn.Conv2d's layout [cout][cin][3][3]
 * (cin = ca + cb), bias may be NULL.  Input channels ci < ca are read from a
 * (B, ca, ., .) through amode: 0 same size H x W; 1 a is 2H x 2W and is
 * averaged over 2x2 blocks (F.avg_pool2d(., 2)); 2 a is H/2 x W/2, nearest
 * upsampling (F.interpolate(scale_factor=2, mode="nearest")); channels
 * ca .. ca + cb - 1 from b (B, cb, H, W) (torch.cat([a', b], 1)).
 * act: 0 none; 1 exact (erf) GELU; 2 res + GELU, res (B, cout, H, W).
 * Replaces the nn.Conv2d / F.gelu / avg_pool2d / interpolate / cat layers of
 * LocalNet and UNet (scripts/scatter_field.py:15-67). */
int augur_conv2d_f32(const float* a, int ca, int amode, const float* b, int cb, const float* w, const float* bias,
                     int64_t B, int64_t H, int64_t W, int cout, int act, const float* res, float* out, void* stream);

/* The learned kernel K on the (2G)^2 min-image distance grid: K[iy][ix] =
 * mlp(r, log(r + h)) (1 - (1 - min(r/pp, 1)^2)^2 if pp > 0, the split taper),
 * r = sqrt(dy^2 + dx^2), d = min(i, 2G - i) h.  mlp: the 4417 FP32
 * parameters of nn.Sequential(Linear(2,64), GELU, Linear(64,64), GELU,
 * Linear(64,1)) in module order (W0, b0, W1, b1, W2, b2).  Replaces the
 * kernel construction of ScatterField._kernel_fft (scripts/scatter_field.py:
 * 263-272); its rfft2 is augur_fft_r2c. */
int augur_sf_kernel_grid(const float* mlp, int64_t G, float h, float pp, float* K, void* stream);

/* -grad by central differences, out (B, 2, G, G) = [-d/dx, -d/dy] (+ addend,
 * same shape, may be NULL).  ext = 0: phi (B, G, G) zero padded
 * (ScatterField.neg_grad, scripts/scatter_field.py:551-555); ext = 1: phi is
 * the full (B, 2G, 2G) zero-padded convolution, differenced at rows and
 * columns -1 (= 2G - 1) .. G unpadded (kernel_acc with kpad "ext", :286-288).
 * The addend is the "acc + ak" of _grid_force (:660-662). */
int augur_sf_neg_grad(const float* phi, int ext, int64_t B, int64_t G, float h, const float* addend, float* out,
                      void* stream);

/* Learned short-range pair term through a cell list: cells of side `cell`
 * (= pp), C per axis, keys and the stable order from augur_cell_keys /
 * augur_sort_by_cell (perm over B*N, starts over B*C*C + 1).  For every
 * unmasked particle, the min(knn, .) nearest unmasked others with
 * r = sqrt(|d|^2 + 1e-8) < pp add ppmlp([r, log(r + 0.05)]) (1 - (r/pp)^2)^2
 * m_j d / r to acc (B, N, 2) (accumulated).  knn <= 16.  Replaces
 * ScatterField.pp_acc_cell (scripts/scatter_field.py:324-359) and, with the
 * same result, the dense pp_acc path (:504-522). */
int augur_pair_knn_f32(const float* pos, const float* mass, const float* mask, int64_t B, int64_t N,
                       const int64_t* perm, const int64_t* starts, int C, float off, float cell, float pp, int knn,
                       const float* mlp, float* acc, void* stream);

/* acc (B, N, D) -= sum(m acc) / max(sum m, 1e-9) per scene (fix != 0;
 * m = mass * mask, FP64 sums in a fixed order), then acc *= mask.  mass and
 * mask may be NULL (ones).  Replaces the momfix and masking of
 * ScatterField._force (scripts/scatter_field.py:612-615). */
int augur_sf_momfix(float* acc, const float* mass, const float* mask, int64_t B, int64_t N, int D, int fix,
                    void* stream);

/* Elementwise step updates over n floats, in the reference's operation order:
 * phase 0 (Verlet drift): pos += dt vel + (dt^2/2) a0, aux = vel + dt a0;
 * phase 1 (Verlet close): aux = (dt/2)(a0 + a1) (aux may be NULL), vel += it;
 * phase 2 (plain step): aux = a0 dt, pos += dt (vel + aux/2), vel += aux.
 * Replaces ScatterField._step (scripts/scatter_field.py:677-688). */
int augur_sf_update(float* pos, float* vel, const float* a0, const float* a1, float* aux, int64_t n, float dt,
                    int phase, void* stream);

/* ------------------------------------------------------------ measurement */

/* Per-launch timing: when enabled, each kernel launch of the library records
 * a HIP event pair on its own stream; augur_timing_collect() waits for them
 * and returns, per kernel name, the summed milliseconds and launch count
 * (names '\n'-separated in `names`).  Disabled by default; no event work is
 * done while disabled.  Replaces the CUDA-event component timers of
 * scripts/scatter_field_2d_cost_blowup_bench.py:64-80. */
int augur_timing_enable(int on);
/* Restricts recording to launches of one name (NULL or "": all), so a timed
 * region carries the events of a single kernel only.  No counterpart in the
 * reference; for timing one kernel inside an otherwise undisturbed run. */
int augur_timing_filter(const char* name);
int augur_timing_collect(char* names, size_t names_cap, double* ms, int64_t* counts, int max_entries, int* n_entries);

#ifdef __cplusplus
}
#endif
#endif /* AUGUR_H */
