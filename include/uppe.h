/*
 * C ABI of the MI355X ERK4(3)-IP UPPE propagators (libuppe.so, gfx950).
 *
 * Plain pointers and sizes only.  Device buffers (fields, workspace, outputs) are allocated by
 * the caller; the library never allocates or frees device memory.  Every entry point returns
 * 0 on success or a negative status; uppe_last_error() gives a thread-local message.  No C++
 * exception crosses this boundary.  `stream` is a hipStream_t passed as void* (NULL = default
 * stream).  Work is enqueued on that stream; only the *_status_get / *_run helpers synchronise.
 *
 * Reference interfaces each group replaces (paths relative to the reference root):
 *   uppe3d_*  ERK43IP_UPPE_3D_Optimized.__init__ / propagate / erk43_step
 *             src/3+1D-UPPE-ERK4(3)IP-envelope/solver_3d.py:8-54, 100-160, 162-233
 *             and the medium's calc_nonlinear_response + plasma RawKernel
 *             src/3+1D-UPPE-ERK4(3)IP-envelope/media_physics.py:11-67, 121-205
 *   uppe1d_*  ERK43IP_UPPE.propagate / erk43_step / linear_operator / nonlinear_operator
 *             src/1D-ERK4(3)IP UPPE gas/erk43ip_method.py:181-369 (argon patch gas.py:11-44)
 */
#ifndef UPPE_H
#define UPPE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    UPPE_OK = 0,
    UPPE_EINVAL = -1,   /* bad argument / unsupported size */
    UPPE_EHIP = -2,     /* a HIP runtime call failed */
    UPPE_ENONFINITE = -3 /* the error norm became NaN/Inf (the reference would loop forever) */
};

/* ------------------------------------------------------------------------------------------ */
/* 3+1D envelope solver                                                                        */
/* ------------------------------------------------------------------------------------------ */
typedef struct uppe3d_plan uppe3d_plan;

typedef struct {
    int nx, ny, nt;        /* grid; nx = ny in {16,...,256} (powers of 2); 128 <= nt <= 2048
                              with prime factors <= 31, at most one of them above 13 (powers of
                              2 on register transforms, others such as the MPC runner's 2046 on
                              mixed-radix LDS transforms); the adjoint takes powers of 2 only */
    double dt;             /* t[1] - t[0] (s) */
    double omega0;         /* carrier angular frequency (rad/s) */
    double a_scale;        /* max|A0| (solver_3d.py:166-168); the state is A / a_scale */
    /* medium scalars (media_physics.py:72-104 and subclasses) */
    double n2, n4;         /* Kerr and HOKE coefficients (m^2/W, m^4/W^2) */
    double Ui;             /* ionisation potential (J) */
    double rho_nt;         /* neutral density (1/m^3) */
    double tau_rec;        /* recombination time (s) */
    double tau_c;          /* collision time (s) */
    double sigma_drude;    /* Drude cross section (m^2) */
    double av_coef;        /* avalanche rate = av_coef * I (0 for the gases) */
    int use_shock;         /* shock operator (omega+omega0)/omega0 on the Kerr term */
    /* ionisation LUT: W at lut_n points of log10(I) evenly spaced on [lut_x0, lut_x1] */
    int lut_n;
    double lut_x0, lut_x1;
} uppe3d_desc;

typedef struct {
    double z, h, err;      /* position, next step size, last attempt's error norm */
    int attempts, accepted, rejected, done, nonfinite, fail_count, cur;
    int plasma_groups;     /* t-pass line groups evaluated with the plasma term since begin */
} uppe3d_status;

/* Bytes of device workspace a plan needs (tables + 7 fields + control block). */
int uppe3d_workspace_size(const uppe3d_desc *d, size_t *bytes);

/* Builds a plan over caller-owned device workspace.  Per-omega arrays (length nt, FFT order,
 * FP64, from the medium's refractive_index hook on the host): k(omega) = n*(omega0+omega)/c,
 * k_ref = k0 + k1*omega, valid = (omega0+omega > 0), anti-alias filter, shock factor,
 * valid-frequency mask.  kx2 (nx), ky2 (ny): squared transverse wavenumbers (FFT order).
 * damp (nx*ny, x-major): absorber rate.  lut_w (lut_n): ionisation rate table. */
int uppe3d_plan_create(const uppe3d_desc *d, void *workspace, size_t workspace_bytes,
                       const double *k_w, const double *kref_w, const double *valid_w,
                       const double *filt_w, const double *shock_w, const double *mask_w,
                       const double *kx2, const double *ky2, const double *damp,
                       const double *lut_w, void *stream, uppe3d_plan **out);
int uppe3d_plan_destroy(uppe3d_plan *p);

/* A0: complex64 (nx, ny, nt), t fastest, device.  Sets the state to FFT_t(A0 / a_scale). */
int uppe3d_load_field(uppe3d_plan *p, const void *a0_xyt, void *stream);
/* Starts an adaptive propagation over length L (solver_3d.py:172-180): z = 0,
 * h = min(max_step, L/1000), controller constants.  Enqueued on `stream`. */
int uppe3d_begin(uppe3d_plan *p, double L, double tol, double max_step, double min_step,
                 void *stream);
/* One ERK4(3)-IP attempt with the on-device controller (no host synchronisation). */
int uppe3d_attempt(uppe3d_plan *p, void *stream);
/* n attempts as uppe3d_attempt, with a HIP event recorded between consecutive launches;
 * synchronises once at the end.  kernel_ms[11] receives each launch slot's summed duration in
 * launch order T1 P2 T2 P3 T3 P4 T4 P5 T5 P6 CTL (Tn: fused t-passes = N evaluations, Pn:
 * xy-plane passes); total_ms the time from the first to the last event. */
int uppe3d_attempt_timed(uppe3d_plan *p, int n, void *stream, float *kernel_ms,
                         float *total_ms);
/* One attempt with a prescribed step h and accept decision (step-sequence replay). */
int uppe3d_replay_attempt(uppe3d_plan *p, double h, int accept, void *stream);
/* Runs attempts until done or max_attempts, reading the status every `check_every` attempts.
 * A window of attempts that deferred no t-pass workgroup to the plasma launch makes the
 * following attempts of the plan launch it on a small grid (it loops over its work list, so the
 * result is the same; uppe3d_begin restores the full grid). */
int uppe3d_run(uppe3d_plan *p, int max_attempts, int check_every, void *stream,
               uppe3d_status *st);
/* Synchronises `stream` and copies the control block. */
int uppe3d_status_get(uppe3d_plan *p, void *stream, uppe3d_status *st);
/* out: complex64 (nx, ny, nt) = fftshift_t(FFT_t(A)) * a_scale (solver_3d.py:215-233). */
int uppe3d_store_spectrum(uppe3d_plan *p, void *out_xyw, void *stream);
/* out: complex64 (nx, ny, nt) = A(x, y, t) * a_scale (time domain). */
int uppe3d_store_field(uppe3d_plan *p, void *out_xyt, void *stream);
/* Device pointer of the current state (complex64, layout [omega][x][y], divided by a_scale). */
int uppe3d_state_ptr(uppe3d_plan *p, void *stream, void **state);
/* Copies the current state (layout [omega][x][y], divided by a_scale) to / from a caller
 * buffer of nx*ny*nt complex64, enqueued on `stream` (no synchronisation). */
int uppe3d_get_state(uppe3d_plan *p, void *dst, void *stream);
int uppe3d_set_state(uppe3d_plan *p, const void *src, void *stream);

/* Diagnostics of the 3+1D MPC runner (src/3+1D-ERK4(3)IP UPPE/sim.py:140-220) as device
 * reductions over the current state, in normalised units (A / a_scale; the caller applies
 * a_scale^2, dx, dy, dt).  out: float64 device buffer of 8 nt + nx ny + 5 values:
 *   [0, nt)        sum_xy |B(x,y,w)|^2                 (B = FFT_t(A / a_scale), FFT order)
 *   [nt, 3 nt)     re, im of sum_xy B(x,y,w) m(x,y)     m = exp(-(x^2+y^2)/w0_mode^2)
 *   [3 nt, 4 nt)   sum_xy |A(x,y,t)|^2                 (t in natural order)
 *   [4 nt, 6 nt)   re, im of sum_xy A(x,y,t) m(x,y)
 *   [6 nt, 7 nt)   |B(ix0,iy0,w)|^2;  [7 nt, 8 nt) |A(ix0,iy0,t)|^2
 *   [8 nt, 8 nt + nx ny)   fluence sum_t |A(x,y,t)|^2, x-major
 *   then sum f, x mean, y mean, x rms, y rms of the fluence (sim.py:145-157).
 * x = x0 + i dx, y = y0 + j dy.  work: uppe3d_diag_workspace_size bytes (device).
 * uppe3d_field_at_time: out_xy complex64 (nx, ny) = A(x, y, t_it) / a_scale.
 * uppe3d_apply_mirror: the curved mirror of sim.py:82-96 and a loss factor on the state,
 * B *= amp exp(-i (omega0 + omega) (x^2 + y^2) / (c R)) (R = 0: amp only).
 * No counterpart for the ABI itself: the reference computes these with CuPy array ops. */
int uppe3d_diag_workspace_size(const uppe3d_desc *d, size_t *bytes);
int uppe3d_diagnostics(uppe3d_plan *p, double w0_mode, double x0, double dx, double y0,
                       double dy, int ix0, int iy0, void *work, void *out, void *stream);
int uppe3d_field_at_time(uppe3d_plan *p, int it, void *out_xy, void *stream);
int uppe3d_apply_mirror(uppe3d_plan *p, double R, double amp, double x0, double dx, double y0,
                        double dy, void *stream);

/* Error norm of the embedded estimate for every following attempt of the plan.
 * UPPE_NORM_LINF (the default of a new plan): max|A4 - A3| / (max|A4| + 1e-20), replaces the
 * envelope solver's norm at src/3+1D-UPPE-ERK4(3)IP-envelope/solver_3d.py:151-153.
 * UPPE_NORM_L2: sqrt(sum|A4 - A3|^2) / (sqrt(sum|A4|^2) + 1e-20) over the whole
 * (kx, ky, omega) array, the norm of the MPC runner's solver,
 * src/3+1D-ERK4(3)IP UPPE/solver_3d.py:252-254 (sums in FP64 over FP32 squares). */
#define UPPE_NORM_LINF 0
#define UPPE_NORM_L2 1
int uppe3d_set_error_norm(uppe3d_plan *p, int norm);

/* Attempt dataflow (no counterpart in the reference, whose step is one CuPy expression
 * sequence, solver_3d.py:100-160).  A plan runs each attempt either as 11 launches over 7
 * fields (31 field passes: the stage combinations in (kx, ky, omega)) or, by default where it
 * applies, restructured: every stage field kept in (x, y, omega), the two N evaluations of
 * stages 2-3 and of stages 4-5 fused into one t-pass each, 6 launch groups over 4 fields
 * (21 field passes).  Same arithmetic up to FP32 rounding.  The restructured attempt applies
 * when nx <= 128, nt <= 1024, every filter value is positive, min filter * exp(-max damp * h)
 * >= 1e-30 for the run's max_step (replayed attempts: their h), and no stage record is set.
 * uppe3d_set_fused(p, 0) forces the 11-launch attempt; uppe3d_attempt_path reports the path
 * the next attempt takes (1 restructured, 0 not).  uppe3d_attempt_timed's kernel_ms then holds
 * T1, XY-A, T23, XY-B, T45, XY-E, controller and four zeros. */
int uppe3d_set_fused(uppe3d_plan *p, int enable);
int uppe3d_attempt_path(uppe3d_plan *p, int *fused);

/* Adjoint of the 3+1D step (reverse mode through a frozen step sequence; no counterpart in
 * the reference).  uppe3d_set_stage_record makes every following attempt copy its stage
 * fields X2, X3, X4 (the inputs of the 2nd..4th nonlinear evaluations, same layout as the
 * state) into the given buffers (NULLs switch it off).  uppe3d_adjoint_step maps Bbar = dL/dB'
 * (B' = state after an accepted step of size h from state B) to dL/dB in place, given B and
 * the recorded X2..X4 of that step (it sets the plan's step size to h for its propagator
 * tables).  Complex gradients: dL/dRe z + i dL/dIm z.
 * adj_ws: uppe3d_adjoint_workspace_size bytes of device memory.  The plasma term is not
 * differentiated: uppe3d_adjoint_check (synchronises) reports whether any line of the steps
 * since the last check had a plasma term above FP32 resolution, and clears the flag. */
int uppe3d_set_stage_record(uppe3d_plan *p, void *x2, void *x3, void *x4);
int uppe3d_adjoint_workspace_size(const uppe3d_desc *d, size_t *bytes);
int uppe3d_adjoint_step(uppe3d_plan *p, double h, const void *B, const void *X2, const void *X3,
                        const void *X4, void *Bbar, void *adj_ws, void *stream);
int uppe3d_adjoint_check(uppe3d_plan *p, void *stream, int *plasma_lines);

/* ------------------------------------------------------------------------------------------ */
/* Real-field (carrier-resolved) 3+1D solver                                                  */
/* ------------------------------------------------------------------------------------------ */
/* Replaces the ERK4(3)-IP step of src/3+1D-UPPE-ERK4(3)IP-real-field/solver_3d.py:113-172
 * (erk43_step, with its low-order error estimate) and the response of media_physics.py:11-118
 * (the plasma / Drude-current kernel) and :258-318 (calc_nonlinear_response).  The field is
 * real float32 [nx][ny][nt]; spectra are complex64 [nx][ny][nt/2+1].  nx, ny: powers of two in
 * 16..256; nt: a power of two in 64..2048.  The step controller stays with the caller
 * (solver_3d.py:174-241): uppe_rf3d_attempt returns err and leaves the candidate field aside,
 * uppe_rf3d_accept makes it the state. */
typedef struct {
    int nx, ny, nt, lut_n;
    double dt;          /* s */
    double i_coef;      /* I = i_coef E^2 (0.5 eps0 c n0) */
    double chi3_coef;   /* eps0 chi3 */
    double chi5_coef;   /* eps0 chi5 (0: no fifth-order term) */
    double av_coef;     /* avalanche rate = av_coef I (0: none) */
    double rate_rec;    /* 1 / tau_rec */
    double rho_nt, Ui, e_charge, m_e, tau_c;   /* SI; rounded to float32 like the reference */
    double coupling;    /* 2 eps0 c^2 (the P / J coefficients' denominator is coupling kz) */
} uppe_rf3d_desc;
typedef struct uppe_rf3d_plan uppe_rf3d_plan;
int uppe_rf3d_workspace_size(const uppe_rf3d_desc *d, size_t *bytes);
/* Host FP64 tables over omega_w = 2 pi rfftfreq (nt/2+1 each): k_w = n(omega) omega / c (omega
 * = 0 replaced by 1e-12), kref_w = k1 omega; and kx2 [nx], ky2 [ny] (squared transverse
 * wavenumbers).  From them the kernels form the propagator exponent -i (kz - kref) and the
 * P / J coupling coefficients of solver_3d.py:54-100 per element (rounded to complex64 as the
 * reference's arrays).  filt: float32 [nt/2+1]; damp: float32 [nx][ny]; lut_x, lut_w: the
 * float32 log10(I) grid and PPT rates (media_physics.py:167-183). */
int uppe_rf3d_plan_create(const uppe_rf3d_desc *d, void *workspace, size_t workspace_bytes,
                          const double *k_w, const double *kref_w, const double *omega_w,
                          const double *kx2, const double *ky2, const float *filt,
                          const float *damp, const float *lut_x, const float *lut_w,
                          void *stream, uppe_rf3d_plan **out);
int uppe_rf3d_plan_destroy(uppe_rf3d_plan *p);
/* state = e0 / e_scale (float32, device e0 [nx][ny][nt]) */
int uppe_rf3d_load(uppe_rf3d_plan *p, const void *e0, double e_scale, void *stream);
/* one attempt of step h from the state; err = max|E4 - E_low| / (max|E4| + 1e-20) (synchronises) */
int uppe_rf3d_attempt(uppe_rf3d_plan *p, double h, void *stream, double *err);
int uppe_rf3d_accept(uppe_rf3d_plan *p);
/* state * e_scale -> out (real [nx][ny][nt]); rFFT_t of it -> out (complex [nx][ny][nt/2+1]) */
int uppe_rf3d_store_field(uppe_rf3d_plan *p, void *out, void *stream);
int uppe_rf3d_store_spectrum(uppe_rf3d_plan *p, void *out, void *stream);

/* ------------------------------------------------------------------------------------------ */
/* 1D envelope solver (single instance and batched, one workgroup per instance)               */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int nt;                /* power of two, 256..8192 */
    double dt;
    double omega0;
    double gamma;          /* nonlinear coefficient (1/(W m)), read at every propagate */
    int use_ss;            /* self-steepening with the UPPE factor (erk43ip_method.py:229-251) */
    int use_raman;         /* I_eff = (1-f_R) I + f_R ifft(fft(I) H_R) (:206-217) */
    double f_R;
    int use_plasma;        /* plasma extension (1D restatement of media_physics.py:156-205) */
    int K;                 /* multiphoton order: W = sigma_K * I^K, I = |A|^2 / area */
    double sigma_K, Ui, rho_nt, tau_rec, tau_c, sigma_drude, area;
    int avalanche;         /* W_ava = sigma_drude / Ui * I when set */
    double L, tol, max_step, min_step;
    int max_hist;          /* capacity of the history buffer (snapshots) */
    int max_attempts;      /* safety cap on attempts (0 = unlimited) */
} uppe1d_desc;

typedef struct {
    double z, h, err;      /* final position, next step, last error norm */
    int attempts, accepted, rejected, forced, nhist, done, nonfinite, fsal_hits;
    int overflow;          /* bit 0: history buffer full, bit 1: step record capacity exceeded */
} uppe1d_status;

/* Optional recording of the accepted step sequence and checkpoints (for the adjoint):
 * hs[k] = size of accepted step k (k < hs_cap); ckpt[j] = state A(omega) before accepted step
 * j*ckpt_every (j < ckpt_cap), complex64 [ckpt_cap][nt].  replay_hs (device, replay_M steps):
 * when set, the propagation takes exactly these steps, all accepted (frozen step sequence). */
typedef struct {
    double *hs;
    int hs_cap;
    void *ckpt;
    int ckpt_every, ckpt_cap;
    const double *replay_hs;
    int replay_M;
} uppe1d_record;

/* Device bytes needed by one instance (tables, state, status). */
int uppe1d_workspace_size(const uppe1d_desc *d, size_t *bytes);

/* Propagates one pulse (erk43ip_method.py:288-369).  D (complex, nt), uppe factor (nt) and
 * H_R (complex, nt, or NULL) are FP64 host arrays in FFT order.  a0: complex64 (nt), time
 * domain, device.  hist: complex64 (max_hist, nt) device, zhist: float64 (max_hist) device:
 * snapshot 0 is A0, then the reference history rule ("append while fewer than 1000 entries or
 * when z advanced by more than L/500"); once the buffer is full further snapshots are dropped
 * and st->overflow is set.  out_w (optional): complex64 (nt), final A(omega).  rec: optional
 * step recording.  Synchronises the stream. */
int uppe1d_propagate(const uppe1d_desc *d, const double *D_re, const double *D_im,
                     const double *uppe_factor, const double *H_re, const double *H_im,
                     const void *a0, void *workspace, void *hist, void *zhist, void *out_w,
                     const uppe1d_record *rec, void *stream, uppe1d_status *st);

/* Allocation-free variant of uppe1d_propagate (same arguments without `workspace`; same
 * kernels and results).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  The instance workspace lives in a static device array of the library
 * (about 0.5 MB per device, held until the process ends), sized for nt up to 8192 (larger nt
 * is rejected); the tables are staged in a static host array.  Calls on one device share that
 * array: calls that overlap on one GPU race (host calls are serialised by a lock, but work
 * enqueued on different streams is not).  Synchronises the stream. */
int uppe1d_propagate_noalloc(const uppe1d_desc *d, const double *D_re, const double *D_im,
                             const double *uppe_factor, const double *H_re, const double *H_im,
                             const void *a0, void *hist, void *zhist, void *out_w,
                             const uppe1d_record *rec, void *stream, uppe1d_status *st);

/* Slice chain (cavity / multipass drivers; replaces the host loop of gas.py:113-147 and
 * "gas copy.py":115-141 that calls propagate once per slice with a new gamma).  All slices run
 * in order in one kernel launch, with no host round trip between them.  Slice i is a full
 * adaptive propagation (z = 0, h = max_step/100, fresh controller and history) over L with
 * its own gamma, tol and step bounds, whose input is the field the reference loop hands on:
 * A_hist[-1] of slice i-1, its last recorded snapshot (the final field unless the history
 * rule skipped the last step).  d: the pulse and model (its L, tol, step bounds and gamma are
 * not used); slices: host array of nslices.  hist/zhist (optional, max_hist snapshots in
 * total): every slice's history, slice i's st[i].nhist snapshots after those of slices
 * 0..i-1 (st[i].overflow bit 0 when the buffer ran out).  out_t (optional): complex64
 * (nslices, nt), time domain, the field slice i hands on.  out_w (optional): complex64 (nt),
 * final A(omega) of the last slice.  st: host array of nslices.  Synchronises the stream. */
typedef struct {
    double L, gamma, tol, max_step, min_step;
} uppe1d_slice;
int uppe1d_slices_workspace_size(const uppe1d_desc *d, int nslices, size_t *bytes);
int uppe1d_propagate_slices(const uppe1d_desc *d, int nslices, const uppe1d_slice *slices,
                            const double *D_re, const double *D_im, const double *uppe_factor,
                            const double *H_re, const double *H_im, const void *a0,
                            void *workspace, void *hist, void *zhist, void *out_t, void *out_w,
                            void *stream, uppe1d_status *st);

/* Batched sweep: n instances of uneven nt, each with its own adaptive step, run by one
 * persistent kernel (workgroups take instances from a device counter).  Arrays are
 * concatenated in instance order: tables and a0/out_w by nt, hist by max_hist*nt, zhist by
 * max_hist, rec by instance.  prepare uploads tables (synchronises); launch enqueues only
 * kernels (capturable in a hipGraph) for instances [first, first+count) with `threads` per
 * workgroup (256/512/1024; need nt/threads in 1..8) and an LDS capacity of lds_cap points
 * (2048 or 8192), using counter `slot` (0..29, one per concurrent launch); status reads the
 * per-instance results (synchronises). */
int uppe1d_batch_workspace_size(int n, const uppe1d_desc *descs, size_t *bytes);
int uppe1d_batch_prepare(int n, const uppe1d_desc *descs, const double *D_re,
                         const double *D_im, const double *uppe_factor, const double *H_re,
                         const double *H_im, const void *a0, void *workspace, void *hist,
                         void *zhist, void *out_w, const uppe1d_record *rec, void *stream);
int uppe1d_batch_launch(void *workspace, int n, int first, int count, int threads, int lds_cap,
                        int slot, void *stream);
int uppe1d_batch_status(void *workspace, int n, void *stream, uppe1d_status *st);

/* Adjoint of a recorded forward propagation (reverse mode through the frozen accepted-step
 * sequence; no counterpart in the reference).  fwd_workspace: the workspace of the forward run
 * (uppe1d_propagate or uppe1d_batch_prepare), instance `index` of n.  hs (device, M accepted
 * step sizes) and ckpt (device, states before steps 0, C, 2C, ...) come from its
 * uppe1d_record.  lam: complex64 (nt) device, in: dL/dA(L) in the omega domain, out: dL/dA(0).
 * Convention for complex gradients: dL/dRe z + i dL/dIm z.  Outputs (device, float64):
 * gD (2 nt, re/im interleaved) = dL/dD(omega), gU (nt) = dL/d(UPPE factor), gG (1) =
 * dL/dgamma.  The segment between checkpoints is recomputed once; the plasma term is
 * differentiated in its affine regime (zero avalanche rate, no density clamp) and the call
 * fails otherwise.  Synchronises the stream. */
int uppe1d_adjoint_workspace_size(int nt, int ckpt_every, size_t *bytes);
int uppe1d_adjoint(void *fwd_workspace, int n, int index, int M, const double *hs,
                   const void *ckpt, int ckpt_every, void *lam, void *adj_workspace, double *gD,
                   double *gU, double *gG, void *stream);

const char *uppe_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* UPPE_H */
