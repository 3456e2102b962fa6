// 1D envelope UPPE / GNLSE, ERK4(3) in the interaction picture, on gfx950.
//
// Replaces ERK43IP_UPPE.propagate / erk43_step / linear_operator / nonlinear_operator
// (src/1D-ERK4(3)IP UPPE gas/erk43ip_method.py:181-369; argon patch gas.py:11-44 acts on the
// host-side tables).  One workgroup owns one pulse instance for the whole propagation: the
// adaptive loop, controller, error norm and history rule all run on the device, so a single
// launch serves one pulse or a batch of pulses of uneven length (persistent kernel, each
// workgroup takes instances from a device counter).
//
// State: A(omega) (FFT order), so the interaction-picture propagator exp(h/2 D) is diagonal.
// One attempt (FSAL: N(A) is kept from the previous attempt while A is unchanged):
//   v = E A, a1 = E N(A), a2 = N(v + h/2 a1), a3 = N(v + h/2 a2), a4 = N(E (v + h a3)),
//   beta = E (v + h/6 (a1 + 2 a2 + 2 a3)), A4 = beta + h/6 a4, a5 = N(A4),
//   err = ||A4 - A3|| / ||A4|| with A4 - A3 = (h/10)(a4 - a5), norms in FP64 (time-domain
//   values by Parseval).  N(F) = FFT(iγ·uppe·FFT... ) as in erk43ip_method.py:197-255 plus the
//   optional plasma term (1D restatement of media_physics.py:156-205).
// Fields live in registers (element e = tid + T*i), FFTs go through two LDS buffers.
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "abi_common.hpp"
#include "fft_common.hpp"

using namespace uppe;

namespace {

constexpr int MAXE = 8;   // elements per thread

struct Stat1D {
    double z, h, err;
    int attempts, accepted, rejected, forced, nhist, done, nonfinite, fsal_hits;
    int hist_overflow, rec_overflow;
};

struct Inst1D {
    int nt, max_hist, max_attempts, use_ss;
    int use_raman, use_plasma, K, avalanche;
    double gamma, f_R, L, tol, max_step, min_step;
    double sigma_K, inv_area, Ui, rho_nt, r_rec, dt, half_sigma, w0tc, av_coef, e_rec, f_rec;
    const double *D_re, *D_im;
    const float *uppe;
    const float2 *H, *tw;
    const float2 *a0;      // input, time domain
    float2 *A, *a5, *v;    // state (omega), FSAL term, IP-transformed state
    float2 *hist;          // [max_hist][nt], time domain
    double *zhist;         // [max_hist]
    float2 *outw;          // final state A(omega) (nt), or null
    const double *replay_hs;   // prescribed step sequence (every attempt accepted), or null
    int replay_M, pad1_;
    double *hs;            // accepted step sizes (record for the adjoint), or null
    float2 *ckpt;          // states A(omega) before accepted steps 0, C, 2C, ... ([ckpt_cap][nt])
    int hs_cap, ckpt_every, ckpt_cap, pad_;
    Stat1D *st;
};

// Length-N FFT (N a compile-time power of two) of the data in a (natural order) through LDS:
// Stockham autosort stages of radix 8 (then 4 or 2 for the remainder), stage size NS a
// template constant, so the butterfly index arithmetic is shifts and masks and the stage loop
// is unrolled.  Input x[j + r N/R] -> output y[(j/NS) NS R + j%NS + r NS].  Returns the buffer
// holding the result.
// Returns p unchanged, hidden from the optimiser: twiddle loads through the result stay in
// their stage instead of being hoisted above the barriers of earlier stages.
__device__ __forceinline__ const float2 *opaque_tw(const float2 *p) {
    asm volatile("" : "+v"(p));
    return p;
}

// Intermediate stages keep their data at fpad(n) = n + n / 32: the outputs of a butterfly
// are NS apart, so without padding 16 lanes writing at stride R NS (8 at NS = 1) shared one
// bank pair (8-way conflicts); padded, two-way at most.  The first stage reads and the last
// stage writes natural order, so callers see unpadded buffers of N points; the buffers hold
// N + N / 32 points (lds_pad).
__device__ __forceinline__ constexpr int fpad(int n) { return n + (n >> 5); }
constexpr int lds_pad(int n) { return n + n / 32; }

template <int T, int N, int NS, int DIR>
struct LdsFFT {
    __device__ __forceinline__ static float2 *run(float2 *src, float2 *dst,
                                                  const float2 *__restrict__ tw0, int tid) {
        const float2 *tw = opaque_tw(tw0);
        constexpr int REM = N / NS;
        constexpr int R = REM >= 8 ? 8 : (REM == 4 ? 4 : 2);
        constexpr int NB = N / R;
        constexpr int STEP = N / (NS * R);   // twiddle stride per unit of k
        constexpr bool PIN = NS > 1;         // input padded (not the first stage)
        constexpr bool POUT = NS * R < N;    // output padded (not the last stage)
#pragma unroll
        for (int j0 = 0; j0 < NB; j0 += T) {
            const int j = j0 + tid;
            if (NB % T == 0 || j < NB) {
                float2 v[R];
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = src[PIN ? fpad(j + r * NB) : j + r * NB];
                const int k = j & (NS - 1);
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const float2 w = tw[k * STEP * r];
                        v[r] = DIR > 0 ? cmul(v[r], w) : cmulc(v[r], w);
                    }
                }
                Dft<R, DIR>::run(v);
                const int b = (j - k) * R + k;   // (j / NS) NS R + j % NS
#pragma unroll
                for (int r = 0; r < R; ++r) dst[POUT ? fpad(b + r * NS) : b + r * NS] = v[r];
            }
        }
        __syncthreads();
        return LdsFFT<T, N, NS * R, DIR>::run(dst, src, tw0, tid);
    }
};
template <int T, int N, int DIR>
struct LdsFFT<T, N, N, DIR> {
    __device__ __forceinline__ static float2 *run(float2 *src, float2 *, const float2 *, int) {
        return src;
    }
};

// Out of line: the callers hold large live state (stage fields), the FFT gets its own
// register allocation.
template <int T, int DIR, int N>
__device__ __noinline__ float2 *lds_fft_c(float2 *a, float2 *b, const float2 *__restrict__ tw,
                                          int tid) {
    static_assert((N & (N - 1)) == 0 && N >= 2, "power-of-two length");
    return LdsFFT<T, N, 1, DIR>::run(a, b, tw, tid);
}

template <int T>
__device__ double block_max(double v, double *red, int tid) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) s = fmax(s, red[w]);
    return s;
}

template <int T>
__device__ double block_sum(double v, double *red, int tid) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) s += red[w];
    return s;
}

// One reference recurrence step (media_physics.py:36-63) with general avalanche rate.
__device__ __noinline__ void k1_step(double &rho, double w, double wav, double I,
                                     const Inst1D &P, float &c1, float &c2) {
    double depl = (P.rho_nt - rho) / P.rho_nt;
    if (depl < 0.0) depl = 0.0;
    const double gen = (w * P.rho_nt + wav * rho) * depl;
    const double S = w * P.rho_nt * depl;
    double e = P.e_rec, f = P.f_rec;
    if (wav != 0.0) {
        const double G = wav * depl - P.r_rec;
        double gdt = G * P.dt;
        gdt = gdt > 10.0 ? 10.0 : (gdt < -10.0 ? -10.0 : gdt);
        e = exp(gdt);
        f = fabs(gdt) < 1e-5 ? P.dt * (1.0 + 0.5 * gdt) : (e - 1.0) / G;
    }
    const double cl = gen * P.Ui / (2.0 * I + 1e-30);
    const double cd = P.half_sigma * rho;
    c1 = (float)(cl + cd);
    c2 = (float)(cd * P.w0tc);
    rho = rho * e + S * f;
    rho = rho > P.rho_nt ? P.rho_nt : (rho < 0.0 ? 0.0 : rho);
}

// Plasma coefficients of the line: on entry buf[t] = (W_ion(t), I(t)); on exit
// buf[t] = (c1, c2) with plasma term -(c1 + i c2) A(t).  Thread tid owns the contiguous
// samples [tid*E, tid*E + E).  Affine prefix scan in FP64 when the avalanche rate is zero
// and no clamp can act; otherwise thread 0 runs the serial recurrence.
template <int T>
__device__ void plasma_coeffs_1d(float2 *buf, int n, const Inst1D &P, double *scan, int tid) {
    const int E = n / T;
    const int t0 = tid * E;
    bool risk = false;
    for (int k = 0; k < E; ++k) risk |= ((double)buf[t0 + k].x * P.f_rec > P.e_rec);
    const bool serial = P.avalanche || __syncthreads_or(risk ? 1 : 0);
    if (!serial) {
        double A = 1.0, B = 0.0;
        for (int k = 0; k < E; ++k) {
            const double w = (double)buf[t0 + k].x;
            const double ak = P.e_rec - w * P.f_rec;
            B = ak * B + w * P.rho_nt * P.f_rec;
            A = ak * A;
        }
        const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double Ao = __shfl_up(A, off, 64), Bo = __shfl_up(B, off, 64);
            if (lane >= off) {
                B = A * Bo + B;
                A = A * Ao;
            }
        }
        double Ae = __shfl_up(A, 1, 64), Be = __shfl_up(B, 1, 64);
        if (lane == 0) {
            Ae = 1.0;
            Be = 0.0;
        }
        if (lane == 63) {
            scan[2 * wv] = A;
            scan[2 * wv + 1] = B;
        }
        __syncthreads();
        if (tid == 0) {   // exclusive prefix over the waves, in place
            double pa = 1.0, pb = 0.0;
            for (int w = 0; w < T / 64; ++w) {
                const double a = scan[2 * w], b = scan[2 * w + 1];
                scan[2 * w] = pa;
                scan[2 * w + 1] = pb;
                pb = a * pb + b;
                pa = a * pa;
            }
        }
        __syncthreads();
        double rho = Ae * scan[2 * wv + 1] + Be;
        for (int k = 0; k < E; ++k) {
            const float2 p = buf[t0 + k];
            float c1, c2;
            k1_step(rho, (double)p.x, 0.0, (double)p.y, P, c1, c2);
            buf[t0 + k] = make_float2(c1, c2);
        }
        __syncthreads();
        return;
    }
    if (tid == 0) {
        double rho = 0.0;
        for (int t = 0; t < n; ++t) {
            const float2 p = buf[t];
            const double wav =
                P.avalanche ? (double)(float)(P.av_coef * (double)p.y) : 0.0;
            float c1, c2;
            k1_step(rho, (double)p.x, wav, (double)p.y, P, c1, c2);
            buf[t] = make_float2(c1, c2);
        }
    }
    __syncthreads();
}

// Fills buf[t] = (W(t), I(t)) for the time samples a (element e = tid + T i) and decides
// whether the plasma term is evaluated: not when every rate is zero, nor when its
// whole-propagation bound is below one FP32 ulp of the field.  Bound: with W = sigma I^K
// (K >= 1) every sample has W <= W(I_max) and W / I <= W(I_max) / I_max, and the recurrence
// gives rho <= f rho_nt sum_t W, so |c1|, |c2| <= Ui rho_nt W(I_max) / (2 I_max) +
// sigma/2 rho_max (1 + omega0 tau_c); the term is left out when L times that is below 2^-24
// (avalanche off).  The forward and the adjoint take the same decision from the same field.
template <int T, int E>
__device__ bool plasma_setup(const Inst1D &P, const float2 (&a)[E], float2 *buf, double *scan,
                             int tid) {
    double wsum = 0.0, imax = 0.0;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const double I = (double)cabs2(a[i]) * P.inv_area;
        double pw = 1.0;   // I^K by repeated multiplication (integer K)
        for (int k = 0; k < P.K; ++k) pw *= I;
        const float w = (float)(P.sigma_K * pw);
        buf[tid + T * i] = make_float2(w, (float)I);
        wsum += (double)w;
        imax = fmax(imax, I);
    }
    wsum = block_sum<T>(wsum, scan, tid);
    imax = block_max<T>(imax, scan + 32, tid);
    if (!(wsum > 0.0)) return false;
    if (P.avalanche) return true;
    double pw = 1.0;
    for (int k = 0; k < P.K; ++k) pw *= imax;
    const double wmax = (double)(float)(P.sigma_K * pw);
    const double rmax = P.rho_nt * fmin(1.0, P.f_rec * wsum);
    const double cb = P.Ui * P.rho_nt * wmax / (2.0 * imax + 1e-30) +
                      P.half_sigma * rmax * (1.0 + P.w0tc);
    return !(P.L * cb < 5.9604644775390625e-8);
}

// N(F): F in omega (registers, e = tid + T i) -> out in omega.  If snap != nullptr the time
// samples of F are also written there (history slot).
template <int T, int NMAX, int E>
__device__ void neval(const Inst1D &P, const float2 (&F)[E], float2 (&out)[E], float2 *bufA,
                      float2 *bufB, double *scan, int tid, float2 *snap) {
    constexpr int n = T * E;
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) bufA[tid + T * i] = F[i];
    __syncthreads();
    float2 *r = lds_fft_c<T, -1, n>(bufA, bufB, P.tw, tid);
    const float invn = 1.0f / (float)n;
    float2 a[E], Ie[E];
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            a[i] = cscale(r[tid + T * i], invn);
            Ie[i] = make_float2(cabs2(a[i]), 0.0f);
            if (snap) snap[tid + T * i] = a[i];
        }
    __syncthreads();
    if (P.use_raman) {
        // I_eff = (1 - f_R) I + f_R ifft(fft(I) H)   (complex, erk43ip_method.py:206-217)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) bufA[tid + T * i] = Ie[i];
        __syncthreads();
        float2 *q = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
        float2 *o = (q == bufA) ? bufB : bufA;
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) o[tid + T * i] = cmul(q[tid + T * i], P.H[tid + T * i]);
        __syncthreads();
        float2 *o2 = (o == bufA) ? bufB : bufA;
        float2 *d = lds_fft_c<T, -1, n>(o, o2, P.tw, tid);
        const float fr = (float)P.f_R;
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                const float2 idl = cscale(d[tid + T * i], invn);
                Ie[i] = make_float2((1.0f - fr) * Ie[i].x + fr * idl.x, fr * idl.y);
            }
        __syncthreads();
    }
    // plasma coefficients (zero unless the model is on and some rate is non-zero)
    float c1[E], c2[E];
    bool pl = false;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        c1[i] = 0.0f;
        c2[i] = 0.0f;
    }
    if (P.use_plasma) {
        pl = plasma_setup<T, E>(P, a, bufA, scan, tid);
        if (pl) {
            plasma_coeffs_1d<T>(bufA, n, P, scan, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const float2 c = bufA[tid + T * i];
                    c1[i] = c.x;
                    c2[i] = c.y;
                }
        }
        __syncthreads();
    }
    // P_NL = A I_eff; plasma term -(c1 + i c2) A
    const float g = (float)P.gamma;
    if (P.use_ss) {
        // out = ifft-free omega form: iγ uppe(ω) FFT(P_NL) + FFT(plasma)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) bufA[tid + T * i] = cmul(a[i], Ie[i]);
        __syncthreads();
        float2 *k = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                const int e = tid + T * i;
                const float s = g * P.uppe[e];
                const float2 v = k[e];
                out[i] = make_float2(-s * v.y, s * v.x);
            }
        __syncthreads();
        if (pl) {
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E)
                    bufA[tid + T * i] = make_float2(-c1[i] * a[i].x + c2[i] * a[i].y,
                                                    -c1[i] * a[i].y - c2[i] * a[i].x);
            __syncthreads();
            float2 *q = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) out[i] = cadd(out[i], q[tid + T * i]);
            __syncthreads();
        }
    } else {
        // iγ P_NL + plasma in time, one FFT
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                const float2 p = cmul(a[i], Ie[i]);
                bufA[tid + T * i] = make_float2(-g * p.y - c1[i] * a[i].x + c2[i] * a[i].y,
                                                g * p.x - c1[i] * a[i].y - c2[i] * a[i].x);
            }
        __syncthreads();
        float2 *k = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) out[i] = k[tid + T * i];
        __syncthreads();
    }
}

// exp(s D) for the thread's elements; the phase s*D_im is formed and reduced in FP64.
__device__ __forceinline__ float2 ipexp(const Inst1D &P, int e, double s) {
    double sn, cs;
    sincos(s * P.D_im[e], &sn, &cs);
    const double m = exp(s * P.D_re[e]);
    return make_float2((float)(m * cs), (float)(m * sn));
}

// One attempt of size h from state A (global) with FSAL term a5 (global).  Produces A4 in
// A4o, N(A4) in a5o, returns err (all threads).  snap: history slot for A4's time samples.
template <int T, int NMAX, int E>
__device__ double attempt(const Inst1D &P, double h, float2 (&A4o)[E], float2 (&a5o)[E],
                          float2 *bufA, float2 *bufB, double *scan, int tid, float2 *snap) {
    
    const float hf = (float)h;
    float2 Ef[E], v[E], S[E], x[E], y[E];
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const int e = tid + T * i;
            Ef[i] = ipexp(P, e, 0.5 * h);
            v[i] = cmul(Ef[i], P.A[e]);
            const float2 a1 = cmul(Ef[i], P.a5[e]);
            S[i] = caxpy(v[i], hf / 6.0f, a1);
            x[i] = caxpy(v[i], 0.5f * hf, a1);
        }
    neval<T, NMAX, E>(P, x, y, bufA, bufB, scan, tid, nullptr);   // a2
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            S[i] = caxpy(S[i], hf / 3.0f, y[i]);
            x[i] = caxpy(v[i], 0.5f * hf, y[i]);
        }
    neval<T, NMAX, E>(P, x, y, bufA, bufB, scan, tid, nullptr);   // a3
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            S[i] = caxpy(S[i], hf / 3.0f, y[i]);
            x[i] = cmul(Ef[i], caxpy(v[i], hf, y[i]));
        }
    neval<T, NMAX, E>(P, x, y, bufA, bufB, scan, tid, nullptr);   // a4
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            A4o[i] = caxpy(cmul(Ef[i], S[i]), hf / 6.0f, y[i]);   // beta + h/6 a4
            x[i] = y[i];                                          // keep a4
        }
    neval<T, NMAX, E>(P, A4o, a5o, bufA, bufB, scan, tid, snap);  // a5
    double sd = 0.0, sa = 0.0;
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const float2 d = csub(x[i], a5o[i]);
            sd += (double)d.x * d.x + (double)d.y * d.y;
            sa += (double)A4o[i].x * A4o[i].x + (double)A4o[i].y * A4o[i].y;
        }
    sd = block_sum<T>(sd, scan, tid);
    sa = block_sum<T>(sa, scan + 32, tid);
    // time-domain norms (Parseval: sum_t |x|^2 = sum_w |X|^2 / n)
    double na = sqrt(sa / (double)P.nt);
    if (na < 1e-20) na = 1.0;
    return (h / 10.0) * sqrt(sd / (double)P.nt) / na;
}

template <int T, int NMAX, int E>
__device__ void run_instance(const Inst1D &P, float2 *bufA, float2 *bufB, double *scan,
                             double *bc, int tid) {
    constexpr int n = T * E;
    // A0 -> history slot 0 and state FFT(A0); a5 = N(A)
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const int e = tid + T * i;
            const float2 a0 = P.a0[e];
            if (P.max_hist > 0) P.hist[e] = a0;
            bufA[e] = a0;
        }
    __syncthreads();
    {
        float2 *r = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
        float2 A[E], N0[E];
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                A[i] = r[tid + T * i];
                P.A[tid + T * i] = A[i];
            }
        __syncthreads();
        neval<T, NMAX, E>(P, A, N0, bufA, bufB, scan, tid, nullptr);
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) P.a5[tid + T * i] = N0[i];
    }
    // controller state (thread 0 owns it; broadcast through bc[])
    double z = 0.0, h = P.max_step / 100.0, eprev = -1.0;   // eprev < 0: None
    int nh = P.max_hist > 0 ? 1 : 0, attempts = 0, acc = 0, rej = 0, forced = 0, nonfinite = 0;
    int overflow = 0, rec_overflow = 0;
    double zlast = 0.0;
    if (nh > 0 && tid == 0) P.zhist[0] = 0.0;
    double err = 0.0;
    float2 A4[E], a5[E];
    while (P.replay_hs ? (acc < P.replay_M) : (z < P.L)) {
        if (P.max_attempts > 0 && attempts >= P.max_attempts) break;
        if (P.replay_hs) h = P.replay_hs[acc];
        else if (z + h > P.L) h = P.L - z;
        const int slot = nh < P.max_hist ? nh : -1;   // no overwrite once the buffer is full
        float2 *snap = slot >= 0 ? P.hist + (size_t)slot * n : nullptr;
        double h_used = h;
        err = attempt<T, NMAX, E>(P, h, A4, a5, bufA, bufB, scan, tid, snap);
        ++attempts;
        if (!(err == err) || isinf(err)) {
            nonfinite = 1;
            break;
        }
        bool accept = err < P.tol || P.replay_hs != nullptr;
        bool record = false;
        if (accept) {
            z += h;
            double fac = 0.95 * pow(P.tol / (err + 1e-30), 0.25);
            if (eprev >= 0.0 && eprev > 1e-30) fac *= pow(eprev / (err + 1e-30), 0.1);
            eprev = err;
            const double hn = h * (fac < 2.0 ? fac : 2.0);
            h = hn < P.max_step ? hn : P.max_step;
            ++acc;
        } else {
            ++rej;
            const double fac = 0.95 * pow(P.tol / (err + 1e-30), 0.25);
            const double hn = h * (fac > 0.1 ? fac : 0.1);
            h = hn > P.min_step ? hn : P.min_step;
            if (h <= P.min_step) {
                // forced step at min_step (erk43ip_method.py:344-367)
                h_used = h;
                err = attempt<T, NMAX, E>(P, h, A4, a5, bufA, bufB, scan, tid, snap);
                ++attempts;
                ++forced;
                if (!(err == err) || isinf(err)) {
                    nonfinite = 1;
                    break;
                }
                z += h;
                eprev = -1.0;
                accept = true;
            }
        }
        if (accept && P.hs) {
            // frozen step sequence and checkpoints for the adjoint (state before the step)
            const int k = acc + forced - 1;
            if (k < P.hs_cap) {
                if (tid == 0) P.hs[k] = h_used;
                if (P.ckpt && P.ckpt_every > 0 && k % P.ckpt_every == 0 &&
                    k / P.ckpt_every < P.ckpt_cap) {
                    float2 *cp = P.ckpt + (size_t)(k / P.ckpt_every) * n;
#pragma unroll
                    for (int i = 0; i < E; ++i)
                        if (i < E) cp[tid + T * i] = P.A[tid + T * i];
                }
            } else {
                rec_overflow = 1;
            }
        }
        if (accept) {
            // history rule of erk43ip_method.py:327-329
            record = (nh < 1000) || (z - zlast > P.L / 500.0);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    P.A[tid + T * i] = A4[i];
                    P.a5[tid + T * i] = a5[i];
                }
            __syncthreads();
            if (record && P.max_hist > 0) {
                if (slot >= 0) {
                    if (tid == 0) P.zhist[slot] = z;
                    zlast = z;
                    ++nh;
                } else {
                    overflow = 1;
                }
            }
            __syncthreads();
        }
    }
    if (P.outw) {
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) P.outw[tid + T * i] = P.A[tid + T * i];
    }
    if (tid == 0) {
        Stat1D s;
        s.z = z;
        s.h = h;
        s.err = err;
        s.attempts = attempts;
        s.accepted = acc;
        s.rejected = rej;
        s.forced = forced;
        s.nhist = nh;
        s.done = (P.replay_hs ? acc >= P.replay_M : z >= P.L) ? 1 : 0;
        s.nonfinite = nonfinite;
        s.fsal_hits = attempts;   // every attempt reuses N(A) (computed once at the start)
        s.hist_overflow = overflow;
        s.rec_overflow = rec_overflow;
        *P.st = s;
    }
    __syncthreads();
}

// instance of nt = T E points (E, the values per thread, fixed at compile time)
template <int T, int NMAX, int E>
__device__ __forceinline__ void run_e(const Inst1D &P, float2 *bufA, float2 *bufB, double *scan,
                                      double *bc, int tid) {
    if constexpr (T * E <= NMAX) run_instance<T, NMAX, E>(P, bufA, bufB, scan, bc, tid);
}

__global__ void zero_counter(int *c) { *c = 0; }

// Persistent kernel: workgroups take instances [first, first + count) from a device counter
// until none is left.  An instance longer than the LDS capacity NMAX is marked as failed.
template <int T, int NMAX>
__global__ __launch_bounds__(T) void uppe1d_kernel(const Inst1D *__restrict__ insts, int count,
                                                   int *counter) {
    __shared__ float2 bufA[lds_pad(NMAX)], bufB[lds_pad(NMAX)];
    __shared__ double scan[64];
    __shared__ double bc[8];
    __shared__ int next;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int ii = next;
        __syncthreads();
        if (ii >= count) break;
        const Inst1D P = insts[ii];
        if (P.nt > NMAX || P.nt < T || P.nt % T != 0 || P.nt / T > MAXE) {
            if (tid == 0) {
                P.st->done = 0;
                P.st->nonfinite = 0;
                P.st->attempts = -1;   // configuration rejected on the device
            }
            continue;
        }
        switch (P.nt / T) {
            case 1: run_e<T, NMAX, 1>(P, bufA, bufB, scan, bc, tid); break;
            case 2: run_e<T, NMAX, 2>(P, bufA, bufB, scan, bc, tid); break;
            case 4: run_e<T, NMAX, 4>(P, bufA, bufB, scan, bc, tid); break;
            default: run_e<T, NMAX, 8>(P, bufA, bufB, scan, bc, tid); break;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Slice chain (cavity / multipass drivers, gas.py:113-147): slices run in order by one
// workgroup.  Each slice is a fresh adaptive propagation (z = 0, h = max_step/100, fresh
// controller, history slot 0 = its input) with its own length, gamma, tol and step bounds,
// whose input is what the reference loop hands on: A_hist[-1] of the previous slice, i.e. the
// state at its last recorded snapshot (kept in lastrec, omega domain).
struct Slice1D {
    double L, gamma, tol, max_step, min_step;
};

struct Chain1D {
    Inst1D P;              // the pulse and model; P.L, tol, step bounds and gamma are not used
    const Slice1D *slices;
    Stat1D *slice_st;      // [nslices]
    float2 *lastrec;       // omega state at the last recorded snapshot [nt]
    float2 *out_t;         // [nslices][nt] time domain (the field each slice hands on), or null
    int nslices, pad_;
};

template <int T, int NMAX, int E>
__device__ void run_chain(const Chain1D &C, float2 *bufA, float2 *bufB, double *scan, int tid) {
    constexpr int n = T * E;
    const Inst1D &Q = C.P;
    int used = 0;   // history snapshots stored by the previous slices
    for (int sl = 0; sl < C.nslices; ++sl) {
        Inst1D P = Q;
        const Slice1D S = C.slices[sl];
        P.L = S.L;
        P.gamma = S.gamma;
        P.tol = S.tol;
        P.max_step = S.max_step;
        P.min_step = S.min_step;
        const int cap = Q.max_hist > used ? Q.max_hist - used : 0;
        float2 *hist = cap > 0 ? Q.hist + (size_t)used * n : nullptr;
        double *zhist = cap > 0 ? Q.zhist + used : nullptr;
        // input: time-domain a0 (slice 0) or the previous slice's last recorded state
        float2 A[E], N0[E];
        if (sl == 0) {
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) bufA[tid + T * i] = Q.a0[tid + T * i];
            __syncthreads();
            float2 *r = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) A[i] = r[tid + T * i];
            __syncthreads();
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) A[i] = C.lastrec[tid + T * i];
        }
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                P.A[tid + T * i] = A[i];
                C.lastrec[tid + T * i] = A[i];
            }
        __syncthreads();
        // N(A); its inverse transform also writes the input's time samples to slot 0
        neval<T, NMAX, E>(P, A, N0, bufA, bufB, scan, tid, hist);
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) P.a5[tid + T * i] = N0[i];
        double z = 0.0, h = P.max_step / 100.0, eprev = -1.0, err = 0.0, zrec = 0.0;
        int nrec = 1, nh = cap > 0 ? 1 : 0, attempts = 0, acc = 0, rej = 0, forced = 0;
        int nonfinite = 0, overflow = (Q.max_hist > 0 && cap == 0) ? 1 : 0;
        if (nh > 0 && tid == 0) zhist[0] = 0.0;
        float2 A4[E], a5[E];
        while (z < P.L) {
            if (P.max_attempts > 0 && attempts >= P.max_attempts) break;
            if (z + h > P.L) h = P.L - z;
            const int slot = nh < cap ? nh : -1;
            float2 *snap = slot >= 0 ? hist + (size_t)slot * n : nullptr;
            err = attempt<T, NMAX, E>(P, h, A4, a5, bufA, bufB, scan, tid, snap);
            ++attempts;
            if (!(err == err) || isinf(err)) {
                nonfinite = 1;
                break;
            }
            bool accept = err < P.tol;
            if (accept) {
                z += h;
                double fac = 0.95 * pow(P.tol / (err + 1e-30), 0.25);
                if (eprev >= 0.0 && eprev > 1e-30) fac *= pow(eprev / (err + 1e-30), 0.1);
                eprev = err;
                const double hn = h * (fac < 2.0 ? fac : 2.0);
                h = hn < P.max_step ? hn : P.max_step;
                ++acc;
            } else {
                ++rej;
                const double fac = 0.95 * pow(P.tol / (err + 1e-30), 0.25);
                const double hn = h * (fac > 0.1 ? fac : 0.1);
                h = hn > P.min_step ? hn : P.min_step;
                if (h <= P.min_step) {
                    // forced step at min_step (erk43ip_method.py:344-367)
                    err = attempt<T, NMAX, E>(P, h, A4, a5, bufA, bufB, scan, tid, snap);
                    ++attempts;
                    ++forced;
                    if (!(err == err) || isinf(err)) {
                        nonfinite = 1;
                        break;
                    }
                    z += h;
                    eprev = -1.0;
                    accept = true;
                }
            }
            if (accept) {
                // history rule of erk43ip_method.py:327-329, counted independently of storage
                const bool record = (nrec < 1000) || (z - zrec > P.L / 500.0);
#pragma unroll
                for (int i = 0; i < E; ++i)
                    if (i < E) {
                        P.A[tid + T * i] = A4[i];
                        P.a5[tid + T * i] = a5[i];
                        if (record) C.lastrec[tid + T * i] = A4[i];
                    }
                if (record) {
                    ++nrec;
                    zrec = z;
                    if (cap > 0) {
                        if (slot >= 0) {
                            if (tid == 0) zhist[slot] = z;
                            ++nh;
                        } else {
                            overflow = 1;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (tid == 0) {
            Stat1D s;
            s.z = z;
            s.h = h;
            s.err = err;
            s.attempts = attempts;
            s.accepted = acc;
            s.rejected = rej;
            s.forced = forced;
            s.nhist = cap > 0 ? nh : 0;
            s.done = z >= P.L ? 1 : 0;
            s.nonfinite = nonfinite;
            s.fsal_hits = attempts;
            s.hist_overflow = overflow;
            s.rec_overflow = 0;
            C.slice_st[sl] = s;
        }
        used += cap > 0 ? nh : 0;
        if (C.out_t) {
            // the field this slice hands on, in the time domain
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) bufA[tid + T * i] = C.lastrec[tid + T * i];
            __syncthreads();
            float2 *r = lds_fft_c<T, -1, n>(bufA, bufB, P.tw, tid);
            const float invn = 1.0f / (float)n;
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) C.out_t[(size_t)sl * n + tid + T * i] = cscale(r[tid + T * i], invn);
        }
        __syncthreads();
        if (__syncthreads_or(nonfinite)) break;
    }
    if (Q.outw) {
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) Q.outw[tid + T * i] = Q.A[tid + T * i];
    }
    __syncthreads();
}

template <int T, int NMAX>
__global__ __launch_bounds__(T) void uppe1d_chain_kernel(const Chain1D *__restrict__ chain) {
    __shared__ float2 bufA[lds_pad(NMAX)], bufB[lds_pad(NMAX)];
    __shared__ double scan[64];
    const Chain1D C = *chain;
    const int nt = C.P.nt, tid = threadIdx.x;
    if (nt > NMAX || nt < T || nt % T != 0 || nt / T > MAXE) {
        if (tid == 0) C.slice_st[0].attempts = -1;   // configuration rejected on the device
        return;
    }
    switch (nt / T) {
        case 1: if constexpr (T <= NMAX) run_chain<T, NMAX, 1>(C, bufA, bufB, scan, tid); break;
        case 2: if constexpr (2 * T <= NMAX) run_chain<T, NMAX, 2>(C, bufA, bufB, scan, tid); break;
        case 4: if constexpr (4 * T <= NMAX) run_chain<T, NMAX, 4>(C, bufA, bufB, scan, tid); break;
        default: if constexpr (8 * T <= NMAX) run_chain<T, NMAX, 8>(C, bufA, bufB, scan, tid); break;
    }
}

// ---------------------------------------------------------------------------------------
// Adjoint (reverse mode) of the frozen accepted-step sequence.  Gradient convention for a
// complex quantity z: zb = dL/dRe z + i dL/dIm z (so y = c z gives zb = conj(c) yb and
// y = M z gives zb = M^H yb; FFT^H = n IFFT, IFFT^H = FFT / n).
struct Adj1D {
    Inst1D P;                 // forward tables and scalars (P.A is used as state scratch)
    const double *hs;         // accepted step sizes [M]
    const float2 *ckpt;       // state A(omega) before steps 0, C, 2C, ... [ceil(M/C)][nt]
    float2 *seg;              // recomputed segment: per step A, N(A), a2, a3, a4 [C][5][nt]
    float2 *lam;              // in: dL/dA(L); out: dL/dA(0)      (omega, [nt])
    float2 *tmp;              // vb, a1b, a2b, a3b                  [4][nt]
    double *pscr;             // plasma adjoint scratch             [4][nt]
    double *gD;               // dL/dD (re, im interleaved)         [2 nt]
    double *gU;               // dL/dU (UPPE factor)                [nt]
    double *gG;               // dL/dgamma                          [1]
    int *flags;               // bit 0: plasma regime outside the affine (differentiable) scan
    int M, C, pad0, pad1;
};

// Exclusive block scan of affine maps x -> A x + B held one per thread, in thread order
// (forward: thread 0 first) or reverse thread order.  Returns the composition of all maps
// before this thread applied to 0, i.e. the value entering this thread's segment.
template <int T, bool REVERSE>
__device__ double block_affine_entry(double A, double B, double *scan, int tid) {
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double Ao, Bo;
        bool take;
        if (!REVERSE) {
            Ao = __shfl_up(A, off, 64);
            Bo = __shfl_up(B, off, 64);
            take = lane >= off;
        } else {
            Ao = __shfl_down(A, off, 64);
            Bo = __shfl_down(B, off, 64);
            take = lane + off < 64;
        }
        if (take) {
            B = A * Bo + B;
            A = A * Ao;
        }
    }
    double Be;
    if (!REVERSE) {
        Be = __shfl_up(B, 1, 64);
        if (lane == 0) Be = 0.0;
    } else {
        Be = __shfl_down(B, 1, 64);
        if (lane == 63) Be = 0.0;
    }
    double Ae = !REVERSE ? __shfl_up(A, 1, 64) : __shfl_down(A, 1, 64);
    if ((!REVERSE && lane == 0) || (REVERSE && lane == 63)) Ae = 1.0;
    __syncthreads();
    if ((!REVERSE && lane == 63) || (REVERSE && lane == 0)) {
        scan[2 * wv] = A;
        scan[2 * wv + 1] = B;
    }
    __syncthreads();
    if (tid == 0) {
        double pa = 1.0, pb = 0.0;
        for (int k = 0; k < T / 64; ++k) {
            const int w = REVERSE ? T / 64 - 1 - k : k;
            const double a = scan[2 * w], b = scan[2 * w + 1];
            scan[2 * w] = pa;
            scan[2 * w + 1] = pb;
            pb = a * pb + b;
            pa = a * pa;
        }
    }
    __syncthreads();
    const double entry = Ae * scan[2 * wv + 1] + Be;
    __syncthreads();
    return entry;
}

// Adds the plasma-term contribution to ab (time, natural mapping) given a (time) and
// Pb = dL/dP (time).  Affine regime only (zero avalanche rate, no clamp); otherwise flags.
template <int T, int NMAX, int E>
__device__ void plasma_adjoint_1d(const Adj1D &J, const float2 (&a)[E], const float2 (&Pb)[E],
                                  float2 (&ab)[E], float2 *bufA, float2 *bufB, double *scan,
                                  int tid) {
    const Inst1D &P = J.P;
    constexpr int n = T * E;
    if (!plasma_setup<T, E>(P, a, bufA, scan, tid)) return;
    bool risk = false;
#pragma unroll
    for (int i = 0; i < E; ++i)
        risk |= ((double)bufA[tid + T * i].x * P.f_rec > P.e_rec);
    if (P.avalanche || __syncthreads_or(risk ? 1 : 0)) {
        if (tid == 0) atomicOr(J.flags, 1);
        return;
    }
    // forward densities rho_t (pre-update), thread-contiguous segments -> bufB as doubles
    double *rho_s = (double *)bufB;
    const int t0 = tid * E;
    {
        double A = 1.0, B = 0.0;
        for (int k = 0; k < E; ++k) {
            const double w = (double)bufA[t0 + k].x;
            const double ak = P.e_rec - w * P.f_rec;
            B = ak * B + w * P.rho_nt * P.f_rec;
            A = ak * A;
        }
        double rho = block_affine_entry<T, false>(A, B, scan, tid);
        for (int k = 0; k < E; ++k) {
            rho_s[t0 + k] = rho;
            const double w = (double)bufA[t0 + k].x;
            double depl = (P.rho_nt - rho) / P.rho_nt;
            if (depl < 0.0) depl = 0.0;
            rho = rho * P.e_rec + (w * P.rho_nt * depl) * P.f_rec;
            rho = rho > P.rho_nt ? P.rho_nt : (rho < 0.0 ? 0.0 : rho);
        }
    }
    __syncthreads();
    double *dd = J.pscr, *wb = J.pscr + n, *Ib = J.pscr + 2 * n;
    // direct terms at every sample (natural mapping)
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const int t = tid + T * i;
            const double rho = rho_s[t];
            const double w = (double)bufA[t].x, I = (double)bufA[t].y;
            const double den = 2.0 * I + 1e-30;
            const double cl = w * (P.rho_nt - rho) * P.Ui / den;
            const double c1 = cl + P.half_sigma * rho, c2 = P.half_sigma * rho * P.w0tc;
            ab[i] = cadd(ab[i], make_float2((float)(-c1 * Pb[i].x - c2 * Pb[i].y),
                                            (float)(-c1 * Pb[i].y + c2 * Pb[i].x)));
            const double c1b = -((double)Pb[i].x * a[i].x + (double)Pb[i].y * a[i].y);
            const double c2b = (double)Pb[i].x * a[i].y - (double)Pb[i].y * a[i].x;
            dd[t] = c1b * (P.half_sigma - w * P.Ui / den) + c2b * P.half_sigma * P.w0tc;
            wb[t] = c1b * (P.rho_nt - rho) * P.Ui / den;
            Ib[t] = -c1b * cl * 2.0 / den;
        }
    __syncthreads();
    // reverse scan: rb_t = d_t + (e - w_t f) rb_{t+1}, rb_n = 0; wb_t += rb_{t+1} (rho_nt - rho_t) f
    {
        double A = 1.0, B = 0.0;   // segment map: rb_{t0} = A rb_{t0+E} + B
        for (int k = E - 1; k >= 0; --k) {
            const double al = P.e_rec - (double)bufA[t0 + k].x * P.f_rec;
            B = dd[t0 + k] + al * B;
            A = al * A;
        }
        double rb_next = block_affine_entry<T, true>(A, B, scan, tid);   // rb_{t0+E}
        for (int k = E - 1; k >= 0; --k) {
            const int t = t0 + k;
            const double w = (double)bufA[t].x, I = (double)bufA[t].y;
            double wbt = wb[t] + rb_next * (P.rho_nt - rho_s[t]) * P.f_rec;
            const double rb = dd[t] + (P.e_rec - w * P.f_rec) * rb_next;
            // w = sigma_K I^K  ->  dw/dI = K w / I
            Ib[t] += (I > 0.0) ? wbt * (double)P.K * w / I : 0.0;
            rb_next = rb;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const double s = 2.0 * Ib[tid + T * i] * P.inv_area;
            ab[i] = cadd(ab[i], make_float2((float)(s * a[i].x), (float)(s * a[i].y)));
        }
    __syncthreads();
}

// Fb = N^*(F; yb): adjoint of y = N(F), accumulating dL/dU (per omega) and dL/dgamma.
template <int T, int NMAX, int E>
__device__ void adj_neval(const Adj1D &J, const float2 (&F)[E], const float2 (&yb)[E],
                          float2 (&Fb)[E], double &gG, float2 *bufA, float2 *bufB,
                          double *scan, int tid) {
    const Inst1D &P = J.P;
    constexpr int n = T * E;
    const float invn = 1.0f / (float)n;
    const float g = (float)P.gamma;
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) bufA[tid + T * i] = F[i];
    __syncthreads();
    float2 *r = lds_fft_c<T, -1, n>(bufA, bufB, P.tw, tid);
    float2 a[E], pb[E], Pb[E], ab[E];
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) a[i] = cscale(r[tid + T * i], invn);
    __syncthreads();
    double gsum = 0.0;
    if (P.use_ss) {
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) bufA[tid + T * i] = cscale(a[i], cabs2(a[i]));
        __syncthreads();
        float2 *q = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
        float2 Qb[E];
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                const int e = tid + T * i;
                const float2 Q = q[e];
                const double re = (double)yb[i].y * Q.x - (double)yb[i].x * Q.y;   // Re(conj(yb) i Q)
                J.gU[e] += (double)P.gamma * re;
                gsum += (double)P.uppe[e] * re;
                const float s = g * P.uppe[e];
                Qb[i] = make_float2(s * yb[i].y, -s * yb[i].x);   // conj(i gamma U) yb
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) bufA[tid + T * i] = Qb[i];
        __syncthreads();
        float2 *r2 = lds_fft_c<T, -1, n>(bufA, bufB, P.tw, tid);   // n IFFT
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) pb[i] = r2[tid + T * i];
        __syncthreads();
        if (P.use_plasma) {
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) bufA[tid + T * i] = yb[i];
            __syncthreads();
            float2 *r3 = lds_fft_c<T, -1, n>(bufA, bufB, P.tw, tid);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) Pb[i] = r3[tid + T * i];
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) bufA[tid + T * i] = yb[i];
        __syncthreads();
        float2 *r2 = lds_fft_c<T, -1, n>(bufA, bufB, P.tw, tid);   // wb = n IFFT(yb)
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) {
                const float2 w = r2[tid + T * i];
                const float2 p = cscale(a[i], cabs2(a[i]));
                gsum += (double)w.y * p.x - (double)w.x * p.y;     // Re(conj(wb) i p)
                pb[i] = make_float2(g * w.y, -g * w.x);           // conj(i gamma) wb
                Pb[i] = w;
            }
        __syncthreads();
    }
    // p = |a|^2 a:  ab = 2 |a|^2 pb + conj(pb) a^2
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const float m = 2.0f * cabs2(a[i]);
            const float2 a2 = cmul(a[i], a[i]);
            const float2 cp = make_float2(pb[i].x, -pb[i].y);
            ab[i] = cadd(cscale(pb[i], m), cmul(cp, a2));
        }
    if (P.use_plasma) plasma_adjoint_1d<T, NMAX, E>(J, a, Pb, ab, bufA, bufB, scan, tid);
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) bufA[tid + T * i] = ab[i];
    __syncthreads();
    float2 *q2 = lds_fft_c<T, 1, n>(bufA, bufB, P.tw, tid);
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) Fb[i] = cscale(q2[tid + T * i], invn);
    __syncthreads();
    gG += block_sum<T>(gsum, scan + 32, tid);
}

// D-gradient of y = E z, E = exp(h/2 D):  dL/dD += (h/2) yb conj(y)
__device__ __forceinline__ void acc_gD(double *gD, int e, double s, float2 yb, float2 y) {
    gD[2 * e] += s * ((double)yb.x * y.x + (double)yb.y * y.y);
    gD[2 * e + 1] += s * ((double)yb.y * y.x - (double)yb.x * y.y);
}

template <int T, int NMAX, int E>
__device__ void run_adjoint(const Adj1D &J, float2 *bufA, float2 *bufB, double *scan, int tid) {
    const Inst1D &P = J.P;
    constexpr int n = T * E;
    const size_t nn = (size_t)n;
    float2 lam[E];
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) {
            const int e = tid + T * i;
            lam[i] = J.lam[e];
            J.gD[2 * e] = 0.0;
            J.gD[2 * e + 1] = 0.0;
            J.gU[e] = 0.0;
        }
    double gG = 0.0;
    float2 *vb = J.tmp, *a1b = J.tmp + nn, *a2b = J.tmp + 2 * nn, *a3b = J.tmp + 3 * nn;
    const int nseg = (J.M + J.C - 1) / J.C;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int j0 = sg * J.C;
        const int j1 = (j0 + J.C < J.M) ? j0 + J.C : J.M;
        // recompute the segment forward from its checkpoint, storing A, N(A), a2, a3, a4
        float2 A[E], k1[E];
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i < E) A[i] = J.ckpt[(size_t)sg * nn + tid + T * i];
        neval<T, NMAX, E>(P, A, k1, bufA, bufB, scan, tid, nullptr);
        for (int j = j0; j < j1; ++j) {
            float2 *s5 = J.seg + (size_t)(j - j0) * 5 * nn;
            const double h = J.hs[j];
            const float hf = (float)h;
            float2 Ef[E], v[E], S[E], x[E], y[E];
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const int e = tid + T * i;
                    s5[e] = A[i];
                    s5[nn + e] = k1[i];
                    Ef[i] = ipexp(P, e, 0.5 * h);
                    v[i] = cmul(Ef[i], A[i]);
                    const float2 a1 = cmul(Ef[i], k1[i]);
                    S[i] = caxpy(v[i], hf / 6.0f, a1);
                    x[i] = caxpy(v[i], 0.5f * hf, a1);
                }
            neval<T, NMAX, E>(P, x, y, bufA, bufB, scan, tid, nullptr);   // a2
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    s5[2 * nn + tid + T * i] = y[i];
                    S[i] = caxpy(S[i], hf / 3.0f, y[i]);
                    x[i] = caxpy(v[i], 0.5f * hf, y[i]);
                }
            neval<T, NMAX, E>(P, x, y, bufA, bufB, scan, tid, nullptr);   // a3
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    s5[3 * nn + tid + T * i] = y[i];
                    S[i] = caxpy(S[i], hf / 3.0f, y[i]);
                    x[i] = cmul(Ef[i], caxpy(v[i], hf, y[i]));
                }
            neval<T, NMAX, E>(P, x, y, bufA, bufB, scan, tid, nullptr);   // a4
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    s5[4 * nn + tid + T * i] = y[i];
                    A[i] = caxpy(cmul(Ef[i], S[i]), hf / 6.0f, y[i]);
                }
            if (j + 1 < j1) neval<T, NMAX, E>(P, A, k1, bufA, bufB, scan, tid, nullptr);
            __syncthreads();
        }
        // backward through the segment
        for (int j = j1 - 1; j >= j0; --j) {
            const float2 *s5 = J.seg + (size_t)(j - j0) * 5 * nn;
            const double h = J.hs[j];
            const float hf = (float)h;
            const double sh = 0.5 * h;
            float2 Ef[E], t1[E], t2[E];
            // stage A4 = E S + h/6 a4
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const int e = tid + T * i;
                    Ef[i] = ipexp(P, e, sh);
                    const float2 Ai = s5[e], k1i = s5[nn + e], a2 = s5[2 * nn + e],
                                 a3 = s5[3 * nn + e];
                    const float2 v = cmul(Ef[i], Ai), a1 = cmul(Ef[i], k1i);
                    float2 S = caxpy(v, hf / 6.0f, a1);
                    S = caxpy(S, hf / 3.0f, a2);
                    S = caxpy(S, hf / 3.0f, a3);
                    acc_gD(J.gD, e, sh, lam[i], cmul(Ef[i], S));
                    const float2 Sb = cmulc(lam[i], Ef[i]);   // conj(E) lam
                    vb[e] = Sb;
                    a1b[e] = cscale(Sb, hf / 6.0f);
                    a2b[e] = cscale(Sb, hf / 3.0f);
                    a3b[e] = cscale(Sb, hf / 3.0f);
                    t1[i] = cmul(Ef[i], caxpy(v, hf, a3));   // x4
                    t2[i] = cscale(lam[i], hf / 6.0f);        // a4b
                }
            float2 xb[E];
            adj_neval<T, NMAX, E>(J, t1, t2, xb, gG, bufA, bufB, scan, tid);   // x4b
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const int e = tid + T * i;
                    acc_gD(J.gD, e, sh, xb[i], t1[i]);
                    const float2 tb = cmulc(xb[i], Ef[i]);
                    vb[e] = cadd(vb[e], tb);
                    a3b[e] = caxpy(a3b[e], hf, tb);
                    const float2 v = cmul(Ef[i], s5[e]);
                    t1[i] = caxpy(v, 0.5f * hf, s5[2 * nn + e]);   // x3
                    t2[i] = a3b[e];
                }
            adj_neval<T, NMAX, E>(J, t1, t2, xb, gG, bufA, bufB, scan, tid);   // x3b
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const int e = tid + T * i;
                    vb[e] = cadd(vb[e], xb[i]);
                    a2b[e] = caxpy(a2b[e], 0.5f * hf, xb[i]);
                    const float2 v = cmul(Ef[i], s5[e]);
                    t1[i] = caxpy(v, 0.5f * hf, cmul(Ef[i], s5[nn + e]));   // x2
                    t2[i] = a2b[e];
                }
            adj_neval<T, NMAX, E>(J, t1, t2, xb, gG, bufA, bufB, scan, tid);   // x2b
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const int e = tid + T * i;
                    vb[e] = cadd(vb[e], xb[i]);
                    const float2 a1bt = caxpy(a1b[e], 0.5f * hf, xb[i]);
                    const float2 a1 = cmul(Ef[i], s5[nn + e]);
                    acc_gD(J.gD, e, sh, a1bt, a1);
                    t1[i] = s5[e];                      // A
                    t2[i] = cmulc(a1bt, Ef[i]);         // k1b
                }
            adj_neval<T, NMAX, E>(J, t1, t2, xb, gG, bufA, bufB, scan, tid);   // from k1 = N(A)
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i < E) {
                    const int e = tid + T * i;
                    const float2 v = cmul(Ef[i], s5[e]);
                    acc_gD(J.gD, e, sh, vb[e], v);
                    lam[i] = cadd(xb[i], cmulc(vb[e], Ef[i]));
                }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < E; ++i)
        if (i < E) J.lam[tid + T * i] = lam[i];
    if (tid == 0) *J.gG = gG;
}

template <int T, int NMAX, int E>
__device__ __forceinline__ void adjoint_e(const Adj1D &J, float2 *bufA, float2 *bufB,
                                          double *scan, int tid) {
    if constexpr (T * E <= NMAX) run_adjoint<T, NMAX, E>(J, bufA, bufB, scan, tid);
}

template <int T, int NMAX>
__global__ __launch_bounds__(T) void uppe1d_adjoint_kernel(const Adj1D *__restrict__ jobs,
                                                           int count) {
    __shared__ float2 bufA[lds_pad(NMAX)], bufB[lds_pad(NMAX)];
    __shared__ double scan[64];
    for (int k = blockIdx.x; k < count; k += gridDim.x) {
        const Adj1D J = jobs[k];
        if (J.P.nt > NMAX || J.P.nt < T || J.P.nt % T != 0 || J.P.nt / T > MAXE) {
            if (threadIdx.x == 0) atomicOr(J.flags, 2);
            continue;
        }
        switch (J.P.nt / T) {
            case 1: adjoint_e<T, NMAX, 1>(J, bufA, bufB, scan, threadIdx.x); break;
            case 2: adjoint_e<T, NMAX, 2>(J, bufA, bufB, scan, threadIdx.x); break;
            case 4: adjoint_e<T, NMAX, 4>(J, bufA, bufB, scan, threadIdx.x); break;
            default: adjoint_e<T, NMAX, 8>(J, bufA, bufB, scan, threadIdx.x); break;
        }
        __syncthreads();
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Host side.  Workspace layout: [header][Inst1D x n][per-instance tables and state].
namespace {

struct Header1D {
    int n, pad_;
    int counters[30];   // one per concurrent launch slot
};

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

constexpr size_t inst_bytes(int nt) {
    return align256(16 * (size_t)nt) + align256(4 * (size_t)nt) + 5 * align256(8 * (size_t)nt) +
           align256(sizeof(Stat1D));
}

constexpr size_t header_bytes(int n) {
    return align256(sizeof(Header1D)) + align256(sizeof(Inst1D) * n);
}

int check1d(const uppe1d_desc &d) {
    UPPE_CHECK(d.nt >= 256 && d.nt <= 8192 && (d.nt & (d.nt - 1)) == 0,
               "uppe1d: nt = %d unsupported (power of two in [256, 8192])", d.nt);
    UPPE_CHECK(d.tol > 0.0 && d.max_step > 0.0 && d.min_step >= 0.0 && d.L >= 0.0,
               "uppe1d: bad tolerance / step arguments");
    UPPE_CHECK(!d.use_plasma || (d.K >= 1 && d.area > 0.0 && d.rho_nt > 0.0 && d.tau_rec > 0.0),
               "uppe1d: plasma needs K >= 1, area > 0, rho_nt > 0, tau_rec > 0");
    return UPPE_OK;
}

// Scalars and tables of one instance: the tables are written to the host staging area hb in
// the device layout at db (the instance's slice of the workspace); P gets the device
// pointers.  Returns the bytes used (inst_bytes(d.nt)).  D_re .. H_im point at this
// instance's nt values (H may be null).
size_t fill_instance(const uppe1d_desc &d, const double *D_re, const double *D_im,
                     const double *uppe_factor, const double *H_re, const double *H_im,
                     char *hb, char *db, Inst1D &P) {
    const int nt = d.nt;
    memset(&P, 0, sizeof P);
    P.nt = nt;
    P.max_hist = d.max_hist > 0 ? d.max_hist : 0;
    P.max_attempts = d.max_attempts;
    P.use_ss = d.use_ss;
    P.use_raman = d.use_raman && H_re && H_im;
    P.use_plasma = d.use_plasma;
    P.K = d.K;
    P.avalanche = d.avalanche;
    P.gamma = d.gamma;
    P.f_R = d.f_R;
    P.L = d.L;
    P.tol = d.tol;
    P.max_step = d.max_step;
    P.min_step = d.min_step;
    if (d.use_plasma) {
        P.sigma_K = d.sigma_K;
        P.inv_area = 1.0 / d.area;
        P.Ui = d.Ui;
        // the K1 launch casts these to float32 (media_physics.py:185-188)
        P.rho_nt = (double)(float)d.rho_nt;
        P.r_rec = (double)(float)(1.0 / d.tau_rec);
        P.dt = (double)(float)d.dt;
        P.half_sigma = d.sigma_drude / 2.0;
        P.w0tc = d.omega0 * d.tau_c;
        P.av_coef = d.avalanche ? d.sigma_drude / d.Ui : 0.0;
        double gdt = -P.r_rec * P.dt;
        gdt = gdt > 10.0 ? 10.0 : (gdt < -10.0 ? -10.0 : gdt);
        P.e_rec = exp(gdt);
        P.f_rec = fabs(gdt) < 1e-5 ? P.dt * (1.0 + 0.5 * gdt) : (P.e_rec - 1.0) / (-P.r_rec);
    }
    size_t o = 0;
    double *dd = (double *)(hb + o);
    for (int k = 0; k < nt; ++k) {
        dd[k] = D_re[k];
        dd[nt + k] = D_im[k];
    }
    P.D_re = (const double *)(db + o);
    P.D_im = (const double *)(db + o) + nt;
    o += align256(16 * (size_t)nt);
    float *uf = (float *)(hb + o);
    for (int k = 0; k < nt; ++k) uf[k] = (float)uppe_factor[k];
    P.uppe = (const float *)(db + o);
    o += align256(4 * (size_t)nt);
    float2 *hv = (float2 *)(hb + o);
    for (int k = 0; k < nt; ++k)
        hv[k] = P.use_raman ? make_float2((float)H_re[k], (float)H_im[k])
                            : make_float2(1.0f, 0.0f);
    P.H = (const float2 *)(db + o);
    o += align256(8 * (size_t)nt);
    float2 *tv = (float2 *)(hb + o);
    for (int k = 0; k < nt; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)nt;
        tv[k] = make_float2((float)cos(ang), (float)sin(ang));
    }
    P.tw = (const float2 *)(db + o);
    o += align256(8 * (size_t)nt);
    P.A = (float2 *)(db + o);
    o += align256(8 * (size_t)nt);
    P.a5 = (float2 *)(db + o);
    o += align256(8 * (size_t)nt);
    P.v = (float2 *)(db + o);
    o += align256(8 * (size_t)nt);
    P.st = (Stat1D *)(db + o);
    o += align256(sizeof(Stat1D));
    return o;
}

int prepare(int n, const uppe1d_desc *descs, const double *D_re, const double *D_im,
            const double *uppe_factor, const double *H_re, const double *H_im, const void *a0,
            void *workspace, void *hist, void *zhist, void *out_w, const uppe1d_record *rec,
            hipStream_t s) {
    UPPE_CHECK(n > 0 && descs && D_re && D_im && uppe_factor && a0 && workspace,
               "uppe1d: null argument");
    for (int i = 0; i < n; ++i) {
        int r = check1d(descs[i]);
        if (r) return r;
        UPPE_CHECK(descs[i].max_hist <= 0 || (hist && zhist), "uppe1d: history buffers missing");
    }
    char *ws = (char *)workspace;
    std::vector<Inst1D> ins(n);
    size_t total_nt = 0;
    for (int i = 0; i < n; ++i) total_nt += descs[i].nt;
    // all tables of all instances are staged in one host buffer and copied once
    size_t tables = 0;
    for (int i = 0; i < n; ++i) tables += inst_bytes(descs[i].nt);
    std::vector<char> stage(tables, 0);
    size_t off = header_bytes(n), toff = 0, hoff = 0, zoff = 0, loc = 0;
    for (int i = 0; i < n; ++i) {
        const uppe1d_desc &d = descs[i];
        const int nt = d.nt;
        Inst1D &P = ins[i];
        const size_t o = fill_instance(d, D_re + toff, D_im + toff, uppe_factor + toff,
                                       H_re ? H_re + toff : nullptr, H_im ? H_im + toff : nullptr,
                                       stage.data() + loc, ws + off, P);
        off += o;
        loc += o;
        P.a0 = (const float2 *)a0 + toff;
        P.hist = P.max_hist > 0 ? (float2 *)hist + hoff : nullptr;
        P.zhist = P.max_hist > 0 ? (double *)zhist + zoff : nullptr;
        P.outw = out_w ? (float2 *)out_w + toff : nullptr;
        if (rec && rec[i].replay_hs) {
            P.replay_hs = rec[i].replay_hs;
            P.replay_M = rec[i].replay_M;
        }
        if (rec) {
            P.hs = rec[i].hs;
            P.hs_cap = rec[i].hs_cap;
            P.ckpt = (float2 *)rec[i].ckpt;
            P.ckpt_every = rec[i].ckpt_every;
            P.ckpt_cap = rec[i].ckpt_cap;
        }
        toff += nt;
        hoff += (size_t)P.max_hist * nt;
        zoff += P.max_hist;
    }
    Header1D h;
    memset(&h, 0, sizeof h);
    h.n = n;
    UPPE_HIP(hipMemcpyAsync(ws, &h, sizeof h, hipMemcpyHostToDevice, s));
    UPPE_HIP(hipMemcpyAsync(ws + align256(sizeof(Header1D)), ins.data(), sizeof(Inst1D) * n,
                            hipMemcpyHostToDevice, s));
    UPPE_HIP(hipMemcpyAsync(ws + header_bytes(n), stage.data(), tables, hipMemcpyHostToDevice, s));
    UPPE_HIP(hipStreamSynchronize(s));   // the staged host buffers go out of scope
    return UPPE_OK;
}

template <int T>
int launch_t(const Inst1D *insts, int count, int *counter, int lds_cap, hipStream_t s) {
    zero_counter<<<1, 1, 0, s>>>(counter);   // a kernel, not a memset node (graph capture)
    if (lds_cap <= 2048) uppe1d_kernel<T, 2048><<<count, T, 0, s>>>(insts, count, counter);
    else uppe1d_kernel<T, 8192><<<count, T, 0, s>>>(insts, count, counter);
    UPPE_HIP(hipGetLastError());
    return UPPE_OK;
}

int launch(void *workspace, int n, int first, int count, int threads, int lds_cap, int slot,
           hipStream_t s) {
    UPPE_CHECK(workspace && first >= 0 && count > 0 && first + count <= n,
               "uppe1d_batch_launch: bad instance range");
    UPPE_CHECK(threads == 256 || threads == 512 || threads == 1024,
               "uppe1d_batch_launch: threads must be 256, 512 or 1024");
    UPPE_CHECK(slot >= 0 && slot < 30, "uppe1d_batch_launch: slot must be in [0, 30)");
    char *ws = (char *)workspace;
    const Inst1D *insts = (const Inst1D *)(ws + align256(sizeof(Header1D))) + first;
    int *counter = ((Header1D *)ws)->counters + slot;
    switch (threads) {
        case 256: return launch_t<256>(insts, count, counter, lds_cap, s);
        case 512: return launch_t<512>(insts, count, counter, lds_cap, s);
        default: return launch_t<1024>(insts, count, counter, lds_cap, s);
    }
}

void to_status(const Stat1D &x, uppe1d_status &o) {
    o.z = x.z;
    o.h = x.h;
    o.err = x.err;
    o.attempts = x.attempts;
    o.accepted = x.accepted + x.forced;
    o.rejected = x.rejected;
    o.forced = x.forced;
    o.nhist = x.nhist;
    o.done = x.done;
    o.nonfinite = x.nonfinite;
    o.fsal_hits = x.fsal_hits;
    o.overflow = x.hist_overflow | (x.rec_overflow << 1);
}

int status(void *workspace, int n, hipStream_t s, uppe1d_status *st) {
    UPPE_CHECK(workspace && st && n > 0, "uppe1d_batch_status: null argument");
    std::vector<Inst1D> ins(n);
    char *ws = (char *)workspace;
    UPPE_HIP(hipMemcpyAsync(ins.data(), ws + align256(sizeof(Header1D)), sizeof(Inst1D) * n,
                            hipMemcpyDeviceToHost, s));
    UPPE_HIP(hipStreamSynchronize(s));
    std::vector<Stat1D> hs(n);
    for (int i = 0; i < n; ++i)
        UPPE_HIP(hipMemcpyAsync(&hs[i], ins[i].st, sizeof(Stat1D), hipMemcpyDeviceToHost, s));
    UPPE_HIP(hipStreamSynchronize(s));
    int bad = 0, rejected_cfg = 0;
    for (int i = 0; i < n; ++i) {
        to_status(hs[i], st[i]);
        if (hs[i].nonfinite) bad = 1;
        if (hs[i].attempts < 0) rejected_cfg = 1;
    }
    if (rejected_cfg) {
        uppe::set_error("uppe1d: an instance does not fit the launch (nt vs threads / LDS cap)");
        return UPPE_EINVAL;
    }
    if (bad) {
        uppe::set_error("uppe1d: error norm became non-finite (NaN/Inf field)");
        return UPPE_ENONFINITE;
    }
    return UPPE_OK;
}

int default_threads(int nt) { return nt <= 2048 ? 256 : nt / MAXE; }

static_assert(sizeof(Slice1D) == sizeof(uppe1d_slice), "slice layout");

size_t chain_bytes(int nt, int nslices) {
    return align256(sizeof(Chain1D)) + align256(sizeof(Slice1D) * (size_t)nslices) +
           align256(sizeof(Stat1D) * (size_t)nslices) + align256(8 * (size_t)nt);
}

int propagate_slices(const uppe1d_desc *d, int nslices, const uppe1d_slice *slices,
                     const double *D_re, const double *D_im, const double *uppe_factor,
                     const double *H_re, const double *H_im, const void *a0, void *workspace,
                     void *hist, void *zhist, void *out_t, void *out_w, hipStream_t s,
                     uppe1d_status *st) {
    UPPE_CHECK(d && slices && st && nslices > 0, "uppe1d_propagate_slices: null argument");
    for (int i = 0; i < nslices; ++i) {
        const uppe1d_slice &c = slices[i];
        UPPE_CHECK(c.tol > 0.0 && c.max_step > 0.0 && c.min_step >= 0.0 && c.L >= 0.0,
                   "uppe1d_propagate_slices: slice %d has bad tolerance / step arguments", i);
    }
    // the instance (tables, state, history pointers) as for one propagation
    int r = prepare(1, d, D_re, D_im, uppe_factor, H_re, H_im, a0, workspace, hist, zhist,
                    out_w, nullptr, s);
    if (r) return r;
    char *ws = (char *)workspace;
    Chain1D C;
    memset(&C, 0, sizeof C);
    UPPE_HIP(hipMemcpy(&C.P, ws + align256(sizeof(Header1D)), sizeof(Inst1D),
                       hipMemcpyDeviceToHost));
    char *cb = ws + header_bytes(1) + inst_bytes(d->nt);
    size_t o = align256(sizeof(Chain1D));
    C.slices = (const Slice1D *)(cb + o);
    o += align256(sizeof(Slice1D) * (size_t)nslices);
    C.slice_st = (Stat1D *)(cb + o);
    o += align256(sizeof(Stat1D) * (size_t)nslices);
    C.lastrec = (float2 *)(cb + o);
    C.out_t = (float2 *)out_t;
    C.nslices = nslices;
    std::vector<Stat1D> zero(nslices);
    memset(zero.data(), 0, sizeof(Stat1D) * zero.size());
    UPPE_HIP(hipMemcpyAsync(cb, &C, sizeof C, hipMemcpyHostToDevice, s));
    UPPE_HIP(hipMemcpyAsync((void *)C.slices, slices, sizeof(Slice1D) * nslices,
                            hipMemcpyHostToDevice, s));
    UPPE_HIP(hipMemcpyAsync(C.slice_st, zero.data(), sizeof(Stat1D) * nslices,
                            hipMemcpyHostToDevice, s));
    const Chain1D *cd = (const Chain1D *)cb;
    switch (default_threads(d->nt)) {
        case 256:
            if (d->nt <= 2048) uppe1d_chain_kernel<256, 2048><<<1, 256, 0, s>>>(cd);
            else uppe1d_chain_kernel<256, 8192><<<1, 256, 0, s>>>(cd);
            break;
        case 512: uppe1d_chain_kernel<512, 8192><<<1, 512, 0, s>>>(cd); break;
        default: uppe1d_chain_kernel<1024, 8192><<<1, 1024, 0, s>>>(cd); break;
    }
    UPPE_HIP(hipGetLastError());
    std::vector<Stat1D> hs(nslices);
    UPPE_HIP(hipMemcpyAsync(hs.data(), C.slice_st, sizeof(Stat1D) * nslices,
                            hipMemcpyDeviceToHost, s));
    UPPE_HIP(hipStreamSynchronize(s));
    UPPE_CHECK(hs[0].attempts >= 0,
               "uppe1d_propagate_slices: the instance does not fit the launch (nt vs threads)");
    int bad = 0;
    for (int i = 0; i < nslices; ++i) {
        uppe1d_status &q = st[i];
        const Stat1D &x = hs[i];
        q.z = x.z;
        q.h = x.h;
        q.err = x.err;
        q.attempts = x.attempts;
        q.accepted = x.accepted + x.forced;
        q.rejected = x.rejected;
        q.forced = x.forced;
        q.nhist = x.nhist;
        q.done = x.done;
        q.nonfinite = x.nonfinite;
        q.fsal_hits = x.fsal_hits;
        q.overflow = x.hist_overflow;
        if (x.nonfinite) bad = 1;
    }
    if (bad) {
        uppe::set_error("uppe1d: error norm became non-finite (NaN/Inf field)");
        return UPPE_ENONFINITE;
    }
    return UPPE_OK;
}

size_t adj_bytes(int nt, int ckpt_every) {
    const size_t n = (size_t)nt;
    return align256(sizeof(Adj1D)) + align256(5 * (size_t)ckpt_every * n * 8) +
           align256(4 * n * 8) + align256(4 * n * 8) + align256(sizeof(int));
}

int adjoint(void *fwd_ws, int n, int index, int M, const double *hs, const void *ckpt,
            int ckpt_every, void *lam, void *adj_ws, double *gD, double *gU, double *gG,
            hipStream_t s) {
    UPPE_CHECK(fwd_ws && adj_ws && hs && ckpt && lam && gD && gU && gG,
               "uppe1d_adjoint: null argument");
    UPPE_CHECK(index >= 0 && index < n && M >= 0 && ckpt_every > 0, "uppe1d_adjoint: bad range");
    Inst1D P;
    UPPE_HIP(hipMemcpyAsync(&P, (char *)fwd_ws + align256(sizeof(Header1D)) + index * sizeof(Inst1D),
                            sizeof(Inst1D), hipMemcpyDeviceToHost, s));
    UPPE_HIP(hipStreamSynchronize(s));
    const size_t nt = (size_t)P.nt;
    Adj1D J;
    memset(&J, 0, sizeof J);
    J.P = P;
    J.hs = hs;
    J.ckpt = (const float2 *)ckpt;
    J.M = M;
    J.C = ckpt_every;
    char *w = (char *)adj_ws;
    size_t o = align256(sizeof(Adj1D));
    J.seg = (float2 *)(w + o);
    o += align256(5 * (size_t)ckpt_every * nt * 8);
    J.tmp = (float2 *)(w + o);
    o += align256(4 * nt * 8);
    J.pscr = (double *)(w + o);
    o += align256(4 * nt * 8);
    J.flags = (int *)(w + o);
    J.lam = (float2 *)lam;
    J.gD = gD;
    J.gU = gU;
    J.gG = gG;
    int zero = 0;
    UPPE_HIP(hipMemcpyAsync(J.flags, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    UPPE_HIP(hipMemcpyAsync(w, &J, sizeof J, hipMemcpyHostToDevice, s));
    const int T = default_threads(P.nt);
    const Adj1D *jd = (const Adj1D *)w;
    if (M > 0) {
        switch (T) {
            case 256:
                if (P.nt <= 2048) uppe1d_adjoint_kernel<256, 2048><<<1, 256, 0, s>>>(jd, 1);
                else uppe1d_adjoint_kernel<256, 8192><<<1, 256, 0, s>>>(jd, 1);
                break;
            case 512: uppe1d_adjoint_kernel<512, 8192><<<1, 512, 0, s>>>(jd, 1); break;
            default: uppe1d_adjoint_kernel<1024, 8192><<<1, 1024, 0, s>>>(jd, 1); break;
        }
        UPPE_HIP(hipGetLastError());
    }
    int flags = 0;
    UPPE_HIP(hipMemcpyAsync(&flags, J.flags, sizeof(int), hipMemcpyDeviceToHost, s));
    UPPE_HIP(hipStreamSynchronize(s));
    UPPE_CHECK(!(flags & 2), "uppe1d_adjoint: instance does not fit the launch configuration");
    UPPE_CHECK(!(flags & 1),
               "uppe1d_adjoint: plasma outside the differentiable regime (avalanche rate or a "
               "density clamp is active)");
    return UPPE_OK;
}

// ---------------------------------------------------------------------------------------
// Allocation-free single-pulse propagation (uppe1d_propagate_noalloc).  The workspace of one
// instance (header, descriptor, tables, state, status) lives in a file-scope device array of
// the largest supported length; the host staging of the tables in a file-scope host array.
// The device address of the array is queried once per device with hipGetSymbolAddress (the
// array's name is not a device address in host code) and kept per device.
constexpr int NOALLOC_MAX_NT = 8192;
constexpr size_t NOALLOC_BYTES = header_bytes(1) + inst_bytes(NOALLOC_MAX_NT);
constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

std::mutex g_noalloc_mu;                                  // serialises the host staging
alignas(256) unsigned char g_noalloc_stage[inst_bytes(NOALLOC_MAX_NT)];
void *g_noalloc_dev[NOALLOC_MAX_DEVICES];                  // device address of scratch

int noalloc_scratch(void **ptr) {
    int dev = 0;
    UPPE_HIP(hipGetDevice(&dev));
    UPPE_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "uppe1d_propagate_noalloc: device %d", dev);
    if (!g_noalloc_dev[dev]) {
        void *p = nullptr;
        UPPE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_dev[dev] = p;
    }
    *ptr = g_noalloc_dev[dev];
    return UPPE_OK;
}

int propagate_noalloc(const uppe1d_desc *d, const double *D_re, const double *D_im,
                      const double *uppe_factor, const double *H_re, const double *H_im,
                      const void *a0, void *hist, void *zhist, void *out_w,
                      const uppe1d_record *rec, hipStream_t s, uppe1d_status *st) {
    UPPE_CHECK(d && st && D_re && D_im && uppe_factor && a0, "uppe1d_propagate_noalloc: null argument");
    UPPE_CHECK(d->nt <= NOALLOC_MAX_NT,
               "uppe1d_propagate_noalloc: nt = %d exceeds the static workspace (at most %d)",
               d->nt, NOALLOC_MAX_NT);
    int r = check1d(*d);
    if (r) return r;
    UPPE_CHECK(d->max_hist <= 0 || (hist && zhist), "uppe1d: history buffers missing");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    void *sp = nullptr;
    r = noalloc_scratch(&sp);
    if (r) return r;
    char *ws = (char *)sp;
    Inst1D P;
    const size_t tables = fill_instance(*d, D_re, D_im, uppe_factor, H_re, H_im,
                                        (char *)g_noalloc_stage, ws + header_bytes(1), P);
    P.a0 = (const float2 *)a0;
    P.hist = P.max_hist > 0 ? (float2 *)hist : nullptr;
    P.zhist = P.max_hist > 0 ? (double *)zhist : nullptr;
    P.outw = (float2 *)out_w;
    if (rec) {
        P.replay_hs = rec->replay_hs;
        P.replay_M = rec->replay_hs ? rec->replay_M : 0;
        P.hs = rec->hs;
        P.hs_cap = rec->hs_cap;
        P.ckpt = (float2 *)rec->ckpt;
        P.ckpt_every = rec->ckpt_every;
        P.ckpt_cap = rec->ckpt_cap;
    }
    Header1D h;
    memset(&h, 0, sizeof h);
    h.n = 1;
    UPPE_HIP(hipMemcpyAsync(ws, &h, sizeof h, hipMemcpyHostToDevice, s));
    UPPE_HIP(hipMemcpyAsync(ws + align256(sizeof(Header1D)), &P, sizeof P, hipMemcpyHostToDevice,
                            s));
    UPPE_HIP(hipMemcpyAsync(ws + header_bytes(1), g_noalloc_stage, tables,
                            hipMemcpyHostToDevice, s));
    // the instance counter is cleared by a kernel inside launch(); the status block is
    // written whole by the kernel
    r = launch(ws, 1, 0, 1, default_threads(d->nt), d->nt, 0, s);
    if (r) return r;
    Stat1D x;
    UPPE_HIP(hipMemcpyAsync(&x, P.st, sizeof x, hipMemcpyDeviceToHost, s));
    UPPE_HIP(hipStreamSynchronize(s));
    UPPE_CHECK(x.attempts >= 0, "uppe1d: the instance does not fit the launch (nt vs threads)");
    to_status(x, *st);
    if (x.nonfinite) {
        uppe::set_error("uppe1d: error norm became non-finite (NaN/Inf field)");
        return UPPE_ENONFINITE;
    }
    return UPPE_OK;
}

}  // namespace

extern "C" {

int uppe1d_workspace_size(const uppe1d_desc *d, size_t *bytes) {
    UPPE_CHECK(d && bytes, "uppe1d_workspace_size: null argument");
    *bytes = header_bytes(1) + inst_bytes(d->nt);
    return UPPE_OK;
}

int uppe1d_batch_workspace_size(int n, const uppe1d_desc *descs, size_t *bytes) {
    UPPE_CHECK(n > 0 && descs && bytes, "uppe1d_batch_workspace_size: null argument");
    size_t b = header_bytes(n);
    for (int i = 0; i < n; ++i) b += inst_bytes(descs[i].nt);
    *bytes = b;
    return UPPE_OK;
}

int uppe1d_batch_prepare(int n, const uppe1d_desc *descs, const double *D_re,
                         const double *D_im, const double *uppe_factor, const double *H_re,
                         const double *H_im, const void *a0, void *workspace, void *hist,
                         void *zhist, void *out_w, const uppe1d_record *rec, void *stream) {
    return prepare(n, descs, D_re, D_im, uppe_factor, H_re, H_im, a0, workspace, hist, zhist,
                   out_w, rec, (hipStream_t)stream);
}

int uppe1d_batch_launch(void *workspace, int n, int first, int count, int threads, int lds_cap,
                        int slot, void *stream) {
    return launch(workspace, n, first, count, threads, lds_cap, slot, (hipStream_t)stream);
}

int uppe1d_batch_status(void *workspace, int n, void *stream, uppe1d_status *st) {
    return status(workspace, n, (hipStream_t)stream, st);
}

int uppe1d_adjoint_workspace_size(int nt, int ckpt_every, size_t *bytes) {
    UPPE_CHECK(bytes && nt > 0 && ckpt_every > 0, "uppe1d_adjoint_workspace_size: bad argument");
    *bytes = adj_bytes(nt, ckpt_every);
    return UPPE_OK;
}

int uppe1d_adjoint(void *fwd_workspace, int n, int index, int M, const double *hs,
                   const void *ckpt, int ckpt_every, void *lam, void *adj_workspace, double *gD,
                   double *gU, double *gG, void *stream) {
    return adjoint(fwd_workspace, n, index, M, hs, ckpt, ckpt_every, lam, adj_workspace, gD, gU,
                   gG, (hipStream_t)stream);
}

int uppe1d_slices_workspace_size(const uppe1d_desc *d, int nslices, size_t *bytes) {
    UPPE_CHECK(d && bytes && nslices > 0, "uppe1d_slices_workspace_size: bad argument");
    *bytes = header_bytes(1) + inst_bytes(d->nt) + chain_bytes(d->nt, nslices);
    return UPPE_OK;
}

int uppe1d_propagate_slices(const uppe1d_desc *d, int nslices, const uppe1d_slice *slices,
                            const double *D_re, const double *D_im, const double *uppe_factor,
                            const double *H_re, const double *H_im, const void *a0,
                            void *workspace, void *hist, void *zhist, void *out_t, void *out_w,
                            void *stream, uppe1d_status *st) {
    return propagate_slices(d, nslices, slices, D_re, D_im, uppe_factor, H_re, H_im, a0,
                            workspace, hist, zhist, out_t, out_w, (hipStream_t)stream, st);
}

int uppe1d_propagate_noalloc(const uppe1d_desc *d, const double *D_re, const double *D_im,
                             const double *uppe_factor, const double *H_re, const double *H_im,
                             const void *a0, void *hist, void *zhist, void *out_w,
                             const uppe1d_record *rec, void *stream, uppe1d_status *st) {
    return propagate_noalloc(d, D_re, D_im, uppe_factor, H_re, H_im, a0, hist, zhist, out_w, rec,
                             (hipStream_t)stream, st);
}

int uppe1d_propagate(const uppe1d_desc *d, const double *D_re, const double *D_im,
                     const double *uppe_factor, const double *H_re, const double *H_im,
                     const void *a0, void *workspace, void *hist, void *zhist, void *out_w,
                     const uppe1d_record *rec, void *stream, uppe1d_status *st) {
    UPPE_CHECK(d != nullptr && st != nullptr, "uppe1d_propagate: null argument");
    hipStream_t s = (hipStream_t)stream;
    int r = prepare(1, d, D_re, D_im, uppe_factor, H_re, H_im, a0, workspace, hist, zhist, out_w,
                    rec, s);
    if (r) return r;
    r = launch(workspace, 1, 0, 1, default_threads(d->nt), d->nt, 0, s);
    if (r) return r;
    return status(workspace, 1, s, st);
}

}  // extern "C"
