// Shared device helpers for the UPPE kernels: complex arithmetic, small in-register DFTs
// and one Stockham autosort stage.  Written for gfx950 (wave64); no other target.
//
// Conventions: forward transforms use exp(-2*pi*i*n*k/N) and are unnormalised; inverse
// transforms use exp(+2*pi*i*n*k/N) and the caller applies 1/N (numpy / scipy "backward"
// norm, the one the reference relies on).  Twiddle tables hold W_N^k = exp(-2*pi*i*k/N),
// k = 0..N-1, computed in FP64 on the host and rounded to float.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace uppe {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a + s * b
__device__ __forceinline__ float2 caxpy(float2 a, float s, float2 b) {
    return make_float2(fmaf(s, b.x, a.x), fmaf(s, b.y, a.y));
}
__device__ __forceinline__ float cabs2(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }
// multiply by -i (DIR=+1, forward) or +i (DIR=-1, inverse)
template <int DIR>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return DIR > 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// cos / sin of pi*m/16, m = 0..15 (decimal expansions rounded to float)
__device__ __forceinline__ float c32(int m) {
    const float t[16] = {1.0f, 0.98078528040323044913f, 0.92387953251128675613f,
                         0.83146961230254523708f, 0.70710678118654752440f,
                         0.55557023301960222474f, 0.38268343236508977173f,
                         0.19509032201612826785f, 0.0f, -0.19509032201612826785f,
                         -0.38268343236508977173f, -0.55557023301960222474f,
                         -0.70710678118654752440f, -0.83146961230254523708f,
                         -0.92387953251128675613f, -0.98078528040323044913f};
    return t[m];
}
__device__ __forceinline__ float s32(int m) {
    const float t[16] = {0.0f, 0.19509032201612826785f, 0.38268343236508977173f,
                         0.55557023301960222474f, 0.70710678118654752440f,
                         0.83146961230254523708f, 0.92387953251128675613f,
                         0.98078528040323044913f, 1.0f, 0.98078528040323044913f,
                         0.92387953251128675613f, 0.83146961230254523708f,
                         0.70710678118654752440f, 0.55557023301960222474f,
                         0.38268343236508977173f, 0.19509032201612826785f};
    return t[m];
}

// In-place DFT of R values held in registers, natural order in and out.
// Radix-2 decimation in time, fully unrolled at compile time.
template <int R, int DIR>
struct Dft {
    __device__ __forceinline__ static void run(float2 (&v)[R]) {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        Dft<R / 2, DIR>::run(e);
        Dft<R / 2, DIR>::run(o);
#pragma unroll
        for (int k = 0; k < R / 2; ++k) {
            // W_R^k = exp(-2 pi i k / R) = cos(pi m / 16) - i sin(pi m / 16), m = 32 k / R
            static_assert(R <= 32, "register DFT supports R <= 32");
            const int m = k * (32 / R);
            float2 w = make_float2(c32(m), -DIR * s32(m));
            float2 t = (k == 0) ? o[k] : cmul(o[k], w);
            v[k] = cadd(e[k], t);
            v[k + R / 2] = csub(e[k], t);
        }
    }
};
template <int DIR>
struct Dft<1, DIR> {
    __device__ __forceinline__ static void run(float2 (&)[1]) {}
};
template <int DIR>
struct Dft<2, DIR> {
    __device__ __forceinline__ static void run(float2 (&v)[2]) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <int DIR>
struct Dft<4, DIR> {
    __device__ __forceinline__ static void run(float2 (&v)[4]) {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

// Twiddle a butterfly's inputs for Stockham stage (N, Ns, R): v[r] *= W_N^(k*r*N/(Ns*R)),
// conjugated for the inverse.  tw holds W_N^m, m < N.
template <int R, int DIR>
__device__ __forceinline__ void twiddle(float2 (&v)[R], int k, int Ns, int N,
                                        const float2* __restrict__ tw) {
    if (Ns == 1) return;
    const int step = k * (N / (Ns * R));
#pragma unroll
    for (int r = 1; r < R; ++r) {
        float2 w = tw[step * r];
        v[r] = DIR > 0 ? cmul(v[r], w) : cmulc(v[r], w);
    }
}

// One radix-R Stockham butterfly j of a length-N transform at stage Ns:
//   inputs  x[j + r*N/R]            (r = 0..R-1, loaded by the caller into v)
//   outputs y[(j/Ns)*Ns*R + j%Ns + r*Ns]   (the caller stores v there)
// Returns the output base index; output r goes to base + r*Ns.
template <int R, int DIR>
__device__ __forceinline__ int stockham_butterfly(float2 (&v)[R], int j, int Ns, int N,
                                                  const float2* __restrict__ tw) {
    const int k = j % Ns;
    twiddle<R, DIR>(v, k, Ns, N, tw);
    Dft<R, DIR>::run(v);
    return (j / Ns) * Ns * R + k;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace uppe
