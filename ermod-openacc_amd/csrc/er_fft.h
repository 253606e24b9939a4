// er_fft.h -- register-resident mixed-radix (2/3/4/5) Stockham FFT for one
// line per thread, fully unrolled at compile time so that every element and
// every twiddle index is a constant: the line lives in VGPRs and the
// twiddles are uniform loads from a per-length table tw[k] = exp(+2 pi i k/N)
// (computed on the host in double precision, rounded once to float).
//
// The PME transforms of erdst (recpcal.F90:401,425,475,499 through
// fft_iface.F90:66-98) are unnormalised; so are these.  SIGN = -1 is the
// forward e^{-i} transform, SIGN = +1 the backward e^{+i} one.
#pragma once
#include <hip/hip_runtime.h>

namespace erfft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)  // a * conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by SIGN * i
template <int SIGN> __device__ __forceinline__ float2 mul_si(float2 a)
{
    return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// a * w^SIGN for a table entry w = exp(+2 pi i e/N)
template <int SIGN> __device__ __forceinline__ float2 twmul(float2 a, float2 w)
{
    return SIGN > 0 ? cmul(a, w) : cmulc(a, w);
}

// y_k = sum_t v_t exp(SIGN 2 pi i t k / R)
template <int R, int SIGN> struct Dft;

template <int SIGN> struct Dft<2, SIGN> {
    __device__ __forceinline__ static void run(float2 *v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int SIGN> struct Dft<3, SIGN> {
    __device__ __forceinline__ static void run(float2 *v)
    {
        const float h = 0.866025403784438646763723170752936183f;  // sin(2 pi/3)
        float2 s = cadd(v[1], v[2]);
        float2 d = csub(v[1], v[2]);
        float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
        float2 r = mul_si<SIGN>(cscale(d, h));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <int SIGN> struct Dft<4, SIGN> {
    __device__ __forceinline__ static void run(float2 *v)
    {
        float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        float2 c = cadd(v[1], v[3]), d = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <int SIGN> struct Dft<5, SIGN> {
    __device__ __forceinline__ static void run(float2 *v)
    {
        const float c1 = 0.309016994374947424102293417182819059f;   // cos(2 pi/5)
        const float c2 = -0.809016994374947424102293417182819059f;  // cos(4 pi/5)
        const float s1 = 0.951056516295153572116439333379382143f;   // sin(2 pi/5)
        const float s2 = 0.587785252292473129168705954639072769f;   // sin(4 pi/5)
        float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        float2 x0 = v[0];
        float2 t1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
        float2 t2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
        float2 u1 = mul_si<SIGN>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
        float2 u2 = mul_si<SIGN>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
        v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
        v[1] = cadd(t1, u1);
        v[4] = csub(t1, u1);
        v[2] = cadd(t2, u2);
        v[3] = csub(t2, u2);
    }
};

// One Stockham autosort stage of radix R with current sub-transform size NS,
// then the remaining stages.
template <int N, int NS, int SIGN, int... Rs> struct Stages;

template <int N, int NS, int SIGN> struct Stages<N, NS, SIGN> {
    __device__ __forceinline__ static void run(float2 (&)[N], const float2 *__restrict__) {}
};

template <int N, int NS, int SIGN, int R, int... Rest> struct Stages<N, NS, SIGN, R, Rest...> {
    __device__ __forceinline__ static void run(float2 (&a)[N], const float2 *__restrict__ tw)
    {
        float2 b[N];
#pragma unroll
        for (int j = 0; j < N / R; ++j) {
            const int k = j % NS;
            float2 v[R];
#pragma unroll
            for (int t = 0; t < R; ++t) v[t] = a[j + t * (N / R)];
#pragma unroll
            for (int t = 1; t < R; ++t) {
                const int e = (t * k * (N / (NS * R))) % N;
                // half and quarter turns are exact sign and component swaps
                if (2 * e == N) v[t] = make_float2(-v[t].x, -v[t].y);
                else if (4 * e == N) v[t] = mul_si<SIGN>(v[t]);
                else if (4 * e == 3 * N) v[t] = mul_si<-SIGN>(v[t]);
                else if (e != 0) v[t] = twmul<SIGN>(v[t], tw[e]);
            }
            Dft<R, SIGN>::run(v);
            const int idx = (j / NS) * NS * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) b[idx + t * NS] = v[t];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) a[i] = b[i];
        Stages<N, NS * R, SIGN, Rest...>::run(a, tw);
    }
};

// Radix plans for the supported line lengths (radix 4 first: fewest stages).
template <int N> struct Plan;
#define ER_FFT_PLAN(N, ...)                                                        \
    template <> struct Plan<N> {                                                   \
        template <int SIGN>                                                        \
        __device__ __forceinline__ static void run(float2 (&a)[N], const float2 *tw) \
        {                                                                          \
            Stages<N, 1, SIGN, __VA_ARGS__>::run(a, tw);                           \
        }                                                                          \
    };
ER_FFT_PLAN(4, 4)
ER_FFT_PLAN(5, 5)
ER_FFT_PLAN(6, 2, 3)
ER_FFT_PLAN(8, 4, 2)
ER_FFT_PLAN(9, 3, 3)
ER_FFT_PLAN(10, 2, 5)
ER_FFT_PLAN(12, 4, 3)
ER_FFT_PLAN(16, 4, 4)
ER_FFT_PLAN(18, 2, 3, 3)
ER_FFT_PLAN(20, 4, 5)
ER_FFT_PLAN(24, 4, 2, 3)
ER_FFT_PLAN(27, 3, 3, 3)
ER_FFT_PLAN(30, 2, 3, 5)
ER_FFT_PLAN(32, 4, 4, 2)
ER_FFT_PLAN(36, 4, 3, 3)
ER_FFT_PLAN(40, 4, 2, 5)
ER_FFT_PLAN(48, 4, 4, 3)
ER_FFT_PLAN(54, 2, 3, 3, 3)
ER_FFT_PLAN(60, 4, 3, 5)
ER_FFT_PLAN(64, 4, 4, 4)
ER_FFT_PLAN(72, 4, 2, 3, 3)
ER_FFT_PLAN(80, 4, 4, 5)
#undef ER_FFT_PLAN

template <int N, int SIGN> __device__ __forceinline__ void fft(float2 (&a)[N], const float2 *tw)
{
    Plan<N>::template run<SIGN>(a, tw);
}

}  // namespace erfft
