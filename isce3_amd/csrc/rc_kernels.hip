// Range compression on gfx950: frequency-domain matched filtering of a batch
// of complex64 pulses (the operation of cxx/isce3/focus/RangeComp.cpp:93-131),
// one workgroup per pulse and one pass over HBM.
//
// A workgroup of L/16 threads loads its pulse zero-padded to L = 2^p samples
// into LDS, transforms it with a Stockham FFT (radix-16 passes and one
// radix-2/4/8 pass, each thread holding 16 values per pass), multiplies by
// the matched filter's spectrum (1/L folded in), conjugates, transforms again
// (conj(FFT(conj(X))) = L IFFT(X)) and writes the conjugated crop of the
// mode.  Twiddles W_L^k come from a table in global memory (L2 resident);
// radix-16/8 butterflies read W^t and W^(4t) and form the other powers by at
// most two complex products.  LDS addresses are padded by one slot every 16
// to spread the strided stores of the early passes over the banks.
#include <hip/hip_runtime.h>

#include "rc_internal.h"

namespace rc {
namespace {

__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (-i)
__device__ inline float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ inline float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

__device__ inline int pad(int i) { return i + (i >> 4); }

// Native 2 x float vector for the non-temporal stores of the compressed
// pulses (written once, not read again by the kernel).
typedef float v2f __attribute__((ext_vector_type(2)));

// Forward DFTs (sign -1) of R values in registers.
__device__ inline void dft2(float2& a, float2& b)
{
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

__device__ inline void dft4(float2& x0, float2& x1, float2& x2, float2& x3)
{
    const float2 a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const float2 a2 = cadd(x1, x3), a3 = mul_mi(csub(x1, x3));
    x0 = cadd(a0, a2);
    x2 = csub(a0, a2);
    x1 = cadd(a1, a3);
    x3 = csub(a1, a3);
}

template<int R>
__device__ inline void dft(float2 (&v)[R]);

template<>
__device__ inline void dft<2>(float2 (&v)[2])
{
    dft2(v[0], v[1]);
}

template<>
__device__ inline void dft<4>(float2 (&v)[4])
{
    dft4(v[0], v[1], v[2], v[3]);
}

template<>
__device__ inline void dft<8>(float2 (&v)[8])
{
    constexpr float c = 0.70710678118654752f;
    // even and odd halves
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    // W8^k o_k
    o1 = make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
    o2 = mul_mi(o2);
    o3 = make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

template<>
__device__ inline void dft<16>(float2 (&v)[16])
{
    // n = 4 n1 + n2: DFT4 over n1, twiddle W16^(k1 n2), DFT4 over n2;
    // output k = k1 + 4 k2
    constexpr float c = 0.70710678118654752f;   // cos(pi/4)
    constexpr float c1 = 0.92387953251128676f;  // cos(pi/8)
    constexpr float s1 = 0.38268343236508977f;  // sin(pi/8)
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // v[4 k1 + n2] now holds A[k1][n2]; multiply by W16^(k1 n2) = exp(-i pi k1 n2 / 8)
    auto rot = [](float2 a, float cs, float sn) {  // a * (cs - i sn)
        return make_float2(a.x * cs + a.y * sn, a.y * cs - a.x * sn);
    };
    v[5] = rot(v[5], c1, s1);     // k1=1, n2=1: W^1
    v[6] = rot(v[6], c, c);       // W^2
    v[7] = rot(v[7], s1, c1);     // W^3
    v[9] = rot(v[9], c, c);       // k1=2, n2=1: W^2
    v[10] = mul_mi(v[10]);        // W^4
    v[11] = rot(v[11], -c, c);    // W^6
    v[13] = rot(v[13], s1, c1);   // k1=3, n2=1: W^3
    v[14] = rot(v[14], -c, c);    // W^6
    v[15] = rot(v[15], -c1, -s1); // W^9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    // v[4 k1 + k2] holds y[k1 + 4 k2]: transpose the 4 x 4 block
    float2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = v[i];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[4 * k1 + k2];
}

// v[r] *= W^(r t) with W = exp(-2 pi i / L), from the table tw[k] = W^k.
template<int R>
__device__ inline void twiddle(float2 (&v)[R], const float2* __restrict__ tw, int t)
{
    const float2 w1 = tw[t];
    if (R == 2) {
        v[1] = cmul(v[1], w1);
        return;
    }
    const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    if (R == 4) {
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        return;
    }
    const float2 w4 = tw[4 * t];
    float2 w[R];
    w[1] = w1;
    w[2] = w2;
    w[3] = w3;
    w[4] = w4;
#pragma unroll
    for (int r = 5; r < 8 && r < R; ++r) w[r] = cmul(w4, w[r - 4]);
    if (R == 16) {
        const float2 w8 = cmul(w4, w4), w12 = cmul(w8, w4);
        w[8] = w8;
        w[12] = w12;
#pragma unroll
        for (int r = 1; r < 4; ++r) {
            w[8 + r] = cmul(w8, w[r]);
            w[12 + r] = cmul(w12, w[r]);
        }
    }
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[r]);
}

// One Stockham pass of radix R at sub-transform size Ns over L values in
// padded LDS, T = L / 16 threads, each taking 16 / R butterflies.
template<int R, int L>
__device__ inline void stockham_pass(float2* s, int Ns, const float2* __restrict__ tw)
{
    constexpr int T = L / 16;
    constexpr int NB = L / R;
    constexpr int Q = NB / T;
    float2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = threadIdx.x + q * T;
#pragma unroll
        for (int r = 0; r < R; ++r) v[q][r] = s[pad(j + r * NB)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = threadIdx.x + q * T;
        const int k = j & (Ns - 1);
        if (Ns > 1) twiddle<R>(v[q], tw, k * (L / (Ns * R)));
        dft<R>(v[q]);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) s[pad(base + r * Ns)] = v[q][r];
    }
    __syncthreads();
}

// Forward FFT of L = 2^LOG2L (10 <= LOG2L <= 14) values in padded LDS.
template<int LOG2L>
__device__ inline void fft(float2* s, const float2* __restrict__ tw)
{
    constexpr int L = 1 << LOG2L;
    constexpr int NR16 = LOG2L / 4;
    constexpr int REM = LOG2L % 4;
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < NR16; ++p) {
        stockham_pass<16, L>(s, Ns, tw);
        Ns *= 16;
    }
    if (REM == 1) stockham_pass<2, L>(s, Ns, tw);
    if (REM == 2) stockham_pass<4, L>(s, Ns, tw);
    if (REM == 3) stockham_pass<8, L>(s, Ns, tw);
}

// A Stockham pass split into its parts so that the first pass can read from
// global memory and the last pass of the forward transform can hand its
// outputs to the first pass of the inverse transform in registers.  Thread t
// holds butterflies j = t + q T (q < 16 / R), element r of butterfly j being
// sample j + r L / R on input; outputs go to (j - k) R + k + r Ns, k = j mod Ns.
template<int R, int L>
struct Regs {
    static constexpr int T = L / 16;
    static constexpr int NB = L / R;
    static constexpr int Q = NB / T;
    float2 v[Q][R];
};

template<int R, int L>
__device__ inline void pass_load_lds(Regs<R, L>& d, const float2* s)
{
#pragma unroll
    for (int q = 0; q < d.Q; ++q) {
        const int j = threadIdx.x + q * d.T;
#pragma unroll
        for (int r = 0; r < R; ++r) d.v[q][r] = s[pad(j + r * d.NB)];
    }
}

// Samples g0 + i of x (zero outside [0, n)).
template<int R, int L>
__device__ inline void pass_load_global(Regs<R, L>& d, const float2* x, int64_t g0, int64_t n)
{
#pragma unroll
    for (int q = 0; q < d.Q; ++q) {
        const int j = threadIdx.x + q * d.T;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t g = g0 + j + r * d.NB;
            d.v[q][r] = (g >= 0 && g < n) ? x[g] : make_float2(0.f, 0.f);
        }
    }
}

template<int R, int L>
__device__ inline void pass_compute(Regs<R, L>& d, int Ns, const float2* __restrict__ tw)
{
#pragma unroll
    for (int q = 0; q < d.Q; ++q) {
        const int j = threadIdx.x + q * d.T;
        const int k = j & (Ns - 1);
        if (Ns > 1) twiddle<R>(d.v[q], tw, k * (L / (Ns * R)));
        dft<R>(d.v[q]);
    }
}

template<int R, int L>
__device__ inline void pass_store_lds(const Regs<R, L>& d, float2* s, int Ns)
{
#pragma unroll
    for (int q = 0; q < d.Q; ++q) {
        const int j = threadIdx.x + q * d.T;
        const int k = j & (Ns - 1);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) s[pad(base + r * Ns)] = d.v[q][r];
    }
}

// One full pass through LDS (read, barrier, compute, write, barrier).
template<int R, int L>
__device__ inline void pass_lds(float2* s, int Ns, const float2* __restrict__ tw)
{
    Regs<R, L> d;
    pass_load_lds(d, s);
    __syncthreads();
    pass_compute(d, Ns, tw);
    pass_store_lds(d, s, Ns);
    __syncthreads();
}

template<int REM>
struct LastRadix {
    static constexpr int R = REM == 1 ? 2 : (REM == 2 ? 4 : (REM == 3 ? 8 : 16));
};

// Convolution of one pulse, 2^LOG2L points.  Forward transform: radix-16
// passes then the remainder radix Rl (16 when LOG2L is a multiple of 4); the
// inverse runs the radices in the opposite order, so the forward transform's
// last pass leaves in each thread exactly the spectrum bins the inverse's
// first pass needs: the spectral product and conjugation happen in
// registers.  The pulse is read from global memory by the first pass and the
// crop is written by the last, so LDS holds the data only between passes.
template<int LOG2L>
__global__ __launch_bounds__((1 << LOG2L) / 16) void conv_kernel(ConvArgs a)
{
    constexpr int L = 1 << LOG2L;
    constexpr int T = L / 16;
    constexpr int Rl = LastRadix<LOG2L % 4>::R;
    constexpr int N16 = (LOG2L - (Rl == 16 ? 4 : 0)) / 4;  // radix-16 passes besides the last
    __shared__ float2 s[L + L / 16];
    const int64_t b = blockIdx.x / a.nblk;
    const int64_t blk = a.blk0 + blockIdx.x % a.nblk;
    const int64_t o0 = blk * a.valid - a.shift;   // full-convolution index of transform sample 0
    const float2* x = a.in + b * a.n_in;

    // forward: first radix-16 pass from global memory
    {
        Regs<16, L> d;
        pass_load_global(d, x, o0, a.n_in);
        pass_compute(d, 1, a.tw);
        pass_store_lds(d, s, 1);
        __syncthreads();
    }
    int Ns = 16;
#pragma unroll
    for (int p = 1; p < N16; ++p) {
        pass_lds<16, L>(s, Ns, a.tw);
        Ns *= 16;
    }
    // forward last pass (radix Rl, Ns = L / Rl) -> product -> inverse first pass (radix Rl, Ns = 1)
    {
        Regs<Rl, L> d;
        pass_load_lds(d, s);
        __syncthreads();
        pass_compute(d, Ns, a.tw);
#pragma unroll
        for (int q = 0; q < d.Q; ++q) {
            const int j = threadIdx.x + q * T;   // j < L / Rl = Ns: bins j + r Ns
#pragma unroll
            for (int r = 0; r < Rl; ++r) d.v[q][r] = conjf2(cmul(d.v[q][r], a.H[j + r * Ns]));
        }
        pass_compute(d, 1, a.tw);
        pass_store_lds(d, s, 1);
        __syncthreads();
    }
    Ns = Rl;
#pragma unroll
    for (int p = 0; p + 1 < N16; ++p) {
        pass_lds<16, L>(s, Ns, a.tw);
        Ns *= 16;
    }
    // inverse last pass (radix 16, Ns = L / 16): conjugated crop straight to global memory
    {
        Regs<16, L> d;
        pass_load_lds(d, s);
        pass_compute(d, Ns, a.tw);
        float2* y = a.out + b * a.n_out;
        const int j = threadIdx.x;              // outputs j + r L / 16
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int t = j + r * Ns;           // transform sample: valid from `shift` on
            const int64_t i = o0 + t - a.offset;
            if (t >= a.shift && i >= 0 && i < a.n_out) {
                const float2 c = conjf2(d.v[0][r]);
                __builtin_nontemporal_store(v2f{c.x, c.y}, reinterpret_cast<v2f*>(y + i));
            }
        }
        // one block per pulse: outputs past the transform are zero
        if (a.nblk == 1 && a.shift == 0)
            for (int i = max(0, L - a.offset) + (int)threadIdx.x; i < a.n_out; i += T)
                y[i] = make_float2(0.f, 0.f);
    }
}

// ------------------------------------------------------------------------
// Lengths L = 32 L2 above 2^14 (four steps; L2 = 2^10 .. 2^14).  Sample n of a
// pulse is element (n1, n2) = (n / L2, n % L2) of a 32 x L2 matrix and
// spectrum bin k = k1 + 32 k2 is element (k1, k2):
//   col_fwd   per column n2: 32-point DFT over n1, times W_L^(n2 k1)   -> Y[k1][n2]
//   row       per row k1: L2-point DFT over n2 -> bins (k1, k2), times the
//             permuted filter spectrum, conjugate, L2-point DFT        -> Z[k1][n2]
//   col_inv   per column n2: Z times W_L^(n2 k1), 32-point DFT over k1,
//             conjugate: sample L2 n1 + n2 of the convolution, cropped
// (the inverse transform is conj(DFT(conj(.))) with 1/L folded into the
// spectrum; the conjugations of the two inverse steps cancel except the last).

template<>
__device__ inline void dft<32>(float2 (&v)[32])
{
    float2 e[16], o[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft<16>(e);
    dft<16>(o);
    // W32^k = exp(-2 pi i k / 32) = cos(pi k / 16) - i sin(pi k / 16)
    constexpr float kc[16] = {1.0f, 0.98078528040323043f, 0.92387953251128674f,
                              0.83146961230254524f, 0.70710678118654752f, 0.55557023301960218f,
                              0.38268343236508977f, 0.19509032201612827f, 0.0f,
                              -0.19509032201612827f, -0.38268343236508977f, -0.55557023301960218f,
                              -0.70710678118654752f, -0.83146961230254524f, -0.92387953251128674f,
                              -0.98078528040323043f};
    constexpr float ks[16] = {0.0f, 0.19509032201612827f, 0.38268343236508977f,
                              0.55557023301960218f, 0.70710678118654752f, 0.83146961230254524f,
                              0.92387953251128674f, 0.98078528040323043f, 1.0f,
                              0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
                              0.70710678118654752f, 0.55557023301960218f, 0.38268343236508977f,
                              0.19509032201612827f};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float cs = kc[k], sn = ks[k];
        const float2 t = make_float2(o[k].x * cs + o[k].y * sn, o[k].y * cs - o[k].x * sn);
        v[k] = cadd(e[k], t);
        v[k + 16] = csub(e[k], t);
    }
}

// v[k] *= W_L^(n k) for k < 32 from the table tw[j] = W_L^j, j < L: three
// reads (W^n, W^4n, W^16n) and at most four products per factor.
__device__ inline void twiddle32(float2 (&v)[32], const float2* __restrict__ tw, int n)
{
    const float2 w1 = tw[n], w4 = tw[4 * n], w16 = tw[16 * n];
    const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    const float2 w8 = cmul(w4, w4), w12 = cmul(w8, w4);
    const float2 lo[4] = {make_float2(1.f, 0.f), w1, w2, w3};
    const float2 mid[4] = {make_float2(1.f, 0.f), w4, w8, w12};
#pragma unroll
    for (int k = 1; k < 32; ++k) {
        float2 w = mid[(k >> 2) & 3];
        if (k & 3) w = (k >> 2) & 3 ? cmul(w, lo[k & 3]) : lo[k & 3];
        if (k & 16) w = cmul(w, w16);
        v[k] = cmul(v[k], w);
    }
}

template<int LOG2L2>
__global__ __launch_bounds__(256) void col_fwd_kernel(ConvArgs a)
{
    constexpr int L2 = 1 << LOG2L2;
    const int n2 = blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y;
    const float2* x = a.in + b * a.n_in;
    float2 v[32];
#pragma unroll
    for (int n1 = 0; n1 < 32; ++n1) {
        const int idx = n1 * L2 + n2;
        v[n1] = idx < a.n_in ? x[idx] : make_float2(0.f, 0.f);
    }
    dft<32>(v);
    twiddle32(v, a.tw, n2);
    float2* y = a.work + b * (int64_t)(32 * L2);
#pragma unroll
    for (int k1 = 0; k1 < 32; ++k1) y[(int64_t)k1 * L2 + n2] = v[k1];
}

template<int LOG2L2>
__global__ __launch_bounds__((1 << LOG2L2) / 16) void row_kernel(ConvArgs a)
{
    constexpr int L2 = 1 << LOG2L2;
    constexpr int T = L2 / 16;
    __shared__ float2 s[L2 + L2 / 16];
    const int k1 = blockIdx.x;
    float2* row = a.work + (int64_t)blockIdx.y * (32 * L2) + (int64_t)k1 * L2;
    for (int i = threadIdx.x; i < L2; i += T) s[pad(i)] = row[i];
    __syncthreads();
    fft<LOG2L2>(s, a.tw2);
    const float2* h = a.H + (int64_t)k1 * L2;   // permuted spectrum: bin k1 + 32 k2 at [k1][k2]
    for (int i = threadIdx.x; i < L2; i += T) s[pad(i)] = conjf2(cmul(s[pad(i)], h[i]));
    __syncthreads();
    fft<LOG2L2>(s, a.tw2);
    for (int i = threadIdx.x; i < L2; i += T) row[i] = s[pad(i)];
}

template<int LOG2L2>
__global__ __launch_bounds__(256) void col_inv_kernel(ConvArgs a)
{
    constexpr int L2 = 1 << LOG2L2;
    const int n2 = blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y;
    const float2* z = a.work + b * (int64_t)(32 * L2);
    float2 v[32];
#pragma unroll
    for (int k1 = 0; k1 < 32; ++k1) v[k1] = z[(int64_t)k1 * L2 + n2];
    twiddle32(v, a.tw, n2);
    dft<32>(v);
    float2* y = a.out + b * a.n_out;
#pragma unroll
    for (int n1 = 0; n1 < 32; ++n1) {
        const int i = n1 * L2 + n2 - a.offset;
        if (i >= 0 && i < a.n_out) y[i] = conjf2(v[n1]);
    }
    // outputs past the transform (valid mode with a chirp over twice the
    // pulse length) are zero, as in the fused kernel
    for (int i = max(0, 32 * L2 - a.offset) + n2; i < a.n_out; i += L2)
        y[i] = make_float2(0.f, 0.f);
}

}  // namespace

hipError_t launch_four_step(const ConvArgs& a, int log2l, int batch, hipStream_t stream)
{
    if (batch <= 0) return hipSuccess;
    const int log2l2 = log2l - 5;
    const dim3 cols((unsigned)((1 << log2l2) / 256), (unsigned)batch);
    const dim3 rows(32u, (unsigned)batch);
    switch (log2l2) {
#define RC_FOUR_STEP(P)                                                                      \
    case P:                                                                                  \
        hipLaunchKernelGGL(col_fwd_kernel<P>, cols, dim3(256), 0, stream, a);                \
        hipLaunchKernelGGL(row_kernel<P>, rows, dim3((1 << P) / 16), 0, stream, a);          \
        hipLaunchKernelGGL(col_inv_kernel<P>, cols, dim3(256), 0, stream, a);                \
        break;
        RC_FOUR_STEP(10)
        RC_FOUR_STEP(11)
        RC_FOUR_STEP(12)
        RC_FOUR_STEP(13)
        RC_FOUR_STEP(14)
#undef RC_FOUR_STEP
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_conv(const ConvArgs& a, int log2l, int batch, hipStream_t stream)
{
    if (batch <= 0) return hipSuccess;
    const dim3 grid((unsigned)((int64_t)batch * a.nblk));
    switch (log2l) {
    case 10: hipLaunchKernelGGL(conv_kernel<10>, grid, dim3(1 << 6), 0, stream, a); break;
    case 11: hipLaunchKernelGGL(conv_kernel<11>, grid, dim3(1 << 7), 0, stream, a); break;
    case 12: hipLaunchKernelGGL(conv_kernel<12>, grid, dim3(1 << 8), 0, stream, a); break;
    case 13: hipLaunchKernelGGL(conv_kernel<13>, grid, dim3(1 << 9), 0, stream, a); break;
    case 14: hipLaunchKernelGGL(conv_kernel<14>, grid, dim3(1 << 10), 0, stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rc
