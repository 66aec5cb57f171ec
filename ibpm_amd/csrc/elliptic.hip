// elliptic.hip -- batched multi-domain DST-I Poisson/Helmholtz solves (FP64).
//
// Replaces EllipticSolver2d::solve (reference EllipticSolver2d.cc:57-185) and
// the multi-domain driver EllipticSolver::solve (EllipticSolver.cc:47-73).
//
// One level solve u = S diag(e) S f * norm (S = DST-I matrix) is three passes,
// each reading and writing the level once (6 words per grid point):
//   pass 1: rows of the right-hand side (coarsify select, sign, boundary-value
//           correction of getRHS fused in) -> DST along j -> transposed T1
//   pass 2: rows of T1 -> DST along i -> eigenvalue scaling -> DST along k
//           -> transposed T2
//   pass 3: rows of T2 -> DST along l -> * 1/(4 nx ny) -> u (or u = W - .)
// Each DST-I of length N is the imaginary part of a complex FFT of length
// M = 2(N+1) of the odd extension; two rows share one FFT (real part row a,
// imaginary part row b).  M = 1024 and 2048 use a radix 16 x 16 x M/256
// transform held in registers (exchanges through LDS); the other lengths a
// mixed-radix (2,3,4,5) Stockham transform held in LDS, with each stage
// staged through registers so one LDS buffer suffices.  Passes may run over
// part of a level when only part of the result is read (EllJob's partial-pass
// fields).  The file also holds the coarsify chain and the local Jacobi solves
// of the force solve's A^-1 (k_fs_helm).
#include "ops.h"

#include <algorithm>

namespace ibpmg {

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// -i * a
__device__ inline double2 cmi(double2 a) { return make_double2(a.y, -a.x); }
// The same for the FP32 transforms (float2).
__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmi(float2 a) { return make_float2(a.y, -a.x); }

// Real type of a complex type, and a complex value from two reals.
template <class C> struct RealOf;
template <> struct RealOf<double2> { typedef double T; };
template <> struct RealOf<float2> { typedef float T; };
template <class C>
__device__ inline C cpx(typename RealOf<C>::T x, typename RealOf<C>::T y) {
    C r;
    r.x = x;
    r.y = y;
    return r;
}

typedef double DoublePair __attribute__((ext_vector_type(2), aligned(8)));

// Radix-R DFT in place (R = 2, 3, 4, 5, 8), FP64 (double2) or FP32 (float2):
// the same operations in the same order for both, constants rounded to the
// type.
template <int R, class C>
__device__ inline void dft(C* v) {
    typedef typename RealOf<C>::T T;
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T h = T(0.86602540378443864676);   // sin(2 pi / 3)
        C t1 = cadd(v[1], v[2]);
        C t2 = csub(v[1], v[2]);
        C m = cpx<C>(v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y);
        C r = cmi(t2);
        v[0] = cadd(v[0], t1);
        v[1] = cpx<C>(m.x + h * r.x, m.y + h * r.y);
        v[2] = cpx<C>(m.x - h * r.x, m.y - h * r.y);
    } else if constexpr (R == 4) {
        C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        C s13 = cadd(v[1], v[3]), d13 = cmi(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, d13);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410);    // cos(2 pi / 5)
        const T c2 = T(-0.80901699437494742410);   // cos(4 pi / 5)
        const T s1 = T(0.95105651629515357212);    // sin(2 pi / 5)
        const T s2 = T(0.58778525229247312917);    // sin(4 pi / 5)
        C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        C a1 = cpx<C>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        C a2 = cpx<C>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        C b1 = cpx<C>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        C b2 = cpx<C>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        b1 = cmi(b1);   // -i b1
        b2 = cmi(b2);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    } else {
        static_assert(R == 8, "radix 2, 3, 4, 5 or 8");
        // radix 8 = two radix-4 DFTs of the even and odd samples + radix-2 merge
        const T c = T(0.70710678118654752440);
        C a[4] = {v[0], v[2], v[4], v[6]};
        C b[4] = {v[1], v[3], v[5], v[7]};
        dft<4>(a);
        dft<4>(b);
        const C t0 = b[0];
        const C t1 = cpx<C>(c * (b[1].x + b[1].y), c * (b[1].y - b[1].x));    // W8^1 b1
        const C t2 = cpx<C>(b[2].y, -b[2].x);                                // W8^2 b2 = -i b2
        const C t3 = cpx<C>(c * (b[3].y - b[3].x), -c * (b[3].x + b[3].y));   // W8^3 b3
        v[0] = cadd(a[0], t0); v[4] = csub(a[0], t0);
        v[1] = cadd(a[1], t1); v[5] = csub(a[1], t1);
        v[2] = cadd(a[2], t2); v[6] = csub(a[2], t2);
        v[3] = cadd(a[3], t3); v[7] = csub(a[3], t3);
    }
}

// 16-point DFT as 4 x 4: X[k1 + 4 k2] = sum_n2 W16^(n2 k1) W4^(n2 k2)
// sum_n1 x[4 n1 + n2] W4^(n1 k1).
__device__ inline double2 mul_w16(double2 a, int k) {   // a * exp(-2 pi i k / 16), k in {1,2,3,6,9}
    const double c1 = 0.92387953251128675613;   // cos(pi / 8)
    const double s1 = 0.38268343236508977173;   // sin(pi / 8)
    const double h = 0.70710678118654752440;    // cos(pi / 4)
    switch (k) {
    case 1: return make_double2(c1 * a.x + s1 * a.y, c1 * a.y - s1 * a.x);
    case 2: return make_double2(h * (a.x + a.y), h * (a.y - a.x));
    case 3: return make_double2(s1 * a.x + c1 * a.y, s1 * a.y - c1 * a.x);
    case 6: return make_double2(h * (a.y - a.x), -h * (a.x + a.y));
    default: return make_double2(-c1 * a.x - s1 * a.y, s1 * a.x - c1 * a.y);   // 9
    }
}

__device__ inline void dft16(double2* v) {
    double2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        double2 c[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
        dft<4>(c);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = c[k1];
    }
    a[1][1] = mul_w16(a[1][1], 1); a[1][2] = mul_w16(a[1][2], 2); a[1][3] = mul_w16(a[1][3], 3);
    a[2][1] = mul_w16(a[2][1], 2); a[2][2] = cmi(a[2][2]);        a[2][3] = mul_w16(a[2][3], 6);
    a[3][1] = mul_w16(a[3][1], 3); a[3][2] = mul_w16(a[3][2], 6); a[3][3] = mul_w16(a[3][3], 9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        double2 c[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft<4>(c);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = c[k2];
    }
}

// FFT lengths with a compiled plan (M = 2(N+1) for N = 199, 255, 383, 399,
// 511, 1023: the configured grids); any other length uses the runtime plan.
__host__ __device__ constexpr bool compiled_len(int M) {
    return M == 400 || M == 512 || M == 768 || M == 800 || M == 1024 || M == 2048;
}

// Twiddle w^k = exp(-2 pi i k / M), k < M.  The longest compiled length
// keeps only the first quarter in LDS, w^(q M/4 + r) = (-i)^q w^r, which
// halves its LDS footprint and doubles the workgroups per CU; shorter
// lengths are not LDS-limited and read the full table.
__host__ __device__ constexpr bool quarter_twiddles(int M) { return M >= 2048 && M % 4 == 0; }

template <int CM, class C>
__device__ inline C twiddle(const C* tw, int k) {
    if constexpr (CM != 0 && quarter_twiddles(CM)) {
        typedef typename RealOf<C>::T T;
        constexpr int Q = CM / 4;
        const int q = k / Q, r = k - q * Q;
        const C t = tw[r];
        T x = t.x, y = t.y;
        if (q & 1) {   // times -i
            const T u = x;
            x = y;
            y = -u;
        }
        if (q & 2) {   // times -1
            x = -x;
            y = -y;
        }
        return cpx<C>(x, y);
    } else {
        return tw[k];
    }
}

// LDS slot of element e when the buffer is held swizzled: the low 3 bits
// are XORed with bits 3..5.  This permutes each aligned group of 8 complex
// values (128 B, one sweep of the 32 banks; compiled lengths are multiples
// of 8), so 8 lanes whose elements share their low 3 bits -- the first
// stage's radix-R writes, R apart -- land on distinct banks, while contiguous
// runs stay conflict-free.  Only the layout between the first and second
// stage is swizzled (the first stage's writes were the conflicting ones).
template <bool SW>
__device__ inline int lds_slot(int e) {
    if constexpr (SW) return e ^ ((e >> 3) & 7);
    else return e;
}

// The 2048-point transforms (largest grids, bandwidth-bound batches) keep the
// whole buffer swizzled, which measured faster there; the shorter compiled
// lengths swizzle only between the first two stages.
template <int CM>
constexpr bool kSwizzleAll = CM >= 2048;

// One Stockham radix-R stage over np FFTs of length M stored back to back in
// LDS; inputs are staged through registers so the stage runs in place.
// Twiddles come from the LDS table (see twiddle()).  Template
// arguments that are 0 are taken from the runtime values.
template <int R, int CM, int CNP, int CNS, bool SWIN = false, bool SWOUT = false, class C = double2>
__device__ inline void fft_stage(C* buf, const C* tw, int Mr, int npr, int Nsr) {
    const int M = CM ? CM : Mr;
    const int np = CM ? CNP : npr;
    const int Ns = CM ? CNS : Nsr;
    constexpr int MAXB = CM ? (CNP * CM / R + kFftThreads - 1) / kFftThreads
                            : (kFftLds / R + kFftThreads - 1) / kFftThreads;
    const int MR = M / R;
    const int total = np * MR;
    const int tstep = M / (Ns * R);
    C v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int g = threadIdx.x + b * kFftThreads;
        if (g < total) {
            const int p = g / MR;
            const int j = g - p * MR;
            const int jm = j % Ns;
            const int x = p * M;
            v[b][0] = buf[lds_slot<SWIN>(x + j)];
            if constexpr (CM != 0 && CNS == 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = buf[lds_slot<SWIN>(x + j + r * MR)];
            } else {
                // w^r by binary splitting, w^r = w^(r/2) w^(r - r/2): at most
                // 3 dependent products for R <= 8
                C pw[R];
                pw[1] = twiddle<CM>(tw, jm * tstep);
#pragma unroll
                for (int r = 2; r < R; ++r) pw[r] = cmul(pw[r / 2], pw[r - r / 2]);
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(buf[lds_slot<SWIN>(x + j + r * MR)], pw[r]);
            }
            dft<R>(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int g = threadIdx.x + b * kFftThreads;
        if (g < total) {
            const int p = g / MR;
            const int j = g - p * MR;
            const int jm = j % Ns;
            const int d = p * M + (j / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lds_slot<SWOUT>(d + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}

__host__ __device__ constexpr int pick_radix(int rem) {
    return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 5 == 0 ? 5 : 3;
}

// Compile-time stage list: radix 8 first, then 4, 2, 5, 3.
template <int M, int NP, int Ns, class C>
__device__ inline void fft_ct(C* buf, const C* tw) {
    if constexpr (Ns < M) {
        constexpr int R = pick_radix(M / Ns);
        constexpr bool first = Ns == 1, second = Ns > 1 && Ns == pick_radix(M);
        constexpr bool all = kSwizzleAll<M>;
        fft_stage<R, M, NP, Ns, all || second, all || (first && R < M), C>(buf, tw, 0, 0, 0);
        fft_ct<M, NP, Ns * R, C>(buf, tw);
    }
}

// Runtime stage list of a host-built plan (any 2(N+1) with factors 2,3,5).
template <class C>
__device__ inline void fft_rt(C* buf, const C* tw, const FftPlan& P) {
    int Ns = 1;
    for (int s = 0; s < P.nst; ++s) {
        const int R = P.rad[s];
        switch (R) {
        case 8: fft_stage<8, 0, 0, 0, false, false, C>(buf, tw, P.M, P.np, Ns); break;
        case 4: fft_stage<4, 0, 0, 0, false, false, C>(buf, tw, P.M, P.np, Ns); break;
        case 2: fft_stage<2, 0, 0, 0, false, false, C>(buf, tw, P.M, P.np, Ns); break;
        case 3: fft_stage<3, 0, 0, 0, false, false, C>(buf, tw, P.M, P.np, Ns); break;
        default: fft_stage<5, 0, 0, 0, false, false, C>(buf, tw, P.M, P.np, Ns); break;
        }
        Ns *= R;
    }
}

__host__ __device__ constexpr int np_of(int M) {
    return kFftLds / M < 1 ? 1 : (kFftLds / M > kFftNpMax ? kFftNpMax : kFftLds / M);
}

template <int CM, class C>
__device__ inline void fft_any(C* buf, const C* tw, const FftPlan& P) {
    if constexpr (CM != 0) fft_ct<CM, np_of(CM), 1, C>(buf, tw);
    else fft_rt<C>(buf, tw, P);
}

// ------------------------------------- 1024- and 2048-point transforms in registers
// The 511- and 1023-point grids (M = 1024, 2048) run one row pair per
// workgroup of T = M/16 threads as three radix stages 16 x 16 x R3 (R3 =
// M/256) held in registers: the first stage's inputs z[t + T r] come straight
// from global memory, the two exchanges between stages go through LDS, and
// the last stage hands Z[t + T m], m < 8 -- the half of the spectrum a DST-I
// needs -- to the caller as it is ready.
// A launch at M = 1024 runs only a quarter wave per CU when it holds few rows
// (one 511-row level: 256 workgroups of one wave), where the LDS passes'
// larger workgroups hide latency better; it takes the register-held
// transform from kRegMinPairs row pairs on.  M = 2048 always takes it.
template <int M>
__host__ __device__ constexpr bool r16_len() { return M == 1024 || M == 2048; }
constexpr long kRegMinPairs = 1024;

// Row pairs per workgroup of the register-held passes 1 and 2 (see
// pass1_r16; powers of two, at most 8); pass 3 writes rows and takes one.
#ifndef IBPM_R16_GROUP1
#define IBPM_R16_GROUP1 4
#endif
#ifndef IBPM_R16_GROUP2
#define IBPM_R16_GROUP2 1
#endif
__host__ __device__ constexpr int r16_group(int pass) {
    return pass == 1 ? IBPM_R16_GROUP1 : pass == 2 ? IBPM_R16_GROUP2 : 1;
}
static_assert((IBPM_R16_GROUP1 & (IBPM_R16_GROUP1 - 1)) == 0 && IBPM_R16_GROUP1 <= 8 &&
                  (IBPM_R16_GROUP2 & (IBPM_R16_GROUP2 - 1)) == 0 && IBPM_R16_GROUP2 <= 8,
              "row pairs per workgroup: a power of two, at most 8");

template <int CM, bool REG, int PASS>
__host__ __device__ constexpr int pass_threads() { return REG ? r16_group(PASS) * (CM / 16) : kFftThreads; }

// LDS slot of element e: one pad slot after every 16, so the first stage's
// writes (16 consecutive elements per lane) spread over the banks, and every
// access of a stage is a per-lane base plus a compile-time offset.  The
// exchanges move real and imaginary parts in two rounds through a buffer of
// M + M/16 doubles, half the footprint of a buffer of complex values, so
// twice as many workgroups share a CU.  r16_buf counts it in complex slots.
__host__ __device__ constexpr int r16_buf(int M) { return (M + M / 16) / 2; }
__device__ inline int pad16(int e) { return e + (e >> 4); }

// One exchange: every lane writes its 16 values at wr[k * ws], reads 16 at
// rd[k * rs] (per-lane bases, compile-time strides), real parts then
// imaginary parts.
template <int WS, int RS>
__device__ __forceinline__ void exchange16(double2* v, double* wr, const double* rd) {
#pragma unroll
    for (int r = 0; r < 16; ++r) wr[WS * r] = v[r].x;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r].x = rd[RS * r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) wr[WS * r] = v[r].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r].y = rd[RS * r];
    __syncthreads();
}

// Second half of the odd extension from the first: v[r] = z[t + T r] is set
// for r < 8; z[e] = -z[M - e] (another lane's) for e > M/2 comes through LDS,
// z[M/2] = 0.  The buffer is free again on return.
template <int M>
__device__ __forceinline__ void mirror_second_half(double2* v, double2* buf) {
    constexpr int T = M / 16;
    const int t = threadIdx.x & (T - 1);
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[t + T * r] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 8; r < 16; ++r) {
        const int e = t + T * r;
        if (e == M / 2) {
            v[r] = make_double2(0.0, 0.0);
        } else {
            const double2 y = buf[M - e];
            v[r] = make_double2(-y.x, -y.y);
        }
    }
    __syncthreads();
}

// The whole odd extension from its first half z[e], e < M/2, left in buf by
// the caller: v[r] = z[t + T r], with z[e] = -z[M - e] for e > M/2 and
// z[M/2] = 0.  The buffer is free again on return.
template <int M>
__device__ __forceinline__ void extension_from_lds(double2* v, double2* buf) {
    constexpr int T = M / 16;
    const int t = threadIdx.x & (T - 1);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = buf[t + T * r];
#pragma unroll
    for (int r = 8; r < 16; ++r) {
        const int e = t + T * r;
        if (e == M / 2) {
            v[r] = make_double2(0.0, 0.0);
        } else {
            const double2 y = buf[M - e];
            v[r] = make_double2(-y.x, -y.y);
        }
    }
    __syncthreads();
}

// First half of the odd extension of two rows a, b (b absent: zeros):
// v[r] = (a, b)[e - 1] at e = t + T r, r < 8, and 0 at e = 0.  Row b is read
// from row a when absent and multiplied by 0.  A select of the loaded value
// would be compiled as a branch around the load, and a load under a branch
// gets a wait for every load before it: the row pair's sixteen loads would
// then complete one after another.
template <int M>
__device__ __forceinline__ void load_first_half(double2* v, const IBPM_GLOBAL double* a,
                                                const IBPM_GLOBAL double* b, bool hb) {
    constexpr unsigned T = M / 16;
    const unsigned t = threadIdx.x & (T - 1);
    const IBPM_GLOBAL double* bb = hb ? b : a;
    const double keep = hb ? 1.0 : 0.0;   // a product, not a select: see above
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const unsigned e = t + T * r, k = e == 0 ? 0u : e - 1;
        const double nz = e == 0 ? 0.0 : 1.0;   // e = 0 holds the extension's zero, also a product
        v[r] = make_double2(a[k] * nz, bb[k] * keep * nz);
    }
}

// In: v[r] = z[t + T r], r < 16.  Out: emit(m, Z[t + T m]) for m < 8, called
// as each output is ready (so callers write results without holding all
// eight).  Twiddles, w = exp(-2 pi i / M), come from global memory (cached):
// the second stage's w^(R3 jm r) from the plan's 256-entry table at 16 r + jm,
// so 16 lanes with consecutive jm read consecutive entries; the third
// stage's w^j (j < 256) from the full table.  The buffer is free again on
// return (the last LDS access precedes the last barrier).
template <int M, class Emit>
__device__ __forceinline__ void fft_regs(double2* v, double2* buf, const IBPM_GLOBAL double* tw,
                                         const IBPM_GLOBAL double* twg, const Emit& emit) {
    constexpr int T = M / 16, R3 = M / 256, H = 256 / T, RS = 17 * T / 16;
    const int t = threadIdx.x & (T - 1);
    double* bd = (double*)buf;
    dft16(v);   // stage 1 (Ns = 1): outputs 16 t + r to slots 17 t + r; reads
                // t + T r from slots pad16(t) + RS r
    exchange16<1, RS>(v, bd + 17 * t, bd + pad16(t));
    const int jm = t & 15;   // stage 2 (Ns = 16): twiddles w^(R3 jm r) from the table
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], make_double2(tw[2 * (16 * r + jm)], tw[2 * (16 * r + jm) + 1]));
    dft16(v);
    // outputs (t/16) 256 + jm + 16 r to slots (t/16) 272 + jm + 17 r; stage 3
    // reads j + 256 r (j = t + T h, h < H) = t + T (h + H r), the exchange16
    // read pattern: lane t reads pad16(t) + RS k for k = 0..15
    exchange16<17, RS>(v, bd + (t >> 4) * 272 + jm, bd + pad16(t));
#pragma unroll
    for (int h = 0; h < H; ++h) {   // stage 3 (Ns = 256): butterflies j = t + T h
        const int j = t + T * h;
        double2 u[R3];
#pragma unroll
        for (int r = 0; r < R3; ++r) u[r] = v[h + H * r];
        const double2 w1 = make_double2(twg[2 * j], twg[2 * j + 1]);
        const double2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        u[1] = cmul(u[1], w1);
        u[2] = cmul(u[2], w2);
        u[3] = cmul(u[3], w3);
        if constexpr (R3 == 8) {
            const double2 w4 = cmul(w2, w2);
            u[4] = cmul(u[4], w4);
            u[5] = cmul(u[5], cmul(w4, w1));
            u[6] = cmul(u[6], cmul(w4, w2));
            u[7] = cmul(u[7], cmul(w4, w3));
        }
        dft<R3>(u);
#pragma unroll
        for (int r = 0; r < R3 / 2; ++r) emit(h + H * r, u[r]);   // Z[j + 256 r] with j + 256 r < M/2
    }
}

// ---------------------------------------------------------------- pass 1
// getBC (reference Scalar.cc:91-127) evaluated at one boundary index.
// (Global-address-space reads: a flat load would also count against the LDS
// counter, and the barrier after the right-hand side would wait for it.)
__device__ inline double bc_bottom(const EllJob& J, int i) {
    if (J.bcring) return i == 0 ? gin(J.bcring)[0] : gin(J.bcring)[2 * (J.nx + J.ny) - i];
    const int ii = i / 2 + nx_ext(J.nx), jj = nx_ext(J.ny);
    const double a = gin(J.uc)[sidx(J.nx, J.ny, ii, jj)];
    return (i & 1) ? 0.5 * (a + gin(J.uc)[sidx(J.nx, J.ny, ii + 1, jj)]) : a;
}
__device__ inline double bc_top(const EllJob& J, int i) {
    if (J.bcring) return gin(J.bcring)[J.ny + i];
    const int ii = i / 2 + nx_ext(J.nx), jj = J.ny / 2 + nx_ext(J.ny);
    const double a = gin(J.uc)[sidx(J.nx, J.ny, ii, jj)];
    return (i & 1) ? 0.5 * (a + gin(J.uc)[sidx(J.nx, J.ny, ii + 1, jj)]) : a;
}
__device__ inline double bc_left(const EllJob& J, int j) {
    if (J.bcring) return gin(J.bcring)[j];
    const int ii = nx_ext(J.nx), jj = j / 2 + nx_ext(J.ny);
    const double a = gin(J.uc)[sidx(J.nx, J.ny, ii, jj)];
    return (j & 1) ? 0.5 * (a + gin(J.uc)[sidx(J.nx, J.ny, ii, jj + 1)]) : a;
}
__device__ inline double bc_right(const EllJob& J, int j) {
    if (J.bcring) return gin(J.bcring)[2 * J.ny + J.nx - j];
    const int ii = J.nx / 2 + nx_ext(J.nx), jj = j / 2 + nx_ext(J.ny);
    const double a = gin(J.uc)[sidx(J.nx, J.ny, ii, jj)];
    return (j & 1) ? 0.5 * (a + gin(J.uc)[sidx(J.nx, J.ny, ii, jj + 1)]) : a;
}

// bc_left (left) or bc_right at j with no load under a lane-dependent branch:
// for even j the second read repeats the first, and 0.5 (a + a) = a exactly.
__device__ inline double bc_side_loads(const EllJob& J, bool left, int j) {
    if (J.bcring) return left ? gin(J.bcring)[j] : gin(J.bcring)[2 * J.ny + J.nx - j];
    const int ii = left ? nx_ext(J.nx) : J.nx / 2 + nx_ext(J.nx), jj = j / 2 + nx_ext(J.ny);
    const double a = gin(J.uc)[sidx(J.nx, J.ny, ii, jj)];
    const double b = gin(J.uc)[sidx(J.nx, J.ny, ii, jj + (j & 1))];
    return 0.5 * (a + b);
}

// Right-hand side of one level at (i, j), 1-based: coarsified copy in the
// central region of levels >= 1 (EllipticSolver.cc:54-55), then getRHS
// (EllipticSolver2d.cc:125-144 / 166-185).
__device__ inline double rhs_value(const EllJob& J, int i, int j) {
    const int ex = nx_ext(J.nx), ey = nx_ext(J.ny);
    const bool central = J.lev >= 1 && i > ex && i < J.nx / 2 + ex && j > ey && j < J.ny / 2 + ey;
    const long k = sidx(J.nx, J.ny, i, j);
    double v = (central ? gin(J.R)[k] : gin(J.f)[k]) * J.in_scale;
    if (J.uc || J.bcring) {
        const double c = J.bc_coef;
        if (j == 1) v -= bc_bottom(J, i) * c;
        if (j == J.ny - 1) v -= bc_top(J, i) * c;
        if (i == 1) v -= bc_left(J, j) * c;
        if (i == J.nx - 1) v -= bc_right(J, j) * c;
    }
    return v;
}

// Points of a coarse level (lev >= 1) that getBC of the next finer level
// reads: the rectangle i = ex..ex+nx/2, j = ey..ey+ny/2 traced by the finer
// domain's boundary (Scalar.cc:91-127).
__device__ inline bool on_bc_ring(const EllJob& J, int i, int j) {
    const int ex = nx_ext(J.nx), ey = nx_ext(J.ny);
    const int i1 = ex + J.nx / 2, j1 = ey + J.ny / 2;
    return ((j == ey || j == j1) && i >= ex && i <= i1) || ((i == ex || i == i1) && j >= ey && j <= j1);
}

// The three passes are written once over a template length CM (0: runtime
// plan).  Row length N, FFT length M and rows per block follow from the plan,
// so with CM != 0 every index computation folds to constants.
template <int CM, class C>
__device__ inline void load_twiddles(C* tw, const FftPlan& P) {
    typedef typename RealOf<C>::T T;
    const int M = (CM && quarter_twiddles(CM)) ? CM / 4 : (CM ? CM : P.M);
    const IBPM_GLOBAL double* g = gin((const double*)P.tw);
    for (int k = threadIdx.x; k < M; k += kFftThreads) tw[k] = cpx<C>(T(g[2 * k]), T(g[2 * k + 1]));
}

// The transposed scratch rows T1, T2 as the transform's real type: FP64, or
// FP32 in the FP32 passes (the same buffers, half of each used).
template <class T>
__device__ inline const IBPM_GLOBAL T* tin(const double* p) { return (const IBPM_GLOBAL T*)p; }
template <class T>
__device__ inline IBPM_GLOBAL T* tout(double* p) { return (IBPM_GLOBAL T*)p; }


// The LDS passes are also instanced over C = float2: FP32 transforms and
// transposed scratch, with the right-hand side, boundary terms, eigenvalue
// scaling and output in FP64.
template <int CM, class C = double2>
__device__ __forceinline__ void pass1_lds(const EllJob& J, C* smem, int bx) {
    typedef typename RealOf<C>::T T;
    const FftPlan& P = J.pj;
    const int M = CM ? CM : P.M, np = CM ? np_of(CM) : P.np;
    const int N1 = J.nx - 1, N2 = M / 2 - 1;
    const int rows = 2 * np;
    const int r0 = bx * rows;
    if (r0 >= N1) return;
    C* buf = smem;
    C* tw = smem + np * M;
    load_twiddles<CM>(tw, P);
    const int n = N2 + 1;
    for (int idx = threadIdx.x; idx < np * N2; idx += kFftThreads) {
        const int p = idx / N2, k = idx - p * N2;
        const int ra = r0 + 2 * p, rb = ra + 1;
        const T va = (ra < N1) ? T(rhs_value(J, ra + 1, k + 1)) : T(0);
        const T vb = (rb < N1) ? T(rhs_value(J, rb + 1, k + 1)) : T(0);
        buf[lds_slot<kSwizzleAll<CM>>(p * M + k + 1)] = cpx<C>(va, vb);
        buf[lds_slot<kSwizzleAll<CM>>(p * M + M - 1 - k)] = cpx<C>(-va, -vb);
    }
    for (int p = threadIdx.x; p < np; p += kFftThreads) {
        buf[lds_slot<kSwizzleAll<CM>>(p * M)] = cpx<C>(T(0), T(0));
        buf[lds_slot<kSwizzleAll<CM>>(p * M + n)] = cpx<C>(T(0), T(0));
    }
    __syncthreads();
    fft_any<CM>(buf, tw, P);
    for (int idx = threadIdx.x; idx < rows * N2; idx += kFftThreads) {
        const int r = idx % rows, k = idx / rows;
        const int row = r0 + r;
        if (row >= N1) continue;
        const C z = buf[lds_slot<kSwizzleAll<CM>>((r >> 1) * M + k + 1)];
        tout<T>(J.T1)[(long)k * J.ld1 + row] = (r & 1) ? z.x : -z.y;
    }
}

template <int CM, class C = double2>
__device__ __forceinline__ void pass2_lds(const EllJob& J, C* smem, int bx) {
    typedef typename RealOf<C>::T T;
    const FftPlan& P = J.pi;
    const int M = CM ? CM : P.M, np = CM ? np_of(CM) : P.np;
    const int N1 = M / 2 - 1, N2 = J.ny - 1;
    const int rows = 2 * np;
    const int r0 = bx * rows;   // rows are l = 0..N2-1
    if (r0 >= N2) return;
    C* buf = smem;
    C* tw = smem + np * M;
    load_twiddles<CM>(tw, P);
    const int n = N1 + 1;
    for (int idx = threadIdx.x; idx < np * N1; idx += kFftThreads) {
        const int p = idx / N1, k = idx - p * N1;
        const int la = r0 + 2 * p, lb = la + 1;
        const T va = (la < N2) ? tin<T>(J.T1)[(long)la * J.ld1 + k] : T(0);
        const T vb = (lb < N2) ? tin<T>(J.T1)[(long)lb * J.ld1 + k] : T(0);
        buf[lds_slot<kSwizzleAll<CM>>(p * M + k + 1)] = cpx<C>(va, vb);
        buf[lds_slot<kSwizzleAll<CM>>(p * M + M - 1 - k)] = cpx<C>(-va, -vb);
    }
    for (int p = threadIdx.x; p < np; p += kFftThreads) {
        buf[lds_slot<kSwizzleAll<CM>>(p * M)] = cpx<C>(T(0), T(0));
        buf[lds_slot<kSwizzleAll<CM>>(p * M + n)] = cpx<C>(T(0), T(0));
    }
    __syncthreads();
    fft_any<CM>(buf, tw, P);
    // spectral scaling, EllipticSolver2d.cc:42-54 (eigenvalues), :89-93 (scale)
    for (int idx = threadIdx.x; idx < np * N1; idx += kFftThreads) {
        const int p = idx / N1, k = idx - p * N1;
        const int la = r0 + 2 * p, lb = la + 1;
        const C z = buf[lds_slot<kSwizzleAll<CM>>(p * M + k + 1)];
        double ya = -z.y, yb = z.x;
        const double cx = gin(J.cosx)[k];
        if (la < N2) {
            const double lam = 2. * (cx + gin(J.cosy)[la] - 2.) * J.bydx2;
            ya *= (J.kind == 0) ? 1. / lam : 1. / (1 + J.alpha * lam);
        }
        if (lb < N2) {
            const double lam = 2. * (cx + gin(J.cosy)[lb] - 2.) * J.bydx2;
            yb *= (J.kind == 0) ? 1. / lam : 1. / (1 + J.alpha * lam);
        }
        buf[lds_slot<kSwizzleAll<CM>>(p * M + k + 1)] = cpx<C>(T(ya), T(yb));
        buf[lds_slot<kSwizzleAll<CM>>(p * M + M - 1 - k)] = cpx<C>(T(-ya), T(-yb));
    }
    for (int p = threadIdx.x; p < np; p += kFftThreads) {
        buf[lds_slot<kSwizzleAll<CM>>(p * M)] = cpx<C>(T(0), T(0));
        buf[lds_slot<kSwizzleAll<CM>>(p * M + n)] = cpx<C>(T(0), T(0));
    }
    __syncthreads();
    fft_any<CM>(buf, tw, P);
    for (int idx = threadIdx.x; idx < rows * N1; idx += kFftThreads) {
        const int r = idx % rows, k = idx / rows;
        const int l = r0 + r;
        if (l >= N2) continue;
        const C z = buf[lds_slot<kSwizzleAll<CM>>((r >> 1) * M + k + 1)];
        tout<T>(J.T2)[(long)k * J.ld2 + l] = (r & 1) ? z.x : -z.y;
    }
}

template <int CM, class C = double2>
__device__ __forceinline__ void pass3_lds(const EllJob& J, C* smem, int bx) {
    typedef typename RealOf<C>::T T;
    const FftPlan& P = J.pj;
    const int M = CM ? CM : P.M, np = CM ? np_of(CM) : P.np;
    const int N1 = J.nx - 1, N2 = M / 2 - 1;
    const int rows = 2 * np;
    const int r0 = bx * rows;
    if (r0 >= N1) return;
    C* buf = smem;
    C* tw = smem + np * M;
    load_twiddles<CM>(tw, P);
    const int n = N2 + 1;
    for (int idx = threadIdx.x; idx < np * N2; idx += kFftThreads) {
        const int p = idx / N2, k = idx - p * N2;
        const int ra = r0 + 2 * p, rb = ra + 1;
        const T va = (ra < N1) ? tin<T>(J.T2)[(long)ra * J.ld2 + k] : T(0);
        const T vb = (rb < N1) ? tin<T>(J.T2)[(long)rb * J.ld2 + k] : T(0);
        buf[lds_slot<kSwizzleAll<CM>>(p * M + k + 1)] = cpx<C>(va, vb);
        buf[lds_slot<kSwizzleAll<CM>>(p * M + M - 1 - k)] = cpx<C>(-va, -vb);
    }
    for (int p = threadIdx.x; p < np; p += kFftThreads) {
        buf[lds_slot<kSwizzleAll<CM>>(p * M)] = cpx<C>(T(0), T(0));
        buf[lds_slot<kSwizzleAll<CM>>(p * M + n)] = cpx<C>(T(0), T(0));
    }
    __syncthreads();
    fft_any<CM>(buf, tw, P);
    for (int idx = threadIdx.x; idx < np * N2; idx += kFftThreads) {
        const int p = idx / N2, k = idx - p * N2;
        const C z = buf[lds_slot<kSwizzleAll<CM>>(p * M + k + 1)];
        const int ra = r0 + 2 * p, rb = ra + 1;
        // sinTransformInv normalisation, EllipticSolver2d.cc:80-83
        // with the epilogue, the finer level's BCs must come from v itself
        // (EllipticSolver.cc:62-66 solves c = A^-1 c before omega = omega* - c)
        if (ra < N1) {
            const long o = (long)ra * N2 + k;
            const double v = -(double)z.y * J.norm;
            const double uv = J.epilogue ? gin(J.W)[o] - v : v;
            gout(J.u)[o] = uv;
            if (J.u2) gout(J.u2)[o] = uv;
            if (J.vside && on_bc_ring(J, ra + 1, k + 1)) gout(J.vside)[o] = v;
        }
        if (rb < N1) {
            const long o = (long)rb * N2 + k;
            const double v = (double)z.x * J.norm;
            const double uv = J.epilogue ? gin(J.W)[o] - v : v;
            gout(J.u)[o] = uv;
            if (J.u2) gout(J.u2)[o] = uv;
            if (J.vside && on_bc_ring(J, rb + 1, k + 1)) gout(J.vside)[o] = v;
        }
    }
}

// Two adjacent doubles of a transposed row pair written as one 16-byte
// store (8-byte aligned), or the first alone for the last, odd row.
__device__ __forceinline__ void store_pair(IBPM_GLOBAL double* p, double a, double b, bool both) {
    if (both) *(IBPM_GLOBAL DoublePair*)p = DoublePair{a, b};   // one global_store_dwordx4
    else p[0] = a;
}

// The three passes for M = 1024, 2048 (N = M/2 - 1), the same arithmetic as the
// LDS passes above around a register-held transform.  A workgroup holds G =
// r16_group(pass) row pairs, one per sub-group of T = M/16 threads; the
// sub-groups share the transform's barriers and each has its own LDS buffer.
// Passes 1 and 2 write their outputs transposed, 2 G neighbouring values per
// output column, so the outputs are staged in LDS (all buffers are free once
// the last exchange is done) and written in pieces of 2 G values: with one
// pair per workgroup each 16-byte piece was a write request of its own, and
// the request rate, not the bytes, bounded those passes.  A sub-group past
// the last pair transforms the last pair again and writes nothing, so every
// thread reaches every barrier.  (Two pairs per workgroup processed one after
// the other, the second's loads issued before the first's transform,
// measured slower: fewer, longer workgroups.)

// Staged output (a, b) of sub-group g for column k: 16-byte slot k G + (g ^
// sw(k)), the swizzle sw spreading one wave's writes (fixed g, consecutive k)
// over the banks; write_staged reads the slots in order.
template <int G>
__device__ __forceinline__ int stage_slot(int k, int g) {
    return k * G + (g ^ ((k / (16 / G)) & (G - 1)));
}

template <int G>
__device__ __forceinline__ void stage_pair(double2* st, int k, int g, double a, double b) {
    st[stage_slot<G>(k, g)] = make_double2(a, b);
}

// Rows row0 + 2 g + s (those < nrows) of the staged columns k < N to
// out[k * ld + row], one 16-byte piece per thread, the G pieces of a column
// from neighbouring threads (with ld a multiple of 2 G, one aligned piece of
// 16 G bytes).
template <int M, int G>
__device__ __forceinline__ void write_staged(IBPM_GLOBAL double* out, const double2* st, int nrows, int ld,
                                             int row0, int k_lo = 0, int k_hi = M) {
    constexpr int N = M / 2 - 1, NT = G * (M / 16);
    for (int idx = threadIdx.x; idx < N * G; idx += NT) {
        const int k = idx / G, slot = idx % G;
        if (k < k_lo || k > k_hi) continue;   // columns k outside [k_lo, k_hi] are not read
        const int g = slot ^ ((k / (16 / G)) & (G - 1));
        const int row = row0 + 2 * g;
        if (row < nrows) {
            const double2 z = st[idx];
            store_pair(out + (long)k * ld + row, z.x, z.y, row + 1 < nrows);
        }
    }
}

// Pass 1 of a workgroup whose rows have no interior right-hand side: the
// rows' only nonzeros are getRHS's terms x_1 = -bottom c and x_N = -top c
// (rhs_value's, in its order), so the DST-I of a row is
// 2 sin(pi k / ny) (x_1 + (-1)^(k+1) x_N), written like write_staged.
template <int M, int G>
__device__ __forceinline__ void pass1_bc_rows(const EllJob& J, double2* smem, int row0) {
    constexpr int N = M / 2 - 1, NT = G * (M / 16);
    const int N1 = J.nx - 1;
    const int t = threadIdx.x;
    double* bv = (double*)smem;   // x_1 of the 2G rows, then x_N
    if (t < 2 * G) {
        const int i = row0 + t + 1;
        double x1 = 0.0, xn = 0.0;
        if (i <= N1 && (J.uc || J.bcring)) {
            x1 = 0.0 - bc_bottom(J, i) * J.bc_coef;
            xn = 0.0 - bc_top(J, i) * J.bc_coef;
        }
        bv[t] = x1;
        bv[2 * G + t] = xn;
    }
    __syncthreads();
    const IBPM_GLOBAL double* s1 = gin(J.sin1);
    IBPM_GLOBAL double* out = gout(J.T1);
    for (int idx = t; idx < N * G; idx += NT) {
        const int k = idx / G, g = idx % G;   // mode k + 1
        const int row = row0 + 2 * g;
        if (row >= N1) continue;
        const double s = s1[k];
        const double a = s * ((k & 1) ? bv[2 * g] - bv[2 * G + 2 * g] : bv[2 * g] + bv[2 * G + 2 * g]);
        const double b = s * ((k & 1) ? bv[2 * g + 1] - bv[2 * G + 2 * g + 1] : bv[2 * g + 1] + bv[2 * G + 2 * g + 1]);
        store_pair(out + (long)k * J.ld1 + row, a, b, row + 1 < N1);
    }
}

template <int M, bool FS>
__device__ __forceinline__ void pass1_r16(const EllJob& J, double2* smem, int bx) {
    constexpr int T = M / 16, G = r16_group(1);
    const int N1 = J.nx - 1, npairs = (N1 + 1) / 2;
    if (G * bx >= npairs) return;
    if (J.sin1) {   // a workgroup of rows off the right-hand side's support, none on the edge
        const int i_first = 2 * G * bx + 1, i_last = min(2 * G * bx + 2 * G, N1);
        if ((i_last < J.sup_lo || i_first > J.sup_hi) && i_first > 1 && i_last < J.nx - 1) {
            pass1_bc_rows<M, G>(J, smem, 2 * G * bx);
            return;
        }
    }
    const int g = threadIdx.x / T, t = threadIdx.x & (T - 1);
    const int q = min(G * bx + g, npairs - 1);
    double2* buf = smem + g * r16_buf(M);
    const IBPM_GLOBAL double* tw = gin((const double*)J.pj.tw16);
    const IBPM_GLOBAL double* twg = gin((const double*)J.pj.tw);
    const int ey = nx_ext(J.ny), ex = nx_ext(J.nx);
    const int jc0 = ey, jc1 = J.ny / 2 + ey;   // central columns: jc0 < j < jc1
    const double sc = J.in_scale;
    const bool bcs = J.uc || J.bcring;
    // rhs_value along rows ra, rb = ra + 1 with the row-invariant parts
    // hoisted: the first half of the odd extension, e = t + T r (r < 8) at
    // column j = e.  The boundary terms are subtracted in rhs_value's order
    // (bottom, top, left, right), the left and right ones on the level's edge
    // rows with all their loads issued before any is used (a row pair whose
    // loads wait one by one makes the slowest workgroup of the launch).
    auto load = [&](int q, double2* v) {
        const int ra = 2 * q, rb = ra + 1;
        const bool hb = rb < N1;
        const bool cra = J.lev >= 1 && ra + 1 > ex && ra + 1 < J.nx / 2 + ex;
        const bool crb = hb && J.lev >= 1 && rb + 1 > ex && rb + 1 < J.nx / 2 + ex;
        const IBPM_GLOBAL double* fa = gin(J.f) + sidx(J.nx, J.ny, ra + 1, 1);
        const IBPM_GLOBAL double* fb = gin(J.f) + sidx(J.nx, J.ny, hb ? rb + 1 : ra + 1, 1);
        const IBPM_GLOBAL double* Ra = gin(J.R) + sidx(J.nx, J.ny, ra + 1, 1);
        const IBPM_GLOBAL double* Rb = gin(J.R) + sidx(J.nx, J.ny, hb ? rb + 1 : ra + 1, 1);
        const double keep = hb ? 1.0 : 0.0;
        const bool ina = ra + 1 >= J.rb_i0 && ra + 1 <= J.rb_i1, inb = rb + 1 >= J.rb_i0 && rb + 1 <= J.rb_i1;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = t + T * r, k = e - 1;
            const bool colc = e > jc0 && e < jc1;
            const bool colr = e >= J.rb_j0 && e <= J.rb_j1;
            const int kk = e == 0 ? 0 : k;
            // row b absent: read row a, times 0; e = 0, the extension's zero:
            // times 0; R off its box: times 0 (products, as in
            // load_first_half, so that the sixteen loads are in flight together)
            const double nz = e == 0 ? 0.0 : 1.0;
            const double la = ((cra && colc) ? Ra : fa)[kk], lb = ((crb && colc) ? Rb : fb)[kk];
            if constexpr (FS) {
                const double ma = (cra && colc && !(ina && colr)) ? 0.0 : 1.0;
                const double mb = (crb && colc && !(inb && colr)) ? 0.0 : 1.0;
                v[r] = make_double2(la * sc * nz * ma, lb * sc * keep * nz * mb);
            } else {
                v[r] = make_double2(la * sc * nz, lb * sc * keep * nz);
            }
        }
        if (!bcs) return;
        const double c = J.bc_coef;
        // getRHS at j = 1 (e = 1) and j = ny - 1 (e = M/2 - 1 = T - 1 + 7 T)
        if (t == 1) {
            v[0].x -= bc_bottom(J, ra + 1) * c;
            if (hb) v[0].y -= bc_bottom(J, rb + 1) * c;
        }
        if (t == T - 1) {
            v[7].x -= bc_top(J, ra + 1) * c;
            if (hb) v[7].y -= bc_top(J, rb + 1) * c;
        }
        // i = 1 (row a of the first pair only) and i = nx - 1 (row a or b)
        const bool left_a = ra + 1 == 1, right_a = ra + 1 == J.nx - 1, right_b = hb && rb + 1 == J.nx - 1;
        if (left_a || right_a || right_b) {
            double gl[8], gr[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int j = max(t + T * r, 1);   // e = 0 holds the extension's zero and takes no term
                gl[r] = left_a ? bc_side_loads(J, true, j) : 0.0;
                gr[r] = (right_a || right_b) ? bc_side_loads(J, false, j) : 0.0;
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                if (t + T * r == 0) continue;
                if (left_a) v[r].x -= gl[r] * c;
                if (right_a) v[r].x -= gr[r] * c;
                if (right_b) v[r].y -= gr[r] * c;
            }
        }
    };
    double2 v[16];
    load(q, v);
    mirror_second_half<M>(v, buf);
    fft_regs<M>(v, buf, tw, twg, [&](int m, double2 z) {
        const int kk = t + T * m;
        if (kk > 0) stage_pair<G>(smem, kk - 1, g, -z.y, z.x);
    });
    __syncthreads();
    write_staged<M, G>(gout(J.T1), smem, N1, J.ld1, 2 * G * bx);
}

template <int M>
__device__ __forceinline__ void pass2_r16(const EllJob& J, double2* smem, int bx) {
    constexpr int T = M / 16, G = r16_group(2);
    const int N2 = J.ny - 1, npairs = (N2 + 1) / 2;
    if (G * bx >= npairs) return;
    const int g = threadIdx.x / T, t = threadIdx.x & (T - 1);
    const int q = min(G * bx + g, npairs - 1);
    double2* buf = smem + g * r16_buf(M);
    const IBPM_GLOBAL double* tw = gin((const double*)J.pi.tw16);
    const IBPM_GLOBAL double* twg = gin((const double*)J.pi.tw);
    const IBPM_GLOBAL double* gcx = gin(J.cosx);
    const int la = 2 * q, lb = la + 1;
    const bool hb = lb < N2;
    double2 v[16];
    load_first_half<M>(v, gin(J.T1) + (long)la * J.ld1, gin(J.T1) + (long)lb * J.ld1, hb);
    mirror_second_half<M>(v, buf);
    // spectral scaling, EllipticSolver2d.cc:42-54 (eigenvalues), :89-93 (scale),
    // into the first half of the odd extension for the second transform,
    // kept in the (free) LDS buffer rather than in registers beside the
    // transform's own
    const double cya = gin(J.cosy)[la], cyb = hb ? gin(J.cosy)[lb] : 0.0;
    fft_regs<M>(v, buf, tw, twg, [&](int m, double2 z) {
        const int kk = t + T * m;
        double ya = 0.0, yb = 0.0;
        if (kk > 0) {
            const double cx = gcx[kk - 1];
            ya = -z.y;
            yb = z.x;
            const double lama = 2. * (cx + cya - 2.) * J.bydx2;
            ya *= (J.kind == 0) ? 1. / lama : 1. / (1 + J.alpha * lama);
            if (hb) {
                const double lamb = 2. * (cx + cyb - 2.) * J.bydx2;
                yb *= (J.kind == 0) ? 1. / lamb : 1. / (1 + J.alpha * lamb);
            }
        }
        buf[kk] = make_double2(ya, yb);
    });
    extension_from_lds<M>(v, buf);
    fft_regs<M>(v, buf, tw, twg, [&](int m, double2 z) {
        const int kk = t + T * m;
        if (kk > 0) stage_pair<G>(smem, kk - 1, g, -z.y, z.x);
    });
    __syncthreads();
    write_staged<M, G>(gout(J.T2), smem, N2, J.ld2, 2 * G * bx, J.t2_lo, J.t2_hi);
}

// Pass 3 of a row pair that is not transformed: u at columns rj0, rj1 of its
// rows in [ring_lo, ring_hi], each the sum over the row's modes l of
// T2[l] 2 sin(pi l j / ny) (the inverse DST-I at one point), times norm.
template <int M>
__device__ __forceinline__ void pass3_ring(const EllJob& J, double2* smem, int bx) {
    constexpr int T = M / 16;
    const int N1 = J.nx - 1, N2 = M / 2 - 1;
    const int t = threadIdx.x;
    const IBPM_GLOBAL double* sn = gin(J.ringsin);
    double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int r = 2 * bx + h, i = r + 1;
        if (r >= N1 || i < J.ring_lo || i > J.ring_hi) continue;
        const IBPM_GLOBAL double* row = gin(J.T2) + (long)r * J.ld2;
        double a = 0.0, b = 0.0;
        for (int l = t; l < N2; l += T) {
            const double x = row[l];
            a += x * sn[l];
            b += x * sn[N2 + l];
        }
        part[2 * h] = a;
        part[2 * h + 1] = b;
    }
    double* red = (double*)smem;   // 4 x T partial sums
#pragma unroll
    for (int q = 0; q < 4; ++q) red[q * T + t] = part[q];
    __syncthreads();
    if (t < 4) {
        const int r = 2 * bx + (t >> 1), i = r + 1;
        if (r < N1 && i >= J.ring_lo && i <= J.ring_hi) {
            double v = 0.0;
            for (int k = 0; k < T; ++k) v += red[t * T + k];
            const int j = (t & 1) ? J.rj1 : J.rj0;
            gout(J.u)[(long)r * N2 + (j - 1)] = v * J.norm;
        }
    }
}

template <int M>
__device__ __forceinline__ void pass3_r16(const EllJob& J, double2* smem, int bx) {
    static_assert(r16_group(3) == 1, "pass 3 writes rows: one pair per workgroup");
    constexpr int T = M / 16;
    const int N1 = J.nx - 1, N2 = M / 2 - 1, npairs = (N1 + 1) / 2;
    if (bx >= npairs) return;
    if (!((bx >= J.full_lo && bx <= J.full_hi) || bx == J.fq0 || bx == J.fq1)) {
        if (J.ring_lo <= J.ring_hi) pass3_ring<M>(J, smem, bx);
        return;
    }
    double2* buf = smem;
    const IBPM_GLOBAL double* tw = gin((const double*)J.pj.tw16);
    const IBPM_GLOBAL double* twg = gin((const double*)J.pj.tw);
    const int t = threadIdx.x;
    const int ra = 2 * bx, rb = ra + 1;
    const bool hb = rb < N1;
    double2 v[16];
    load_first_half<M>(v, gin(J.T2) + (long)ra * J.ld2, gin(J.T2) + (long)rb * J.ld2, hb);
    mirror_second_half<M>(v, buf);
    fft_regs<M>(v, buf, tw, twg, [&](int m, double2 z) {
        const int kk = t + T * m;
        if (kk == 0) return;
        const int k = kk - 1;
        // sinTransformInv normalisation, EllipticSolver2d.cc:80-83; with the
        // epilogue the finer level's BCs come from v (EllipticSolver.cc:62-66)
        {
            const long o2 = (long)ra * N2 + k;
            const double val_a = -z.y * J.norm;
            const double ua = J.epilogue ? gin(J.W)[o2] - val_a : val_a;
            gout(J.u)[o2] = ua;
            if (J.u2) gout(J.u2)[o2] = ua;
            if (J.vside && on_bc_ring(J, ra + 1, k + 1)) gout(J.vside)[o2] = val_a;
        }
        if (hb) {
            const long o2 = (long)rb * N2 + k;
            const double val_b = z.x * J.norm;
            const double ub = J.epilogue ? gin(J.W)[o2] - val_b : val_b;
            gout(J.u)[o2] = ub;
            if (J.u2) gout(J.u2)[o2] = ub;
            if (J.vside && on_bc_ring(J, rb + 1, k + 1)) gout(J.vside)[o2] = val_b;
        }
    });
}

// Logical (row block, job) of this workgroup.  Workgroups are handed to the
// 8 XCDs round-robin in launch order, so logical neighbours are mapped to the
// same XCD: consecutive row blocks write neighbouring 16-byte pieces of the
// same cache lines of the transposed output, and those pieces then meet in
// one XCD's L2 instead of being written back partially from eight.
__device__ inline int2 xcd_block() {
    constexpr int X = 8;
    const int gx = gridDim.x, total = gx * gridDim.y;
    const int flat = blockIdx.x + blockIdx.y * gx;
    const int xcd = flat % X, idx = flat / X, q = total / X, r = total % X;
    const int lf = (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
    return make_int2(lf % gx, lf / gx);
}

// Pass bodies: the register-held transform (REG, M = 1024 or 2048) or the LDS
// passes.
// FS: an instance for the force solve, whose register-held pass 1 also masks
// R off its box (EllJob::rb_*); the step's own solves never restrict it.
#define IBPM_PASS_BODY(N, R16)                                                             \
    template <int CM, bool REG, bool FS>                                                   \
    __device__ __forceinline__ void pass##N##_body(const EllJob& J, double2* smem, int bx) { \
        if constexpr (REG) R16;                                                            \
        else pass##N##_lds<CM>(J, smem, bx);                                               \
    }
IBPM_PASS_BODY(1, (pass1_r16<CM, FS>(J, smem, bx)))
IBPM_PASS_BODY(2, (pass2_r16<CM>(J, smem, bx)))
IBPM_PASS_BODY(3, (pass3_r16<CM>(J, smem, bx)))

// One kernel per pass and FFT length (compiled lengths, 0 = runtime plan), so
// each length gets its own register budget; a launch holds jobs of one length.
// The step's own solves use k_ell_pass*; the solves inside the force solve
// (M applications) use k_fs_pass*, the same code under another name so that
// profiles separate the two.  A job whose flow has finished its force solve
// does nothing (skipped()).
#define IBPM_PASS_KERNEL(NAME, BODY, PASS, FS)                                   \
    template <int CM, bool REG>                                                  \
    __global__ void __launch_bounds__((pass_threads<CM, REG, PASS>())) NAME(const EllJob* __restrict__ jobs) { \
        static_assert(!REG || r16_len<CM>(), "no register-held transform at this length"); \
        extern __shared__ double2 smem[];                                         \
        const int2 blk = xcd_block();                                             \
        const EllJob& J = jobs[blk.y];                                            \
        if (skipped(J.skip)) return;                                              \
        BODY<CM, REG, FS>(J, smem, blk.x);                                        \
    }
IBPM_PASS_KERNEL(k_ell_pass1, pass1_body, 1, false)
IBPM_PASS_KERNEL(k_ell_pass2, pass2_body, 2, false)
IBPM_PASS_KERNEL(k_ell_pass3, pass3_body, 3, false)
IBPM_PASS_KERNEL(k_fs_pass1, pass1_body, 1, true)
IBPM_PASS_KERNEL(k_fs_pass2, pass2_body, 2, true)
IBPM_PASS_KERNEL(k_fs_pass3, pass3_body, 3, true)

// FP32 instances of the LDS passes (every length, runtime plan for lengths
// without a compiled one): the step's own solves (k32_ell_pass*) and those
// inside the force solve (k32_fs_pass*).
#define IBPM_PASS32_KERNEL(NAME, LDSFN)                                                        \
    template <int CM>                                                                          \
    __global__ void __launch_bounds__(kFftThreads) NAME(const EllJob* __restrict__ jobs) {    \
        extern __shared__ float2 smem32[];                                                     \
        const int2 blk = xcd_block();                                                          \
        const EllJob& J = jobs[blk.y];                                                         \
        if (skipped(J.skip)) return;                                                           \
        LDSFN<CM, float2>(J, smem32, blk.x);                                                   \
    }
IBPM_PASS32_KERNEL(k32_ell_pass1, pass1_lds)
IBPM_PASS32_KERNEL(k32_ell_pass2, pass2_lds)
IBPM_PASS32_KERNEL(k32_ell_pass3, pass3_lds)
IBPM_PASS32_KERNEL(k32_fs_pass1, pass1_lds)
IBPM_PASS32_KERNEL(k32_fs_pass2, pass2_lds)
IBPM_PASS32_KERNEL(k32_fs_pass3, pass3_lds)

// --------------------------------------------------------------- coarsify
// Scalar::coarsify (reference Scalar.cc:46-63) for one level, writing only
// the central region into R; neighbours at level lev-1 come from R where the
// finer level is itself coarsified (central region of a level >= 1) and from
// the original field elsewhere.
// One workgroup per coarse row i of the central region; each thread makes two
// neighbouring points (j, j + 1) from the 3 x 5 fine values they share, with
// the same sums in the same order as Scalar::coarsify.
__device__ inline void coarsify_body(const CoarsifyJob& J) {
    if (skipped(J.skip)) return;
    const int nx = J.nx, ny = J.ny;
    const int ex = nx_ext(nx), ey = nx_ext(ny);
    const int wi = nx / 2 - 1, wj = ny / 2 - 1;
    const int row = blockIdx.x;
    if (row >= wi) return;
    const int i = ex + 1 + row, ii = (i - ex) * 2;   // Grid::c2f
    const int lf = J.lev - 1;
    const long lofs = (long)lf * (nx - 1) * (ny - 1);
    // fine rows ii - 1, ii, ii + 1: base pointers (value at column c is [c]) and
    // whether the row crosses the finer level's own central region
    const IBPM_GLOBAL double* fr[3];
    const IBPM_GLOBAL double* Rr[3];
    bool crow[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int r = ii - 1 + d;
        fr[d] = gin(J.f) + lofs + sidx(nx, ny, r, 1) - 1;
        Rr[d] = gin(J.R) + lofs + sidx(nx, ny, r, 1) - 1;
        crow[d] = lf >= 1 && r > ex && r < nx / 2 + ex;
    }
    IBPM_GLOBAL double* out = gout(J.R) + (long)J.lev * (nx - 1) * (ny - 1) + sidx(nx, ny, i, 1) - 1;
    const int npairs = (wj + 1) / 2;
    for (int q = threadIdx.x; q < npairs; q += blockDim.x) {
        const int j0 = ey + 1 + 2 * q;
        const bool two = j0 + 1 <= ey + wj;
        const int jj = (j0 - ey) * 2;
        double x[3][5];   // fine (ii - 1 + d, jj - 1 + c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const int col = jj - 1 + c;
                if (c >= 3 && !two) {
                    x[d][c] = 0.0;
                    continue;
                }
                const bool central = crow[d] && col > ey && col < ny / 2 + ey;
                x[d][c] = central ? Rr[d][col] : fr[d][col];
            }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            if (o == 1 && !two) break;
            const int c = 1 + 2 * o;   // centre column of this output
            const double v = 0.25 * x[1][c]
                + 0.125 * (x[2][c] + x[1][c + 1] + x[0][c] + x[1][c - 1])
                + 0.0625 * (x[2][c + 1] + x[2][c - 1] + x[0][c + 1] + x[0][c - 1]);
            out[j0 + o] = v;
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_coarsify(const CoarsifyJob* __restrict__ jobs) {
    coarsify_body(jobs[blockIdx.y]);
}
__global__ void __launch_bounds__(kThreads) k_fs_coarsify(const CoarsifyJob* __restrict__ jobs) {
    coarsify_body(jobs[blockIdx.y]);
}

// ------------------------------------------------- local Helmholtz sweeps
// Jacobi sweeps of A u = r, A = 1 + alpha Lap, over a job's box, for the
// force solve's A^-1.  With |alpha| / dx^2 small the solution decays by a
// factor 4|c| / (1 - 4c) per cell away from the right-hand side, so a box
// around the body's smeared force holds it to rounding and the host chooses
// the sweep count for an error below 1e-18 of the solution; it uses the DST
// solve wherever the box would reach a level's boundary ring.  The first
// sweep starts from u = 0, u is 0 off the box, and the right-hand side is
// rhs_value's without boundary terms, which the box keeps away from.
//
// One launch runs up to kHelmDepth sweeps on tiles of kHelmTile^2 points:
// a workgroup loads its tile with a kHelmDepth-point halo into LDS and
// sweeps there, the valid region shrinking by one point a sweep, then writes
// the tile.  Each point's update is the same arithmetic in the same order as
// a sweep over global memory, so the result does not depend on the blocking.
// Launches alternate between tmp and out so that the last writes out.
#ifndef IBPM_HELM_DEPTH
#define IBPM_HELM_DEPTH 6
#endif
constexpr int kHelmTile = 32, kHelmDepth = IBPM_HELM_DEPTH, kHelmReg = kHelmTile + 2 * kHelmDepth;
static_assert(kHelmDepth >= 1 && kHelmDepth <= 12, "sweeps per launch: 1..12 (LDS tiles with the halo)");

__global__ void __launch_bounds__(kThreads) k_fs_helm(const HelmJob* __restrict__ jobs, int sweep0, int nsw,
                                                     bool to_out) {
    const HelmJob& J = jobs[blockIdx.y];
    if (skipped(J.skip)) return;
    const int w = J.j1 - J.j0 + 1, h = J.i1 - J.i0 + 1;
    if (w <= 0 || h <= 0) return;
    const int tw = (w + kHelmTile - 1) / kHelmTile, th = (h + kHelmTile - 1) / kHelmTile;
    if ((int)blockIdx.x >= tw * th) return;
    const int ti0 = J.i0 + (blockIdx.x / tw) * kHelmTile - kHelmDepth;   // region origin
    const int tj0 = J.j0 + (blockIdx.x % tw) * kHelmTile - kHelmDepth;
    __shared__ double rr[kHelmReg * kHelmReg], ua[kHelmReg * kHelmReg], ub[kHelmReg * kHelmReg];
    const int ex = nx_ext(J.nx), ey = nx_ext(J.ny);
    const IBPM_GLOBAL double* src = gin(to_out ? J.tmp : J.out);
    for (int q = threadIdx.x; q < kHelmReg * kHelmReg; q += kThreads) {
        const int i = ti0 + q / kHelmReg, j = tj0 + q % kHelmReg;
        const bool inbox = i >= J.i0 && i <= J.i1 && j >= J.j0 && j <= J.j1;
        double r = 0.0, u = 0.0;
        if (inbox) {
            const long k = sidx(J.nx, J.ny, i, j);
            const bool central = J.lev >= 1 && i > ex && i < J.nx / 2 + ex && j > ey && j < J.ny / 2 + ey;
            const bool support = i >= J.s0 && i <= J.s1 && j >= J.t0 && j <= J.t1;
            r = support ? (central ? gin(J.R)[k] : gin(J.f)[k]) * J.in_scale : 0.0;
            if (sweep0 > 0) u = src[k];
        }
        rr[q] = r;
        ua[q] = u;
    }
    __syncthreads();
    double* cur = ua;
    double* nxt = ub;
    for (int d = 0; d < nsw; ++d) {
        const int m = d + 1;   // points within m of the region's edge are no longer valid
        for (int q = threadIdx.x; q < kHelmReg * kHelmReg; q += kThreads) {
            const int a = q / kHelmReg, b = q % kHelmReg;
            if (a < m || b < m || a >= kHelmReg - m || b >= kHelmReg - m) continue;
            const int i = ti0 + a, j = tj0 + b;
            const bool inbox = i >= J.i0 && i <= J.i1 && j >= J.j0 && j <= J.j1;
            double v = 0.0;
            if (inbox) {
                if (sweep0 + d == 0) {
                    v = rr[q] * J.inv_diag;
                } else {
                    const double s4 = (cur[q - kHelmReg] + cur[q + kHelmReg]) + (cur[q - 1] + cur[q + 1]);
                    v = (rr[q] - J.c * s4) * J.inv_diag;
                }
            }
            nxt[q] = v;
        }
        __syncthreads();
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    IBPM_GLOBAL double* dst = gout(to_out ? J.out : J.tmp);
    for (int q = threadIdx.x; q < kHelmTile * kHelmTile; q += kThreads) {
        const int a = kHelmDepth + q / kHelmTile, b = kHelmDepth + q % kHelmTile;
        const int i = ti0 + a, j = tj0 + b;
        if (i > J.i1 || j > J.j1) continue;
        dst[sidx(J.nx, J.ny, i, j)] = cur[a * kHelmReg + b];
    }
}

// u = W - v over each job's box, v the local result (J.out): the epilogue of
// the step's omega = omega* - A^-1 c, whose omega already holds omega* off the
// boxes (the omega* solve writes both).
__global__ void __launch_bounds__(kThreads) k_helm_epilogue(const HelmJob* __restrict__ jobs) {
    const HelmJob& J = jobs[blockIdx.y];
    const int w = J.j1 - J.j0 + 1, h = J.i1 - J.i0 + 1;
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (w <= 0 || h <= 0 || p >= w * h) return;
    const long k = sidx(J.nx, J.ny, J.i0 + p / w, J.j0 + p % w);
    gout(J.eo)[k] = gin(J.ei)[k] - gin(J.out)[k];
}

// The coarsify chain of the local solves' right-hand side, level lev of every
// job, over the job's support only: Scalar::coarsify's weights and order (as
// coarsify_body) with the finer level's value taken as 0 in its central
// region off its support (R holds other solves' values there) and from f
// outside the central region.
__global__ void __launch_bounds__(kThreads) k_fs_helm_coarsify(const HelmJob* __restrict__ jobs, int lev) {
    const HelmJob& J = jobs[blockIdx.y];
    if (J.lev != lev || skipped(J.skip)) return;
    const int w = J.t1 - J.t0 + 1, h = J.s1 - J.s0 + 1;
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (w <= 0 || h <= 0 || p >= w * h) return;
    const int i = J.s0 + p / w, j = J.t0 + p % w;
    const int ex = nx_ext(J.nx), ey = nx_ext(J.ny);
    const int ii = (i - ex) * 2, jj = (j - ey) * 2;   // Grid::c2f
    const int lf = lev - 1;
    auto fine = [&](int fi, int fj) {
        const long k = sidx(J.nx, J.ny, fi, fj);
        const bool central = lf >= 1 && fi > ex && fi < J.nx / 2 + ex && fj > ey && fj < J.ny / 2 + ey;
        if (!central) return gin(J.f_fine)[k];
        const bool sup = fi >= J.fs0 && fi <= J.fs1 && fj >= J.ft0 && fj <= J.ft1;
        return sup ? gin(J.R_fine)[k] : 0.0;
    };
    double x[3][3];   // fine (ii - 1 + d, jj - 1 + c)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) x[d][c] = fine(ii - 1 + d, jj - 1 + c);
    const double v = 0.25 * x[1][1] + 0.125 * (x[2][1] + x[1][2] + x[0][1] + x[1][0])
        + 0.0625 * (x[2][2] + x[2][0] + x[0][2] + x[0][0]);
    gout(J.R)[sidx(J.nx, J.ny, i, j)] = v;
}

// ------------------------------------------------ dense sine-matrix GEMM on MFMA
// A level solve as four dense products with the DST-I matrix S (S[a][b] =
// 2 sin(pi (a+1)(b+1) / n), symmetric), each C = A S of an Np x Np operand A
// (Np = N rounded up to 64, zero padded) held as two bf16 planes, hi =
// bf16(x) and lo = bf16(x - hi): C = Ah Sh + Ah Sl + Al Sh accumulated in FP32
// on the 16x16x32 bf16 MFMA (split-BF16 compensated products; the dropped
// Al Sl term and FP32 accumulation leave a relative error near 1e-6).
//   prep: F = the level's right-hand side (rhs_value) -> planes X0
//   G1: X1^T = (X0 S)^T              along j, written transposed
//   G2: X2^T = X1^T S, / eigenvalue  along i, the spectral scaling
//   G3: X3 = (X2^T S)^T              inverse along i, written transposed
//   G4: u = norm X3 S                inverse along j, FP64 output (pass 3's epilogue)
// A and S fragments are read straight from global memory (16 bytes per lane:
// 8 consecutive bf16 along k, as the MFMA takes them; S symmetric makes its
// columns rows); one wave computes a 32 x 64 tile of C.
typedef __bf16 Bf16x8 __attribute__((ext_vector_type(8)));
typedef float FloatX4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int gemm_np(int n) { return (n + 63) / 64 * 64; }

// The two planes of a level's operand inside the job's transposed scratch:
// plane pair 0 in T1, pair 1 in T2 (each 2 Np^2 bf16 <= the FP64 scratch).
__device__ inline __bf16* gemm_plane(const EllJob& J, int pair, int lohi) {
    const int Np = gemm_np(J.nx - 1);
    __bf16* base = (__bf16*)(pair == 0 ? J.T1 : J.T2);
    return base + (size_t)lohi * Np * Np;
}

__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// prep: X0 = rhs (FP64 -> FP32 -> hi, lo), zero in the padding.
__global__ void __launch_bounds__(kThreads) k_gemm_prep(const EllJob* __restrict__ jobs) {
    const EllJob& J = jobs[blockIdx.y];
    if (skipped(J.skip)) return;
    const int N1 = J.nx - 1, N2 = J.ny - 1, Np = gemm_np(N1);
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= (long)Np * Np) return;
    const int i = (int)(t / Np), j = (int)(t % Np);
    const float v = (i < N1 && j < N2) ? (float)rhs_value(J, i + 1, j + 1) : 0.0f;
    __bf16 hi, lo;
    split_bf16(v, hi, lo);
    gemm_plane(J, 0, 0)[t] = hi;
    gemm_plane(J, 0, 1)[t] = lo;
}

// One C = A S product of every job (stage 1..4 as above).  Grid: (Np / 64
// column tiles, Np / 32 row tiles, jobs); one wave per workgroup.
template <int STAGE>
__global__ void __launch_bounds__(64) k_gemm_dst(const EllJob* __restrict__ jobs, const __bf16* __restrict__ Sh,
                                                 const __bf16* __restrict__ Sl, int Np_max) {
    const EllJob& J = jobs[blockIdx.z];
    if (skipped(J.skip)) return;
    const int N1 = J.nx - 1, N2 = J.ny - 1, Np = gemm_np(N1);
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 32;
    if (c0 >= Np || r0 >= Np) return;
    const int in_pair = (STAGE == 1 || STAGE == 3) ? 0 : 1;
    const __bf16* Ah = gemm_plane(J, in_pair, 0);
    const __bf16* Al = gemm_plane(J, in_pair, 1);
    const int l = threadIdx.x, lr = l & 15, lk = (l >> 4) * 8;
    FloatX4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = FloatX4{0.f, 0.f, 0.f, 0.f};
    // fragments of step k0 are loaded while step k0 - 32 multiplies
    Bf16x8 ah[2], al[2], sh[4], sl[4];
    auto load = [&](int k0, Bf16x8* xh, Bf16x8* xl, Bf16x8* yh, Bf16x8* yl) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const size_t o = (size_t)(r0 + 16 * a + lr) * Np + k0 + lk;
            xh[a] = *(const Bf16x8*)(Ah + o);
            xl[a] = *(const Bf16x8*)(Al + o);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {   // S[k][c] = S[c][k]: row c, 8 consecutive k
            const size_t o = (size_t)(c0 + 16 * b + lr) * Np_max + k0 + lk;
            yh[b] = *(const Bf16x8*)(Sh + o);
            yl[b] = *(const Bf16x8*)(Sl + o);
        }
    };
    load(0, ah, al, sh, sl);
    for (int k0 = 0; k0 < Np; k0 += 32) {
        Bf16x8 nah[2], nal[2], nsh[4], nsl[4];
        const int kn = k0 + 32 < Np ? k0 + 32 : k0;   // the last step reloads its own (in bounds)
        load(kn, nah, nal, nsh, nsl);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], sh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], sl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], sh[b], acc[a][b], 0, 0, 0);
            }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            ah[a] = nah[a];
            al[a] = nal[a];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            sh[b] = nsh[b];
            sl[b] = nsl[b];
        }
    }
    // C[r][c] of block (a, b): r = r0 + 16 a + 4 (l / 16) + e, c = c0 + 16 b + l % 16
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + 16 * a + 4 * (l >> 4) + e, c = c0 + 16 * b + lr;
                float v = acc[a][b][e];
                if constexpr (STAGE == 4) {   // u = norm X3 S (+ pass 3's epilogue), rows i = r, columns j = c
                    if (r < N1 && c < N2) {
                        const long o = (long)r * N2 + c;
                        const double vv = (double)v * J.norm;
                        const double uv = J.epilogue ? gin(J.W)[o] - vv : vv;
                        gout(J.u)[o] = uv;
                        if (J.u2) gout(J.u2)[o] = uv;
                        if (J.vside && on_bc_ring(J, r + 1, c + 1)) gout(J.vside)[o] = vv;
                    }
                } else {
                    if constexpr (STAGE == 2) {   // X2^T[l][k] / eigenvalue (EllipticSolver2d.cc:42-54, 89-93)
                        if (r < N2 && c < N1) {
                            const double lam = 2. * (gin(J.cosx)[c] + gin(J.cosy)[r] - 2.) * J.bydx2;
                            v = (float)((double)v * ((J.kind == 0) ? 1. / lam : 1. / (1 + J.alpha * lam)));
                        } else {
                            v = 0.0f;
                        }
                    }
                    __bf16 hi, lo;
                    split_bf16(v, hi, lo);
                    const size_t o = (STAGE == 2) ? (size_t)r * Np + c : (size_t)c * Np + r;   // 1, 3 transposed
                    gemm_plane(J, 1 - in_pair, 0)[o] = hi;
                    gemm_plane(J, 1 - in_pair, 1)[o] = lo;
                }
            }
}

// GEMM-DST of every job of one group (N1 = N2 = N of the plan, <= 511).
// Sh, Sl: the split sine matrix of that N, Np_max columns per row.
void launch_gemm_dst(const EllJob* d_jobs, int njobs, int n, const __bf16* Sh, const __bf16* Sl, hipStream_t s) {
    const int Np = gemm_np(n);
    hipLaunchKernelGGL(k_gemm_prep, dim3((unsigned)(((long)Np * Np + kThreads - 1) / kThreads), njobs),
                       dim3(kThreads), 0, s, d_jobs);
    const dim3 grid(Np / 64, Np / 32, njobs);
    hipLaunchKernelGGL(k_gemm_dst<1>, grid, dim3(64), 0, s, d_jobs, Sh, Sl, Np);
    hipLaunchKernelGGL(k_gemm_dst<2>, grid, dim3(64), 0, s, d_jobs, Sh, Sl, Np);
    hipLaunchKernelGGL(k_gemm_dst<3>, grid, dim3(64), 0, s, d_jobs, Sh, Sl, Np);
    hipLaunchKernelGGL(k_gemm_dst<4>, grid, dim3(64), 0, s, d_jobs, Sh, Sl, Np);
}

// ------------------------------------------------------------ host launch
bool pass_in_registers(int M, long row_pairs) {
    return M == 2048 || (M == 1024 && row_pairs >= kRegMinPairs);
}

int lds_for_pass(const EllJob& J, int pass, bool regs) {
    const FftPlan& P = (pass == 2) ? J.pi : J.pj;
    if (regs) return (int)sizeof(double2) * r16_group(pass) * r16_buf(P.M);   // 8.5 or 17 KB per pair
    const int ntw = (compiled_len(P.M) && quarter_twiddles(P.M)) ? P.M / 4 : P.M;
    return (int)sizeof(double2) * (P.np * P.M + ntw);
}

int blocks_for_pass(const EllJob& J, int pass, bool regs) {
    const int rows_total = (pass == 2) ? J.ny - 1 : J.nx - 1;
    const FftPlan& P = (pass == 2) ? J.pi : J.pj;
    const int rows = regs ? 2 * r16_group(pass) : 2 * P.np;
    return (rows_total + rows - 1) / rows;
}

#define IBPM_LAUNCH_LEN(KER, PASS)                                                             \
    switch (M) {                                                                               \
    case 400: hipLaunchKernelGGL((KER<400, false>), grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;  \
    case 512: hipLaunchKernelGGL((KER<512, false>), grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;  \
    case 768: hipLaunchKernelGGL((KER<768, false>), grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;  \
    case 800: hipLaunchKernelGGL((KER<800, false>), grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;  \
    case 1024:                                                                                 \
        if (regs) hipLaunchKernelGGL((KER<1024, true>), grid, dim3(r16_group(PASS) * 1024 / 16), lds_bytes, s, d_jobs); \
        else hipLaunchKernelGGL((KER<1024, false>), grid, dim3(kFftThreads), lds_bytes, s, d_jobs);     \
        break;                                                                                 \
    case 2048: hipLaunchKernelGGL((KER<2048, true>), grid, dim3(r16_group(PASS) * 2048 / 16), lds_bytes, s, d_jobs); break; \
    default: hipLaunchKernelGGL((KER<0, false>), grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;      \
    }

int lds_for_pass32(const EllJob& J, int pass) {
    const FftPlan& P = (pass == 2) ? J.pi : J.pj;
    const int ntw = (compiled_len(P.M) && quarter_twiddles(P.M)) ? P.M / 4 : P.M;
    return (int)sizeof(float2) * (P.np * P.M + ntw);
}

#define IBPM_LAUNCH32_LEN(KER)                                                                    \
    switch (M) {                                                                                  \
    case 400: hipLaunchKernelGGL(KER<400>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;   \
    case 512: hipLaunchKernelGGL(KER<512>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;   \
    case 768: hipLaunchKernelGGL(KER<768>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;   \
    case 800: hipLaunchKernelGGL(KER<800>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;   \
    case 1024: hipLaunchKernelGGL(KER<1024>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break; \
    case 2048: hipLaunchKernelGGL(KER<2048>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break; \
    default: hipLaunchKernelGGL(KER<0>, grid, dim3(kFftThreads), lds_bytes, s, d_jobs); break;      \
    }

// FP32 passes (LDS transforms); max_blocks from blocks_for_pass(J, pass,
// false), lds_bytes from lds_for_pass32.
void launch_ell_pass32(int pass, const EllJob* d_jobs, int njobs, int max_blocks, int lds_bytes, hipStream_t s,
                       bool force_solve, int M) {
    dim3 grid(max_blocks, njobs);
    if (pass == 1) {
        if (force_solve) { IBPM_LAUNCH32_LEN(k32_fs_pass1) } else { IBPM_LAUNCH32_LEN(k32_ell_pass1) }
    } else if (pass == 2) {
        if (force_solve) { IBPM_LAUNCH32_LEN(k32_fs_pass2) } else { IBPM_LAUNCH32_LEN(k32_ell_pass2) }
    } else {
        if (force_solve) { IBPM_LAUNCH32_LEN(k32_fs_pass3) } else { IBPM_LAUNCH32_LEN(k32_ell_pass3) }
    }
}

// M: the FFT length of this pass (J.pj.M for passes 1 and 3, J.pi.M for
// pass 2), the same for every job of the launch; regs: pass_in_registers for
// the launch, with max_blocks and lds_bytes sized for that choice.
void launch_ell_pass(int pass, const EllJob* d_jobs, int njobs, int max_blocks, int lds_bytes, hipStream_t s,
                     bool force_solve, int M, bool regs) {
    static_assert(compiled_len(400) && compiled_len(512) && compiled_len(768) && compiled_len(800) &&
                      compiled_len(1024) && compiled_len(2048),
                  "the launch switch and compiled_len must list the same lengths");
    dim3 grid(max_blocks, njobs);
    if (pass == 1) {
        if (force_solve) { IBPM_LAUNCH_LEN(k_fs_pass1, 1) } else { IBPM_LAUNCH_LEN(k_ell_pass1, 1) }
    } else if (pass == 2) {
        if (force_solve) { IBPM_LAUNCH_LEN(k_fs_pass2, 2) } else { IBPM_LAUNCH_LEN(k_ell_pass2, 2) }
    } else {
        if (force_solve) { IBPM_LAUNCH_LEN(k_fs_pass3, 3) } else { IBPM_LAUNCH_LEN(k_ell_pass3, 3) }
    }
}

void launch_helm_coarsify(const HelmJob* d_jobs, int njobs, int max_blocks, int lev, hipStream_t s) {
    hipLaunchKernelGGL(k_fs_helm_coarsify, dim3(max_blocks, njobs), dim3(kThreads), 0, s, d_jobs, lev);
}

// out = tmp = 0 over each job's box.
__global__ void __launch_bounds__(kThreads) k_helm_clear(const HelmJob* __restrict__ jobs) {
    const HelmJob& J = jobs[blockIdx.y];
    const int w = J.j1 - J.j0 + 1, h = J.i1 - J.i0 + 1;
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (w <= 0 || h <= 0 || p >= w * h) return;
    const long k = sidx(J.nx, J.ny, J.i0 + p / w, J.j0 + p % w);
    gout(J.out)[k] = 0.0;
    gout(J.tmp)[k] = 0.0;
}

void launch_helm_clear(const HelmJob* d_jobs, int njobs, int max_blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_helm_clear, dim3(max_blocks, njobs), dim3(kThreads), 0, s, d_jobs);
}

void launch_helm_epilogue(const HelmJob* d_jobs, int njobs, int max_blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_helm_epilogue, dim3(max_blocks, njobs), dim3(kThreads), 0, s, d_jobs);
}

int r16_pass_group(int pass) { return r16_group(pass); }

int helm_launches(int nsweeps) { return (nsweeps + kHelmDepth - 1) / kHelmDepth; }

int helm_tiles(int h, int w) { return ((h + kHelmTile - 1) / kHelmTile) * ((w + kHelmTile - 1) / kHelmTile); }

void launch_helm_sweeps(const HelmJob* d_jobs, int njobs, int max_tiles, int nsweeps, hipStream_t s,
                        const std::function<void(const std::function<void()>&)>& timed_launch) {
    const int nl = (nsweeps + kHelmDepth - 1) / kHelmDepth;
    for (int n = 0; n < nl; ++n) {
        const int s0 = n * kHelmDepth, nsw = std::min(kHelmDepth, nsweeps - s0);
        const bool to_out = ((nl - 1 - n) & 1) == 0;
        timed_launch([&] {
            hipLaunchKernelGGL(k_fs_helm, dim3(max_tiles, njobs), dim3(kThreads), 0, s, d_jobs, s0, nsw, to_out);
        });
    }
}

void launch_coarsify(const CoarsifyJob* d_jobs, int njobs, int max_rows, hipStream_t s, bool force_solve) {
    dim3 grid(max_rows, njobs);
    hipLaunchKernelGGL(force_solve ? k_fs_coarsify : k_coarsify, grid, dim3(kThreads), 0, s, d_jobs);
}

}  // namespace ibpmg
