// fft.hpp -- batched Stockham FFT engine in LDS for gfx950.
//
// A workgroup holds C independent length-N complex columns in LDS, laid
// out [a][LD] (element a of column c at a*LD + c), and transforms them in
// place with TPC threads per column (thread id = t*C + c, so the C threads
// that share t touch C adjacent words: bank-conflict free).  Stages are
// radix-3/8/4/2 Stockham autosort steps (natural order in and out):
//   stage of radix R after NS points are done, butterfly b in [0, N/R):
//     k = b mod NS, inputs x[b + r N/R] * W_{NS R}^{r k},
//     DFT_R, outputs y[(b div NS) NS R + k + r NS].
// Twiddles come from a per-N table W_N^j = exp(-2 pi i j/N) computed on the
// host in double and staged in LDS by each workgroup; the inverse uses the
// conjugate.  Transforms are
// unnormalised, forward e^{-i}, inverse e^{+i}, as cuFFT's used by the
// reference (multi/main.f90:853-912).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace mhit {

template <typename T> struct CplxOf;
template <> struct CplxOf<float> { using type = float2; };
template <> struct CplxOf<double> { using type = double2; };
template <typename T> using cplx_t = typename CplxOf<T>::type;

template <typename V> __device__ __forceinline__ V c_add(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __device__ __forceinline__ V c_sub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V> __device__ __forceinline__ V c_mul(V a, V b) {
    V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// a b with two fused multiply-adds per component product pair (the
// wavefront-form transforms' twiddles): 4 operations instead of 6
template <typename V> __device__ __forceinline__ V c_mul_fma(V a, V b) {
    V r; r.x = fma(a.x, b.x, -(a.y * b.y)); r.y = fma(a.x, b.y, a.y * b.x); return r;
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename V> __device__ __forceinline__ V c_rot(V a) {
    V r;
    if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
    return r;
}

// Radix plan 0 (the default): factors of 3 first, then radix 8 while a
// radix-8 stage can still end the plan, a 2 or a 4 in the middle, and 8
// last whenever the power-of-two part allows (512 = 8.8.8, 256 = 8.4.8,
// 128 = 8.2.8, 96 = 3.4.8).  A radix-8 first or last stage with N/8
// threads per column takes its inputs from / leaves its outputs in
// registers in the order of the global loads / stores (see k_fft_cols).
// Plan 1 takes radix 4 until 64 points remain, then 8.8 (512 = 4.4.4.8,
// 256 = 4.8.8, 128 = 4.4.8, 384 = 3.4.4.8): a stage more at 512 and 128,
// with two radix-4 butterflies per thread instead of one radix-8
// (Tuning::fft_xplan, the x passes only).
constexpr int next_radix(int rem, int plan = 0) {
    return (rem % 3 == 0) ? 3
         : (rem == 8 || (plan == 0 ? rem % 64 == 0 : rem == 64)) ? 8
         : (plan == 1 && rem % 64 == 0) ? 4
         : (rem == 16) ? 2
         : (rem == 32) ? 4
         : (rem == 4) ? 4 : 2;
}
constexpr int last_radix(int n, int plan = 0) {
    int ns = 1, r = 1;
    while (ns < n) { r = next_radix(n / ns, plan); ns *= r; }
    return r;
}
// whether the two plans differ for n
constexpr bool plans_differ(int n) {
    int a = 1, b = 1;
    while (a < n && b < n) {
        const int ra = next_radix(n / a, 0), rb = next_radix(n / b, 1);
        if (a != b || ra != rb) return true;
        a *= ra;
        b *= rb;
    }
    return a != b;
}
template <int REM, int PLAN = 0> struct NextRadix {
    static constexpr int value = next_radix(REM, PLAN);
};
template <bool INV, typename V> __device__ __forceinline__ void dft2(V* a) {
    V t = a[0];
    a[0] = c_add(t, a[1]);
    a[1] = c_sub(t, a[1]);
}

template <bool INV, typename V> __device__ __forceinline__ void dft3(V* a) {
    using T = decltype(a[0].x);
    const T s = (T)0.86602540378443864676;
    V t1 = c_add(a[1], a[2]);
    V t2 = c_sub(a[1], a[2]);
    V m1; m1.x = a[0].x - (T)0.5 * t1.x; m1.y = a[0].y - (T)0.5 * t1.y;
    V st; st.x = s * t2.x; st.y = s * t2.y;
    V r = c_rot<INV>(st);
    a[0] = c_add(a[0], t1);
    a[1] = c_add(m1, r);
    a[2] = c_sub(m1, r);
}

template <bool INV, typename V> __device__ __forceinline__ void dft4v(V& a0, V& a1, V& a2, V& a3) {
    V s02 = c_add(a0, a2), d02 = c_sub(a0, a2);
    V s13 = c_add(a1, a3), d13 = c_rot<INV>(c_sub(a1, a3));
    a0 = c_add(s02, s13);
    a2 = c_sub(s02, s13);
    a1 = c_add(d02, d13);
    a3 = c_sub(d02, d13);
}

template <bool INV, typename V> __device__ __forceinline__ void dft4(V* a) {
    dft4v<INV>(a[0], a[1], a[2], a[3]);
}

template <bool INV, typename V> __device__ __forceinline__ void dft8(V* a) {
    using T = decltype(a[0].x);
    const T c = (T)0.70710678118654752440;
    V e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    V o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4v<INV>(e0, e1, e2, e3);
    dft4v<INV>(o0, o1, o2, o3);
    V t1, t3;
    if (!INV) {  // W8^1 = c(1-i), W8^3 = -c(1+i)
        t1.x = c * (o1.x + o1.y); t1.y = c * (o1.y - o1.x);
        t3.x = c * (o3.y - o3.x); t3.y = -c * (o3.x + o3.y);
    } else {     // conjugates
        t1.x = c * (o1.x - o1.y); t1.y = c * (o1.x + o1.y);
        t3.x = -c * (o3.x + o3.y); t3.y = c * (o3.x - o3.y);
    }
    V t2 = c_rot<INV>(o2);
    a[0] = c_add(e0, o0); a[4] = c_sub(e0, o0);
    a[1] = c_add(e1, t1); a[5] = c_sub(e1, t1);
    a[2] = c_add(e2, t2); a[6] = c_sub(e2, t2);
    a[3] = c_add(e3, t3); a[7] = c_sub(e3, t3);
}

template <int R, bool INV, typename V> __device__ __forceinline__ void dft(V* a) {
    if constexpr (R == 2) dft2<INV>(a);
    else if constexpr (R == 3) dft3<INV>(a);
    else if constexpr (R == 4) dft4<INV>(a);
    else dft8<INV>(a);
}

// LDS index of element a of column c: rows of LD values.  With SW != 0
// the column is XOR-swizzled by the row (c ^ ((a >> 1) & SW), SW = C - 1,
// LD = C): the values a 16-lane group reads together (8-byte values; 8 lanes
// for 16-byte values) then sit in distinct banks both when the lanes walk
// along a row and when they walk down a column (the x passes do both).
template <int LD, int SW>
__device__ __forceinline__ int lidx(int a, int c) {
    if constexpr (SW != 0) return a * LD + (c ^ ((a >> 1) & SW));
    else return a * LD + c;
}

// One Stockham stage on C columns in LDS.
template <typename V, int N, int R, int NS, int C, int TPC, int LD, bool INV, int SW = 0>
__device__ __forceinline__ void fft_stage(V* lds, int c, int t, const V* __restrict__ tw) {
    constexpr int NBF = N / R;
    constexpr int NB = (NBF + TPC - 1) / TPC;
    V v[NB][R];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = t + q * TPC;
        if (NBF % TPC == 0 || b < NBF) {
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = lds[lidx<LD, SW>(b + r * NBF, c)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = t + q * TPC;
        if (NBF % TPC == 0 || b < NBF) {
            const int k = b % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    V w = tw[(r * k) * (N / (NS * R))];
                    if (INV) w.y = -w.y;
                    v[q][r] = c_mul(v[q][r], w);
                }
            }
            dft<R, INV>(v[q]);
            const int j = (b / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<LD, SW>(j + r * NS, c)] = v[q][r];
        }
    }
    __syncthreads();
}

// Stages from NS (points already combined) up to NS_END, each in LDS, in
// radix plan PLAN.
template <typename V, int N, int NS, int C, int TPC, int LD, bool INV, int NS_END = N, int SW = 0, int PLAN = 0>
__device__ __forceinline__ void fft_stages(V* lds, int c, int t, const V* __restrict__ tw) {
    if constexpr (NS < NS_END) {
        constexpr int R = NextRadix<N / NS, PLAN>::value;
        fft_stage<V, N, R, NS, C, TPC, LD, INV, SW>(lds, c, t, tw);
        fft_stages<V, N, NS * R, C, TPC, LD, INV, NS_END, SW, PLAN>(lds, c, t, tw);
    }
}

// Register-direct stages for N/8 threads per column (thread t of column c).
// First stage (radix 8, NS = 1): inputs x[t + r N/8] in v[r], outputs to
// LDS at 8t + r.
template <typename V, int N, int LD, bool INV>
__device__ __forceinline__ void fft_first8_to_lds(V* lds, int c, int t, V (&v)[8]) {
    dft8<INV>(v);
#pragma unroll
    for (int r = 0; r < 8; r++) lds[(8 * t + r) * LD + c] = v[r];
}
// Last stage (radix 8, NS = N/8): inputs from LDS at t + r N/8, outputs
// X[t + r N/8] in v[r].  The caller places a barrier before LDS is written
// again.
template <typename V, int N, int LD, bool INV>
__device__ __forceinline__ void fft_last8_from_lds(const V* lds, int c, int t, V (&v)[8],
                                                   const V* __restrict__ tw) {
    constexpr int NS = N / 8;
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = lds[(t + r * NS) * LD + c];
    if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < 8; r++) {
            V w = tw[r * t];            // W_N^{r k}, k = t (b = t < NS)
            if (INV) w.y = -w.y;
            v[r] = c_mul(v[r], w);
        }
    }
    dft8<INV>(v);
}

// ------------------------------------------------ wavefront exchanges
// Transpose of the lane digit b = lane bits 3-5 with the register digit r
// (index 0..7) of eight values per lane, inside the wavefront: afterwards
// lane (b, l) holds in v[r] what lane (r, l) held in v[b] (l = lane bits
// 0-2).  Three bit swaps: lane bit 5 with index bit 2 by
// v_permlane32_swap (the upper 32 lanes of the first operand trade places
// with the lower 32 of the second), lane bit 4 with index bit 1 by
// v_permlane16_swap (the same on 16-lane rows), lane bit 3 with index bit
// 0 by a DPP row rotation of 8 lanes (lane ^ 8 within each 16-lane row)
// and a select.  No LDS, no barrier.
template <typename V>
__device__ __forceinline__ void xpose_lane_hi(V (&v)[8]) {
    constexpr int DW = sizeof(V) / 4;
    unsigned d[8][DW];
#pragma unroll
    for (int r = 0; r < 8; r++) __builtin_memcpy(d[r], &v[r], sizeof(V));
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < DW; q++) {
            const auto p = __builtin_amdgcn_permlane32_swap(d[r][q], d[r + 4][q], false, false);
            d[r][q] = p[0];
            d[r + 4][q] = p[1];
        }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        if (r & 2) continue;
#pragma unroll
        for (int q = 0; q < DW; q++) {
            const auto p = __builtin_amdgcn_permlane16_swap(d[r][q], d[r + 2][q], false, false);
            d[r][q] = p[0];
            d[r + 2][q] = p[1];
        }
    }
    const bool odd = (__lane_id() >> 3) & 1;
#pragma unroll
    for (int r = 0; r < 8; r += 2)
#pragma unroll
        for (int q = 0; q < DW; q++) {
            // lanes with bit 3 clear send their odd value and take the
            // partner's even one into it; lanes with bit 3 set the reverse
            const unsigned s = odd ? d[r][q] : d[r + 1][q];
            const unsigned x = (unsigned)__builtin_amdgcn_update_dpp(0, (int)s, 0x128 /* row_ror:8 */,
                                                                     0xF, 0xF, false);
            if (odd) d[r][q] = x;
            else d[r + 1][q] = x;
        }
#pragma unroll
    for (int r = 0; r < 8; r++) __builtin_memcpy(&v[r], d[r], sizeof(V));
}

// Copy the twiddle table W_N^j, j < N, into LDS.  Stages read their
// twiddles from there: the C threads of one butterfly row share an address
// (broadcast), and no stage waits on a global load.
template <typename V, int N, int NT>
__device__ __forceinline__ void stage_twiddles(V* stw, const MHIT_GLOBAL V* __restrict__ tw, int tid) {
#pragma unroll
    for (int r = 0; r < (N + NT - 1) / NT; r++) {
        const int e = tid + r * NT;
        if (N % NT == 0 || e < N) stw[e] = gload(tw + e);
    }
}
// ---------------------------------------------------------------- config
constexpr int pow2_divisor_cap(int x, int cap) {
    int c = 1;
    while (c * 2 <= cap && x % (c * 2) == 0) c *= 2;
    return c;
}

template <typename T, int N> struct FftCfg {
    static constexpr int H = N / 2;
    static constexpr bool F32 = sizeof(T) == 4;
    // columns per workgroup for the strided (y, z) passes: 128-byte rows
    static constexpr int CY = pow2_divisor_cap(H, F32 ? 16 : 8);
    // row pairs per workgroup for the x passes
    static constexpr int PX = pow2_divisor_cap(H, F32 ? 8 : 4);
    // threads per column
    static constexpr int TPC = (N % 8 == 0) ? N / 8 : ((N % 4 == 0) ? N / 4 : N / 2);
};

// Whether a column pass of C columns can take the wavefront form (eight
// strided columns, N = 64 M with M = 1, 2, 4 or 8, N/8 threads per column).
template <typename T, int N, int C, bool AXIS_CONTIG>
constexpr bool wave_form_fits() {
    return !AXIS_CONTIG && C == 8 && N % 64 == 0 && FftCfg<T, N>::TPC == N / 8 &&
           (N == 64 || N == 128 || N == 256 || N == 512);
}

}  // namespace mhit
