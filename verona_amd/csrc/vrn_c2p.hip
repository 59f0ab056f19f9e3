// vrn_c2p.hip -- batched con2prim with a learned Newton initial guess
// (gfx950), and the C ABI of include/verona_c2p.h.
//
// Recovery.  Every cell runs the 1DW Newton iteration of Verona's UtoP
// (src/3dim/structs.jl:273-343: W_min clamp, v^2 clamp, halving towards
// W_min) on W = w Gamma^2, stopped when |dW| < rtol * W.  Only the initial
// W differs between the guess modes.  A cell the iteration leaves unfinished
// goes through the rest of the reference cascade (c2p_cascade_tail).
//
// Learned guess.  For an ideal gas W/E (E = -T^t_t) depends only on D/E and
// |S|/E, and W = E + p.  The network predicts y0 = log(q / z) with
// q = log(W/E) and z = (E^2 - |S|^2 - D^2) / E^2 >= 0, which to first order
// in p is 2 rho p Gamma^2 / ((gamma - 1) E^2): y0 stays O(1) over the whole
// state space, and the guess W0 = E (1 + expm1(z e^y0)) never falls below the
// lower bound E, so the relative guess error is (error of y0) * q -- small
// exactly where the reference's W_min clamp makes Newton slow (Gamma >> 1).
// The second output is log Gamma.  Inputs: D/E, |S|/E, sqrt(z) and ln(z)/10;
// none depends on the direction of S.
//
// Inference runs inside the solve kernel: a wave takes 64 cells, computes
// their features, and evaluates the 4 -> 16 -> 16 -> 2 network on the matrix
// cores (v_mfma_f32_16x16x16_bf16) in the transposed orientation H^T = W X^T,
// one 16-cell column tile at a time: every lane supplies its own cell's
// inputs, and the accumulator tile of one layer (rows = hidden units,
// columns = cells) is, element for element, the B operand of the next -- no
// LDS round trip and no shuffle before the outputs.  BF16 values: inputs,
// W1, W2, W3 and both hidden layers; FP32: the biases (accumulator
// initialisation) and accumulation.
//
// Training (FP32 VALU): one thread per sample runs forward and backward with
// the parameters in LDS and caches per-sample activations and deltas; weight
// gradients are sums over samples, formed by a thread per parameter over a
// fixed slice of the batch and then a fixed-order sum over slices, so the
// result is deterministic; Adam runs one thread per parameter.
#include "../../include/verona_c2p.h"
#define VRN_NS vrn_c4
#include "vrn_physics.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace vrn_c4 {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = VRN_MLP_NIN, NH = VRN_MLP_NHID, NOUT = VRN_MLP_NOUT, NPARAM = VRN_MLP_NPARAM;
constexpr int OFF_W1 = 0, OFF_B1 = NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NOUT * NH;
static_assert(OFF_B3 + NOUT == NPARAM, "parameter packing");

constexpr int RECORD_FLAG = 0x100;

// ---------------------------------------------------------------- BF16
// Round to nearest even (the rounding of v_cvt_pk_bf16_f32); NaN stays NaN.
__host__ __device__ __forceinline__ unsigned short bf16_bits(float f)
{
    unsigned u = __builtin_bit_cast(unsigned, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_float(unsigned short b)
{
    return __builtin_bit_cast(float, (unsigned)b << 16);
}
__host__ __device__ __forceinline__ float bf16_round(float f) { return bf16_float(bf16_bits(f)); }

// ---------------------------------------------------------------- features
struct Feat {
    float x[NIN];
    double E, z;
    bool ok;
};

// Features of one cell of conserved variables (D, T^t_t, S^2): D/E, |S|/E,
// sqrt(z) and ln(z)/10 -- W/E depends on (D/E, |S|/E) alone, and z is the
// distance from the p = 0 boundary where W/E -> 1.
__device__ __forceinline__ Feat c4_features(double D, double Qt, double S2)
{
    Feat f;
    f.E = -Qt;
    const double tau = f.E - D;
    // 1/E: hardware reciprocal + two Newton-Raphson steps (about 1 ulp; E is
    // a normal positive number wherever the features are used)
    double rE = __builtin_amdgcn_rcp(f.E);
    rE = __builtin_fma(__builtin_fma(-f.E, rE, 1.0), rE, rE);
    rE = __builtin_fma(__builtin_fma(-f.E, rE, 1.0), rE, rE);
    // E^2 - S^2 - D^2 = tau (E + D) - S^2, in FP64: it is a difference of
    // nearly equal terms when p << E
    f.z = (tau * (f.E + D) - S2) * (rE * rE);
    // (an infinite or NaN E or S^2 makes z NaN or -inf, so 0 < z < 1 also
    // rejects non-finite inputs)
    f.ok = D > 0.0 && tau > 0.0 && f.z > 0.0 && f.z < 1.0;
    if (f.ok) {
        // single hardware instructions (v_sqrt_f32, v_log_f32 = log2; ~1 ulp):
        // the features only steer the initial guess
        const float LN2 = 0.693147181f;
        const float zf = (float)f.z;
        f.x[0] = (float)(D * rE);
        f.x[1] = __builtin_amdgcn_sqrtf((float)S2) * (float)rE;
        f.x[2] = __builtin_amdgcn_sqrtf(zf);
        f.x[3] = (0.1f * LN2) * __builtin_amdgcn_logf(zf);
    } else {
#pragma unroll
        for (int k = 0; k < NIN; k++) f.x[k] = 0.0f;
    }
    return f;
}

// Training targets (log(q/z), log Gamma) of a (U, P) pair; false outside the
// model's domain.
__device__ __forceinline__ bool c4_targets(const double *P, double g, const Feat &f, float *t)
{
    const double p = (g - 1.0) * P[1];
    const double gam = vrn_sqrt1(1.0 + P[2] * P[2] + P[3] * P[3] + P[4] * P[4]);
    const double q = log1p(p / f.E);
    const double t0 = log(q / f.z), t1 = log(gam);
    t[0] = (float)t0;
    t[1] = (float)t1;
    return f.ok && q > 0.0 && __builtin_isfinite(t0) && __builtin_isfinite(t1);
}

// Initial W from the network's first output.
__device__ __forceinline__ double learned_w0(const Feat &f, float y0)
{
    const float LOG2E = 1.44269504f;
    float qh = (float)f.z * __builtin_amdgcn_exp2f(y0 * LOG2E);
    qh = qh < 0.0f ? 0.0f : (qh > 0.5f ? 0.5f : qh);
    // expm1(qh) on [0, 1/2]: Taylor polynomial to 6th order, relative error
    // below qh^6 / 5040 (3e-6 at the clamp, 1e-7 at q = 0.29, the largest
    // log(W/E) of an ideal gas with gamma = 4/3), with no branch
    const float em =
        qh * (1.0f + qh * (0.5f + qh * (1.0f / 6.0f + qh * (1.0f / 24.0f + qh * (1.0f / 120.0f + qh * (1.0f / 720.0f))))));
    return f.E + f.E * (double)em;
}

// ---------------------------------------------------------------- Newton
// 1DW iteration of src/3dim/structs.jl:292-342 with the relative stop
// |dW| < rtol * W.  W holds the initial guess on entry and the root on exit.
__device__ __forceinline__ bool newton_1dw_rel(double D, double Qt, double S2, double g, int n_iter,
                                               double rtol, double &W, int &iters)
{
    const double c = (g - 1.0) / g;
    const double W_max = 1e30;
    const double W_min = vrn_sqrt(S2) * (1.0 + 1e-15);
    for (int t10 = 0; t10 < 1100 && S2 / (W * W) >= 1.0 && W < W_max; t10++) W *= 10.0;
    for (int it = 0; it < n_iter; it++) {
        iters++;
        if (W < W_min) W = W_min;
        double v2 = S2 / (W * W);
        if (v2 < 0.0)
            v2 = 0.0;
        else if (v2 > 1.0 - 1e-15)
            v2 = 1.0 - 1e-15;
        const double sv = vrn_sqrt(1.0 - v2);
        const double fun = Qt + W - c * (W * (1.0 - v2) - D * sv);
        const double W3 = W * W * W;
        const double jac = 1.0 - c * (1.0 - v2) - c * (W * (2.0 * S2 / W3) - D * (S2 / (W3 * sv)));
        double dW = fun / jac;
        const double Wp = W - dW;
        W = (Wp < W_min) ? 0.5 * (W + W_min) : Wp;
        if (dW < 0.0) dW = -dW;
        if (dW < rtol * W) return true;
    }
    return false;
}

// Primitives from the converged W (src/3dim/structs.jl:598-609 without the
// Gamma <= 50 cap).
__device__ __forceinline__ void recover(double D, double Qx, double Qy, double Qz, double S2, double W,
                                        double g, double *P)
{
    const double v2 = S2 / (W * W);
    const double gam = 1.0 / vrn_sqrt(1.0 - v2);
    P[0] = jmax(1e-8, D / gam);
    P[1] = jmax(1e-8, (W / (gam * gam) - D / gam) / g);
    P[2] = gam * Qx / W;
    P[3] = gam * Qy / W;
    P[4] = gam * Qz / W;
}

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// Wave maximum of non-negative counts, built bit by bit from the top with one
// ballot per bit (counts below 16 take five compares).
__device__ __forceinline__ int wave_max_count(int v)
{
    if (__ballot(v >= 16)) return wave_max_i(v);
    int m = 0;
#pragma unroll
    for (int b = 3; b >= 0; b--) {
        const int c = m | (1 << b);
        if (__ballot(v >= c)) m = c;
    }
    return m;
}

// Per-lane histogram of iteration counts 1..15: 8-bit fields, bins 0..7 in
// w0 and 8..15 in w1.  flush() adds the wave's totals to the block bins; it
// must run (wave-uniformly) before any lane has counted 256 cells.
struct LaneHist {
    unsigned long long w0 = 0, w1 = 0;
    __device__ __forceinline__ void add(int it)
    {
        const unsigned long long inc = 1ull << (8 * (it & 7));
        if (it < 8)
            w0 += inc;
        else
            w1 += inc;
    }
    __device__ __forceinline__ void flush(unsigned *bins, int lane)
    {
        // widen to 16-bit fields (even and odd bytes) so that the wave sums
        // (at most 64 x 255) cannot carry into the next field
        const unsigned long long M = 0x00ff00ff00ff00ffull;
        const unsigned long long t[4] = {
            (unsigned long long)wave_sum_ll((long long)(w0 & M)), (unsigned long long)wave_sum_ll((long long)((w0 >> 8) & M)),
            (unsigned long long)wave_sum_ll((long long)(w1 & M)), (unsigned long long)wave_sum_ll((long long)((w1 >> 8) & M))};
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int bin = 8 * (k >> 1) + 2 * j + (k & 1);
                    const unsigned c = (unsigned)((t[k] >> (16 * j)) & 0xffffu);
                    if (bin >= 1 && c) atomicAdd(&bins[bin], c);
                }
        }
        w0 = w1 = 0;
    }
};

struct Counters {
    unsigned long long n_unrecovered, n_fallback, n_guess_bad, iters_sum, wave_max_sum, n_waves;
    unsigned long long iters_max;
    unsigned long long hist[64];
    unsigned long long n_rescue;  // entries in the rescue list
};

// Per-block partial counters: row of PART_STRIDE values, columns
// 0 unrecovered, 1 fallback, 2 guess_bad, 3 iters_sum, 4 wave_max_sum,
// 5 n_waves, 6 iters_max, 8 + b histogram bin b.
constexpr int PART_STRIDE = 72;

struct Rescue {
    long i;     // cell
    double W0;  // its initial W
};

struct NetRegs;

struct SolveArgs {
    long n;
    const double *U;      // SoA: plane v at U + v * n
    const double *Pprev;  // SoA (warm guess)
    double *P;            // SoA output
    double *W0;           // per-cell initial W (record mode)
    unsigned char *iters; // per-cell iterations (record mode)
    Counters *cnt;
    unsigned long long *part;  // [gridDim.x][PART_STRIDE]
    Rescue *rescue;       // cells the 1DW iteration did not finish
    double g, rtol;
    int n_iter;
    int record;
    const NetRegs *net;   // per-lane weight fragments (learned guess), k_pack_net
};

typedef short s16x2 __attribute__((ext_vector_type(2)));

// Two floats -> two packed BF16, round to nearest even (compiler conversion:
// v_cvt_pk_bf16_f32 on gfx950, with the hazard wait states the compiler
// inserts after a matrix-core result).
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned cvt_pk_bf16(float lo, float hi)
{
    const f32x2 v = {lo, hi};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2v));
}
// ReLU of two packed BF16: negative values (sign bit set) are negative as
// int16, so a signed max with 0 clears them (v_pk_max_i16).
__device__ __forceinline__ unsigned relu_pk_bf16(unsigned x)
{
    const s16x2 z = {0, 0};
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(s16x2, x), z));
}
// ReLU + BF16 of an accumulator tile's four values, as one 64-bit word.
__device__ __forceinline__ unsigned long long relu_pack4(const f32x4 &v)
{
    // pairwise: a 4-wide vector conversion is lowered to one conversion per
    // value plus permutes
    const unsigned lo = relu_pk_bf16(cvt_pk_bf16(v[0], v[1]));
    const unsigned hi = relu_pk_bf16(cvt_pk_bf16(v[2], v[3]));
    return ((unsigned long long)hi << 32) | lo;
}

// Lane l of v_mfma_f32_16x16x16_bf16 supplies A[m = l%16][k = 4 (l/16) + j]
// and B[k = 4 (l/16) + j][n = l%16] (j = 0..3) and receives D[m = 4 (l/16) +
// j][n = l%16].  The layers are evaluated as H^T = W X^T (rows = the 16
// hidden units, columns = 16 cells), one column tile t per 16 cells of the
// wave (cells 16 t + n = lanes of group t), so:
//  * layer 1: lane l supplies its own cell's four inputs as B rows
//    4 (l/16) + j.  For tile t only lane group t carries the tile's cells, so
//    the A operand of tile t holds W1 in k-block t and zeros elsewhere; the
//    other groups' inputs meet zero weights.  The bias b1 is the FP32
//    accumulator initialisation (rows 4 (l/16) + j in lane l).
//  * layers 2 and 3: an accumulator tile (rows 4 (l/16) + j of column l%16)
//    is, after ReLU and BF16 packing, exactly the B operand rows the lane
//    supplies next: no data movement between layers.
//  * outputs: rows 0 and 1 of the last tile land in lanes 0..15 (cells
//    16 t + lane), one shuffle per tile delivers them to their owners.
// 12 MFMA per 64 cells, all operands in 12 VGPRs of weights and 12 of biases.
struct NetRegs {
    bf16x4 w1[4];  // layer 1 A operand of column tile t: W1 in k-block t only
    bf16x4 w2;     // layer 2 A operand
    bf16x4 w3;     // output layer A operand, rows 0 and 1
    f32x4 b1, b2;  // accumulator initialisation (biases)
    f32x4 b3;      // output bias (rows 0, 1 live in lane group 0)
};

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    bf16x4 r;
    r[0] = (short)bf16_bits(a);
    r[1] = (short)bf16_bits(b);
    r[2] = (short)bf16_bits(c);
    r[3] = (short)bf16_bits(d);
    return r;
}

__device__ __forceinline__ void load_net(const float *prm, NetRegs &nr, int lane)
{
    const int m = lane & 15, g = lane >> 4, kq = 4 * g;
    static_assert(NIN == 4 && NH == 16 && NOUT <= 4, "layout of the 16-wide network");
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const float *w = prm + OFF_W1 + m * NIN;
        nr.w1[t] = g == t ? pack4(w[0], w[1], w[2], w[3]) : pack4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const float *w2 = prm + OFF_W2 + m * NH + kq;
    nr.w2 = pack4(w2[0], w2[1], w2[2], w2[3]);
    const float *w3 = prm + OFF_W3 + m * NH + kq;
    nr.w3 = m < NOUT ? pack4(w3[0], w3[1], w3[2], w3[3]) : pack4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        nr.b1[j] = prm[OFF_B1 + kq + j];
        nr.b2[j] = prm[OFF_B2 + kq + j];
        nr.b3[j] = (g == 0 && j < NOUT) ? prm[OFF_B3 + j] : 0.0f;
    }
}

// The fragments are formed once per parameter set (one wave) and read by
// every wave of the inference kernels.
__global__ void __launch_bounds__(64) k_pack_net(const float *__restrict__ prm, NetRegs *__restrict__ net)
{
    NetRegs nr;
    load_net(prm, nr, threadIdx.x);
    net[threadIdx.x] = nr;
}

// Lane 16 t + n receives lane n of r_t (t = 0..3, n = 0..15).
// v_permlane32_swap(a, b) leaves a = [a.lo32, b.lo32] and b = [a.hi32, b.hi32];
// v_permlane16_swap(a, b) leaves a = [a.row0, b.row0, a.row2, b.row2] and
// b = [a.row1, b.row1, a.row3, b.row3] (rows of 16 lanes).
__device__ __forceinline__ float deliver(float r0, float r1, float r2, float r3)
{
    const auto s02 = __builtin_amdgcn_permlane32_swap(__float_as_uint(r0), __float_as_uint(r2), false, false);
    const auto s13 = __builtin_amdgcn_permlane32_swap(__float_as_uint(r1), __float_as_uint(r3), false, false);
    // s02[0] = [r0.row0, r0.row1, r2.row0, r2.row1], s13[0] likewise for r1, r3
    const auto s = __builtin_amdgcn_permlane16_swap(s02[0], s13[0], false, false);
    return __uint_as_float(s[0]);
}

// Network outputs for the 64 cells of a wave; lane c passes the features of
// cell c and receives that cell's outputs (y1 only when BOTH).  All 64 lanes
// must be active.
template <bool BOTH>
__device__ __forceinline__ void net_eval(const NetRegs &nr, const float *x, float &y0, float &y1)
{
    const bf16x4 bx = __builtin_bit_cast(
        bf16x4, ((unsigned long long)cvt_pk_bf16(x[2], x[3]) << 32) | cvt_pk_bf16(x[0], x[1]));
    f32x4 out[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const f32x4 z1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nr.w1[t], bx, nr.b1, 0, 0, 0);
        const bf16x4 h1 = __builtin_bit_cast(bf16x4, relu_pack4(z1));
        const f32x4 z2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nr.w2, h1, nr.b2, 0, 0, 0);
        const bf16x4 h2 = __builtin_bit_cast(bf16x4, relu_pack4(z2));
        out[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nr.w3, h2, nr.b3, 0, 0, 0);
    }
    // Row 0 (lanes 0..15) of out[t][0] holds y0 of cells 16 t + lane.  Two
    // 32-lane swaps put tiles 0 | 2 and 1 | 3 side by side, a 16-lane swap
    // interleaves them: lane 16 t + n then holds out[t][0] of lane n.
    y0 = deliver(out[0][0], out[1][0], out[2][0], out[3][0]);
    if constexpr (BOTH) y1 = deliver(out[0][1], out[1][1], out[2][1], out[3][1]);
}

// ---------------------------------------------------------------- solve
template <int GUESS>
__global__ void __launch_bounds__(256, 3) k_solve(SolveArgs a)
{
    __shared__ unsigned hist[64];
    __shared__ unsigned long long wtot[4][8];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 64; i += blockDim.x) hist[i] = 0;
    LaneHist lh;
    NetRegs nr;
    if constexpr (GUESS == VRN_GUESS_LEARNED) nr = a.net[lane];
    __syncthreads();
    const long n = a.n;
    const long nchunks = (n + 63) / 64;
    unsigned long long unrec = 0, fall = 0, bad = 0, isum = 0, wmax_sum = 0, nw = 0;
    int imax = 0;
    const long stride = (long)gridDim.x * 4;
    long ch = (long)blockIdx.x * 4 + wid;
    // the next chunk's inputs are loaded while this chunk computes
    double un[5], pn[5];
    {
        const long i = ch * 64 + lane;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            un[v] = (ch < nchunks && i < n) ? a.U[v * n + i] : 0.0;
            if constexpr (GUESS == VRN_GUESS_WARM) pn[v] = (ch < nchunks && i < n) ? a.Pprev[v * n + i] : 1.0;
        }
    }
    for (; ch < nchunks; ch += stride) {
        const long i = ch * 64 + lane;
        const bool act = i < n;
        double u[5];
        double Pg[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            u[v] = un[v];
            if constexpr (GUESS == VRN_GUESS_WARM) Pg[v] = pn[v];
        }
        {
            const long cn = ch + stride, in = cn * 64 + lane;
            const bool an = cn < nchunks && in < n;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                un[v] = an ? a.U[v * n + in] : 0.0;
                if constexpr (GUESS == VRN_GUESS_WARM) pn[v] = an ? a.Pprev[v * n + in] : 1.0;
            }
        }
        const double S2 = u[2] * u[2] + u[3] * u[3] + u[4] * u[4];
        double W;
        if constexpr (GUESS == VRN_GUESS_GENERIC) {
            Pg[0] = 1.0;
            Pg[1] = 1.0;
            Pg[2] = Pg[3] = Pg[4] = 0.0;
        }
        if constexpr (GUESS == VRN_GUESS_GENERIC || GUESS == VRN_GUESS_WARM) {
            // W from the guess as UtoP forms it (src/3dim/structs.jl:276-287)
            const double gm = vrn_sqrt1(Pg[2] * Pg[2] + Pg[3] * Pg[3] + Pg[4] * Pg[4] + 1.0);
            W = (Pg[0] + a.g * Pg[1]) * (gm * gm);
        } else if constexpr (GUESS == VRN_GUESS_ANALYTIC) {
            W = -u[1];
        } else {
            const Feat f = c4_features(u[0], u[1], S2);
            float y0 = 0.0f, y1 = 0.0f;
            net_eval<false>(nr, f.x, y0, y1);
            if (f.ok && __builtin_isfinite(y0)) {
                W = learned_w0(f, y0);
            } else {
                W = f.E;
                if (act) bad++;
            }
        }
        const double W0 = W;
        int it = 0;
        if (act) {
            const bool conv = newton_1dw_rel(u[0], u[1], S2, a.g, a.n_iter, a.rtol, W, it);
            if (conv) {
                double P[5];
                recover(u[0], u[2], u[3], u[4], S2, W, a.g, P);
                bool fin = true;
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    fin = fin && __builtin_isfinite(P[v]);
                    a.P[v * n + i] = P[v];
                }
                if (!fin) unrec++;
            } else {
                fall++;
                const unsigned long long r = atomicAdd(&a.cnt->n_rescue, 1ull);
                a.rescue[r].i = i;
                a.rescue[r].W0 = W0;
            }
            if (a.record) {
                a.W0[i] = W0;
                a.iters[i] = (unsigned char)(it > 255 ? 255 : it);
            }
            isum += it;
            imax = it > imax ? it : imax;
            if (it >= 16)
                atomicAdd(&hist[it < 63 ? it : 63], 1u);
            else if (it >= 1)
                lh.add(it);
        }
        const int mine = act ? it : 0;
        const int wm = wave_max_count(mine);
        wmax_sum += wm;
        nw++;
        if ((nw & 127) == 0) lh.flush(hist, lane);  // nw is wave-uniform; fields stay below 256
    }
    lh.flush(hist, lane);
    // per-wave totals -> block row (no global atomics on shared addresses)
    const long long t_unrec = wave_sum_ll((long long)unrec), t_fall = wave_sum_ll((long long)fall),
                    t_bad = wave_sum_ll((long long)bad), t_isum = wave_sum_ll((long long)isum);
    const int t_imax = wave_max_i(imax);
    if (lane == 0) {
        wtot[wid][0] = (unsigned long long)t_unrec;
        wtot[wid][1] = (unsigned long long)t_fall;
        wtot[wid][2] = (unsigned long long)t_bad;
        wtot[wid][3] = (unsigned long long)t_isum;
        wtot[wid][4] = wmax_sum;
        wtot[wid][5] = nw;
        wtot[wid][6] = (unsigned long long)t_imax;
        wtot[wid][7] = 0;
    }
    __syncthreads();
    unsigned long long *row = a.part + (long)blockIdx.x * PART_STRIDE;
    if (threadIdx.x < 8) {
        const int c = threadIdx.x;
        const int nwv = (int)(blockDim.x >> 6);
        unsigned long long v = wtot[0][c];
        for (int w = 1; w < nwv; w++) v = c == 6 ? (wtot[w][c] > v ? wtot[w][c] : v) : v + wtot[w][c];
        row[c] = v;
    } else if (threadIdx.x >= 64 && threadIdx.x < 128) {
        row[8 + threadIdx.x - 64] = hist[threadIdx.x - 64];
    }
}

// Sum of the per-block rows (column max for iters_max): one block per column.
__global__ void __launch_bounds__(256) k_reduce_counters(const unsigned long long *__restrict__ part, int rows,
                                                         Counters *cnt)
{
    __shared__ unsigned long long red[256];
    const int c = blockIdx.x;
    if (c == 7) return;
    const bool is_max = c == 6;
    unsigned long long v = 0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        const unsigned long long x = part[(long)r * PART_STRIDE + c];
        v = is_max ? (x > v ? x : v) : v + x;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const unsigned long long y = red[threadIdx.x + o];
            red[threadIdx.x] = is_max ? (y > red[threadIdx.x] ? y : red[threadIdx.x]) : red[threadIdx.x] + y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned long long t = red[0];
        switch (c) {
        case 0: cnt->n_unrecovered = t; break;
        case 1: cnt->n_fallback = t; break;
        case 2: cnt->n_guess_bad = t; break;
        case 3: cnt->iters_sum = t; break;
        case 4: cnt->wave_max_sum = t; break;
        case 5: cnt->n_waves = t; break;
        case 6: cnt->iters_max = t; break;
        default: cnt->hist[c - 8] = t; break;
        }
    }
}

// Cells the 1DW iteration left unfinished: the rest of the reference cascade
// (2D Newton, 5D Newton, bisection) from the primitive guess of the mode.
template <int GUESS>
__global__ void __launch_bounds__(256) k_rescue(SolveArgs a)
{
    const long nr = (long)a.cnt->n_rescue;
    const long n = a.n;
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < nr; r += (long)gridDim.x * blockDim.x) {
        const long i = a.rescue[r].i;
        double u[5], P[5];
#pragma unroll
        for (int v = 0; v < 5; v++) u[v] = a.U[v * n + i];
        if constexpr (GUESS == VRN_GUESS_GENERIC) {
            P[0] = 1.0;
            P[1] = 1.0;
            P[2] = P[3] = P[4] = 0.0;
        } else if constexpr (GUESS == VRN_GUESS_WARM) {
#pragma unroll
            for (int v = 0; v < 5; v++) P[v] = a.Pprev[v * n + i];
        } else {
            // a state at rest with the guessed W
            P[0] = u[0];
            P[1] = jmax(1e-8, (a.rescue[r].W0 - u[0]) / a.g);
            P[2] = P[3] = P[4] = 0.0;
        }
        int it = 0;
        c2p_cascade_tail(u, P, a.g, a.n_iter, a.rtol * (-u[1]), it);
        bool fin = true;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            fin = fin && __builtin_isfinite(P[v]);
            a.P[v * n + i] = P[v];
        }
        if (!fin) atomicAdd(&a.cnt->n_unrecovered, 1ull);
        atomicAdd(&a.cnt->iters_sum, (unsigned long long)it);
    }
}

// ---------------------------------------------------------------- data movement
__global__ void k_aos_to_soa(long n, const double *__restrict__ src, double *__restrict__ dst)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < 5 * n; t += (long)gridDim.x * blockDim.x) {
        const long i = t / 5, v = t % 5;
        dst[v * n + i] = src[t];
    }
}
__global__ void k_soa_to_aos(long n, const double *__restrict__ src, double *__restrict__ dst)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < 5 * n; t += (long)gridDim.x * blockDim.x) {
        const long i = t / 5, v = t % 5;
        dst[t] = src[v * n + i];
    }
}
__global__ void k_zero_counters(Counters *c)
{
    unsigned long long *p = reinterpret_cast<unsigned long long *>(c);
    for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 8); i += blockDim.x) p[i] = 0;
}

// Features (and targets) of AoS (U, P) pairs.
__global__ void k_features(long n, const double *__restrict__ U, const double *__restrict__ P, double g,
                           float *__restrict__ X, float *__restrict__ T, unsigned char *__restrict__ valid)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double *u = U + 5 * i;
        const double S2 = u[2] * u[2] + u[3] * u[3] + u[4] * u[4];
        const Feat f = c4_features(u[0], u[1], S2);
#pragma unroll
        for (int k = 0; k < NIN; k++) X[NIN * i + k] = f.x[k];
        bool ok = f.ok;
        if (P) {
            float t[2];
            ok = c4_targets(P + 5 * i, g, f, t);
            T[2 * i] = t[0];
            T[2 * i + 1] = t[1];
        }
        valid[i] = ok ? 1 : 0;
    }
}

// Inference of the network alone on given features (64 samples per wave).
__global__ void __launch_bounds__(256) k_infer(long n, const float *__restrict__ X, const NetRegs *net,
                                               float *__restrict__ Y)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const NetRegs nr = net[lane];
    const long nchunks = (n + 63) / 64;
    for (long ch = (long)blockIdx.x * 4 + wid; ch < nchunks; ch += (long)gridDim.x * 4) {
        const long i = ch * 64 + lane;
        float x[NIN];
#pragma unroll
        for (int k = 0; k < NIN; k++) x[k] = i < n ? X[NIN * i + k] : 0.0f;
        float y0 = 0.0f, y1 = 0.0f;
        net_eval<true>(nr, x, y0, y1);
        if (i < n) {
            Y[2 * i] = y0;
            Y[2 * i + 1] = y1;
        }
    }
}

// ---------------------------------------------------------------- training
// Per-sample cache layout (floats): x[8] h1[64] h2[64] dz1[64] dz2[64] dy[2] loss[1]
constexpr int C_X = 0, C_H1 = 8, C_H2 = C_H1 + NH, C_DZ1 = C_H2 + NH, C_DZ2 = C_DZ1 + NH,
              C_DY = C_DZ2 + NH, C_L = C_DY + NOUT, C_STRIDE = 272;
static_assert(C_L < C_STRIDE, "cache stride");

__device__ __forceinline__ unsigned long long splitmix(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct TrainArgs {
    const float *X, *T;
    const int *idx;      // valid-sample list (sampling) or NULL (identity)
    long n_idx;
    int B;
    unsigned long long seed;
    int step;
    const float *prm;
    float *cache;
    float w1;
    float q_ref;  // > 0: the first output's squared error weighted by 1 + (q / q_ref)^2
    int qat;
};

__global__ void __launch_bounds__(256) k_train_fwd_bwd(TrainArgs a)
{
    __shared__ float w[NPARAM];
    for (int p = threadIdx.x; p < NPARAM; p += blockDim.x) {
        float v = a.prm[p];
        // BF16 in inference: W1, W2, W3 (the biases are FP32 accumulator inputs)
        const bool is_w = p < OFF_B1 || (p >= OFF_W2 && p < OFF_B2) || (p >= OFF_W3 && p < OFF_B3);
        w[p] = (a.qat && is_w) ? bf16_round(v) : v;
    }
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.B) return;
    long id = s;
    if (a.idx) {
        unsigned long long r = splitmix(a.seed ^ ((unsigned long long)a.step * 0x100000001B3ull + (unsigned long long)s));
        id = a.idx[(long)(r % (unsigned long long)a.n_idx)];
    }
    float *c = a.cache + (long)s * C_STRIDE;
    float x[NIN];
#pragma unroll
    for (int k = 0; k < NIN; k++) {
        x[k] = a.X[(long)NIN * id + k];
        if (a.qat) x[k] = bf16_round(x[k]);
        c[C_X + k] = x[k];
    }
    float h1[NH];
#pragma unroll
    for (int j = 0; j < NH; j++) {
        float z = w[OFF_B1 + j];
#pragma unroll
        for (int k = 0; k < NIN; k++) z = z + w[OFF_W1 + j * NIN + k] * x[k];
        float h = fmaxf(z, 0.0f);
        if (a.qat) h = bf16_round(h);
        h1[j] = h;
        c[C_H1 + j] = h;
    }
    float y0 = w[OFF_B3], y1 = w[OFF_B3 + 1];
    for (int j = 0; j < NH; j++) {
        float z = w[OFF_B2 + j];
#pragma unroll
        for (int k = 0; k < NH; k++) z = z + w[OFF_W2 + j * NH + k] * h1[k];
        float h = fmaxf(z, 0.0f);
        if (a.qat) h = bf16_round(h);
        c[C_H2 + j] = h;
        y0 = y0 + w[OFF_W3 + j] * h;
        y1 = y1 + w[OFF_W3 + NH + j] * h;
    }
    const float t0 = a.T[2 * id];
    const float e0 = y0 - t0, e1 = y1 - a.T[2 * id + 1];
    const float invB = 1.0f / (float)a.B;
    // q = z e^{t0} = log(W/E): the guess W0 = E e^q misses W by q times the
    // error of y0, so the hot states (large q) are weighted up
    float wq = 1.0f;
    if (a.q_ref > 0.0f) {
        const float sz = a.X[(long)NIN * id + 2];
        const float r = sz * sz * expf(t0) / a.q_ref;
        wq = 1.0f + r * r;
    }
    const float dy0 = 2.0f * wq * e0 * invB, dy1 = 2.0f * a.w1 * e1 * invB;
    c[C_DY] = dy0;
    c[C_DY + 1] = dy1;
    c[C_L] = (wq * e0 * e0 + a.w1 * e1 * e1) * invB;
    // dz2 = (W3^T dy) * [z2 > 0]   (h2 > 0 <=> z2 > 0)
    float dh1[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) dh1[k] = 0.0f;
    for (int j = 0; j < NH; j++) {
        const float h = c[C_H2 + j];
        const float d = h > 0.0f ? (w[OFF_W3 + j] * dy0 + w[OFF_W3 + NH + j] * dy1) : 0.0f;
        c[C_DZ2 + j] = d;
#pragma unroll
        for (int k = 0; k < NH; k++) dh1[k] = dh1[k] + w[OFF_W2 + j * NH + k] * d;
    }
#pragma unroll
    for (int k = 0; k < NH; k++) c[C_DZ1 + k] = h1[k] > 0.0f ? dh1[k] : 0.0f;
}

// Partial gradient sums: thread = parameter (+1 for the loss), blockIdx.y =
// slice of the batch.  part[slice][NPARAM + 1].
__global__ void __launch_bounds__(256) k_train_grad_partial(const float *__restrict__ cache, int B, int nsplit,
                                                            float *__restrict__ part)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    const int per = (B + nsplit - 1) / nsplit;
    const int s0 = blockIdx.y * per, s1 = min(B, s0 + per);
    int oa, ob;  // cache offsets of the two factors (ob < 0: factor 1)
    if (p < OFF_B1) { oa = C_DZ1 + p / NIN; ob = C_X + p % NIN; }
    else if (p < OFF_W2) { oa = C_DZ1 + (p - OFF_B1); ob = -1; }
    else if (p < OFF_B2) { oa = C_DZ2 + (p - OFF_W2) / NH; ob = C_H1 + (p - OFF_W2) % NH; }
    else if (p < OFF_W3) { oa = C_DZ2 + (p - OFF_B2); ob = -1; }
    else if (p < OFF_B3) { oa = C_DY + (p - OFF_W3) / NH; ob = C_H2 + (p - OFF_W3) % NH; }
    else if (p < NPARAM) { oa = C_DY + (p - OFF_B3); ob = -1; }
    else { oa = C_L; ob = -1; }
    float acc = 0.0f;
    if (ob >= 0) {
        for (int s = s0; s < s1; s++) acc = acc + cache[(long)s * C_STRIDE + oa] * cache[(long)s * C_STRIDE + ob];
    } else {
        for (int s = s0; s < s1; s++) acc = acc + cache[(long)s * C_STRIDE + oa];
    }
    part[(long)blockIdx.y * (NPARAM + 1) + p] = acc;
}

__global__ void k_train_grad_final(const float *__restrict__ part, int nsplit, float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    float acc = 0.0f;
    for (int s = 0; s < nsplit; s++) acc = acc + part[(long)s * (NPARAM + 1) + p];
    grad[p] = acc;
}

__global__ void k_adam(float *__restrict__ prm, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ grad, float lr, float b1, float b2, float eps, float bc1,
                       float bc2, float *loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == NPARAM && loss_out) *loss_out = grad[NPARAM];
    if (p >= NPARAM) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    prm[p] = prm[p] - lr * (mm / bc1) / (__builtin_sqrtf(vv / bc2) + eps);
}

}  // namespace vrn_c4

// ================================================================ host
using namespace vrn_c4;

namespace {
thread_local std::string t_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(-2, "%s: %s", #x, hipGetErrorString(e_));            \
    } while (0)

int grid_for(long work, int block, int cap = 8192)
{
    long g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

constexpr int MAX_BATCH = 65536;
constexpr int MAX_SOLVE_GRID = 4096;
constexpr int NSPLIT = 64;
}  // namespace

struct vrn_c2p {
    long n_max = 0, n = 0;
    double g = 0;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double *U = nullptr, *Pprev = nullptr, *P = nullptr, *W0 = nullptr, *stage = nullptr;
    unsigned char *iters = nullptr;
    Counters *cnt = nullptr;
    Rescue *rescue = nullptr;
    unsigned long long *crow = nullptr;  // per-block counter rows
    float *params = nullptr;  // device FP32 parameters
    NetRegs *net = nullptr;   // their per-lane MFMA fragments (64 lanes)
    bool net_ok = false;      // net matches params
    bool have_params = false, have_prev = false, have_P = false, recorded = false;
    // training / test buffers (allocated on first use)
    float *cache = nullptr, *part = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *tprm = nullptr;
    float *X = nullptr, *T = nullptr;
    int *idx = nullptr;
    unsigned char *valid = nullptr;
    double *pairU = nullptr, *pairP = nullptr;
    long train_cap = 0;
};

namespace {

// Weight fragments of the current parameters, formed on the stream before
// the first inference that needs them.
int ensure_net(vrn_c2p *h)
{
    if (h->net_ok) return 0;
    hipLaunchKernelGGL(k_pack_net, dim3(1), dim3(64), 0, h->s, (const float *)h->params, h->net);
    HIPCHK(hipGetLastError());
    h->net_ok = true;
    return 0;
}

int ensure_train(vrn_c2p *h)
{
    if (h->cache) return 0;
    HIPCHK(hipMalloc(&h->cache, sizeof(float) * (size_t)C_STRIDE * MAX_BATCH));
    HIPCHK(hipMalloc(&h->part, sizeof(float) * (size_t)NSPLIT * (NPARAM + 1)));
    HIPCHK(hipMalloc(&h->grad, sizeof(float) * (NPARAM + 1)));
    HIPCHK(hipMalloc(&h->m, sizeof(float) * NPARAM));
    HIPCHK(hipMalloc(&h->v, sizeof(float) * NPARAM));
    HIPCHK(hipMalloc(&h->tprm, sizeof(float) * NPARAM));
    return 0;
}

int ensure_pairs(vrn_c2p *h, long n)
{
    if (n <= h->train_cap) return 0;
    (void)hipFree(h->X);
    (void)hipFree(h->T);
    (void)hipFree(h->idx);
    (void)hipFree(h->valid);
    (void)hipFree(h->pairU);
    (void)hipFree(h->pairP);
    h->X = h->T = nullptr;
    h->idx = nullptr;
    h->valid = nullptr;
    h->pairU = h->pairP = nullptr;
    h->train_cap = 0;
    HIPCHK(hipMalloc(&h->X, sizeof(float) * NIN * (size_t)n));
    HIPCHK(hipMalloc(&h->T, sizeof(float) * 2 * (size_t)n));
    HIPCHK(hipMalloc(&h->idx, sizeof(int) * (size_t)n));
    HIPCHK(hipMalloc(&h->valid, (size_t)n));
    HIPCHK(hipMalloc(&h->pairU, sizeof(double) * 5 * (size_t)n));
    HIPCHK(hipMalloc(&h->pairP, sizeof(double) * 5 * (size_t)n));
    h->train_cap = n;
    return 0;
}

// One forward/backward + gradient reduction over B samples (cache -> grad).
int grad_pass(vrn_c2p *h, const float *X, const float *T, const int *idx, long n_idx, int B,
              unsigned long long seed, int step, const float *prm, float w1, float q_ref, int qat)
{
    TrainArgs ta;
    ta.X = X;
    ta.T = T;
    ta.idx = idx;
    ta.n_idx = n_idx;
    ta.B = B;
    ta.seed = seed;
    ta.step = step;
    ta.prm = prm;
    ta.cache = h->cache;
    ta.w1 = w1;
    ta.q_ref = q_ref;
    ta.qat = qat;
    hipLaunchKernelGGL(k_train_fwd_bwd, dim3((B + 255) / 256), dim3(256), 0, h->s, ta);
    const int nsplit = B < NSPLIT ? B : NSPLIT;
    hipLaunchKernelGGL(k_train_grad_partial, dim3((NPARAM + 1 + 255) / 256, nsplit), dim3(256), 0, h->s,
                       (const float *)h->cache, B, nsplit, h->part);
    hipLaunchKernelGGL(k_train_grad_final, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, h->s,
                       (const float *)h->part, nsplit, h->grad);
    HIPCHK(hipGetLastError());
    return 0;
}

// Seeded He initialisation on the host (SplitMix64 + Box-Muller).
void he_init(std::vector<float> &w, unsigned long long seed)
{
    unsigned long long st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() {
        double u1 = next(), u2 = next();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    w.assign(NPARAM, 0.0f);
    for (int i = 0; i < NH * NIN; i++) w[OFF_W1 + i] = (float)(normal() * std::sqrt(2.0 / NIN));
    for (int i = 0; i < NH * NH; i++) w[OFF_W2 + i] = (float)(normal() * std::sqrt(2.0 / NH));
    for (int i = 0; i < NOUT * NH; i++) w[OFF_W3 + i] = (float)(normal() * std::sqrt(1.0 / NH));
}

}  // namespace

extern "C" {

const char *vrn_c2p_last_error(void) { return t_err.c_str(); }

int vrn_c2p_create(int64_t n_max, double gamma, vrn_c2p **out)
{
    if (!out) return fail(-1, "vrn_c2p_create: out is NULL");
    *out = nullptr;
    if (n_max < 1 || n_max > (int64_t)1 << 31) return fail(-1, "vrn_c2p_create: n_max %lld out of range", (long long)n_max);
    if (!(gamma > 1.0 && gamma <= 2.0)) return fail(-1, "vrn_c2p_create: gamma %g outside (1, 2]", gamma);
    vrn_c2p *h = new vrn_c2p();
    h->n_max = n_max;
    h->g = gamma;
    auto cleanup = [&](int rc) {
        vrn_c2p_destroy(h);
        return rc;
    };
    const size_t nb = sizeof(double) * 5 * (size_t)n_max;
    if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess ||
        hipMalloc(&h->U, nb) != hipSuccess || hipMalloc(&h->P, nb) != hipSuccess ||
        hipMalloc(&h->stage, nb) != hipSuccess || hipMalloc(&h->W0, sizeof(double) * n_max) != hipSuccess ||
        hipMalloc(&h->iters, n_max) != hipSuccess || hipMalloc(&h->cnt, sizeof(Counters)) != hipSuccess ||
        hipMalloc(&h->rescue, sizeof(Rescue) * n_max) != hipSuccess ||
        hipMalloc(&h->crow, sizeof(unsigned long long) * PART_STRIDE * MAX_SOLVE_GRID) != hipSuccess ||
        hipMalloc(&h->params, sizeof(float) * NPARAM) != hipSuccess ||
        hipMalloc(&h->net, sizeof(NetRegs) * 64) != hipSuccess)
        return cleanup(fail(-2, "vrn_c2p_create: device allocation of %zu bytes failed", 4 * nb));
    *out = h;
    return 0;
}

int vrn_c2p_destroy(vrn_c2p *h)
{
    if (!h) return 0;
    if (h->s) (void)hipStreamSynchronize(h->s);
    (void)hipFree(h->U);
    (void)hipFree(h->Pprev);
    (void)hipFree(h->P);
    (void)hipFree(h->W0);
    (void)hipFree(h->stage);
    (void)hipFree(h->iters);
    (void)hipFree(h->cnt);
    (void)hipFree(h->rescue);
    (void)hipFree(h->crow);
    (void)hipFree(h->params);
    (void)hipFree(h->net);
    (void)hipFree(h->cache);
    (void)hipFree(h->part);
    (void)hipFree(h->grad);
    (void)hipFree(h->m);
    (void)hipFree(h->v);
    (void)hipFree(h->tprm);
    (void)hipFree(h->X);
    (void)hipFree(h->T);
    (void)hipFree(h->idx);
    (void)hipFree(h->valid);
    (void)hipFree(h->pairU);
    (void)hipFree(h->pairP);
    if (h->e0) (void)hipEventDestroy(h->e0);
    if (h->e1) (void)hipEventDestroy(h->e1);
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h;
    return 0;
}

int vrn_c2p_upload(vrn_c2p *h, int64_t n, const double *U, const double *P_prev)
{
    if (!h || !U) return fail(-1, "vrn_c2p_upload: NULL argument");
    if (n < 1 || n > h->n_max) return fail(-1, "vrn_c2p_upload: n = %lld outside [1, %ld]", (long long)n, h->n_max);
    const size_t nb = sizeof(double) * 5 * (size_t)n;
    HIPCHK(hipMemcpyAsync(h->stage, U, nb, hipMemcpyHostToDevice, h->s));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(5 * n, 256)), dim3(256), 0, h->s, (long)n,
                       (const double *)h->stage, h->U);
    HIPCHK(hipGetLastError());
    h->have_prev = false;
    if (P_prev) {
        if (!h->Pprev) HIPCHK(hipMalloc(&h->Pprev, sizeof(double) * 5 * (size_t)h->n_max));
        HIPCHK(hipStreamSynchronize(h->s));
        HIPCHK(hipMemcpyAsync(h->stage, P_prev, nb, hipMemcpyHostToDevice, h->s));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(5 * n, 256)), dim3(256), 0, h->s, (long)n,
                           (const double *)h->stage, h->Pprev);
        HIPCHK(hipGetLastError());
        h->have_prev = true;
    }
    HIPCHK(hipStreamSynchronize(h->s));
    h->n = n;
    h->have_P = false;
    return 0;
}

int vrn_c2p_solve(vrn_c2p *h, int guess, int n_iter, double rtol, vrn_c2p_stats *stats)
{
    if (!h) return fail(-1, "vrn_c2p_solve: NULL handle");
    if (h->n < 1) return fail(-1, "vrn_c2p_solve: nothing uploaded");
    const int record = (guess & RECORD_FLAG) ? 1 : 0;
    guess &= ~RECORD_FLAG;
    if (guess < VRN_GUESS_GENERIC || guess > VRN_GUESS_LEARNED) return fail(-1, "vrn_c2p_solve: unknown guess %d", guess);
    if (guess == VRN_GUESS_WARM && !h->have_prev) return fail(-1, "vrn_c2p_solve: warm guess needs P_prev in vrn_c2p_upload");
    if (guess == VRN_GUESS_LEARNED && !h->have_params) return fail(-1, "vrn_c2p_solve: learned guess needs trained parameters (vrn_mlp_set / vrn_mlp_train)");
    if (n_iter < 1 || n_iter > 10000) return fail(-1, "vrn_c2p_solve: n_iter %d out of range", n_iter);
    if (!(rtol > 0.0 && rtol < 1.0)) return fail(-1, "vrn_c2p_solve: rtol %g out of range", rtol);
    if (guess == VRN_GUESS_LEARNED)
        if (int rc = ensure_net(h)) return rc;
    SolveArgs a;
    a.n = h->n;
    a.U = h->U;
    a.Pprev = h->Pprev;
    a.P = h->P;
    a.W0 = h->W0;
    a.iters = h->iters;
    a.cnt = h->cnt;
    a.rescue = h->rescue;
    a.part = h->crow;
    a.g = h->g;
    a.rtol = rtol;
    a.n_iter = n_iter;
    a.record = record;
    a.net = h->net;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(128), 0, h->s, h->cnt);
    const long nchunks = (h->n + 63) / 64;
    const int grid = grid_for(nchunks, 4, MAX_SOLVE_GRID);
    HIPCHK(hipEventRecord(h->e0, h->s));
    const int rgrid = 1024;
    switch (guess) {
    case VRN_GUESS_GENERIC:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_GENERIC>, dim3(grid), dim3(256), 0, h->s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, h->s,
                           (const unsigned long long *)h->crow, grid, h->cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_GENERIC>, dim3(rgrid), dim3(256), 0, h->s, a);
        break;
    case VRN_GUESS_WARM:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_WARM>, dim3(grid), dim3(256), 0, h->s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, h->s,
                           (const unsigned long long *)h->crow, grid, h->cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_WARM>, dim3(rgrid), dim3(256), 0, h->s, a);
        break;
    case VRN_GUESS_ANALYTIC:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_ANALYTIC>, dim3(grid), dim3(256), 0, h->s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, h->s,
                           (const unsigned long long *)h->crow, grid, h->cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_ANALYTIC>, dim3(rgrid), dim3(256), 0, h->s, a);
        break;
    default:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_LEARNED>, dim3(grid), dim3(256), 0, h->s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, h->s,
                           (const unsigned long long *)h->crow, grid, h->cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_LEARNED>, dim3(rgrid), dim3(256), 0, h->s, a);
        break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(h->e1, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    h->have_P = true;
    h->recorded = record;
    if (stats) {
        Counters c;
        HIPCHK(hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
        stats->n = h->n;
        stats->n_unrecovered = (int64_t)c.n_unrecovered;
        stats->n_fallback = (int64_t)c.n_fallback;
        stats->n_guess_bad = (int64_t)c.n_guess_bad;
        stats->iters_sum = (int64_t)c.iters_sum;
        stats->iters_max = (int64_t)c.iters_max;
        stats->wave_max_sum = (int64_t)c.wave_max_sum;
        stats->n_waves = (int64_t)c.n_waves;
        for (int b = 0; b < 64; b++) stats->hist[b] = (int64_t)c.hist[b];
        stats->ms = ms;
    }
    return 0;
}

int vrn_c2p_download(vrn_c2p *h, double *P, double *W0, uint8_t *iters)
{
    if (!h) return fail(-1, "vrn_c2p_download: NULL handle");
    if (!h->have_P) return fail(-1, "vrn_c2p_download: no solve since the last upload");
    if ((W0 || iters) && !h->recorded)
        return fail(-1, "vrn_c2p_download: W0/iters need a solve with the record flag (guess | 0x100)");
    const long n = h->n;
    if (P) {
        hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(5 * n, 256)), dim3(256), 0, h->s, n, (const double *)h->P,
                           h->stage);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(P, h->stage, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost, h->s));
    }
    if (W0) HIPCHK(hipMemcpyAsync(W0, h->W0, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, h->s));
    if (iters) HIPCHK(hipMemcpyAsync(iters, h->iters, (size_t)n, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return 0;
}

int vrn_mlp_set(vrn_c2p *h, const float *params)
{
    if (!h || !params) return fail(-1, "vrn_mlp_set: NULL argument");
    for (int i = 0; i < NPARAM; i++)
        if (!std::isfinite(params[i])) return fail(-1, "vrn_mlp_set: parameter %d is not finite", i);
    HIPCHK(hipMemcpyAsync(h->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    h->have_params = true;
    h->net_ok = false;
    return 0;
}

int vrn_mlp_get(vrn_c2p *h, float *params)
{
    if (!h || !params) return fail(-1, "vrn_mlp_get: NULL argument");
    if (!h->have_params) return fail(-1, "vrn_mlp_get: no parameters set or trained");
    HIPCHK(hipMemcpyAsync(params, h->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return 0;
}

int vrn_mlp_features(vrn_c2p *h, int64_t n, const double *U, const double *P, float *X, float *T,
                     uint8_t *valid)
{
    if (!h || !U || !X || !valid || (P && !T)) return fail(-1, "vrn_mlp_features: NULL argument");
    if (n < 1 || n > h->n_max) return fail(-1, "vrn_mlp_features: n = %lld out of range", (long long)n);
    if (int rc = ensure_pairs(h, n)) return rc;
    HIPCHK(hipMemcpyAsync(h->pairU, U, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice, h->s));
    if (P) HIPCHK(hipMemcpyAsync(h->pairP, P, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice, h->s));
    hipLaunchKernelGGL(k_features, dim3(grid_for(n, 256)), dim3(256), 0, h->s, (long)n, (const double *)h->pairU,
                       P ? (const double *)h->pairP : (const double *)nullptr, h->g, h->X, h->T, h->valid);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(X, h->X, sizeof(float) * NIN * (size_t)n, hipMemcpyDeviceToHost, h->s));
    if (P) HIPCHK(hipMemcpyAsync(T, h->T, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipMemcpyAsync(valid, h->valid, (size_t)n, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return 0;
}

int vrn_mlp_infer(vrn_c2p *h, int64_t n, const float *X, float *Y)
{
    if (!h || !X || !Y) return fail(-1, "vrn_mlp_infer: NULL argument");
    if (n < 1 || n > h->n_max) return fail(-1, "vrn_mlp_infer: n = %lld out of range", (long long)n);
    if (!h->have_params) return fail(-1, "vrn_mlp_infer: no parameters set or trained");
    if (int rc = ensure_pairs(h, n)) return rc;
    if (int rc = ensure_net(h)) return rc;
    HIPCHK(hipMemcpyAsync(h->X, X, sizeof(float) * NIN * (size_t)n, hipMemcpyHostToDevice, h->s));
    const long nchunks = (n + 63) / 64;
    hipLaunchKernelGGL(k_infer, dim3(grid_for(nchunks, 4, 4096)), dim3(256), 0, h->s, (long)n, (const float *)h->X,
                       (const NetRegs *)h->net, h->T);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(Y, h->T, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return 0;
}

int vrn_mlp_grad(vrn_c2p *h, int64_t n, const float *X, const float *T, const float *params, int qat,
                 float q_ref, float *loss, float *grad)
{
    if (!h || !X || !T || !params || !grad) return fail(-1, "vrn_mlp_grad: NULL argument");
    if (n < 1 || n > MAX_BATCH) return fail(-1, "vrn_mlp_grad: n = %lld outside [1, %d]", (long long)n, MAX_BATCH);
    if (int rc = ensure_train(h)) return rc;
    if (int rc = ensure_pairs(h, n)) return rc;
    HIPCHK(hipMemcpyAsync(h->X, X, sizeof(float) * NIN * (size_t)n, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->T, T, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->tprm, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, h->s));
    if (int rc = grad_pass(h, h->X, h->T, nullptr, n, (int)n, 0, 0, h->tprm, 1.0f, q_ref, qat ? 1 : 0)) return rc;
    std::vector<float> g(NPARAM + 1);
    HIPCHK(hipMemcpyAsync(g.data(), h->grad, sizeof(float) * (NPARAM + 1), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    std::memcpy(grad, g.data(), sizeof(float) * NPARAM);
    if (loss) *loss = g[NPARAM];
    return 0;
}

int vrn_mlp_train(vrn_c2p *h, int64_t n, const double *U, const double *P, const float *init,
                  const vrn_mlp_train_params *tp, float *loss_curve, double *ms)
{
    if (!h || !U || !P || !tp) return fail(-1, "vrn_mlp_train: NULL argument");
    if (n < 1 || n > (int64_t)1 << 30) return fail(-1, "vrn_mlp_train: n = %lld out of range", (long long)n);
    if (tp->steps < 1 || tp->batch < 1 || tp->batch > MAX_BATCH)
        return fail(-1, "vrn_mlp_train: steps %d / batch %d out of range", tp->steps, tp->batch);
    if (int rc = ensure_train(h)) return rc;
    if (int rc = ensure_pairs(h, n)) return rc;
    HIPCHK(hipEventRecord(h->e0, h->s));
    HIPCHK(hipMemcpyAsync(h->pairU, U, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->pairP, P, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice, h->s));
    hipLaunchKernelGGL(k_features, dim3(grid_for(n, 256)), dim3(256), 0, h->s, (long)n, (const double *)h->pairU,
                       (const double *)h->pairP, h->g, h->X, h->T, h->valid);
    HIPCHK(hipGetLastError());
    // list of the valid pairs (host side: a one-off O(n) pass)
    std::vector<unsigned char> ok((size_t)n);
    HIPCHK(hipMemcpyAsync(ok.data(), h->valid, (size_t)n, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    std::vector<int> ids;
    ids.reserve((size_t)n);
    for (long i = 0; i < n; i++)
        if (ok[i]) ids.push_back((int)i);
    if (ids.empty()) return fail(-1, "vrn_mlp_train: no pair lies in the model's domain");
    HIPCHK(hipMemcpyAsync(h->idx, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, h->s));
    std::vector<float> w0;
    if (init) {
        w0.assign(init, init + NPARAM);
    } else {
        he_init(w0, tp->seed);
    }
    h->net_ok = false;
    HIPCHK(hipMemcpyAsync(h->params, w0.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemsetAsync(h->m, 0, sizeof(float) * NPARAM, h->s));
    HIPCHK(hipMemsetAsync(h->v, 0, sizeof(float) * NPARAM, h->s));
    float *curve = nullptr;
    HIPCHK(hipMalloc(&curve, sizeof(float) * (size_t)tp->steps));
    double bc1 = 1.0, bc2 = 1.0;
    int rc = 0;
    for (int s = 0; s < tp->steps && rc == 0; s++) {
        const double lr = tp->lr_min + 0.5 * (tp->lr - tp->lr_min) * (1.0 + std::cos(3.141592653589793 * s / tp->steps));
        bc1 *= tp->beta1;
        bc2 *= tp->beta2;
        rc = grad_pass(h, h->X, h->T, h->idx, (long)ids.size(), tp->batch, tp->seed, s, h->params, tp->w_out1,
                       tp->q_ref, tp->qat);
        if (rc) break;
        hipLaunchKernelGGL(k_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, h->s, h->params, h->m, h->v,
                           (const float *)h->grad, (float)lr, tp->beta1, tp->beta2, tp->eps, (float)(1.0 - bc1),
                           (float)(1.0 - bc2), curve + s);
        if (hipGetLastError() != hipSuccess) rc = fail(-2, "vrn_mlp_train: launch failed at step %d", s);
    }
    if (rc == 0 && loss_curve)
        rc = hipMemcpyAsync(loss_curve, curve, sizeof(float) * (size_t)tp->steps, hipMemcpyDeviceToHost, h->s) ==
                     hipSuccess
                 ? 0
                 : fail(-2, "vrn_mlp_train: loss copy failed");
    if (hipEventRecord(h->e1, h->s) != hipSuccess || hipStreamSynchronize(h->s) != hipSuccess)
        rc = rc ? rc : fail(-2, "vrn_mlp_train: stream error");
    (void)hipFree(curve);
    if (rc) return rc;
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, h->e0, h->e1));
    if (ms) *ms = t;
    std::vector<float> w(NPARAM);
    HIPCHK(hipMemcpy(w.data(), h->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    for (int i = 0; i < NPARAM; i++)
        if (!std::isfinite(w[i])) {
            h->have_params = false;
            return fail(-3, "vrn_mlp_train: training diverged (parameter %d not finite)", i);
        }
    h->have_params = true;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- no-allocation solve
// vrn_c2p_solve_noalloc: the solve of vrn_c2p_solve on caller-owned device
// arrays, with its working memory (counters, per-block counter rows, rescue
// list) in the file-scope array below instead of a handle's allocations.
// Every GPU the library runs on holds one copy of it until the process ends;
// the host takes its device address once per device with
// hipGetSymbolAddress, since the array's name in host code is only a
// host-side stand-in.
namespace {
constexpr size_t NA_CNT_BYTES = 1024;  // Counters (576 B), padded
constexpr size_t NA_PART_BYTES = sizeof(unsigned long long) * PART_STRIDE * MAX_SOLVE_GRID;
constexpr size_t NA_RESCUE_BYTES = sizeof(Rescue) * (size_t)VRN_C2P_NOALLOC_MAX_CELLS;
constexpr size_t NA_SCRATCH_BYTES = NA_CNT_BYTES + NA_PART_BYTES + NA_RESCUE_BYTES;
static_assert(sizeof(Counters) <= NA_CNT_BYTES, "counter block");
static_assert(NA_PART_BYTES % 256 == 0, "rescue list alignment");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
unsigned char *na_ptr[NA_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
int noalloc_scratch(unsigned char **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(-1, "vrn_c2p_solve_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(na_mutex);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(-2, "vrn_c2p_solve_noalloc: no device address for the scratch array");
        na_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = na_ptr[dev];
    return 0;
}
}  // namespace

extern "C" int vrn_c2p_solve_noalloc(int64_t n, const double *U, const double *P_prev, double *P, double gamma,
                                     int guess, int n_iter, double rtol, void *stream, vrn_c2p_stats *stats)
{
    if (n < 1 || n > VRN_C2P_NOALLOC_MAX_CELLS)
        return fail(-1, "vrn_c2p_solve_noalloc: n = %lld outside [1, %d] (the scratch array's capacity)",
                    (long long)n, VRN_C2P_NOALLOC_MAX_CELLS);
    if (!U || !P) return fail(-1, "vrn_c2p_solve_noalloc: NULL U or P");
    if (!(gamma > 1.0 && gamma <= 2.0)) return fail(-1, "vrn_c2p_solve_noalloc: gamma %g outside (1, 2]", gamma);
    if (guess != VRN_GUESS_GENERIC && guess != VRN_GUESS_WARM && guess != VRN_GUESS_ANALYTIC)
        return fail(-1, "vrn_c2p_solve_noalloc: guess %d not available (GENERIC, WARM or ANALYTIC)", guess);
    if (guess == VRN_GUESS_WARM && !P_prev) return fail(-1, "vrn_c2p_solve_noalloc: warm guess needs P_prev");
    if (n_iter < 1 || n_iter > 10000) return fail(-1, "vrn_c2p_solve_noalloc: n_iter %d out of range", n_iter);
    if (!(rtol > 0.0 && rtol < 1.0)) return fail(-1, "vrn_c2p_solve_noalloc: rtol %g out of range", rtol);
    unsigned char *base = nullptr;
    if (int rc = noalloc_scratch(&base)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Counters *cnt = reinterpret_cast<Counters *>(base);
    SolveArgs a;
    a.n = n;
    a.U = U;
    a.Pprev = P_prev;
    a.P = P;
    a.W0 = nullptr;
    a.iters = nullptr;
    a.cnt = cnt;
    a.part = reinterpret_cast<unsigned long long *>(base + NA_CNT_BYTES);
    a.rescue = reinterpret_cast<Rescue *>(base + NA_CNT_BYTES + NA_PART_BYTES);
    a.g = gamma;
    a.rtol = rtol;
    a.n_iter = n_iter;
    a.record = 0;
    a.net = nullptr;
    // the counters keep the previous call's values: cleared by a kernel
    // (no memset node, so the sequence can also be captured into a graph)
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(128), 0, s, cnt);
    const int grid = grid_for((n + 63) / 64, 4, MAX_SOLVE_GRID);
    const int rgrid = 1024;
    switch (guess) {
    case VRN_GUESS_GENERIC:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_GENERIC>, dim3(grid), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, s,
                           (const unsigned long long *)a.part, grid, cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_GENERIC>, dim3(rgrid), dim3(256), 0, s, a);
        break;
    case VRN_GUESS_WARM:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_WARM>, dim3(grid), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, s,
                           (const unsigned long long *)a.part, grid, cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_WARM>, dim3(rgrid), dim3(256), 0, s, a);
        break;
    default:
        hipLaunchKernelGGL(k_solve<VRN_GUESS_ANALYTIC>, dim3(grid), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_reduce_counters, dim3(PART_STRIDE), dim3(256), 0, s,
                           (const unsigned long long *)a.part, grid, cnt);
        hipLaunchKernelGGL(k_rescue<VRN_GUESS_ANALYTIC>, dim3(rgrid), dim3(256), 0, s, a);
        break;
    }
    HIPCHK(hipGetLastError());
    if (stats) {
        Counters c;
        HIPCHK(hipMemcpyAsync(&c, cnt, sizeof c, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        stats->n = n;
        stats->n_unrecovered = (int64_t)c.n_unrecovered;
        stats->n_fallback = (int64_t)c.n_fallback;
        stats->n_guess_bad = (int64_t)c.n_guess_bad;
        stats->iters_sum = (int64_t)c.iters_sum;
        stats->iters_max = (int64_t)c.iters_max;
        stats->wave_max_sum = (int64_t)c.wave_max_sum;
        stats->n_waves = (int64_t)c.n_waves;
        for (int b = 0; b < 64; b++) stats->hist[b] = (int64_t)c.hist[b];
        stats->ms = 0.0;
    }
    return 0;
}
