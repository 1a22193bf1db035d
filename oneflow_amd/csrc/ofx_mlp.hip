// ofx_mlp.hip -- the warm-start predictor: a BF16 multilayer perceptron
// (inputs 9 -> 5 hidden layers of 256 with SiLU -> 5 outputs) trained and
// evaluated with hand-written MFMA kernels on the GPU.  No counterpart in the
// reference: it predicts a starting flow field, and the solver then iterates
// from it to the same tolerance as from freestream.
//
// Numerics: weights are kept in FP32 (Adam master copy) and used in BF16;
// activations between layers are BF16; every product is accumulated in FP32
// by v_mfma_f32_16x16x32_bf16.  Pre-activations are kept in FP32 for the
// backward pass.
//
// Layout: activations row-major [samples][features]; weights [in][out]
// row-major, with a transposed bf16 copy [out][in] for the forward products
// (the GEMM reads both operands along k).  Feature counts are padded: input 9 -> 16, output 5 -> 64 (the
// padding weights are zero and stay zero: their gradients are exactly zero);
// the sample count is padded to a multiple of 64 with zero rows that carry no
// loss.
#include "ofx_common.h"

#include <mutex>

namespace ofx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint16_t bf16;

constexpr int NIN = 9, NIN_P = 16, NHID = 256, NLAYER = 6, NOUT = 5, NOUT_P = 64;
constexpr int MLP_MAXN = 1 << 20;  // samples per call
constexpr int SPLITK_CHUNK = 1024;

__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((uint32_t)h) << 16); }
// SiLU and its derivative with the hardware reciprocal (v_rcp_f32, 1 ulp)
// instead of a correctly rounded division: the epilogues evaluate one per
// output element, and a division is ten instructions against one.  The
// values feed BF16 activations, whose rounding is 2^-9 relative.
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

enum { EPI_PARTIAL = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_DSILU = 3 };

// C[M x N] = A[M x K] Bt[N x K]^T: both operands bf16 with k contiguous
// (row strides lda, ldb), so every global and LDS access moves 8 elements
// (16 bytes) at a time.  BM x BN block tile (64 or 128 each; M and N
// multiples of them); blockIdx.z selects the K range [z*kchunk,
// (z+1)*kchunk), a multiple of 16 (an 8-element vector at or past its end is
// zero-filled, so K = 16 runs through the 32-deep MFMA with zero products
// added).  256 threads: wave w owns rows (BM/2)*(w>>1).. and columns
// (BN/2)*(w&1).. as (BM/32) x (BN/32) tiles of gfx950's
// v_mfma_f32_16x16x32_bf16.  MFMA operands: lane l supplies A[l%16][8*(l/16)
// + 0..7] and Bt[l%16][8*(l/16) + 0..7] and receives D[4*(l/16)+r][l%16],
// r = 0..3.  K advances KC at a time through LDS (one barrier pair per KC/32
// MFMA steps); the next chunk's global loads are issued into registers before
// the current chunk's MFMAs, so their latency overlaps the arithmetic.  The
// MFMA chain runs over k in ascending order of 32-wide steps.  The larger
// tiles re-read each operand from L2 fewer times (the forward of many samples).
template <int EPI, int KC, int BM, int BN>
__global__ void __launch_bounds__(256) k_gemm(const bf16* __restrict__ A, int lda, const bf16* __restrict__ Bt, int ldb,
                                              int kchunk, int M, float* __restrict__ Cf, bf16* __restrict__ Cb, int ldc,
                                              const float* __restrict__ bias, const float* __restrict__ zin,
                                              float* __restrict__ zout) {
    constexpr int PAD = 8;              // bf16 elements of row padding (LDS bank spread)
    constexpr int VR = KC / 8;          // 8-element vectors per tile row
    constexpr int NVA = BM * VR / 256;  // vectors per thread of the A tile
    constexpr int NVB = BN * VR / 256;  // and of the Bt tile
    constexpr int MT = BM / 32, NT = BN / 32;  // MFMA tiles per wave
    static_assert(KC % 32 == 0 && NVA >= 1 && NVB >= 1, "tile shape");
    __shared__ __align__(16) bf16 sA[BM][KC + PAD];
    __shared__ __align__(16) bf16 sB[BN][KC + PAD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (BM / 2) * (w >> 1), wn = (BN / 2) * (w & 1);
    const int kb = blockIdx.z * kchunk, ke = kb + kchunk;
    v4f acc[MT][NT];
    for (int i = 0; i < MT; ++i)
        for (int j = 0; j < NT; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    // vector e = u*256 + t of a tile: row e / VR, k offset 8*(e % VR)
    v4u ra[NVA], rb[NVB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NVA; ++u) {
            const int e = u * 256 + t, r = e / VR, kv = 8 * (e % VR);
            ra[u] = (k0 + kv < ke) ? *(const v4u*)&A[(size_t)(m0 + r) * lda + k0 + kv] : v4u{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < NVB; ++u) {
            const int e = u * 256 + t, r = e / VR, kv = 8 * (e % VR);
            rb[u] = (k0 + kv < ke) ? *(const v4u*)&Bt[(size_t)(n0 + r) * ldb + k0 + kv] : v4u{0u, 0u, 0u, 0u};
        }
    };
    gload(kb);
    for (int k0 = kb; k0 < ke; k0 += KC) {
        __syncthreads();  // the previous chunk's MFMAs have read LDS
#pragma unroll
        for (int u = 0; u < NVA; ++u) {
            const int e = u * 256 + t, r = e / VR, kv = 8 * (e % VR);
            *(v4u*)&sA[r][kv] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NVB; ++u) {
            const int e = u * 256 + t, r = e / VR, kv = 8 * (e % VR);
            *(v4u*)&sB[r][kv] = rb[u];
        }
        __syncthreads();
        if (k0 + KC < ke) gload(k0 + KC);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 32) {
            v8bf a[MT], b[NT];
            for (int i = 0; i < MT; ++i) a[i] = *(const v8bf*)&sA[wm + 16 * i + (lane & 15)][ks + 8 * (lane >> 4)];
            for (int j = 0; j < NT; ++j) b[j] = *(const v8bf*)&sB[wn + 16 * j + (lane & 15)][ks + 8 * (lane >> 4)];
            for (int i = 0; i < MT; ++i)
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    for (int i = 0; i < MT; ++i)
        for (int j = 0; j < NT; ++j)
            for (int r = 0; r < 4; ++r) {
                int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                int col = n0 + wn + 16 * j + (lane & 15);
                size_t o = (size_t)row * ldc + col;
                float v = acc[i][j][r];
                if (EPI == EPI_PARTIAL) {
                    Cf[(size_t)blockIdx.z * M * ldc + o] = v;
                } else if (EPI == EPI_BIAS_SILU) {
                    float z = v + bias[col];
                    if (zout) zout[o] = z;  // pre-activations kept for a backward pass only
                    Cb[o] = f2bf(silu(z));
                } else if (EPI == EPI_BIAS) {
                    Cf[o] = v + bias[col];
                } else {  // EPI_DSILU: gradient through the activation
                    float g = v * dsilu(zin[o]);
                    Cf[o] = g;
                    Cb[o] = f2bf(g);
                }
            }
}

// The network's weights and biases as the one-kernel forward pass reads them.
struct MlpFwd {
    const bf16* wt[NLAYER];  // transposed bf16 weights [kout][kin]
    const float* b[NLAYER];
};

// The whole forward pass of 64 samples in one workgroup: the activations of
// every hidden layer stay in LDS (64 x 256 bf16, 32 KiB), so HBM sees only the
// packed inputs, the weights (from L2) and the outputs, and several
// workgroups share a CU, so one's layer barriers do not idle the MFMA units.
// Four waves; wave w computes all 64 rows and columns 64w..64w+63 of a
// hidden layer (4 x 4 tiles of v_mfma_f32_16x16x32_bf16, operands as in
// k_gemm: A from LDS, Bt straight from the weights), then all waves write
// their tiles back over the layer's input once every wave has read it.  The
// output layer needs its first 16 (of 64 padded) columns only: wave w
// computes rows 16w..16w+15 of them.
// Every product, sum order and epilogue is k_gemm's (the MFMA chain over k in
// ascending 32-wide steps from zero, the 16-wide input layer through one MFMA
// with k 16..31 zero), so the outputs are those of the layer-by-layer pass
// bit for bit.  LDS rows are 512 bytes with no padding; the 16-byte vector v
// of row r is stored at v ^ (r & 15), so the 16 rows an MFMA operand load
// reads fall in different banks.  Mp is a multiple of 64; rows at or past it
// would read zero inputs and not be stored.
__device__ __forceinline__ int swz(int r, int k) { return (((k >> 3) ^ (r & 15)) << 3) | (k & 7); }

__global__ void __launch_bounds__(256) k_mlp_forward(const bf16* __restrict__ x16, int Mp, MlpFwd P,
                                                     float* __restrict__ y) {
    constexpr int BM = 64;
    __shared__ __align__(16) bf16 sH[BM][NHID];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.x * BM;
    const int r0 = 0, c0 = 64 * w;  // this wave's tile of a hidden layer
    const int lr = lane & 15, lk = 8 * (lane >> 4);    // operand row within a 16-tile; k offset of the lane's 8
    const v8bf zero8 = {};
    v4f acc[4][4];
    // bias + SiLU of this wave's tile into the activation rows
    auto store_silu = [&](const float* bias) {
        for (int j = 0; j < 4; ++j) {
            const int col = c0 + 16 * j + lr;
            const float bc = bias[col];
            for (int i = 0; i < 4; ++i)
                for (int r = 0; r < 4; ++r) {
                    const int row = r0 + 16 * i + 4 * (lane >> 4) + r;
                    sH[row][swz(row, col)] = f2bf(silu(acc[i][j][r] + bc));
                }
        }
    };
    // layer 0: k = 16 packed inputs, k 16..31 of the MFMA zero
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    {
        v8bf b[4];
        for (int j = 0; j < 4; ++j)
            b[j] = lk < NIN_P ? *(const v8bf*)&P.wt[0][(size_t)(c0 + 16 * j + lr) * NIN_P + lk] : zero8;
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + r0 + 16 * i + lr;
            const v8bf a = (lk < NIN_P && row < Mp) ? *(const v8bf*)&x16[(size_t)row * NIN_P + lk] : zero8;
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
    }
    store_silu(P.b[0]);
    __syncthreads();
    // hidden layers 1..4: 256 -> 256, 8 k steps of 32 each.  The weights are
    // loaded from L2 two steps ahead (across layer ends: they do not depend
    // on the activations) and the activations from LDS one step ahead, into
    // register rings indexed by compile-time step numbers (both loops
    // unrolled).
    v8bf B[3][4], Am[2][4];
    auto load_b = [&](const bf16* wt, int ks, v8bf* dst) {
        for (int j = 0; j < 4; ++j) dst[j] = *(const v8bf*)&wt[(size_t)(c0 + 16 * j + lr) * NHID + ks + lk];
    };
    auto load_a = [&](int ks, v8bf* dst) {
        for (int i = 0; i < 4; ++i) {
            const int row = r0 + 16 * i + lr;
            dst[i] = *(const v8bf*)&sH[row][swz(row, ks + lk)];
        }
    };
    constexpr int NSTEP = NHID / 32, NHL = NLAYER - 2;  // k steps per layer, hidden 256 -> 256 layers
    load_b(P.wt[1], 0, B[0]);
    load_b(P.wt[1], 32, B[1]);
#pragma unroll
    for (int h = 0; h < NHL; ++h) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        load_a(0, Am[0]);
#pragma unroll
        for (int st = 0; st < NSTEP; ++st) {
            const int g = h * NSTEP + st, g2 = g + 2;
            if (g2 < NHL * NSTEP) load_b(P.wt[1 + g2 / NSTEP], 32 * (g2 % NSTEP), B[g2 % 3]);
            if (st + 1 < NSTEP) load_a(32 * (st + 1), Am[(st + 1) & 1]);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am[st & 1][i], B[g % 3][j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();  // every wave has read this layer's input
        store_silu(P.b[1 + h]);
        __syncthreads();
    }
    // output layer: columns 0..15 of the padded 64, rows 16w..16w+15
    {
        const int L = NLAYER - 1;
        v4f o = v4f{0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < NHID; ks += 32) {
            const v8bf b = *(const v8bf*)&P.wt[L][(size_t)lr * NHID + ks + lk];
            const int row = 16 * w + lr;
            const v8bf a = *(const v8bf*)&sH[row][swz(row, ks + lk)];
            o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, o, 0, 0, 0);
        }
        const float bc = P.b[L][lr];
        for (int r = 0; r < 4; ++r) {
            const int row = m0 + 16 * w + 4 * (lane >> 4) + r;
            if (row < Mp) y[(size_t)row * NOUT_P + lr] = o[r] + bc;
        }
    }
}

// out[c][r] = in[r][c] for an R x Cc bf16 matrix; rows R..Rp-1 of the output's
// transposed shape (columns of `in` beyond Cc) are zero-filled up to Cp.
__global__ void k_transpose(const bf16* in, int R, int Cc, bf16* out, int Cp) {
    __shared__ bf16 tile[32][33];
    int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        int r = by + k, c = bx + threadIdx.x;
        tile[k][threadIdx.x] = (r < R && c < Cc) ? in[(size_t)r * Cc + c] : (bf16)0;
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        int c = bx + k, r = by + threadIdx.x;
        if (c < Cp && r < R) out[(size_t)c * R + r] = tile[threadIdx.x][k];
    }
}

// Sum of the split-K partials in split order.
__global__ void k_splitk_sum(const float* part, int nsplit, size_t n, float* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; ++z) s += part[(size_t)z * n + i];
    out[i] = s;
}

// Column sums of an M x N fp32 matrix (bias gradients), two stages in a fixed
// order: block (x, y) sums columns 64x..64x+63 over rows COLSUM_ROWS*y ..
// (4 row lanes of COLSUM_ROWS/4 rows, then the 4 lanes in order) into
// part[y][N]; then one block per 64 columns adds the row-chunk partials: row
// lane r takes chunks r, r+4, r+8, ... in order, then the 4 lanes in order.
constexpr int COLSUM_ROWS = 64;
__global__ void __launch_bounds__(256) k_colsum_part(const float* g, int M, int N, float* part) {
    __shared__ float sm[4][64];
    int c = blockIdx.x * 64 + (threadIdx.x & 63), lane4 = threadIdx.x >> 6;
    int r0 = blockIdx.y * COLSUM_ROWS + lane4 * (COLSUM_ROWS / 4);
    float s = 0.f;
    if (c < N)
#pragma unroll
        for (int r = r0; r < r0 + COLSUM_ROWS / 4; ++r)
            if (r < M) s += g[(size_t)r * N + c];
    sm[lane4][threadIdx.x & 63] = s;
    __syncthreads();
    if (threadIdx.x < 64 && c < N)
        part[(size_t)blockIdx.y * N + c] = ((sm[0][threadIdx.x] + sm[1][threadIdx.x]) + sm[2][threadIdx.x]) + sm[3][threadIdx.x];
}
__global__ void __launch_bounds__(256) k_colsum_final(const float* part, int chunks, int N, float* out) {
    __shared__ float sm[4][64];
    int c = blockIdx.x * 64 + (threadIdx.x & 63), lane4 = threadIdx.x >> 6;
    float s = 0.f;
    if (c < N)
        for (int k = lane4; k < chunks; k += 4) s += part[(size_t)k * N + c];
    sm[lane4][threadIdx.x & 63] = s;
    __syncthreads();
    if (threadIdx.x < 64 && c < N)
        out[c] = ((sm[0][threadIdx.x] + sm[1][threadIdx.x]) + sm[2][threadIdx.x]) + sm[3][threadIdx.x];
}

// Inputs fp32 [n][NIN] -> bf16 [Mp][NIN_P] (zero padded).
__global__ void k_pack_inputs(const float* x, int n, int Mp, bf16* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Mp * NIN_P) return;
    int r = i / NIN_P, c = i % NIN_P;
    out[i] = (r < n && c < NIN) ? f2bf(x[(size_t)r * NIN + c]) : (bf16)0;
}

// MSE over the real outputs: loss partials and the bf16 / fp32 output gradient.
__global__ void k_mse(const float* y, const float* tgt, int n, int Mp, bf16* dyb, float* dyf, float* lpart) {
    __shared__ float sm[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < Mp * NOUT_P) {
        int r = i / NOUT_P, c = i % NOUT_P;
        float g = 0.f;
        if (r < n && c < NOUT) {
            float d = y[i] - tgt[(size_t)r * NOUT + c];
            l = d * d;
            g = 2.0f * d / (float)(n * NOUT);
        }
        dyb[i] = f2bf(g);
        dyf[i] = g;
    }
    sm[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = sm[0];
}

// Sum of n partials by one block (strided, then a tree in LDS; fixed order).
__global__ void __launch_bounds__(256) k_sum_partials(const float* p, int n, float scale, float* out) {
    __shared__ double sm[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)(sm[0] * scale);
}

// Adam on fp32 master weights, bf16 copy refreshed.
__global__ void k_adam(float* w, const float* g, float* m, float* v, bf16* wb, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// Mini-batch rows idx[0..b) of the device training set into the input/target buffers.
__global__ void k_gather_batch(const float* x, const float* t, const int* idx, int b, float* xo, float* to) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b) return;
    int r = idx[i];
    for (int c = 0; c < NIN; ++c) xo[(size_t)i * NIN + c] = x[(size_t)r * NIN + c];
    for (int c = 0; c < NOUT; ++c) to[(size_t)i * NOUT + c] = t[(size_t)r * NOUT + c];
}

__global__ void k_copy1(const float* a, float* b) {
    if (threadIdx.x == 0 && blockIdx.x == 0) b[0] = a[0];
}

// y[i][0..NOUT) of the padded output rows, packed densely for one host copy
__global__ void k_pack_outputs(const float* __restrict__ yp, int nsamp, float* __restrict__ y) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)nsamp * NOUT) y[i] = yp[(i / NOUT) * NOUT_P + i % NOUT];
}

__global__ void k_to_bf16(const float* a, size_t n, bf16* b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

}  // namespace ofx

using namespace ofx;

struct ofx_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int kin[NLAYER], kout[NLAYER];        // padded shapes of each layer
    float *w[NLAYER], *b[NLAYER];          // fp32 master
    bf16 *wb[NLAYER], *wt[NLAYER];         // bf16 copy and its transpose
    float *mw[NLAYER], *vw[NLAYER], *mb[NLAYER], *vb[NLAYER];
    float *gw[NLAYER], *gb[NLAYER];        // gradients
    int step = 0;
    // per-call buffers, sized for `cap` samples
    int cap = 0, max_split = 1;
    bf16* h[NLAYER];   // layer inputs: h[0] = packed inputs, h[l] = silu(z[l-1])
    float* z[NLAYER - 1];
    float* y = nullptr;
    float* tgt = nullptr;
    float* xin = nullptr;
    bf16 *dyb = nullptr, *ht = nullptr, *gt = nullptr;  // ht, gt: transposed factors of a weight gradient
    float *dyf = nullptr, *gz = nullptr;  // fp32 gradient of the current layer's pre-activation
    bf16 *gzb = nullptr, *gzb2 = nullptr;
    float *part = nullptr, *lpart = nullptr, *loss = nullptr, *cpart = nullptr;
    int fwd_variant = 1;  // inference forward pass: 0 layer by layer, 1 one kernel (k_mlp_forward)
    std::vector<void*> allocs;
};

namespace {

int mlp_alloc(ofx_mlp* n, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) { set_error("mlp: out of device memory"); return OFX_ENOMEM; }
    n->allocs.push_back(*p);
    OFX_HIP(hipMemsetAsync(*p, 0, bytes, n->stream));
    return 0;
}

int mlp_reserve(ofx_mlp* n, int samples) {
    int Mp = (samples + 63) / 64 * 64;
    if (Mp <= n->cap) return 0;
    // grow: drop the old per-call buffers (parameters are the first allocations)
    OFX_HIP(hipStreamSynchronize(n->stream));
    size_t keep = NLAYER * 12;
    for (size_t i = keep; i < n->allocs.size(); ++i) (void)hipFree(n->allocs[i]);
    n->allocs.resize(std::min(keep, n->allocs.size()));
    auto A = [&](void** p, size_t bytes) { return mlp_alloc(n, p, bytes); };
    for (int l = 0; l < NLAYER; ++l) OFX_TRY(A((void**)&n->h[l], sizeof(bf16) * (size_t)Mp * n->kin[l]));
    for (int l = 0; l < NLAYER - 1; ++l) OFX_TRY(A((void**)&n->z[l], sizeof(float) * (size_t)Mp * NHID));
    OFX_TRY(A((void**)&n->y, sizeof(float) * (size_t)Mp * NOUT_P));
    OFX_TRY(A((void**)&n->tgt, sizeof(float) * (size_t)Mp * NOUT));
    OFX_TRY(A((void**)&n->xin, sizeof(float) * (size_t)Mp * NIN));
    OFX_TRY(A((void**)&n->dyb, sizeof(bf16) * (size_t)Mp * NOUT_P));
    OFX_TRY(A((void**)&n->dyf, sizeof(float) * (size_t)Mp * NOUT_P));
    OFX_TRY(A((void**)&n->ht, sizeof(bf16) * (size_t)Mp * NHID));
    OFX_TRY(A((void**)&n->gt, sizeof(bf16) * (size_t)Mp * NHID));
    OFX_TRY(A((void**)&n->gz, sizeof(float) * (size_t)Mp * NHID));
    OFX_TRY(A((void**)&n->gzb, sizeof(bf16) * (size_t)Mp * NHID));
    OFX_TRY(A((void**)&n->gzb2, sizeof(bf16) * (size_t)Mp * NHID));
    // weight_grad doubles the split count while a split exceeds SPLITK_CHUNK
    // samples, so it stays below 2 Mp / SPLITK_CHUNK; one more slot holds the sum
    n->max_split = 2 * (Mp / SPLITK_CHUNK) + 1;
    OFX_TRY(A((void**)&n->part, sizeof(float) * (size_t)(n->max_split + 1) * NHID * NHID));
    OFX_TRY(A((void**)&n->lpart, sizeof(float) * ((size_t)Mp * NOUT_P / 256 + 1)));
    OFX_TRY(A((void**)&n->loss, sizeof(float) * 4));
    OFX_TRY(A((void**)&n->cpart, sizeof(float) * ((size_t)Mp / COLSUM_ROWS * NHID + NHID)));
    n->cap = Mp;
    return 0;
}

// C = A Bt^T (see k_gemm): A [M][K], Bt [N][K].
template <int EPI>
int gemm_on(hipStream_t st, const bf16* A, int lda, const bf16* Bt, int ldb, int M, int N, int K, int nsplit, float* Cf,
            bf16* Cb, int ldc, const float* bias, const float* zin, float* zout) {
    if (M % 64 || N % 64 || K % 16 || K % nsplit || (K / nsplit) % 16) {
        set_error("mlp: GEMM shape not tiled");
        return OFX_EINVAL;
    }
    int kc = K / nsplit;
    // 128-row tiles (and 128 columns where N allows) once there are rows
    // enough to fill the GPU with them: fewer L2 re-reads of both operands
    if (nsplit == 1 && M % 128 == 0 && M >= 65536 && kc % 64 == 0) {
        if (N % 128 == 0)
            k_gemm<EPI, 64, 128, 128><<<dim3(N / 128, M / 128, 1), 256, 0, st>>>(A, lda, Bt, ldb, kc, M, Cf, Cb, ldc,
                                                                                 bias, zin, zout);
        else
            k_gemm<EPI, 64, 128, 64><<<dim3(N / 64, M / 128, 1), 256, 0, st>>>(A, lda, Bt, ldb, kc, M, Cf, Cb, ldc,
                                                                               bias, zin, zout);
    } else {
        dim3 grid(N / 64, M / 64, nsplit);
        if (kc % 64 == 0)
            k_gemm<EPI, 64, 64, 64><<<grid, 256, 0, st>>>(A, lda, Bt, ldb, kc, M, Cf, Cb, ldc, bias, zin, zout);
        else
            k_gemm<EPI, 32, 64, 64><<<grid, 256, 0, st>>>(A, lda, Bt, ldb, kc, M, Cf, Cb, ldc, bias, zin, zout);
    }
    OFX_HIP(hipGetLastError());
    return 0;
}

template <int EPI>
int gemm(ofx_mlp* n, const bf16* A, int lda, const bf16* Bt, int ldb, int M, int N, int K, int nsplit, float* Cf, bf16* Cb,
         int ldc, const float* bias, const float* zin, float* zout) {
    return gemm_on<EPI>(n->stream, A, lda, Bt, ldb, M, N, K, nsplit, Cf, Cb, ldc, bias, zin, zout);
}

int transpose(ofx_mlp* n, const bf16* in, int R, int C, bf16* out, int Cp) {
    dim3 grid((std::max(C, Cp) + 31) / 32, (R + 31) / 32);
    k_transpose<<<grid, dim3(32, 8), 0, n->stream>>>(in, R, C, out, Cp);
    OFX_HIP(hipGetLastError());
    return 0;
}

// Weight gradient dW[kin x kout] = H^T[kin x Mp] G[Mp x kout], split over
// samples: both factors transposed so that k (the sample) is contiguous.
int weight_grad(ofx_mlp* n, int l, const bf16* H, const bf16* G, int Mp) {
    int ki = n->kin[l], ko = n->kout[l];
    int rows = (ki + 63) / 64 * 64;  // the input layer's 16 rows padded to one tile
    OFX_TRY(transpose(n, H, Mp, ki, n->ht, rows));
    OFX_TRY(transpose(n, G, Mp, ko, n->gt, ko));
    int nsplit = 1;
    while (Mp / nsplit > SPLITK_CHUNK && Mp % (nsplit * 2) == 0 && (Mp / (nsplit * 2)) % 16 == 0) nsplit *= 2;
    if (nsplit > n->max_split || Mp > n->cap) { set_error("mlp: split-K buffer too small"); return OFX_EINVAL; }
    OFX_TRY(gemm<EPI_PARTIAL>(n, n->ht, Mp, n->gt, Mp, rows, ko, Mp, nsplit, n->part, nullptr, ko, nullptr, nullptr, nullptr));
    size_t cnt = (size_t)rows * ko;
    // rows beyond kin are zero inputs (their sums are zero); keep the first kin rows
    k_splitk_sum<<<(unsigned)((cnt + 255) / 256), 256, 0, n->stream>>>(n->part, nsplit, cnt, n->part + (size_t)nsplit * cnt);
    OFX_HIP(hipGetLastError());
    OFX_HIP(hipMemcpyAsync(n->gw[l], n->part + (size_t)nsplit * cnt, sizeof(float) * (size_t)ki * ko,
                           hipMemcpyDeviceToDevice, n->stream));
    return 0;
}

int colsum(ofx_mlp* n, const float* g, int M, int N, float* out) {
    int chunks = (M + COLSUM_ROWS - 1) / COLSUM_ROWS;
    if ((size_t)chunks * N > (size_t)n->cap / COLSUM_ROWS * NHID + NHID) { set_error("mlp: column-sum buffer too small"); return OFX_EINVAL; }
    k_colsum_part<<<dim3((N + 63) / 64, chunks), 256, 0, n->stream>>>(g, M, N, n->cpart);
    k_colsum_final<<<(N + 63) / 64, 256, 0, n->stream>>>(n->cpart, chunks, N, out);
    OFX_HIP(hipGetLastError());
    return 0;
}

// All layers in one kernel (k_mlp_forward): packed inputs x16 [Mp][16] ->
// padded outputs y [Mp][64] (columns 0..15 written).
int forward_fused(const ofx_mlp* n, hipStream_t st, const bf16* x16, int Mp, float* y) {
    MlpFwd P;
    for (int l = 0; l < NLAYER; ++l) { P.wt[l] = n->wt[l]; P.b[l] = n->b[l]; }
    k_mlp_forward<<<(Mp + 63) / 64, 256, 0, st>>>(x16, Mp, P, y);
    OFX_HIP(hipGetLastError());
    return 0;
}

// keep_z: store the pre-activations (a backward pass follows).
int forward(ofx_mlp* n, int Mp, bool keep_z) {
    if (!keep_z && n->fwd_variant == 1) return forward_fused(n, n->stream, n->h[0], Mp, n->y);
    for (int l = 0; l < NLAYER - 1; ++l)
        OFX_TRY(gemm<EPI_BIAS_SILU>(n, n->h[l], n->kin[l], n->wt[l], n->kin[l], Mp, n->kout[l], n->kin[l], 1, nullptr,
                                    n->h[l + 1], n->kout[l], n->b[l], nullptr, keep_z ? n->z[l] : nullptr));
    int L = NLAYER - 1;
    OFX_TRY(gemm<EPI_BIAS>(n, n->h[L], n->kin[L], n->wt[L], n->kin[L], Mp, n->kout[L], n->kin[L], 1, n->y, nullptr,
                           n->kout[L], n->b[L], nullptr, nullptr));
    return 0;
}

// The transposed bf16 weights [kout][kin] that the forward products read,
// refreshed whenever the bf16 weights change.
int refresh_wt(ofx_mlp* n) {
    for (int l = 0; l < NLAYER; ++l) OFX_TRY(transpose(n, n->wb[l], n->kin[l], n->kout[l], n->wt[l], n->kout[l]));
    return 0;
}

int backward(ofx_mlp* n, int nsamp, int Mp) {
    int L = NLAYER - 1;
    int nb = (Mp * NOUT_P + 255) / 256;
    k_mse<<<nb, 256, 0, n->stream>>>(n->y, n->tgt, nsamp, Mp, n->dyb, n->dyf, n->lpart);
    k_sum_partials<<<1, 256, 0, n->stream>>>(n->lpart, nb, 1.0f / (float)(nsamp * NOUT), n->loss);
    OFX_HIP(hipGetLastError());
    // output layer
    OFX_TRY(weight_grad(n, L, n->h[L], n->dyb, Mp));
    OFX_TRY(colsum(n, n->dyf, Mp, NOUT_P, n->gb[L]));
    const bf16* g = n->dyb;
    bf16* gnext = n->gzb;
    for (int l = L; l >= 1; --l) {
        // gradient of layer l-1's pre-activation: (G W_l^T) * silu'(z_{l-1});
        // the second factor read along k is W_l itself
        OFX_TRY(gemm<EPI_DSILU>(n, g, n->kout[l], n->wb[l], n->kout[l], Mp, n->kin[l], n->kout[l], 1, n->gz, gnext,
                                n->kin[l], nullptr, n->z[l - 1], nullptr));
        OFX_TRY(weight_grad(n, l - 1, n->h[l - 1], gnext, Mp));
        OFX_TRY(colsum(n, n->gz, Mp, NHID, n->gb[l - 1]));
        OFX_HIP(hipGetLastError());
        g = gnext;
        gnext = (gnext == n->gzb) ? n->gzb2 : n->gzb;
    }
    return 0;
}

int load_batch(ofx_mlp* n, const float* x, const float* t, int nsamp, bool on_device) {
    OFX_TRY(mlp_reserve(n, nsamp));
    int Mp = (nsamp + 63) / 64 * 64;
    hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    OFX_HIP(hipMemcpyAsync(n->xin, x, sizeof(float) * (size_t)nsamp * NIN, kind, n->stream));
    if (t) OFX_HIP(hipMemcpyAsync(n->tgt, t, sizeof(float) * (size_t)nsamp * NOUT, kind, n->stream));
    k_pack_inputs<<<(Mp * NIN_P + 255) / 256, 256, 0, n->stream>>>(n->xin, nsamp, Mp, n->h[0]);
    OFX_HIP(hipGetLastError());
    return 0;
}

// ---- allocation-free forward pass (ofx_mlp_forward_noalloc)
// Its work arrays live in one file-scope device array, `scratch`, of a fixed
// size: packed inputs [MAXN][16] bf16, two activation buffers [MAXN][256]
// bf16 (layers alternate between them) and the padded outputs [MAXN][64]
// fp32 -- 10.25 MiB, held by every GPU the library runs on until the process
// ends.  Every call on one GPU uses that GPU's copy; each element a call reads
// it has written first, so nothing needs clearing, and calls that overlap on
// one GPU race.
constexpr int NOALLOC_MAXN = 8192;
constexpr size_t NOALLOC_IN = sizeof(bf16) * NOALLOC_MAXN * NIN_P;
constexpr size_t NOALLOC_ACT = sizeof(bf16) * NOALLOC_MAXN * NHID;
constexpr size_t NOALLOC_OUT = sizeof(float) * NOALLOC_MAXN * NOUT_P;
constexpr size_t NOALLOC_BYTES = NOALLOC_IN + 2 * NOALLOC_ACT + NOALLOC_OUT;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // device id -> address of its `scratch`

// The device address of `scratch` on `device`, looked up once per device with
// that device current (host code must never pass the array's name itself).
int scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        OFX_HIP(hipGetDevice(&cur));
        OFX_HIP(hipSetDevice(device));
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            set_error(std::string("mlp: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return OFX_EHIP;
        }
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[device];
    return 0;
}

}  // namespace

extern "C" {

int ofx_mlp_create(ofx_mlp** out, int device, unsigned long long seed) {
    if (!out) { set_error("null out"); return OFX_EINVAL; }
    *out = nullptr;
    ofx_mlp* n = new (std::nothrow) ofx_mlp();
    if (!n) { set_error("out of host memory"); return OFX_ENOMEM; }
    n->device = device;
    int rc = 0;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("mlp: no such HIP device");
        delete n;
        return OFX_EHIP;
    }
    for (int l = 0; l < NLAYER; ++l) {
        n->kin[l] = (l == 0) ? NIN_P : NHID;
        n->kout[l] = (l == NLAYER - 1) ? NOUT_P : NHID;
    }
    unsigned char* sp = nullptr;  // the allocation-free forward's work array on this device
    rc = scratch_address(device, &sp);
    // parameters first (mlp_reserve keeps the first NLAYER * 12 allocations)
    std::vector<float> hw, hb;
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto urand = [&]() {  // uniform in [-1, 1)
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    };
    for (int l = 0; l < NLAYER && !rc; ++l) {
        size_t nw = (size_t)n->kin[l] * n->kout[l];
        int fan_in = (l == 0) ? NIN : NHID;
        int real_out = (l == NLAYER - 1) ? NOUT : NHID;
        float a = std::sqrt(6.0f / (float)fan_in);  // uniform He-style initialisation
        hw.assign(nw, 0.0f);
        for (int i = 0; i < ((l == 0) ? NIN : NHID); ++i)
            for (int j = 0; j < real_out; ++j) hw[(size_t)i * n->kout[l] + j] = a * urand();
        hb.assign(n->kout[l], 0.0f);
        void* p;
        rc = mlp_alloc(n, &p, sizeof(float) * nw); n->w[l] = (float*)p;
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * n->kout[l]); n->b[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(bf16) * nw); n->wb[l] = (bf16*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(bf16) * nw); n->wt[l] = (bf16*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * nw); n->mw[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * nw); n->vw[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * n->kout[l]); n->mb[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * n->kout[l]); n->vb[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * nw); n->gw[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, sizeof(float) * n->kout[l]); n->gb[l] = (float*)p; }
        if (!rc) { rc = mlp_alloc(n, &p, 16); }  // spare slots keep the count at 12 per layer
        if (!rc) { rc = mlp_alloc(n, &p, 16); }
        if (!rc && hipMemcpyAsync(n->w[l], hw.data(), sizeof(float) * nw, hipMemcpyHostToDevice, n->stream) != hipSuccess)
            rc = OFX_EHIP;
        if (!rc) {
            k_to_bf16<<<(unsigned)((nw + 255) / 256), 256, 0, n->stream>>>(n->w[l], nw, n->wb[l]);
            if (hipGetLastError() != hipSuccess) rc = OFX_EHIP;
        }
        if (!rc && hipStreamSynchronize(n->stream) != hipSuccess) rc = OFX_EHIP;
    }
    if (!rc) rc = refresh_wt(n);
    if (rc) {
        ofx_mlp_destroy(n);
        return rc;
    }
    *out = n;
    return 0;
}

int ofx_mlp_destroy(ofx_mlp* n) {
    if (!n) return 0;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void* p : n->allocs) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return 0;
}

int ofx_mlp_param_count(ofx_mlp* n, int* count) {
    if (!n || !count) { set_error("null argument"); return OFX_EINVAL; }
    int c = 0;
    for (int l = 0; l < NLAYER; ++l) c += n->kin[l] * n->kout[l] + n->kout[l];
    *count = c;
    return 0;
}

// Parameters (or gradients) as one fp32 vector: for each layer W[kin_p][kout_p]
// then b[kout_p] (padded shapes: 16/256/.../256 -> 256/.../64).
static int mlp_copy(ofx_mlp* n, float* host, bool grads, bool to_device) {
    size_t o = 0;
    for (int l = 0; l < NLAYER; ++l) {
        size_t nw = (size_t)n->kin[l] * n->kout[l];
        float* dw = grads ? n->gw[l] : n->w[l];
        float* db = grads ? n->gb[l] : n->b[l];
        if (to_device) {
            OFX_HIP(hipMemcpyAsync(dw, host + o, sizeof(float) * nw, hipMemcpyHostToDevice, n->stream));
            OFX_HIP(hipMemcpyAsync(db, host + o + nw, sizeof(float) * n->kout[l], hipMemcpyHostToDevice, n->stream));
            if (!grads) k_to_bf16<<<(unsigned)((nw + 255) / 256), 256, 0, n->stream>>>(n->w[l], nw, n->wb[l]);
        } else {
            OFX_HIP(hipMemcpyAsync(host + o, dw, sizeof(float) * nw, hipMemcpyDeviceToHost, n->stream));
            OFX_HIP(hipMemcpyAsync(host + o + nw, db, sizeof(float) * n->kout[l], hipMemcpyDeviceToHost, n->stream));
        }
        o += nw + n->kout[l];
    }
    if (to_device && !grads) OFX_TRY(refresh_wt(n));
    OFX_HIP(hipStreamSynchronize(n->stream));
    return 0;
}

int ofx_mlp_get_params(ofx_mlp* n, float* host) {
    if (!n || !host) { set_error("null argument"); return OFX_EINVAL; }
    return mlp_copy(n, host, false, false);
}

int ofx_mlp_set_params(ofx_mlp* n, const float* host) {
    if (!n || !host) { set_error("null argument"); return OFX_EINVAL; }
    n->step = 0;
    return mlp_copy(n, const_cast<float*>(host), false, true);
}

int ofx_mlp_forward(ofx_mlp* n, const float* x, int nsamp, float* y) {
    if (!n || !x || !y || nsamp < 1 || nsamp > MLP_MAXN) { set_error("mlp_forward: bad arguments"); return OFX_EINVAL; }
    OFX_HIP(hipSetDevice(n->device));
    OFX_TRY(load_batch(n, x, nullptr, nsamp, false));
    int Mp = (nsamp + 63) / 64 * 64;
    OFX_TRY(forward(n, Mp, false));
    // tgt (Mp x NOUT floats) is free in a forward-only call: pack there, copy once
    size_t cnt = (size_t)nsamp * NOUT;
    k_pack_outputs<<<(unsigned)((cnt + 255) / 256), 256, 0, n->stream>>>(n->y, nsamp, n->tgt);
    OFX_HIP(hipGetLastError());
    OFX_HIP(hipMemcpyAsync(y, n->tgt, sizeof(float) * cnt, hipMemcpyDeviceToHost, n->stream));
    OFX_HIP(hipStreamSynchronize(n->stream));
    return 0;
}

int ofx_mlp_forward_noalloc(ofx_mlp* n, const float* x_dev, int nsamp, float* y_dev, void* stream) {
    if (!n || !x_dev || !y_dev || nsamp < 1) { set_error("mlp_forward_noalloc: bad arguments"); return OFX_EINVAL; }
    if (nsamp > NOALLOC_MAXN) {
        set_error("mlp_forward_noalloc: at most 8192 samples per call (fixed work array)");
        return OFX_EINVAL;
    }
    OFX_HIP(hipSetDevice(n->device));
    unsigned char* base = nullptr;
    OFX_TRY(scratch_address(n->device, &base));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : n->stream;
    bf16* in = reinterpret_cast<bf16*>(base);
    bf16* act[2] = {reinterpret_cast<bf16*>(base + NOALLOC_IN), reinterpret_cast<bf16*>(base + NOALLOC_IN + NOALLOC_ACT)};
    float* yp = reinterpret_cast<float*>(base + NOALLOC_IN + 2 * NOALLOC_ACT);
    const int Mp = (nsamp + 63) / 64 * 64;
    k_pack_inputs<<<(Mp * NIN_P + 255) / 256, 256, 0, st>>>(x_dev, nsamp, Mp, in);
    OFX_HIP(hipGetLastError());
    if (n->fwd_variant == 1) {
        OFX_TRY(forward_fused(n, st, in, Mp, yp));
    } else {
        const bf16* h = in;
        for (int l = 0; l < NLAYER - 1; ++l) {
            OFX_TRY(gemm_on<EPI_BIAS_SILU>(st, h, n->kin[l], n->wt[l], n->kin[l], Mp, n->kout[l], n->kin[l], 1,
                                           nullptr, act[l & 1], n->kout[l], n->b[l], nullptr, nullptr));
            h = act[l & 1];
        }
        const int L = NLAYER - 1;
        OFX_TRY(gemm_on<EPI_BIAS>(st, h, n->kin[L], n->wt[L], n->kin[L], Mp, n->kout[L], n->kin[L], 1, yp, nullptr,
                                  n->kout[L], n->b[L], nullptr, nullptr));
    }
    size_t cnt = (size_t)nsamp * NOUT;
    k_pack_outputs<<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(yp, nsamp, y_dev);
    OFX_HIP(hipGetLastError());
    return 0;
}

int ofx_mlp_set_forward_variant(ofx_mlp* n, int variant) {
    if (!n || variant < 0 || variant > 1) { set_error("mlp_set_forward_variant: variant 0 or 1"); return OFX_EINVAL; }
    n->fwd_variant = variant;
    return 0;
}

int ofx_mlp_gradients(ofx_mlp* n, const float* x, const float* t, int nsamp, float* loss, float* grads) {
    if (!n || !x || !t || nsamp < 1 || nsamp > MLP_MAXN) { set_error("mlp_gradients: bad arguments"); return OFX_EINVAL; }
    OFX_HIP(hipSetDevice(n->device));
    OFX_TRY(load_batch(n, x, t, nsamp, false));
    int Mp = (nsamp + 63) / 64 * 64;
    OFX_TRY(forward(n, Mp, true));
    OFX_TRY(backward(n, nsamp, Mp));
    if (loss) OFX_HIP(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    OFX_HIP(hipStreamSynchronize(n->stream));
    if (grads) OFX_TRY(mlp_copy(n, grads, true, false));
    return 0;
}

static int train_step(ofx_mlp* n, int nsamp, float lr) {
    int Mp = (nsamp + 63) / 64 * 64;
    OFX_TRY(forward(n, Mp, true));
    OFX_TRY(backward(n, nsamp, Mp));
    ++n->step;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f / (1.0f - std::pow(b1, (float)n->step));
    float c2 = 1.0f / (1.0f - std::pow(b2, (float)n->step));
    for (int l = 0; l < NLAYER; ++l) {
        size_t nw = (size_t)n->kin[l] * n->kout[l];
        k_adam<<<(unsigned)((nw + 255) / 256), 256, 0, n->stream>>>(n->w[l], n->gw[l], n->mw[l], n->vw[l], n->wb[l], nw,
                                                                   lr, b1, b2, eps, c1, c2);
        k_adam<<<(n->kout[l] + 255) / 256, 256, 0, n->stream>>>(n->b[l], n->gb[l], n->mb[l], n->vb[l], nullptr,
                                                               (size_t)n->kout[l], lr, b1, b2, eps, c1, c2);
    }
    OFX_HIP(hipGetLastError());
    return refresh_wt(n);
}

int ofx_mlp_train(ofx_mlp* n, const float* x, const float* t, int nsamp, int batch, int steps, float lr,
                  unsigned long long seed, float* loss_out) {
    if (!n || !x || !t || nsamp < 1 || batch < 64 || batch > MLP_MAXN || steps < 1) {
        set_error("mlp_train: bad arguments");
        return OFX_EINVAL;
    }
    OFX_HIP(hipSetDevice(n->device));
    int bsz = std::min(batch, nsamp);
    OFX_TRY(mlp_reserve(n, bsz));
    // training set, epoch permutation and per-step losses live on the device
    float *dx = nullptr, *dt = nullptr, *dl = nullptr;
    int* dperm = nullptr;
    int rc = 0;
    if (hipMalloc(&dx, sizeof(float) * (size_t)nsamp * NIN) != hipSuccess ||
        hipMalloc(&dt, sizeof(float) * (size_t)nsamp * NOUT) != hipSuccess ||
        hipMalloc(&dperm, sizeof(int) * (size_t)nsamp) != hipSuccess ||
        hipMalloc(&dl, sizeof(float) * (size_t)steps) != hipSuccess) {
        set_error("mlp_train: out of device memory");
        rc = OFX_ENOMEM;
    }
    if (!rc && (hipMemcpyAsync(dx, x, sizeof(float) * (size_t)nsamp * NIN, hipMemcpyHostToDevice, n->stream) != hipSuccess ||
                hipMemcpyAsync(dt, t, sizeof(float) * (size_t)nsamp * NOUT, hipMemcpyHostToDevice, n->stream) != hipSuccess))
        rc = OFX_EHIP;
    std::vector<int> perm(nsamp);
    for (int i = 0; i < nsamp; ++i) perm[i] = i;
    uint64_t s = seed * 2862933555777941757ull + 3037000493ull;
    int pos = nsamp;
    int Mp = (bsz + 63) / 64 * 64;
    for (int st = 0; st < steps && !rc; ++st) {
        if (pos + bsz > nsamp) {  // new epoch: reshuffle (Fisher-Yates) and upload
            for (int i = nsamp - 1; i > 0; --i) {
                s = s * 6364136223846793005ull + 1442695040888963407ull;
                int j = (int)((s >> 33) % (uint64_t)(i + 1));
                std::swap(perm[i], perm[j]);
            }
            if (hipStreamSynchronize(n->stream) != hipSuccess ||
                hipMemcpy(dperm, perm.data(), sizeof(int) * (size_t)nsamp, hipMemcpyHostToDevice) != hipSuccess) {
                rc = OFX_EHIP;
                break;
            }
            pos = 0;
        }
        k_gather_batch<<<(bsz + 255) / 256, 256, 0, n->stream>>>(dx, dt, dperm + pos, bsz, n->xin, n->tgt);
        k_pack_inputs<<<(Mp * NIN_P + 255) / 256, 256, 0, n->stream>>>(n->xin, bsz, Mp, n->h[0]);
        if (hipGetLastError() != hipSuccess) { rc = OFX_EHIP; break; }
        pos += bsz;
        rc = train_step(n, bsz, lr);
        if (!rc) k_copy1<<<1, 64, 0, n->stream>>>(n->loss, dl + st);
    }
    if (!rc && hipStreamSynchronize(n->stream) != hipSuccess) rc = OFX_EHIP;
    if (!rc && loss_out && hipMemcpy(loss_out, dl, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost) != hipSuccess)
        rc = OFX_EHIP;
    if (rc == OFX_EHIP) set_error("mlp_train: HIP error");
    (void)hipStreamSynchronize(n->stream);
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (dperm) (void)hipFree(dperm);
    if (dl) (void)hipFree(dl);
    return rc;
}

}  // extern "C"
