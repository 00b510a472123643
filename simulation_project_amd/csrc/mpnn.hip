// mpnn.hip -- forward, backward and Adam of the PPE message-passing network on gfx950.
//
// Dense layers: k_nn_gemm computes C = act([A1 | A2] B + c) for 64 rows per block (one 16-row
// tile per wave) with v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation.  B is staged
// transposed in LDS so every lane reads its four consecutive K values with one 8-byte load.
// Weight gradients (sums over rows) come from k_nn_wgrad: each block reduces a chunk of rows with
// the same MFMA and writes a partial; k_nn_sum_partials adds the partials in block order, so the
// gradients are deterministic.
// Message aggregation (SpMM over the SELL matrix, width 32): half a wave per row, four
// neighbours in flight, 8-byte BF16 loads; features: half a wave per row over its entries.
// Loss and gradient of the outputs use FP64 SELL matrix-vector products.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "mpnn.h"

namespace mpsr {

typedef short nn_s4 __attribute__((ext_vector_type(4)));
typedef float nn_f4 __attribute__((ext_vector_type(4)));

// BF16 <-> FP32: round to nearest even (NaN kept quiet), widening is exact
__host__ __device__ inline uint16_t f2bf(float f) {
    union { float f; uint32_t u; } c;
    c.f = f;
    uint32_t u = c.u;
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    union { uint32_t u; float f; } c;
    c.u = (uint32_t)h << 16;
    return c.f;
}

static constexpr int NB = 256;   // block of the row kernels
static constexpr int WG_CHUNK = 512;

struct LayerDims {
    int F, FO, FOP;   // input width, output width, output width padded to 16
};
static const LayerDims kDims[3] = {{NN_F, NN_H, NN_H}, {NN_H, NN_H, NN_H}, {NN_H, NN_OUT, 16}};

struct ParamOffsets {
    int U[3], W[3], V[3], c[3], total;
};
static ParamOffsets param_offsets() {
    ParamOffsets o{};
    int k = 0;
    for (int l = 0; l < 3; ++l) {
        const LayerDims& L = kDims[l];
        o.U[l] = k; k += L.F * NN_M;
        o.W[l] = k; k += L.F * L.FO;
        o.V[l] = k; k += NN_M * L.FO;
        o.c[l] = k; k += L.FO;
    }
    o.total = k;
    return o;
}

struct Mpnn {
    int cap = 0;
    ParamOffsets off{};
    float* w = nullptr;      // FP32 master parameters (flat)
    float* g = nullptr;      // gradients (flat)
    float* am = nullptr;     // Adam moments
    float* av = nullptr;
    int adam_t = 0;
    int has_w = 0;
    // packed BF16 operands, per layer
    uint16_t* Bz[3] = {};    // [FOP][F + 32]   forward z
    uint16_t* Bu[3] = {};    // [32][F]         forward u
    uint16_t* Bdm[3] = {};   // [32][FOP]       dM = dZ V^T
    uint16_t* Bdh[3] = {};   // [F][FOP + 32]   dH = dZ W^T + dU U^T
    float* bias[3] = {};     // [FOP] (zero padded)
    // activations (row-major, rows = cap)
    uint16_t* X = nullptr;   // [cap][16]
    uint16_t* H[3] = {};     // H[0] = X, H[1], H[2]: [cap][128]
    uint16_t* U = nullptr;   // [cap][32] message before aggregation
    uint16_t* Mg[3] = {};    // [cap][32] aggregated messages per layer
    float* Y = nullptr;      // [cap][16] outputs (first two used)
    // backward
    uint16_t* dZ = nullptr;  // [cap][128]
    uint16_t* dZ2 = nullptr; // [cap][128]
    float* dMf = nullptr;    // [cap][32]
    uint16_t* dU = nullptr;  // [cap][32]
    double* vec[11] = {};    // FP64 row vectors for the loss
    float* part = nullptr;   // weight-gradient partials
    size_t part_cap = 0;
    double* dpart = nullptr; // reduction partials
    double* dscal = nullptr; // [16] device scalars: 0 s, 1 lx, 2 lr, 3 rr, 4 ee, 5 n
    double* hscal = nullptr; // pinned
};

// ---------------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------------
__device__ inline double nn_block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0;
    if (w == 0) {
        r = lane < (int)(blockDim.x >> 6) ? lds[lane] : 0.0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}

// sum K interleaved partial vectors part[b * K + k] over nb blocks into dscal[slot + k]
__global__ __launch_bounds__(1024) void k_nn_finish(const double* __restrict__ part, int nb, int K, double* dscal,
                                                    int slot) {
    __shared__ double lds[16];
    for (int k = 0; k < K; ++k) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[(size_t)b * K + k];
        v = nn_block_sum(v, lds);
        if (threadIdx.x == 0) dscal[slot + k] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(NB) void k_nn_jac(int N, const int* __restrict__ dir, const double* __restrict__ rhs,
                                               const double* __restrict__ diag, double* part) {
    __shared__ double lds[NB / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, c = 0.0;
    if (t < N && !dir[t]) {
        const double j = rhs[t] / diag[t];
        a = j * j;
        c = 1.0;
    }
    a = nn_block_sum(a, lds);
    c = nn_block_sum(c, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = c;
    }
}
// s = sqrt(mean) + 1e-30 into dscal[0], the active row count into dscal[5]
__global__ void k_nn_scale(double* dscal) {
    const double n = dscal[7];
    dscal[0] = (n > 0 ? std::sqrt(dscal[6] / n) : 0.0) + 1e-30;
    dscal[5] = n;
}

__device__ inline double hw_sum(double v) {
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 32);
    return v;
}

// half a wave per row: lane h takes the row's stored entries h, h + 32, ...; the eight neighbour
// moments are summed over the half-wave, lane 0 writes the row's 16 features
template <typename R, typename VA>
__global__ __launch_bounds__(NB) void k_nn_feat(MpnnSys sys, const VA* __restrict__ vals, MpnnGeo<R> geo,
                                                const double* __restrict__ dscal, uint16_t* X) {
    const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const int h = threadIdx.x & 31;
    if (t >= sys.N) return;
    float f[NN_F];
    for (int k = 0; k < NN_F; ++k) f[k] = 0.f;
    const bool live = !sys.dir[t];
    double sa = 0, sj = 0, sx = 0, rx = 0, ry = 0, rz = 0, rd = 0, dv = 0;
    int n = 0;
    R4<R> P{}, V{};
    if (live) {
        const double dg = sys.diag[t];
        const int sp = sys.pos[t], sb = sys.slice_off[sp >> 6];
        n = sys.cnt[t];
        P = geo.spos[t];
        V = geo.svel[t];
        for (int k = h; k < n; k += 32) {
            const int e = sell_at(sb, sp & 63, k);
            const int j = sys.cols[e];
            const double a = -(double)vals[e] / dg;
            sa += a;
            sj += a * (sys.rhs[j] / sys.diag[j]);
            sx += a * sys.xprev[j];
            const R4<R> Q = geo.spos[j], W = geo.svel[j];
            const double dx = (double)Q.x - (double)P.x, dy = (double)Q.y - (double)P.y, dz = (double)Q.z - (double)P.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            rx += a * dx; ry += a * dy; rz += a * dz;
            rd += a * std::sqrt(r2);
            if (r2 > 0)
                dv += a * (((double)W.x - (double)V.x) * dx + ((double)W.y - (double)V.y) * dy +
                           ((double)W.z - (double)V.z) * dz) / r2;
        }
    }
    sa = hw_sum(sa); sj = hw_sum(sj); sx = hw_sum(sx); rx = hw_sum(rx);
    ry = hw_sum(ry); rz = hw_sum(rz); rd = hw_sum(rd); dv = hw_sum(dv);
    if (h != 0) return;
    if (live) {
        const double s = dscal[0];
        const double dg = sys.diag[t];
        const double il0 = 1.0 / geo.l0, vs = geo.dt / geo.l0;
        f[0] = (float)(sys.rhs[t] / dg / s);
        f[1] = (float)(sys.xprev[t] / s);
        f[2] = (float)(1.0 - sa);
        f[3] = (float)(n / 128.0);
        f[4] = (float)(sj / s);
        f[5] = (float)(sx / s);
        f[6] = (float)(rx * il0);
        f[7] = (float)(ry * il0);
        f[8] = (float)(rz * il0);
        f[9] = (float)(rd * il0);
        f[10] = (float)((double)V.x * vs);
        f[11] = (float)((double)V.y * vs);
        f[12] = (float)((double)V.z * vs);
        f[13] = (float)(dv * geo.dt);
        const double wd = (double)geo.stl_mag[geo.perm[t]] / geo.re;
        f[14] = (float)(wd >= 0.0 && wd < 1.0 ? wd : 1.0);
        f[15] = 1.f;
    }
    uint16_t* o = X + (size_t)t * NN_F;
    for (int k = 0; k < NN_F; ++k) o[k] = f2bf(f[k]);
}

// ---------------------------------------------------------------------------------------------
// dense layers on MFMA
// ---------------------------------------------------------------------------------------------
// C[rows][fo_store] = act([A1 | A2] B + bias); Bt = B transposed, [16 NT][K1 + K2] BF16.
// Rows with dir set, and (with mask) entries where mask <= 0, are written as 0.
template <int NT>
__global__ __launch_bounds__(256) void k_nn_gemm(int rows, const uint16_t* __restrict__ A1, int lda1, int K1,
                                                  const uint16_t* __restrict__ A2, int lda2, int K2,
                                                  const uint16_t* __restrict__ Bt, const float* __restrict__ bias,
                                                  int relu, const int* __restrict__ dir,
                                                  const uint16_t* __restrict__ mask, int ldm, int fo_store,
                                                  uint16_t* Cb, float* Cf, int ldc) {
    extern __shared__ __align__(16) uint16_t sB[];
    const int K = K1 + K2;
    const int FO = 16 * NT;
    for (int i = threadIdx.x; i < FO * K; i += blockDim.x) sB[i] = Bt[i];
    __syncthreads();
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    const int r0 = blockIdx.x * 64 + w * 16;
    const int ar = r0 + (lane & 15), kq = 4 * (lane >> 4);
    nn_f4 acc[NT];
    for (int j = 0; j < NT; ++j) acc[j] = nn_f4{0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < K; kk += 16) {
        nn_s4 a = {0, 0, 0, 0};
        if (ar < rows) {
            const uint16_t* src = kk < K1 ? A1 + (size_t)ar * lda1 + kk + kq : A2 + (size_t)ar * lda2 + (kk - K1) + kq;
            const uint2 v = *reinterpret_cast<const uint2*>(src);
            a = nn_s4{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
        }
        for (int j = 0; j < NT; ++j) {
            const uint16_t* bp = sB + (size_t)(16 * j + (lane & 15)) * K + kk + kq;
            const uint2 v = *reinterpret_cast<const uint2*>(bp);
            const nn_s4 b = {(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
    }
    for (int j = 0; j < NT; ++j) {
        const int n = 16 * j + (lane & 15);
        if (n >= fo_store) continue;
        for (int i = 0; i < 4; ++i) {
            const int row = r0 + 4 * (lane >> 4) + i;
            if (row >= rows) continue;
            float v = acc[j][i] + (bias ? bias[n] : 0.f);
            if (relu && v < 0.f) v = 0.f;
            if (mask && bf2f(mask[(size_t)row * ldm + n]) <= 0.f) v = 0.f;
            if (dir && dir[row]) v = 0.f;
            if (Cb) Cb[(size_t)row * ldc + n] = f2bf(v);
            if (Cf) Cf[(size_t)row * ldc + n] = v;
        }
    }
}

// The same product on gfx950's v_mfma_f32_16x16x32_bf16 (32 K values per instruction).  A block
// of four waves stages B^T once in LDS and then walks 64-row tiles (grid-stride): each tile's
// [A1 | A2] rows come in with 16-byte coalesced loads, zero-padded to a multiple of 32 in K, every
// LDS row 8 BF16 longer than that so the 16 rows a wave reads at once fall in distinct banks.
// Lane l holds row l % 16 and K values 8 (l / 16) .. + 7 of its A and B operands; the accumulator
// holds rows 4 (l / 16) .. + 3 of column l % 16.  BF16 results go back through the wave's own LDS
// rows (whose stride also fits a row of 16 NT results) and leave in 16-byte stores.  LDS: 16 NT
// (Kp + 8) + 64 max(Kp + 8, 16 NT + 8) BF16, 63 KB at NT = 8, K = 160.
static constexpr int MM_ROWS = 64;    // rows per tile (4 waves of 16)
static constexpr int MM_BLOCKS = 1024;
typedef __bf16 nn_bf8 __attribute__((ext_vector_type(8)));

template <int NT>
__global__ __launch_bounds__(256) void k_nn_mm(int rows, const uint16_t* __restrict__ A1, int lda1, int K1,
                                                const uint16_t* __restrict__ A2, int lda2, int K2,
                                                const uint16_t* __restrict__ Bt, const float* __restrict__ bias,
                                                int relu, const int* __restrict__ dir,
                                                const uint16_t* __restrict__ mask, int ldm, int fo_store,
                                                uint16_t* Cb, float* Cf, int ldc) {
    extern __shared__ __align__(16) uint16_t lds[];
    const int K = K1 + K2;
    const int Kp = (K + 31) & ~31;     // K padded to the instruction's 32
    const int ST = Kp + 8;             // LDS row stride of B^T (BF16)
    const int FO = 16 * NT;
    const int SA = ST > FO + 8 ? ST : FO + 8;   // row stride of the A tile, which also holds a row of results
    uint16_t* sB = lds;                // [FO][ST]
    uint16_t* sA = lds + FO * ST;      // [MM_ROWS][SA]
    const int kc = Kp / 8;
    // B^T: rows of K BF16, zero padding to Kp (8-BF16 chunks; K is a multiple of 8)
    for (int i = threadIdx.x; i < FO * kc; i += blockDim.x) {
        const int n = i / kc, c = i % kc;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (8 * c < K) v = *reinterpret_cast<const uint4*>(Bt + (size_t)n * K + 8 * c);
        *reinterpret_cast<uint4*>(sB + n * ST + 8 * c) = v;
    }
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    const int lr = lane & 15, kg = 8 * (lane >> 4), rq = 4 * (lane >> 4);
    const int tiles = (rows + MM_ROWS - 1) / MM_ROWS;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int row0 = tile * MM_ROWS;
        __syncthreads();   // the last tile's results have left sA
        for (int i = threadIdx.x; i < MM_ROWS * kc; i += blockDim.x) {
            const int r = i / kc, k = 8 * (i % kc);
            const int row = row0 + r;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (row < rows && k < K)
                v = k < K1 ? *reinterpret_cast<const uint4*>(A1 + (size_t)row * lda1 + k)
                           : *reinterpret_cast<const uint4*>(A2 + (size_t)row * lda2 + (k - K1));
            *reinterpret_cast<uint4*>(sA + r * SA + k) = v;
        }
        __syncthreads();
        nn_f4 acc[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = nn_f4{0.f, 0.f, 0.f, 0.f};
        const uint16_t* arow = sA + (16 * w + lr) * SA + kg;
        for (int kk = 0; kk < Kp; kk += 32) {
            const nn_bf8 a = *reinterpret_cast<const nn_bf8*>(arow + kk);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const nn_bf8 b = *reinterpret_cast<const nn_bf8*>(sB + (16 * j + lr) * ST + kg + kk);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
            }
        }
        auto value = [&](int j, int i, int& row, int& n) -> float {
            n = 16 * j + lr;
            row = row0 + 16 * w + rq + i;
            float v = acc[j][i] + (bias ? bias[n] : 0.f);
            if (relu && v < 0.f) v = 0.f;
            if (row < rows) {
                if (mask && bf2f(mask[(size_t)row * ldm + n]) <= 0.f) v = 0.f;
                if (dir && dir[row]) v = 0.f;
            }
            return v;
        };
        if (Cf) {
#pragma unroll
            for (int j = 0; j < NT; ++j)
                for (int i = 0; i < 4; ++i) {
                    int row, n;
                    const float v = value(j, i, row, n);
                    if (row < rows && n < fo_store) Cf[(size_t)row * ldc + n] = v;
                }
        }
        if (Cb) {
            // the wave's 16 rows of sA are its own: they take its BF16 results
            uint16_t* out = sA + (16 * w) * SA;
#pragma unroll
            for (int j = 0; j < NT; ++j)
                for (int i = 0; i < 4; ++i) {
                    int row, n;
                    const float v = value(j, i, row, n);
                    out[(rq + i) * SA + n] = f2bf(v);
                }
            // the 16-bit writes above reach LDS before the 16-byte reads below (a barrier also for
            // the compiler, which could otherwise move loads of another type ahead of the stores)
            __syncthreads();
            // 16 rows of fo_store BF16 (a multiple of 8) in 16-byte chunks
            const int cpr = fo_store / 8;
            for (int i = lane; i < 16 * cpr; i += 64) {
                const int r = i / cpr, c = i % cpr;
                const int row = row0 + 16 * w + r;
                if (row < rows)
                    *reinterpret_cast<uint4*>(Cb + (size_t)row * ldc + 8 * c) =
                        *reinterpret_cast<const uint4*>(out + r * SA + 8 * c);
            }
        }
    }
}

// part[blk][Ka][16 NT] = sum over the block's rows of [A1 | A2]^T D   (Ka = K1 + K2)
template <int NT>
__global__ __launch_bounds__(256) void k_nn_wgrad(int rows, const uint16_t* __restrict__ A1, int lda1, int K1,
                                                   const uint16_t* __restrict__ A2, int lda2, int K2,
                                                   const uint16_t* __restrict__ D, int ldd, float* part) {
    const int Ka = K1 + K2, FO = 16 * NT;
    const int mtiles = Ka / 16;
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    const int row0 = blockIdx.x * WG_CHUNK;
    const int row1 = min(rows, row0 + WG_CHUNK);
    float* out = part + (size_t)blockIdx.x * Ka * FO;
    // wave w owns m-tiles w, w + 4, ... (at most 3 for Ka <= 160)
    nn_f4 acc[3][NT];
    for (int q = 0; q < 3; ++q)
        for (int j = 0; j < NT; ++j) acc[q][j] = nn_f4{0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4);
    for (int kk = row0; kk < row1; kk += 16) {
        nn_s4 b[NT];
        for (int j = 0; j < NT; ++j) {
            const int n = 16 * j + (lane & 15);
            short v[4];
            for (int i = 0; i < 4; ++i) {
                const int r = kk + kq + i;
                v[i] = r < row1 ? (short)D[(size_t)r * ldd + n] : (short)0;
            }
            b[j] = nn_s4{v[0], v[1], v[2], v[3]};
        }
        for (int q = 0; q < 3; ++q) {
            const int mt = w + 4 * q;
            if (mt >= mtiles) break;
            const int m = mt * 16 + (lane & 15);
            short v[4];
            for (int i = 0; i < 4; ++i) {
                const int r = kk + kq + i;
                uint16_t x = 0;
                if (r < row1) x = m < K1 ? A1[(size_t)r * lda1 + m] : A2[(size_t)r * lda2 + (m - K1)];
                v[i] = (short)x;
            }
            const nn_s4 a = {v[0], v[1], v[2], v[3]};
            for (int j = 0; j < NT; ++j) acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[j], acc[q][j], 0, 0, 0);
        }
    }
    for (int q = 0; q < 3; ++q) {
        const int mt = w + 4 * q;
        if (mt >= mtiles) break;
        for (int j = 0; j < NT; ++j) {
            const int n = 16 * j + (lane & 15);
            for (int i = 0; i < 4; ++i) {
                const int m = mt * 16 + 4 * (lane >> 4) + i;
                out[(size_t)m * FO + n] = acc[q][j][i];
            }
        }
    }
}

// g[dst_row * dst_ld + col] = sum over blocks of part[blk][row][col] for rows [0, nrows), cols
// [0, ncols); part rows are FO wide
__global__ __launch_bounds__(256) void k_nn_sum_partials(const float* __restrict__ part, int nblk, int Ka, int FO,
                                                          int row_lo, int nrows, int ncols, float* dst, int dst_ld) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nrows * ncols) return;
    const int r = idx / ncols, c = idx % ncols;
    float acc = 0.f;
    for (int b = 0; b < nblk; ++b) acc += part[((size_t)b * Ka + row_lo + r) * FO + c];
    dst[r * dst_ld + c] = acc;
}

// column sums of D (bf16 [rows][ldd]) for cols [0, ncols): one block per column, fixed order
__global__ __launch_bounds__(256) void k_nn_colsum(int rows, const uint16_t* __restrict__ D, int ldd, int ncols,
                                                   float* dst) {
    __shared__ double lds[4];
    const int c = blockIdx.x;
    if (c >= ncols) return;
    double a = 0.0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) a += (double)bf2f(D[(size_t)r * ldd + c]);
    a = nn_block_sum(a, lds);
    if (threadIdx.x == 0) dst[c] = (float)a;
}

// ---------------------------------------------------------------------------------------------
// aggregation over the PPE graph (width 32): half a wave per row
// mode 0: out_i = sum_j (-A_ij / A_ii) in_j       mode 1: out_j = sum_i (-A_ji / A_ii) in_i
// ---------------------------------------------------------------------------------------------
template <typename VA, typename TI>
__global__ __launch_bounds__(256) void k_nn_agg(MpnnSys sys, const VA* __restrict__ vals, const TI* __restrict__ in,
                                                int mode, uint16_t* out) {
    // half a wave per row: 4 groups of 8 lanes take every 4th stored entry, lane f of a group
    // owns features 4f..4f+3 (one 8-byte load of 4 BF16 per neighbour); the groups' partial sums
    // meet through two shuffles at the end
    const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const int h = threadIdx.x & 31, g = h >> 3, f = h & 7;
    if (t >= sys.N) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    const bool live = !sys.dir[t];
    if (live) {
        const int P = sys.pos[t], sb = sys.slice_off[P >> 6], n = sys.cnt[t];
        const double dg = sys.diag[t];
        for (int q = g; q < n; q += 4) {
            const int e = sell_at(sb, P & 63, q);
            const int j = sys.cols[e];
            const float a = (float)(-(double)vals[e] / (mode == 0 ? dg : sys.diag[j]));
            float x0, x1, x2, x3;
            if constexpr (sizeof(TI) == 2) {
                const uint2 v = *reinterpret_cast<const uint2*>(in + (size_t)j * NN_M + 4 * f);
                x0 = bf2f((uint16_t)(v.x & 0xffff)); x1 = bf2f((uint16_t)(v.x >> 16));
                x2 = bf2f((uint16_t)(v.y & 0xffff)); x3 = bf2f((uint16_t)(v.y >> 16));
            } else {
                const float4 v = *reinterpret_cast<const float4*>(in + (size_t)j * NN_M + 4 * f);
                x0 = v.x; x1 = v.y; x2 = v.z; x3 = v.w;
            }
            a0 += a * x0; a1 += a * x1; a2 += a * x2; a3 += a * x3;
        }
    }
    for (int o = 8; o <= 16; o <<= 1) {
        a0 += __shfl_xor(a0, o, 32);
        a1 += __shfl_xor(a1, o, 32);
        a2 += __shfl_xor(a2, o, 32);
        a3 += __shfl_xor(a3, o, 32);
    }
    if (g == 0) {
        uint2 v;
        v.x = (uint32_t)f2bf(a0) | ((uint32_t)f2bf(a1) << 16);
        v.y = (uint32_t)f2bf(a2) | ((uint32_t)f2bf(a3) << 16);
        *reinterpret_cast<uint2*>(out + (size_t)t * NN_M + 4 * f) = v;
    }
}

// ---------------------------------------------------------------------------------------------
// FP64 matrix-vector product over the active rows: y = A (scale .* x), scale optional
// ---------------------------------------------------------------------------------------------
template <typename VA>
__global__ __launch_bounds__(NB) void k_nn_mv(MpnnSys sys, const VA* __restrict__ vals, const double* __restrict__ x,
                                              const double* __restrict__ scale, double* y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sys.N) return;
    if (sys.dir[t]) {
        y[t] = 0.0;
        return;
    }
    const int P = sys.pos[t], sb = sys.slice_off[P >> 6], n = sys.cnt[t];
    double a = sys.diag[t] * (scale ? scale[t] * x[t] : x[t]);
    for (int q = 0; q < n; ++q) {
        const int e = sell_at(sb, P & 63, q);
        const int j = sys.cols[e];
        a += (double)vals[e] * (scale ? scale[j] * x[j] : x[j]);
    }
    y[t] = a;
}

__global__ __launch_bounds__(NB) void k_nn_y0(MpnnSys sys, const float* __restrict__ Y, double* y0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < sys.N) y0[t] = sys.dir[t] ? 0.0 : (double)Y[(size_t)t * 16];
}

// step 1 of the loss: y0, tgt, rt, r and partial sums (lx, lr, rr)
__global__ __launch_bounds__(NB) void k_nn_loss1(MpnnSys sys, const float* __restrict__ Y, const double* __restrict__ Ay0,
                                                 const double* __restrict__ dscal, double* y0, double* rt, double* r,
                                                 double* part) {
    __shared__ double lds[NB / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0, b = 0, c = 0;
    if (t < sys.N) {
        if (!sys.dir[t]) {
            const double s = dscal[0], dg = sys.diag[t];
            const double v = (double)Y[(size_t)t * 16];
            const double tg = sys.xsol[t] / s;
            const double rtv = sys.rhs[t] / (dg * s) - Ay0[t] / dg;
            const double rv = sys.rhs[t] / s - Ay0[t];
            y0[t] = v;
            rt[t] = rtv;
            r[t] = rv;
            a = (v - tg) * (v - tg);
            b = rtv * rtv;
            c = rv * rv;
        } else {
            y0[t] = 0.0; rt[t] = 0.0; r[t] = 0.0;
        }
    }
    a = nn_block_sum(a, lds);
    b = nn_block_sum(b, lds);
    c = nn_block_sum(c, lds);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = a;
        part[3 * blockIdx.x + 1] = b;
        part[3 * blockIdx.x + 2] = c;
    }
}
// step 2: q = A r / D, e = r - omega q, w1 = omega e / D, w2 = rt / D; partial e.e
__global__ __launch_bounds__(NB) void k_nn_loss2(MpnnSys sys, const float* __restrict__ Y, const double* __restrict__ Ar,
                                                 const double* __restrict__ r, const double* __restrict__ rt,
                                                 double* q, double* e, double* w1, double* w2, double* part) {
    __shared__ double lds[NB / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0;
    if (t < sys.N) {
        if (!sys.dir[t]) {
            const double dg = sys.diag[t];
            const double om = 2.0 / (1.0 + std::exp(-(double)Y[(size_t)t * 16 + 1]));
            const double qv = Ar[t] / dg;
            const double ev = r[t] - om * qv;
            q[t] = qv;
            e[t] = ev;
            w1[t] = om * ev / dg;
            w2[t] = rt[t] / dg;
            a = ev * ev;
        } else {
            q[t] = 0.0; e[t] = 0.0; w1[t] = 0.0; w2[t] = 0.0;
        }
    }
    a = nn_block_sum(a, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// step 3: dr = (2/n) ((e - A w1) / rr - ee / rr^2 r)      (dscal: 5 n, 1 lx, 2 lr, 3 rr, 4 ee)
__global__ __launch_bounds__(NB) void k_nn_loss3(MpnnSys sys, const double* __restrict__ Aw1, const double* __restrict__ e,
                                                 const double* __restrict__ r, const double* __restrict__ dscal, double* dr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sys.N) return;
    if (sys.dir[t]) {
        dr[t] = 0.0;
        return;
    }
    const double n = dscal[5], rr = dscal[3] / n + 1e-300, ee = dscal[4] / n;
    dr[t] = (2.0 / n) * ((e[t] - Aw1[t]) / rr - ee / (rr * rr) * r[t]);
}
// step 4: dY (bf16, [rows][16], columns 2.. zero)
__global__ __launch_bounds__(NB) void k_nn_loss4(MpnnSys sys, const float* __restrict__ Y, const double* __restrict__ y0,
                                                 const double* __restrict__ Aw2, const double* __restrict__ Adr,
                                                 const double* __restrict__ q, const double* __restrict__ r,
                                                 const double* __restrict__ dscal, double beta, uint16_t* dY) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sys.N) return;
    float d0 = 0.f, d1 = 0.f;
    if (!sys.dir[t]) {
        const double n = dscal[5], s = dscal[0], rr = dscal[3] / n + 1e-300;
        const double tg = sys.xsol[t] / s;
        const double g0 = 2.0 * (y0[t] - tg) / n - beta * (2.0 / n) * Aw2[t] - Adr[t];
        const double om = 2.0 / (1.0 + std::exp(-(double)Y[(size_t)t * 16 + 1]));
        const double dom = 2.0 * (om * q[t] - r[t]) * q[t] / (n * rr);
        d0 = (float)g0;
        d1 = (float)(dom * om * (1.0 - om / 2.0));
    }
    uint16_t* o = dY + (size_t)t * 16;
    o[0] = f2bf(d0);
    o[1] = f2bf(d1);
    for (int k = 2; k < 16; ++k) o[k] = 0;
}

__global__ __launch_bounds__(NB) void k_nn_apply(MpnnSys sys, const float* __restrict__ Y,
                                                 const double* __restrict__ dscal, double* x0, double* omega) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sys.N) return;
    if (sys.dir[t]) {
        if (x0) x0[t] = 0.0;
        if (omega) omega[t] = 1.0;
        return;
    }
    if (x0) x0[t] = dscal[0] * (double)Y[(size_t)t * 16];
    if (omega) omega[t] = fmin(1.5, fmax(0.5, 2.0 / (1.0 + std::exp(-(double)Y[(size_t)t * 16 + 1]))));
}

// ---------------------------------------------------------------------------------------------
// parameters: packing to BF16 operands, Adam
// ---------------------------------------------------------------------------------------------
// Bz[n][k] = (k < F ? W[k][n] : V[k-F][n]) for n < FO, 0 for FO <= n < FOP
// Bu[n][k] = U[k][n];  Bdm[n][k] = V[n][k] (k < FO);  Bdh[n][k] = (k < FOP ? W[n][k] : U[n][k-FOP])
__global__ __launch_bounds__(256) void k_nn_pack(const float* __restrict__ w, int F, int FO, int FOP, int oU, int oW,
                                                 int oV, int oc, uint16_t* Bz, uint16_t* Bu, uint16_t* Bdm,
                                                 uint16_t* Bdh, float* bias) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nz = FOP * (F + NN_M), nu = NN_M * F, ndm = NN_M * FOP, ndh = F * (FOP + NN_M);
    if (i < nz) {
        const int n = i / (F + NN_M), k = i % (F + NN_M);
        float v = 0.f;
        if (n < FO) v = k < F ? w[oW + k * FO + n] : w[oV + (k - F) * FO + n];
        Bz[i] = f2bf(v);
    }
    if (i < nu) {
        const int n = i / F, k = i % F;
        Bu[i] = f2bf(w[oU + k * NN_M + n]);
    }
    if (i < ndm) {
        const int n = i / FOP, k = i % FOP;
        Bdm[i] = f2bf(k < FO ? w[oV + n * FO + k] : 0.f);
    }
    if (i < ndh) {
        const int n = i / (FOP + NN_M), k = i % (FOP + NN_M);
        float v = 0.f;
        if (k < FOP) v = k < FO ? w[oW + n * FO + k] : 0.f;
        else v = w[oU + n * NN_M + (k - FOP)];
        Bdh[i] = f2bf(v);
    }
    if (i < FOP) bias[i] = i < FO ? w[oc + i] : 0.f;
}

__global__ __launch_bounds__(256) void k_nn_adam(int n, float* w, const float* __restrict__ g, float* m, float* v,
                                                 float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (std::sqrt(vi / c2) + eps);
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
template <typename T> static bool nn_alloc(T** p, size_t n) {
    return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess;
}

Mpnn* mpnn_create(int cap_rows) {
    Mpnn* n = new Mpnn();
    n->cap = ((std::max(cap_rows, 1) + 63) / 64) * 64;
    n->off = param_offsets();
    const size_t C = n->cap;
    bool ok = nn_alloc(&n->w, n->off.total) && nn_alloc(&n->g, n->off.total) && nn_alloc(&n->am, n->off.total) &&
              nn_alloc(&n->av, n->off.total);
    for (int l = 0; l < 3; ++l) {
        const LayerDims& L = kDims[l];
        ok = ok && nn_alloc(&n->Bz[l], (size_t)L.FOP * (L.F + NN_M)) && nn_alloc(&n->Bu[l], (size_t)NN_M * L.F) &&
             nn_alloc(&n->Bdm[l], (size_t)NN_M * L.FOP) && nn_alloc(&n->Bdh[l], (size_t)L.F * (L.FOP + NN_M)) &&
             nn_alloc(&n->bias[l], L.FOP) && nn_alloc(&n->Mg[l], C * NN_M);
    }
    ok = ok && nn_alloc(&n->X, C * NN_F) && nn_alloc(&n->H[1], C * NN_H) && nn_alloc(&n->H[2], C * NN_H) &&
         nn_alloc(&n->U, C * NN_M) && nn_alloc(&n->Y, C * 16) && nn_alloc(&n->dZ, C * NN_H) &&
         nn_alloc(&n->dZ2, C * NN_H) && nn_alloc(&n->dMf, C * NN_M) && nn_alloc(&n->dU, C * NN_M);
    for (auto& v : n->vec) ok = ok && nn_alloc(&v, C);
    n->part_cap = ((C + WG_CHUNK - 1) / WG_CHUNK) * (size_t)(NN_H + NN_M) * NN_H;
    ok = ok && nn_alloc(&n->part, n->part_cap) && nn_alloc(&n->dpart, 3 * (C / NB + 1)) && nn_alloc(&n->dscal, 16);
    ok = ok && hipHostMalloc((void**)&n->hscal, 16 * sizeof(double)) == hipSuccess;
    n->H[0] = n->X;
    if (!ok) {
        mpnn_destroy(n);
        return nullptr;
    }
    (void)hipMemset(n->am, 0, sizeof(float) * n->off.total);
    (void)hipMemset(n->av, 0, sizeof(float) * n->off.total);
    (void)hipMemset(n->g, 0, sizeof(float) * n->off.total);
    (void)hipMemset(n->dscal, 0, sizeof(double) * 16);
    return n;
}

void mpnn_destroy(Mpnn* n) {
    if (!n) return;
    void* ps[] = {n->w, n->g, n->am, n->av, n->X, n->H[1], n->H[2], n->U, n->Y, n->dZ, n->dZ2, n->dMf, n->dU,
                  n->part, n->dpart, n->dscal};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (int l = 0; l < 3; ++l) {
        void* q[] = {n->Bz[l], n->Bu[l], n->Bdm[l], n->Bdh[l], n->bias[l], n->Mg[l]};
        for (void* p : q)
            if (p) (void)hipFree(p);
    }
    for (double* v : n->vec)
        if (v) (void)hipFree(v);
    if (n->hscal) (void)hipHostFree(n->hscal);
    delete n;
}

int mpnn_param_count() { return param_offsets().total; }
bool mpnn_has_weights(const Mpnn* n) { return n && n->has_w; }

static void nn_pack(Mpnn* n, hipStream_t st) {
    for (int l = 0; l < 3; ++l) {
        const LayerDims& L = kDims[l];
        const int cnt = std::max({L.FOP * (L.F + NN_M), NN_M * L.F, NN_M * L.FOP, L.F * (L.FOP + NN_M), L.FOP});
        hipLaunchKernelGGL(k_nn_pack, dim3(grid_for(cnt, 256)), dim3(256), 0, st, n->w, L.F, L.FO, L.FOP,
                           n->off.U[l], n->off.W[l], n->off.V[l], n->off.c[l], n->Bz[l], n->Bu[l], n->Bdm[l],
                           n->Bdh[l], n->bias[l]);
    }
}

mpsr_status mpnn_set_weights(Mpnn* n, const float* w, hipStream_t st) {
    if (hipMemcpyAsync(n->w, w, sizeof(float) * n->off.total, hipMemcpyHostToDevice, st) != hipSuccess)
        return MPSR_GPUERR;
    nn_pack(n, st);
    (void)hipMemsetAsync(n->am, 0, sizeof(float) * n->off.total, st);
    (void)hipMemsetAsync(n->av, 0, sizeof(float) * n->off.total, st);
    n->adam_t = 0;
    n->has_w = 1;
    return hipStreamSynchronize(st) == hipSuccess ? MPSR_NOERR : MPSR_GPUERR;
}
mpsr_status mpnn_get_weights(Mpnn* n, float* w, hipStream_t st) {
    if (hipMemcpyAsync(w, n->w, sizeof(float) * n->off.total, hipMemcpyDeviceToHost, st) != hipSuccess)
        return MPSR_GPUERR;
    return hipStreamSynchronize(st) == hipSuccess ? MPSR_NOERR : MPSR_GPUERR;
}
mpsr_status mpnn_get_grads(Mpnn* n, float* g, hipStream_t st) {
    if (hipMemcpyAsync(g, n->g, sizeof(float) * n->off.total, hipMemcpyDeviceToHost, st) != hipSuccess)
        return MPSR_GPUERR;
    return hipStreamSynchronize(st) == hipSuccess ? MPSR_NOERR : MPSR_GPUERR;
}

template <typename R>
void mpnn_features(Mpnn* n, const MpnnSys& sys, const MpnnGeo<R>& geo, hipStream_t st) {
    const int N = sys.N;
    const int nb = (int)grid_for(N, NB);
    hipLaunchKernelGGL(k_nn_jac, dim3(std::max(nb, 1)), dim3(NB), 0, st, N, sys.dir, sys.rhs, sys.diag, n->dpart);
    hipLaunchKernelGGL(k_nn_finish, dim3(1), dim3(1024), 0, st, n->dpart, std::max(nb, 1), 2, n->dscal, 6);
    hipLaunchKernelGGL(k_nn_scale, dim3(1), dim3(1), 0, st, n->dscal);
    if (N <= 0) return;
    const dim3 gf(grid_for((long long)N * 32, NB));
    if (sys.v32)
        hipLaunchKernelGGL((k_nn_feat<R, float>), gf, dim3(NB), 0, st, sys, sys.v32, geo, n->dscal, n->X);
    else
        hipLaunchKernelGGL((k_nn_feat<R, double>), gf, dim3(NB), 0, st, sys, sys.v64, geo, n->dscal, n->X);
}

template <int NT>
static void gemm(int rows, const uint16_t* A1, int lda1, int K1, const uint16_t* A2, int lda2, int K2,
                 const uint16_t* Bt, const float* bias, int relu, const int* dir, const uint16_t* mask, int ldm,
                 int fo_store, uint16_t* Cb, float* Cf, int ldc, hipStream_t st) {
    if (rows <= 0) return;
    // every operand width here is a multiple of 8 BF16 (16 bytes), as the 16-byte staging needs
    const int Kp = ((K1 + K2) + 31) & ~31;
    const int ST = Kp + 8, SA = std::max(ST, 16 * NT + 8);
    const size_t lds = ((size_t)16 * NT * ST + (size_t)MM_ROWS * SA) * sizeof(uint16_t);
    const int tiles = (int)grid_for(rows, MM_ROWS);
    hipLaunchKernelGGL(k_nn_mm<NT>, dim3(std::min(tiles, MM_BLOCKS)), dim3(256), lds, st, rows, A1, lda1, K1, A2,
                       lda2, K2, Bt, bias, relu, dir, mask, ldm, fo_store, Cb, Cf, ldc);
}

static void agg(Mpnn* n, const MpnnSys& sys, const void* in, bool in_bf16, int mode, uint16_t* out, hipStream_t st) {
    if (sys.N <= 0) return;
    const dim3 grid(grid_for((long long)sys.N * 32, 256));
    if (sys.v32) {
        if (in_bf16) hipLaunchKernelGGL((k_nn_agg<float, uint16_t>), grid, dim3(256), 0, st, sys, sys.v32, (const uint16_t*)in, mode, out);
        else hipLaunchKernelGGL((k_nn_agg<float, float>), grid, dim3(256), 0, st, sys, sys.v32, (const float*)in, mode, out);
    } else {
        if (in_bf16) hipLaunchKernelGGL((k_nn_agg<double, uint16_t>), grid, dim3(256), 0, st, sys, sys.v64, (const uint16_t*)in, mode, out);
        else hipLaunchKernelGGL((k_nn_agg<double, float>), grid, dim3(256), 0, st, sys, sys.v64, (const float*)in, mode, out);
    }
    (void)n;
}

static void mv(const MpnnSys& sys, const double* x, const double* scale, double* y, hipStream_t st) {
    if (sys.N <= 0) return;
    const dim3 grid(grid_for(sys.N, NB));
    if (sys.v32) hipLaunchKernelGGL((k_nn_mv<float>), grid, dim3(NB), 0, st, sys, sys.v32, x, scale, y);
    else hipLaunchKernelGGL((k_nn_mv<double>), grid, dim3(NB), 0, st, sys, sys.v64, x, scale, y);
}

void mpnn_forward(Mpnn* n, const MpnnSys& sys, hipStream_t st) {
    const int N = sys.N;
    for (int l = 0; l < 3; ++l) {
        const LayerDims& L = kDims[l];
        const uint16_t* h = n->H[l];
        // u = h U (32 wide), m = Agg(u)
        gemm<2>(N, h, L.F, L.F, nullptr, 0, 0, n->Bu[l], nullptr, 0, sys.dir, nullptr, 0, NN_M, n->U, nullptr, NN_M, st);
        agg(n, sys, n->U, true, 0, n->Mg[l], st);
        // z = [h | m] [W; V] + c
        if (l < 2)
            gemm<8>(N, h, L.F, L.F, n->Mg[l], NN_M, NN_M, n->Bz[l], n->bias[l], 1, sys.dir, nullptr, 0, NN_H,
                    n->H[l + 1], nullptr, NN_H, st);
        else
            gemm<1>(N, h, L.F, L.F, n->Mg[l], NN_M, NN_M, n->Bz[l], n->bias[l], 0, sys.dir, nullptr, 0, 16, nullptr,
                    n->Y, 16, st);
    }
}

void mpnn_apply(Mpnn* n, const MpnnSys& sys, double* x0, double* omega, hipStream_t st) {
    if (sys.N <= 0) return;
    hipLaunchKernelGGL(k_nn_apply, dim3(grid_for(sys.N, NB)), dim3(NB), 0, st, sys, n->Y, n->dscal, x0, omega);
}

// weight gradient of one operand pair into the flat gradient vector
template <int NT>
static void wgrad(Mpnn* n, int rows, const uint16_t* A1, int lda1, int K1, const uint16_t* A2, int lda2, int K2,
                  const uint16_t* D, int ldd, int row_lo, int nrows, int ncols, float* dst, int dst_ld, hipStream_t st) {
    const int nblk = std::max(1, (int)((rows + WG_CHUNK - 1) / WG_CHUNK));
    const int Ka = K1 + K2;
    hipLaunchKernelGGL(k_nn_wgrad<NT>, dim3(nblk), dim3(256), 0, st, rows, A1, lda1, K1, A2, lda2, K2, D, ldd, n->part);
    hipLaunchKernelGGL(k_nn_sum_partials, dim3(grid_for(nrows * ncols, 256)), dim3(256), 0, st, n->part, nblk, Ka,
                       16 * NT, row_lo, nrows, ncols, dst, dst_ld);
}

mpsr_status mpnn_loss_grad(Mpnn* n, const MpnnSys& sys, double beta, double* parts, hipStream_t st) {
    const int N = sys.N;
    if (N <= 0 || !sys.xsol) return MPSR_PARAMERR;
    const int nb = (int)grid_for(N, NB);
    double *y0 = n->vec[0], *Av = n->vec[1], *rt = n->vec[2], *r = n->vec[3], *q = n->vec[4], *e = n->vec[5],
           *w1 = n->vec[6], *w2 = n->vec[7], *dr = n->vec[8], *Aw2 = n->vec[9], *Adr = n->vec[10];
    // ---- loss and dL/dY ----
    hipLaunchKernelGGL(k_nn_y0, dim3(nb), dim3(NB), 0, st, sys, n->Y, y0);
    mv(sys, y0, nullptr, Av, st);
    hipLaunchKernelGGL(k_nn_loss1, dim3(nb), dim3(NB), 0, st, sys, n->Y, Av, n->dscal, y0, rt, r, n->dpart);
    hipLaunchKernelGGL(k_nn_finish, dim3(1), dim3(1024), 0, st, n->dpart, nb, 3, n->dscal, 1);
    mv(sys, r, nullptr, Av, st);
    hipLaunchKernelGGL(k_nn_loss2, dim3(nb), dim3(NB), 0, st, sys, n->Y, Av, r, rt, q, e, w1, w2, n->dpart);
    hipLaunchKernelGGL(k_nn_finish, dim3(1), dim3(1024), 0, st, n->dpart, nb, 1, n->dscal, 4);
    mv(sys, w1, nullptr, Av, st);
    hipLaunchKernelGGL(k_nn_loss3, dim3(nb), dim3(NB), 0, st, sys, Av, e, r, n->dscal, dr);
    mv(sys, w2, nullptr, Aw2, st);
    mv(sys, dr, nullptr, Adr, st);
    uint16_t* dY = n->dZ2;  // [rows][16] view of the spare gradient buffer
    hipLaunchKernelGGL(k_nn_loss4, dim3(nb), dim3(NB), 0, st, sys, n->Y, y0, Aw2, Adr, q, r, n->dscal, beta, dY);
    // ---- backward through the layers ----
    const ParamOffsets& o = n->off;
    // layer 3 (H[2] -> Y): dZ = dY (16 wide)
    wgrad<1>(n, N, n->H[2], NN_H, NN_H, n->Mg[2], NN_M, NN_M, dY, 16, 0, NN_H, NN_OUT, n->g + o.W[2], NN_OUT, st);
    wgrad<1>(n, N, n->H[2], NN_H, NN_H, n->Mg[2], NN_M, NN_M, dY, 16, NN_H, NN_M, NN_OUT, n->g + o.V[2], NN_OUT, st);
    hipLaunchKernelGGL(k_nn_colsum, dim3(NN_OUT), dim3(256), 0, st, N, dY, 16, NN_OUT, n->g + o.c[2]);
    gemm<2>(N, dY, 16, 16, nullptr, 0, 0, n->Bdm[2], nullptr, 0, sys.dir, nullptr, 0, NN_M, n->U, nullptr, NN_M, st);
    agg(n, sys, n->U, true, 1, n->dU, st);
    wgrad<2>(n, N, n->H[2], NN_H, NN_H, nullptr, 0, 0, n->dU, NN_M, 0, NN_H, NN_M, n->g + o.U[2], NN_M, st);
    gemm<8>(N, dY, 16, 16, n->dU, NN_M, NN_M, n->Bdh[2], nullptr, 0, sys.dir, n->H[2], NN_H, NN_H, n->dZ, nullptr,
            NN_H, st);
    // layer 2 (H[1] -> H[2]): dZ
    wgrad<8>(n, N, n->H[1], NN_H, NN_H, n->Mg[1], NN_M, NN_M, n->dZ, NN_H, 0, NN_H, NN_H, n->g + o.W[1], NN_H, st);
    wgrad<8>(n, N, n->H[1], NN_H, NN_H, n->Mg[1], NN_M, NN_M, n->dZ, NN_H, NN_H, NN_M, NN_H, n->g + o.V[1], NN_H, st);
    hipLaunchKernelGGL(k_nn_colsum, dim3(NN_H), dim3(256), 0, st, N, n->dZ, NN_H, NN_H, n->g + o.c[1]);
    gemm<2>(N, n->dZ, NN_H, NN_H, nullptr, 0, 0, n->Bdm[1], nullptr, 0, sys.dir, nullptr, 0, NN_M, n->U, nullptr, NN_M,
            st);
    agg(n, sys, n->U, true, 1, n->dU, st);
    wgrad<2>(n, N, n->H[1], NN_H, NN_H, nullptr, 0, 0, n->dU, NN_M, 0, NN_H, NN_M, n->g + o.U[1], NN_M, st);
    gemm<8>(N, n->dZ, NN_H, NN_H, n->dU, NN_M, NN_M, n->Bdh[1], nullptr, 0, sys.dir, n->H[1], NN_H, NN_H, n->dZ2,
            nullptr, NN_H, st);
    // layer 1 (X -> H[1]): dZ2
    wgrad<8>(n, N, n->X, NN_F, NN_F, n->Mg[0], NN_M, NN_M, n->dZ2, NN_H, 0, NN_F, NN_H, n->g + o.W[0], NN_H, st);
    wgrad<8>(n, N, n->X, NN_F, NN_F, n->Mg[0], NN_M, NN_M, n->dZ2, NN_H, NN_F, NN_M, NN_H, n->g + o.V[0], NN_H, st);
    hipLaunchKernelGGL(k_nn_colsum, dim3(NN_H), dim3(256), 0, st, N, n->dZ2, NN_H, NN_H, n->g + o.c[0]);
    gemm<2>(N, n->dZ2, NN_H, NN_H, nullptr, 0, 0, n->Bdm[0], nullptr, 0, sys.dir, nullptr, 0, NN_M, n->U, nullptr,
            NN_M, st);
    agg(n, sys, n->U, true, 1, n->dU, st);
    wgrad<2>(n, N, n->X, NN_F, NN_F, nullptr, 0, 0, n->dU, NN_M, 0, NN_F, NN_M, n->g + o.U[0], NN_M, st);
    // ---- loss values ----
    if (hipMemcpyAsync(n->hscal, n->dscal, sizeof(double) * 16, hipMemcpyDeviceToHost, st) != hipSuccess)
        return MPSR_GPUERR;
    if (hipStreamSynchronize(st) != hipSuccess) return MPSR_GPUERR;
    const double cnt = n->hscal[5] > 0 ? n->hscal[5] : 1.0;
    const double lx = n->hscal[1] / cnt, lr = n->hscal[2] / cnt, rr = n->hscal[3] / cnt + 1e-300;
    const double lw = (n->hscal[4] / cnt) / rr;
    if (parts) {
        parts[0] = lx + beta * lr + lw;
        parts[1] = lx;
        parts[2] = lr;
        parts[3] = lw;
    }
    return MPSR_NOERR;
}

void mpnn_adam(Mpnn* n, float lr, hipStream_t st) {
    n->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_nn_adam, dim3(grid_for(n->off.total, 256)), dim3(256), 0, st, n->off.total, n->w, n->g, n->am,
                       n->av, lr, b1, b2, 1e-8f, c1, c2);
    nn_pack(n, st);
}

mpsr_status mpnn_get_rows(Mpnn* n, int N, float* X, float* Y, double* s, hipStream_t st) {
    std::vector<uint16_t> xb((size_t)std::max(N, 1) * NN_F);
    std::vector<float> yb((size_t)std::max(N, 1) * 16);
    if (N > 0) {
        if (hipMemcpyAsync(xb.data(), n->X, sizeof(uint16_t) * N * NN_F, hipMemcpyDeviceToHost, st) != hipSuccess)
            return MPSR_GPUERR;
        if (hipMemcpyAsync(yb.data(), n->Y, sizeof(float) * N * 16, hipMemcpyDeviceToHost, st) != hipSuccess)
            return MPSR_GPUERR;
    }
    if (hipMemcpyAsync(n->hscal, n->dscal, sizeof(double) * 16, hipMemcpyDeviceToHost, st) != hipSuccess)
        return MPSR_GPUERR;
    if (hipStreamSynchronize(st) != hipSuccess) return MPSR_GPUERR;
    for (int t = 0; t < N; ++t) {
        if (X)
            for (int k = 0; k < NN_F; ++k) X[(size_t)t * NN_F + k] = bf2f(xb[(size_t)t * NN_F + k]);
        if (Y) {
            Y[2 * (size_t)t] = yb[(size_t)t * 16];
            Y[2 * (size_t)t + 1] = yb[(size_t)t * 16 + 1];
        }
    }
    if (s) *s = n->hscal[0];
    return MPSR_NOERR;
}

template void mpnn_features<float>(Mpnn*, const MpnnSys&, const MpnnGeo<float>&, hipStream_t);
template void mpnn_features<double>(Mpnn*, const MpnnSys&, const MpnnGeo<double>&, hipStream_t);

}  // namespace mpsr
