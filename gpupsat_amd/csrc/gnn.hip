// gnn.hip -- the learned branching network: message passing over the literal-clause
// incidence graph, trained and run on the GPU (no reference counterpart).
//
// Nodes: 2n literal nodes and m clause nodes per instance, instances back to back;
// an edge per clause occurrence.  Hidden size 64, R rounds, weights shared by rounds:
//   H_l^0 = occ(l)/8 u_l + v_l                       H_c^0 = len(c)/4 u_c + v_c
//   M_c = 1/4 sum_{l in c} H_l     H_c' = relu([H_c | M_c] W_c^T + b_c)
//   M_l = 1/8 sum_{c ni l} H_c'    H_l' = relu([H_l | M_l | H_{not l}] W_l^T + b_l)
//   Z_v = [H_{2v+1} | H_{2v}]      R_v = relu(Z_v W_o^T + b_o)   Y_v = R_v W_y^T + b_y
// Y_v = (branching score, phase logit).  The three 64-wide products run on MFMA
// (v_mfma_f32_16x16x32_bf16): BF16 weights and BF16 inputs, FP32 accumulation.
// Aggregations, biases, the 2-output head, the loss and Adam are FP32.  Gradients and the
// loss are summed from per-wavefront (or per-workgroup) partials in a fixed order, never
// with floating-point atomics, so they are reproducible bit for bit.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gpsat.h"
#include "common.hpp"
#include "gpsat_internal.hpp"

namespace gpsat {
namespace gnn {

constexpr int HID = 64;
typedef uint16_t bf16_t;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// parameter layout (FP32 master copy), in this order
constexpr int O_UL = 0, O_VL = 64, O_UC = 128, O_VC = 192;
constexpr int O_WC = 256, O_BC = O_WC + 64 * 128, O_WL = O_BC + 64, O_BL = O_WL + 64 * 192;
constexpr int O_WO = O_BL + 64, O_BO = O_WO + 64 * 128, O_WY = O_BO + 64, O_BY = O_WY + 2 * 64;
constexpr int N_PARAMS = O_BY + 2;

__device__ __forceinline__ bf16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------------------
// MFMA tiles.  v_mfma_f32_16x16x32_bf16 (gfx950) on a wave of 64: lane l supplies the 8
// consecutive-k values A[l%16][8(l/16) .. 8(l/16)+7] and B[8(l/16) .. 8(l/16)+7][l%16] as one
// 16-byte read each, and receives D[4(l/16)+j][l%16] (j = 0..3).  LDS tiles are stored
// with the k index contiguous (transposed where the source is not), rows padded by 16 bytes
// so the 16 rows a wave reads at once fall in different banks.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p)); }
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// out[N x 64] = relu(X[N x K] W[64 x K]^T + b); one wave per 16 rows, W staged in LDS.
template <int K>
__global__ __launch_bounds__(256) void k_dense_fwd(const bf16_t* __restrict__ X, const bf16_t* __restrict__ Wq,
                                                   const float* __restrict__ bias, float* __restrict__ out, int N)
{
    constexpr int SK = K + 8;
    __shared__ __align__(16) bf16_t sW[64 * SK];
    for (int i = threadIdx.x; i < 64 * K / 8; i += blockDim.x) {
        const int r = (8 * i) / K, c = (8 * i) % K;
        *reinterpret_cast<uint4*>(sW + r * SK + c) = reinterpret_cast<const uint4*>(Wq)[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + wave) * 16;
    if (r0 >= N) return;
    const int ar = r0 + (lane & 15), kq = 8 * (lane >> 4);
    f32x4 acc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ct++) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
        if (ar < N) a = ld8(X + (size_t)ar * K + k0 + kq);
#pragma unroll
        for (int ct = 0; ct < 4; ct++) acc[ct] = mfma32(a, ld8(sW + (ct * 16 + (lane & 15)) * SK + k0 + kq), acc[ct]);
    }
    const int dq = 4 * (lane >> 4);
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = r0 + dq + j;
            if (row < N) out[(size_t)row * 64 + col] = fmaxf(acc[ct][j] + bias[col], 0.0f);
        }
    }
}

// dX[N x K] = dP[N x 64] W[64 x K]; W staged transposed (k = hidden index contiguous).
template <int K>
__global__ __launch_bounds__(256) void k_dense_bwd_dx(const bf16_t* __restrict__ dP, const bf16_t* __restrict__ Wq,
                                                      float* __restrict__ dX, int N)
{
    constexpr int SH = 64 + 8;
    __shared__ __align__(16) bf16_t sWt[K * SH];
    for (int i = threadIdx.x; i < 64 * K; i += blockDim.x) {
        const int h = i / K, c = i % K;
        sWt[c * SH + h] = Wq[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + wave) * 16;
    if (r0 >= N) return;
    const int ar = r0 + (lane & 15), kq = 8 * (lane >> 4);
    constexpr int CT = K / 16;
    f32x4 acc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int h0 = 0; h0 < 64; h0 += 32) {
        bf16x8 a = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
        if (ar < N) a = ld8(dP + (size_t)ar * 64 + h0 + kq);
#pragma unroll
        for (int ct = 0; ct < CT; ct++) acc[ct] = mfma32(a, ld8(sWt + (ct * 16 + (lane & 15)) * SH + h0 + kq), acc[ct]);
    }
    const int dq = 4 * (lane >> 4);
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = r0 + dq + j;
            if (row < N) dX[(size_t)row * K + col] = acc[ct][j];
        }
    }
}

// gfx950's transposing LDS read: within each 16-lane group, lane i receives element i % 4 of
// the 8-byte chunks that lanes i/4, i/4 + 4, i/4 + 8 and i/4 + 12 address.  With lane L of
// the group pointing at row L/4, columns 4(L%4) .. 4(L%4)+3 of a row-major 4 x 16 tile, lane
// i receives column i of those 4 rows.  Two reads (rows r .. r+3, r+4 .. r+7) give the 8
// consecutive-k values of an MFMA operand from a tile stored with k as its row index.
typedef short v4s __attribute__((ext_vector_type(4)));
__device__ __forceinline__ bf16x8 ld8_tr(const bf16_t* p, int row_stride)
{
    typedef __attribute__((address_space(3))) v4s lds_v4s;
    const v4s r0 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s*)p);
    const v4s r1 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s*)(p + 4 * row_stride));
    typedef short v8s __attribute__((ext_vector_type(8)));
    const v8s v = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    return __builtin_bit_cast(bf16x8, v);
}

// Partial weight gradient of this workgroup: part[blockIdx.x][64 x K] = sum over its
// 64-node chunks of dP[N x 64]^T X[N x K] (the reduction adds the workgroups' partials in
// order).  A chunk is staged row-major in LDS (node index = row, 16-byte stores, rows padded
// by 32 bytes) by all four wavefronts, and the MFMA operands, whose k index is the node, come
// out of it through transposing reads; wavefront w computes output rows h = 16w .. 16w+15.
template <int K>
__global__ __launch_bounds__(256) void k_dense_bwd_dw(const bf16_t* __restrict__ dP, const bf16_t* __restrict__ X,
                                                      float* __restrict__ part, int N)
{
    constexpr int SP = 64 + 16, SX = K + 16;   // row strides (elements) of the staged tiles
    __shared__ __align__(16) bf16_t sP[64 * SP];
    __shared__ __align__(16) bf16_t sX[64 * SX];
    constexpr int CT = K / 16;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // this lane's place in a transposing read: row 8(lane/16) + (lane%16)/4 of the k block,
    // columns 4(lane%4) .. +3 of the 16-column tile
    const int tr_row = 8 * (lane >> 4) + ((lane & 15) >> 2), tr_col = 4 * (lane & 3);
    f32x4 acc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int n0 = blockIdx.x * 64; n0 < N; n0 += gridDim.x * 64) {
        const int rows = min(64, N - n0);
        __syncthreads();
        for (int q = t; q < 64 * 8; q += 256) {
            const int r = q >> 3, c = (q & 7) * 8;
            uint4 v = {0u, 0u, 0u, 0u};
            if (r < rows) v = *reinterpret_cast<const uint4*>(dP + (size_t)(n0 + r) * 64 + c);
            *reinterpret_cast<uint4*>(sP + r * SP + c) = v;
        }
        for (int q = t; q < 64 * (K / 8); q += 256) {
            const int r = q / (K / 8), c = (q % (K / 8)) * 8;
            uint4 v = {0u, 0u, 0u, 0u};
            if (r < rows) v = *reinterpret_cast<const uint4*>(X + (size_t)(n0 + r) * K + c);
            *reinterpret_cast<uint4*>(sX + r * SX + c) = v;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 64; s += 32) {
            // A[h][n] = dP[n][h] (16 h x 32 n), B[n][col] = X[n][col] (32 n x 16 col)
            const bf16x8 a = ld8_tr(sP + (s + tr_row) * SP + 16 * wave + tr_col, SP);
#pragma unroll
            for (int ct = 0; ct < CT; ct++)
                acc[ct] = mfma32(a, ld8_tr(sX + (s + tr_row) * SX + 16 * ct + tr_col, SX), acc[ct]);
        }
    }
    float* out = part + (size_t)blockIdx.x * 64 * K;
    const int dq = 4 * (lane >> 4);
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; j++) out[(wave * 16 + dq + j) * K + col] = acc[ct][j];
    }
}

// Fixed-order reductions of per-slot partials: out[i] (+)= the sum over slots of
// part[slot * width + i].  Every gradient and the loss are summed this way, never with
// floating-point atomics, so a gradient is the same bit for bit on every run.
// Wide (weight matrices, few slots): one thread per element adds its slots in order.
template <bool ACCUMULATE>
__global__ void k_reduce_wide(const float* __restrict__ part, int slots, int width, float* __restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < width; i += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int k = 0; k < slots; k++) s += part[(size_t)k * width + i];
        out[i] = ACCUMULATE ? out[i] + s : s;
    }
}

// Wide with many slots (weight-gradient partials of up to 1024 workgroups): first each
// of `groups` workgroup rows adds its own contiguous run of slots, four columns per
// thread, into part2[group * width + i]; k_reduce_wide then adds the groups in order.
__global__ __launch_bounds__(256) void k_reduce_wide_split(const float* __restrict__ part, int slots, int width,
                                                           float* __restrict__ part2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;  // column quad
    if (4 * q >= width) return;
    const int per = (slots + gridDim.y - 1) / gridDim.y;
    const int k0 = blockIdx.y * per, k1 = min(slots, k0 + per);
    const float4* p = reinterpret_cast<const float4*>(part) + q;
    const size_t stride = (size_t)width / 4;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 8
    for (int k = k0; k < k1; k++) {
        const float4 v = p[(size_t)k * stride];
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    reinterpret_cast<float4*>(part2)[(size_t)blockIdx.y * stride + q] = s;
}

// Narrow (biases, embeddings, the loss; thousands of slots): one workgroup per element,
// each thread adds a fixed stride of slots, then a fixed binary tree in LDS.
template <bool ACCUMULATE>
__global__ __launch_bounds__(256) void k_reduce_tree(const float* __restrict__ part, int slots, int width,
                                                     float* __restrict__ out)
{
    __shared__ float sh[256];
    const int t = threadIdx.x;
    for (int i = blockIdx.x; i < width; i += gridDim.x) {
        float s = 0.f;
        for (int k = t; k < slots; k += 256) s += part[(size_t)k * width + i];
        sh[t] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) sh[t] += sh[t + o];
            __syncthreads();
        }
        if (t == 0) out[i] = ACCUMULATE ? out[i] + sh[0] : sh[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Aggregation, embedding, head and bookkeeping kernels: one wave per node row, lane = dim.

__global__ void k_embed(const int32_t* __restrict__ l_off, const int32_t* __restrict__ c_off, int L, int C,
                        const float* __restrict__ th, float* __restrict__ Hl, float* __restrict__ Hc)
{
    const int64_t total = (int64_t)(L + C) * HID;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int d = (int)(i & 63);
        const int64_t r = i >> 6;
        if (r < L) {
            const float occ = (float)(l_off[r + 1] - l_off[r]);
            Hl[i] = occ * 0.125f * th[O_UL + d] + th[O_VL + d];
        } else {
            const int64_t c = r - L;
            const float len = (float)(c_off[c + 1] - c_off[c]);
            Hc[c * HID + d] = len * 0.25f * th[O_UC + d] + th[O_VC + d];
        }
    }
}

__device__ __forceinline__ int wave_global() { return (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int wave_count() { return (int)((gridDim.x * blockDim.x) >> 6); }

// The gather kernels give each wavefront 4 rows at once, 16 lanes per row, each lane on 4
// consecutive dimensions: a 256-byte row is one float4 load per lane, and a wavefront
// keeps 4 rows' dependent offset -> edge -> neighbour loads in flight together.
__device__ __forceinline__ int row_slot() { return (threadIdx.x & 63) >> 4; }
__device__ __forceinline__ int col4() { return 4 * (threadIdx.x & 15); }
__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void add4(float4& s, const float4& x)
{
    s.x += x.x;
    s.y += x.y;
    s.z += x.z;
    s.w += x.w;
}
__device__ __forceinline__ float4 mul4(float a, const float4& x) { return make_float4(a * x.x, a * x.y, a * x.z, a * x.w); }
__device__ __forceinline__ void st_bf16x4(bf16_t* p, const float4& v)
{
    uint2 w;
    w.x = (uint32_t)to_bf16(v.x) | (uint32_t)to_bf16(v.y) << 16;
    w.y = (uint32_t)to_bf16(v.z) | (uint32_t)to_bf16(v.w) << 16;
    *reinterpret_cast<uint2*>(p) = w;
}

// Sum over edges e0 .. e1-1 of X[edges[e] * stride + col .. col + 3], four neighbour rows in
// flight at a time (a slot past e1 re-reads edge e0 of its quad and is not added), added in
// edge order.
__device__ __forceinline__ float4 gather_sum(const int32_t* __restrict__ edges, int e0, int e1,
                                             const float* __restrict__ X, int stride, int col)
{
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int e = e0; e < e1; e += 4) {
        const int k = e1 - e;
        const int i0 = edges[e], i1 = edges[k > 1 ? e + 1 : e], i2 = edges[k > 2 ? e + 2 : e],
                  i3 = edges[k > 3 ? e + 3 : e];
        const float4 x0 = ld4(X + (size_t)i0 * stride + col), x1 = ld4(X + (size_t)i1 * stride + col);
        const float4 x2 = ld4(X + (size_t)i2 * stride + col), x3 = ld4(X + (size_t)i3 * stride + col);
        add4(s, x0);
        if (k > 1) add4(s, x1);
        if (k > 2) add4(s, x2);
        if (k > 3) add4(s, x3);
    }
    return s;
}

// Xc = bf16([Hc | 1/4 sum_{l in c} Hl])
__global__ void k_gather_clause(const int32_t* __restrict__ c_off, const int32_t* __restrict__ c_edges, int C,
                                const float* __restrict__ Hl, const float* __restrict__ Hc, bf16_t* __restrict__ Xc)
{
    const int d = col4();
    for (int c0 = 4 * wave_global(); c0 < C; c0 += 4 * wave_count()) {
        const int c = c0 + row_slot();
        if (c >= C) continue;
        const float4 s = gather_sum(c_edges, c_off[c], c_off[c + 1], Hl, HID, d);
        st_bf16x4(Xc + (size_t)c * 128 + d, ld4(Hc + (size_t)c * HID + d));
        st_bf16x4(Xc + (size_t)c * 128 + 64 + d, mul4(0.25f, s));
    }
}

// Xl = bf16([Hl | 1/8 sum_{c ni l} Hc' | H_{l^1}])
__global__ void k_gather_lit(const int32_t* __restrict__ l_off, const int32_t* __restrict__ l_edges, int L,
                             const float* __restrict__ Hl, const float* __restrict__ Hc1, bf16_t* __restrict__ Xl)
{
    const int d = col4();
    for (int l0 = 4 * wave_global(); l0 < L; l0 += 4 * wave_count()) {
        const int l = l0 + row_slot();
        if (l >= L) continue;
        const float4 s = gather_sum(l_edges, l_off[l], l_off[l + 1], Hc1, HID, d);
        st_bf16x4(Xl + (size_t)l * 192 + d, ld4(Hl + (size_t)l * HID + d));
        st_bf16x4(Xl + (size_t)l * 192 + 64 + d, mul4(0.125f, s));
        st_bf16x4(Xl + (size_t)l * 192 + 128 + d, ld4(Hl + (size_t)(l ^ 1) * HID + d));
    }
}

// Z = bf16([H_{2v+1} | H_{2v}])
__global__ void k_build_z(int V, const float* __restrict__ Hl, bf16_t* __restrict__ Z)
{
    const int d = threadIdx.x & 63;
    for (int v = wave_global(); v < V; v += wave_count()) {
        Z[(size_t)v * 128 + d] = to_bf16(Hl[(size_t)(2 * v + 1) * HID + d]);
        Z[(size_t)v * 128 + 64 + d] = to_bf16(Hl[(size_t)(2 * v) * HID + d]);
    }
}

// Y = R W_y^T + b_y (W_y in BF16); with labels also the loss and dY.
__global__ void k_head(int V, const float* __restrict__ Rv, const bf16_t* __restrict__ Wyq, const float* __restrict__ th,
                       float* __restrict__ Y, const float* __restrict__ target, const float* __restrict__ phase,
                       const float* __restrict__ mask, float alpha, float* __restrict__ dY, float* __restrict__ loss_part)
{
    const int d = threadIdx.x & 63;
    float lsum = 0.f;
    for (int v = wave_global(); v < V; v += wave_count()) {
        const float r = Rv[(size_t)v * HID + d];
        float y0 = r * from_bf16(Wyq[d]);
        float y1 = r * from_bf16(Wyq[64 + d]);
        for (int s = 32; s > 0; s >>= 1) {
            y0 += __shfl_xor(y0, s);
            y1 += __shfl_xor(y1, s);
        }
        y0 += th[O_BY];
        y1 += th[O_BY + 1];
        if (d == 0) {
            Y[2 * (size_t)v] = y0;
            Y[2 * (size_t)v + 1] = y1;
            if (target) {
                const float t = target[v], p = phase[v], mk = mask[v];
                const float sig = 1.0f / (1.0f + expf(-y1));
                const float bce = fmaxf(y1, 0.f) - y1 * p + log1pf(expf(-fabsf(y1)));
                lsum += ((y0 - t) * (y0 - t) + alpha * mk * bce) / (float)V;
                dY[2 * (size_t)v] = 2.0f * (y0 - t) / (float)V;
                dY[2 * (size_t)v + 1] = alpha * mk * (sig - p) / (float)V;
            }
        }
    }
    // this wavefront's share of the loss (k_reduce sums the shares in order)
    if (target && d == 0) loss_part[wave_global()] = lsum;
}

// dPo = (dY W_y) * [R > 0] -> BF16, plus this wavefront's partial gradients of b_o, W_y
// and b_y, laid out as the parameters are (b_o at O_BO, then W_y, then b_y).
constexpr int HEAD_PART = O_BY + 2 - O_BO;
__global__ void k_head_bwd(int V, const float* __restrict__ Rv, const float* __restrict__ dY,
                           const bf16_t* __restrict__ Wyq, bf16_t* __restrict__ dPo, float* __restrict__ part)
{
    const int d = threadIdx.x & 63;
    const float w0 = from_bf16(Wyq[d]), w1 = from_bf16(Wyq[64 + d]);
    float gw0 = 0.f, gw1 = 0.f, gb = 0.f, gy0 = 0.f, gy1 = 0.f;
    for (int v = wave_global(); v < V; v += wave_count()) {
        const float r = Rv[(size_t)v * HID + d];
        const float a = dY[2 * (size_t)v], b = dY[2 * (size_t)v + 1];
        gw0 += a * r;
        gw1 += b * r;
        const float dp = r > 0.f ? a * w0 + b * w1 : 0.f;
        gb += dp;
        dPo[(size_t)v * HID + d] = to_bf16(dp);
        if (d == 0) {
            gy0 += a;
            gy1 += b;
        }
    }
    float* o = part + (size_t)wave_global() * HEAD_PART;
    o[d] = gb;
    o[O_WY - O_BO + d] = gw0;
    o[O_WY - O_BO + 64 + d] = gw1;
    if (d == 0) {
        o[O_BY - O_BO] = gy0;
        o[O_BY - O_BO + 1] = gy1;
    }
}

// dH_l^R from dZ
__global__ void k_bwd_z(int V, const float* __restrict__ dZ, float* __restrict__ dHl)
{
    const int d = threadIdx.x & 63;
    for (int v = wave_global(); v < V; v += wave_count()) {
        dHl[(size_t)(2 * v + 1) * HID + d] = dZ[(size_t)v * 128 + d];
        dHl[(size_t)(2 * v) * HID + d] = dZ[(size_t)v * 128 + 64 + d];
    }
}

// dP = dH * [H > 0] -> BF16, and this wavefront's partial bias gradient.
__global__ void k_bwd_mask(int N, const float* __restrict__ dH, const float* __restrict__ H, bf16_t* __restrict__ dP,
                           float* __restrict__ part)
{
    const int d = threadIdx.x & 63;
    float gb = 0.f;
    for (int r = wave_global(); r < N; r += wave_count()) {
        const size_t i = (size_t)r * HID + d;
        const float g = H[i] > 0.f ? dH[i] : 0.f;
        gb += g;
        dP[i] = to_bf16(g);
    }
    part[(size_t)wave_global() * HID + d] = gb;
}

// dH_c' = self part (from the next round, stride sstride) + 1/8 sum_{l in c} dXl[l][64 + d];
// then masked by [H_c' > 0] into dPc (BF16) with the bias gradient.  Row layout as in the
// gather kernels; the 4 row slots' bias partials are added (slot 0+1, 2+3, then both)
// before the wavefront writes its partial.
__global__ void k_bwd_clause(const int32_t* __restrict__ c_off, const int32_t* __restrict__ c_edges, int C,
                             const float* __restrict__ self, int sstride, const float* __restrict__ dXl,
                             const float* __restrict__ Hc1, bf16_t* __restrict__ dPc, float* __restrict__ part)
{
    const int d = col4();
    float4 gb = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int c0 = 4 * wave_global(); c0 < C; c0 += 4 * wave_count()) {
        const int c = c0 + row_slot();
        if (c >= C) continue;
        float4 s = self ? ld4(self + (size_t)c * sstride + d) : make_float4(0.f, 0.f, 0.f, 0.f);
        add4(s, mul4(0.125f, gather_sum(c_edges, c_off[c], c_off[c + 1], dXl, 192, 64 + d)));
        const float4 h = ld4(Hc1 + (size_t)c * HID + d);
        const float4 g = make_float4(h.x > 0.f ? s.x : 0.f, h.y > 0.f ? s.y : 0.f, h.z > 0.f ? s.z : 0.f,
                                     h.w > 0.f ? s.w : 0.f);
        add4(gb, g);
        st_bf16x4(dPc + (size_t)c * HID + d, g);
    }
    for (int o = 16; o <= 32; o <<= 1) {
        gb.x += __shfl_xor(gb.x, o);
        gb.y += __shfl_xor(gb.y, o);
        gb.z += __shfl_xor(gb.z, o);
        gb.w += __shfl_xor(gb.w, o);
    }
    if (row_slot() == 0) *reinterpret_cast<float4*>(part + (size_t)wave_global() * HID + d) = gb;
}

// dH_l = dXl[l][d] + dXl[l^1][128 + d] + 1/4 sum_{c ni l} dXc[c][64 + d]
__global__ void k_bwd_lit(const int32_t* __restrict__ l_off, const int32_t* __restrict__ l_edges, int L,
                          const float* __restrict__ dXl, const float* __restrict__ dXc, float* __restrict__ dHl)
{
    const int d = col4();
    for (int l0 = 4 * wave_global(); l0 < L; l0 += 4 * wave_count()) {
        const int l = l0 + row_slot();
        if (l >= L) continue;
        const float4 m = gather_sum(l_edges, l_off[l], l_off[l + 1], dXc, 128, 64 + d);
        float4 h = ld4(dXl + (size_t)l * 192 + d);
        add4(h, ld4(dXl + (size_t)(l ^ 1) * 192 + 128 + d));
        add4(h, mul4(0.25f, m));
        *reinterpret_cast<float4*>(dHl + (size_t)l * HID + d) = h;
    }
}

// this wavefront's partial gradients of the embeddings u_l, v_l, u_c, v_c (parameters
// 0..255) from dH^0 (the self part of dXc for clauses, stride 128)
__global__ void k_bwd_embed(const int32_t* __restrict__ l_off, const int32_t* __restrict__ c_off, int L, int C,
                            const float* __restrict__ dHl, const float* __restrict__ dXc, float* __restrict__ part)
{
    const int d = threadIdx.x & 63;
    float gu = 0.f, gv = 0.f, cu = 0.f, cv = 0.f;
    for (int r = wave_global(); r < L + C; r += wave_count()) {
        if (r < L) {
            const float g = dHl[(size_t)r * HID + d];
            gu += (float)(l_off[r + 1] - l_off[r]) * 0.125f * g;
            gv += g;
        } else {
            const int c = r - L;
            const float g = dXc[(size_t)c * 128 + d];
            cu += (float)(c_off[c + 1] - c_off[c]) * 0.25f * g;
            cv += g;
        }
    }
    float* o = part + (size_t)wave_global() * O_WC;
    o[O_UL + d] = gu;
    o[O_VL + d] = gv;
    o[O_UC + d] = cu;
    o[O_VC + d] = cv;
}

__global__ void k_quantize(const float* __restrict__ th, bf16_t* __restrict__ Wc, bf16_t* __restrict__ Wl,
                           bf16_t* __restrict__ Wo, bf16_t* __restrict__ Wy)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 64 * 192; i += gridDim.x * blockDim.x) {
        if (i < 64 * 128) Wc[i] = to_bf16(th[O_WC + i]);
        Wl[i] = to_bf16(th[O_WL + i]);
        if (i < 64 * 128) Wo[i] = to_bf16(th[O_WO + i]);
        if (i < 128) Wy[i] = to_bf16(th[O_WY + i]);
    }
}

__global__ void k_adam(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

}  // namespace gnn
}  // namespace gpsat

// ---------------------------------------------------------------------------------------
// Host orchestration and C ABI.
using namespace gpsat;
using namespace gpsat::gnn;

// Kernels that leave per-wavefront partials run on a fixed grid of RED_BLOCKS workgroups
// of 4 wavefronts (one partial slot per wavefront; 16 wavefronts per CU to hide the gather
// latency); the weight-gradient products on at most DW_BLOCKS workgroups of 4 wavefronts
// (one partial matrix each).
constexpr int RED_BLOCKS = 4096;
constexpr int RED_WAVES = 4 * RED_BLOCKS;
constexpr int DW_BLOCKS = 1024;
constexpr size_t PART_FLOATS = std::max<size_t>((size_t)DW_BLOCKS * 64 * 192, (size_t)RED_WAVES * O_WC);
// Slot groups of the two-stage weight-gradient reduction, and their partials after
// PART_FLOATS in the same buffer.
constexpr int WIDE_GROUPS = 32;
constexpr size_t PART2_FLOATS = (size_t)WIDE_GROUPS * 64 * 192;

// One batch's activations, carved from the model's arena (see gpsat_gnn_model).
struct Acts {
    std::vector<float*> Hl, Hc;
    std::vector<bf16_t*> Xc, Xl;
    bf16_t *Z = nullptr, *dP = nullptr;
    float *Rv = nullptr, *dY = nullptr, *dXa = nullptr, *dXb = nullptr, *dHl = nullptr, *loss = nullptr;
};

// A model: FP32 master parameters, gradient, Adam moments, BF16 weight copies, the partials
// of the gradient reductions, and a grow-only arena that holds the activations of the graph
// being run.  Graphs own only their incidence structure and labels, so a training loop keeps
// every batch's graph across steps without holding every batch's activations; calls on one
// model therefore run one at a time.
struct gpsat_gnn_model {
    float* th = nullptr;
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* part = nullptr;
    bf16_t *Wc = nullptr, *Wl = nullptr, *Wo = nullptr, *Wy = nullptr;
    char* arena = nullptr;
    size_t arena_bytes = 0;
    int64_t step = 0;
    bool dirty = true;
};

struct gpsat_gnn_graph {
    int32_t V = 0, L = 0, C = 0, R = 3;
    int64_t E = 0;
    int32_t *c_off = nullptr, *c_edges = nullptr, *l_off = nullptr, *l_edges = nullptr;
    float *Y = nullptr, *target = nullptr, *phase = nullptr, *mask = nullptr;
    hipStream_t st = nullptr;
    std::vector<void*> owned;
    ~gpsat_gnn_graph()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : owned) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    template <class T>
    hipError_t alloc(T*& p, size_t n)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) owned.push_back(q);
        p = (T*)q;
        return e;
    }
};

#define GNN_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));          \
            return _e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;    \
        }                                                                          \
    } while (0)

namespace {

int grid_rows(int64_t rows, int waves_per_block = 4)
{
    const int64_t b = (rows + waves_per_block - 1) / waves_per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

// Lay a graph's activations out in `base` (nullptr: only count); returns the bytes used.
// Every piece starts 256-byte aligned.
size_t carve(const gpsat_gnn_graph* g, char* base, Acts& a)
{
    size_t off = 0;
    auto take = [&](auto*& p, size_t count) {
        using T = std::remove_reference_t<decltype(*p)>;
        p = base ? (T*)(base + off) : nullptr;
        off += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    };
    const size_t L = g->L, C = g->C, V = g->V, R = g->R;
    a.Hl.assign(R + 1, nullptr);
    a.Hc.assign(R + 1, nullptr);
    a.Xc.assign(R, nullptr);
    a.Xl.assign(R, nullptr);
    for (size_t t = 0; t <= R; t++) {
        take(a.Hl[t], L * HID);
        take(a.Hc[t], C * HID);
        if (t < R) take(a.Xc[t], C * 128);
        if (t < R) take(a.Xl[t], L * 192);
    }
    const size_t maxn = std::max({L, C, V});
    take(a.Z, V * 128);
    take(a.dP, maxn * HID);
    take(a.Rv, V * HID);
    take(a.dY, V * 2);
    take(a.dXa, std::max(L * 192, V * 128));
    take(a.dXb, C * 128);
    take(a.dHl, L * HID);
    take(a.loss, 1);
    return off;
}

// The activations of graph g in model m's arena (grown if g needs more).
hipError_t acts_for(gpsat_gnn_model* m, gpsat_gnn_graph* g, Acts& a)
{
    const size_t need = carve(g, nullptr, a);
    if (need > m->arena_bytes) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
        (void)hipFree(m->arena);
        m->arena = nullptr;
        m->arena_bytes = 0;
        e = hipMalloc(&m->arena, need);
        if (e != hipSuccess) return e;
        m->arena_bytes = need;
    }
    carve(g, m->arena, a);
    return hipSuccess;
}

template <bool ACC>
void reduce(const float* part, int slots, int width, float* out, hipStream_t st)
{
    if (width >= 4096 && slots > 2 * WIDE_GROUPS && width % 4 == 0 && (size_t)width * WIDE_GROUPS <= PART2_FLOATS) {
        float* part2 = const_cast<float*>(part) + PART_FLOATS;
        hipLaunchKernelGGL(k_reduce_wide_split, dim3((width / 4 + 255) / 256, WIDE_GROUPS), dim3(256), 0, st, part,
                           slots, width, part2);
        hipLaunchKernelGGL(k_reduce_wide<ACC>, dim3((width + 255) / 256), dim3(256), 0, st, part2, WIDE_GROUPS, width,
                           out);
    } else if (width >= 4096)
        hipLaunchKernelGGL(k_reduce_wide<ACC>, dim3((width + 255) / 256), dim3(256), 0, st, part, slots, width, out);
    else
        hipLaunchKernelGGL(k_reduce_tree<ACC>, dim3(width), dim3(256), 0, st, part, slots, width, out);
}

// gW += dP^T X over N nodes (weight gradient of one dense layer)
template <int K>
void dense_dw(gpsat_gnn_model* m, const bf16_t* dP, const bf16_t* X, int N, float* gW, hipStream_t st)
{
    const int nb = std::max(1, std::min(DW_BLOCKS, (N + 63) / 64));
    hipLaunchKernelGGL((k_dense_bwd_dw<K>), dim3(nb), dim3(256), 0, st, dP, X, m->part, N);
    reduce<true>(m->part, nb, 64 * K, gW, st);
}

hipError_t forward_pass(gpsat_gnn_model* m, gpsat_gnn_graph* g, const Acts& a)
{
    hipStream_t st = g->st;
    if (m->dirty) {
        hipLaunchKernelGGL(k_quantize, dim3(48), dim3(256), 0, st, m->th, m->Wc, m->Wl, m->Wo, m->Wy);
        m->dirty = false;
    }
    const int64_t rows = (int64_t)(g->L + g->C) * HID;
    hipLaunchKernelGGL(k_embed, dim3((int)std::min<int64_t>((rows + 255) / 256, 65536)), dim3(256), 0, st, g->l_off,
                       g->c_off, g->L, g->C, m->th, a.Hl[0], a.Hc[0]);
    for (int t = 0; t < g->R; t++) {
        hipLaunchKernelGGL(k_gather_clause, dim3(grid_rows((g->C + 3) / 4)), dim3(256), 0, st, g->c_off, g->c_edges, g->C,
                           a.Hl[t], a.Hc[t], a.Xc[t]);
        hipLaunchKernelGGL((k_dense_fwd<128>), dim3((g->C + 63) / 64), dim3(256), 0, st, a.Xc[t], m->Wc,
                           m->th + O_BC, a.Hc[t + 1], g->C);
        hipLaunchKernelGGL(k_gather_lit, dim3(grid_rows((g->L + 3) / 4)), dim3(256), 0, st, g->l_off, g->l_edges, g->L, a.Hl[t],
                           a.Hc[t + 1], a.Xl[t]);
        hipLaunchKernelGGL((k_dense_fwd<192>), dim3((g->L + 63) / 64), dim3(256), 0, st, a.Xl[t], m->Wl,
                           m->th + O_BL, a.Hl[t + 1], g->L);
    }
    hipLaunchKernelGGL(k_build_z, dim3(grid_rows(g->V)), dim3(256), 0, st, g->V, a.Hl[g->R], a.Z);
    hipLaunchKernelGGL((k_dense_fwd<128>), dim3((g->V + 63) / 64), dim3(256), 0, st, a.Z, m->Wo, m->th + O_BO, a.Rv,
                       g->V);
    return hipGetLastError();
}

hipError_t head(gpsat_gnn_model* m, gpsat_gnn_graph* g, const Acts& a, bool labels, float alpha)
{
    hipLaunchKernelGGL(k_head, dim3(RED_BLOCKS), dim3(256), 0, g->st, g->V, a.Rv, m->Wy, m->th, g->Y,
                       labels ? g->target : nullptr, g->phase, g->mask, alpha, a.dY, m->part);
    if (labels) reduce<false>(m->part, RED_WAVES, 1, a.loss, g->st);
    return hipGetLastError();
}

hipError_t backward_pass(gpsat_gnn_model* m, gpsat_gnn_graph* g, const Acts& a)
{
    hipStream_t st = g->st;
    // head and readout
    hipLaunchKernelGGL(k_head_bwd, dim3(RED_BLOCKS), dim3(256), 0, st, g->V, a.Rv, a.dY, m->Wy, a.dP, m->part);
    reduce<true>(m->part, RED_WAVES, HEAD_PART, m->grad + O_BO, st);
    dense_dw<128>(m, a.dP, a.Z, g->V, m->grad + O_WO, st);
    hipLaunchKernelGGL((k_dense_bwd_dx<128>), dim3((g->V + 63) / 64), dim3(256), 0, st, a.dP, m->Wo, a.dXa, g->V);
    hipLaunchKernelGGL(k_bwd_z, dim3(grid_rows(g->V)), dim3(256), 0, st, g->V, a.dXa, a.dHl);
    // dXa holds dXl (L x 192), dXb holds dXc (C x 128)
    const float* self = nullptr;
    for (int t = g->R - 1; t >= 0; t--) {
        hipLaunchKernelGGL(k_bwd_mask, dim3(RED_BLOCKS), dim3(256), 0, st, g->L, a.dHl, a.Hl[t + 1], a.dP, m->part);
        reduce<true>(m->part, RED_WAVES, HID, m->grad + O_BL, st);
        dense_dw<192>(m, a.dP, a.Xl[t], g->L, m->grad + O_WL, st);
        hipLaunchKernelGGL((k_dense_bwd_dx<192>), dim3((g->L + 63) / 64), dim3(256), 0, st, a.dP, m->Wl, a.dXa, g->L);
        // the self part of dH_c^{t+1} comes from round t+1's clause update (dXc of that round)
        hipLaunchKernelGGL(k_bwd_clause, dim3(RED_BLOCKS), dim3(256), 0, st, g->c_off, g->c_edges, g->C, self, 128,
                           a.dXa, a.Hc[t + 1], a.dP, m->part);
        reduce<true>(m->part, RED_WAVES, HID, m->grad + O_BC, st);
        dense_dw<128>(m, a.dP, a.Xc[t], g->C, m->grad + O_WC, st);
        hipLaunchKernelGGL((k_dense_bwd_dx<128>), dim3((g->C + 63) / 64), dim3(256), 0, st, a.dP, m->Wc, a.dXb, g->C);
        hipLaunchKernelGGL(k_bwd_lit, dim3(grid_rows((g->L + 3) / 4)), dim3(256), 0, st, g->l_off, g->l_edges, g->L, a.dXa,
                           a.dXb, a.dHl);
        self = a.dXb;
    }
    hipLaunchKernelGGL(k_bwd_embed, dim3(RED_BLOCKS), dim3(256), 0, st, g->l_off, g->c_off, g->L, g->C, a.dHl, a.dXb,
                       m->part);
    reduce<true>(m->part, RED_WAVES, O_WC, m->grad, st);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int gpsat_gnn_param_count(int32_t* n)
{
    if (!n) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    *n = N_PARAMS;
    return GPSAT_OK;
}

int gpsat_gnn_model_create(const float* theta, gpsat_gnn_model** out)
{
    if (!out || !theta) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return GPSAT_ERR_NODEVICE;
    }
    gpsat_gnn_model* m = new gpsat_gnn_model();
    hipError_t e = hipMalloc(&m->th, 4 * N_PARAMS);
    if (e == hipSuccess) e = hipMalloc(&m->grad, 4 * N_PARAMS);
    if (e == hipSuccess) e = hipMalloc(&m->part, 4 * (PART_FLOATS + PART2_FLOATS));
    if (e == hipSuccess) e = hipMalloc(&m->m, 4 * N_PARAMS);
    if (e == hipSuccess) e = hipMalloc(&m->v, 4 * N_PARAMS);
    if (e == hipSuccess) e = hipMalloc(&m->Wc, 2 * 64 * 128);
    if (e == hipSuccess) e = hipMalloc(&m->Wl, 2 * 64 * 192);
    if (e == hipSuccess) e = hipMalloc(&m->Wo, 2 * 64 * 128);
    if (e == hipSuccess) e = hipMalloc(&m->Wy, 2 * 128);
    if (e == hipSuccess) e = hipMemcpy(m->th, theta, 4 * N_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m, 0, 4 * N_PARAMS);
    if (e == hipSuccess) e = hipMemset(m->v, 0, 4 * N_PARAMS);
    if (e != hipSuccess) {
        set_error(std::string("gnn model: ") + hipGetErrorString(e));
        gpsat_gnn_model_destroy(m);
        return GPSAT_ERR_HIP;
    }
    *out = m;
    return GPSAT_OK;
}

int gpsat_gnn_model_destroy(gpsat_gnn_model* m)
{
    if (!m) return GPSAT_OK;
    (void)hipDeviceSynchronize();
    for (void* p : {(void*)m->th, (void*)m->grad, (void*)m->m, (void*)m->v, (void*)m->part, (void*)m->arena,
                    (void*)m->Wc, (void*)m->Wl, (void*)m->Wo, (void*)m->Wy})
        (void)hipFree(p);
    delete m;
    return GPSAT_OK;
}

int gpsat_gnn_model_params(gpsat_gnn_model* m, float* theta)
{
    if (!m || !theta) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    GNN_TRY(hipDeviceSynchronize());
    GNN_TRY(hipMemcpy(theta, m->th, 4 * N_PARAMS, hipMemcpyDeviceToHost));
    return GPSAT_OK;
}

int gpsat_gnn_graph_create(gpsat_batch* b, int32_t rounds, gpsat_gnn_graph** out)
{
    if (!b || !out || rounds < 1 || rounds > 16) {
        set_error("bad argument");
        return GPSAT_ERR_ARG;
    }
    *out = nullptr;
    DeviceBatch& d = b->d;
    gpsat_gnn_graph* g = new gpsat_gnn_graph();
    g->R = rounds;
    // CSR both ways over global node ids
    std::vector<int64_t> var_base(d.n_inst + 1, 0);
    for (int32_t i = 0; i < d.n_inst; i++) var_base[i + 1] = var_base[i] + d.h_meta[i].n;
    g->V = (int32_t)var_base[d.n_inst];
    g->L = 2 * g->V;
    g->C = (int32_t)d.n_clauses;
    g->E = d.n_lits;
    std::vector<int32_t> c_off(g->C + 1, 0), c_edges(g->E), l_off(g->L + 1, 0), l_edges(g->E);
    int32_t cg = 0;
    for (int32_t i = 0; i < d.n_inst; i++) {
        const InstMeta& mi = d.h_meta[i];
        const int32_t* co = d.h_coff.data() + mi.coff_base;
        for (int32_t c = 0; c < mi.m; c++, cg++) {
            c_off[cg + 1] = (int32_t)(mi.lit_base + co[c + 1]);
            for (int32_t k = co[c]; k < co[c + 1]; k++) {
                const int32_t ln = (int32_t)(2 * var_base[i] + d.h_lits[mi.lit_base + k]);
                c_edges[mi.lit_base + k] = ln;
                l_off[ln + 1]++;
            }
        }
    }
    for (int32_t l = 0; l < g->L; l++) l_off[l + 1] += l_off[l];
    std::vector<int32_t> fill(l_off.begin(), l_off.end() - 1);
    for (int32_t c = 0; c < g->C; c++)
        for (int32_t e = c_off[c]; e < c_off[c + 1]; e++) l_edges[fill[c_edges[e]]++] = c;
    hipError_t e = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking);
    const size_t L = g->L, C = g->C, V = g->V;
    if (e == hipSuccess) e = g->alloc(g->c_off, C + 1);
    if (e == hipSuccess) e = g->alloc(g->c_edges, g->E);
    if (e == hipSuccess) e = g->alloc(g->l_off, L + 1);
    if (e == hipSuccess) e = g->alloc(g->l_edges, g->E);
    if (e == hipSuccess) e = g->alloc(g->Y, V * 2);
    if (e == hipSuccess) e = g->alloc(g->target, V);
    if (e == hipSuccess) e = g->alloc(g->phase, V);
    if (e == hipSuccess) e = g->alloc(g->mask, V);
    if (e == hipSuccess) e = hipMemcpy(g->c_off, c_off.data(), 4 * (C + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && g->E) e = hipMemcpy(g->c_edges, c_edges.data(), 4 * g->E, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->l_off, l_off.data(), 4 * (L + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && g->E) e = hipMemcpy(g->l_edges, l_edges.data(), 4 * g->E, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("gnn graph: ") + hipGetErrorString(e));
        delete g;
        return e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;
    }
    *out = g;
    return GPSAT_OK;
}

int gpsat_gnn_graph_destroy(gpsat_gnn_graph* g)
{
    delete g;
    return GPSAT_OK;
}

int gpsat_gnn_graph_labels(gpsat_gnn_graph* g, const float* target, const float* phase, const float* mask)
{
    if (!g || !target || !phase || !mask) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    GNN_TRY(hipMemcpyAsync(g->target, target, 4 * (size_t)g->V, hipMemcpyHostToDevice, g->st));
    GNN_TRY(hipMemcpyAsync(g->phase, phase, 4 * (size_t)g->V, hipMemcpyHostToDevice, g->st));
    GNN_TRY(hipMemcpyAsync(g->mask, mask, 4 * (size_t)g->V, hipMemcpyHostToDevice, g->st));
    GNN_TRY(hipStreamSynchronize(g->st));
    return GPSAT_OK;
}

int gpsat_gnn_forward(gpsat_gnn_model* m, gpsat_gnn_graph* g, float* y)
{
    if (!m || !g || !y) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    Acts a;
    GNN_TRY(acts_for(m, g, a));
    GNN_TRY(forward_pass(m, g, a));
    GNN_TRY(head(m, g, a, false, 0.f));
    GNN_TRY(hipMemcpyAsync(y, g->Y, 8 * (size_t)g->V, hipMemcpyDeviceToHost, g->st));
    GNN_TRY(hipStreamSynchronize(g->st));
    return GPSAT_OK;
}

int gpsat_gnn_loss_grad(gpsat_gnn_model* m, gpsat_gnn_graph* g, float alpha, float* loss, float* grad, float* y)
{
    if (!m || !g) {
        set_error("null argument");
        return GPSAT_ERR_ARG;
    }
    Acts a;
    GNN_TRY(acts_for(m, g, a));
    GNN_TRY(hipMemsetAsync(m->grad, 0, 4 * N_PARAMS, g->st));
    GNN_TRY(forward_pass(m, g, a));
    GNN_TRY(head(m, g, a, true, alpha));
    GNN_TRY(backward_pass(m, g, a));
    if (loss) GNN_TRY(hipMemcpyAsync(loss, a.loss, 4, hipMemcpyDeviceToHost, g->st));
    if (grad) GNN_TRY(hipMemcpyAsync(grad, m->grad, 4 * N_PARAMS, hipMemcpyDeviceToHost, g->st));
    if (y) GNN_TRY(hipMemcpyAsync(y, g->Y, 8 * (size_t)g->V, hipMemcpyDeviceToHost, g->st));
    GNN_TRY(hipStreamSynchronize(g->st));
    return GPSAT_OK;
}

int gpsat_gnn_train_step(gpsat_gnn_model* m, gpsat_gnn_graph* g, float alpha, float lr, float* loss)
{
    if (!m || !g || !(lr > 0.f)) {
        set_error("bad argument");
        return GPSAT_ERR_ARG;
    }
    Acts a;
    GNN_TRY(acts_for(m, g, a));
    GNN_TRY(hipMemsetAsync(m->grad, 0, 4 * N_PARAMS, g->st));
    GNN_TRY(forward_pass(m, g, a));
    GNN_TRY(head(m, g, a, true, alpha));
    GNN_TRY(backward_pass(m, g, a));
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, g->st, m->th, m->grad, m->m, m->v, N_PARAMS, lr, b1, b2, 1e-8f,
                       c1, c2);
    m->dirty = true;
    GNN_TRY(hipGetLastError());
    if (loss) GNN_TRY(hipMemcpyAsync(loss, a.loss, 4, hipMemcpyDeviceToHost, g->st));
    GNN_TRY(hipStreamSynchronize(g->st));
    return GPSAT_OK;
}

}  // extern "C"
