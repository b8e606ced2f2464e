// Guiding MLP on the gfx950 matrix cores: BF16 operands, FP32 accumulation, with
// v_mfma_f32_16x16x32_bf16.  Operand layout of that instruction (lane l of a wave64):
//   A (16 x 32): row l % 16, k = 8 * (l / 16) .. + 7
//   B (32 x 16): column l % 16, k = 8 * (l / 16) .. + 7
//   D (16 x 16): column l % 16, rows 4 * (l / 16) .. + 3
// Activations live in LDS as bf16 [row][64] so that an A operand is one 16-byte read;
// weights are kept twice, transposed ([out][in], B operand of the forward product) and
// plain ([in][out], B operand of the backward product dX = dZ W^T), both 16-byte reads.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "guide.h"
#include "guide_device.h"

namespace phg {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__host__ __device__ __forceinline__ int layer_out(int l) { return l == kGuideLayers - 1 ? kGuideOut : kGuideHid; }

// fp32 master parameters -> bf16 transposed / plain copies (rows padded to 64) and fp32 biases
// Inference keeps activations in registers: a layer is computed transposed, Z^T = W^T X^T,
// so that the lane holding output row 16 mt + 4 g + i of query r (the matrix-core result
// layout) feeds it to the next layer as the B operand.  The input (k) order of the next layer
// is therefore permuted: position j (0..7) of lane group g in k-chunk kc is input
// 16 (2 kc + j / 4) + 4 g + j % 4, and the weights are stored in that order.  The last
// layer's 40 lobe outputs are moved to rows where lane group g holds lobes 2g and 2g + 1 whole
// (slots s = 0..9 at rows 16 (s / 4) + 4 g + s % 4), and the selection logit (output 40) to
// slot 10 of lane group 0; the other 7 of the 48 rows stay zero.
__host__ __device__ inline int infer_in_index(int kc, int g, int j) { return 16 * (2 * kc + (j >> 2)) + 4 * g + (j & 3); }
// The first layer's input is the encoding, so its k order is free: lane group g encodes
// feature coordinates 2g (k-chunk 0) and 2g + 1 (k-chunk 1) whole — slot 0 the raw value,
// slots 1..6 sin and cos at the three bands — and slot 7 holds raw coordinate 8 and the
// constant (g = 0) or the sine and cosine of coordinate 8 at band g - 1 (g = 1..3).  A lane
// then evaluates two coordinates' encodings instead of 16 scattered entries.
__host__ __device__ inline int infer_in0_index(int kc, int g, int j) {
    if (j == 7) return g == 0 ? (kc == 0 ? 8 : kGuideIn - 1) : 9 + 18 * (g - 1) + 16 + kc;
    const int c = 2 * g + kc;
    if (j == 0) return c;
    return 9 + 18 * ((j - 1) >> 1) + 2 * c + ((j - 1) & 1);
}
// logical output (0..40) at padded output row R of the last layer, or -1
__host__ __device__ inline int infer_out_logical(int R) {
    const int g = (R & 15) >> 2, s = 4 * (R >> 4) + (R & 3);
    if (s == 10 && g == 0) return kGuideSel;
    if (s >= 10) return -1;
    const int lobe = 2 * g + (s >= 5 ? 1 : 0), p = s >= 5 ? s - 5 : s;
    return 5 * lobe + p;
}

__global__ void k_guide_prepare(const float* __restrict__ params, uint16_t* __restrict__ wt_bits,
                                uint16_t* __restrict__ wn_bits, float* __restrict__ bias, uint16_t* __restrict__ wp_bits,
                                float* __restrict__ bp) {
    __bf16* wt = reinterpret_cast<__bf16*>(wt_bits);
    __bf16* wn = reinterpret_cast<__bf16*>(wn_bits);
    __bf16* wp = reinterpret_cast<__bf16*>(wp_bits);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kGuideLayers * 64 * 64) {
        // wp index ((((l * 4 + mt) * 2 + kc) * 4 + g) * 16 + m) * 8 + j: output row 16 mt + m
        // of layer l, input infer_in0_index(kc, g, j) (l = 0) or infer_in_index(kc, g, j); lane
        // 16 g + m reads its 16-byte A operand at 16 (16 g + m) bytes from the (mt, kc) block, so
        // the lanes of a wave read consecutive bytes (no LDS bank conflicts)
        const int l = i / 4096, rem = i % 4096;
        const int mt = rem >> 10, kc = (rem >> 9) & 1, g = (rem >> 7) & 3, m = (rem >> 3) & 15, j = rem & 7;
        const int R = 16 * mt + m, in = l == 0 ? infer_in0_index(kc, g, j) : infer_in_index(kc, g, j);
        const int nout = layer_out(l);
        const int o = l == kGuideLayers - 1 ? infer_out_logical(R) : R;
        const float* W = params + kGuideParamOffsets[2 * l];
        wp[i] = (__bf16)(o >= 0 && o < nout ? W[in * nout + o] : 0.0f);
    }
    if (i < kGuideLayers * 64) {
        const int l = i / 64, R = i % 64;
        const int o = l == kGuideLayers - 1 ? infer_out_logical(R) : R;
        bp[i] = o >= 0 && o < layer_out(l) ? params[kGuideParamOffsets[2 * l + 1] + o] : 0.0f;
    }
    if (i < kGuideLayers * 64 * 64) {
        const int l = i / 4096, rem = i % 4096, a = rem / 64, b = rem % 64;  // wt: a = out, b = in
        const int nout = layer_out(l);
        const float* W = params + kGuideParamOffsets[2 * l];
        const float v = a < nout ? W[b * nout + a] : 0.0f;
        wt[i] = (__bf16)v;
        // wn: [in][out]: a = in, b = out.  The selection output is a head on the last hidden
        // layer: its column is zero here, so its gradient does not flow back into the hidden
        // layers (they learn from the lobes' gradients alone).
        const bool head = l == kGuideLayers - 1 && b == kGuideSel;
        const float v2 = b < nout && !head ? W[a * nout + b] : 0.0f;
        wn[i] = (__bf16)v2;
    }
    if (i < kGuideLayers * 64) {
        const int l = i / 64, o = i % 64;
        bias[i] = o < layer_out(l) ? params[kGuideParamOffsets[2 * l + 1] + o] : 0.0f;
    }
}

// Features of one query: from a feature array, or from a queued hit (position, shading
// normal and outgoing direction of the surface point, as the shade stage computes them).
struct InferSource {
    const float* features;   // 9 per query, or null
    const float4* hits;      // tri (bits), bary a, b, c
    const phg_ray* rays;     // queued extension rays (direction)
    const TriRec* tris;
    GuideScene gs;
};

// The inputs of one query as loaded from memory, fetched an iteration ahead of their use:
// a feature row (a = x0..3, b = x4..7, c = x8), or a queued hit (a) and its ray's direction (b).
struct QueryRaw {
    float4 a, b;
    float c;
};

__device__ __forceinline__ QueryRaw query_fetch(const InferSource& src, uint32_t q, uint32_t n, const QueueMap& qm) {
    QueryRaw r;
    r.a = make_float4(__int_as_float(-1), 0.0f, 0.0f, 0.0f);
    r.b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r.c = 0.0f;
    if (q >= n) {
        if (src.features) r.a.x = 0.0f;
        return r;
    }
    if (src.features) {
        const float* f = src.features + 9 * (size_t)q;
        r.a = make_float4(f[0], f[1], f[2], f[3]);
        r.b = make_float4(f[4], f[5], f[6], f[7]);
        r.c = f[8];
    } else {
        const uint32_t qs = qm.slot(q);
        r.a = src.hits[qs];
        r.b = reinterpret_cast<const float4*>(src.rays + qs)[1];
    }
    return r;
}

// features of a fetched query (zeros for a miss or a query past the end)
__device__ __forceinline__ void query_features_of(const InferSource& src, const QueryRaw& r, float* x) {
    if (src.features) {
        x[0] = r.a.x; x[1] = r.a.y; x[2] = r.a.z; x[3] = r.a.w;
        x[4] = r.b.x; x[5] = r.b.y; x[6] = r.b.z; x[7] = r.b.w; x[8] = r.c;
        return;
    }
    const int32_t tri = __float_as_int(r.a.x);
    if (tri < 0) { for (int k = 0; k < 9; ++k) x[k] = 0.0f; return; }
    V3 a, b, c;
    load_tri(src.tris, (uint32_t)tri, &a, &b, &c);
    const SurfaceHit X = make_hit(a, b, c, (uint32_t)tri, r.a.y, r.a.z, r.a.w);
    guide_features(src.gs, X.P, X.Ns, mk(-r.b.x, -r.b.y, -r.b.z), x);
}

// Inference: decoded mixture (8 x (mu, kappa, weight), selection probability) per query, or
// raw outputs.  Each wave takes 64 queries per iteration: every lane computes one query's
// 9 features (kept in LDS), then the four 16-query tiles go through the 5 layers together,
// so that each weight operand read from LDS feeds four matrix-core products.  Activations
// stay in registers: a layer is computed transposed (see infer_in_index), the bias enters
// as the accumulator of the first product, and ReLU acts on the packed bf16 results (the
// sign bit decides, so max(bf16(v), 0) as 16-bit integers equals bf16(max(v, 0))).  The last
// layer leaves lobes 2g and 2g + 1 of query r in lane 16 g + r, which decodes them; the
// softmax over the 8 lobe weights reduces across the 4 lanes of a query.
constexpr int kInferWaves = 8;   // waves per block
constexpr int kInferGroup = 4;   // 16-query tiles that go through the layers together
typedef __attribute__((ext_vector_type(8))) short s16x8;

typedef __attribute__((ext_vector_type(4))) unsigned u32x4;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) float f32x2;

// two floats to packed bf16, round to nearest even: converting them as one 2-vector makes the
// compiler emit one v_cvt_pk_bf16_f32 per pair instead of two conversions and a re-pack
__device__ __forceinline__ unsigned cvt_pk_bf16(float lo, float hi) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
__device__ __forceinline__ bf16x8 pack_bf16(f32x4 lo, f32x4 hi) {
    const u32x4 u = {cvt_pk_bf16(lo[0], lo[1]), cvt_pk_bf16(lo[2], lo[3]), cvt_pk_bf16(hi[0], hi[1]),
                     cvt_pk_bf16(hi[2], hi[3])};
    return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ bf16x8 relu_pack(f32x4 lo, f32x4 hi) {
    s16x8 s = __builtin_bit_cast(s16x8, pack_bf16(lo, hi));
    s = __builtin_elementwise_max(s, (s16x8)0);
    return __builtin_bit_cast(bf16x8, s);
}

__global__ __launch_bounds__(64 * kInferWaves) void k_guide_infer(GuideNetDev net, InferSource src,
                                                                  const uint32_t* regions, const uint32_t* bounce_base,
                                                                  uint32_t bk, uint32_t count, float* __restrict__ out,
                                                                  int raw, unsigned long long* queries) {
    __shared__ __attribute__((aligned(16))) __bf16 s_w[kGuideLayers * 64 * 64];
    __shared__ __attribute__((aligned(16))) float s_b[kGuideLayers * 64];
    __shared__ float s_x[kInferWaves][9 * 64];   // features of the wave's 64 queries, [coordinate][query]
    // queued hits of bounce bounce_base[0] + bk: the entries of the extension queue's regions
    // (query d is the d-th of them, regions in order), or `count` features in a plain array
    QueueMap qm;
    if (regions) qm.load(regions + fetch_slot(kPushExt, bounce_base[0] + bk), bounce_base[1]);
    const uint32_t n = regions ? qm.n : count;
    if (queries && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(queries, (unsigned long long)n);
    if (n == 0) return;   // whole block (a captured chunk runs past the last bounce of its paths)
    for (int i = threadIdx.x; i < kGuideLayers * 64 * 64 / 8; i += 64 * kInferWaves)
        reinterpret_cast<bf16x8*>(s_w)[i] = reinterpret_cast<const bf16x8*>(net.wp)[i];  // bf16 bit patterns
    for (int i = threadIdx.x; i < kGuideLayers * 64; i += 64 * kInferWaves) s_b[i] = net.bp[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    float* xs = s_x[wave];
    // band g - 1 of coordinate 8 (slot 7 of lane groups 1..3), in revolutions
    const float scale8 = g == 0 ? 0.0f : (float)(1 << (g - 1)) * 0.5f;
    const uint32_t step = gridDim.x * kInferWaves * 64u;
    const uint32_t first = (blockIdx.x * kInferWaves + wave) * 64u;
    // Software pipeline over the wave's iterations.  Loads and stores retire in issue order, so
    // a wait for a load also waits for every store issued before it: the next iteration's
    // features (the triangle loads) are therefore made after this iteration's layers and before
    // its stores, and the raw inputs of the iteration after that are fetched one iteration ahead.
    QueryRaw next = query_fetch(src, first + (uint32_t)lane, n, qm);
    if (first < n) {
        float x[9];
        query_features_of(src, next, x);
        next = query_fetch(src, first + step + (uint32_t)lane, n, qm);
#pragma unroll
        for (int k = 0; k < 9; ++k) xs[k * 64 + lane] = x[k];
        wave_sync();
    }
    for (uint32_t base = first; base < n; base += step) {
        // the four 16-query tiles in groups of kInferGroup: the layer-0 operands
        // (infer_in0_index), the five layers, then the decode
#pragma unroll
        for (int T0 = 0; T0 < 4; T0 += kInferGroup) {
            if (base + 16u * T0 >= n) break;  // whole wave
            bf16x8 bop[kInferGroup][2];
#pragma unroll
            for (int t = 0; t < kInferGroup; ++t) {
                const int qi = 16 * (T0 + t) + r;
                const float v8 = xs[8 * 64 + qi];
                const float r8 = scale8 * v8;
                const float e0 = g == 0 ? v8 : __builtin_amdgcn_sinf(r8);
                const float e1 = g == 0 ? 1.0f : __builtin_amdgcn_cosf(r8);
#pragma unroll
                for (int kc = 0; kc < 2; ++kc) {
                    const float v = xs[(2 * g + kc) * 64 + qi];
                    float e[8];
                    e[0] = v;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const float rev = (float)(1 << b) * 0.5f * v;
                        e[1 + 2 * b] = __builtin_amdgcn_sinf(rev);
                        e[2 + 2 * b] = __builtin_amdgcn_cosf(rev);
                    }
                    e[7] = kc == 0 ? e0 : e1;
                    bop[t][kc] = pack_bf16(f32x4{e[0], e[1], e[2], e[3]}, f32x4{e[4], e[5], e[6], e[7]});
                }
            }
            f32x4 acc[kInferGroup][4];
#pragma unroll
            for (int l = 0; l < kGuideLayers; ++l) {
                const int nmt = l == kGuideLayers - 1 ? kGuideOutPad / 16 : 4;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    if (mt >= nmt) continue;
                    const f32x4 bias = *reinterpret_cast<const f32x4*>(s_b + l * 64 + 16 * mt + 4 * g);
#pragma unroll
                    for (int kc = 0; kc < 2; ++kc) {
                        const bf16x8 a = *reinterpret_cast<const bf16x8*>(s_w + l * 4096 + ((mt * 2 + kc) * 64 + lane) * 8);
#pragma unroll
                        for (int t = 0; t < kInferGroup; ++t)
                            acc[t][mt] = mfma16(a, bop[t][kc], kc == 0 ? bias : acc[t][mt]);
                    }
                }
                if (l < kGuideLayers - 1) {
#pragma unroll
                    for (int t = 0; t < kInferGroup; ++t) {
                        bop[t][0] = relu_pack(acc[t][0], acc[t][1]);
                        bop[t][1] = relu_pack(acc[t][2], acc[t][3]);
                    }
                }
            }
            if (T0 + kInferGroup >= 4 && base + step < n) {
                // the next iteration's features, staged in LDS once every lane has encoded
                float x[9];
                query_features_of(src, next, x);
                next = query_fetch(src, base + 2u * step + (uint32_t)lane, n, qm);
                wave_sync();
#pragma unroll
                for (int k = 0; k < 9; ++k) xs[k * 64 + lane] = x[k];
                wave_sync();
            }
#pragma unroll
            for (int t = 0; t < kInferGroup; ++t) {
                if (base + 16u * (T0 + t) >= n) break;  // whole wave
                const uint32_t q = base + 16u * (T0 + t) + (uint32_t)r;
                const size_t qs = qm.slot(q);   // the query's queue slot (its output row)
                // slots s = 0..9 of this lane: logical outputs 10 g + s (lobes 2g, 2g + 1); slot 10
                // of lane group 0: the selection logit
                float o[11];
#pragma unroll
                for (int s2 = 0; s2 < 11; ++s2) o[s2] = acc[t][s2 >> 2][s2 & 3];
                if (raw) {
                    float* dst = out + qs * kGuideOut + 10 * g;
                    if (q < n) {
                        for (int s2 = 0; s2 < 10; ++s2) dst[s2] = o[s2];
                        if (g == 0) out[qs * kGuideOut + kGuideSel] = o[10];
                    }
                    continue;
                }
                float* dst = out + qs * kGuideDec + 10 * g;
                if (q < n && g == 0) out[qs * kGuideDec + kGuideSel] = guide_sel(o[10]);
                float mx = fmaxf(o[4], o[9]);
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                mx = fmaxf(mx, __shfl_xor(mx, 32));
                const float ea = __expf(o[4] - mx), eb = __expf(o[9] - mx);
                float sum = ea + eb;
                sum += __shfl_xor(sum, 16);
                sum += __shfl_xor(sum, 32);
                // hardware reciprocal and reciprocal square root (about 1 ulp): the lobes are
                // stored once and the shade stage samples and evaluates them from the same values
                const float inv_sum = __builtin_amdgcn_rcpf(sum);
                if (q < n) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const float* oh = o + 5 * h;
                        const float r2 = oh[0] * oh[0] + oh[1] * oh[1] + oh[2] * oh[2];
                        float* L = dst + 5 * h;
                        if (r2 > 0x1p-100f) {
                            const float inv = __builtin_amdgcn_rsqf(r2);
                            L[0] = oh[0] * inv; L[1] = oh[1] * inv; L[2] = oh[2] * inv;
                        } else {
                            L[0] = 0.0f; L[1] = 0.0f; L[2] = 1.0f;
                        }
                        L[3] = __expf(clampf(oh[3], kKappaLogMin, kKappaLogMax));
                        L[4] = (h == 0 ? ea : eb) * inv_sum;
                    }
                }
            }
        }
    }
}

// Training step as batch GEMMs over the whole minibatch (N records):
//   forward   X_0 = encode(features); Z_l = X_l W_l + b_l; X_{l+1} = relu(Z_l)   (store X_l)
//   output    dZ_4 = d(mean record loss) / dZ_4                                  (per record)
//   backward  dW_l = X_l^T dZ_l, db_l = column sums of dZ_l,
//             dZ_{l-1} = (dZ_l W_l^T) * [X_l > 0]                                (layer by layer)
// X_l and dZ_l are bf16 [record][64] in global memory (the matrix-core operands); every wave
// of the backward kernel reduces dW over its own contiguous run of records in registers and
// writes one partial, and k_train_reduce sums the partials, so no atomics are involved and
// the result does not depend on scheduling.

// Forward over 16-record tiles (4 waves per block, weights in LDS), storing every layer's
// input and the output-layer gradient.  Rows of 64 bf16 in LDS (activations, weights) keep
// 16-byte chunk c of row r at position c ^ (r & 7) (lds_chunk), so that the lanes reading
// chunk g of 16 consecutive rows fall on 8 different bank groups instead of one.
__device__ __forceinline__ int lds_chunk(int row, int chunk) { return row * 8 + (chunk ^ (row & 7)); }

// One dense layer for a 16-row tile held by one wave: act_out = act(act_in . W + b).
// act_in/act_out: bf16 [16][64] (swizzled rows) in LDS; wt_l: bf16 [64 out][64 in] (swizzled
// rows); out_f32 (last layer): [16][48].
__device__ __forceinline__ void layer_forward(const __bf16* act_in, const __bf16* wt_l, const float* bias_l,
                                              __bf16* act_out, float* out_f32, bool last) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const bf16x8* ain = reinterpret_cast<const bf16x8*>(act_in);
    const bf16x8* win = reinterpret_cast<const bf16x8*>(wt_l);
    const bf16x8 a0 = ain[lds_chunk(r, g)];
    const bf16x8 a1 = ain[lds_chunk(r, 4 + g)];
    const int ntiles = last ? kGuideOutPad / 16 : 4;
    for (int nt = 0; nt < ntiles; ++nt) {
        const int col = nt * 16 + r;
        const bf16x8 b0 = win[lds_chunk(col, g)];
        const bf16x8 b1 = win[lds_chunk(col, 4 + g)];
        const float bv = bias_l[col];
        f32x4 c = {bv, bv, bv, bv};
        c = mfma16(a0, b0, c);
        c = mfma16(a1, b1, c);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * g + i;
            if (last) out_f32[row * kGuideOutPad + col] = c[i];
            else act_out[lds_chunk(row, col >> 3) * 8 + (col & 7)] = (__bf16)fmaxf(c[i], 0.0f);
        }
    }
}

// Output gradients of one record computed by its 4 lanes (lanes r, r + 16, r + 32, r + 48 of
// the tile): lane group g takes lobes 2g and 2g + 1 (outputs 10 g .. 10 g + 9) and lane group 0
// also the selection logit (output 40).  The arithmetic is guide_output_grad's; the softmax and
// the mixture density, sums over the 8 lobes, are reduced across the 4 lanes.
__device__ __forceinline__ void output_grad_lanes(const float* o, int g, V3 w, float weight, float bsdf_pdf, float sel,
                                                  float samp_pdf, float* gl, float* gsel) {
    const float* ol = o + 10 * g;
    float mx = fmaxf(ol[4], ol[9]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float e[2] = {expf(ol[4] - mx), expf(ol[9] - mx)};
    float sum = e[0] + e[1];
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    float mu[2][3], kap[2], wt[2], c[2], p[2], inv[2];
    float qp = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float* oh = ol + 5 * h;
        const float rl = sqrtf(oh[0] * oh[0] + oh[1] * oh[1] + oh[2] * oh[2]);
        if (rl > 0.0f) { mu[h][0] = oh[0] / rl; mu[h][1] = oh[1] / rl; mu[h][2] = oh[2] / rl; }
        else { mu[h][0] = 0.0f; mu[h][1] = 0.0f; mu[h][2] = 1.0f; }
        inv[h] = rl > 0.0f ? 1.0f / rl : 1.0f;
        kap[h] = expf(clampf(oh[3], kKappaLogMin, kKappaLogMax));
        wt[h] = e[h] / sum;
        c[h] = mu[h][0] * w.x + mu[h][1] * w.y + mu[h][2] * w.z;
        p[h] = vmf_norm(kap[h]) * expf(kap[h] * (c[h] - 1.0f));
        qp += wt[h] * p[h];
    }
    float q = qp + __shfl_xor(qp, 16);
    q += __shfl_xor(q, 32);
    const bool learned = samp_pdf > 0.0f && sel > 0.0f;
    const float a = learned ? guide_sel(o[kGuideSel]) : sel;
    const bool mix = a > 0.0f && bsdf_pdf > 0.0f;
    const float dens = mix ? a * q + (1.0f - a) * bsdf_pdf : q;
    const float share = mix ? a * q / dens : 1.0f;
    const float scale = q > 0.0f ? -weight * share / q : 0.0f;
    *gsel = (learned && mix && dens > 0.0f)
                ? -weight * weight * samp_pdf * (q - bsdf_pdf) / (dens * dens) *
                      ((a - kSelMin) * (kSelMax - a) / (kSelMax - kSelMin))
                : 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float k = kap[h];
        const float wp = wt[h] * p[h];
        const float f = scale * wp * k * inv[h];
        gl[5 * h + 0] = f * (w.x - c[h] * mu[h][0]);
        gl[5 * h + 1] = f * (w.y - c[h] * mu[h][1]);
        gl[5 * h + 2] = f * (w.z - c[h] * mu[h][2]);
        const float lk = ol[5 * h + 3];
        float dl;  // 1/k - 2/expm1(2k) - 1
        if (k < 1e-2f) dl = -k / 3.0f + k * k * k / 45.0f;
        else dl = 1.0f / k - 2.0f / expm1f(2.0f * k) - 1.0f;
        gl[5 * h + 3] = (lk > kKappaLogMin && lk < kKappaLogMax) ? scale * wp * (dl + c[h]) * k : 0.0f;
        gl[5 * h + 4] = scale * (wp - wt[h] * q);
    }
}

__global__ __launch_bounds__(256) void k_train_forward(GuideNetDev net, const float* __restrict__ rec, uint32_t nrec,
                                                       uint32_t npad, float inv_batch, __bf16* __restrict__ X,
                                                       __bf16* __restrict__ dz_out) {
    __shared__ __attribute__((aligned(16))) __bf16 s_wt[kGuideLayers * 64 * 64];
    __shared__ float s_bias[kGuideLayers * 64];
    __shared__ __attribute__((aligned(16))) __bf16 s_act[4][2][16 * 64];
    __shared__ float s_out[4][16 * kGuideOutPad];
    for (int i = threadIdx.x; i < kGuideLayers * 64 * 64 / 8; i += 256)
        reinterpret_cast<bf16x8*>(s_wt)[(i >> 9) * 512 + lds_chunk((i >> 3) & 63, i & 7)] =
            reinterpret_cast<const bf16x8*>(net.wt)[i];
    for (int i = threadIdx.x; i < kGuideLayers * 64; i += 256) s_bias[i] = net.bias[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (uint32_t tile = blockIdx.x * 4 + wave; tile * 16 < npad; tile += gridDim.x * 4) {
        const uint32_t q = tile * 16 + r;
        const bool valid = q < nrec;
        float x[9];
        if (valid) for (int k = 0; k < 9; ++k) x[k] = rec[kGuideRecordFloats * (size_t)q + k];
        else for (int k = 0; k < 9; ++k) x[k] = 0.0f;
        // encoding entries 16 g .. 16 g + 15 of record r: chunks 2 g and 2 g + 1 of its row
        bf16x8* a0 = reinterpret_cast<bf16x8*>(s_act[wave][0]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 v;
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = valid ? (__bf16)guide_encode(x, 16 * g + 8 * h + j) : (__bf16)0.0f;
            a0[lds_chunk(r, 2 * g + h)] = v;
        }
        wave_sync();
        for (int l = 0; l < kGuideLayers; ++l) {
            // store this layer's input tile: 16 rows x 64 bf16 = 128 x 16 bytes, two per lane
            const __bf16* src = s_act[wave][l & 1];
            const bf16x8* sv = reinterpret_cast<const bf16x8*>(src);
            __bf16* dst = X + ((size_t)l * npad + (size_t)tile * 16) * 64;
            reinterpret_cast<bf16x8*>(dst)[lane] = sv[lds_chunk(lane >> 3, lane & 7)];
            reinterpret_cast<bf16x8*>(dst)[lane + 64] = sv[lds_chunk(8 + (lane >> 3), lane & 7)];
            const bool last = l == kGuideLayers - 1;
            layer_forward(src, s_wt + l * 4096, s_bias + l * 64, s_act[wave][(l + 1) & 1], s_out[wave], last);
            wave_sync();
        }
        // output-layer gradients: the 4 lanes of record r share its 8 lobes; the results go back
        // into the record's row of s_out (outputs 0..40, zeros up to 47) and leave as bf16 rows
        // of 64 (two 16-byte stores per lane)
        float* orow = s_out[wave] + r * kGuideOutPad;
        float gl[10], gsel;
        {
            const float* R = rec + kGuideRecordFloats * (size_t)(valid ? q : 0u);
            const V3 w = valid ? mk(R[9], R[10], R[11]) : mk(0.0f, 0.0f, 1.0f);
            output_grad_lanes(orow, g, w, valid ? R[12] : 0.0f, valid ? R[13] : 0.0f, valid ? R[14] : 0.0f,
                              valid ? R[15] : 0.0f, gl, &gsel);
        }
        wave_sync();
#pragma unroll
        for (int s2 = 0; s2 < 10; ++s2) orow[10 * g + s2] = valid ? gl[s2] * inv_batch : 0.0f;
        if (g == 0) orow[kGuideSel] = valid ? gsel * inv_batch : 0.0f;
        if (g == 1)
            for (int k = kGuideOut; k < kGuideOutPad; ++k) orow[k] = 0.0f;
        wave_sync();
        bf16x8 d0, d1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k0 = 16 * g + j, k1 = 16 * g + 8 + j;
            d0[j] = (__bf16)(k0 < kGuideOutPad ? orow[k0] : 0.0f);
            d1[j] = (__bf16)(k1 < kGuideOutPad ? orow[k1] : 0.0f);
        }
        bf16x8* d = reinterpret_cast<bf16x8*>(dz_out + (size_t)q * 64 + 16 * g);
        d[0] = d0;
        d[1] = d1;
        wave_sync();
    }
}

// Backward through layer l for a contiguous run of 32-record sub-tiles per wave: dW_l and db_l
// partials in registers, and (l > 0) dZ_{l-1}.  Sub-tiles of X_l and dZ_l are staged in LDS
// so that the transposed operands of X_l^T dZ_l are LDS gathers.
constexpr int kBackWaves = 4;
__global__ __launch_bounds__(64 * kBackWaves) void k_train_backward(const __bf16* __restrict__ wn_l,
                                                                    const __bf16* __restrict__ X_l,
                                                                    const __bf16* __restrict__ dz,
                                                                    __bf16* __restrict__ dz_prev, uint32_t npad,
                                                                    uint32_t chunk, int l, float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) __bf16 s_w[64 * 64];
    __shared__ __attribute__((aligned(16))) __bf16 s_x[kBackWaves][32 * 64];
    __shared__ __attribute__((aligned(16))) __bf16 s_dz[kBackWaves][32 * 64];
    for (int i = threadIdx.x; i < 64 * 64 / 8; i += 64 * kBackWaves)
        reinterpret_cast<bf16x8*>(s_w)[i] = reinterpret_cast<const bf16x8*>(wn_l)[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const uint32_t wid = blockIdx.x * kBackWaves + wave;
    const int nout = layer_out(l), ntiles = (nout + 15) / 16;
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float acc_b = 0.0f;
    __bf16* sx = s_x[wave];
    __bf16* sd = s_dz[wave];
    const uint32_t begin = wid * chunk, end = min(begin + chunk, npad);
    // 32 records of X_l and dZ_l per sub-tile: 32 x 64 bf16 = 256 x 16 bytes each, 4 per lane;
    // the next sub-tile is loaded into registers while this one is computed
    bf16x8 px[4], pd[4];
    if (begin < end) {
        const bf16x8* gx = reinterpret_cast<const bf16x8*>(X_l + (size_t)begin * 64);
        const bf16x8* gd = reinterpret_cast<const bf16x8*>(dz + (size_t)begin * 64);
#pragma unroll
        for (int k = 0; k < 4; ++k) { px[k] = gx[lane + 64 * k]; pd[k] = gd[lane + 64 * k]; }
    }
    for (uint32_t base = begin; base < end; base += 32) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            reinterpret_cast<bf16x8*>(sx)[lane + 64 * k] = px[k];
            reinterpret_cast<bf16x8*>(sd)[lane + 64 * k] = pd[k];
        }
        if (base + 32 < end) {
            const bf16x8* gx = reinterpret_cast<const bf16x8*>(X_l + (size_t)(base + 32) * 64);
            const bf16x8* gd = reinterpret_cast<const bf16x8*>(dz + (size_t)(base + 32) * 64);
#pragma unroll
            for (int k = 0; k < 4; ++k) { px[k] = gx[lane + 64 * k]; pd[k] = gd[lane + 64 * k]; }
        }
        wave_sync();
        // dW_l += X^T dZ: A = X^T (16 inputs x 32 records), B = dZ (32 records x 16 outputs)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = sx[(8 * g + j) * 64 + mt * 16 + r];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if (nt >= ntiles) continue;
                bf16x8 b;
#pragma unroll
                for (int j = 0; j < 8; ++j) b[j] = sd[(8 * g + j) * 64 + nt * 16 + r];
                acc[mt][nt] = mfma16(a, b, acc[mt][nt]);
            }
        }
        // db_l: column sums of dZ (one column per lane)
        for (int s = 0; s < 32; ++s) acc_b += (float)sd[s * 64 + lane];
        if (l > 0) {
            // dZ_{l-1} = (dZ W^T) masked by X_l > 0, computed transposed (W dZ^T: rows = inputs,
            // columns = records), so that a lane holds 4 consecutive inputs of one record: one
            // 8-byte mask read and one 8-byte store per product
            typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
            for (int t = 0; t < 2; ++t) {
                const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(sd + (t * 16 + r) * 64 + 8 * g);
                const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(sd + (t * 16 + r) * 64 + 32 + 8 * g);
                const int rec_row = t * 16 + r;
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const int in_row = nt * 16 + r;
                    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(s_w + in_row * 64 + 8 * g);
                    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(s_w + in_row * 64 + 32 + 8 * g);
                    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
                    c = mfma16(a0, b0, c);
                    c = mfma16(a1, b1, c);
                    // inputs nt * 16 + 4 g + i of record rec_row
                    const int col = nt * 16 + 4 * g;
                    const bf16x4 xm = *reinterpret_cast<const bf16x4*>(sx + rec_row * 64 + col);
                    bf16x4 o;
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = (__bf16)((float)xm[i] > 0.0f ? c[i] : 0.0f);
                    *reinterpret_cast<bf16x4*>(dz_prev + (size_t)(base + rec_row) * 64 + col) = o;
                }
            }
        }
        wave_sync();
    }
    // the block's partial, dW [64 in][64 out] then db [64]: the waves' sums added in LDS in
    // wave order (a fixed order, so the result does not depend on scheduling), then written once
    __shared__ float s_part[64 * 64 + 64];
    for (int w = 0; w < kBackWaves; ++w) {
        if (wave == w) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int e = (mt * 16 + 4 * g + i) * 64 + nt * 16 + r;
                        s_part[e] = w == 0 ? acc[mt][nt][i] : s_part[e] + acc[mt][nt][i];
                    }
            s_part[64 * 64 + lane] = w == 0 ? acc_b : s_part[64 * 64 + lane] + acc_b;
        }
        __syncthreads();
    }
    float* P = partial + (size_t)blockIdx.x * (64 * 64 + 64);
    for (int e = threadIdx.x; e < 64 * 64 + 64; e += 64 * kBackWaves) P[e] = s_part[e];
}

// First-level sums of the waves' partials, all layers in one launch: grid (17, slices, 5);
// slice s of layer l adds partials s, s + slices, ... in that order.
__global__ void k_train_reduce_parts(const float* __restrict__ partial, size_t layer_stride, uint32_t nparts,
                                     float* __restrict__ partial2) {
    const int l = blockIdx.z;
    const uint32_t s = blockIdx.y, slices = gridDim.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 64 * 64 + 64) return;
    const float* P = partial + (size_t)l * layer_stride;
    float acc = 0.0f;
    for (uint32_t p = s; p < nparts; p += slices) acc += P[(size_t)p * (64 * 64 + 64) + e];
    partial2[((size_t)l * slices + s) * (64 * 64 + 64) + e] = acc;
}

// grads[layer l] = sum of the slices (dW entries of real outputs, then biases); grid (17, 1, 5).
__global__ void k_train_reduce(const float* __restrict__ partial2, uint32_t slices, float* __restrict__ grads) {
    const int l = blockIdx.z;
    const int nout = layer_out(l);
    const int n = 64 * nout + nout;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int src = i < 64 * nout ? (i / nout) * 64 + (i % nout) : 64 * 64 + (i - 64 * nout);
    float s = 0.0f;
    for (uint32_t p = 0; p < slices; ++p) s += partial2[((size_t)l * slices + p) * (64 * 64 + 64) + src];
    grads[kGuideParamOffsets[2 * l] + i] = s;
}

// Adam (Kingma & Ba 2015) on the fp32 master parameters; clears the gradient.
__global__ void k_guide_adam(float* __restrict__ params, float* __restrict__ grads, float* __restrict__ m,
                             float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGuideParams) return;
    const float gr = __builtin_isfinite(grads[i]) ? grads[i] : 0.0f;  // a non-finite gradient skips the update
    const float mi = b1 * m[i] + (1.0f - b1) * gr;
    const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi; v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grads[i] = 0.0f;
}

__global__ void k_fill_float(float* p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Training records of a finished pass.  For each recorded vertex, (final path radiance -
// radiance before the direction was followed) / (throughput before the vertex), channel by
// channel, is f cos L_i / pdf of the sampled direction: a one-sample estimate of the target
// density (BSDF x cosine x incident radiance, which depends on the outgoing direction the
// network sees) over the sampling density.  Its luminance is the record's weight in the
// KL-divergence gradient.  Records are kept with probability keep_prob.
//
// The records a path keeps this pass, as a mask over its recorded bounces (and how many of
// its records are valid before the keep test).
__device__ __forceinline__ uint32_t collect_mask(const PathBuffers& pb, const PassParams& pp, uint32_t pid,
                                                 float keep_prob, uint32_t salt, uint32_t* valid) {
    const float4 L = pb.L[pid];
    uint32_t mask = 0, nvalid = 0;
    for (uint32_t b = 0; b < pp.rec_bounces; ++b) {
        const float4* rc = pb.rec + ((size_t)b * pp.paths + pid) * kRecFloat4;
        const float4 r0 = rc[0];
        if (r0.w < 0.0f) continue;       // a vertex without a record (a material the lobes do not serve)
        if (!(r0.w > 0.0f)) break;
        // (L - Lb) / T = f cos L_i / pdf along the sampled direction: the sample's weight for
        // fitting the lobes to the product of BSDF, cosine and incident radiance
        const float4 T = rc[4], lb = rc[5];
        const float lr = T.x > 0.0f ? (L.x - lb.x) / T.x : 0.0f;
        const float lg = T.y > 0.0f ? (L.y - lb.y) / T.y : 0.0f;
        const float lbb = T.z > 0.0f ? (L.z - lb.z) / T.z : 0.0f;
        const float w = luminance(rgb3(fmaxf(lr, 0.0f), fmaxf(lg, 0.0f), fmaxf(lbb, 0.0f)));
        if (!(w > 0.0f) || !__builtin_isfinite(w)) continue;
        ++nvalid;
        uint32_t h = pid * 0x9E3779B1u ^ (b + 1u) * 0x85EBCA77u ^ salt * 0xC2B2AE3Du;
        h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
        if ((float)(h >> 8) * 0x1p-24f >= keep_prob) continue;
        mask |= 1u << b;
    }
    *valid = nvalid;
    return mask;
}

// exclusive prefix sum over a 256-thread block (returns this thread's prefix; *total = block sum)
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t v, uint32_t* s, uint32_t* total) {
    const uint32_t t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t a = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    const uint32_t incl = s[t];
    *total = s[255];
    __syncthreads();
    return incl - v;
}

// Records are collected in path order, bounce by bounce within a path (no atomics, so the
// training batches and therefore the trained network do not depend on scheduling):
// count per block, scan the block counts, write.
__global__ __launch_bounds__(256) void k_guide_collect_count(PathBuffers pb, PassParams pp, float keep_prob, uint32_t salt,
                                                             uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t s[256];
    const uint32_t pid = blockIdx.x * 256 + threadIdx.x;
    uint32_t valid = 0, kept = 0;
    if (pid < pp.paths) kept = (uint32_t)__popc(collect_mask(pb, pp, pid, keep_prob, salt, &valid));
    uint32_t tk = 0, tv = 0;
    block_exclusive_scan256(kept, s, &tk);
    block_exclusive_scan256(valid, s, &tv);
    if (threadIdx.x == 0) { block_counts[blockIdx.x] = tk; block_counts[gridDim.x + blockIdx.x] = tv; }
}

// One block: block_counts[0, n) become exclusive offsets; counters[0] = records kept,
// counters[1] = records valid before the keep test.
__global__ __launch_bounds__(256) void k_guide_collect_scan(uint32_t* __restrict__ block_counts, uint32_t n,
                                                            uint32_t* __restrict__ counters) {
    __shared__ uint32_t s[256];
    const uint32_t per = (n + 255u) / 256u, b0 = min(threadIdx.x * per, n), b1 = min(b0 + per, n);
    uint32_t sum = 0, vsum = 0;
    for (uint32_t k = b0; k < b1; ++k) { sum += block_counts[k]; vsum += block_counts[n + k]; }
    uint32_t total = 0, vtotal = 0;
    uint32_t run = block_exclusive_scan256(sum, s, &total);
    block_exclusive_scan256(vsum, s, &vtotal);
    for (uint32_t k = b0; k < b1; ++k) { const uint32_t c = block_counts[k]; block_counts[k] = run; run += c; }
    if (threadIdx.x == 0) { counters[0] = total; counters[1] = vtotal; }
}

__global__ __launch_bounds__(256) void k_guide_collect_write(PathBuffers pb, PassParams pp, GuideScene gs,
                                                             float* __restrict__ out, uint32_t capacity,
                                                             const uint32_t* __restrict__ block_offsets,
                                                             float keep_prob, uint32_t salt) {
    __shared__ uint32_t s[256];
    const uint32_t pid = blockIdx.x * 256 + threadIdx.x;
    uint32_t valid = 0, mask = 0;
    if (pid < pp.paths) mask = collect_mask(pb, pp, pid, keep_prob, salt, &valid);
    uint32_t tk = 0;
    uint32_t slot = block_offsets[blockIdx.x] + block_exclusive_scan256((uint32_t)__popc(mask), s, &tk);
    for (uint32_t b = 0; mask != 0; ++b, mask >>= 1) {
        if (!(mask & 1u)) continue;
        const uint32_t k = slot++;
        if (k >= capacity) break;
        const float4* rc = pb.rec + ((size_t)b * pp.paths + pid) * kRecFloat4;
        const float4 L = pb.L[pid];
        const float4 r0 = rc[0], n4 = rc[1], v4 = rc[2], d4 = rc[3], T = rc[4], lb = rc[5];
        const float lr = T.x > 0.0f ? (L.x - lb.x) / T.x : 0.0f;
        const float lg = T.y > 0.0f ? (L.y - lb.y) / T.y : 0.0f;
        const float lbb = T.z > 0.0f ? (L.z - lb.z) / T.z : 0.0f;
        const float w = luminance(rgb3(fmaxf(lr, 0.0f), fmaxf(lg, 0.0f), fmaxf(lbb, 0.0f)));
        float x[9];
        guide_features(gs, mk(r0.x, r0.y, r0.z), mk(n4.x, n4.y, n4.z), mk(v4.x, v4.y, v4.z), x);
        float* o = out + (size_t)k * kGuideRecordFloats;
        for (int q = 0; q < 9; ++q) o[q] = x[q];
        o[9] = d4.x; o[10] = d4.y; o[11] = d4.z; o[12] = w;
        // BSDF density of the direction, the lobes' selection probability used there (0 where
        // the BSDF sampled alone) and, where that was the network's own, the density the
        // direction was drawn with (else 0)
        o[13] = n4.w; o[14] = v4.w; o[15] = d4.w;
    }
}

uint32_t guide_collect_blocks(uint32_t paths) { return (paths + 255u) / 256u; }

hipError_t guide_collect(const PathBuffers& pb, const PassParams& pp, const GuideScene& gs, float* out_rec,
                         uint32_t capacity, uint32_t* counters, uint32_t* block_counts, float keep_prob,
                         uint32_t salt, hipStream_t st) {
    const uint32_t nb = guide_collect_blocks(pp.paths);
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_collect_count, dim3(nb), dim3(256), 0, st, pb, pp, keep_prob, salt, block_counts);
    hipLaunchKernelGGL(k_guide_collect_scan, dim3(1), dim3(256), 0, st, block_counts, nb, counters);
    hipLaunchKernelGGL(k_guide_collect_write, dim3(nb), dim3(256), 0, st, pb, pp, gs, out_rec, capacity, block_counts,
                       keep_prob, salt);
    return hipGetLastError();
}

hipError_t guide_prepare(const GuideNetDev& net, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_prepare, dim3((kGuideLayers * 4096 + 255) / 256), dim3(256), 0, st,
                       net.params, net.wt_mut, net.wn_mut, net.bias_mut, net.wp_mut, net.bp_mut);
    return hipGetLastError();
}

hipError_t guide_infer_features(const GuideNetDev& net, const float* features, uint32_t n, float* out, int raw,
                                uint32_t grid, hipStream_t st) {
    InferSource src{};
    src.features = features;
    const uint32_t need = (n + 64u * kInferWaves - 1u) / (64u * kInferWaves);
    if (need == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_infer, dim3(std::min(grid, need)), dim3(64 * kInferWaves), 0, st, net, src,
                       (const uint32_t*)nullptr, (const uint32_t*)nullptr, 0u, n, out, raw,
                       (unsigned long long*)nullptr);
    return hipGetLastError();
}

hipError_t guide_infer_hits(const GuideNetDev& net, const float4* hits, const phg_ray* rays, const TriRec* tris,
                            const GuideScene& gs, const uint32_t* regions, const uint32_t* bounce_base, uint32_t bk,
                            float* out, unsigned long long* queries, uint32_t grid, hipStream_t st) {
    InferSource src{};
    src.hits = hits; src.rays = rays; src.tris = tris; src.gs = gs;
    hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(64 * kInferWaves), 0, st, net, src, regions, bounce_base, bk, 0u,
                       out, 0, queries);
    return hipGetLastError();
}

hipError_t guide_gradients(const GuideNetDev& net, const float* rec, uint32_t nrec, float* grads,
                           const GuideWorkspace& ws, hipStream_t st) {
    if (nrec == 0) return hipSuccess;
    if (nrec > ws.cap_records) return hipErrorInvalidValue;
    const uint32_t npad = (nrec + 31u) & ~31u;
    __bf16* X = reinterpret_cast<__bf16*>(ws.X);
    __bf16* dz[2] = {reinterpret_cast<__bf16*>(ws.dz0), reinterpret_cast<__bf16*>(ws.dz1)};
    const uint32_t fwd_blocks = std::min<uint32_t>(ws.max_blocks, (npad / 16 + 3) / 4);
    hipLaunchKernelGGL(k_train_forward, dim3(fwd_blocks), dim3(256), 0, st, net, rec, nrec, npad, 1.0f / (float)nrec, X,
                       dz[0]);
    // backward: waves x chunk covers npad in whole 32-record sub-tiles, at least 4 per wave
    const uint32_t tiles = npad / 32;
    const uint32_t waves = std::max<uint32_t>(1u, std::min<uint32_t>(ws.max_parts, tiles / 4));
    const uint32_t chunk = ((tiles + waves - 1) / waves) * 32;
    const uint32_t used = (npad + chunk - 1) / chunk;
    const uint32_t blocks = (used + kBackWaves - 1) / kBackWaves;
    int cur = 0;
    for (int l = kGuideLayers - 1; l >= 0; --l) {
        hipLaunchKernelGGL(k_train_backward, dim3(blocks), dim3(64 * kBackWaves), 0, st,
                           reinterpret_cast<const __bf16*>(net.wn) + l * 4096, X + (size_t)l * npad * 64, dz[cur],
                           dz[cur ^ 1], npad, chunk, l, ws.partial + (size_t)l * (ws.max_parts + 4) * (64 * 64 + 64));
        cur ^= 1;
    }
    const uint32_t eblocks = (64 * 64 + 64 + 255) / 256;
    hipLaunchKernelGGL(k_train_reduce_parts, dim3(eblocks, ws.slices, kGuideLayers), dim3(256), 0, st, ws.partial,
                       (size_t)(ws.max_parts + 4) * (64 * 64 + 64), blocks, ws.partial2);
    hipLaunchKernelGGL(k_train_reduce, dim3(eblocks, 1, kGuideLayers), dim3(256), 0, st, ws.partial2, ws.slices, grads);
    return hipGetLastError();
}

hipError_t guide_adam(const GuideNetDev& net, float* grads, float* m, float* v, float lr, int step, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_guide_adam, dim3((kGuideParams + 255) / 256), dim3(256), 0, st, net.params_mut, grads, m, v, lr,
                       b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t guide_fill(float* p, uint32_t n, float v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_float, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace phg
