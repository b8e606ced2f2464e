// 4-wide BVH traversal of one ray (host and device): 128-byte node fetches, child boxes
// tested against the shrinking segment with the reference's robust slab test
// (Math/Geometry/TAABB3D.ipp:390-420) widened by the watertight test's error bound,
// near-first order, leaves tested with the watertight triangle test.
// Traversal stack: the newest RING entries live in a ring addressed as stk[slot * stride]
// (LDS on the device, laid out [slot][lane]: bank-conflict free); older entries spill to a
// per-thread global area of spill_cap entries, so any tree depth stays exact while the LDS
// footprint per lane stays small (occupancy).  Closest hit keeps shrinking the segment to the
// accepted hit as TLinearDepthFirstBinaryBvh::nearestTraversal does
// (Math/Algorithm/BVH/TLinearDepthFirstBinaryBvh.ipp:43-139); any hit returns at the first.
#pragma once

#include "pt_device.h"

namespace phg {

struct TraceResult { float t; int32_t tri; float a, b, c; };

template <int RING>
struct TraversalStack {
    int32_t* ring;        // RING slots, slot k at ring[k * stride]
    uint32_t stride;
    int32_t* spill;       // spill_cap entries
    uint32_t spill_cap;
    int sp;
    PHG_HD void push(int32_t v, uint32_t* overflow) {
        if (sp >= RING) {
            const int j = sp - RING;  // oldest ring entry moves to the spill area
            if ((uint32_t)j < spill_cap) spill[j] = ring[(j & (RING - 1)) * stride];
            else *overflow = 1u;
        }
        ring[(sp & (RING - 1)) * stride] = v;
        ++sp;
    }
    PHG_HD int32_t pop() {
        --sp;
        const int32_t v = ring[(sp & (RING - 1)) * stride];
        if (sp >= RING) {
            const int j = sp - RING;  // bring the newest spilled entry back into the freed slot
            if ((uint32_t)j < spill_cap) ring[(sp & (RING - 1)) * stride] = spill[j];
        }
        return v;
    }
};

// Traversal of one ray as a resumable state machine: begin() sets the ray up, each step()
// visits one inner node or one leaf and returns true once the ray is finished (result in r).
// Kernels that stream many rays per lane start the lane's next ray as soon as the current one
// finishes, instead of waiting for the slowest ray of the wave.
template <int RING, bool ANY, bool COUNT>
struct Traversal {
    V3 o, d;
    // slab setup, axes permuted so that axis 0 is the dominant axis of d: origin and
    // reciprocal direction per permuted axis, and the byte offsets (within a node) of the
    // near and far bound rows of each permuted axis -- the rows of the low bounds when the
    // direction component is >= 0, of the high bounds when it is negative
    float op0, op1, op2, rp0, rp1, rp2;
    uint32_t nr0, fr0, nr1, fr1, nr2, fr2;
    float tmin, tb;
    int32_t cur;
    TraceResult r;
    TraversalStack<RING> st;

    PHG_HD void attach(int32_t* stk, uint32_t stride, int32_t* spill, uint32_t spill_cap) {
        st.ring = stk; st.stride = stride; st.spill = spill; st.spill_cap = spill_cap; st.sp = 0;
    }

    PHG_HD void begin(V3 o_, V3 d_, float tmin_, float tmax_, int32_t root) {
        o = o_; d = d_; tmin = tmin_; tb = tmax_;
        // dominant axis of the direction, chosen as the watertight test permutes it
        const float adx = fabsf(d.x), ady = fabsf(d.y), adz = fabsf(d.z);
        const int kdom = adx > ady ? (adx > adz ? 0 : 2) : (ady > adz ? 1 : 2);
        // permuted axes: kdom, then the other two in increasing order
        const uint32_t a0 = (uint32_t)kdom, a1 = kdom == 0 ? 1u : 0u, a2 = kdom == 2 ? 1u : 2u;
        op0 = kdom == 0 ? o.x : (kdom == 1 ? o.y : o.z);
        op1 = kdom == 0 ? o.y : o.x;
        op2 = kdom == 2 ? o.y : o.z;
        const float d0 = kdom == 0 ? d.x : (kdom == 1 ? d.y : d.z);
        const float d1 = kdom == 0 ? d.y : d.x;
        const float d2 = kdom == 2 ? d.y : d.z;
        rp0 = 1.0f / d0; rp1 = 1.0f / d1; rp2 = 1.0f / d2;
        // TAABB3D.ipp:405-407: the near slab is the low bound unless the component is negative
        nr0 = 32u * a0 + (d0 < 0.0f ? 16u : 0u); fr0 = 32u * a0 + (d0 < 0.0f ? 0u : 16u);
        nr1 = 32u * a1 + (d1 < 0.0f ? 16u : 0u); fr1 = 32u * a1 + (d1 < 0.0f ? 0u : 16u);
        nr2 = 32u * a2 + (d2 < 0.0f ? 16u : 0u); fr2 = 32u * a2 + (d2 < 0.0f ? 0u : 16u);
        r.t = 0.0f; r.tri = -1; r.a = 0.0f; r.b = 0.0f; r.c = 0.0f;
        st.sp = 0;
        cur = root;
    }

    PHG_HD bool step(const WideNode* __restrict__ nodes, const TriRec* __restrict__ tris, uint32_t* nvis,
                     uint32_t* ntri, uint32_t* overflow) {
        constexpr float mult = 1.1920928955078125e-07f * 2.0f + 1.0f;  // TAABB3D.ipp:417
        if (cur >= 0) {
            // node rows: lx, hx, ly, hy, lz, hz (4 children each), children; the near and far
            // rows of each axis are fetched directly through the ray's row offsets
            const char* nb = reinterpret_cast<const char*>(nodes);
            const uint32_t off = (uint32_t)cur * (uint32_t)sizeof(WideNode);
            const float4 n0 = *reinterpret_cast<const float4*>(nb + (off + nr0));
            const float4 f0 = *reinterpret_cast<const float4*>(nb + (off + fr0));
            const float4 n1 = *reinterpret_cast<const float4*>(nb + (off + nr1));
            const float4 f1 = *reinterpret_cast<const float4*>(nb + (off + fr1));
            const float4 n2 = *reinterpret_cast<const float4*>(nb + (off + nr2));
            const float4 f2 = *reinterpret_cast<const float4*>(nb + (off + fr2));
            const int4 ch = *reinterpret_cast<const int4*>(nb + (off + 96u));
            if (COUNT) (*nvis)++;
            const float near0[4] = {n0.x, n0.y, n0.z, n0.w}, far0[4] = {f0.x, f0.y, f0.z, f0.w};
            const float near1[4] = {n1.x, n1.y, n1.z, n1.w}, far1[4] = {f1.x, f1.y, f1.z, f1.w};
            const float near2[4] = {n2.x, n2.y, n2.z, n2.w}, far2[4] = {f2.x, f2.y, f2.z, f2.w};
            const int32_t cc[4] = {ch.x, ch.y, ch.z, ch.w};
            float key[4];
            int32_t kid[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                // robust slab (TAABB3D.ipp:390-420): entry t0 over [tmin, ...), exit t1.
                // fmaxf / fminf return the other operand when one is NaN, which is the
                // reference's "safe max / min"; the axis order does not change the result.
                const float m0 = (near0[k] - op0) * rp0, M0 = (far0[k] - op0) * rp0;
                const float m1 = (near1[k] - op1) * rp1, M1 = (far1[k] - op1) * rp1;
                const float m2 = (near2[k] - op2) * rp2, M2 = (far2[k] - op2) * rp2;
                const float t0 = fmaxf(fmaxf(fmaxf(tmin, m0), m1), m2);
                const float t1 = fminf(fminf(fminf(kFltMax, M0), M1), M2);
                // The watertight t of a triangle in this box is a convex combination of its
                // vertices' distances along the dominant axis, so its rounding error scales
                // with the box's parametric span on that axis, not with t itself: widen the
                // test against the current closest hit by 2^-19 of that span.
                const float slack = fmaxf(fabsf(m0), fabsf(M0)) * 0x1p-19f;
                // t0 <= t1 * mult + slack and t0 <= tb + slack, as one comparison: adding the
                // slack rounds monotonically, so min(a + s, b + s) = min(a, b) + s exactly, and
                // t1 is never NaN (the fminf chain starts at FLT_MAX)
                const bool hit = (t0 <= fminf(t1 * mult, tb) + slack) && cc[k] != kEmptyChild;
                key[k] = hit ? t0 : INFINITY;
                kid[k] = cc[k];
            }
            // sorting network on (key, kid), ascending
#define PHG_CSWAP(i, j)                                                   \
    if (key[j] < key[i]) {                                                \
        const float tk = key[i]; key[i] = key[j]; key[j] = tk;            \
        const int32_t ti = kid[i]; kid[i] = kid[j]; kid[j] = ti;          \
    }
            if (ANY) {
                // any hit: order does not matter; move the hits to the front
                PHG_CSWAP(0, 1) PHG_CSWAP(2, 3) PHG_CSWAP(0, 2) PHG_CSWAP(1, 3)
            } else {
                PHG_CSWAP(0, 1) PHG_CSWAP(2, 3) PHG_CSWAP(0, 2) PHG_CSWAP(1, 3) PHG_CSWAP(1, 2)
            }
#undef PHG_CSWAP
            if (key[0] == INFINITY) {
                if (st.sp == 0) return true;
                cur = st.pop();
                return false;
            }
#pragma unroll
            for (int k = 3; k >= 1; --k) {
                if (key[k] != INFINITY) st.push(kid[k], overflow);
            }
            cur = kid[0];
            return false;
        }
        // one triangle of the leaf per step, so that a wave whose lanes are split between
        // inner nodes and leaves pays for one triangle test per iteration, not a whole leaf;
        // the rest of the leaf stays in cur (leaf flag, triangles left - 1, next triangle)
        const uint32_t u = (uint32_t)cur;
        const uint32_t first = u & ((1u << kLeafFirstBits) - 1u);
        const uint32_t rest = (u >> kLeafFirstBits) & 7u;
        if (COUNT) (*ntri)++;
        {
            V3 a, b, c;
            load_tri(tris, first, &a, &b, &c);
            float t, ba, bb, bc;
            if (wt_intersect(a, b, c, o, d, tmin, tb, &t, &ba, &bb, &bc)) {
                tb = t;
                r.t = t; r.tri = (int32_t)first; r.a = ba; r.b = bb; r.c = bc;
                if (ANY) return true;
            }
        }
        if (rest > 0) {
            cur = (int32_t)((u & 0x80000000u) | ((rest - 1u) << kLeafFirstBits) | (first + 1u));
            return false;
        }
        if (st.sp == 0) return true;
        cur = st.pop();
        return false;
    }
};

template <int RING, bool ANY, bool COUNT>
PHG_HD TraceResult traverse(const WideNode* __restrict__ nodes, const TriRec* __restrict__ tris, int32_t root,
                            V3 o, V3 d, float tmin, float tmax, int32_t* __restrict__ stk, uint32_t stride,
                            int32_t* __restrict__ spill, uint32_t spill_cap,
                            uint32_t* nvis, uint32_t* ntri, uint32_t* overflow) {
    static_assert((RING & (RING - 1)) == 0, "ring size must be a power of two");
    Traversal<RING, ANY, COUNT> tv;
    tv.attach(stk, stride, spill, spill_cap);
    tv.begin(o, d, tmin, tmax, root);
    while (!tv.step(nodes, tris, nvis, ntri, overflow)) {
    }
    return tv.r;
}


}  // namespace phg
