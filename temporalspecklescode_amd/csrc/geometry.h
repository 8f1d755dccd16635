// geometry.h -- medium lookups, free flight, scattering sampling and the NEE voxel walk.
// Device functions templated on the real type R (float for transport and the FP32 NEE
// variant, double for the FP64 NEE variant) and on DIMS (2 or 3).
#pragma once
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "common.h"

// Hook on every walk-grid cell index whose word the fast walk reads; empty in the library. A
// host-side test defines it to check that the word lies inside the padded grid.
#ifndef SSTMC_WALK_INDEX_CHECK
#define SSTMC_WALK_INDEX_CHECK(m, cell)
#endif

namespace sstmc {

template <typename R> struct Real;
template <> struct Real<float> {
    typedef f3 V;
    __host__ __device__ static const float* edges(const MediumView& m, int a) { return m.ef[a]; }
    __host__ __device__ static float sigt(const MediumView& m, int i) { return m.sigtf[i]; }
    __host__ __device__ static float bmin(const MediumView& m, int a) { return m.bminf[a]; }
    __host__ __device__ static float bmax(const MediumView& m, int a) { return m.bmaxf[a]; }
};
template <> struct Real<double> {
    typedef d3 V;
    __host__ __device__ static const double* edges(const MediumView& m, int a) { return m.ed[a]; }
    __host__ __device__ static double sigt(const MediumView& m, int i) { return m.sigtd[i]; }
    __host__ __device__ static double bmin(const MediumView& m, int a) { return m.bmind[a]; }
    __host__ __device__ static double bmax(const MediumView& m, int a) { return m.bmaxd[a]; }
};

__device__ inline float vget(const f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ inline double vget(const d3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// MediumNS::isInsideKernel (MediumInterface.cuh:204-213): closed box.
template <int DIMS, typename R, typename V>
__host__ __device__ inline bool inside(const MediumView& m, const V& p) {
    bool in = p.x >= Real<R>::bmin(m, 0) && p.x <= Real<R>::bmax(m, 0) &&
              p.y >= Real<R>::bmin(m, 1) && p.y <= Real<R>::bmax(m, 1);
    if (DIMS == 3) in = in && p.z >= Real<R>::bmin(m, 2) && p.z <= Real<R>::bmax(m, 2);
    return in;
}

// Cell index along one axis with the semantics of binarySearchKernel (Simulation.cuh:325-367):
// e[b] <= x < e[b+1], clamped to a valid cell. Equally spaced axes take a multiply and at most a
// one-cell correction against the stored edges, so both paths return the same cell.
template <typename R>
__host__ __device__ inline int locate(const MediumView& m, int a, R x) {
    const R* e = Real<R>::edges(m, a);
    int n = (int)m.n[a];
    int b;
    if (m.uniformAxis[a]) {
        float f = ((float)x - m.e0[a]) * m.invH[a];
        b = f <= 0.f ? 0 : (f >= m.nf[a] ? n - 1 : (int)f);
        if (b > 0 && x < e[b]) --b;
        if (b < n - 1 && x >= e[b + 1]) ++b;
        return b;
    }
    if (x < e[0]) return 0;
    if (x >= e[n]) return n - 1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (e[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// locate for an equally spaced axis (m.uniformAxis[a]): the same cell as locate, without the
// general binary search (the fast walk only runs on grids whose every axis is equally spaced).
template <typename R>
__host__ __device__ inline int locateUniform(const MediumView& m, int a, R x) {
    const R* e = Real<R>::edges(m, a);
    const int n = (int)m.n[a];
    const float f = ((float)x - m.e0[a]) * m.invH[a];
    int b = f <= 0.f ? 0 : (f >= m.nf[a] ? n - 1 : (int)f);
    if (b > 0 && x < e[b]) --b;
    if (b < n - 1 && x >= e[b + 1]) ++b;
    return b;
}

__host__ __device__ inline uint32_t cellIndex(const MediumView& m, int bx, int by, int bz) {
    return (uint32_t)bx + m.n[0] * ((uint32_t)by + m.n[1] * (uint32_t)bz);
}

// bdDistance (HeterogeneousMedium.cuh:269-383) reduced to its use in Woodcock tracking: is the
// ray from a point outside the box heading into it.
template <int DIMS>
__device__ inline bool pointsToward(const MediumView& m, const f3& p, const f3& d) {
    float r[3], o[3], dd[3] = {d.x, d.y, d.z}, pp[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r[a] = 0.5f * (m.bmaxf[a] - m.bminf[a]);
        o[a] = pp[a] - 0.5f * (m.bmaxf[a] + m.bminf[a]);
    }
    float mx = -3.0e38f;
#pragma unroll
    for (int a = 0; a < DIMS; ++a) {
        float ir = fabsf(r[a]) < 1e-12f ? 1e12f : 1.0f / r[a];
        mx = fmaxf(mx, fabsf(o[a]) * ir);
    }
    int winding = mx <= 1.0f ? -1 : 1;
#pragma unroll
    for (int a = 0; a < DIMS; ++a) {
        float inv = fabsf(dd[a]) < 1e-12f ? 1e12f : 1.0f / dd[a];
        int sgn = dd[a] < 0.f ? 1 : -1;
        float dist = inv * (r[a] * (float)(winding * sgn) - o[a]);
        bool ok = dist >= 0.0f;
#pragma unroll
        for (int b = 0; b < DIMS; ++b)
            if (b != a) ok = ok && fabsf(fmaf(dd[b], dist, o[b])) < r[b];
        if (ok) return true;
    }
    return false;
}

template <int DIMS>
__device__ inline uint32_t materialAt(const MediumView& m, const f3& x) {
    int bx = locate<float>(m, 0, x.x);
    int by = locate<float>(m, 1, x.y);
    int bz = DIMS == 3 ? locate<float>(m, 2, x.z) : 0;
    return m.grid[cellIndex(m, bx, by, bz)];
}

// heterogeneousSampleKernel (HeterogeneousMedium.cuh:576-678): Woodcock tracking against the
// global majorant, absorption as termination. Returns the material of the sampled scatterer or 0
// (escaped / absorbed). Distances use -ln(u) with u uniform on (0,1) (the reference writes
// -ln(1-u); same law).
template <int DIMS>
__device__ inline uint32_t woodcock(const MediumView& m, const f3& o, const f3& d, Rng& rng, f3* out) {
    float dist = 0.f;
    f3 x = o;
    uint32_t mat = 0;
    for (;;) {
        dist += -__logf(rng.uniform()) * m.invMaj;
        x = mk(fmaf(dist, d.x, o.x), fmaf(dist, d.y, o.y), DIMS == 3 ? fmaf(dist, d.z, o.z) : 0.f);
        if (!inside<DIMS, float>(m, x)) {
            if (!pointsToward<DIMS>(m, x, d)) break;
        } else {
            mat = materialAt<DIMS>(m, x);
            if (rng.uniform() <= m.accept[mat]) break;
        }
    }
    *out = x;
    if (!inside<DIMS, float>(m, x)) return 0;
    if (rng.uniform() > m.albedo[mat]) return 0;
    return mat;
}

// randomDirection (VectorType.cuh:316-362).
template <int DIMS>
__device__ inline f3 randomDirection(Rng& rng) {
    if (DIMS == 2) {
        float s, c;
        sincospif(2.0f * rng.uniform(), &s, &c);
        return mk(s, c, 0.f);
    }
    float sp, cp;
    sincospif(2.0f * rng.uniform(), &sp, &cp);
    float ct = fmaf(rng.uniform(), 2.0f, -1.0f);
    float st = sqrtf(fmaxf(0.f, 1.0f - ct * ct));
    return mk(st * cp, st * sp, ct);
}

// random_vMF_cosine (VectorType.cuh:169-205), 3D: Wood's rejection sampler for the cosine of a
// von Mises-Fisher law of concentration kappa, in FP64.
__device__ inline double vmfCosine3(double kappa, Rng& rng) {
    const double b = -kappa + hypot(kappa, 1.0);
    const double x0 = (1.0 - b) / (1.0 + b);
    const double sc = kappa * x0 + 2.0 * log(1.0 - x0 * x0);
    double w, t, u;
    do {
        const double z = rng.uniform();
        u = rng.uniform();
        w = (1.0 - (1.0 + b) * z) / (1.0 - (1.0 - b) * z);
        t = kappa * w + 2.0 * log(1.0 - x0 * w) - sc;
    } while (t < log(u));
    return w;
}

// sampleDeviceImplementation (Scattering.cuh:1091-1109): HG inverse CDF (:1039-1063) and
// vector_rotation (VectorType.cuh:261-282) about the incoming direction. TAB = false compiles the
// tabular and vMF samplers out (no material uses them).
template <int DIMS, bool TAB = true>
__device__ inline f3 sampleDirection(const ScatterView& s, uint32_t mat, const f3& prev, Rng& rng) {
    if (s.type[mat] == 1 || (TAB && DIMS == 2 && s.type[mat] == 4)) return randomDirection<DIMS>(rng);
    float g = s.g[mat], c;
    if (TAB && s.type[mat] == 4) {
        // vMF mixture: a pair by its cumulative weight, then its cosine (Scattering.cuh:1066-1077)
        const PhaseTables& t = *s.tab;
        const uint32_t j = t.vmfPairOff[mat] +
                           cdfSearch(t.vmfPairCdf + t.vmfPairOff[mat], t.vmfM[mat] * t.vmfM[mat], (double)rng.uniform());
        c = (float)vmfCosine3(t.vmfPairMu[j], rng);
    } else if (TAB && s.type[mat] == 2) {
        // tabular: the inverse cdf at round(u n) (Scattering.cuh:1030-1037); u has 25 significant
        // bits, so u n is exact in FP64 and the entry is the one an FP64 evaluation picks
        const PhaseTables& t = *s.tab;
        const uint32_t n = t.n[mat];
        const double x = rint((double)rng.uniform() * (double)n);
        const uint32_t i = x >= (double)n ? n - 1 : (uint32_t)x;
        c = cosf(t.icdf[t.off[mat] + i]);
    } else if (fabsf(g) < 0.001f) {
        c = rng.uniform() * 2.0f - 1.0f;
    } else {
        float u = rng.uniform();
        if (DIMS == 2) {
            c = cosf(2.0f * atanf(((1.0f - g) / (1.0f + g)) * tanf(1.57079632679489662f * (1.0f - 2.0f * u))));
        } else {
            float sv = (1.0f - g * g) / (1.0f - g + 2.0f * g * u);
            c = (1.0f + g * g - sv * sv) / (2.0f * g);
        }
    }
    c = fminf(1.0f, fmaxf(-1.0f, c));
    float sa = (rng.uniform() < 0.5f ? -1.0f : 1.0f) * sqrtf(fmaxf(0.f, 1.0f - c * c));
    if (DIMS == 2) return mk(prev.x * c - prev.y * sa, prev.x * sa + prev.y * c, 0.f);
    float sp, cp;
    sincospif(2.0f * rng.uniform(), &sp, &cp);
    float az = fabsf(prev.z);
    if (az >= 1.0f) return mk(sa * cp, sa * sp, prev.z < 0.f ? -c : c);
    float stw = sqrtf((1.0f - az) * (1.0f + az));
    float inv = 1.0f / stw;
    return mk(sa * (prev.x * prev.z * cp - prev.y * sp) * inv + prev.x * c,
              sa * (prev.y * prev.z * cp + prev.x * sp) * inv + prev.y * c,
              -sa * cp * stw + prev.z * c);
}

// ---------------------------------------------------------------- NEE voxel walk
// heterogeneousAttenuationKernel (HeterogeneousMedium.cuh:386-573): exact walk from p along d to
// the box boundary, tau = sum(segment * sigt). The crossing parameter of each face is computed
// from the fixed origin p (one rounding per crossing instead of an accumulated position) and the
// axis choice keeps the reference's strict-< tie order (x, then y, else z). The walk stops early
// once tau > tauCut; callers pass a cut beyond which the result they form is exactly zero in R.
// `sig` is the per-material sigt table (an LDS copy in the kernels).
template <int DIMS, typename R, typename V>
__host__ __device__ inline R opticalDepth(const MediumView& m, const R* sig, const V& p, const V& d, R tauCut,
                                          uint32_t* steps) {
    if (!inside<DIMS, R>(m, p)) return (R)0;
    const R* e0 = Real<R>::edges(m, 0);
    const R* e1 = Real<R>::edges(m, 1);
    const R* e2 = Real<R>::edges(m, 2);
    int b0 = locate<R>(m, 0, p.x), b1 = locate<R>(m, 1, p.y), b2 = DIMS == 3 ? locate<R>(m, 2, p.z) : 0;
    const R big = (R)1e12;
    R i0 = fabs(d.x) < (R)1e-12 ? (d.x >= 0 ? big : -big) : (R)1 / d.x;
    R i1 = fabs(d.y) < (R)1e-12 ? (d.y >= 0 ? big : -big) : (R)1 / d.y;
    R i2 = DIMS == 3 ? (fabs(d.z) < (R)1e-12 ? (d.z >= 0 ? big : -big) : (R)1 / d.z) : big;
    int s0 = d.x < 0 ? -1 : 1, s1 = d.y < 0 ? -1 : 1, s2 = (DIMS == 3 && d.z < 0) ? -1 : 1;
    int u0 = s0 > 0, u1 = s1 > 0, u2 = s2 > 0;
    int n0 = (int)m.n[0], n1 = (int)m.n[1], n2 = (int)m.n[2];
    R t0 = (e0[b0 + u0] - p.x) * i0;
    R t1 = (e1[b1 + u1] - p.y) * i1;
    R t2 = DIMS == 3 ? (e2[b2 + u2] - p.z) * i2 : (R)3e38;
    R tcur = 0, tau = 0;
    R st = sig[m.grid[cellIndex(m, b0, b1, b2)]];
    uint32_t cnt = 0;
    for (;;) {
        ++cnt;
        int ax;
        if (DIMS == 2) ax = t0 < t1 ? 0 : 1;
        else ax = (t0 < t1 && t0 < t2) ? 0 : (t1 < t2 ? 1 : 2);
        R tn = ax == 0 ? t0 : (ax == 1 ? t1 : t2);
        tau += (tn - tcur) * st;
        tcur = tn;
        if (ax == 0) {
            b0 += s0;
            if (b0 < 0 || b0 >= n0) break;
            t0 = (e0[b0 + u0] - p.x) * i0;
        } else if (ax == 1) {
            b1 += s1;
            if (b1 < 0 || b1 >= n1) break;
            t1 = (e1[b1 + u1] - p.y) * i1;
        } else {
            b2 += s2;
            if (b2 < 0 || b2 >= n2) break;
            t2 = (e2[b2 + u2] - p.z) * i2;
        }
        if (tau > tauCut) break;
        st = sig[m.grid[cellIndex(m, b0, b1, b2)]];
    }
    if (steps) *steps += cnt;
    return tau;
}

// The `bits`-bit field at bit (cell mod 32) of a little-endian word. On the device one bit-field
// extract, whose offset operand is taken mod 32, so the walk's bit offset needs no mask.
__host__ __device__ inline uint32_t cellField(uint32_t word, uint32_t cell, uint32_t bits) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ubfe(word, cell, bits);
#else
    return (word >> (cell & 31u)) & ((1u << bits) - 1u);
#endif
}

// Reads of the walk grid's 32-bit words by word index. On the device a buffer load with index
// addressing (descriptor stride 4 bytes, range = the grid's words), so the word index (bit offset
// >> 5) is the whole address computation; on the host a plain array read.
#if defined(__HIP_DEVICE_COMPILE__)
typedef int32_t sstmc_i32x4 __attribute__((ext_vector_type(4)));
__device__ int32_t sstmc_struct_buffer_load_i32(sstmc_i32x4 rsrc, int32_t vindex, int32_t voffset, int32_t soffset,
                                                int32_t aux) __asm("llvm.amdgcn.struct.buffer.load.i32");
struct WalkWords {
    sstmc_i32x4 rsrc;
    __device__ explicit WalkWords(const MediumView& m) {
        const uint64_t a = (uint64_t)m.walkGrid;
        rsrc.x = (int32_t)(uint32_t)a;
        rsrc.y = (int32_t)(((uint32_t)(a >> 32) & 0xffffu) | (4u << 16));   // base[47:32], stride 4
        rsrc.z = (int32_t)m.walkWords;                                      // records
        rsrc.w = 0x00020000;                                                // 32-bit data format
    }
    __device__ uint32_t operator()(uint32_t bit) const {
        return (uint32_t)sstmc_struct_buffer_load_i32(rsrc, (int32_t)(bit >> 5), 0, 0, 0);
    }
};
#else
struct WalkWords {
    const uint32_t* w;
    explicit WalkWords(const MediumView& m) : w(m.walkGrid) {}
    uint32_t operator()(uint32_t bit) const { return w[bit >> 5]; }
};
#endif

// Branch-free FP32 walk for equally spaced grids (the common case). Per step: the nearest face
// crossing is min3 of the three pending crossings; the axis is the highest one whose crossing
// equals it, which is the reference's strict-< tie order (x only when strictly nearest, then y,
// else z). Add segment * sigt, advance the cell index by that axis' stride. The crossing
// parameter of the k-th face along an axis is t_first + k * t_delta (one rounding, no
// accumulation), recomputed for every axis from its crossing count. The walk ends at the first
// crossing at or beyond tExit, the parameter of the box face the ray leaves through, computed by
// the same expression as the crossings, so the exit test needs no per-axis counters and does not
// wait for memory.
//
// Memory: the walk grid (m.walkGrid: each cell a 1-, 4- or 8-bit field holding 4 x its index in
// the palette of the materials present, i.e. the byte offset of its sigt in the table `sig`, in
// LDS in the kernels; axes ordered with the dominant walk direction fastest) is read one aligned
// 4-byte word at a time; a 128^3 grid of four materials packs into 1 MB. The walk's cell position
// is a bit offset; each walk keeps the word
// holding its current cell and reads a new one only when its next cell lies in another word, i.e.
// after a step along a slow axis or every 32 / bits steps along the fast one, so the 64 walks of a
// wavefront issue few reads and touch few cache lines per step. The offset includes the grid's
// leading padding, so the cell one step past any face is still inside the allocation and never
// below 0 (that word is read but not used).
//
// The stages are software-pipelined one step apart: step j reads the word of cell j+1 (when it
// changes), takes cell j's field from its word and looks up its sigt, and adds segment(j-1) *
// sigt(j-1); each read has a whole step of independent work to hide behind. The sum is formed in
// the same order as a sequential walk and the tauCut test sees the same partial sums, so the
// result is bit-identical to one.
// The direction-only part of the walk's set-up: per axis the inverse direction, the parameter
// step between faces, the walk-grid bit stride (two's complement) and the stepping sense. A
// kernel whose lanes each walk one direction from many points forms it once per lane.
struct WalkDir {
    float inv[3], td[3];
    uint32_t stride[3];
    int up[3];
};
template <int DIMS>
__host__ __device__ inline WalkDir walkDir(const MediumView& m, const f3& d) {
    const float dd[3] = {d.x, d.y, DIMS == 3 ? d.z : 0.f};
    WalkDir w;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w.inv[a] = fabsf(dd[a]) < 1e-12f ? (dd[a] >= 0.f ? 1e12f : -1e12f) : 1.0f / dd[a];
        w.up[a] = dd[a] >= 0.f;
        w.td[a] = m.h[a] * fabsf(w.inv[a]);
        w.stride[a] = w.up[a] ? m.walkStride[a] : 0u - m.walkStride[a];
    }
    return w;
}

template <int DIMS>
__host__ __device__ inline float opticalDepthFastDir(const MediumView& m, const float* sig, const f3& p,
                                                     const WalkDir& wd, float tauCut, uint32_t* steps);

template <int DIMS>
__host__ __device__ inline float opticalDepthFast(const MediumView& m, const float* sig, const f3& p, const f3& d,
                                                  float tauCut, uint32_t* steps) {
    return opticalDepthFastDir<DIMS>(m, sig, p, walkDir<DIMS>(m, d), tauCut, steps);
}

// The walk's start at p: per axis the cell index b, the offsets of the cell's low and high faces
// from p (FP32), and the walk-grid bit offset of the cell. Per vertex, independent of the
// direction; the far-field view kernel reads it from a per-vertex record (FarRec).
struct WalkStart {
    float dlo[3], dhi[3], b[3];
    uint32_t cell;
};
template <int DIMS>
__host__ __device__ inline WalkStart walkStart(const MediumView& m, const f3& p) {
    const float pp[3] = {p.x, p.y, DIMS == 3 ? p.z : 0.f};
    WalkStart s;
    s.cell = m.walkOrigin;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a >= DIMS) {
            s.dlo[a] = s.dhi[a] = s.b[a] = 0.f;
            continue;
        }
        const int b = locateUniform<float>(m, a, pp[a]);   // the fast walk runs on equally spaced grids
        // both faces of the cell (edges hold n + 1 entries): when p is shared by a wavefront's
        // walks these are two scalar loads rather than a gather
        s.dlo[a] = m.ef[a][b] - pp[a];
        s.dhi[a] = m.ef[a][b + 1] - pp[a];
        s.b[a] = (float)b;
        s.cell += (uint32_t)b * m.walkStride[a];
    }
    return s;
}

template <int DIMS>
__host__ __device__ inline float walkFastCore(const MediumView& m, const float* sig, const float* tf,
                                              const float* td, const uint32_t* stride, uint32_t cell, float tExit,
                                              float tauCut, uint32_t* steps);

// The walk from a start (walkStart) in the direction part wd (walkDir). Per axis the first face
// crossing ahead is (face ahead - p) / d, and the box is left at the crossing of the face
// n - 1 - b (up) or b (down) faces further on.
template <int DIMS>
__host__ __device__ inline float walkFromStart(const MediumView& m, const float* sig, const float* dlo,
                                               const float* dhi, const float* bf, uint32_t cell, const WalkDir& wd,
                                               float tauCut, uint32_t* steps) {
    float tf[3] = {3.0e38f, 3.0e38f, 3.0e38f}, td[3] = {0.f, 0.f, 0.f};
    uint32_t stride[3] = {0u, 0u, 0u};   // walk-grid bit strides in two's complement
    float tExit = 3.0e38f;
#pragma unroll
    for (int a = 0; a < DIMS; ++a) {
        const int up = wd.up[a];
        tf[a] = (up ? dhi[a] : dlo[a]) * wd.inv[a];
        td[a] = wd.td[a];
        // faces crossed before the one that leaves the box along this axis (exact integers)
        const float ahead = up ? (float)(m.n[a] - 1) - bf[a] : bf[a];
        tExit = fminf(tExit, fmaf(ahead, td[a], tf[a]));
        stride[a] = wd.stride[a];
    }
    return walkFastCore<DIMS>(m, sig, tf, td, stride, cell, tExit, tauCut, steps);
}

// walkFromStart for a lane that walks one direction from many starts: the direction's selects
// folded into per-axis factors once, so a start costs two multiply-adds and one multiply per axis.
// Bit-identical to walkFromStart: with up, fma(dhi, inv, dlo * 0) rounds dhi * inv once and adds a
// zero; without, fma(dhi, 0, dlo * inv) adds a zero to the rounded dlo * inv (dhi >= 0 >= dlo and
// the sign of inv is the sense, so no product is a negative zero that could flip a +0 result); and
// the faces ahead, -b + (n - 1) or b + 0, are exact small integers either way.
struct WalkLane {
    WalkDir d;
    float invUp[3], invDn[3], aheadSgn[3], aheadBase[3];
};
template <int DIMS>
__host__ __device__ inline WalkLane walkLane(const MediumView& m, const f3& dir) {
    WalkLane w;
    w.d = walkDir<DIMS>(m, dir);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w.invUp[a] = w.d.up[a] ? w.d.inv[a] : 0.f;
        w.invDn[a] = w.d.up[a] ? 0.f : w.d.inv[a];
        w.aheadSgn[a] = w.d.up[a] ? -1.f : 1.f;
        w.aheadBase[a] = w.d.up[a] ? (float)(m.n[a] - 1) : 0.f;
    }
    return w;
}
template <int DIMS>
__host__ __device__ inline float walkFromStartLane(const MediumView& m, const float* sig, const float* dlo,
                                                   const float* dhi, const float* bf, uint32_t cell,
                                                   const WalkLane& wl, float tauCut, uint32_t* steps) {
    float tf[3] = {3.0e38f, 3.0e38f, 3.0e38f}, td[3] = {0.f, 0.f, 0.f};
    uint32_t stride[3] = {0u, 0u, 0u};
    float tExit = 3.0e38f;
#pragma unroll
    for (int a = 0; a < DIMS; ++a) {
        tf[a] = fmaf(dhi[a], wl.invUp[a], dlo[a] * wl.invDn[a]);
        td[a] = wl.d.td[a];
        tExit = fminf(tExit, fmaf(fmaf(wl.aheadSgn[a], bf[a], wl.aheadBase[a]), td[a], tf[a]));
        stride[a] = wl.d.stride[a];
    }
    return walkFastCore<DIMS>(m, sig, tf, td, stride, cell, tExit, tauCut, steps);
}

// The walk from p with the direction part wd already formed (walkDir).
template <int DIMS>
__host__ __device__ inline float opticalDepthFastDir(const MediumView& m, const float* sig, const f3& p,
                                                     const WalkDir& wd, float tauCut, uint32_t* steps) {
    if (!inside<DIMS, float>(m, p)) return 0.f;
    const WalkStart s = walkStart<DIMS>(m, p);
    return walkFromStart<DIMS>(m, sig, s.dlo, s.dhi, s.b, s.cell, wd, tauCut, steps);
}

template <int DIMS>
__host__ __device__ inline float walkFastCore(const MediumView& m, const float* sig, const float* tf,
                                              const float* td, const uint32_t* stride, uint32_t cell, float tExit,
                                              float tauCut, uint32_t* steps) {
    const WalkWords words(m);
    const uint32_t bits = m.walkBits;
    const char* sigb = (const char*)sig;
    float t0 = tf[0], t1 = tf[1], t2 = tf[2];
    float k0 = 0.f, k1 = 0.f, k2 = 0.f;
    float tcur = 0.f, tau = 0.f;
    // One step. In: CELL and WORD (the word holding CELL, read last step or kept), SEGI/SIGI =
    // segment and sigt of the previous cell. Out: CELLN, WORDN (read issued when the word
    // changes), SEGO/SIGO for this cell, DONE when tau passed the cut or this cell is the last one
    // (tau is not changed after the cut test, so after the loop "tau <= tauCut" means the last
    // cell's segment is still to be added; it is added there, so no step waits for its own reads).
#define SSTMC_WALK_STEP(CELL, WORD, SEGI, SIGI, CELLN, WORDN, SEGO, SIGO, DONE)                    \
    {                                                                                               \
        const float tn = DIMS == 3 ? fminf(fminf(t0, t1), t2) : fminf(t0, t1);                      \
        const bool m2 = DIMS == 3 && tn == t2;                                                      \
        const bool m1 = !m2 && tn == t1;                                                            \
        const bool m0 = !m2 && !m1;                                                                 \
        CELLN = CELL + (m2 ? stride[2] : (m1 ? stride[1] : stride[0]));                             \
        WORDN = WORD;                                                                               \
        if ((CELLN ^ CELL) > 31u) {                                                                 \
            SSTMC_WALK_INDEX_CHECK(m, CELLN);                                                       \
            WORDN = words(CELLN);                                                                   \
        }                                                                                           \
        SEGO = tn - tcur;                                                                           \
        tcur = tn;                                                                                  \
        k0 += m0 ? 1.f : 0.f;                                                                       \
        k1 += m1 ? 1.f : 0.f;                                                                       \
        if (DIMS == 3) k2 += m2 ? 1.f : 0.f;                                                        \
        t0 = fmaf(k0, td[0], tf[0]);                                                                \
        t1 = fmaf(k1, td[1], tf[1]);                                                                \
        if (DIMS == 3) t2 = fmaf(k2, td[2], tf[2]);                                                 \
        SIGO = *(const float*)(sigb + cellField(WORD, CELL, bits));                                 \
        tau = fmaf(SEGI, SIGI, tau);  /* previous cell */                                           \
        DONE = (tau > tauCut) | (tn >= tExit);                                                      \
    }
    // two steps per iteration with the pipeline registers swapping roles (no register copies)
    SSTMC_WALK_INDEX_CHECK(m, cell);
    uint32_t ca = cell, cb, wa = words(cell), wb;
    float sga = 0.f, sia = 0.f, sgb, sib;       // nothing before the first cell: adds exactly 0
    float lastSeg = 0.f, lastSig = 0.f;
    bool done;
    for (;;) {
        SSTMC_WALK_STEP(ca, wa, sga, sia, cb, wb, sgb, sib, done)
        if (done) { lastSeg = sgb; lastSig = sib; break; }
        SSTMC_WALK_STEP(cb, wb, sgb, sib, ca, wa, sga, sia, done)
        if (done) { lastSeg = sga; lastSig = sia; break; }
    }
#undef SSTMC_WALK_STEP
    if (!(tau > tauCut)) tau = fmaf(lastSeg, lastSig, tau);   // left the box: last cell
    if (steps) *steps += (uint32_t)(k0 + k1 + k2);
    return tau;
}

// The skip threshold of the far-field view kernel for a walk cut `cut`: a vertex whose lower bound
// over a set of sources exceeds it has every term of those sources past the cut, whatever the FP32 walk's
// rounding (a relative 1e-3 margin over the walk's and the bound's own rounding), so skipping it
// leaves the terms as the walks would: zero.
__host__ __device__ inline float farSkip(float cut) { return cut + 1e-3f * (1.f + cut); }

// Least sigt over cells [b0, b1] x [c0, c1] of one layer: the smallest pyramid level at which the
// range spans at most 2 x 2 blocks, min over those blocks (a superset of the range: a lower bound).
__host__ __device__ inline float pyrMin(const MinPyramid& mp, uint32_t layer, uint32_t b0, uint32_t b1, uint32_t c0,
                               uint32_t c1) {
    uint32_t L = 0;
    while (L + 1 < mp.levels && ((b1 >> L) - (b0 >> L) > 1u || (c1 >> L) - (c0 >> L) > 1u)) ++L;
    const uint32_t nb = mp.nb[L];
    const float* q = mp.p + mp.off[L] + (uint64_t)layer * nb * mp.nc[L];
    const uint32_t bl = b0 >> L, bh = b1 >> L, cl = c0 >> L, ch = c1 >> L;
    return fminf(fminf(q[cl * nb + bl], q[cl * nb + bh]), fminf(q[ch * nb + bl], q[ch * nb + bh]));
}

// A lower bound on the optical depth from p to the box boundary along every direction of the set
// bd, stopping once past `stop`. Layer by layer along the set's axis (cells [i, i + 1] along it): a
// ray advancing along the axis crosses a layer's thickness over a path at least as long, at lateral
// positions p_q + s r_q with s in [near, far] (distances along the axis) and r_q within [rmin_q,
// rmax_q], i.e. inside a rectangle of cells whose least sigt bounds the layer's from below. A
// rectangle reaching the box's side ends the bound (rays may leave there); so does the box's face
// along the axis. Indices carry a 1e-3-cell margin against rounding, and the result is scaled by
// (1 - 1e-5). 0 for a point outside the box or no set.
__host__ __device__ inline float farLowerBound(const MediumView& m, const MinPyramid& mp, const BoundDirs& bd, const f3& pf,
                                      float stop) {
    if (!bd.valid || mp.axis < 0 || !inside<3, float>(m, pf)) return 0.f;
    const float p[3] = {pf.x, pf.y, pf.z};
    const int a = mp.axis, qs[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
    int layer = locateUniform<float>(m, a, p[a]);
    float near = 0.f, far = bd.sgn > 0 ? m.ef[a][layer + 1] - p[a] : p[a] - m.ef[a][layer];
    float lb = 0.f;
    for (;;) {
        uint32_t lo[2], hi[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int q = qs[j];
            const float xlo = p[q] + fminf(near * bd.rmin[j], far * bd.rmin[j]);
            const float xhi = p[q] + fmaxf(near * bd.rmax[j], far * bd.rmax[j]);
            const float mg = 1e-3f * m.h[q];
            if (xlo <= m.bminf[q] + mg || xhi >= m.bmaxf[q] - mg) return lb * (1.f - 1e-5f);
            const float flo = (xlo - m.e0[q]) * m.invH[q] - 1e-3f, fhi = (xhi - m.e0[q]) * m.invH[q] + 1e-3f;
            const int n1 = (int)m.n[q] - 1;
            const int il = (int)floorf(flo), ih = (int)floorf(fhi);
            lo[j] = (uint32_t)(il < 0 ? 0 : (il > n1 ? n1 : il));
            hi[j] = (uint32_t)(ih < 0 ? 0 : (ih > n1 ? n1 : ih));
        }
        lb = fmaf(far - near, pyrMin(mp, (uint32_t)layer, lo[0], hi[0], lo[1], hi[1]), lb);
        if (lb * (1.f - 1e-5f) > stop) break;
        layer += bd.sgn;
        if (layer < 0 || layer >= (int)m.n[a]) break;
        near = far;
        far += m.h[a];
    }
    return lb * (1.f - 1e-5f);
}

// Faces along axis a of sub-box k of cell c (sub boxes a cell along each axis): the cell's own
// faces at its ends, equal divisions inside.
__host__ __device__ inline void farSubBoxFaces(const MediumView& m, int a, uint32_t c, uint32_t k, uint32_t sub,
                                               float* lo, float* hi) {
    const float e0 = m.ef[a][c], e1 = m.ef[a][c + 1], d = (e1 - e0) / (float)sub;
    *lo = k == 0 ? e0 : e0 + (float)k * d;
    *hi = k + 1 == sub ? e1 : e0 + (float)(k + 1) * d;
}

// Least sigt of layer `layer` over cells [b0, b1] x [c0, c1] from the finest pyramid level at which
// the range covers at most span + 1 blocks along each axis (span 1: at most four lookups).
__host__ __device__ inline float pyrMinSpan(const MinPyramid& mp, uint32_t layer, uint32_t b0, uint32_t b1,
                                            uint32_t c0, uint32_t c1, uint32_t span) {
    uint32_t L = 0;
    while (L + 1 < mp.levels && ((b1 >> L) - (b0 >> L) > span || (c1 >> L) - (c0 >> L) > span)) ++L;
    const uint32_t nb = mp.nb[L];
    const float* q = mp.p + mp.off[L] + (uint64_t)layer * nb * mp.nc[L];
    float r = 3.0e38f;
    for (uint32_t c = c0 >> L; c <= (c1 >> L); ++c)
        for (uint32_t b = b0 >> L; b <= (b1 >> L); ++b) r = fminf(r, q[c * nb + b]);
    return r;
}

// farLowerBound for every point of a box at once: a lower bound on the optical depth from any
// point of the box [blo, bhi] (inside cell c, indices along the three axes) to the box boundary
// along every direction of bd, with no stop, from pyrMinSpan(span). With the axis sense sgn, the
// walks cross the rest of the cell's own layer over at least its face distance from the box's far
// side, and layer c_a + s sgn (s >= 1) at distances along the axis between those of its faces from
// the box's far and near sides; the lateral rectangles widen the box by those distances times the
// lateral slopes. Same margins as farLowerBound.
__host__ __device__ inline float farBoxBound(const MediumView& m, const MinPyramid& mp, const BoundDirs& bd,
                                             const uint32_t c[3], const float blo[3], const float bhi[3],
                                             uint32_t span) {
    if (!bd.valid || mp.axis < 0) return 0.f;
    const int a = mp.axis, qs[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
    float lb = 0.f;
    for (int L = (int)c[a]; L >= 0 && L < (int)m.n[a]; L += bd.sgn) {
        const bool own = L == (int)c[a];
        // distances along the axis to the layer's entry and exit faces, least and greatest over the box
        float tn, tf, w;
        if (bd.sgn > 0) {
            tn = own ? 0.f : m.ef[a][L] - bhi[a];
            tf = m.ef[a][L + 1] - blo[a];
            w = own ? m.ef[a][L + 1] - bhi[a] : m.h[a];
        } else {
            tn = own ? 0.f : blo[a] - m.ef[a][L + 1];
            tf = bhi[a] - m.ef[a][L];
            w = own ? blo[a] - m.ef[a][L] : m.h[a];
        }
        tn = fmaxf(tn, 0.f);
        w = fmaxf(w, 0.f);
        uint32_t lo[2], hi[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int q = qs[j];
            const float xlo = blo[q] + fminf(tn * bd.rmin[j], tf * bd.rmin[j]);
            const float xhi = bhi[q] + fmaxf(tn * bd.rmax[j], tf * bd.rmax[j]);
            const float mg = 1e-3f * m.h[q];
            if (xlo <= m.bminf[q] + mg || xhi >= m.bmaxf[q] - mg) return lb * (1.f - 1e-5f);
            const float flo = (xlo - m.e0[q]) * m.invH[q] - 1e-3f, fhi = (xhi - m.e0[q]) * m.invH[q] + 1e-3f;
            const int n1 = (int)m.n[q] - 1;
            const int il = (int)floorf(flo), ih = (int)floorf(fhi);
            lo[j] = (uint32_t)(il < 0 ? 0 : (il > n1 ? n1 : il));
            hi[j] = (uint32_t)(ih < 0 ? 0 : (ih > n1 ? n1 : ih));
        }
        if (w > 0.f) lb = fmaf(w, pyrMinSpan(mp, (uint32_t)L, lo[0], hi[0], lo[1], hi[1], span), lb);
    }
    return lb * (1.f - 1e-5f);
}

// The sub-box of an FP32 point inside the medium's box for tables of `sub` x `sub` x `sub` boxes a
// cell (cell c, box [blo, bhi]); returns the box's linear index, x fastest over the n * sub boxes
// of each axis.
__host__ __device__ inline uint64_t farSubBoxOf(const MediumView& m, const f3& pf, uint32_t sub, uint32_t c[3],
                                                float blo[3], float bhi[3]) {
    const float p[3] = {pf.x, pf.y, pf.z};
    uint32_t k[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = (uint32_t)locateUniform<float>(m, a, p[a]);
        const float fs = (p[a] - m.ef[a][c[a]]) * m.invH[a] * (float)sub;
        const int ks = (int)floorf(fs);
        k[a] = (uint32_t)(ks < 0 ? 0 : (ks >= (int)sub ? (int)sub - 1 : ks));
        farSubBoxFaces(m, a, c[a], k[a], sub, &blo[a], &bhi[a]);
    }
    return (((uint64_t)c[2] * sub + k[2]) * ((uint64_t)m.n[1] * sub) + c[1] * sub + k[1]) * ((uint64_t)m.n[0] * sub) +
           c[0] * sub + k[0];
}

// The order in which the far-field view kernel takes n sources (FP32 directions, 3 each) walked
// along sg * d: by Morton code of the ranks of their lateral slopes about `axis`, so that each run
// of 64 is a compact cone (a regular 32 x 32 grid gives 8 x 8 squares).
inline std::vector<uint32_t> farOrderHost(const float* d, size_t n, double sg, int axis) {
    const int qs[2] = {axis == 0 ? 1 : 0, axis == 2 ? 1 : 2};
    std::vector<double> r[2];
    std::vector<uint32_t> rank[2];
    for (int j = 0; j < 2; ++j) {
        r[j].resize(n);
        for (size_t i = 0; i < n; ++i)
            r[j][i] = sg * (double)d[3 * i + qs[j]] / std::fabs((double)d[3 * i + axis] + 1e-300);
        std::vector<double> u(r[j]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end(), [](double x, double y) { return std::fabs(x - y) <= 1e-9 * (1.0 + std::fabs(x)); }), u.end());
        rank[j].resize(n);
        for (size_t i = 0; i < n; ++i) {
            const size_t k = std::lower_bound(u.begin(), u.end(), r[j][i] - 1e-9 * (1.0 + std::fabs(r[j][i]))) - u.begin();
            rank[j][i] = (uint32_t)(k < u.size() ? k : u.size() - 1);
        }
    }
    auto spread = [](uint64_t x) {
        x &= 0xffffffffull;
        x = (x | (x << 16)) & 0x0000ffff0000ffffull;
        x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
        x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
        x = (x | (x << 2)) & 0x3333333333333333ull;
        x = (x | (x << 1)) & 0x5555555555555555ull;
        return x;
    };
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    for (size_t i = 0; i < n; ++i) key[i] = {spread(rank[0][i]) | (spread(rank[1][i]) << 1), (uint32_t)i};
    std::stable_sort(key.begin(), key.end(), [](const std::pair<uint64_t, uint32_t>& x, const std::pair<uint64_t, uint32_t>& y) { return x.first < y.first; });
    std::vector<uint32_t> perm(n);
    for (size_t i = 0; i < n; ++i) perm[i] = key[i].second;
    return perm;
}

// boundDirsHost for a subset of directions about a given axis and sense (a block of the kernel's
// sources about the whole set's pyramid axis): valid = 0 when some walk advances less than 0.1
// along it.
inline BoundDirs boundDirsAxisHost(const float* d, const uint32_t* idx, size_t n, double sg, int axis, int sgn) {
    BoundDirs bd;
    double mn = 1e300;
    for (size_t i = 0; i < n; ++i) {
        const double c = sgn * sg * (double)d[3 * idx[i] + axis];
        mn = c < mn ? c : mn;
    }
    if (!(mn >= 0.1)) return bd;
    bd.valid = 1;
    bd.sgn = sgn;
    const int qs[2] = {axis == 0 ? 1 : 0, axis == 2 ? 1 : 2};
    for (int j = 0; j < 2; ++j) {
        double lo = 1e300, hi = -1e300;
        for (size_t i = 0; i < n; ++i) {
            const double da = std::fabs((double)d[3 * idx[i] + axis]);
            const double r = sg * (double)d[3 * idx[i] + qs[j]] / da;
            lo = r < lo ? r : lo;
            hi = r > hi ? r : hi;
        }
        bd.rmin[j] = (float)(lo - 1e-6 * (1.0 + std::fabs(lo)));
        bd.rmax[j] = (float)(hi + 1e-6 * (1.0 + std::fabs(hi)));
    }
    return bd;
}

// Host builders of the bound's inputs. minPyramidHost: the pyramid of a grid of material ids
// (x fastest) with layers along `axis`, FP32 sigt per id; mp's offsets and sizes filled, mp->p left
// for the caller (the returned values' device copy); an empty vector when levels exceed kPyrLevels.
inline std::vector<float> minPyramidHost(const uint8_t* g, const uint32_t n[3], const float* sigt, int axis,
                                         MinPyramid* mp) {
    *mp = MinPyramid{};
    const int qb = axis == 0 ? 1 : 0, qc = axis == 2 ? 1 : 2;
    uint32_t levels = 1;
    while ((1u << (levels - 1)) < (n[qb] > n[qc] ? n[qb] : n[qc])) ++levels;
    if (levels > (uint32_t)kPyrLevels) return {};
    mp->axis = axis;
    mp->nl = n[axis];
    mp->levels = levels;
    uint64_t total = 0;
    for (uint32_t L = 0; L < levels; ++L) {
        mp->nb[L] = (n[qb] + (1u << L) - 1) >> L;
        mp->nc[L] = (n[qc] + (1u << L) - 1) >> L;
        mp->off[L] = total;
        total += (uint64_t)mp->nl * mp->nb[L] * mp->nc[L];
    }
    std::vector<float> h(total);
    uint32_t idx[3];
    for (uint32_t l = 0; l < mp->nl; ++l)
        for (uint32_t ic = 0; ic < n[qc]; ++ic)
            for (uint32_t ib = 0; ib < n[qb]; ++ib) {
                idx[axis] = l;
                idx[qb] = ib;
                idx[qc] = ic;
                const size_t cell = idx[0] + (size_t)n[0] * (idx[1] + (size_t)n[1] * idx[2]);
                h[((uint64_t)l * n[qc] + ic) * n[qb] + ib] = sigt[g[cell]];
            }
    for (uint32_t L = 1; L < levels; ++L) {
        const uint32_t pb = mp->nb[L - 1], pc = mp->nc[L - 1], nb = mp->nb[L], nc = mp->nc[L];
        for (uint32_t l = 0; l < mp->nl; ++l)
            for (uint32_t ic = 0; ic < nc; ++ic)
                for (uint32_t ib = 0; ib < nb; ++ib) {
                    float v = 3.0e38f;
                    for (uint32_t dc = 0; dc < 2; ++dc)
                        for (uint32_t db = 0; db < 2; ++db) {
                            const uint32_t cb = 2 * ib + db, cc = 2 * ic + dc;
                            if (cb < pb && cc < pc) {
                                const float w = h[mp->off[L - 1] + ((uint64_t)l * pc + cc) * pb + cb];
                                v = w < v ? w : v;
                            }
                        }
                    h[mp->off[L] + ((uint64_t)l * nc + ic) * nb + ib] = v;
                }
    }
    return h;
}

// The bound's direction set for walks along sg * d_i (d: n FP32 directions, 3 each): the axis
// and sense along which every walk advances most (the largest least component), the lateral slopes'
// range widened by a relative 1e-6. valid = 0 when some walk advances less than 0.1 along it.
inline BoundDirs boundDirsHost(const float* d, size_t n, double sg, int* axisOut) {
    BoundDirs bd;
    double best = -1.0;
    int axis = -1, sgn = 1;
    for (int a = 0; a < 3; ++a)
        for (int s = -1; s <= 1; s += 2) {
            double mn = 1e300;
            for (size_t i = 0; i < n; ++i) {
                const double c = s * sg * (double)d[3 * i + a];
                mn = c < mn ? c : mn;
            }
            if (mn > best) {
                best = mn;
                axis = a;
                sgn = s;
            }
        }
    *axisOut = axis;
    if (!(best >= 0.1)) return bd;
    bd.valid = 1;
    bd.sgn = sgn;
    const int qs[2] = {axis == 0 ? 1 : 0, axis == 2 ? 1 : 2};
    for (int j = 0; j < 2; ++j) {
        double lo = 1e300, hi = -1e300;
        for (size_t i = 0; i < n; ++i) {
            const double da = std::fabs((double)d[3 * i + axis]);
            const double r = sg * (double)d[3 * i + qs[j]] / da;
            lo = r < lo ? r : lo;
            hi = r > hi ? r : hi;
        }
        bd.rmin[j] = (float)(lo - 1e-6 * (1.0 + std::fabs(lo)));
        bd.rmax[j] = (float)(hi + 1e-6 * (1.0 + std::fabs(hi)));
    }
    return bd;
}

}  // namespace sstmc
