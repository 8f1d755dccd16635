// nee.hip -- (b) next-event estimation terms of every path of a batch.
//
// One wavefront per path. Lanes stride over illuminations / views (x frames), so the 64 voxel
// walks issued together start at the same vertex and fan out over neighbouring directions: their
// material-grid reads stay in the same cache lines. The vertex list of the path is read as
// wave-uniform broadcasts. The per-path outputs are exactly what the reference materialises per
// bounce before its cuBLAS reductions (NextEventEstimation.cuh:1016-1458):
//   a[t][p][l] = g_l(x1^t) w1 exp(i 2 pi rho_1)          (single scattering, illumination side)
//   b[t][p][v] = g_v(x1^t)                                (single scattering, view side)
//   A[t][p][l] = g_l(x1^t) f_l(x1^t -> x2^t) w2           (multiple scattering, illumination side)
//   B[t][p][v] = sum_{k>=2} g_v(x_k^t) f_v(x_{k-1}^t -> x_k^t) exp(i 2 pi (rho_k + L_k^t / lambda))
// with w1 = mean_l(pr_l)^-1/2, w2 = mean_l(pr_l pdf_l)^-1/2 (pathProbabilitySingle/Multiple,
// NextEventEstimation.cuh:853-1009). For the correlation output the same slots hold
// a(0) conj a(t), b(0) conj b(t), A(0) conj A(t) and sum_k G_k(0) conj G_k(t).
//
// Precision: voxel walks, exponentials and phase functions run in R (FP32 or FP64); every
// phase (source phase, rho, path length / lambda) is summed and reduced mod 1 in FP64, and path
// lengths are accumulated in FP64 from the FP32 transport positions.
#include "common.h"
#include "gauss.h"
#include "geometry.h"
#include "kernels.h"

namespace sstmc {

template <typename R> struct Cx;
template <> struct Cx<float> {
    typedef float2 T;
    __device__ static float2 make(float a, float b) { return make_float2(a, b); }
};
template <> struct Cx<double> {
    typedef double2 T;
    __device__ static double2 make(double a, double b) { return make_double2(a, b); }
};

template <typename R> __device__ inline R sqrtR(R x);
template <> __device__ inline float sqrtR<float>(float x) { return sqrtf(x); }
template <> __device__ inline double sqrtR<double>(double x) { return sqrt(x); }
template <typename R> __device__ inline R expR(R x);
template <> __device__ inline float expR<float>(float x) { return __expf(x); }
template <> __device__ inline double expR<double>(double x) { return exp(x); }
template <typename R> __device__ inline void sincospiR(R x, R* s, R* c);
template <> __device__ inline void sincospiR<float>(float x, float* s, float* c) { sincospif(x, s, c); }
template <> __device__ inline void sincospiR<double>(double x, double* s, double* c) { sincospi(x, s, c); }

// Position / direction in the walk's real type.
template <typename R> struct Pos;
template <> struct Pos<float> {
    typedef f3 V;
    __device__ static f3 pick(const f3& pf, const d3& pd) { return pf; }
    __device__ static f3 fromD(const d3& d) { return mk((float)d.x, (float)d.y, (float)d.z); }
};
template <> struct Pos<double> {
    typedef d3 V;
    __device__ static d3 pick(const f3& pf, const d3& pd) { return pd; }
    __device__ static d3 fromD(const d3& d) { return d; }
};

__device__ inline d3 dsub(const d3& a, const d3& b) { return mkd(a.x - b.x, a.y - b.y, a.z - b.z); }
template <int DIMS> __device__ inline double ddot(const d3& a, const d3& b) {
    return DIMS == 3 ? a.x * b.x + a.y * b.y + a.z * b.z : a.x * b.x + a.y * b.y;
}
template <int DIMS> __device__ inline d3 dnormalize(const d3& a) {
    double n2 = ddot<DIMS>(a, a);
    if (n2 == 0.0) return DIMS == 3 ? mkd(0, 0, 1) : mkd(0, 1, 0);
    double inv = 1.0 / sqrt(n2);
    return mkd(a.x * inv, a.y * inv, a.z * inv);
}

// Scattering amplitude / pdf (Scattering.cuh:919-931, 983-995; isotropic constants
// Simulation.cuh:266-272).
// The cosine arrives in FP64 (it is formed from FP64 unit vectors) and 1 + g^2 - 2 g c is
// evaluated in FP64: for forward-peaked HG it cancels to (1 - g)^2 near c = 1, which FP32 would
// resolve to ~1e-5 relative only.
template <int DIMS, typename R>
__device__ inline R hgPdf(double g, double c) {
    const R up = (R)(DIMS == 3 ? (1.0 - g * g) / 12.566370614359172954 : (1.0 - g * g) / 6.283185307179586477);
    const R den = (R)fma(-2.0 * g, c, 1.0 + g * g);
    return DIMS == 3 ? up / (den * sqrtR<R>(den)) : up / den;
}
template <int DIMS, typename R, bool TAB = true>
__device__ inline R phasePdf(const ScatterView& s, uint32_t m, double c) {
    if (s.type[m] == 1) return DIMS == 3 ? (R)0.079577471545947668 : (R)0.159154943091895336;
    if (TAB && s.type[m] == 2) {   // tabular evaluation pdf (Scattering.cuh:971-980)
        const PhaseTables& t = *s.tab;
        const uint32_t i = t.off[m] + tabularEntry(t.n[m], DIMS, c);
        if constexpr (sizeof(R) == 4) return t.f[i].w;
        else return t.d[i].w;
    }
    if (TAB && s.type[m] == 4)   // vMF mixture: isotropic in 2D (Scattering.cuh:963-968)
        return DIMS == 3 ? (R)vmfPdf(*s.tab, m, c) : (R)0.159154943091895336;
    return hgPdf<DIMS, R>(s.gd[m], c);
}
// Amplitude |f| at cosine c; a tabular amplitude's phase (in cycles) is added to *ph
// (Scattering.cuh:907-916: the tabulated complex amplitude). TAB = false compiles the tabular and
// vMF cases out (no material uses them), keeping their FP64 code out of the view kernels'
// registers.
template <int DIMS, typename R, bool TAB = true>
__device__ inline R phaseAmp(const ScatterView& s, uint32_t m, double c, double* ph) {
    if (s.type[m] == 1) return DIMS == 3 ? (R)0.282094791773878143 : (R)0.398942280401432678;
    if (TAB && s.type[m] == 2) {
        const PhaseTables& t = *s.tab;
        const uint32_t i = t.off[m] + tabularEntry(t.n[m], DIMS, c);
        *ph += t.ph[i];
        if constexpr (sizeof(R) == 4) return t.f[i].x;
        else return t.d[i].x;
    }
    if (TAB && s.type[m] == 4)   // vMF mixture: isotropic in 2D (Scattering.cuh:898-903, 932-945)
        return DIMS == 3 ? (R)vmfAmplitude(*s.tab, m, c) : (R)0.398942280401432678;
    return sqrtR<R>(hgPdf<DIMS, R>(s.gd[m], c));
}

// Voxel walk dispatch: the branch-free FP32 walk over the packed walk grid (equally spaced
// grids), the general walk (edge tables, FP64 when R is double) otherwise. `sig` is the block's
// LDS table: per-material sigt, then (FP32) the walk grid's palette sigt.
template <int DIMS, typename R, typename V>
__device__ inline R walk(const MediumView& m, const R* sig, const V& p, const V& d, R tauCut, uint32_t* steps) {
    if constexpr (sizeof(R) == 4) {
        if (m.walkFast) return opticalDepthFast<DIMS>(m, sig + kMaxMaterials, p, d, tauCut, steps);
    }
    return opticalDepth<DIMS, R>(m, sig, p, d, tauCut, steps);
}

// Per-block LDS sigt table in R: kMaxMaterials per-material entries, then kWalkPalette entries
// of the walk grid's palette. Every thread of the block must call it.
constexpr int kSigLds = kMaxMaterials + kWalkPalette;
template <typename R>
__device__ inline const R* loadSigTable(const MediumView& m, R* lds) {
    if (threadIdx.x < kMaxMaterials) lds[threadIdx.x] = Real<R>::sigt(m, threadIdx.x);
    else if (threadIdx.x < kSigLds) lds[threadIdx.x] = (R)m.walkSig[threadIdx.x - kMaxMaterials];
    __syncthreads();
    return lds;
}
template <int DIMS, typename R> __device__ inline R isoAmp() {
    return DIMS == 3 ? (R)0.282094791773878143 : (R)0.398942280401432678;
}

struct Geo {
    f3 pf;    // position in FP32 (x + delta rounded)
    d3 pd;    // position in FP64 (x + delta exact)
};

__device__ inline Geo vertexFrame(const VertexBatch& vb, const FramesView& fr, uint64_t o, uint32_t t,
                                  const float4& xm) {
    Geo g;
    if (fr.trivial) {
        g.pf = mk(xm.x, xm.y, xm.z);
        g.pd = mkd(xm.x, xm.y, xm.z);
    } else {
        const float* d = vb.delta + (o * fr.nt + t) * 3;
        float dx = d[0], dy = d[1], dz = d[2];
        g.pf = mk(xm.x + dx, xm.y + dy, xm.z + dz);
        g.pd = mkd((double)xm.x + (double)dx, (double)xm.y + (double)dy, (double)xm.z + (double)dz);
    }
    return g;
}

// g of one source at one position: amplitude (with sqrt of the transmittance) and phase in cycles.
// FarField: gDirectionKernel (FarField.cuh:223-263), from the source's direction dd (FP64) and df
// (FP32, the walk direction's rounding); the illumination walks against the direction and takes
// the positive phase, the view walks along it with the negative phase.
template <int DIMS, typename R>
__device__ inline void gFar(const MediumView& m, const R* sig, const d3& dd, const f3& df, bool illum, const Geo& x,
                            double invLambda, R tauCut, uint32_t* steps, R* amp, double* ph) {
    typedef typename Pos<R>::V V;
    const V p = Pos<R>::pick(x.pf, x.pd);
    const double phase = ddot<DIMS>(dd, x.pd) * invLambda;
    *ph = illum ? phase : -phase;
    V w;
    if constexpr (sizeof(R) == 4) w = illum ? mk(-df.x, -df.y, -df.z) : df;
    else w = illum ? mkd(-dd.x, -dd.y, -dd.z) : dd;
    const R tau = walk<DIMS, R>(m, sig, p, w, tauCut, steps);
    *amp = tau > tauCut ? (R)0 : expR<R>(-(R)0.5 * tau);
}
// PointSource: gPointKernel (PointSource.cuh:117-172), from the source location sp (z = 0 in 2D).
template <int DIMS, typename R>
__device__ inline void gPoint(const MediumView& m, const R* sig, const d3& sp, bool area, const Geo& x,
                              double invLambda, R tauCut, uint32_t* steps, R* amp, double* ph) {
    typedef typename Pos<R>::V V;
    const V p = Pos<R>::pick(x.pf, x.pd);
    const d3 q = dsub(sp, x.pd);
    const double r = sqrt(ddot<DIMS>(q, q));
    const double rr = 1.0 / r;
    *ph = r * invLambda;
    const V w = Pos<R>::fromD(mkd(q.x * rr, q.y * rr, q.z * rr));
    const R tau = walk<DIMS, R>(m, sig, p, w, tauCut, steps);
    double fac = DIMS == 3 ? rr * 0.531125966013598 /* sqrt(ISOTROPIC_AMPLITUDE) */ : sqrt(rr);
    if (area) fac *= sqrt(rr * fabs(DIMS == 3 ? x.pd.z - sp.z : x.pd.y - sp.y));
    *amp = tau > tauCut ? (R)0 : expR<R>(-(R)0.5 * tau) * (R)fac;
}
// FP64 direction or location of source i (z = 0 for a 2D location) and its FP32 direction.
template <int DIMS>
__device__ inline d3 sourceD(const SourceView& s, uint32_t i) {
    const double* q = s.pd + 3 * (size_t)i;
    return mkd(q[0], q[1], DIMS == 3 || s.type == 2 ? q[2] : 0.0);
}
__device__ inline f3 sourceF(const SourceView& s, uint32_t i) {
    const float* q = s.pf + 3 * (size_t)i;
    return mk(q[0], q[1], q[2]);
}
// GaussianBeam: gGaussianBeam with the attenuation along the beam (GaussianBeam.cuh:795-816,
// 1942-1992), the square root of the transmittance toward the beam's source side (illumination
// against the beam direction, view along it), zero outside the medium; no phase.
template <int DIMS, typename R>
__device__ inline void gGauss(const MediumView& m, const R* sig, const d3& dir, bool illum, const Geo& x, R tauCut,
                              uint32_t* steps, R* amp, double* ph) {
    const d3 w = illum ? mkd(-dir.x, -dir.y, -dir.z) : dir;
    const R tau = walk<DIMS, R>(m, sig, Pos<R>::pick(x.pf, x.pd), Pos<R>::fromD(w), tauCut, steps);
    *amp = (gbMaterial<DIMS>(m, x.pd) && !(tau > tauCut)) ? expR<R>(-(R)0.5 * tau) : (R)0;
    *ph = 0;
}
template <int DIMS, typename R, bool GB = false>
__device__ inline void gTerm(const MediumView& m, const R* sig, const SourceView& s, uint32_t i, const Geo& x,
                             double invLambda, R tauCut, uint32_t* steps, R* amp, double* ph) {
    if constexpr (GB) {
        gGauss<DIMS, R>(m, sig, gbDir(s.g, i), s.illum != 0, x, tauCut, steps, amp, ph);
        return;
    }
    if (s.type == 2)
        gFar<DIMS, R>(m, sig, sourceD<DIMS>(s, i), sizeof(R) == 4 ? sourceF(s, i) : mk(0.f, 0.f, 0.f), s.illum != 0,
                      x, invLambda, tauCut, steps, amp, ph);
    else gPoint<DIMS, R>(m, sig, sourceD<DIMS>(s, i), s.area != 0, x, invLambda, tauCut, steps, amp, ph);
}

// Incoming unit direction at x from source i as the three-point functions use it: FarField
// illumination d, FarField view -d (FarField.cuh:546-556), PointSource n(x - s_i)
// (PointSource.cuh:347-361).
template <int DIMS>
__device__ inline d3 incoming(const SourceView& s, uint32_t i, const d3& x) {
    if (s.type == 2) {
        const d3 d = mkd(s.pd[3 * i], s.pd[3 * i + 1], s.pd[3 * i + 2]);
        return s.illum ? d : mkd(-d.x, -d.y, -d.z);
    }
    return dnormalize<DIMS>(dsub(x, mkd(s.pd[3 * i], s.pd[3 * i + 1], DIMS == 3 ? s.pd[3 * i + 2] : 0.0)));
}

// Incoming unit direction at x1 from illumination l (FarField d_l; PointSource n(x1 - s_l)).
template <int DIMS>
__device__ inline d3 illumIn(const SourceView& il, uint32_t l, const d3& x1) {
    if (il.type == 2) return mkd(il.pd[3 * l], il.pd[3 * l + 1], il.pd[3 * l + 2]);
    return dnormalize<DIMS>(dsub(x1, mkd(il.pd[3 * l], il.pd[3 * l + 1], DIMS == 3 ? il.pd[3 * l + 2] : 0.0)));
}

template <typename T> __device__ inline T waveSum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Sum over an aligned group of lpp lanes (lpp a power of two <= 64): xor partners stay inside
// the group, so groups working on different paths never mix.
template <typename T> __device__ inline T groupSum(T v, uint32_t lpp) {
    for (uint32_t o = lpp >> 1; o > 0; o >>= 1) v += __shfl_xor(v, (int)o, 64);
    return v;
}

__device__ inline double frac1(double x) { return x - floor(x); }

// The cut of the unweighted view-side terms (b, B): exact (FP32 210, past which exp(-tau/2) is
// 0 in FP32; FP64 none), or the relative cut tauRef + delta of precision.tauCut.
template <typename R>
__device__ inline R viewCut(const NeeParams& prm) {
    const R exact = sizeof(R) == 4 ? (R)prm.exactCutF : (R)1e300;
    if (!prm.tauRef) return exact;
    const R c = (R)(*prm.tauRef) + (R)prm.tauCutRel;
    return c < exact ? c : exact;
}

template <typename R>
__device__ inline typename Cx<R>::T cis(R amp, double phaseCycles) {
    R s, c;
    sincospiR<R>((R)(2.0 * frac1(phaseCycles)), &s, &c);
    return Cx<R>::make(amp * c, amp * s);
}

// First-vertex work of one path, on a group of lpp lanes: path probabilities, illumination terms
// a and A, single-scattering view terms b (and zero B rows when the path has one vertex).
template <int DIMS, typename R, bool TAB, bool GB = false>
__device__ inline void nee_path(const MediumView& m, const R* sig, const ScatterView& sc, const SourceView& il,
                                const SourceView& vw,
                                const FramesView& fr, const VertexBatch& vb, const NeeParams& prm, const NeeOut& out,
                                uint32_t p, uint32_t sub, uint32_t lpp, uint32_t* stepsOut, float* tauCache) {
    typedef typename Cx<R>::T C;
    const uint32_t P = prm.npaths, L = il.n, V = vw.n, NT = fr.nt;
    const bool corr = prm.output == 1;
    const double invLambda = 1.0 / prm.lambda;
    const R tauCutFF = viewCut<R>(prm);
    C* A = (C*)out.A;
    C* B = (C*)out.B;
    C* a = (C*)out.a;
    C* b = (C*)out.b;
    C* Ar = (C*)out.Ar;
    C* Br = (C*)out.Br;
    // a pool overflow leaves the vertex offsets unusable; the batch is then re-run by the host
    const bool poolOk = (uint64_t)vb.off[prm.npaths] <= vb.cap;
    const uint32_t K = poolOk ? vb.k[p] : 0u;
    const uint64_t off = vb.off[p];
    const C zero = Cx<R>::make(0, 0);
    if (K == 0) {
        for (uint32_t i = sub; i < L * NT; i += lpp) {
            uint32_t l = i % L, t = i / L;
            A[((size_t)t * P + p) * L + l] = zero;
            a[((size_t)t * P + p) * L + l] = zero;
        }
        for (uint32_t i = sub; i < V * NT; i += lpp) {
            uint32_t v = i % V, t = i / V;
            B[((size_t)t * P + p) * V + v] = zero;
            b[((size_t)t * P + p) * V + v] = zero;
            if (Ar) Ar[((size_t)t * P + p) * V + v] = zero;
        }
        if (Br)
            for (uint32_t i = sub; i < L * NT; i += lpp) Br[((size_t)(i / L) * P + p) * L + i % L] = zero;
        return;
    }
    uint32_t steps = 0;
    const float4 x1m = vb.xm[off];
    const uint32_t m1 = __float_as_uint(x1m.w);
    const double rho1 = (double)vb.rho[off];
    const bool multi = K >= 2;
    const float4 x2m = multi ? vb.xm[off + 1] : x1m;
    const Geo x1c = {mk(x1m.x, x1m.y, x1m.z), mkd(x1m.x, x1m.y, x1m.z)};
    const d3 x2cd = mkd(x2m.x, x2m.y, x2m.z);
    const d3 seg12 = dnormalize<DIMS>(dsub(x2cd, x1c.pd));
    const R tauInf = (R)3e38;
    // Illumination terms at x1 (a, A) and the reversed path's view term at x1 (Ar) carry the path
    // weight 1 / sqrt(p), which the same attenuations set: a deep x1 has tiny g and a huge weight,
    // and its terms are O(1). They are walked exactly. The cut of precision.tauCut applies to the
    // unweighted terms only: b (and the view kernels' B, Br).
    const R tauExact = sizeof(R) == 4 ? (R)prm.exactCutF : tauInf;
    constexpr bool gauss = GB && TAB;   // GaussianBeam illumination and view

    if (!multi) {
        for (uint32_t i = sub; i < V * NT; i += lpp) {
            const uint32_t v = i % V, t = i / V;
            B[((size_t)t * P + p) * V + v] = zero;
        }
        if (Br)
            for (uint32_t i = sub; i < L * NT; i += lpp) Br[((size_t)(i / L) * P + p) * L + i % L] = zero;
    }
    const bool isoPointVw = vw.type == 1 && sc.allIsotropic;
    bool anyB = false;   // a nonzero b among this lane's views
    // (w2v: the path weight of the reversed path's term Ar, formed with the probabilities)
    auto viewTerms = [&](R w2v) {

    // ---- single-scattering view terms b = g_v(x1^t)
    // Far-field views under the view cut with the far-field view kernel's records: the views in
    // its order, and a view in a block whose every walk from x1 lies past the cut (x1's live-block
    // mask) takes g = 0 without walking, as its walk would give.
    const bool blockSkip = sizeof(R) == 4 && !gauss && !corr && !Ar && vw.type == 2 && prm.farPerm && vb.live;
    for (uint32_t i = sub; i < V * NT; i += lpp) {
        const uint32_t pos = i % V, t = i / V;
        const uint32_t v = blockSkip ? prm.farPerm[pos] : pos;
        R ga[2];
        double gp[2];
        const int nf = corr ? 2 : 1;
#pragma unroll
        for (int f = 0; f < 2; ++f) {   // constant bound: the two-entry arrays stay in registers
            if (f >= nf) break;
            uint32_t tt = (corr && f == 0) ? 0 : t;
            Geo x1t = vertexFrame(vb, fr, off, tt, x1m);
            const uint32_t blk = pos / 64u;
            if (blockSkip && !((vb.live[2 * (off * NT + t)] >> (blk < 31u ? blk : 31u)) & 1u)) {
                ga[f] = (R)0;   // the term is zero whatever its phase (b below)
                gp[f] = 0.0;
            } else {
                gTerm<DIMS, R, gauss>(m, sig, vw, v, x1t, invLambda, Ar ? tauExact : tauCutFF, &steps, &ga[f],
                                      &gp[f]);
            }
        }
        const size_t o = ((size_t)t * P + p) * V + v;
        b[o] = corr ? cis<R>(ga[0] * ga[1], gp[0] - gp[1]) : (ga[0] == (R)0 ? zero : cis<R>(ga[0], gp[0]));
        anyB = anyB || ga[0] != (R)0;
        if (Ar) {   // reversed path: the view connected at x1 with f toward x2 (field output only)
            R fv = 0;
            double fvp = 0;
            if (multi) {
                const Geo x1t = vertexFrame(vb, fr, off, t, x1m);
                const Geo x2t = vertexFrame(vb, fr, off + 1, t, x2m);
                const d3 pa = fr.temporalAmp ? x1t.pd : x1c.pd;
                const d3 pb = fr.temporalAmp ? x2t.pd : x2cd;
                fv = isoPointVw ? isoAmp<DIMS, R>()
                                : phaseAmp<DIMS, R, TAB>(sc, m1, ddot<DIMS>(incoming<DIMS>(vw, v, pa), dnormalize<DIMS>(dsub(pb, pa))),
                                                    &fvp);
            }
            Ar[o] = multi ? cis<R>(ga[0] * fv * w2v, gp[0] + fvp) : zero;
        }
    }
    };
    // with viewFirst the single-scattering view terms come first: a path whose b and B rows are
    // zero adds nothing to either contraction, which reads only the rows of paths with a nonzero
    // view-side row, so its probabilities and illumination terms are not formed
    if (prm.viewFirst) {
        viewTerms((R)0);   // no Ar with viewFirst
        const bool keep = groupSum<R>(anyB ? (R)1 : (R)0, lpp) > (R)0 || (vb.rowNz && vb.rowNz[p]);
        if (!keep) {
            *stepsOut += steps;
            return;
        }
    }

    // ---- path probabilities from the central first vertex (pathProbabilitySingle/Multiple)
    R s1 = 0, s2 = 0;
    // x1's far-field record (static frames: the record of frame 0 is the central x1), when the
    // fast walk applies
    const FarRec* x1Rec = (sizeof(R) == 4 && DIMS == 3 && vb.far && fr.trivial && m.walkFast) ? &vb.far[off * NT] : nullptr;
    const uint32_t m1g = gauss ? gbMaterial<DIMS>(m, x1c.pd) : 0u;
    for (uint32_t l = sub; l < L; l += lpp) {
        R pr;
        // 3D beams have their own first- and second-point densities; a 2D beam's first point is
        // drawn as from a point source at its focal point, and so are its densities
        // (GaussianBeam.cuh:1667-1683, 1810-1832, 1933-1937)
        if constexpr (gauss && DIMS == 3) {
            // firstPointProbability and secondPointProbability (GaussianBeam.cuh:420-487,
            // 1835-1940): a zero second-point density becomes 1e20 unless every material scatters
            // isotropically
            const double q = gbPr(il.g, *sc.tab, l, x1c.pd, m1g);
            s1 += (R)q;
            if (multi) {
                const double q2 = q * gbPdir(il.g, *sc.tab, sc.allIsotropic != 0, l, x1c.pd, seg12);
                s2 += (q2 == 0 && !sc.allIsotropic) ? (R)1e20 : (R)q2;
            }
            continue;
        }
        if (il.type == 2) {
            typename Pos<R>::V p1 = Pos<R>::pick(x1c.pf, x1c.pd);
            typename Pos<R>::V w;
            if (sizeof(R) == 4) w = Pos<R>::fromD(mkd(-(double)il.pf[3 * l], -(double)il.pf[3 * l + 1], -(double)il.pf[3 * l + 2]));
            else w = Pos<R>::fromD(mkd(-il.pd[3 * l], -il.pd[3 * l + 1], -il.pd[3 * l + 2]));
            R tau;
            if constexpr (sizeof(R) == 4 && DIMS == 3) {
                if (x1Rec) {
                    // the walk's start from x1's far-field record (walkStart of the same FP32
                    // point), the direction part per illumination: the same walk, the same bits
                    const WalkLane wl = walkLane<3>(m, w);
                    tau = x1Rec->cell == kFarOutside
                              ? 0.f
                              : walkFromStartLane<3>(m, sig + kMaxMaterials, x1Rec->dlo, x1Rec->dhi, x1Rec->b,
                                                     x1Rec->cell, wl, (R)prm.exactCutF, &steps);
                } else {
                    tau = walk<DIMS, R>(m, sig, p1, w, (R)prm.exactCutF, &steps);
                }
            } else {
                tau = walk<DIMS, R>(m, sig, p1, w, sizeof(R) == 4 ? (R)prm.exactCutF : tauInf, &steps);
            }
            if (tauCache) tauCache[l] = (float)tau;
            pr = expR<R>(-tau) * (sizeof(R) == 4 ? (R)il.probf[l] : (R)il.probd[l]);
        } else {
            d3 sp = mkd(il.pd[3 * l], il.pd[3 * l + 1], DIMS == 3 ? il.pd[3 * l + 2] : 0.0);
            d3 q = dsub(sp, x1c.pd);
            double rr = 1.0 / sqrt(ddot<DIMS>(q, q));
            typename Pos<R>::V w = Pos<R>::fromD(mkd(q.x * rr, q.y * rr, q.z * rr));
            R tau = walk<DIMS, R>(m, sig, Pos<R>::pick(x1c.pf, x1c.pd), w, tauInf, &steps);
            // PointSource.cuh:92-115 (2D factor rr * 8 pi kept as in the reference, quirk Q7)
            double fac = DIMS == 3 ? rr * rr / 12.566370614359172954 : rr * 25.132741228718345908;
            pr = expR<R>(-tau) * (R)fac;
        }
        s1 += pr;
        if (multi) {
            d3 in = illumIn<DIMS>(il, l, x1c.pd);
            s2 += pr * phasePdf<DIMS, R, TAB>(sc, m1, ddot<DIMS>(in, seg12));
        }
    }
    s1 = groupSum(s1, lpp);
    s2 = groupSum(s2, lpp);
    if (tauCache) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // the group's cached walks
    const R p1 = s1 / (R)L, p2 = s2 / (R)L;
    const R w1 = p1 > 0 ? (R)1 / sqrtR<R>(p1) : (R)0;
    const R w2 = p2 > 0 ? (R)1 / sqrtR<R>(p2) : (R)0;
    const bool isoPointIl = il.type == 1 && sc.allIsotropic;

    // ---- illumination terms a (single) and A (multiple)
    for (uint32_t i = sub; i < L * NT; i += lpp) {
        const uint32_t l = i % L, t = i / L;
        R ga[2], fa[2] = {0, 0};
        double gp[2], fp[2] = {0, 0};   // fp: phase of a tabular amplitude, in cycles
        const int nf = corr ? 2 : 1;
#pragma unroll
        for (int f = 0; f < 2; ++f) {   // constant bound: the two-entry arrays stay in registers
            if (f >= nf) break;
            uint32_t tt = (corr && f == 0) ? 0 : t;
            Geo x1t = vertexFrame(vb, fr, off, tt, x1m);
            if (tauCache) {
                // the probability loop's walk from x1 toward illumination l (gFar, the same FP32
                // start, direction and cut: frames are static, so x1t is x1)
                const R tau = (R)tauCache[l];
                const double phase = ddot<DIMS>(sourceD<DIMS>(il, l), x1t.pd) * invLambda;
                gp[f] = il.illum != 0 ? phase : -phase;
                ga[f] = tau > tauExact ? (R)0 : expR<R>(-(R)0.5 * tau);
            } else {
                gTerm<DIMS, R, gauss>(m, sig, il, l, x1t, invLambda, tauExact, &steps, &ga[f], &gp[f]);
            }
            if (multi) {
                Geo x2t = vertexFrame(vb, fr, off + 1, tt, x2m);
                const d3 pa = fr.temporalAmp ? x1t.pd : x1c.pd;
                const d3 pb = fr.temporalAmp ? x2t.pd : x2cd;
                if (gauss) {   // fGaussianBeam (GaussianBeam.cuh:818-892) of x1's material toward x2
                    const uint32_t mg = gbMaterial<DIMS>(m, pa);
                    if (mg) {
                        const Cd fz = gbF<DIMS>(il.g, *sc.tab, gbFocal(il.g, l), gbDir(il.g, l), mg, pa,
                                          dnormalize<DIMS>(dsub(pb, pa)), il.g.waveK);
                        fa[f] = (R)cabsd(fz);
                        fp[f] = atan2(fz.im, fz.re) * 0.15915494309189533577;
                    }
                } else if (isoPointIl) fa[f] = isoAmp<DIMS, R>();
                else fa[f] = phaseAmp<DIMS, R, TAB>(sc, m1, ddot<DIMS>(illumIn<DIMS>(il, l, pa), dnormalize<DIMS>(dsub(pb, pa))),
                                               &fp[f]);
            }
        }
        const size_t o = ((size_t)t * P + p) * L + l;
        if (!corr) {
            a[o] = cis<R>(ga[0] * w1, gp[0] + rho1);
            A[o] = multi ? cis<R>(ga[0] * fa[0] * w2, gp[0] + fp[0]) : zero;
        } else {
            a[o] = cis<R>(ga[0] * ga[1] * w1 * w1, gp[0] - gp[1]);
            A[o] = multi ? cis<R>(ga[0] * fa[0] * ga[1] * fa[1] * w2 * w2, gp[0] - gp[1] + fp[0] - fp[1]) : zero;
        }
    }

    if (!prm.viewFirst) viewTerms(w2);
    *stepsOut += steps;
}

// Illuminations whose optical depth from x1 k_nee keeps per 64-lane path (16 KiB of LDS a block).
constexpr uint32_t kTauCacheL = 1024;

// GB: GaussianBeam sources (their FP64 beam functions kept out of the other instantiations)
template <int DIMS, typename R, bool TAB, bool GB>
__global__ void __launch_bounds__(256) k_nee(MediumView m, ScatterView sc, SourceView il, SourceView vw, FramesView fr,
                                             VertexBatch vb, NeeParams prm, NeeOut out) {
    typedef typename Cx<R>::T C;
    // lanes are split into groups of lpp lanes, one path per group (several paths per wavefront
    // when L and V are small)
    const uint32_t lane = threadIdx.x & 63, lpp = prm.lpp, sub = lane & (lpp - 1);
    const uint32_t p = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 / lpp) + lane / lpp;
    __shared__ R sigLds[kSigLds];
    // far-field illuminations, FP32 terms, field output, static frames: each illumination's walk
    // from x1 is kept from the path probabilities for the illumination terms (kTauCacheL per
    // 64-lane path; smaller groups hold at most lpp illuminations)
    __shared__ float tauLds[4 * kTauCacheL];
    const uint32_t cap = lpp == 64 ? kTauCacheL : lpp;
    float* tauCache = (sizeof(R) == 4 && !GB && il.type == 2 && prm.output == 0 && fr.trivial && il.n <= cap)
                          ? tauLds + (threadIdx.x / lpp) * cap
                          : nullptr;
    const R* sig = loadSigTable<R>(m, sigLds);
    uint32_t steps = 0;
    if (p < prm.npaths)
        nee_path<DIMS, R, TAB, GB>(m, sig, sc, il, vw, fr, vb, prm, out, p, sub, lpp, &steps, tauCache);
    uint32_t tot = waveSum(steps);
    if (lane == 0 && out.steps) slotAdd(out.steps, tot);
}



// ---------------------------------------------------------------- TemporalCorrelationSampler
// Position of a vertex seen at time s by the correlation sampler: x + s U_m, FP64 (exact sum of
// the FP32 position and the FP64 shift) with its FP32 rounding for the walk.
__device__ inline Geo shiftGeo(const float4& xm, const double* U, double s) {
    Geo g;
    g.pd = mkd((double)xm.x + s * U[0], (double)xm.y + s * U[1], (double)xm.z + s * U[2]);
    g.pf = mk((float)g.pd.x, (float)g.pd.y, (float)g.pd.z);
    return g;
}

// First-vertex work of one path for the TemporalCorrelationSampler, on a group of lpp lanes. The
// sources hold the u1 sources (i < L) then the u2 ones (L + i). For lag tau: u1's g sees x1 at
// -tau/2, u2's g at +tau/2 (copyPointsToBuffersThroughput, TemporalSamplerCorrelation.cuh:228-250);
// the three-point functions of both see x1 and x2 at -tau/2 (copyPointsToBuffersThreePoints,
// :253-313, quirk Q3); the decorrelation at x1 is exp(-k^2 |D tau| |q|^2) with q the mean u1/u2
// incoming direction minus n(x2 - x1), from the unshifted positions (C1_l1l2,
// temporalCorrelationSourceContributionMultiple, SourceInterface.cuh:280-352). Path
// probabilities: means over the 2L sources (pathSingleProbability, :783-829), the second-point
// factor with the u1 directions for both halves (pathMultipleProbability, :833-856), no square
// root (NextEventEstimation.cuh:851-880, 930-977). Outputs:
//   a[t][p][l] = g_l1 conj(g_l2) / p1,     b[t][p][v] = g_v1 conj(g_v2)
//   A[t][p][l] = g_l1 f_l1 conj(g_l2 f_l2) exp(-k^2 |D1 tau| |q|^2) / p2
template <int DIMS, typename R, bool TAB>
__device__ inline void nee_path_corr(const MediumView& m, const R* sig, const ScatterView& sc, const SourceView& il,
                                     const SourceView& vw, const FramesView& fr, const VertexBatch& vb,
                                     const NeeParams& prm, const NeeOut& out, uint32_t p, uint32_t sub, uint32_t lpp,
                                     uint32_t* stepsOut) {
    typedef typename Cx<R>::T C;
    const uint32_t P = prm.npaths, L = il.n / 2, V = vw.n / 2, NT = fr.nt;
    const double invLambda = 1.0 / prm.lambda;
    const double kw2 = 39.478417604357434 * invLambda * invLambda;   // (2 pi / lambda)^2
    const R tauCutFF = viewCut<R>(prm);
    const R tauInf = (R)3e38;
    C* A = (C*)out.A;
    C* B = (C*)out.B;
    C* a = (C*)out.a;
    C* b = (C*)out.b;
    const bool poolOk = (uint64_t)vb.off[prm.npaths] <= vb.cap;
    const uint32_t K = poolOk ? vb.k[p] : 0u;
    const uint64_t off = vb.off[p];
    const C zero = Cx<R>::make(0, 0);
    // coherent backscattering (TemporalCorrelationSampler): Ar / Br and the separate first-vertex
    // terms (NextEventEstimation.cuh:1237-1378)
    C* Ar = (C*)out.Ar;
    C* Br = (C*)out.Br;
    if (K == 0) {
        for (uint32_t i = sub; i < L * NT; i += lpp) {
            const size_t o = ((size_t)(i / L) * P + p) * L + i % L;
            A[o] = zero;
            a[o] = zero;
            if (Br) Br[o] = zero;
        }
        for (uint32_t i = sub; i < V * NT; i += lpp) {
            const size_t o = ((size_t)(i / V) * P + p) * V + i % V;
            B[o] = zero;
            b[o] = zero;
            if (Ar) Ar[o] = zero;
        }
        return;
    }
    uint32_t steps = 0;
    const float4 x1m = vb.xm[off];
    const uint32_t m1 = __float_as_uint(x1m.w);
    const bool multi = K >= 2;
    const float4 x2m = multi ? vb.xm[off + 1] : x1m;
    const uint32_t m2 = __float_as_uint(x2m.w);
    const Geo x1c = {mk(x1m.x, x1m.y, x1m.z), mkd(x1m.x, x1m.y, x1m.z)};
    const d3 seg12 = dnormalize<DIMS>(dsub(mkd(x2m.x, x2m.y, x2m.z), x1c.pd));

    R s1 = 0, s2 = 0;
    const FarRec* x1Rec = nullptr;   // (the plain walk: x1 moves with the lag)
    for (uint32_t l = sub; l < 2 * L; l += lpp) {
        R pr;
        if (il.type == 2) {
            typename Pos<R>::V p1 = Pos<R>::pick(x1c.pf, x1c.pd);
            typename Pos<R>::V w;
            if (sizeof(R) == 4) w = Pos<R>::fromD(mkd(-(double)il.pf[3 * l], -(double)il.pf[3 * l + 1], -(double)il.pf[3 * l + 2]));
            else w = Pos<R>::fromD(mkd(-il.pd[3 * l], -il.pd[3 * l + 1], -il.pd[3 * l + 2]));
            R tau;
            if constexpr (sizeof(R) == 4 && DIMS == 3) {
                if (x1Rec) {
                    // the walk's start from x1's far-field record (walkStart of the same FP32
                    // point), the direction part per illumination: the same walk, the same bits
                    const WalkLane wl = walkLane<3>(m, w);
                    tau = x1Rec->cell == kFarOutside
                              ? 0.f
                              : walkFromStartLane<3>(m, sig + kMaxMaterials, x1Rec->dlo, x1Rec->dhi, x1Rec->b,
                                                     x1Rec->cell, wl, (R)prm.exactCutF, &steps);
                } else {
                    tau = walk<DIMS, R>(m, sig, p1, w, (R)prm.exactCutF, &steps);
                }
            } else {
                tau = walk<DIMS, R>(m, sig, p1, w, sizeof(R) == 4 ? (R)prm.exactCutF : tauInf, &steps);
            }
            pr = expR<R>(-tau) * (sizeof(R) == 4 ? (R)il.probf[l] : (R)il.probd[l]);
        } else {
            d3 sp = mkd(il.pd[3 * l], il.pd[3 * l + 1], DIMS == 3 ? il.pd[3 * l + 2] : 0.0);
            d3 q = dsub(sp, x1c.pd);
            double rr = 1.0 / sqrt(ddot<DIMS>(q, q));
            typename Pos<R>::V w = Pos<R>::fromD(mkd(q.x * rr, q.y * rr, q.z * rr));
            R tau = walk<DIMS, R>(m, sig, Pos<R>::pick(x1c.pf, x1c.pd), w, tauInf, &steps);
            double fac = DIMS == 3 ? rr * rr / 12.566370614359172954 : rr * 25.132741228718345908;
            pr = expR<R>(-tau) * (R)fac;
        }
        s1 += pr;
        if (multi) s2 += pr * phasePdf<DIMS, R, TAB>(sc, m1, ddot<DIMS>(illumIn<DIMS>(il, l % L, x1c.pd), seg12));
    }
    s1 = groupSum(s1, lpp);
    s2 = groupSum(s2, lpp);
    const R p1 = s1 / (R)(2 * L), p2 = s2 / (R)(2 * L);
    const R w1 = p1 > 0 ? (R)1 / p1 : (R)0;
    const R w2 = p2 > 0 ? (R)1 / p2 : (R)0;
    const bool isoPointIl = il.type == 1 && sc.allIsotropic;
    // illumination terms carry the path weight: walked exactly; view terms take the cut, except
    // with coherent backscattering, where the view terms at x1 carry it too
    const R tauIl = sizeof(R) == 4 ? (R)prm.exactCutF : tauInf, tauVw = Ar ? tauIl : tauCutFF;

    for (uint32_t i = sub; i < L * NT; i += lpp) {
        const uint32_t l = i % L, t = i / L;
        const double tau = fr.ticks[t], s = 0.5 * tau;
        const Geo x1a = shiftGeo(x1m, fr.Ud[m1], -s), x1b = shiftGeo(x1m, fr.Ud[m1], s);
        R g1, g2;
        double gp1, gp2;
        gTerm<DIMS, R>(m, sig, il, l, x1a, invLambda, tauIl, &steps, &g1, &gp1);
        gTerm<DIMS, R>(m, sig, il, L + l, x1b, invLambda, tauIl, &steps, &g2, &gp2);
        const size_t o = ((size_t)t * P + p) * L + l;
        a[o] = cis<R>(g1 * g2 * w1, gp1 - gp2);
        if (multi) {
            const Geo x2a = shiftGeo(x2m, fr.Ud[m2], -s);
            const d3 dir = dnormalize<DIMS>(dsub(x2a.pd, x1a.pd));
            R f1, f2;
            double fp1 = 0, fp2 = 0;
            if (isoPointIl) {
                f1 = f2 = isoAmp<DIMS, R>();
            } else {
                f1 = phaseAmp<DIMS, R, TAB>(sc, m1, ddot<DIMS>(illumIn<DIMS>(il, l, x1a.pd), dir), &fp1);
                f2 = phaseAmp<DIMS, R, TAB>(sc, m1, ddot<DIMS>(illumIn<DIMS>(il, L + l, x1a.pd), dir), &fp2);
            }
            const d3 i1 = illumIn<DIMS>(il, l, x1c.pd), i2 = illumIn<DIMS>(il, L + l, x1c.pd);
            const d3 q = mkd(0.5 * (i1.x + i2.x) - seg12.x, 0.5 * (i1.y + i2.y) - seg12.y, 0.5 * (i1.z + i2.z) - seg12.z);
            const double e = kw2 * fabs(fr.Dd[m1] * tau) * ddot<DIMS>(q, q);
            A[o] = cis<R>(g1 * f1 * g2 * f2 * w2 * expR<R>((R)-e), gp1 + fp1 - gp2 - fp2);
            if (out.F1l) {
                ((C*)out.F1l)[o] = cis<R>(g1 * f1 * w2, gp1 + fp1);
                ((C*)out.F2l)[o] = cis<R>(g2 * f2, gp2 + fp2);
            }
        } else {
            A[o] = zero;
        }
    }
    if (!multi) {
        for (uint32_t i = sub; i < V * NT; i += lpp) B[((size_t)(i / V) * P + p) * V + i % V] = zero;
        if (Br)
            for (uint32_t i = sub; i < L * NT; i += lpp) Br[((size_t)(i / L) * P + p) * L + i % L] = zero;
    }
    const bool isoPointVw = vw.type == 1 && sc.allIsotropic;
    for (uint32_t i = sub; i < V * NT; i += lpp) {
        const uint32_t v = i % V, t = i / V;
        const double s = 0.5 * fr.ticks[t];
        const Geo x1a = shiftGeo(x1m, fr.Ud[m1], -s), x1b = shiftGeo(x1m, fr.Ud[m1], s);
        R g1, g2;
        double gp1, gp2;
        gTerm<DIMS, R>(m, sig, vw, v, x1a, invLambda, tauVw, &steps, &g1, &gp1);
        gTerm<DIMS, R>(m, sig, vw, V + v, x1b, invLambda, tauVw, &steps, &g2, &gp2);
        const size_t o = ((size_t)t * P + p) * V + v;
        b[o] = cis<R>(g1 * g2, gp1 - gp2);
        if (Ar) {
            // the reversed path's first connection: the views at x1 with f toward x2, u1 with the
            // path weight; both amplitudes at the u1 positions (quirk Q3); C4_v1v2 at x1
            if (multi) {
                const Geo x2a = shiftGeo(x2m, fr.Ud[m2], -s);
                const d3 dir = dnormalize<DIMS>(dsub(x2a.pd, x1a.pd));
                R f1, f2;
                double fp1 = 0, fp2 = 0;
                if (isoPointVw) {
                    f1 = f2 = isoAmp<DIMS, R>();
                } else {
                    f1 = phaseAmp<DIMS, R, TAB>(sc, m1, ddot<DIMS>(incoming<DIMS>(vw, v, x1a.pd), dir), &fp1);
                    f2 = phaseAmp<DIMS, R, TAB>(sc, m1, ddot<DIMS>(incoming<DIMS>(vw, V + v, x1a.pd), dir), &fp2);
                }
                const d3 e1 = incoming<DIMS>(vw, v, x1c.pd), e2 = incoming<DIMS>(vw, V + v, x1c.pd);
                const d3 q = mkd(0.5 * (e1.x + e2.x) - seg12.x, 0.5 * (e1.y + e2.y) - seg12.y,
                                 0.5 * (e1.z + e2.z) - seg12.z);
                const double ex = kw2 * fabs(fr.Dd[m1] * fr.ticks[t]) * ddot<DIMS>(q, q);
                Ar[o] = cis<R>(g1 * f1 * w2 * g2 * f2 * expR<R>((R)-ex), gp1 + fp1 - gp2 - fp2);
                ((C*)out.F1v)[o] = cis<R>(g1 * f1 * w2, gp1 + fp1);
                ((C*)out.F2v)[o] = cis<R>(g2 * f2, gp2 + fp2);
            } else {
                Ar[o] = zero;
            }
        }
    }
    *stepsOut += steps;
}

template <int DIMS, typename R, bool TAB>
__global__ void __launch_bounds__(256) k_nee_corr(MediumView m, ScatterView sc, SourceView il, SourceView vw,
                                                  FramesView fr, VertexBatch vb, NeeParams prm, NeeOut out) {
    const uint32_t lane = threadIdx.x & 63, lpp = prm.lpp, sub = lane & (lpp - 1);
    const uint32_t p = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 / lpp) + lane / lpp;
    __shared__ R sigLds[kSigLds];
    const R* sig = loadSigTable<R>(m, sigLds);
    uint32_t steps = 0;
    if (p < prm.npaths) nee_path_corr<DIMS, R, TAB>(m, sig, sc, il, vw, fr, vb, prm, out, p, sub, lpp, &steps);
    uint32_t tot = waveSum(steps);
    if (lane == 0 && out.steps) slotAdd(out.steps, tot);
}

// The far-field view kernel's record of a vertex at one frame (FarRec): the walk's start (cell and
// face offsets, or kFarOutside), the material, the FP32 back direction, the FP64 position and
// phase base; its live-block masks start all set (k_far_live narrows them).
template <int DIMS>
__device__ inline void writeFarRec(const MediumView& m, const VertexBatch& vb, uint64_t idx, const Geo& xk,
                                   uint32_t mat, const d3& back, double phase) {
    FarRec f;
    if (inside<DIMS, float>(m, xk.pf)) {
        const WalkStart w = walkStart<DIMS>(m, xk.pf);
        for (int a = 0; a < 3; ++a) {
            f.dlo[a] = w.dlo[a];
            f.dhi[a] = w.dhi[a];
            f.b[a] = w.b[a];
        }
        f.cell = w.cell;
    } else {
        for (int a = 0; a < 3; ++a) f.dlo[a] = f.dhi[a] = f.b[a] = 0.f;
        f.cell = kFarOutside;
    }
    f.mat = mat;
    f.back[0] = (float)back.x;
    f.back[1] = (float)back.y;
    f.back[2] = (float)back.z;
    f.pad[0] = f.pad[1] = 0u;
    f.px = xk.pd.x;
    f.py = xk.pd.y;
    f.pz = xk.pd.z;
    f.phase = phase;
    vb.far[idx] = f;
    if (vb.live) vb.live[2 * idx] = vb.live[2 * idx + 1] = 0xffffffffu;
}

// Vertex records for the view terms: for every vertex k >= 2 and frame, the unit vector back to
// the previous vertex (the direction the view-side HG amplitude is evaluated against,
// FarField.cuh:546-599 / PointSource.cuh:347-446) and the phase base rho_k + L_k^t / lambda with
// L_k^t the FP64 path length from x_1^t (pathContriobutionKernel, TemporalSampler.cuh:668-719).
// Done once per (vertex, frame) here instead of once per (vertex, view, frame) in k_nee_view.
// With vb.far (3D) also the far-field view kernel's record of every vertex, the first included
// (k_nee's single-scattering view walks read its live-block masks). kPrepLanes lanes take a path,
// one vertex each in runs of kPrepLanes, and form L_k as an inclusive scan of the segment lengths
// over the run plus the run before's total, so that a path of K vertices takes K / kPrepLanes
// dependent steps (path lengths are heavy-tailed; the longest paths of a batch set its duration).
constexpr uint32_t kPrepLanes = 16;
template <int DIMS>
__global__ void __launch_bounds__(256) k_vertex_prep(MediumView m, VertexBatch vb, FramesView fr, uint32_t npaths,
                                                     double invLambda) {
    const uint32_t lane = threadIdx.x % kPrepLanes;
    const uint32_t p = (blockIdx.x * blockDim.x + threadIdx.x) / kPrepLanes;
    // every exit below is uniform over a path's lanes, which the scan's shuffles stay within
    if (p >= npaths) return;
    if ((uint64_t)vb.off[npaths] > vb.cap) return;   // pool overflow: the batch is re-run
    if (lane == 0 && vb.rowNz) vb.rowNz[p] = 0u;
    const uint32_t K = vb.k[p], NT = fr.nt;
    if (K == 0) return;
    const uint64_t off = vb.off[p];
    for (uint32_t t = 0; t < NT; ++t) {
        double carry = 0.0;
        for (uint32_t k0 = 0; k0 < K; k0 += kPrepLanes) {
            const uint32_t k = k0 + lane;
            const bool act = k < K;
            const uint64_t ok = off + (act ? k : 0u);
            const float4 xkm = vb.xm[ok];
            const Geo xk = vertexFrame(vb, fr, ok, t, xkm);
            double seg = 0.0;
            d3 b = mkd(0.0, 0.0, 0.0);
            if (act && k > 0) {
                const float4 xpm = vb.xm[ok - 1];
                const Geo xp = vertexFrame(vb, fr, ok - 1, t, xpm);
                const d3 sg = dsub(xk.pd, xp.pd);
                seg = sqrt(ddot<DIMS>(sg, sg));
                const d3 pk = fr.temporalAmp ? xk.pd : mkd(xkm.x, xkm.y, xkm.z);
                const d3 pp = fr.temporalAmp ? xp.pd : mkd(xpm.x, xpm.y, xpm.z);
                b = dnormalize<DIMS>(dsub(pp, pk));
            }
            double len = seg;
#pragma unroll
            for (uint32_t d = 1; d < kPrepLanes; d <<= 1) {
                const double u = __shfl_up(len, d, kPrepLanes);
                if (lane >= d) len += u;
            }
            len += carry;
            carry = __shfl(len, kPrepLanes - 1, kPrepLanes);
            if (!act) continue;
            if (k == 0) {
                if (DIMS == 3 && vb.far)
                    writeFarRec<DIMS>(m, vb, off * NT + t, xk, __float_as_uint(xkm.w), mkd(0.0, 0.0, 0.0),
                                      (double)vb.rho[off]);
                continue;
            }
            VertexRec r;
            r.bx = b.x; r.by = b.y; r.bz = b.z;
            r.phase = (double)vb.rho[ok] + len * invLambda;
            vb.rec[ok * NT + t] = r;
            if (DIMS == 3 && vb.far) writeFarRec<DIMS>(m, vb, ok * NT + t, xk, __float_as_uint(xkm.w), b, r.phase);
        }
    }
}

// Source kinds of the per-vertex sums, fixed at compile time so that the FarField role (which
// flips the walk direction, the phase sign and the incoming direction) is no runtime branch in the
// hot loop: point sources; FarField views (the NEE view terms); FarField illuminations (the
// reversed paths of coherent backscattering).
enum : int { kSrcPoint = 0, kSrcFarView = 1, kSrcFarIllum = 2, kSrcGauss = 3 };

// Beam direction of GaussianBeam source i (other kinds: unused, vd returned).
template <int SRC>
__device__ inline d3 sourceDir(const SourceView& s, uint32_t i, const d3& vd) {
    if constexpr (SRC == kSrcGauss) return gbDir(s.g, i);
    else return vd;
}

// G_k of one source at one frame: attenuated source term g(x_k^t) times the HG amplitude against
// the record's back direction, with the record's phase base added to the source phase. The
// source's direction or location arrives in registers (vd FP64, vf FP32 direction), loaded once
// per lane before the vertex loop.
template <int DIMS, typename R, int SRC, bool TAB>
__device__ inline typename Cx<R>::T viewTerm(const MediumView& m, const R* sig, const ScatterView& sc,
                                            const SourceView& vw, const FramesView& fr, const VertexBatch& vb,
                                            const d3& vd, const d3& vdir, const f3& vf, uint64_t ok, uint32_t t,
                                            const float4& xkm, double invLambda, R tauCut, bool isoPoint,
                                            uint32_t* steps) {
    const Geo xk = vertexFrame(vb, fr, ok, t, xkm);
    const VertexRec rec = vb.rec[ok * fr.nt + t];
    if constexpr (SRC == kSrcGauss) {
        // g along the view beam times fGaussianBeam of x_k's material against the direction from
        // x_{k-1} (focal point vd, beam direction vdir, negative wave number)
        R ga;
        double gp;
        gGauss<DIMS, R>(m, sig, vdir, false, xk, tauCut, steps, &ga, &gp);
        const d3 pk = fr.temporalAmp ? xk.pd : mkd(xkm.x, xkm.y, xkm.z);
        const uint32_t mg = gbMaterial<DIMS>(m, pk);
        if (mg == 0 || ga == (R)0) return Cx<R>::make(0, 0);
        const Cd fz = gbF<DIMS>(vw.g, *sc.tab, vd, vdir, mg, pk, mkd(-rec.bx, -rec.by, -rec.bz), -vw.g.waveK);
        return cis<R>(ga * (R)cabsd(fz), atan2(fz.im, fz.re) * 0.15915494309189533577 + rec.phase);
    }
    R ga;
    double gp;
    if constexpr (SRC == kSrcPoint) gPoint<DIMS, R>(m, sig, vd, vw.area != 0, xk, invLambda, tauCut, steps, &ga, &gp);
    else gFar<DIMS, R>(m, sig, vd, vf, SRC == kSrcFarIllum, xk, invLambda, tauCut, steps, &ga, &gp);
    R fv;
    if (isoPoint) {
        fv = isoAmp<DIMS, R>();
    } else {
        const d3 pk = fr.temporalAmp ? xk.pd : mkd(xkm.x, xkm.y, xkm.z);
        // incoming direction at x_k (FarField: illumination d, view -d; points: from the source)
        const d3 av = SRC == kSrcFarIllum  ? vd
                      : SRC == kSrcFarView ? mkd(-vd.x, -vd.y, -vd.z)
                                           : dnormalize<DIMS>(dsub(pk, vd));
        fv = phaseAmp<DIMS, R, TAB>(sc, __float_as_uint(xkm.w), ddot<DIMS>(av, mkd(rec.bx, rec.by, rec.bz)), &gp);
    }
    return cis<R>(ga * fv, gp + rec.phase);
}

// Work unit of the view kernels: (path, frame chunk, block of lppView sources), one source per
// lane. With WAVE (lppView == 64) a unit fills the wavefront and its index goes through
// readfirstlane, so the path, its vertex count and offset, the frame and every per-vertex read
// (position, record, displacement) are wave-uniform: scalar registers and scalar loads, leaving
// the vector memory path to the walks' scattered word reads.
constexpr uint32_t kCorrLags = 16;   // frames per correlation work unit
struct ViewUnit {
    uint32_t p, chunk, blk, sub;
};
template <bool WAVE>
__device__ inline ViewUnit viewUnit(const NeeParams& prm) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint64_t unit;
    ViewUnit u;
    if constexpr (WAVE) {
        unit = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave);
        u.sub = lane;
    } else {
        const uint32_t lpp = prm.lppView;
        unit = (uint64_t)wave * (64 / lpp) + lane / lpp;
        u.sub = lane & (lpp - 1);
    }
    const uint32_t perPath = prm.viewBlocks * prm.frameChunks;
    u.p = (uint32_t)(unit / perPath);
    const uint32_t rem = (uint32_t)(unit % perPath);
    u.blk = rem % prm.viewBlocks;
    u.chunk = rem / prm.viewBlocks;
    return u;
}

// Multiple-scattering view terms B (the hot loop). Work unit = (path, frame t, block of lppView
// sources): one source per lane, looping over the path's vertices k >= 2. Splitting a path over
// its blocks keeps the longest unit short: vertex counts are heavy-tailed (the longest path of a
// batch can hold ~15x the mean), and a unit spanning all sources of such a path would outlast a
// whole SIMD's share of the batch. Templated on the source kind so the loop carries only what it
// needs. Waves per SIMD (measured on 1024 x 1024 far-field and 64 x 64 point-source workloads): 8
// for FP32 far-field sources (64 VGPRs; the latency of the walk's scattered word reads wants the
// occupancy), 6 for FP32 point sources (80 VGPRs; their longer per-vertex set-up spills in the
// vertex loop at 64), 4 for FP64 (no spills; 1.6x faster than 6 waves with spills).
template <int DIMS, typename R, int SRC, bool WAVE, bool TAB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(SRC == kSrcGauss ? 2 : (sizeof(R) == 8 ? 4 : (SRC != kSrcPoint ? 8 : 6)))))
k_nee_view(MediumView m, ScatterView sc, SourceView vw, FramesView fr, VertexBatch vb, NeeParams prm, NeeOut out) {
    typedef typename Cx<R>::T C;
    __shared__ R sigLds[kSigLds];
    const R* sig = loadSigTable<R>(m, sigLds);
    const ViewUnit u = viewUnit<WAVE>(prm);
    const uint32_t p = u.p, t = u.chunk, P = prm.npaths, V = vw.n;
    const bool poolOk = (uint64_t)vb.off[P] <= vb.cap;
    const uint32_t K = (p < P && poolOk) ? vb.k[p] : 0u;   // K < 2: k_nee wrote the zero rows
    const uint64_t off = p < P ? vb.off[p] : 0;
    const double invLambda = 1.0 / prm.lambda;
    const R tauCut = viewCut<R>(prm);
    C* B = (C*)out.B;
    const bool isoPoint = SRC == kSrcPoint && sc.allIsotropic;
    uint32_t steps = 0;
    const uint32_t v = u.blk * (WAVE ? 64u : prm.lppView) + u.sub;
    const bool active = v < V && K >= 2;
    // direction (far field) or location (point) of this lane's source
    const d3 vd = sourceD<DIMS>(vw, active ? v : 0);
    const f3 vf = SRC != kSrcPoint && SRC != kSrcGauss && sizeof(R) == 4 ? sourceF(vw, active ? v : 0)
                                                                         : mk(0.f, 0.f, 0.f);
    const d3 vdir = sourceDir<SRC>(vw, active ? v : 0, vd);
    if (active) {
        R accRe = 0, accIm = 0;
        for (uint32_t k = 1; k < K; ++k) {
            const uint64_t ok = off + k;
            const C g = viewTerm<DIMS, R, SRC, TAB>(m, sig, sc, vw, fr, vb, vd, vdir, vf, ok, t, vb.xm[ok], invLambda,
                                                    tauCut, isoPoint, &steps);
            accRe += g.x;
            accIm += g.y;
        }
        B[((size_t)t * P + p) * V + v] = Cx<R>::make(accRe, accIm);
    }
    uint32_t tot = waveSum(steps);
    if ((threadIdx.x & 63) == 0 && out.steps) slotAdd(out.steps + kCountSlots * kCountStride, tot);
}

// The live-block masks of the far-field view kernel (VertexBatch::live) for every record of the batch:
// live[0] for the views under the launch's cut (viewCut), live[1] for the reversed paths'
// illuminations under the exact cut; bit j is clear when the cell bound of block j (FarBound::table)
// passes the cut's skip threshold. One thread per record, a gather of one table row; records of
// first vertices, which only k_nee's single-scattering walks read, included.
__device__ inline float farRoleSkip(const NeeParams& prm, int role) {
    return role == 0 ? farSkip(viewCut<float>(prm)) : farSkip(prm.exactCutF);
}
__global__ void __launch_bounds__(256) k_far_live(MediumView m, VertexBatch vb, FramesView fr, NeeParams prm,
                                                  FarBound fb, uint32_t doIllum) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t P = prm.npaths;
    if ((uint64_t)vb.off[P] > vb.cap || i >= (uint64_t)vb.off[P] * fr.nt) return;
    const FarRec& r = vb.far[i];
    const f3 pf = mk((float)r.px, (float)r.py, (float)r.pz);
    const bool in = inside<3, float>(m, pf);
    uint32_t c[3];
    float blo[3], bhi[3];
    const uint64_t box = in ? farSubBoxOf(m, pf, fb.sub, c, blo, bhi) : 0;
#pragma unroll
    for (int role = 0; role < 2; ++role) {
        uint32_t live = 0xffffffffu;
        if (in && fb.table[role] && (role == 0 || doIllum)) {
            const float skip = farRoleSkip(prm, role);
            const uint32_t nc = fb.cols[role];
            const float* t = fb.table[role] + box * nc;
            live = 0u;
            for (uint32_t j = 0; j < nc; ++j) live |= (t[j] > skip ? 0u : 1u) << j;
        }
        vb.live[2 * i + role] = live;
    }
}

// FarBound::table of one role: column j of every sub-box, one thread per (box, column).
__global__ void __launch_bounds__(256) k_far_table(MediumView m, MinPyramid mp, const BoundDirs* blockDirs,
                                                   uint32_t blocks, uint32_t cols, uint32_t sub, uint32_t span,
                                                   float* table) {
    const uint64_t nx = (uint64_t)m.n[0] * sub, ny = (uint64_t)m.n[1] * sub, nz = (uint64_t)m.n[2] * sub;
    const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= nx * ny * nz * cols) return;
    const uint64_t box = j / cols;
    const uint32_t col = (uint32_t)(j % cols), last = col + 1 < cols ? col + 1 : blocks;
    const uint64_t bx[3] = {box % nx, box / nx % ny, box / (nx * ny)};
    uint32_t c[3];
    float blo[3], bhi[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = (uint32_t)(bx[a] / sub);
        farSubBoxFaces(m, a, c[a], (uint32_t)(bx[a] % sub), sub, &blo[a], &bhi[a]);
    }
    float lb = 3.0e38f;
    for (uint32_t b = col; b < last; ++b) lb = fminf(lb, farBoxBound(m, mp, blockDirs[b], c, blo, bhi, span));
    table[j] = lb;
}

hipError_t launchFarTable(const MediumView& m, const MinPyramid& mp, const BoundDirs* blockDirs, uint32_t blocks,
                          uint32_t cols, uint32_t sub, uint32_t span, float* table, hipStream_t st) {
    const uint64_t n = (uint64_t)m.n[0] * m.n[1] * m.n[2] * sub * sub * sub * cols;
    if (!n || cols > 32 || cols > blocks || !sub || !span) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_far_table, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, m, mp, blockDirs, blocks,
                       cols, sub, span, table);
    return hipGetLastError();
}

// HG amplitude constants per material for the far-field view kernel: sqrt((1 - g^2) / 4 pi),
// (1 - g)^2 and 2 g, so that |f| = C den^(-3/4) with den = (1 - g)^2 + 2 g (1 - c). An isotropic
// material is g = 0.
struct HgConst {
    float c, a, b;
};

// Multiple-scattering view terms B for far-field sources on equally spaced grids, FP32 terms,
// HG or isotropic materials, one wavefront per work unit (the common case of the measured
// transmission-matrix workload; other cases take k_nee_view). Same unit and sums as k_nee_view,
// with the per-(vertex, source) work cut to what depends on both:
//   * the walk's direction part (inverse direction, face step, bit stride), the FP64 phase
//     direction d / lambda and the HG incoming direction are formed once per lane, before the
//     vertex loop; the vertex part (the walk's start cell and face offsets, the material, the back
//     direction, the FP64 position and phase base) is one wave-uniform record (FarRec, written by
//     k_vertex_prep), so a vertex costs the lanes two selects and two multiply-adds per axis;
//   * 1 - cos of the HG amplitude is 0.5 |a - b|^2 in FP32, which keeps its relative accuracy in
//     the forward peak where 1 + g^2 - 2 g c cancels; |f| = C rsqrt(den) sqrt(rsqrt(den));
//   * exp(i 2 pi phase) is the hardware sine and cosine (argument in turns) of the FP64-reduced
//     phase (measured on gfx950: within 1.3e-7 of sin, cos over [0, 1));
//   * a vertex whose every lane's term is zero (past the attenuation cut, or outside) skips the
//     amplitude and phase.
template <int SRC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_nee_view_far(MediumView m, ScatterView sc, SourceView vw, FramesView fr, VertexBatch vb, NeeParams prm, NeeOut out,
               FarBound fb) {
    static_assert(SRC == kSrcFarView || SRC == kSrcFarIllum, "far-field sources only");
    constexpr int role = SRC == kSrcFarIllum ? 1 : 0;
    __shared__ float sigLds[kSigLds];
    __shared__ HgConst hg[kMaxMaterials];
    if (threadIdx.x < kMaxMaterials) {
        const double g = sc.type[threadIdx.x] == 3 ? sc.gd[threadIdx.x] : 0.0;
        hg[threadIdx.x] = HgConst{(float)sqrt((1.0 - g * g) / 12.566370614359172954), (float)((1.0 - g) * (1.0 - g)),
                                  (float)(2.0 * g)};
    }
    const float* sig = loadSigTable<float>(m, sigLds);   // synchronises the block
    const ViewUnit u = viewUnit<true>(prm);
    const uint32_t p = u.p, t = u.chunk, P = prm.npaths, V = vw.n;
    const bool poolOk = (uint64_t)vb.off[P] <= vb.cap;
    const uint32_t K = (p < P && poolOk) ? vb.k[p] : 0u;
    const uint64_t off = p < P ? vb.off[p] : 0;
    const float tauCut = viewCut<float>(prm);
    float2* B = (float2*)out.B;
    uint32_t steps = 0;
    // this lane's source: position blk * 64 + lane of the kernel's order (a compact cone per wavefront)
    const uint32_t pos = u.blk * 64u + u.sub;
    const bool act = pos < V;
    const uint32_t v = act ? (fb.perm[role] ? fb.perm[role][pos] : pos) : 0u;
    if (K >= 2 && u.blk * 64u < V) {
        const d3 vd = sourceD<3>(vw, v);
        const f3 vf = sourceF(vw, v);
        // walk along the view direction (view) or against it (illumination); the phase is
        // -d.x / lambda (view) or +d.x / lambda (illumination); the HG incoming direction is -d
        // (view) or d (illumination) (FarField.cuh:223-263, 546-599)
        const float sgn = SRC == kSrcFarIllum ? -1.f : 1.f;
        const WalkLane wl = walkLane<3>(m, mk(sgn * vf.x, sgn * vf.y, sgn * vf.z));
        const double ks = (SRC == kSrcFarIllum ? 1.0 : -1.0) / prm.lambda;
        const d3 kd = mkd(vd.x * ks, vd.y * ks, vd.z * ks);
        const float as = SRC == kSrcFarIllum ? 1.f : -1.f;
        const f3 av = mk(as * (float)vd.x, as * (float)vd.y, as * (float)vd.z);
        float accRe = 0.f, accIm = 0.f;
        // vertices in runs of 64: every lane reads one vertex's live-block mask, and the run's
        // vertices whose block is alive are visited in increasing k (the same order and sums as a
        // visit of every vertex)
        const uint32_t bsel = u.blk < 31u ? u.blk : 31u;
        for (uint32_t c0 = 1; c0 < K; c0 += 64) {
            const uint32_t kl = c0 + u.sub;
            const uint32_t bits = kl < K ? vb.live[2 * ((off + kl) * fr.nt + t) + role] : 0u;
            uint64_t alive = __ballot((bits >> bsel) & 1u);
            while (alive) {
                const uint32_t k = c0 + (uint32_t)__builtin_ctzll(alive);
                alive &= alive - 1;
                const FarRec rec = vb.far[(off + k) * fr.nt + t];
                float tau = 3.0e38f;
                if (act)
                    tau = rec.cell == kFarOutside ? 0.f
                                                  : walkFromStartLane<3>(m, sig + kMaxMaterials, rec.dlo, rec.dhi, rec.b,
                                                                         rec.cell, wl, tauCut, &steps);
                const float amp = tau > tauCut ? 0.f : __expf(-0.5f * tau);
                if (__ballot(amp != 0.f) == 0) continue;   // every term of this vertex is zero
                const HgConst h = hg[rec.mat];
                const float dx = av.x - rec.back[0], dy = av.y - rec.back[1], dz = av.z - rec.back[2];
                const float den = fmaf(h.b, 0.5f * fmaf(dx, dx, fmaf(dy, dy, dz * dz)), h.a);
                const float r = __builtin_amdgcn_rsqf(den);
                const float fv = h.c * r * __builtin_amdgcn_sqrtf(r);
                const double ph = kd.x * rec.px + kd.y * rec.py + kd.z * rec.pz + rec.phase;
                const float turns = (float)(ph - floor(ph));
                const float a = amp * fv;
                accRe = fmaf(a, __builtin_amdgcn_cosf(turns), accRe);
                accIm = fmaf(a, __builtin_amdgcn_sinf(turns), accIm);
            }
        }
        if (act) B[((size_t)t * P + p) * V + v] = make_float2(accRe, accIm);
        if (role == 0 && vb.rowNz && __ballot(accRe != 0.f || accIm != 0.f) && u.sub == 0) vb.rowNz[p] = 1u;
    }
    uint32_t tot = waveSum(steps);
    if ((threadIdx.x & 63) == 0 && out.steps) slotAdd(out.steps + kCountSlots * kCountStride, tot);
}

// Correlation view terms sum_k G_k(0) conj G_k(t) (the per-sub-path estimator). Work unit =
// (path, chunk of up to kCorrLags frames, block of lppView views): one view per lane; per vertex,
// G_k(0) is formed once and reused for every frame of the chunk, whose running sums live in LDS
// (one slot per thread and frame).
template <int DIMS, typename R, int SRC, bool WAVE, bool TAB>
__global__ void __launch_bounds__(256) k_nee_view_corr(MediumView m, ScatterView sc, SourceView vw, FramesView fr,
                                                       VertexBatch vb, NeeParams prm, NeeOut out) {
    typedef typename Cx<R>::T C;
    __shared__ R sigLds[kSigLds];
    __shared__ C acc[kCorrLags][256];
    const R* sig = loadSigTable<R>(m, sigLds);
    const ViewUnit u = viewUnit<WAVE>(prm);
    const uint32_t p = u.p, P = prm.npaths, V = vw.n, NT = fr.nt;
    const bool poolOk = (uint64_t)vb.off[P] <= vb.cap;
    const uint32_t K = (p < P && poolOk) ? vb.k[p] : 0u;
    const uint64_t off = p < P ? vb.off[p] : 0;
    const uint32_t t0 = u.chunk * kCorrLags, nt = NT - t0 < kCorrLags ? NT - t0 : kCorrLags;
    const double invLambda = 1.0 / prm.lambda;
    const R tauCut = viewCut<R>(prm);
    C* B = (C*)out.B;
    const bool isoPoint = SRC == kSrcPoint && sc.allIsotropic;
    uint32_t steps = 0;
    const uint32_t v = u.blk * (WAVE ? 64u : prm.lppView) + u.sub;
    const bool active = v < V && K >= 2 && t0 < NT;
    const uint32_t vv = active ? v : 0;
    const d3 vd = sourceD<DIMS>(vw, vv);
    const f3 vf = SRC != kSrcPoint && SRC != kSrcGauss && sizeof(R) == 4 ? sourceF(vw, vv) : mk(0.f, 0.f, 0.f);
    const d3 vdir = sourceDir<SRC>(vw, vv, vd);
    if (active) {
        for (uint32_t j = 0; j < nt; ++j) acc[j][threadIdx.x] = Cx<R>::make(0, 0);
        for (uint32_t k = 1; k < K; ++k) {
            const uint64_t ok = off + k;
            const float4 xkm = vb.xm[ok];
            const C g0 = viewTerm<DIMS, R, SRC, TAB>(m, sig, sc, vw, fr, vb, vd, vdir, vf, ok, 0, xkm, invLambda,
                                                     tauCut, isoPoint, &steps);
            for (uint32_t j = 0; j < nt; ++j) {
                const uint32_t t = t0 + j;
                const C gt = t == 0 ? g0
                                    : viewTerm<DIMS, R, SRC, TAB>(m, sig, sc, vw, fr, vb, vd, vdir, vf, ok, t, xkm,
                                                                  invLambda, tauCut, isoPoint, &steps);
                C a = acc[j][threadIdx.x];
                a.x += g0.x * gt.x + g0.y * gt.y;      // G(0) * conj(G(t))
                a.y += g0.y * gt.x - g0.x * gt.y;
                acc[j][threadIdx.x] = a;
            }
        }
        for (uint32_t j = 0; j < nt; ++j) B[((size_t)(t0 + j) * P + p) * V + v] = acc[j][threadIdx.x];
    }
    uint32_t tot = waveSum(steps);
    if ((threadIdx.x & 63) == 0 && out.steps) slotAdd(out.steps + kCountSlots * kCountStride, tot);
}

// One view term of the correlation sampler: g of the source at xg times the amplitude at pf
// against the back direction, as amplitude and phase in cycles (the path phase, common to u1 and
// u2, is left out: it cancels in u1 conj(u2)).
template <int DIMS, typename R, int SRC, bool TAB>
__device__ inline void corrViewG(const MediumView& m, const R* sig, const ScatterView& sc, const SourceView& vw,
                                 const d3& vd, const f3& vf, const Geo& xg, const d3& pf, const d3& back, uint32_t mat,
                                 double invLambda, R tauCut, bool isoPoint, uint32_t* steps, R* amp, double* ph) {
    R ga;
    double gp;
    if constexpr (SRC == kSrcPoint) gPoint<DIMS, R>(m, sig, vd, vw.area != 0, xg, invLambda, tauCut, steps, &ga, &gp);
    else gFar<DIMS, R>(m, sig, vd, vf, SRC == kSrcFarIllum, xg, invLambda, tauCut, steps, &ga, &gp);
    R fv;
    if (isoPoint) {
        fv = isoAmp<DIMS, R>();
    } else {
        const d3 av = SRC == kSrcFarIllum  ? vd
                      : SRC == kSrcFarView ? mkd(-vd.x, -vd.y, -vd.z)
                                           : dnormalize<DIMS>(dsub(pf, vd));
        fv = phaseAmp<DIMS, R, TAB>(sc, mat, ddot<DIMS>(av, back), &gp);
    }
    *amp = ga * fv;
    *ph = gp;
}

// Multiple-scattering view terms of the TemporalCorrelationSampler: per (path, chunk of up to
// kCorrLags lags, block of views), one view per lane,
//   B[t][p][v] = sum_{k>=2} G_k,v1(-tau/2) conj G_k,v2(+tau/2) exp(-k^2 |tau| (D_k |q_k|^2 + mt_k))
// with q_k = (e_v1 + e_v2) / 2 - n(x_{k-1} - x_k) (e: incoming direction, minus the direction
// toward the view; C1_v1v2) and
// mt_k = sum_{j=2}^{k-1} D_j |n(x_j - x_{j-1}) - n(x_{j+1} - x_j)|^2 (momentTransfer, applied on u1
// only: pathContriobutionKernel, TemporalSamplerCorrelation.cuh:384-450), all from unshifted
// positions. u1's g sees x_k at -tau/2, u2's at +tau/2; both amplitudes see x_{k-1} and x_k at
// -tau/2 (quirk Q3). Vertices whose own and previous materials are at rest are lag-independent
// except for the exponential: their G products are formed once for all the chunk's lags (once,
// not twice, when view2 holds the same sources as view). With coherent backscattering the same
// kernel runs over the illuminations (SRC kSrcFarIllum or kSrcPoint) into Br (C4_l1l2), and each
// vertex's u1 and u2 terms G_k,1 (without the inner transfer) and G_k,2 go to out.Gv1 / Gv2 for
// the mixed pairings.
template <int DIMS, typename R, int SRC, bool WAVE, bool TAB>
__global__ void __launch_bounds__(256) k_nee_view_corrsampler(MediumView m, ScatterView sc, SourceView vw,
                                                              FramesView fr, VertexBatch vb, NeeParams prm,
                                                              NeeOut out) {
    typedef typename Cx<R>::T C;
    __shared__ R sigLds[kSigLds];
    __shared__ C acc[kCorrLags][256];
    const R* sig = loadSigTable<R>(m, sigLds);
    const ViewUnit u = viewUnit<WAVE>(prm);
    const uint32_t p = u.p, P = prm.npaths, V = vw.n / 2, NT = fr.nt;
    const bool poolOk = (uint64_t)vb.off[P] <= vb.cap;
    const uint32_t K = (p < P && poolOk) ? vb.k[p] : 0u;
    const uint64_t off = p < P ? vb.off[p] : 0;
    const uint32_t t0 = u.chunk * kCorrLags;
    const uint32_t nt = t0 < NT ? (NT - t0 < kCorrLags ? NT - t0 : kCorrLags) : 0u;
    const double invLambda = 1.0 / prm.lambda;
    const double kw2 = 39.478417604357434 * invLambda * invLambda;
    const R tauCut = viewCut<R>(prm);
    C* B = (C*)out.B;
    C* G1out = (C*)out.Gv1;
    C* G2out = (C*)out.Gv2;
    const bool isoPoint = SRC == kSrcPoint && sc.allIsotropic;
    const bool same = prm.sameViews != 0;
    uint32_t steps = 0;
    const uint32_t v = u.blk * (WAVE ? 64u : prm.lppView) + u.sub;
    const bool active = v < V && K >= 2 && nt > 0;
    const uint32_t vv = active ? v : 0;
    const d3 vd1 = sourceD<DIMS>(vw, vv), vd2 = sourceD<DIMS>(vw, V + vv);
    const f3 vf1 = SRC != kSrcPoint && sizeof(R) == 4 ? sourceF(vw, vv) : mk(0.f, 0.f, 0.f);
    const f3 vf2 = SRC != kSrcPoint && sizeof(R) == 4 ? sourceF(vw, V + vv) : mk(0.f, 0.f, 0.f);
    if (active) {
        for (uint32_t j = 0; j < nt; ++j) acc[j][threadIdx.x] = Cx<R>::make(0, 0);
        double mt = 0;
        d3 prevB = mkd(0.0, 0.0, 0.0);
        float4 xpm = vb.xm[off];
        for (uint32_t k = 1; k < K; ++k) {
            const uint64_t ok = off + k;
            const float4 xkm = vb.xm[ok];
            const uint32_t mat = __float_as_uint(xkm.w), mp = __float_as_uint(xpm.w);
            const VertexRec rec = vb.rec[ok];
            const d3 bk = mkd(rec.bx, rec.by, rec.bz);   // n(x_{k-1} - x_k)
            if (k >= 2) {
                const d3 dq = dsub(bk, prevB);
                mt += fr.Dd[mp] * ddot<DIMS>(dq, dq);
            }
            prevB = bk;
            const d3 xk = mkd(xkm.x, xkm.y, xkm.z);
            const d3 e1 = SRC == kSrcFarIllum  ? vd1
                          : SRC == kSrcFarView ? mkd(-vd1.x, -vd1.y, -vd1.z)
                                               : dnormalize<DIMS>(dsub(xk, vd1));
            const d3 e2 = SRC == kSrcFarIllum  ? vd2
                          : SRC == kSrcFarView ? mkd(-vd2.x, -vd2.y, -vd2.z)
                                               : dnormalize<DIMS>(dsub(xk, vd2));
            const d3 q = mkd(0.5 * (e1.x + e2.x) - bk.x, 0.5 * (e1.y + e2.y) - bk.y, 0.5 * (e1.z + e2.z) - bk.z);
            const double rate = kw2 * (fr.Dd[mat] * ddot<DIMS>(q, q) + mt);   // times |tau|
            const double* Uk = fr.Ud[mat];
            const double* Up = fr.Ud[mp];
            const bool moving = Uk[0] != 0.0 || Uk[1] != 0.0 || Uk[2] != 0.0 || Up[0] != 0.0 || Up[1] != 0.0 ||
                                Up[2] != 0.0;
            if (!moving) {
                const Geo xg = {mk(xkm.x, xkm.y, xkm.z), xk};
                R a1, a2;
                double ph1, ph2;
                corrViewG<DIMS, R, SRC, TAB>(m, sig, sc, vw, vd1, vf1, xg, xk, bk, mat, invLambda, tauCut, isoPoint,
                                             &steps, &a1, &ph1);
                if (same) {
                    a2 = a1;
                    ph2 = ph1;
                } else {
                    corrViewG<DIMS, R, SRC, TAB>(m, sig, sc, vw, vd2, vf2, xg, xk, bk, mat, invLambda, tauCut, isoPoint,
                                                 &steps, &a2, &ph2);
                }
                const C g = cis<R>(a1 * a2, ph1 - ph2);
                if (G1out) {
                    const C g1 = cis<R>(a1, ph1), g2 = cis<R>(a2, ph2);
                    for (uint32_t j = 0; j < nt; ++j) {
                        const size_t o = ((size_t)(t0 + j) * out.gcap + ok) * V + v;
                        G1out[o] = g1;
                        G2out[o] = g2;
                    }
                }
                for (uint32_t j = 0; j < nt; ++j) {
                    const R w = expR<R>((R)(-rate * fabs(fr.ticks[t0 + j])));
                    C c = acc[j][threadIdx.x];
                    c.x += g.x * w;
                    c.y += g.y * w;
                    acc[j][threadIdx.x] = c;
                }
            } else {
                for (uint32_t j = 0; j < nt; ++j) {
                    const double tau = fr.ticks[t0 + j], s = 0.5 * tau;
                    const Geo xa = shiftGeo(xkm, Uk, -s), xb = shiftGeo(xkm, Uk, s);
                    const Geo xpa = shiftGeo(xpm, Up, -s);
                    const d3 back = dnormalize<DIMS>(dsub(xpa.pd, xa.pd));
                    R a1, a2;
                    double ph1, ph2;
                    corrViewG<DIMS, R, SRC, TAB>(m, sig, sc, vw, vd1, vf1, xa, xa.pd, back, mat, invLambda, tauCut,
                                                 isoPoint, &steps, &a1, &ph1);
                    corrViewG<DIMS, R, SRC, TAB>(m, sig, sc, vw, vd2, vf2, xb, xa.pd, back, mat, invLambda, tauCut,
                                                 isoPoint, &steps, &a2, &ph2);
                    const C g = cis<R>(a1 * a2 * expR<R>((R)(-rate * fabs(tau))), ph1 - ph2);
                    if (G1out) {
                        const size_t o = ((size_t)(t0 + j) * out.gcap + ok) * V + v;
                        G1out[o] = cis<R>(a1, ph1);
                        G2out[o] = cis<R>(a2, ph2);
                    }
                    C c = acc[j][threadIdx.x];
                    c.x += g.x;
                    c.y += g.y;
                    acc[j][threadIdx.x] = c;
                }
            }
            xpm = xkm;
        }
        for (uint32_t j = 0; j < nt; ++j) B[((size_t)(t0 + j) * P + p) * V + v] = acc[j][threadIdx.x];
    }
    uint32_t tot = waveSum(steps);
    if ((threadIdx.x & 63) == 0 && out.steps) slotAdd(out.steps + kCountSlots * kCountStride, tot);
}

// Mixed pairings of coherent backscattering under the TemporalCorrelationSampler
// (NextEventEstimation.cuh:1237-1330), which do not factor over (l, v): per path,
//   CBS02 = conj(u2Gl0_l) Gv0_v C2_v1l2(x1) * sum_k Gl_k,l conj(u2Gv_k,v) C2_l1v2(x_k)
//   CBS03 = Gl0_l conj(u2Gv0_v) C3_l1v2(x1) * sum_k conj(u2Gl_k,l) Gv_k,v C3_v1l2(x_k)
// each u1 term at x_k times the inner transfer exp(-k^2 |tau| mt_k), the C factors
// exp(-k^2 |D tau| |(e1 + e2) / 2 - n|^2) of the two sources' incoming directions e at the
// connected vertex and n toward its neighbour on the path (x2 for x1, x_{k-1} for x_k),
// temporalCorrelationSourceContributionMultiple (SourceInterface.cuh:280-352). One thread per
// (l, v) of a tile of tl x (256 / tl), one lag per grid row, paths split over grid rows into
// `part` slices; FP64 sums.
template <int DIMS, typename R>
__global__ void __launch_bounds__(256) k_cbs_cross(SourceView il, SourceView vw, FramesView fr, VertexBatch vb,
                                                   CbsCross c, uint32_t npaths, uint32_t tl, uint32_t splits,
                                                   uint32_t pathsPerSplit, double kw2, double2* __restrict__ part) {
    typedef typename Cx<R>::T C;
    const uint32_t L = il.n / 2, V = vw.n / 2;
    const uint32_t l = blockIdx.x * tl + threadIdx.x % tl;
    const uint32_t v = blockIdx.y * (256 / tl) + threadIdx.x / tl;
    const uint32_t t = blockIdx.z / splits, split = blockIdx.z % splits;
    if (l >= L || v >= V) return;
    const bool poolOk = (uint64_t)vb.off[npaths] <= vb.cap;
    const uint32_t pBegin = split * pathsPerSplit;
    const uint32_t pEnd = poolOk ? min(npaths, pBegin + pathsPerSplit) : 0u;
    const double tau = fr.ticks[t];
    const C* F1l = (const C*)c.F1l;
    const C* F2l = (const C*)c.F2l;
    const C* F1v = (const C*)c.F1v;
    const C* F2v = (const C*)c.F2v;
    const C* Gl1 = (const C*)c.Gl1;
    const C* Gl2 = (const C*)c.Gl2;
    const C* Gv1 = (const C*)c.Gv1;
    const C* Gv2 = (const C*)c.Gv2;
    auto decay = [&](const d3& e1, const d3& e2, const d3& n, double dtD) {
        const d3 q = mkd(0.5 * (e1.x + e2.x) - n.x, 0.5 * (e1.y + e2.y) - n.y, 0.5 * (e1.z + e2.z) - n.z);
        return exp(-dtD * ddot<DIMS>(q, q));
    };
    double accR = 0, accI = 0;
    for (uint32_t p = pBegin; p < pEnd; ++p) {
        const uint32_t K = vb.k[p];
        if (K < 2) continue;
        const uint64_t off = vb.off[p];
        const float4 x1m = vb.xm[off], x2m = vb.xm[off + 1];
        const d3 x1 = mkd(x1m.x, x1m.y, x1m.z);
        const d3 n12 = dnormalize<DIMS>(dsub(mkd(x2m.x, x2m.y, x2m.z), x1));
        const double D1 = kw2 * fabs(fr.Dd[__float_as_uint(x1m.w)] * tau);
        // first connections C2_v1l2 (view u1, illumination u2) and C3_l1v2 at x1 toward x2
        const size_t ol = ((size_t)t * npaths + p) * L + l, ov = ((size_t)t * npaths + p) * V + v;
        const C a1 = F1l[ol], a2 = F2l[ol], c1 = F1v[ov], c2 = F2v[ov];
        const double t2 = decay(incoming<DIMS>(vw, v, x1), incoming<DIMS>(il, L + l, x1), n12, D1);
        const double t3 = decay(incoming<DIMS>(il, l, x1), incoming<DIMS>(vw, V + v, x1), n12, D1);
        // conj(a2) c1 t2 and a1 conj(c2) t3
        const double f2r = ((double)a2.x * c1.x + (double)a2.y * c1.y) * t2;
        const double f2i = ((double)a2.x * c1.y - (double)a2.y * c1.x) * t2;
        const double f3r = ((double)a1.x * c2.x + (double)a1.y * c2.y) * t3;
        const double f3i = ((double)a1.y * c2.x - (double)a1.x * c2.y) * t3;
        double s2r = 0, s2i = 0, s3r = 0, s3i = 0, mt = 0;
        d3 prevB = mkd(0.0, 0.0, 0.0);
        uint32_t mp = __float_as_uint(x1m.w);
        for (uint32_t k = 1; k < K; ++k) {
            const uint64_t ok = off + k;
            const float4 xkm = vb.xm[ok];
            const uint32_t mk = __float_as_uint(xkm.w);
            const VertexRec rec = vb.rec[ok];
            const d3 bk = mkd(rec.bx, rec.by, rec.bz);   // n(x_{k-1} - x_k)
            if (k >= 2) {
                const d3 dq = dsub(bk, prevB);
                mt += fr.Dd[mp] * ddot<DIMS>(dq, dq);
            }
            prevB = bk;
            mp = mk;
            const d3 xk = mkd(xkm.x, xkm.y, xkm.z);
            const double inner = exp(-kw2 * fabs(tau) * mt);
            const double Dk = kw2 * fabs(fr.Dd[mk] * tau);
            const double u2 = decay(incoming<DIMS>(il, l, xk), incoming<DIMS>(vw, V + v, xk), bk, Dk) * inner;
            const double u3 = decay(incoming<DIMS>(vw, v, xk), incoming<DIMS>(il, L + l, xk), bk, Dk) * inner;
            const size_t gl = ((size_t)t * c.gcap + ok) * L + l, gv = ((size_t)t * c.gcap + ok) * V + v;
            const C l1 = Gl1[gl], l2 = Gl2[gl], v1 = Gv1[gv], v2 = Gv2[gv];
            // Gl_k conj(u2Gv_k) and conj(u2Gl_k) Gv_k
            s2r += ((double)l1.x * v2.x + (double)l1.y * v2.y) * u2;
            s2i += ((double)l1.y * v2.x - (double)l1.x * v2.y) * u2;
            s3r += ((double)l2.x * v1.x + (double)l2.y * v1.y) * u3;
            s3i += ((double)l2.x * v1.y - (double)l2.y * v1.x) * u3;
        }
        accR += f2r * s2r - f2i * s2i + f3r * s3r - f3i * s3i;
        accI += f2r * s2i + f2i * s2r + f3r * s3i + f3i * s3r;
    }
    part[((size_t)blockIdx.z * V + v) * L + l] = make_double2(accR, accI);
}

// U[t][v][l] += 1/2 sum over slices (in order) of part[t][slice][v][l]
__global__ void __launch_bounds__(256) k_cbs_cross_reduce(const double2* __restrict__ part, uint32_t splits,
                                                          uint32_t LV, uint32_t nt, double* __restrict__ U) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)LV * nt) return;
    const uint32_t t = (uint32_t)(i / LV), e = (uint32_t)(i % LV);
    double re = 0, im = 0;
    for (uint32_t s = 0; s < splits; ++s) {
        const double2 x = part[((size_t)t * splits + s) * LV + e];
        re += x.x;
        im += x.y;
    }
    U[2 * i] += 0.5 * re;
    U[2 * i + 1] += 0.5 * im;
}

uint32_t cbsCrossSplits(uint32_t L, uint32_t V, uint32_t nt, uint32_t npaths) {
    uint32_t tl = 1;
    while (tl < 16 && tl < L) tl *= 2;
    const uint64_t tiles = (uint64_t)((L + tl - 1) / tl) * ((V + 256 / tl - 1) / (256 / tl)) * nt;
    uint64_t s = (2048 + tiles - 1) / tiles;
    if (s > 64) s = 64;
    if (s > npaths / 16) s = npaths / 16;
    return s < 1 ? 1 : (uint32_t)s;
}

template <int DIMS, typename R>
static void launchCbsCrossT(const SourceView& il, const SourceView& vw, const FramesView& fr, const VertexBatch& vb,
                            const CbsCross& c, uint32_t npaths, double lambda, double* U, double2* part,
                            uint32_t maxSplits, hipStream_t st) {
    const uint32_t L = il.n / 2, V = vw.n / 2, nt = fr.nt;
    uint32_t tl = 1;
    while (tl < 16 && tl < L) tl *= 2;
    uint32_t splits = cbsCrossSplits(L, V, nt, npaths);
    if (splits > maxSplits) splits = maxSplits;
    const uint32_t pps = (npaths + splits - 1) / splits;
    const double kw2 = 39.478417604357434 / (lambda * lambda);
    const dim3 grid((L + tl - 1) / tl, (V + 256 / tl - 1) / (256 / tl), nt * splits);
    hipLaunchKernelGGL((k_cbs_cross<DIMS, R>), grid, dim3(256), 0, st, il, vw, fr, vb, c, npaths, tl, splits, pps,
                       kw2, part);
    const uint64_t n = (uint64_t)L * V * nt;
    hipLaunchKernelGGL(k_cbs_cross_reduce, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, part, splits, L * V,
                       nt, U);
}

hipError_t launchCbsCross(int dims, int fp64, const SourceView& il, const SourceView& vw, const FramesView& fr,
                          const VertexBatch& vb, const CbsCross& c, uint32_t npaths, double lambda, double* U,
                          double2* part, uint32_t maxSplits, hipStream_t st) {
    if (npaths == 0) return hipSuccess;
    if (dims == 2) {
        if (fp64) launchCbsCrossT<2, double>(il, vw, fr, vb, c, npaths, lambda, U, part, maxSplits, st);
        else launchCbsCrossT<2, float>(il, vw, fr, vb, c, npaths, lambda, U, part, maxSplits, st);
    } else {
        if (fp64) launchCbsCrossT<3, double>(il, vw, fr, vb, c, npaths, lambda, U, part, maxSplits, st);
        else launchCbsCrossT<3, float>(il, vw, fr, vb, c, npaths, lambda, U, part, maxSplits, st);
    }
    return hipGetLastError();
}

// The reference of the relative attenuation cut: the batch's least view optical depth. One
// wavefront per sampled vertex (every vertexStride-th of the batch's pool, frame 0), one view per
// lane (every view when there are at most 64, else 64 spread evenly), FP32 walks. A walk stops once
// it passes the least depth the batch has published when its wavefront starts, which cannot change
// the minimum: the least pair is walked to its end. The result is the exact minimum over the pairs
// evaluated (whatever order the wavefronts run in), hence at or above the least depth over all
// pairs, so the cut tauRef + delta never drops a term that the same cut about the true least depth
// would keep.
template <int DIMS, int SRC>
__global__ void __launch_bounds__(256) k_tau_ref(MediumView m, SourceView vw, FramesView fr, VertexBatch vb,
                                                 NeeParams prm, uint32_t vertexStride) {
    __shared__ float sigLds[kSigLds];
    const float* sig = loadSigTable<float>(m, sigLds);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t unit = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    const uint64_t o = (uint64_t)unit * vertexStride;
    const uint64_t nv = vb.off[prm.npaths];
    if (nv > vb.cap || o >= nv) return;
    const uint32_t n = vw.n;
    unsigned int* ref = (unsigned int*)prm.tauRef;   // non-negative floats order as their bits
    const float best = __uint_as_float(__hip_atomic_load(ref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const bool active = lane < n;
    const uint32_t v = !active ? 0u : (n <= 64 ? lane : (uint32_t)(((uint64_t)lane * n) / 64u));
    const Geo x = vertexFrame(vb, fr, o, 0, vb.xm[o]);
    float tau = 3.0e38f;
    if (active) {
        const d3 vd = sourceD<DIMS>(vw, v);
        f3 w;
        if constexpr (SRC == kSrcPoint) {
            const d3 q = dsub(vd, x.pd);
            const double rr = 1.0 / sqrt(ddot<DIMS>(q, q));
            w = mk((float)(q.x * rr), (float)(q.y * rr), (float)(q.z * rr));
        } else if constexpr (SRC == kSrcGauss) {
            const d3 b = gbDir(vw.g, v);
            w = mk((float)b.x, (float)b.y, (float)b.z);
        } else {
            w = sourceF(vw, v);
        }
        tau = walk<DIMS, float>(m, sig, x.pf, w, best, nullptr);
    }
#pragma unroll
    for (int s2 = 32; s2 > 0; s2 >>= 1) tau = fminf(tau, __shfl_xor(tau, s2, 64));
    if (lane == 0 && tau < best) atomicMin(ref, __float_as_uint(tau));
}

template <int DIMS>
static void launchTauRef(const MediumView& m, const SourceView& vw, const FramesView& fr, const VertexBatch& vb,
                         const NeeParams& p, hipStream_t st) {
    // about as many vertices as 4096 whole paths hold
    const uint32_t stride = p.npaths > 4096 ? p.npaths / 4096 : 1u;
    const uint64_t units = (vb.cap + stride - 1) / stride;
    const dim3 grid((uint32_t)((units + 3) / 4)), block(256);
    if (vw.type == 2) hipLaunchKernelGGL((k_tau_ref<DIMS, kSrcFarView>), grid, block, 0, st, m, vw, fr, vb, p, stride);
    else if (vw.type == 1) hipLaunchKernelGGL((k_tau_ref<DIMS, kSrcPoint>), grid, block, 0, st, m, vw, fr, vb, p, stride);
    else hipLaunchKernelGGL((k_tau_ref<DIMS, kSrcGauss>), grid, block, 0, st, m, vw, fr, vb, p, stride);
}

// Lanes per view-kernel work unit for n sources: a wavefront, halved (down to 8) while half still
// covers them, so that several units share a wavefront when there are few sources.
static inline void viewLanes(uint32_t n, NeeParams* q) {
    q->lppView = 64;
    while (q->lppView > 8 && q->lppView / 2 >= n) q->lppView /= 2;
    q->viewBlocks = (n + q->lppView - 1) / q->lppView;
}

template <int DIMS, typename R, int SRC>
static void launchView(bool corr, const MediumView& m, const ScatterView& sc, const SourceView& src,
                       const FramesView& fr, const VertexBatch& vb, const NeeParams& q, const NeeOut& out,
                       hipStream_t st, const FarBound* fb = nullptr) {
    const uint64_t units = (uint64_t)q.npaths * q.viewBlocks * q.frameChunks, upb = 4 * (64 / q.lppView);
    const dim3 grid((uint32_t)((units + upb - 1) / upb)), block(256);
    const bool wave = q.lppView == 64, tab = sc.tab != nullptr;
#define SSTMC_VIEW(K, W, T) hipLaunchKernelGGL((K<DIMS, R, SRC, W, T>), grid, block, 0, st, m, sc, src, fr, vb, q, out)
    if (corr) {
        if constexpr (SRC == kSrcGauss) {   // GaussianBeam sources imply vMF tables
            if (wave) SSTMC_VIEW(k_nee_view_corr, true, true);
            else SSTMC_VIEW(k_nee_view_corr, false, true);
        } else if constexpr (SRC != kSrcFarIllum) {
            if (wave) { if (tab) SSTMC_VIEW(k_nee_view_corr, true, true); else SSTMC_VIEW(k_nee_view_corr, true, false); }
            else { if (tab) SSTMC_VIEW(k_nee_view_corr, false, true); else SSTMC_VIEW(k_nee_view_corr, false, false); }
        }
    } else if constexpr (SRC == kSrcGauss) {
        if (wave) SSTMC_VIEW(k_nee_view, true, true);
        else SSTMC_VIEW(k_nee_view, false, true);
    } else if constexpr (DIMS == 3 && sizeof(R) == 4 && (SRC == kSrcFarView || SRC == kSrcFarIllum)) {
        // far-field sources, FP32 terms: the specialised kernel when its preconditions hold
        if (wave && !tab && m.walkFast && vb.far)
            hipLaunchKernelGGL((k_nee_view_far<SRC>), grid, block, 0, st, m, sc, src, fr, vb, q, out,
                               fb ? *fb : FarBound{});
        else if (wave) { if (tab) SSTMC_VIEW(k_nee_view, true, true); else SSTMC_VIEW(k_nee_view, true, false); }
        else { if (tab) SSTMC_VIEW(k_nee_view, false, true); else SSTMC_VIEW(k_nee_view, false, false); }
    } else if (wave) {
        if (tab) SSTMC_VIEW(k_nee_view, true, true);
        else SSTMC_VIEW(k_nee_view, true, false);
    } else {
        if (tab) SSTMC_VIEW(k_nee_view, false, true);
        else SSTMC_VIEW(k_nee_view, false, false);
    }
#undef SSTMC_VIEW
}

template <int DIMS, typename R, int SRC>
static void launchViewCorrSampler(const MediumView& m, const ScatterView& sc, const SourceView& src,
                                  const FramesView& fr, const VertexBatch& vb, const NeeParams& q, const NeeOut& out,
                                  hipStream_t st) {
    const uint64_t units = (uint64_t)q.npaths * q.viewBlocks * q.frameChunks, upb = 4 * (64 / q.lppView);
    const dim3 grid((uint32_t)((units + upb - 1) / upb)), block(256);
    const bool wave = q.lppView == 64, tab = sc.tab != nullptr;
#define SSTMC_VIEWC(W, T) \
    hipLaunchKernelGGL((k_nee_view_corrsampler<DIMS, R, SRC, W, T>), grid, block, 0, st, m, sc, src, fr, vb, q, out)
    if (wave) { if (tab) SSTMC_VIEWC(true, true); else SSTMC_VIEWC(true, false); }
    else { if (tab) SSTMC_VIEWC(false, true); else SSTMC_VIEWC(false, false); }
#undef SSTMC_VIEWC
}

template <int DIMS, typename R>
static void launchNeeT(const MediumView& m, const ScatterView& sc, const SourceView& il, const SourceView& vw,
                       const FramesView& fr, const VertexBatch& vb, const NeeParams& p, const NeeOut& out,
                       hipStream_t st, hipEvent_t between, const FarBound* fb, int parts) {
    const uint32_t ppb = 4 * (64 / p.lpp);   // paths per 256-thread block
    dim3 grid((p.npaths + ppb - 1) / ppb), block(256);
    // the cut's reference, the per-vertex records and, for the far-field view kernel, which blocks
    // of its sources each record's walks can reach inside the cut: all read the batch's vertices only
    if ((parts & kNeeRefPart) && p.tauRef) launchTauRef<DIMS>(m, vw, fr, vb, p, st);
    if (parts & kNeeRecordPart) {
        FramesView rf = fr;
        if (fr.analytic) rf.nt = 1;   // TemporalCorrelationSampler: vertices never move, one record each
        hipLaunchKernelGGL(k_vertex_prep<DIMS>, dim3((p.npaths + 256 / kPrepLanes - 1) / (256 / kPrepLanes)), block,
                           0, st, m, vb, rf, p.npaths, 1.0 / p.lambda);
        if (DIMS == 3 && !fr.analytic && vb.far && fb && (fb->table[0] || fb->table[1])) {
            const uint64_t n = vb.cap * fr.nt;
            hipLaunchKernelGGL(k_far_live, dim3((uint32_t)((n + 255) / 256)), block, 0, st, m, vb, fr, p, *fb,
                               out.Br ? 1u : 0u);
        }
    }
    if (!(parts & kNeeTermPart)) return;
    if (fr.analytic) {
        if (sc.tab) hipLaunchKernelGGL((k_nee_corr<DIMS, R, true>), grid, block, 0, st, m, sc, il, vw, fr, vb, p, out);
        else hipLaunchKernelGGL((k_nee_corr<DIMS, R, false>), grid, block, 0, st, m, sc, il, vw, fr, vb, p, out);
        if (between) hipEventRecord(between, st);
        if (vw.type == 2) launchViewCorrSampler<DIMS, R, kSrcFarView>(m, sc, vw, fr, vb, p, out, st);
        else launchViewCorrSampler<DIMS, R, kSrcPoint>(m, sc, vw, fr, vb, p, out, st);
        if (out.Br) {
            // coherent backscattering: the same per-vertex sums over the illuminations connected at
            // x_k (the reversed path's last connection) into Br, their u1 / u2 terms into Gl1 / Gl2
            NeeParams q = p;
            viewLanes(il.n / 2, &q);
            q.tauRef = nullptr;
            q.sameViews = p.sameIllums;
            NeeOut o = out;
            o.B = out.Br;
            o.Gv1 = out.Gl1;
            o.Gv2 = out.Gl2;
            if (il.type == 2) launchViewCorrSampler<DIMS, R, kSrcFarIllum>(m, sc, il, fr, vb, q, o, st);
            else launchViewCorrSampler<DIMS, R, kSrcPoint>(m, sc, il, fr, vb, q, o, st);
        }
        return;
    }
    auto first = [&]() {
        if (il.type == 3 && sc.tab)
            hipLaunchKernelGGL((k_nee<DIMS, R, true, true>), grid, block, 0, st, m, sc, il, vw, fr, vb, p, out);
        else if (sc.tab)
            hipLaunchKernelGGL((k_nee<DIMS, R, true, false>), grid, block, 0, st, m, sc, il, vw, fr, vb, p, out);
        else hipLaunchKernelGGL((k_nee<DIMS, R, false, false>), grid, block, 0, st, m, sc, il, vw, fr, vb, p, out);
    };
    const bool corr = p.output == 1;
    auto views = [&]() {
        if (vw.type == 3) launchView<DIMS, R, kSrcGauss>(corr, m, sc, vw, fr, vb, p, out, st);
        else if (vw.type == 2) launchView<DIMS, R, kSrcFarView>(corr, m, sc, vw, fr, vb, p, out, st, fb);
        else launchView<DIMS, R, kSrcPoint>(corr, m, sc, vw, fr, vb, p, out, st);
    };
    // with viewFirst the view terms (which flag the paths with a nonzero B row) precede k_nee;
    // `between` follows the first of the two
    if (p.viewFirst) {
        views();
        if (between) hipEventRecord(between, st);
        first();
    } else {
        first();
        if (between) hipEventRecord(between, st);
        views();
    }
    if (out.Br) {
        // coherent backscattering: the same per-vertex sums with the illumination connected at
        // x_k (the reversed path), into Br
        NeeParams q = p;
        viewLanes(il.n, &q);
        q.frameChunks = fr.nt;
        q.tauRef = nullptr;   // the reversed path's Br stays exact
        NeeOut o = out;
        o.B = out.Br;
        if (il.type == 2) launchView<DIMS, R, kSrcFarIllum>(false, m, sc, il, fr, vb, q, o, st, fb);
        else launchView<DIMS, R, kSrcPoint>(false, m, sc, il, fr, vb, q, o, st);
    }
}

hipError_t launchNee(int dims, int fp64, const MediumView& m, const ScatterView& sc, const SourceView& il,
                     const SourceView& vw, const FramesView& fr, const VertexBatch& vb, const NeeParams& p,
                     const NeeOut& out, hipStream_t st, hipEvent_t between, const FarBound* fb, int parts) {
    // lane groups: k_nee gives each path lpp lanes (its L and V items), the view kernels give each
    // work unit lppView lanes (one source each); both shrink below a wavefront when there are few
    // items so that several paths share one. A view unit is (path, frame, block of views) for
    // field output and (path, chunk of up to kCorrLags frames, block of views) for correlation.
    NeeParams q = p;
    uint32_t want = il.n > vw.n * fr.nt ? il.n : vw.n * fr.nt;
    if (fr.analytic) {   // sources hold u1 then u2; items: 2L probabilities, L x T and V x T terms
        const uint32_t L = il.n / 2, V = vw.n / 2;
        want = il.n;
        if (L * fr.nt > want) want = L * fr.nt;
        if (V * fr.nt > want) want = V * fr.nt;
    }
    q.lpp = 64;
    while (q.lpp > 8 && q.lpp / 2 >= want) q.lpp /= 2;
    const bool corr = p.output == 1 || fr.analytic;
    viewLanes(fr.analytic ? vw.n / 2 : vw.n, &q);
    q.frameChunks = corr ? (fr.nt + kCorrLags - 1) / kCorrLags : fr.nt;
    q.farPerm = (fb && vb.far && vb.live) ? fb->perm[0] : nullptr;
    if (dims == 2) {
        if (fp64) launchNeeT<2, double>(m, sc, il, vw, fr, vb, q, out, st, between, fb, parts);
        else launchNeeT<2, float>(m, sc, il, vw, fr, vb, q, out, st, between, fb, parts);
    } else {
        if (fp64) launchNeeT<3, double>(m, sc, il, vw, fr, vb, q, out, st, between, fb, parts);
        else launchNeeT<3, float>(m, sc, il, vw, fr, vb, q, out, st, between, fb, parts);
    }
    return hipGetLastError();
}

}  // namespace sstmc
