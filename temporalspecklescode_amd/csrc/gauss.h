// gauss.h -- GaussianBeam sources (GaussianBeam.cuh) on the device: the beam's complex
// throughput convolved with the vMF lobes of the scattering material (f, fs; 3D through
// sinh(s) / s, 2D through the modified Bessel function I_0), and in 3D the first-point sampler and
// probability from the per-beam z tables and the second-point proposal (a mixture of vMF lobes
// over lobe pairs; in 2D both points are drawn as from a point source at the focal point). FP64
// throughout; complex arithmetic in a small local type.
#pragma once

#include "common.h"
#include "geometry.h"

namespace sstmc {

struct Cd {
    double re, im;
};
__host__ __device__ inline Cd cd(double r, double i = 0.0) { return Cd{r, i}; }
__host__ __device__ inline Cd operator+(Cd a, Cd b) { return cd(a.re + b.re, a.im + b.im); }
__host__ __device__ inline Cd operator-(Cd a, Cd b) { return cd(a.re - b.re, a.im - b.im); }
__host__ __device__ inline Cd operator*(Cd a, Cd b) { return cd(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__host__ __device__ inline Cd operator*(double s, Cd a) { return cd(s * a.re, s * a.im); }
__host__ __device__ inline Cd cdiv(Cd a, Cd b) {
    const double d = b.re * b.re + b.im * b.im;
    return cd((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
__host__ __device__ inline double cabsd(Cd a) { return sqrt(a.re * a.re + a.im * a.im); }
__host__ __device__ inline Cd cexpd(Cd a) {
    const double e = exp(a.re);
    double sn, cs;
    sincos(a.im, &sn, &cs);
    return cd(e * cs, e * sn);
}
// principal square root (ComplexType.cuh:116-131) and its reciprocal
__host__ __device__ inline Cd csqrtd(Cd a) {
    if (a.re == 0 && a.im == 0) return cd(0, 0);
    const double s = sqrt((fabs(a.re) + sqrt(a.re * a.re + a.im * a.im)) / 2.0);
    if (a.re >= 0) return cd(s, a.im / (2.0 * s));
    return cd(fabs(a.im) / (2.0 * s), copysign(s, a.im));
}
__host__ __device__ inline Cd rcsqrtd(Cd a) { return cdiv(cd(1.0), csqrtd(a)); }
// exp(r + z) - exp(r - z), each term dropped when its real part is at most -10 (ComplexType.cuh:170-203)
__host__ __device__ inline Cd twoSinhd(Cd z, Cd r) {
    const Cd p = r + z, n = r - z;
    Cd out = cd(0, 0);
    if (p.re > -10.0) out = out + cexpd(p);
    if (n.re > -10.0) out = out - cexpd(n);
    return out;
}
__host__ __device__ inline Cd clogd(Cd a) { return cd(log(cabsd(a)), atan2(a.im, a.re)); }

// log I_0(z) for complex z (I_0 is even: the argument is taken with Re z >= 0). For |z| <= 12 the
// power series sum_k (z^2 / 4)^k / (k!)^2, summed until a term is below 1e-17 of the sum (near the
// imaginary axis, where I_0 oscillates like J_0, its terms cancel to about e^|z| 2^-53 relative);
// beyond, the large-argument expansion (DLMF 10.40.5)
//   I_0(z) ~ e^z / sqrt(2 pi z) [ sum_k c_k z^-k + s i e^(-2z) sum_k (-1)^k c_k z^-k ],
//   c_k = ((2k - 1)!!)^2 / (k! 8^k), s = sign(Im z) (0 on the real axis),
// both sums stopped before their least term (about e^(-2 |z|) relative), and taken in log form so
// that large Re z does not overflow.
__host__ __device__ inline Cd logI0(Cd z) {
    if (z.re < 0) z = cd(-z.re, -z.im);
    const double r = cabsd(z);
    if (r <= 12.0) {
        const Cd q = 0.25 * (z * z);
        Cd term = cd(1.0), sum = cd(1.0);
        for (int k = 1; k < 100; ++k) {
            term = (1.0 / ((double)k * (double)k)) * (term * q);
            sum = sum + term;
            if (cabsd(term) < 1e-17 * cabsd(sum)) break;
        }
        return clogd(sum);
    }
    const Cd iz = cdiv(cd(1.0), z);
    Cd s1 = cd(1.0), s2 = cd(1.0), t = cd(1.0);
    double prev = 1.0;
    for (int k = 1; k < 200; ++k) {
        t = ((2.0 * k - 1.0) * (2.0 * k - 1.0) / (8.0 * k)) * (t * iz);
        const double mag = cabsd(t);
        if (mag >= prev || mag < 1e-18) break;
        prev = mag;
        s1 = s1 + t;
        s2 = (k & 1) ? s2 - t : s2 + t;
    }
    Cd sum = s1;
    if (z.im != 0.0) sum = sum + (cd(0.0, z.im > 0 ? 1.0 : -1.0) * cexpd(-2.0 * z)) * s2;
    return z - 0.5 * clogd(6.283185307179586477 * z) + clogd(sum);
}

// complex 3-vectors: real part a, imaginary part b
struct CV {
    d3 a, b;
};
__host__ __device__ inline Cd sumSq(const CV& v) {
    return cd(v.a.x * v.a.x - v.b.x * v.b.x + v.a.y * v.a.y - v.b.y * v.b.y + v.a.z * v.a.z - v.b.z * v.b.z,
              2.0 * (v.a.x * v.b.x + v.a.y * v.b.y + v.a.z * v.b.z));
}
__host__ __device__ inline CV cfmaR(double s, const d3& w, const CV& t) {   // s w + t
    return CV{mkd(fma(s, w.x, t.a.x), fma(s, w.y, t.a.y), fma(s, w.z, t.a.z)), t.b};
}
__host__ __device__ inline CV cfmaC(Cd s, const CV& v, const CV& t) {       // s v + t
    return CV{mkd(s.re * v.a.x - s.im * v.b.x + t.a.x, s.re * v.a.y - s.im * v.b.y + t.a.y,
                  s.re * v.a.z - s.im * v.b.z + t.a.z),
              mkd(s.re * v.b.x + s.im * v.a.x + t.b.x, s.re * v.b.y + s.im * v.a.y + t.b.y,
                  s.re * v.b.z + s.im * v.a.z + t.b.z)};
}
__host__ __device__ inline d3 realMultV(Cd s, const CV& v) {               // Re(s v)
    return mkd(s.re * v.a.x - s.im * v.b.x, s.re * v.a.y - s.im * v.b.y, s.re * v.a.z - s.im * v.b.z);
}
__host__ __device__ inline Cd dotCV(const CV& v, const d3& w) {
    return cd(v.a.x * w.x + v.a.y * w.y + v.a.z * w.z, v.b.x * w.x + v.b.y * w.y + v.b.z * w.z);
}
__host__ __device__ inline double dot3(const d3& a, const d3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 add3(const d3& a, const d3& b) { return mkd(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline d3 sub3(const d3& a, const d3& b) { return mkd(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline d3 scale3(double s, const d3& a) { return mkd(s * a.x, s * a.y, s * a.z); }
__host__ __device__ inline d3 unit3(const d3& a) { return scale3(1.0 / sqrt(dot3(a, a)), a); }

__host__ __device__ inline d3 gbFocal(const GaussBeams& g, uint32_t i) {
    return mkd(g.focal[3 * i], g.focal[3 * i + 1], g.focal[3 * i + 2]);
}
__host__ __device__ inline d3 gbDir(const GaussBeams& g, uint32_t i) {
    return mkd(g.dir[3 * i], g.dir[3 * i + 1], g.dir[3 * i + 2]);
}
// the beam's complex throughput gamma_a d + i k (x0 - focal)
__host__ __device__ inline CV gbThroughput(const GaussBeams& g, uint32_t i, const d3& x0, double k) {
    return CV{scale3(g.gammaA, gbDir(g, i)), scale3(k, sub3(x0, gbFocal(g, i)))};
}

// fGaussianBeam (GaussianBeam.cuh:818-892): sum over the lobes of material m of
// exp(apNorm + log 2 pi + C_j) 2 sinh(s_j) / s_j (3D; in 2D exp(log I_0(s_j) + apNorm + log 2 pi +
// C_j), :876-878) with s_j = sqrt((gamma_j w0 + T).(gamma_j w0 + T)),
// T = gamma_a d + i k (x0 - focal) the beam's complex throughput (k negative for a view).
template <int DIMS = 3>
__device__ inline Cd gbF(const GaussBeams& g, const PhaseTables& t, const d3& focal, const d3& dir, uint32_t m,
                         const d3& x0, const d3& w0, double k) {
    const CV T = CV{scale3(g.gammaA, dir), scale3(k, sub3(x0, focal))};
    const double tc = g.apNorm + 1.8378770664093454836;   // + log(2 pi)
    Cd f = cd(0, 0);
    const uint32_t o = t.vmfOff[m], M = t.vmfM[m];
    for (uint32_t j = 0; j < M; ++j) {
        const double lognu = tc + t.vmfC[o + j];
        const Cd s = csqrtd(sumSq(cfmaR(t.vmfKappa[o + j], w0, T)));
        if constexpr (DIMS == 2) f = f + cexpd(logI0(s) + cd(lognu));
        else if (fabs(s.re) < 0.0001 && fabs(s.im) < 0.0001) f = f + cd(2.0 * exp(lognu));
        else f = f + cdiv(twoSinhd(s, cd(lognu)), s);
    }
    return f;
}

// fsGaussianBeam (GaussianBeam.cuh:894-999): illumination beam l of `a` and view beam v of `b`
// through the lobes of material m at x0 (2D: exp(log I_0(C) + log I_0(sqrt mu) + c + log nu),
// :989-991).
template <int DIMS = 3>
__device__ inline Cd gbFs(const GaussBeams& a, const GaussBeams& b, const PhaseTables& t, uint32_t l, uint32_t v,
                          uint32_t m, const d3& x0, double k) {
    const CV aT = CV{scale3(a.gammaA, gbDir(a, l)), scale3(k, sub3(x0, gbFocal(a, l)))};
    const CV bT = CV{scale3(b.gammaA, gbDir(b, v)), scale3(k, sub3(gbFocal(b, v), x0))};
    const double tc = a.apNorm + b.apNorm + 2.0 * 1.8378770664093454836;
    Cd fs = cd(0, 0);
    const uint32_t o = t.vmfOff[m], M = t.vmfM[m];
    for (uint32_t j = 0; j < M; ++j) {
        double gs = t.vmfKappa[o + j];
        const double lognu = tc + t.vmfC[o + j];
        Cd C, sm, c = cd(0, 0);
        if (fabs(gs) < 0.000001) {
            C = csqrtd(sumSq(aT));
            sm = csqrtd(sumSq(bT));
        } else {
            const Cd gob = gs * rcsqrtd(sumSq(cfmaR(gs, gbDir(b, v), aT)));
            sm = csqrtd(sumSq(cfmaC(gob, aT, bT)));
            const d3 w = realMultV(gob, aT);
            const double ra = 1.0 / sqrt(dot3(w, w));
            gs *= ra;
            C = csqrtd(sumSq(cfmaR(gs, w, aT)));
            c = (-ra) * (gob * dotCV(aT, w));
        }
        if constexpr (DIMS == 2) {
            fs = fs + cexpd(logI0(C) + logI0(sm) + c + cd(lognu));
            continue;
        }
        const bool smZero = fabs(sm.re) <= 0.0001 && fabs(sm.im) <= 0.0001;
        const bool cZero = fabs(C.re) <= 0.0001 && fabs(C.im) <= 0.0001;
        if (!smZero && !cZero) fs = fs + cdiv(twoSinhd(C + sm, c + cd(lognu)), C * sm);
        else if (smZero && cZero) fs = fs + 4.0 * cexpd(c + cd(lognu));
        else if (smZero) fs = fs + cdiv(2.0 * twoSinhd(C, c + cd(lognu)), C);
        else fs = fs + cdiv(2.0 * twoSinhd(sm, cd(lognu)), sm);
    }
    return fs;
}

// log of the 3D vMF normalisation kappa / (4 pi sinh kappa) (vMFnormalizationLog,
// SpecialFunctions.cuh:187-201), with the reference's small- and large-kappa branches.
__host__ __device__ inline double vmfNormLog3(double x) {
    if (x < 1e-12) return -2.5310242469692907930;               // -log(4 pi)
    if (x < 5.0) return log(x) - log(12.566370614359172954 * sinh(x));
    return log(x) - 1.8378770664093454836 - x;
}
// the 2D one, 1 / (2 pi I_0(kappa)) (SpecialFunctions.cuh:202-204)
__host__ __device__ inline double vmfNormLog2(double x) { return -1.8378770664093454836 - logI0(cd(x)).re; }

// One term of the second-point proposal (computeCdfBuffer, GaussianBeam.cuh:489-621): the lobe's
// real axis contribution and log-weight about the mean view direction.
__device__ inline void gbConvLobe(const CV& T, const d3& meanView, double gs, double c, d3* cm, double* rc) {
    if (fabs(gs) < 0.000000001) {
        const Cd ct = csqrtd(sumSq(T));
        *rc = ct.re - log(cabsd(ct)) + c;
        *cm = mkd(0.0, 0.0, 0.0);
    } else {
        const Cd gob = gs * rcsqrtd(sumSq(cfmaR(gs, meanView, T)));
        *cm = realMultV(gob, T);
        const d3 w = scale3(1.0 / sqrt(dot3(*cm, *cm)), *cm);
        const Cd ct = csqrtd(sumSq(cfmaR(gs, w, T)));
        *rc = ct.re - log(cabsd(ct)) + c - dot3(w, *cm);
    }
}

// The lobes of beam l's focal material at x0 (at most kGaussLobes), then the pair weights: the
// caller passes per-lobe buffers.
constexpr uint32_t kGaussLobes = 16;
__device__ inline uint32_t gbConvLobes(const GaussBeams& g, const PhaseTables& t, uint32_t l, const d3& x0, double k,
                                       d3* cm, double* rc) {
    const uint32_t fm = g.focalMat[l];
    const uint32_t o = t.vmfOff[fm], M = t.vmfM[fm];
    const CV T = gbThroughput(g, l, x0, k);
    for (uint32_t j = 0; j < M; ++j) gbConvLobe(T, g.meanView, t.vmfKappa[o + j], t.vmfC[o + j], &cm[j], &rc[j]);
    return M;
}
__device__ inline void gbPair(const d3* cm, const double* rc, uint32_t M, uint32_t q, double gammaA, d3* mu,
                              double* logC, double* w) {
    const uint32_t i = q % M, j = q / M;
    *mu = add3(cm[i], cm[j]);
    const double kap = sqrt(dot3(*mu, *mu));
    *logC = kap < 0.0001 ? 0.079577471545947668 : vmfNormLog3(kap);   // 1 / (4 pi) as written
    *w = exp(rc[i] + rc[j] - *logC - 2.0 * gammaA);
}

// sampleDirectionProbability (GaussianBeam.cuh:693-752): the proposal density toward unit d;
// isotropicGaussianProbability (:1013-1023) when every material scatters isotropically.
__device__ inline double gbPdir(const GaussBeams& g, const PhaseTables& t, bool iso, uint32_t l, const d3& x0,
                                const d3& d) {
    if (iso) return 0.079577471545947668;
    d3 cm[kGaussLobes];
    double rc[kGaussLobes];
    const uint32_t M = gbConvLobes(g, t, l, x0, g.waveK, cm, rc);
    double sum = 0;
    for (uint32_t q = 0; q < M * M; ++q) {
        d3 mu;
        double lc, w;
        gbPair(cm, rc, M, q, g.gammaA, &mu, &lc, &w);
        sum += w;
    }
    double p = 0;
    for (uint32_t q = 0; q < M * M; ++q) {
        d3 mu;
        double lc, w;
        gbPair(cm, rc, M, q, g.gammaA, &mu, &lc, &w);
        p += w / sum * exp(dot3(mu, d) + lc);
    }
    return p;
}

// Material at an FP64 position, 0 outside the box (getMaterialInfo, GaussianBeam.cuh:1642).
template <int DIMS = 3>
__device__ inline uint32_t gbMaterial(const MediumView& m, const d3& p) {
    if (!inside<DIMS, double>(m, p)) return 0;
    return m.grid[cellIndex(m, locate<double>(m, 0, p.x), locate<double>(m, 1, p.y),
                            DIMS == 3 ? locate<double>(m, 2, p.z) : 0)];
}

// firstPointProbabilityKernel (GaussianBeam.cuh:420-487) for beam l at x1 (material m1): the
// mixture-weighted z-table value at the point's bin times the transverse Gaussian, over sigs(m1)
// times the box depth in z. The bin is clamped to B - 1 at dz = zmax (the reference reads one past).
__device__ inline double gbPr(const GaussBeams& g, const PhaseTables& t, uint32_t l, const d3& x1, uint32_t m1) {
    const d3 f = gbFocal(g, l), d = gbDir(g, l);
    const d3 q = sub3(x1, f);
    const d3 o = mkd(d.x * d.x * q.x + d.x * d.y * q.y + d.x * d.z * q.z,
                     d.y * d.x * q.x + d.y * d.y * q.y + d.y * d.z * q.z,
                     d.z * d.x * q.x + d.z * d.y * q.y + d.z * d.z * q.z);
    const double dz = dot3(o, d);
    const d3 zv = sub3(add3(f, scale3(dz, d)), x1);
    const double dxy2 = dot3(zv, zv);
    uint32_t zb;
    if (dz < g.zmin[l]) zb = 0;
    else if (dz > g.zmax[l]) zb = g.B - 1;
    else {
        const double r = (dz - g.zmin[l]) / g.zdt[l];
        zb = r >= (double)(g.B - 1) ? g.B - 1 : (uint32_t)r;
    }
    const double k = g.waveK;
    const uint32_t fm = g.focalMat[l];
    const uint32_t o0 = t.vmfOff[fm], M = t.vmfM[fm], mo = g.mixOff[l];
    double pdf = 0;
    for (uint32_t i = 0; i < M; ++i) {
        const double wz2 = (g.gammaA + fabs(t.vmfKappa[o0 + i])) / (k * k) + dz * dz / g.gammaA;
        pdf += g.apdf[mo + i] * g.zpdf[(size_t)(mo + i) * g.B + zb] * exp(-dxy2 / (2.0 * wz2)) /
               (2.0 * wz2 * 3.14159265358979323846);
    }
    return pdf / g.prDen[m1];
}

// sampleFirstBeam (GaussianBeam.cuh:357-418): a beam, a mixture of its focal material by weight, a
// z bin by that mixture's table moved uniformly within the bin, then a Gaussian transverse offset
// of width w(z) = sqrt((gamma_a + |kappa|) / k^2 + dz^2 / gamma_a) (Box-Muller).
__device__ inline d3 gbSampleFirst(const GaussBeams& g, const PhaseTables& t, uint32_t n, Rng& rng, uint32_t* beam) {
    const uint32_t l = rng.index(n);
    *beam = l;
    const uint32_t fm = g.focalMat[l], mo = g.mixOff[l], M = t.vmfM[fm];
    const uint32_t mix = cdfSearch(g.aicdf + mo, M, (double)rng.uniform());
    const uint32_t zb = cdfSearch(g.zicdf + (size_t)(mo + mix) * g.B, g.B, (double)rng.uniform());
    const d3 f = gbFocal(g, l), d = gbDir(g, l);
    const double u = rng.uniform();
    const d3 z = add3(add3(f, scale3(g.zmin[l] + (double)zb * g.zdt[l], d)), scale3(u * g.zdt[l], d));
    const double k = g.waveK;
    const double gs = fabs(t.vmfKappa[t.vmfOff[fm] + mix]);
    const double dz = dot3(sub3(z, f), d);
    const double wz = sqrt((g.gammaA + gs) / (k * k) + dz * dz / g.gammaA);
    const double u1 = rng.uniform(), u2 = rng.uniform();
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    const d3 b0 = mkd(g.xyb[6 * l], g.xyb[6 * l + 1], g.xyb[6 * l + 2]);
    const d3 b1 = mkd(g.xyb[6 * l + 3], g.xyb[6 * l + 4], g.xyb[6 * l + 5]);
    return add3(z, scale3(wz, add3(scale3(r * cs, b0), scale3(r * sn, b1))));
}

// vector_rotation (VectorType.cuh:261-282) in FP64: a direction at cosine c about prev.
__device__ inline f3 gbRotate(const d3& prev, double c, Rng& rng) {
    const double sa = (rng.uniform() < 0.5f ? -1.0 : 1.0) * sqrt(1.0 - c * c);
    double sp, cp;
    sincospi(2.0 * (double)rng.uniform(), &sp, &cp);
    if (fabs(prev.z) > 1.0 - 1e-12) return mk((float)(sa * cp), (float)(sa * sp), (float)(prev.z < 0 ? -c : c));
    const double stw = sqrt(1.0 - prev.z * prev.z);
    return mk((float)(sa * (prev.x * prev.z * cp - prev.y * sp) / stw + prev.x * c),
              (float)(sa * (prev.y * prev.z * cp + prev.x * sp) / stw + prev.y * c),
              (float)(-sa * cp * stw + prev.z * c));
}

// sampleSecondPoint's direction (sampleCdfBuffer, GaussianBeam.cuh:636-691): a lobe pair by the
// first inclusive partial sum of the normalised weights above u, then a vMF direction about its
// axis (the z axis when |mu| < 1e-4); isotropicGaussianSampling (:1001-1011) when every material
// scatters isotropically.
__device__ inline f3 gbSampleSecond(const GaussBeams& g, const PhaseTables& t, bool iso, uint32_t l, const d3& x0,
                                    Rng& rng) {
    if (iso) return randomDirection<3>(rng);
    d3 cm[kGaussLobes];
    double rc[kGaussLobes];
    const uint32_t M = gbConvLobes(g, t, l, x0, g.waveK, cm, rc);
    double sum = 0;
    for (uint32_t q = 0; q < M * M; ++q) {
        d3 mu;
        double lc, w;
        gbPair(cm, rc, M, q, g.gammaA, &mu, &lc, &w);
        sum += w;
    }
    const double u = rng.uniform();
    double acc = 0;
    uint32_t pick = M * M - 1;
    for (uint32_t q = 0; q < M * M; ++q) {
        d3 mu;
        double lc, w;
        gbPair(cm, rc, M, q, g.gammaA, &mu, &lc, &w);
        acc += w / sum;
        if (u < acc) { pick = q; break; }
    }
    d3 mu;
    double lc, w;
    gbPair(cm, rc, M, pick, g.gammaA, &mu, &lc, &w);
    const double kap = sqrt(dot3(mu, mu));
    mu = kap < 0.0001 ? mkd(0.0, 0.0, 1.0) : scale3(1.0 / kap, mu);
    return gbRotate(mu, vmfCosine3(kap, rng), rng);
}

}  // namespace sstmc
