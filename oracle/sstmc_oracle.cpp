// sstmc_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Scalar FP64 restatement of the reference's next-event-estimation field tracer
// (TemporalSampler + NextEventEstimation, one wavelength; coherent backscattering for field
// output), used only by tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker and the CPU
// baseline. It shares nothing with the HIP product except the config struct declared in
// include/sstmc.h (the boundary format).
//
// Random numbers: the reference draws from per-thread curand XORWOW states whose streams
// depend on compaction order (TracerInterface.cuh:56-63, quirk Q2), so its per-path outputs
// are not reproducible. This restatement and the GPU path both use the counter-based stream
// spec below instead; everything else follows the reference files cited per function.
//
//   Philox4x32-10 (Salmon et al., SC'11): key = (seed lo, seed hi),
//   counter = (path lo, path hi, event, purpose<<24 | block), 4 words per block, words used
//   in order.  event = a for the a-th attempt at the first vertex, 0x80000000|k for the
//   sampling of vertex k >= 2.  purpose 0 = transport, 1 = frame displacements (slot s reads
//   block s: normal u1,u2 then direction words), 2 = random phase (block 0 word 0).
//   uniform(w) = ((w>>8)+0.5)*2^-24 in (0,1);  index(w,N) = ((2*(w>>8)+1)*N) >> 25
//   (= ceil(u*N)-1 of Simulation.cuh:306-309 for that u).
//
// Parity status of each part is recorded in the test files (tests/test_oracle_*.py).

#include "../include/sstmc.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;
const double PI = 3.14159265358979323846;
const double EPS = 1e-12;  // EPSILON for PRECISION==DOUBLE (Simulation.cuh:36-40)

// ------------------------------------------------------------------ Philox4x32-10
struct Philox {
    static void block(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
        uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
        uint32_t k0 = key_in[0], k1 = key_in[1];
        for (int r = 0; r < 10; ++r) {
            uint64_t p0 = (uint64_t)0xD2511F53u * c0;
            uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
            uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
            uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
            uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
            c0 = n0; c1 = n1; c2 = n2; c3 = n3;
            k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }
};

struct Stream {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int pos;
    Stream(uint64_t seed, uint64_t path, uint32_t event, uint32_t purpose) {
        key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32);
        ctr[0] = (uint32_t)path; ctr[1] = (uint32_t)(path >> 32); ctr[2] = event;
        ctr[3] = purpose << 24;
        pos = 4;
    }
    void seek_block(uint32_t b) { ctr[3] = (ctr[3] & 0xff000000u) | b; pos = 4; }
    uint32_t word() {
        if (pos == 4) { Philox::block(ctr, key, buf); ctr[3]++; pos = 0; }
        return buf[pos++];
    }
    double uniform() { return ((double)(word() >> 8) + 0.5) * (1.0 / 16777216.0); }
    uint32_t index(uint32_t n) { return (uint32_t)(((2ull * (word() >> 8) + 1ull) * n) >> 25); }
};

// ------------------------------------------------------------------ small vectors
struct V3 {
    double c[3];
    V3() { c[0] = c[1] = c[2] = 0; }
    V3(double x, double y, double z) { c[0] = x; c[1] = y; c[2] = z; }
    double& operator[](int i) { return c[i]; }
    double operator[](int i) const { return c[i]; }
};
inline V3 add(const V3& a, const V3& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 sub(const V3& a, const V3& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 scale(double s, const V3& a) { return V3(s * a[0], s * a[1], s * a[2]); }

// ------------------------------------------------------------------ complex helpers
// ComplexType.cuh:84-215 and VectorType.cuh:389-520 (complex 3-vectors), as GaussianBeam uses them.
struct CV3 {
    cplx c[3];
};
inline CV3 cvec(const V3& re, const V3& im) { return {{cplx(re[0], im[0]), cplx(re[1], im[1]), cplx(re[2], im[2])}}; }
// cfma(real a, real vector b, complex vector c) = a b + c
inline CV3 cfma(double a, const V3& b, const CV3& c) {
    return {{cplx(std::fma(a, b[0], c.c[0].real()), c.c[0].imag()), cplx(std::fma(a, b[1], c.c[1].real()), c.c[1].imag()),
             cplx(std::fma(a, b[2], c.c[2].real()), c.c[2].imag())}};
}
// cfma(complex a, complex vector b, complex vector c) = a b + c
inline CV3 cfma(cplx a, const CV3& b, const CV3& c) { return {{a * b.c[0] + c.c[0], a * b.c[1] + c.c[1], a * b.c[2] + c.c[2]}}; }
inline cplx sumSquare(const CV3& v) { return v.c[0] * v.c[0] + v.c[1] * v.c[1] + v.c[2] * v.c[2]; }
inline cplx dotR(const CV3& v, const V3& w) { return v.c[0] * w[0] + v.c[1] * w[1] + v.c[2] * w[2]; }
// realMult(a, v): the real part of a v, per component
inline V3 realMult(cplx a, const CV3& v) {
    return V3((a * v.c[0]).real(), (a * v.c[1]).real(), (a * v.c[2]).real());
}
// complexSqrt / rComplexSqrt: the principal square root (boost's eval_sqrt) and its reciprocal
inline cplx csqrtRef(cplx a) {
    if (a.real() == 0 && a.imag() == 0) return cplx(0, 0);
    double s = std::sqrt((std::fabs(a.real()) + std::sqrt(a.real() * a.real() + a.imag() * a.imag())) / 2.0);
    if (a.real() >= 0) return cplx(s, a.imag() / (2.0 * s));
    return cplx(std::fabs(a.imag()) / (2.0 * s), std::copysign(s, a.imag()));
}
inline cplx rcsqrtRef(cplx a) { return 1.0 / csqrtRef(a); }
// complex2timesSinh(z, r) = exp(r + z) - exp(r - z), each exponential dropped when its real part
// is at most -10 (ComplexType.cuh:170-203)
inline cplx twoSinh(cplx z, cplx r) {
    cplx p = r + z, n = r - z, out(0, 0);
    if (p.real() > -10.0) out += std::exp(p);
    if (n.real() > -10.0) out -= std::exp(n);
    return out;
}

// log I_0(z) (the modified Bessel function of the first kind, order 0; the reference's
// logbesseli(0, z), SpecialFunctions.cuh:42-180, takes its own series and asymptotic forms):
// I_0 is even, so Re z >= 0 is taken; |z| <= 12 by the series sum_k (z^2 / 4)^k / (k!)^2, beyond by
// the large-argument expansion I_0(z) ~ e^z / sqrt(2 pi z) [sum c_k z^-k + s i e^(-2z) sum (-1)^k
// c_k z^-k], c_k = ((2k - 1)!!)^2 / (k! 8^k), s = sign(Im z), stopped before the least term
// (DLMF 10.40.5), in log form.
inline cplx logI0(cplx z) {
    if (z.real() < 0) z = -z;
    if (std::abs(z) <= 12.0) {
        const cplx q = 0.25 * z * z;
        cplx term(1.0, 0.0), sum(1.0, 0.0);
        for (int k = 1; k < 100; ++k) {
            term *= q / ((double)k * (double)k);
            sum += term;
            if (std::abs(term) < 1e-17 * std::abs(sum)) break;
        }
        return std::log(sum);
    }
    const cplx iz = 1.0 / z;
    cplx s1(1.0, 0.0), s2(1.0, 0.0), t(1.0, 0.0);
    double prev = 1.0;
    for (int k = 1; k < 200; ++k) {
        t *= iz * ((2.0 * k - 1.0) * (2.0 * k - 1.0) / (8.0 * k));
        if (std::abs(t) >= prev || std::abs(t) < 1e-18) break;
        prev = std::abs(t);
        s1 += t;
        s2 += (k & 1) ? -t : t;
    }
    cplx sum = s1;
    if (z.imag() != 0.0) sum += cplx(0.0, z.imag() > 0 ? 1.0 : -1.0) * std::exp(-2.0 * z) * s2;
    return z - 0.5 * std::log(2.0 * PI * z) + std::log(sum);
}

struct Problem;

// VectorType.cuh:95-165 device semantics: dot, norm, normalize (zero vector -> +z / +y).
struct Geo {
    int dims;
    double dot(const V3& a, const V3& b) const {
        return dims == 3 ? a[0] * b[0] + a[1] * b[1] + a[2] * b[2] : a[0] * b[0] + a[1] * b[1];
    }
    double norm(const V3& a) const { return std::sqrt(dot(a, a)); }
    V3 normalize(const V3& a) const {
        double n = norm(a);
        if (n == 0) return dims == 3 ? V3(0, 0, 1) : V3(0, 1, 0);
        return scale(1.0 / n, a);
    }
};

// ------------------------------------------------------------------ medium
// HeterogeneousMedium.cuh + MediumInterface.cuh.
struct Medium {
    int dims;
    std::vector<double> ax[3];
    std::vector<uint32_t> grid;
    std::vector<double> sigs, siga;
    V3 bmin, bmax;
    double maxSigt;
    uint32_t nbin(int a) const { return (uint32_t)ax[a].size() - 1; }

    // MediumNS::isInsideKernel (MediumInterface.cuh:204-213): closed box.
    bool inside(const V3& p) const {
        for (int a = 0; a < dims; ++a)
            if (!(p[a] >= bmin[a] && p[a] <= bmax[a])) return false;
        return true;
    }
    // binarySearchKernel (Simulation.cuh:325-367), clamped to a valid cell: the reference
    // returns the edge count minus one for a point exactly on the max face (quirk Q13).
    uint32_t bin(int a, double x) const {
        const std::vector<double>& e = ax[a];
        uint32_t n = (uint32_t)e.size();
        if (e[0] > x) return 0;
        if (e[n - 1] <= x) return n - 2;
        uint32_t lo = 0, hi = n - 1;  // e[lo] <= x < e[hi]
        while (hi - lo > 1) {
            uint32_t m = (lo + hi) / 2;
            if (e[m] <= x) lo = m; else hi = m;
        }
        return lo;
    }
    uint32_t cell(const uint32_t* b) const {
        uint32_t idx = b[0] + b[1] * nbin(0);
        if (dims == 3) idx += b[2] * nbin(0) * nbin(1);
        return idx;
    }
    uint32_t material(const V3& p) const {
        uint32_t b[3] = {bin(0, p[0]), bin(1, p[1]), dims == 3 ? bin(2, p[2]) : 0u};
        return grid[cell(b)];
    }
    double sigt(uint32_t m) const { return sigs[m] + siga[m]; }

    // bdDistance (HeterogeneousMedium.cuh:269-383): does the ray from o along d hit the box.
    bool pointsToward(const V3& o_in, const V3& d) const {
        V3 r, o;
        for (int a = 0; a < 3; ++a) {
            r[a] = 0.5 * (bmax[a] - bmin[a]);
            o[a] = o_in[a] - 0.5 * (bmax[a] + bmin[a]);
        }
        double invR[3], invD[3];
        for (int a = 0; a < dims; ++a) {
            invR[a] = std::fabs(r[a]) < EPS ? 1e12 : 1.0 / r[a];
            invD[a] = std::fabs(d[a]) < EPS ? 1e12 : 1.0 / d[a];
        }
        double mx = -1e300;
        for (int a = 0; a < dims; ++a) mx = std::max(mx, std::fabs(o[a]) * invR[a]);
        int winding = mx <= 1.0 ? -1 : 1;
        for (int a = 0; a < dims; ++a) {
            int sgn = d[a] < 0 ? 1 : -1;
            double dist = invD[a] * (r[a] * (winding * sgn) - o[a]);
            bool ok = dist >= 0.0;
            for (int b = 0; b < dims && ok; ++b)
                if (b != a) ok = std::fabs(std::fma(d[b], dist, o[b])) < r[b];
            if (ok) return true;
        }
        return false;
    }

    // heterogeneousSampleKernel (HeterogeneousMedium.cuh:576-678): Woodcock tracking with the
    // global majorant (setMaterial, :820-860), absorption as termination. Returns material (0 =
    // escaped or absorbed) and writes the sampled point.
    int woodcock(const V3& o, const V3& d, Stream& s, V3* out) const {
        double dist = 0.0;
        V3 x;
        uint32_t m = 0;
        for (;;) {
            dist += -std::log(s.uniform()) / maxSigt;
            x = add(o, scale(dist, d));
            if (!inside(x)) {
                if (!pointsToward(x, d)) break;
            } else {
                m = material(x);
                if (s.uniform() <= sigt(m) / maxSigt) break;
            }
        }
        *out = x;
        if (!inside(x)) return 0;
        if (s.uniform() > sigs[m] / sigt(m)) return 0;
        return (int)m;
    }

    // heterogeneousAttenuationKernel (HeterogeneousMedium.cuh:386-573): exact voxel walk from
    // p along d to the box boundary; returns the optical depth. Outside -> 0.
    double opticalDepth(const V3& p_in, const V3& d, uint64_t* steps) const {
        if (!inside(p_in)) return 0.0;
        V3 p = p_in;
        double invD[3];
        int dirsgn[3];
        for (int a = 0; a < dims; ++a) {
            invD[a] = std::fabs(d[a]) < EPS ? (d[a] >= 0 ? 1e12 : -1e12) : 1.0 / d[a];
            dirsgn[a] = d[a] < 0 ? -1 : 1;  // winding * sgn with winding = -1
        }
        uint32_t b[3] = {bin(0, p[0]), bin(1, p[1]), dims == 3 ? bin(2, p[2]) : 0u};
        double st = sigt(grid[cell(b)]);
        double tau = 0.0;
        for (;;) {
            if (steps) ++*steps;
            double dist[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < dims; ++a) {
                double lo = ax[a][b[a]], hi = ax[a][b[a] + 1];
                double r = 0.5 * (hi - lo), o = p[a] - 0.5 * (hi + lo);
                dist[a] = invD[a] * (r * dirsgn[a] - o);
            }
            int ax_hit;
            if (dims == 2) ax_hit = dist[0] < dist[1] ? 0 : 1;
            else if (dist[0] < dist[1] && dist[0] < dist[2]) ax_hit = 0;
            else if (dist[1] < dist[2]) ax_hit = 1;
            else ax_hit = 2;
            double bd = dist[ax_hit];
            tau += bd * st;
            int hit = dirsgn[ax_hit];
            if ((b[ax_hit] == 0 && hit == -1) || (b[ax_hit] == nbin(ax_hit) - 1 && hit == 1)) break;
            p = add(p, scale(bd, d));
            b[ax_hit] = (uint32_t)((int)b[ax_hit] + hit);
            st = sigt(grid[cell(b)]);
        }
        return tau;
    }
};

// ------------------------------------------------------------------ scattering
// Scattering.cuh: isotropic, tabular and Henyey-Greenstein (setup :272-429, :500-572; eval
// :895-1110).
struct Scatter {
    int dims;
    std::vector<int> type;  // per material
    std::vector<double> g;
    bool allIsotropic;      // Scattering::isIsotropicScattering() (Scattering.cuh:1505-1527)
    // tabular: amplitude normalised by the cross section, evaluation pdf and inverse cdf angle
    struct Table {
        std::vector<cplx> f;
        std::vector<double> pdf, icdf;
    };
    std::vector<Table> tab;
    // von Mises-Fisher mixtures (Scattering.cuh:613-884): per mixture kappa and C + log(alpha);
    // per pair i + M j the summed concentration, its log normalisation, the normalised weight and
    // the exclusive cumulative weight
    struct Vmf {
        std::vector<double> kappa, c, pairMu, pairLogNorm, pairPdf, pairCdf;
        bool iso = true;
    };
    std::vector<Vmf> vmf;

    // vMFnormalizationLog in 3D (SpecialFunctions.cuh:187-201): log(kappa / (4 pi sinh kappa))
    // with the reference's small- and large-kappa branches.
    static double vmfNormLog(double x) {
        if (x < EPS) return -std::log(4.0 * PI);
        if (x < 5.0) return std::log(x) - std::log(4.0 * PI * std::sinh(x));
        return std::log(x) - std::log(2.0 * PI) - x;
    }
    void setVmf(uint32_t m, const double* mu, const double* C, const double* alpha, uint32_t M) {
        Vmf& v = vmf[m];
        std::vector<double> w(M * M);
        double sum = 0;
        for (uint32_t i = 0; i < M; ++i) {
            v.kappa.push_back(mu[i]);
            v.c.push_back(C[i] + std::log(alpha[i]));
            if (std::fabs(mu[i]) > 0.0001) v.iso = false;
        }
        for (uint32_t k = 0; k < M * M; ++k) {
            uint32_t i = k % M, j = k / M;
            v.pairMu.push_back(mu[i] + mu[j]);
            v.pairLogNorm.push_back(vmfNormLog(mu[i] + mu[j]));
            w[k] = std::exp(C[i] + std::log(alpha[i]) + C[j] + std::log(alpha[j]) - vmfNormLog(mu[i] + mu[j]));
            sum += w[k];
        }
        double acc = 0;
        for (uint32_t k = 0; k < M * M; ++k) {
            v.pairPdf.push_back(w[k] / sum);
            v.pairCdf.push_back(acc);
            acc += w[k] / sum;
        }
    }
    // binarySearchKernel (Simulation.cuh:325-362)
    static size_t cdfSearch(const std::vector<double>& a, double x) {
        size_t r = a.size() - 1, l = 0;
        if (a[0] > x) return 0;
        if (a[r] < x) return r;
        while (l <= r) {
            size_t m = l + (r - l) / 2;
            if (m == 0) return 0;
            if (a[m] == x) return m;
            if (a[m - 1] <= x && a[m] > x) return m - 1;
            if (a[m] < x) l = m + 1;
            else r = m - 1;
        }
        return 0;
    }
    // random_vMF_cosine in 3D (VectorType.cuh:169-205): Wood's rejection sampler
    static double vmfCosine(double kappa, Stream& s) {
        double b = -kappa + std::hypot(kappa, 1.0);
        double x0 = (1.0 - b) / (1.0 + b);
        double sc = kappa * x0 + 2.0 * std::log(1.0 - x0 * x0);
        double w, t, u;
        do {
            double z = s.uniform();
            u = s.uniform();
            w = (1.0 - (1.0 + b) * z) / (1.0 - (1.0 - b) * z);
            t = kappa * w + 2.0 * std::log(1.0 - x0 * w) - sc;
        } while (t < std::log(u));
        return w;
    }

    // setTabularScattering (Scattering.cuh:272-429): n complex samples over [0, pi] (3D) or
    // [0, 2 pi] (2D). getSamplePdfRaw (:208-222): |f|^2, times sin(theta) in 3D; the sum
    // normalises the sampling pdf (:344-348); the cross section (sum / n) 2 pi (times pi in 3D)
    // normalises f (:351-358); the inverse cdf is lower_bound(cdf, i / (n-1)) scaled to the
    // angle range (:361-373); the evaluation pdf is getSamplePdfRaw of the normalised f (:377).
    void setTabular(uint32_t m, const double* fin, uint32_t n) {
        Table& t = tab[m];
        std::vector<double> raw(n), cdf(n);
        double sum = 0;
        for (uint32_t i = 0; i < n; ++i) {
            double th = ((double)i / (double)(n - 1)) * PI;
            raw[i] = (fin[2 * i] * fin[2 * i] + fin[2 * i + 1] * fin[2 * i + 1]) * (dims == 2 ? 1.0 : std::sin(th));
            sum += raw[i];
        }
        double acc = 0;
        for (uint32_t i = 0; i < n; ++i) { acc += raw[i] / sum; cdf[i] = acc; }
        double cross = (sum / (double)n) * 2.0 * PI * (dims == 2 ? 1.0 : PI);
        double k = std::sqrt(1.0 / cross), range = dims == 2 ? 2.0 * PI : PI;
        t.f.resize(n); t.pdf.resize(n); t.icdf.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            double th = ((double)i / (double)(n - 1)) * PI;
            t.f[i] = cplx(fin[2 * i] * k, fin[2 * i + 1] * k);
            t.pdf[i] = std::norm(t.f[i]) * (dims == 2 ? 1.0 : std::sin(th));
            size_t lb = (size_t)(std::lower_bound(cdf.begin(), cdf.end(), (double)i / (double)(n - 1)) - cdf.begin());
            t.icdf[i] = (double)lb * (range / (double)(n - 1));
        }
    }
    // entry of cosine c: round(safeAcos(c) (n-1) / range) half to even, clamped (Scattering.cuh:
    // 908-915; safeAcos, round_num Simulation.cuh:274-297)
    size_t tabEntry(uint32_t m, double c) const {
        size_t n = tab[m].f.size();
        double a = c < -1.0 + EPS ? PI : (c > 1.0 - EPS ? 0.0 : std::acos(c));
        double r = std::nearbyint(a * (double)(n - 1) / (dims == 2 ? 2.0 * PI : PI));
        return r < 0 ? 0 : (r >= (double)n ? n - 1 : (size_t)r);
    }

    double isoAmp() const { return dims == 3 ? 0.282094791773878143 : 0.398942280401432678; }
    double isoPdf() const { return dims == 3 ? 0.079577471545947668 : 0.159154943091895336; }
    double pdf(uint32_t m, double c) const {
        if (type[m] == SSTMC_SCATTERING_ISOTROPIC || (type[m] == SSTMC_SCATTERING_VMF && dims == 2)) return isoPdf();
        if (type[m] == SSTMC_SCATTERING_TABULAR) return tab[m].pdf[tabEntry(m, c)];
        if (type[m] == SSTMC_SCATTERING_VMF) {   // Scattering.cuh:1003-1017
            const Vmf& v = vmf[m];
            double p = 0;
            for (size_t k = 0; k < v.pairMu.size(); ++k) p += v.pairPdf[k] * std::exp(c * v.pairMu[k] + v.pairLogNorm[k]);
            return p;
        }
        double gg = g[m];
        double up = dims == 3 ? (1.0 - gg * gg) / (4.0 * PI) : 2.0 * ((1.0 - gg * gg) / (4.0 * PI));
        double den = std::fma(-2.0 * gg, c, 1.0 + gg * gg);
        return dims == 3 ? up / (den * std::sqrt(den)) : up / den;
    }
    cplx amp(uint32_t m, double c) const {
        if (type[m] == SSTMC_SCATTERING_ISOTROPIC || (type[m] == SSTMC_SCATTERING_VMF && dims == 2)) return cplx(isoAmp());
        if (type[m] == SSTMC_SCATTERING_TABULAR) return tab[m].f[tabEntry(m, c)];
        if (type[m] == SSTMC_SCATTERING_VMF) {   // Scattering.cuh:932-945
            double a = 0;
            for (size_t i = 0; i < vmf[m].kappa.size(); ++i) a += std::exp(c * vmf[m].kappa[i] + vmf[m].c[i]);
            return cplx(a);
        }
        return cplx(std::sqrt(pdf(m, c)));
    }
    // randomDirection (VectorType.cuh:316-362).
    V3 randomDirection(Stream& s) const {
        if (dims == 2) {
            double th = 2.0 * s.uniform();
            return V3(std::sin(PI * th), std::cos(PI * th), 0);
        }
        double ph = 2.0 * s.uniform();
        double sp = std::sin(PI * ph), cp = std::cos(PI * ph);
        double ct = std::fma(s.uniform(), 2.0, -1.0);
        double st = std::sqrt(1.0 - ct * ct);
        return V3(st * cp, st * sp, ct);
    }
    // sampleDeviceImplementation (Scattering.cuh:1091-1109) with sampleCosine (:1009-1063)
    // and vector_rotation (VectorType.cuh:261-282).
    V3 sample(uint32_t m, const V3& prev, Stream& s) const {
        if (type[m] == SSTMC_SCATTERING_ISOTROPIC || (type[m] == SSTMC_SCATTERING_VMF && dims == 2))
            return randomDirection(s);
        double gg = g[m], c;
        if (type[m] == SSTMC_SCATTERING_VMF) {   // Scattering.cuh:1066-1077
            size_t j = cdfSearch(vmf[m].pairCdf, s.uniform());
            c = vmfCosine(vmf[m].pairMu[j], s);
        } else if (type[m] == SSTMC_SCATTERING_TABULAR) {
            // inverse cdf at round_num(u n) (Scattering.cuh:1030-1037)
            size_t n = tab[m].icdf.size();
            double r = std::nearbyint(s.uniform() * (double)n);
            c = std::cos(tab[m].icdf[r >= (double)n ? n - 1 : (size_t)r]);
        } else if (std::fabs(gg) < 0.001) {
            c = s.uniform() * 2.0 - 1.0;
        } else {
            double u = s.uniform();
            if (dims == 2) {
                c = std::cos(2.0 * std::atan(((1.0 - gg) / (1.0 + gg)) * std::tan((PI / 2.0) * (1 - 2.0 * u))));
            } else {
                double sv = (1.0 - gg * gg) / (1.0 - gg + 2.0 * gg * u);
                c = (1.0 + gg * gg - sv * sv) / (2.0 * gg);
            }
        }
        return rotate(prev, c, s);
    }
    // vector_rotation (VectorType.cuh:261-282): a direction at cosine c about prev.
    V3 rotate(const V3& prev, double c, Stream& s) const {
        double sa = (s.uniform() < 0.5 ? -1.0 : 1.0) * std::sqrt(1.0 - c * c);
        if (dims == 2)
            return V3(prev[0] * c - prev[1] * sa, prev[0] * sa + prev[1] * c, 0);
        double ph = 2.0 * s.uniform();
        double sp = std::sin(PI * ph), cp = std::cos(PI * ph);
        if (std::fabs(prev[2]) > 1.0 - EPS)
            return V3(sa * cp, sa * sp, c * (std::signbit(prev[2]) ? -1.0 : 1.0));
        double stw = std::sqrt(1.0 - prev[2] * prev[2]);
        return V3(sa * (prev[0] * prev[2] * cp - prev[1] * sp) / stw + prev[0] * c,
                  sa * (prev[1] * prev[2] * cp + prev[0] * sp) / stw + prev[1] * c,
                  -sa * cp * stw + prev[2] * c);
    }
};

// ------------------------------------------------------------------ sources
// FarField.cuh (:71-263, :439-599) and PointSource.cuh (:72-172, :236-446).
struct Source {
    int type = SSTMC_SOURCE_NONE;
    bool illum = true;
    bool area = false;
    std::vector<V3> pt;        // FarField: normalised directions d; PointSource: locations
    // FarField sampling tables (fillBuffers, FarField.cuh:71-127)
    std::vector<V3> invDir, tangent, perpT, perpS;
    std::vector<double> minT, maxT, minS, maxS, prob;
    // GaussianBeam (GaussianBeam.cuh, 3D): pt holds the focal points, dir the beam directions as
    // given; gammaA = 1 / aperture^2, apNorm = vMFnormalizationLog(gammaA) (:1053-1062)
    std::vector<V3> dir;
    double gammaA = 0, apNorm = 0;
    // illumination sampling tables, per beam (updateSampler, :1281-1613; zSamplesNum = 1000)
    uint32_t zB = 1000;
    std::vector<uint32_t> focalMat;
    std::vector<double> zmin, zmax, zdt;
    std::vector<V3> xyb0, xyb1;
    std::vector<std::vector<double>> zpdf, zicdf, apdf, aicdf;   // [beam][mix * zB + b], [beam][mix]
    uint32_t n() const { return (uint32_t)pt.size(); }
};

struct Frames {
    uint32_t nt = 1;
    uint32_t neg = 0;
    std::vector<double> dt;    // per chain slot (TemporalSampler.cuh:1093-1132)
    std::vector<double> D;     // per material
    std::vector<V3> U;
    bool temporalAmp = true;
};

struct Vertex {
    V3 x;                      // central position
    V3 dirIn;                  // free-flight direction that reached x
    uint32_t m;
    double rho;
    std::vector<V3> xt;        // per output frame
    std::vector<double> len;   // path length from x1 per frame
};

struct Problem {
    int dims;
    Geo geo;
    Medium med;
    Scatter sc;
    Source il, vw;
    Frames fr;
    double lambda;
    uint32_t P, fullIter;
    uint64_t seed;
    int output;                // SSTMC_OUTPUT_FIELD / _CORRELATION
    bool cbs = false;          // coherent backscattering: add the reversed path of every path
    // precision.tauCut = delta: an unweighted view-side term (g of a view) whose optical depth
    // exceeds tauRef + delta is dropped (delta = 0: none), tauRef being the batch's least view
    // optical depth over the sample of tauReference; terms that carry the path weight
    // (illumination terms at x1, the reversed path's view term at x1) and the reversed path's
    // illumination terms are never cut. Not in the reference (it walks every term to the box);
    // restated from include/sstmc.h so the GPU's cut mode can be checked term by term.
    double tauCut = 0.0;
    double tauRef = 0.0;
    double cutT(double tau, bool cut) const {
        return (cut && tauCut > 0 && tau > tauRef + tauCut) ? 0.0 : std::exp(-tau);
    }
    // The batch's least view optical depth over every pathStride-th path (pathStride = P / 4096
    // for batches of more than 4096 paths, else 1), its vertices at frame 0, and every view (at
    // most 64) or 64 views spread evenly (view lane * n / 64).
    double tauReference(const std::vector<std::vector<Vertex>>& batch) const {
        const size_t P = batch.size();
        const size_t stride = P > 4096 ? P / 4096 : 1;
        const uint32_t n = vw.n();
        const uint32_t lanes = n < 64 ? n : 64;
        double best = 3.0e38;
        for (size_t i = 0; i < P; i += stride)
            for (const Vertex& x : batch[i])
                for (uint32_t lane = 0; lane < lanes; ++lane) {
                    const uint32_t v = n <= 64 ? lane : (uint32_t)(((uint64_t)lane * n) / 64);
                    V3 w;
                    if (vw.type == SSTMC_SOURCE_FARFIELD) w = vw.pt[v];
                    else if (vw.type == SSTMC_SOURCE_GAUSSIAN_BEAM) w = vw.dir[v];
                    else w = geo.normalize(sub(vw.pt[v], x.xt[0]));
                    best = std::min(best, med.opticalDepth(x.xt[0], w, nullptr));
                }
        return best;
    }
    // TemporalCorrelationSampler: il holds the illumination (u1) sources then the illumination2
    // (u2) ones, vw the view then the view2 ones; ticks are the time lags
    bool acorr = false;
    std::vector<double> ticks;
    uint32_t maxAttempts = 1u << 20;

    uint32_t L() const { return acorr ? il.n() / 2 : il.n(); }
    uint32_t V() const { return acorr ? vw.n() / 2 : vw.n(); }
    uint32_t T() const { return fr.nt; }

    // projectBoxToLine (MediumInterface.cuh:156-199) with pointToLineProjectionDistance (:137-153).
    void projectBox(const V3& p0, const V3& v, double* tmin, double* tmax) const {
        int nc = dims == 3 ? 8 : 4;
        double lo = 0, hi = 0;
        for (int c = 0; c < nc; ++c) {
            V3 p((c & 1) ? med.bmax[0] : med.bmin[0], (c & 2) ? med.bmax[1] : med.bmin[1],
                 (c & 4) ? med.bmax[2] : med.bmin[2]);
            V3 q = sub(p, p0);
            V3 o;
            for (int a = 0; a < dims; ++a) {
                double acc = 0;
                for (int b = 0; b < dims; ++b) acc += v[a] * v[b] * q[b];
                o[a] = acc;
            }
            double t = geo.dot(o, v);
            if (c == 0 || t < lo) lo = t;
            if (c == 0 || t > hi) hi = t;
        }
        *tmin = lo; *tmax = hi;
    }

    // ---------------------------------------------------------------- GaussianBeam (3D)
    V3 meanView;   // mean view beam direction (GaussianBeam.cuh:1176-1194), the proposal's axis
    uint32_t matAt(const V3& p) const { return med.inside(p) ? med.material(p) : 0u; }
    double waveK() const { return 2.0 * PI / lambda; }
    // orthogonalBase (VectorType.cuh:293-311)
    static void orthoBase(const V3& v, V3* b0, V3* b1) {
        if (std::fabs(1.0 - std::fabs(v[2])) <= EPS) { *b0 = V3(1, 0, 0); *b1 = V3(0, 1, 0); return; }
        V3 a(-v[1], v[0], 0.0);
        double n = std::sqrt(a[0] * a[0] + a[1] * a[1]);
        a = V3(a[0] / n, a[1] / n, 0.0);
        *b0 = a;
        *b1 = V3(a[1] * v[2] - a[2] * v[1], a[2] * v[0] - a[0] * v[2], a[0] * v[1] - a[1] * v[0]);
    }
    // pointToLineProjectionDistance (MediumInterface.cuh:138-153)
    double lineDistance(const V3& p, const V3& p0, const V3& v) const {
        V3 o;
        for (int a = 0; a < 3; ++a) {
            double acc = 0;
            for (int b = 0; b < 3; ++b) acc += v[a] * v[b] * (p[b] - p0[b]);
            o[a] = acc;
        }
        return o[0] * v[0] + o[1] * v[1] + o[2] * v[2];
    }
    // updateSampler (GaussianBeam.cuh:1281-1613): per beam, the focal material's mixture weights
    // exp(C + log-normalisation(|mu|)) (alphaPdfIcdfKernel, :254-295, with its large-kappa term
    // log(2 pi) * kappa), zB points along the beam across the box (initSamplingTable,
    // fillTimesInSamplingTable, :177-224) and per mixture the z density attenuation / w(z)^2
    // (samplingPdfKernel, :297-324), exclusive-scanned and normalised (:336-355).
    int buildGaussian(Source& s) {
        const uint32_t nb = s.n(), B = s.zB;
        const double k = waveK();
        s.focalMat.resize(nb); s.zmin.resize(nb); s.zmax.resize(nb); s.zdt.resize(nb);
        s.xyb0.resize(nb); s.xyb1.resize(nb);
        s.zpdf.assign(nb, {}); s.zicdf.assign(nb, {}); s.apdf.assign(nb, {}); s.aicdf.assign(nb, {});
        for (uint32_t l = 0; l < nb; ++l) {
            const V3 f = s.pt[l], d = s.dir[l];
            const uint32_t fm = matAt(f);
            if (fm == 0 || sc.type[fm] != SSTMC_SCATTERING_VMF) return SSTMC_INVALID_CONFIG;
            s.focalMat[l] = fm;
            projectBox(f, d, &s.zmin[l], &s.zmax[l]);
            s.zdt[l] = (s.zmax[l] - s.zmin[l]) / (double)B;
            orthoBase(d, &s.xyb0[l], &s.xyb1[l]);
            const Scatter::Vmf& v = sc.vmf[fm];
            const uint32_t M = (uint32_t)v.kappa.size();
            std::vector<double>& ap = s.apdf[l];
            double asum = 0;
            for (uint32_t i = 0; i < M; ++i) {
                const double kap = std::fabs(v.kappa[i]);
                const double ln = kap < 1e-5 ? -std::log(4.0 * PI)
                                             : std::log(kap) - (kap > 10.0 ? std::log(2.0 * PI) * kap
                                                                           : std::log(4.0 * PI * std::sinh(kap)));
                ap.push_back(std::exp(v.c[i] + ln));
                asum += ap.back();
            }
            double acc = 0;
            for (uint32_t i = 0; i < M; ++i) {
                ap[i] /= asum;
                s.aicdf[l].push_back(acc);
                acc += ap[i];
            }
            std::vector<double>& zp = s.zpdf[l];
            std::vector<double>& zi = s.zicdf[l];
            zp.assign((size_t)M * B, 0.0);
            zi.assign((size_t)M * B, 0.0);
            for (uint32_t b = 0; b < B; ++b) {
                const double t = s.zmin[l] + (double)b * s.zdt[l];
                const V3 z = add(f, scale(t, d));
                const double att = std::exp(-med.opticalDepth(z, scale(-1.0, d), nullptr));
                const double dz = geo.dot(sub(z, f), d);
                for (uint32_t i = 0; i < M; ++i) {
                    const double wz2 = (s.gammaA + std::fabs(v.kappa[i])) / (k * k) + dz * dz / s.gammaA;
                    zp[(size_t)i * B + b] = att / wz2;
                }
            }
            for (uint32_t i = 0; i < M; ++i) {
                double run = 0;
                for (uint32_t b = 0; b < B; ++b) { zi[(size_t)i * B + b] = run; run += zp[(size_t)i * B + b]; }
                const double sum = zp[(size_t)i * B + B - 1] + zi[(size_t)i * B + B - 1];
                for (uint32_t b = 0; b < B; ++b) {
                    zp[(size_t)i * B + b] *= (double)(B - 1) / sum;
                    zi[(size_t)i * B + b] /= sum;
                }
            }
        }
        return SSTMC_NO_ERROR;
    }
    // sampleFirstBeam (GaussianBeam.cuh:357-418): a beam, a mixture by its weight, a z bin by the
    // mixture's table moved uniformly within the bin, then a Gaussian transverse offset of width
    // w(z); the material is the one at the point (getMaterialInfo, :1642; 0 outside: a miss).
    V3 sampleGaussFirst(Stream& s, uint32_t* beam) const {
        const Source& g = il;
        const uint32_t l = s.index(g.n());
        *beam = l;
        const uint32_t B = g.zB;
        const uint32_t fm = g.focalMat[l];
        const uint32_t mix = (uint32_t)Scatter::cdfSearch(g.aicdf[l], s.uniform());
        const std::vector<double> zi(g.zicdf[l].begin() + (size_t)mix * B, g.zicdf[l].begin() + (size_t)(mix + 1) * B);
        const uint32_t zb = (uint32_t)Scatter::cdfSearch(zi, s.uniform());
        const V3 d = g.dir[l];
        const V3 z = add(add(g.pt[l], scale(g.zmin[l] + (double)zb * g.zdt[l], d)), scale(s.uniform() * g.zdt[l], d));
        const double k = waveK();
        const double gs = std::fabs(sc.vmf[fm].kappa[mix]);
        const double dz = geo.dot(sub(z, g.pt[l]), d);
        const double wz = std::sqrt((g.gammaA + gs) / (k * k) + dz * dz / g.gammaA);
        const double u1 = s.uniform(), u2 = s.uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double n1 = r * std::cos(2.0 * PI * u2), n2 = r * std::sin(2.0 * PI * u2);
        return add(z, scale(wz, add(scale(n1, g.xyb0[l]), scale(n2, g.xyb1[l]))));
    }
    // firstPointProbabilityKernel (GaussianBeam.cuh:420-487): the mixture-weighted z-table value
    // at the point's bin times the transverse Gaussian, over sigs(m) times the box depth in z.
    double prGauss(uint32_t l, const V3& x1) const {
        const Source& g = il;
        const uint32_t B = g.zB, fm = g.focalMat[l];
        const V3 f = g.pt[l], d = g.dir[l];
        const double dz = lineDistance(x1, f, d);
        const V3 zv = sub(add(f, scale(dz, d)), x1);
        const double dxy2 = geo.dot(zv, zv);
        uint32_t zb;
        if (dz < g.zmin[l]) zb = 0;
        else if (dz > g.zmax[l]) zb = B - 1;
        else zb = std::min(B - 1, (uint32_t)((dz - g.zmin[l]) / g.zdt[l]));   // the reference reads bin B at dz = tMax
        const double k = waveK();
        const Scatter::Vmf& v = sc.vmf[fm];
        double pdf = 0;
        for (size_t i = 0; i < v.kappa.size(); ++i) {
            const double wz2 = (g.gammaA + std::fabs(v.kappa[i])) / (k * k) + dz * dz / g.gammaA;
            pdf += g.apdf[l][i] * g.zpdf[l][i * B + zb] * std::exp(-dxy2 / (2.0 * wz2)) / (2.0 * wz2 * PI);
        }
        const uint32_t m = matAt(x1);
        return pdf / (med.sigs[m] * (med.bmax[2] - med.bmin[2]));
    }
    // computeCdfBuffer (GaussianBeam.cuh:489-621): the second-point proposal of beam l at x0, a
    // mixture over pairs (i, j) of the focal material's lobes of vMF lobes with axis mu_ij and
    // weight alpha_ij (normalised), from the beam's complex throughput gamma_a d + i k (x0 - focal)
    // convolved about the mean view direction; log_C = 1 / (4 pi) when |mu_ij| < 1e-4 (as written).
    struct Conv {
        std::vector<V3> mu;
        std::vector<double> w, logC;
    };
    void gaussConv(uint32_t l, const V3& x0, Conv* out) const {
        const Source& g = il;
        const Scatter::Vmf& v = sc.vmf[g.focalMat[l]];
        const uint32_t M = (uint32_t)v.kappa.size();
        const double k = waveK();
        const CV3 T = cvec(scale(g.gammaA, g.dir[l]), scale(k, sub(x0, g.pt[l])));
        std::vector<V3> cm(M);
        std::vector<double> rc(M);
        for (uint32_t i = 0; i < M; ++i) {
            const double gs = v.kappa[i];
            if (std::fabs(gs) < 0.000000001) {
                const cplx ct = csqrtRef(sumSquare(T));
                rc[i] = ct.real() - std::log(std::abs(ct)) + v.c[i];
                cm[i] = V3(0, 0, 0);
            } else {
                const cplx gob = gs * rcsqrtRef(sumSquare(cfma(gs, meanView, T)));
                cm[i] = realMult(gob, T);
                const double rk = 1.0 / geo.norm(cm[i]);
                const V3 w = scale(rk, cm[i]);
                const cplx ct = csqrtRef(sumSquare(cfma(gs, w, T)));
                rc[i] = ct.real() - std::log(std::abs(ct)) + v.c[i] - geo.dot(w, cm[i]);
            }
        }
        out->mu.resize((size_t)M * M);
        out->w.resize((size_t)M * M);
        out->logC.resize((size_t)M * M);
        double sum = 0;
        for (uint32_t q = 0; q < M * M; ++q) {
            const uint32_t i = q % M, j = q / M;
            const V3 mu = add(cm[i], cm[j]);
            const double kap = geo.norm(mu);
            const double lc = kap < 0.0001 ? 1.0 / (4.0 * PI) : Scatter::vmfNormLog(kap);
            out->mu[q] = mu;
            out->logC[q] = lc;
            out->w[q] = std::exp(rc[i] + rc[j] - lc - 2.0 * g.gammaA);
            sum += out->w[q];
        }
        for (double& w : out->w) w /= sum;
    }
    // sampleCdfBuffer (GaussianBeam.cuh:636-691): a pair by the first inclusive partial sum above
    // u, then a vMF direction about its axis (the z axis when |mu| < 1e-4).
    V3 sampleGaussSecond(uint32_t l, const V3& x0, Stream& s) const {
        if (sc.allIsotropic) return sc.randomDirection(s);   // isotropicGaussianSampling (:1001-1011)
        Conv cv;
        gaussConv(l, x0, &cv);
        const double u = s.uniform();
        size_t q = 0;
        double acc = 0;
        for (q = 0; q < cv.w.size(); ++q) {
            if (u < acc) break;
            acc += cv.w[q];
        }
        q -= 1;
        V3 mu = cv.mu[q];
        const double kap = geo.norm(mu);
        mu = kap < 0.0001 ? V3(0, 0, 1) : scale(1.0 / kap, mu);
        return sc.rotate(mu, Scatter::vmfCosine(kap, s), s);
    }
    // sampleDirectionProbability (GaussianBeam.cuh:693-752): the proposal density of the direction
    // x0 -> x1 under beam l's pair mixture; isotropicGaussianProbability (:1013-1023) otherwise.
    double pdirGauss(uint32_t l, const V3& x0, const V3& x1) const {
        if (sc.allIsotropic) return 1.0 / (4.0 * PI);
        Conv cv;
        gaussConv(l, x0, &cv);
        const V3 d = geo.normalize(sub(x1, x0));
        double p = 0;
        for (size_t q = 0; q < cv.w.size(); ++q) p += cv.w[q] * std::exp(geo.dot(cv.mu[q], d) + cv.logC[q]);
        return p;
    }
    // gGaussianBeam with the attenuation along the beam (GaussianBeam.cuh:795-816, 1942-1992): the
    // square root of the transmittance toward the beam's source side (illumination: -d; view: +d).
    cplx gGauss(const Source& s, uint32_t i, const V3& x, uint64_t* steps, bool cut = false) const {
        if (matAt(x) == 0) return cplx(0, 0);
        const V3 w = s.illum ? scale(-1.0, s.dir[i]) : s.dir[i];
        return cplx(std::sqrt(cutT(med.opticalDepth(x, w, steps), cut)), 0.0);
    }
    // fGaussianBeam (GaussianBeam.cuh:818-892): the vMF lobes of x0's material convolved with the
    // beam's complex throughput gamma_a d + i k (x0 - focal) (k negative for a view), w0 the
    // direction x0 -> p2 (illumination) or p2 -> x0 (view):
    //   sum_j exp(apNorm + log 2 pi + C_j) 2 sinh(s_j) / s_j,  s_j = sqrt((gamma_j w0 + T).(gamma_j w0 + T))
    cplx fGauss(const Source& s, uint32_t i, const V3& x0, const V3& p2) const {
        const uint32_t m = matAt(x0);
        if (m == 0) return cplx(0, 0);
        const V3 w0 = s.illum ? geo.normalize(sub(p2, x0)) : geo.normalize(sub(x0, p2));
        const double k = s.illum ? waveK() : -waveK();
        const CV3 T = cvec(scale(s.gammaA, s.dir[i]), scale(k, sub(x0, s.pt[i])));
        const double tc = s.apNorm + std::log(2.0 * PI);
        const Scatter::Vmf& v = sc.vmf[m];
        cplx f(0, 0);
        for (size_t j = 0; j < v.kappa.size(); ++j) {
            const double lognu = tc + v.c[j];
            const cplx sm = csqrtRef(sumSquare(cfma(v.kappa[j], w0, T)));
            if (dims == 2) f += std::exp(logI0(sm) + lognu);   // :876-878
            else if (std::fabs(sm.real()) < 0.0001 && std::fabs(sm.imag()) < 0.0001) f += 2.0 * std::exp(lognu);
            else f += twoSinh(sm, cplx(lognu, 0)) / sm;
        }
        return f;
    }
    // fsGaussianBeam (GaussianBeam.cuh:894-999): illumination beam a and view beam b through x0's
    // lobes, the inner convolution taken about b's direction.
    cplx fsGauss(uint32_t la, uint32_t vb, const V3& x0) const {
        const Source& A = il;
        const Source& Bs = vw;
        const uint32_t m = matAt(x0);
        const double k = waveK();
        const CV3 aT = cvec(scale(A.gammaA, A.dir[la]), scale(k, sub(x0, A.pt[la])));
        const CV3 bT = cvec(scale(Bs.gammaA, Bs.dir[vb]), scale(k, sub(Bs.pt[vb], x0)));
        const double tc = A.apNorm + Bs.apNorm + 2.0 * std::log(2.0 * PI);
        const Scatter::Vmf& v = sc.vmf[m];
        cplx fs(0, 0);
        for (size_t j = 0; j < v.kappa.size(); ++j) {
            double gs = v.kappa[j];
            const double lognu = tc + v.c[j];
            cplx C, sm, c(0, 0);
            if (std::fabs(gs) < 0.000001) {
                C = csqrtRef(sumSquare(aT));
                sm = csqrtRef(sumSquare(bT));
            } else {
                const cplx gob = gs * rcsqrtRef(sumSquare(cfma(gs, Bs.dir[vb], aT)));
                sm = csqrtRef(sumSquare(cfma(gob, aT, bT)));
                const V3 w = realMult(gob, aT);
                const double ra = 1.0 / geo.norm(w);
                gs *= ra;
                C = csqrtRef(sumSquare(cfma(gs, w, aT)));
                c = (-ra) * gob * dotR(aT, w);
            }
            if (dims == 2) {   // :989-991
                fs += std::exp(logI0(C) + logI0(sm) + c + lognu);
                continue;
            }
            const bool smZero = std::fabs(sm.real()) <= 0.0001 && std::fabs(sm.imag()) <= 0.0001;
            const bool cZero = std::fabs(C.real()) <= 0.0001 && std::fabs(C.imag()) <= 0.0001;
            if (!smZero && !cZero) fs += twoSinh(C + sm, c + lognu) / (C * sm);
            else if (smZero && cZero) fs += 4.0 * std::exp(c + lognu);
            else if (smZero) fs += 2.0 * twoSinh(C, c + lognu) / C;
            else fs += 2.0 * twoSinh(sm, cplx(lognu, 0)) / sm;
        }
        return fs;
    }
    bool gauss() const { return il.type == SSTMC_SOURCE_GAUSSIAN_BEAM; }

    // FarFieldNS::fillBuffers (FarField.cuh:71-127).
    void buildFarField(Source& s) {
        V3 r, ctr;
        double rmax = -1e300;
        for (int a = 0; a < dims; ++a) {
            r[a] = 0.5 * (med.bmax[a] - med.bmin[a]);
            ctr[a] = 0.5 * (med.bmax[a] + med.bmin[a]);
            rmax = std::max(rmax, r[a]);
        }
        double rad = rmax + EPS;
        uint32_t n = s.n();
        s.invDir.resize(n); s.tangent.resize(n); s.perpT.resize(n); s.perpS.resize(n);
        s.minT.resize(n); s.maxT.resize(n); s.minS.resize(n); s.maxS.resize(n); s.prob.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            V3 d = s.pt[i];
            V3 opp = scale(-1.0, d);
            for (int a = 0; a < dims; ++a) s.invDir[i][a] = std::fabs(d[a]) < EPS ? 1e12 : 1.0 / d[a];
            s.tangent[i] = add(ctr, scale(rad, opp));
            if (dims == 3) {
                V3 b0, b1;
                if (std::fabs(1.0 - std::fabs(opp[2])) <= EPS) {
                    b0 = V3(1, 0, 0); b1 = V3(0, 1, 0);
                } else {
                    b0 = geo.normalize(V3(-opp[1], opp[0], 0));
                    b1 = V3(b0[1] * opp[2] - b0[2] * opp[1], b0[2] * opp[0] - b0[0] * opp[2],
                            b0[0] * opp[1] - b0[1] * opp[0]);
                }
                s.perpT[i] = b0; s.perpS[i] = b1;
                projectBox(s.tangent[i], b0, &s.minT[i], &s.maxT[i]);
                projectBox(s.tangent[i], b1, &s.minS[i], &s.maxS[i]);
                s.prob[i] = 1.0 / ((s.maxT[i] - s.minT[i]) * (s.maxS[i] - s.minS[i]));
            } else {
                s.perpT[i] = V3(-opp[1], opp[0], 0);
                projectBox(s.tangent[i], s.perpT[i], &s.minT[i], &s.maxT[i]);
                s.prob[i] = 1.0 / (s.maxT[i] - s.minT[i]);
            }
        }
    }

    // sampleFarFieldKernel (FarField.cuh:129-205) / PointSourceNS::sampleFirstPointKernel
    // (PointSource.cuh:72-90): source point x0 and direction of the first flight.
    // With the TemporalCorrelationSampler the source is drawn from the u1 and u2 sources together:
    // u1 or u2 with probability 1/2 each, as the reference's random split of a batch between
    // u1IlluminationsHandle and u2IlluminationsHandle (TemporalSamplerCorrelation.cuh:611-676)
    // gives every path slot on average.
    void sampleSource(Stream& s, V3* x0, V3* dir) const {
        uint32_t l = s.index(il.n());
        // a 2D GaussianBeam starts from its focal point as a point source does (sampleFirstBeam,
        // GaussianBeam.cuh:754-773, then the medium's sample)
        if (il.type == SSTMC_SOURCE_POINT || il.type == SSTMC_SOURCE_GAUSSIAN_BEAM) {
            *x0 = il.pt[l];
            *dir = sc.randomDirection(s);
            return;
        }
        V3 d = il.pt[l];
        V3 p = add(il.tangent[l], scale(il.minT[l] + s.uniform() * (il.maxT[l] - il.minT[l]), il.perpT[l]));
        if (dims == 3)
            p = add(p, scale(il.minS[l] + s.uniform() * (il.maxS[l] - il.minS[l]), il.perpS[l]));
        V3 r, sh;
        for (int a = 0; a < 3; ++a) {
            r[a] = 0.5 * (med.bmax[a] - med.bmin[a]);
            sh[a] = p[a] - 0.5 * (med.bmax[a] + med.bmin[a]);
        }
        double bd = 0;
        for (int a = 0; a < dims; ++a) {
            int sgn = d[a] < 0 ? 1 : -1;
            double dist = il.invDir[l][a] * (r[a] * sgn - sh[a]);
            bool ok = dist >= 0.0;
            for (int b = 0; b < dims && ok; ++b)
                if (b != a) ok = std::fabs(std::fma(d[b], dist, sh[b])) < r[b];
            if (ok) { bd = dist; break; }
        }
        *x0 = add(p, scale(bd + EPS, d));
        *dir = d;
    }

    // TemporalSampler dtshift_host / dtShiftKernel / copyPointsFromBuffers*Temporal
    // (TemporalSampler.cuh:376-436, 540-647, 778-937): displaced position per output frame.
    void frames(uint64_t path, uint32_t event, Vertex& v) const {
        uint32_t nt = fr.nt;
        std::vector<V3> slot(nt);
        Stream s(seed, path, event, 1);
        for (uint32_t k = 0; k < nt; ++k) {
            double dt = fr.dt[k];
            double Dm = fr.D[v.m];
            V3 Um = fr.U[v.m];
            s.seek_block(k);
            double u1 = s.uniform(), u2 = s.uniform();
            double nrm = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
            V3 w = sc.randomDirection(s);
            double dr = nrm * std::sqrt(2.0 * dims * Dm * std::fabs(dt));
            for (int a = 0; a < dims; ++a) slot[k][a] = std::fma(dr, w[a], dt * Um[a]);
        }
        // cumulative sums along the negative chain [0, neg) and positive chain [neg, nt)
        for (uint32_t k = 0; k < nt; ++k) {
            bool first = (k == 0) || (k == fr.neg);
            V3 base = first ? v.x : slot[k - 1];
            slot[k] = add(base, slot[k]);
        }
        v.xt.assign(nt, V3());
        for (uint32_t t = 0; t < nt; ++t) v.xt[t] = slot[t < fr.neg ? fr.neg - 1 - t : t];
    }

    // One path slot: NextEventEstimation::sapmleFirstPathPoints retry loop
    // (NextEventEstimation.cuh:643-679) then sapmleNextPathPoints until the path dies (:681-759).
    int transport(uint64_t path, std::vector<Vertex>& vs, uint32_t* misses) const {
        vs.clear();
        *misses = 0;
        Vertex v1;
        uint32_t event = 0;
        uint32_t beam = 0;
        for (;; ++event) {
            if (event >= maxAttempts) return SSTMC_NO_HIT;
            Stream s(seed, path, event, 0);
            if (gauss() && dims == 3) {   // the first point drawn from the beam profile (sampleFirstBeam)
                const V3 x = sampleGaussFirst(s, &beam);
                const uint32_t m = matAt(x);
                if (m > 0) { v1.x = x; v1.dirIn = il.dir[beam]; v1.m = m; break; }
                ++*misses;
                continue;
            }
            V3 x0, d;
            sampleSource(s, &x0, &d);
            V3 x;
            int m = med.woodcock(x0, d, s, &x);
            if (m > 0) { v1.x = x; v1.dirIn = d; v1.m = (uint32_t)m; break; }
            ++*misses;
        }
        v1.rho = Stream(seed, path, event, 2).uniform();
        if (acorr) v1.xt.assign(fr.nt, v1.x);   // the correlation sampler displaces analytically
        else frames(path, event, v1);
        v1.len.assign(fr.nt, 0.0);
        vs.push_back(v1);
        for (uint32_t k = 2;; ++k) {
            const Vertex& prev = vs.back();
            uint32_t ev = 0x80000000u | k;
            Stream s(seed, path, ev, 0);
            // a GaussianBeam illumination proposes the second point's direction itself
            // (sampleSecondPoint, GaussianBeam.cuh:1673-1748)
            V3 d = (gauss() && dims == 3 && k == 2) ? sampleGaussSecond(beam, prev.x, s) : sc.sample(prev.m, prev.dirIn, s);
            V3 x;
            int m = med.woodcock(prev.x, d, s, &x);
            if (m <= 0) break;
            Vertex v;
            v.x = x; v.dirIn = d; v.m = (uint32_t)m;
            v.rho = Stream(seed, path, ev, 2).uniform();
            if (acorr) v.xt.assign(fr.nt, v.x);
            else frames(path, ev, v);
            v.len.resize(fr.nt);
            for (uint32_t t = 0; t < fr.nt; ++t)
                v.len[t] = prev.len[t] + geo.norm(sub(v.xt[t], prev.xt[t]));
            vs.push_back(v);
        }
        return SSTMC_NO_ERROR;
    }

    // ---------------------------------------------------------------- NEE terms
    // g: FarField gDirectionKernel (FarField.cuh:223-263) / gPointKernel (PointSource.cuh:117-172)
    // cut: an unweighted term, dropped past precision.tauCut
    cplx g(const Source& s, uint32_t i, const V3& x, uint64_t* steps, bool cut = false) const {
        if (s.type == SSTMC_SOURCE_GAUSSIAN_BEAM) return gGauss(s, i, x, steps, cut);
        if (s.type == SSTMC_SOURCE_FARFIELD) {
            const V3& d = s.pt[i];
            V3 wd = s.illum ? scale(-1.0, d) : d;
            double T = cutT(med.opticalDepth(x, wd, steps), cut);
            double a = std::sqrt(T);
            double ph = (s.illum ? 1.0 : -1.0) * 2.0 * geo.dot(d, x) / lambda;
            return cplx(a * std::cos(PI * ph), a * std::sin(PI * ph));
        }
        V3 q = sub(s.pt[i], x);
        double r = geo.norm(q);
        double rr = 1.0 / r;
        double T = cutT(med.opticalDepth(x, geo.normalize(q), steps), cut);
        double at = T * rr;
        if (dims == 3) at *= rr * sc.isoAmp();
        if (s.area) at *= rr * std::fabs(dims == 3 ? x[2] - s.pt[i][2] : x[1] - s.pt[i][1]);
        double a = std::sqrt(at);
        double ph = 2.0 * r / lambda;
        return cplx(a * std::cos(PI * ph), a * std::sin(PI * ph));
    }
    // firstPointProbability (FarField.cuh:469-499, PointSource.cuh:263-294, quirk Q7 in 2D).
    double pr(uint32_t l, const V3& x1, uint64_t* steps) const {
        // a 2D GaussianBeam: the point-source density about its focal point (attenuationPoint and
        // probabilityPointNormalizeKernel, GaussianBeam.cuh:775-792, 1810-1832)
        if (il.type == SSTMC_SOURCE_GAUSSIAN_BEAM && dims == 3) return prGauss(l, x1);
        if (il.type == SSTMC_SOURCE_FARFIELD)
            return std::exp(-med.opticalDepth(x1, scale(-1.0, il.pt[l]), steps)) * il.prob[l];
        V3 q = sub(il.pt[l], x1);
        double rr = 1.0 / geo.norm(q);
        double T = std::exp(-med.opticalDepth(x1, geo.normalize(q), steps));
        return dims == 3 ? T * (rr * rr) / (4.0 * PI) : T * (rr * (8.0 * PI));
    }
    // incoming direction at x1 from illumination l (FarField d_l; PointSource n(x1 - s_l)).
    V3 inDir(uint32_t l, const V3& x1) const {
        if (il.type == SSTMC_SOURCE_FARFIELD) return il.pt[l];
        return geo.normalize(sub(x1, il.pt[l]));
    }
    // outgoing direction from x toward view v (FarField d_v; PointSource n(s_v - x)).
    V3 outDir(uint32_t v, const V3& x) const {
        if (vw.type == SSTMC_SOURCE_FARFIELD) return vw.pt[v];
        return geo.normalize(sub(vw.pt[v], x));
    }
    // threePointFunctionSingle (FarField.cuh:546-599, PointSource.cuh:363-446)
    cplx fsingle(uint32_t l, uint32_t v, uint32_t m, const V3& x1) const {
        if (il.type == SSTMC_SOURCE_GAUSSIAN_BEAM) return fsGauss(l, v, x1);
        if (il.type == SSTMC_SOURCE_POINT && sc.allIsotropic) return sc.isoAmp();
        return sc.amp(m, geo.dot(inDir(l, x1), outDir(v, x1)));
    }
    // illumination f at pL=2 (FarField.cuh:546-556; PointSource.cuh:347-361)
    cplx fil(uint32_t l, uint32_t m, const V3& x1, const V3& x2) const {
        if (il.type == SSTMC_SOURCE_GAUSSIAN_BEAM) return fGauss(il, l, x1, x2);
        if (il.type == SSTMC_SOURCE_POINT && sc.allIsotropic) return sc.isoAmp();
        return sc.amp(m, geo.dot(inDir(l, x1), geo.normalize(sub(x2, x1))));
    }
    // view f at pL>=2: amplitude(-d_v | s_v, x_k, x_{k-1}) (NextEventEstimation.cuh:1560-1563)
    cplx fvw(uint32_t v, uint32_t m, const V3& xprev, const V3& xk) const {
        if (vw.type == SSTMC_SOURCE_GAUSSIAN_BEAM) return fGauss(vw, v, xk, xprev);
        if (vw.type == SSTMC_SOURCE_POINT && sc.allIsotropic) return sc.isoAmp();
        V3 a = vw.type == SSTMC_SOURCE_FARFIELD ? scale(-1.0, vw.pt[v])
                                                  : geo.normalize(sub(xk, vw.pt[v]));
        return sc.amp(m, geo.dot(a, geo.normalize(sub(xprev, xk))));
    }

    // Incoming unit direction at x from source i of s, as the three-point functions use it
    // (FarField: illumination d, view -d, FarField.cuh:546-556; PointSource: n(x - s),
    // PointSource.cuh:347-361).
    V3 incoming(const Source& s, uint32_t i, const V3& x) const {
        if (s.type == SSTMC_SOURCE_FARFIELD) return s.illum ? s.pt[i] : scale(-1.0, s.pt[i]);
        return geo.normalize(sub(x, s.pt[i]));
    }
    // Three-point function of source i of s at vertex p1 with the path continuing to p2
    // (threePointsThroughput, NextEventEstimation.cuh:773-800).
    cplx f3(const Source& s, uint32_t i, uint32_t m, const V3& p1, const V3& p2) const {
        if (s.type == SSTMC_SOURCE_POINT && sc.allIsotropic) return sc.isoAmp();
        return sc.amp(m, geo.dot(incoming(s, i, p1), geo.normalize(sub(p2, p1))));
    }

    struct Terms {
        std::vector<cplx> A;   // [t][l]
        std::vector<cplx> B;   // [t][v]   (field) or sum_k G_k[v,0] conj G_k[v,t] (correlation)
        std::vector<cplx> S;   // [t][v][l]
        // coherent backscattering, reversed path v -> x1 -> ... -> xK -> l
        // (NextEventEstimation.cuh:1212-1235, 1432-1443): Ar = view connected at x1 (Gv0 with f
        // toward x2 and the multiple-scattering weight), Br = illumination connected at every
        // later vertex (Gl with f toward the previous vertex and the path phase).
        std::vector<cplx> Ar;  // [t][v]
        std::vector<cplx> Br;  // [t][l]
        // TemporalCorrelationSampler with coherent backscattering: the two mixed pairings
        // CBS02 + CBS03, which do not factor over (l, v), summed over the path's vertices
        std::vector<cplx> X;   // [t][v][l]
        bool multiple = false;
    };

    // Per-path terms of Appendix A (NextEventEstimation.cuh:1016-1136,1138-1458;
    // TemporalSampler.cuh:668-719).
    void terms(const std::vector<Vertex>& vs, Terms& out, uint64_t* steps) const {
        if (acorr) { termsAcorr(vs, out, steps); return; }
        uint32_t nl = L(), nv = V(), nt = T();
        const Vertex& v1 = vs[0];
        out.A.assign((size_t)nt * nl, cplx(0));
        out.B.assign((size_t)nt * nv, cplx(0));
        out.S.assign((size_t)nt * nv * nl, cplx(0));
        out.multiple = vs.size() >= 2;
        std::vector<double> prl(nl);
        double p1 = 0;
        for (uint32_t l = 0; l < nl; ++l) { prl[l] = pr(l, v1.x, steps); p1 += prl[l]; }
        p1 /= nl;
        double w1 = 1.0 / std::sqrt(p1);
        std::vector<cplx> gl((size_t)nt * nl), gv1((size_t)nt * nv);
        for (uint32_t t = 0; t < nt; ++t) {
            for (uint32_t l = 0; l < nl; ++l) gl[t * nl + l] = g(il, l, v1.xt[t], steps);
            for (uint32_t v = 0; v < nv; ++v) gv1[t * nv + v] = g(vw, v, v1.xt[t], steps, !cbs);
        }
        cplx e1(std::cos(2.0 * PI * v1.rho), std::sin(2.0 * PI * v1.rho));
        for (uint32_t t = 0; t < nt; ++t) {
            const V3& xf = fr.temporalAmp ? v1.xt[t] : v1.x;
            for (uint32_t v = 0; v < nv; ++v)
                for (uint32_t l = 0; l < nl; ++l)
                    out.S[((size_t)t * nv + v) * nl + l] =
                        gl[t * nl + l] * w1 * e1 * gv1[t * nv + v] * fsingle(l, v, v1.m, xf);
        }
        if (!out.multiple) return;
        const Vertex& v2 = vs[1];
        double p2 = 0;
        for (uint32_t l = 0; l < nl; ++l) {
            // 2D: the scattering density about the focal point's direction (multiplyPdf, :1933-1937)
            if (gauss() && dims == 3) {   // secondPointProbability (GaussianBeam.cuh:1835-1940); 0 becomes 1e20
                double q = prl[l] * pdirGauss(l, v1.x, v2.x);
                p2 += (q == 0 && !sc.allIsotropic) ? 1e20 : q;
            } else {
                p2 += prl[l] * sc.pdf(v1.m, geo.dot(inDir(l, v1.x), geo.normalize(sub(v2.x, v1.x))));
            }
        }
        p2 /= nl;
        double w2 = 1.0 / std::sqrt(p2);
        for (uint32_t t = 0; t < nt; ++t) {
            const V3& a1 = fr.temporalAmp ? v1.xt[t] : v1.x;
            const V3& a2 = fr.temporalAmp ? v2.xt[t] : v2.x;
            for (uint32_t l = 0; l < nl; ++l)
                out.A[t * nl + l] = gl[t * nl + l] * fil(l, v1.m, a1, a2) * w2;
        }
        if (cbs) {
            out.Ar.assign((size_t)nt * nv, cplx(0));
            out.Br.assign((size_t)nt * nl, cplx(0));
            for (uint32_t t = 0; t < nt; ++t) {
                const V3& a1 = fr.temporalAmp ? v1.xt[t] : v1.x;
                const V3& a2 = fr.temporalAmp ? v2.xt[t] : v2.x;
                for (uint32_t v = 0; v < nv; ++v)
                    out.Ar[t * nv + v] = gv1[t * nv + v] * f3(vw, v, v1.m, a1, a2) * w2;
            }
            for (size_t k = 1; k < vs.size(); ++k) {
                const Vertex& xk = vs[k];
                const Vertex& xp = vs[k - 1];
                for (uint32_t t = 0; t < nt; ++t) {
                    const V3& pk = fr.temporalAmp ? xk.xt[t] : xk.x;
                    const V3& pp = fr.temporalAmp ? xp.xt[t] : xp.x;
                    double ph = xk.rho + xk.len[t] / lambda;
                    cplx e(std::cos(2.0 * PI * ph), std::sin(2.0 * PI * ph));
                    for (uint32_t l = 0; l < nl; ++l)
                        out.Br[t * nl + l] += g(il, l, xk.xt[t], steps) * f3(il, l, xk.m, pk, pp) * e;
                }
            }
        }
        std::vector<cplx> G((size_t)nt * nv);
        for (size_t k = 1; k < vs.size(); ++k) {
            const Vertex& xk = vs[k];
            const Vertex& xp = vs[k - 1];
            for (uint32_t t = 0; t < nt; ++t) {
                const V3& pk = fr.temporalAmp ? xk.xt[t] : xk.x;
                const V3& pp = fr.temporalAmp ? xp.xt[t] : xp.x;
                double ph = xk.rho + xk.len[t] / lambda;
                cplx e(std::cos(2.0 * PI * ph), std::sin(2.0 * PI * ph));
                for (uint32_t v = 0; v < nv; ++v)
                    G[t * nv + v] = g(vw, v, xk.xt[t], steps, true) * fvw(v, xk.m, pp, pk) * e;
            }
            for (uint32_t t = 0; t < nt; ++t)
                for (uint32_t v = 0; v < nv; ++v) {
                    if (output == SSTMC_OUTPUT_FIELD) out.B[t * nv + v] += G[t * nv + v];
                    else out.B[t * nv + v] += G[v] * std::conj(G[t * nv + v]);
                }
        }
        if (output == SSTMC_OUTPUT_CORRELATION) {
            for (uint32_t t = nt; t-- > 0;)  // frame 0 last: it is the reference of every lag
                for (uint32_t l = 0; l < nl; ++l)
                    out.A[t * nl + l] = out.A[l] * std::conj(out.A[t * nl + l]);
        }
    }
    // ---------------------------------------------------------------- TemporalCorrelationSampler
    // Per-path terms of the analytic correlation u1(-tau/2) conj u2(+tau/2)
    // (TemporalSamplerCorrelation.cuh; NextEventEstimation.cuh:1016-1136 single scattering,
    // :1138-1200 and :1400-1425 multiple scattering, :1461-1601 order of evaluation). For lag tau_t:
    //  * g of u1 sees a vertex of material m at x - (tau/2) U_m, g of u2 at x + (tau/2) U_m
    //    (copyPointsToBuffersThroughput, TemporalSamplerCorrelation.cuh:228-250);
    //  * the three-point functions of u1 and u2 both see x - (tau/2) U_m: the sign test of
    //    copyPointsToBuffersThreePoints (:253-313) is always true (quirk Q3);
    //  * Brownian motion enters as exp(-k^2 |D_m tau| |q|^2), q the momentum transfer at a vertex
    //    from the unshifted positions: at x1 against the mean u1/u2 illumination direction (C1_l1l2),
    //    at x_k against the mean view direction (C1_v1v2), both through
    //    temporalCorrelationSourceContributionMultiple (SourceInterface.cuh:280-352); at the inner
    //    vertices through momentTransfer (copyBuffersToPoints_momentumTransfer, :144-179), applied
    //    once, on u1 (pathContriobutionKernel, :384-450); single scattering through
    //    temporalCorrelationSourceContributionSingle (SourceInterface.cuh:217-277);
    //  * the path phase 2 pi L_k / lambda multiplies u1 and u2 alike and cancels;
    //  * the path probabilities are the means of the u1 and u2 first-point probabilities, without the
    //    square root (pathSingleProbability, :783-829; NextEventEstimation.cuh:851-880, 930-977); the
    //    second-point factor uses the u1 illuminations' directions for both (:833-856,
    //    secondPointProbability of illuminationsHandle);
    //  * single scattering with a scalar Fs (point illumination and view, isotropic scattering,
    //    SamplerInterface.cuh:155, PointSource.cuh:62) is Fs^2 / p1 g_l1 g_v1 conj(g_l2 g_v2):
    //    u2's Fs and the single-scattering temporal factor are not used (NextEventEstimation.cuh:
    //    308-324, 1020-1029, 1094-1118).
    V3 shifted(const V3& x, uint32_t m, double s) const { return add(x, scale(s, fr.U[m])); }
    // Temporal factor of a connection of u1 source (s1, i1) and u2 source (s2, i2) at vertex p1
    // whose path continues to p2: exp(-dtD |(e1 + e2) / 2 - n(p2 - p1)|^2), e the incoming unit
    // direction of each source at p1 (a far-field view enters as minus its direction), all from
    // unshifted positions (temporalCorrelationSourceContributionMultiple,
    // SourceInterface.cuh:280-352; connection types C1-C4, TemporalSamplerCorrelation.cuh:975-1043).
    double temporalT(const Source& s1, uint32_t i1, const Source& s2, uint32_t i2, const V3& p1, const V3& p2,
                     double dtD) const {
        const V3 e1 = incoming(s1, i1, p1), e2 = incoming(s2, i2, p1);
        const V3 q = sub(scale(0.5, add(e1, e2)), geo.normalize(sub(p2, p1)));
        return std::exp(-dtD * geo.dot(q, q));
    }
    void termsAcorr(const std::vector<Vertex>& vs, Terms& out, uint64_t* steps) const {
        const uint32_t nl = L(), nv = V(), nt = T();
        const double kw = 2.0 * PI / lambda, k2 = kw * kw;
        const Vertex& v1 = vs[0];
        out.A.assign((size_t)nt * nl, cplx(0));
        out.B.assign((size_t)nt * nv, cplx(0));
        out.S.assign((size_t)nt * nv * nl, cplx(0));
        out.multiple = vs.size() >= 2;
        std::vector<double> prl(2 * nl);
        double p1 = 0;
        for (uint32_t i = 0; i < 2 * nl; ++i) { prl[i] = pr(i, v1.x, steps); p1 += prl[i]; }
        p1 /= 2.0 * nl;
        const bool scalarFs = il.type == SSTMC_SOURCE_POINT && vw.type == SSTMC_SOURCE_POINT && sc.allIsotropic;
        const double iso = sc.isoAmp();
        std::vector<cplx> gl1(nl), gl2(nl), gv1(nv), gv2(nv);
        for (uint32_t t = 0; t < nt; ++t) {
            const double s = 0.5 * ticks[t];
            const double dtD = k2 * std::fabs(fr.D[v1.m] * ticks[t]);
            const V3 xa = shifted(v1.x, v1.m, -s), xb = shifted(v1.x, v1.m, s);
            for (uint32_t l = 0; l < nl; ++l) { gl1[l] = g(il, l, xa, steps); gl2[l] = g(il, nl + l, xb, steps); }
            for (uint32_t v = 0; v < nv; ++v) { gv1[v] = g(vw, v, xa, steps, true); gv2[v] = g(vw, nv + v, xb, steps, true); }
            for (uint32_t v = 0; v < nv; ++v)
                for (uint32_t l = 0; l < nl; ++l) {
                    const cplx u1 = gl1[l] * gv1[v], u2 = gl2[l] * gv2[v];
                    cplx sv;
                    if (scalarFs) {
                        sv = iso * iso / p1 * u1 * std::conj(u2);
                    } else {
                        const cplx f1 = fsingle(l, v, v1.m, xa), f2 = fsingle(nl + l, nv + v, v1.m, xa);
                        const V3 q = scale(0.5, sub(add(inDir(l, v1.x), inDir(nl + l, v1.x)),
                                                    add(outDir(v, v1.x), outDir(nv + v, v1.x))));
                        sv = u1 * f1 / p1 * std::conj(u2 * f2 * std::exp(-dtD * geo.dot(q, q)));
                    }
                    out.S[((size_t)t * nv + v) * nl + l] = sv;
                }
        }
        if (!out.multiple) return;
        const Vertex& v2 = vs[1];
        const V3 d12 = geo.normalize(sub(v2.x, v1.x));
        double p2 = 0;
        for (uint32_t i = 0; i < 2 * nl; ++i) p2 += prl[i] * sc.pdf(v1.m, geo.dot(inDir(i % nl, v1.x), d12));
        p2 /= 2.0 * nl;
        for (uint32_t t = 0; t < nt; ++t) {
            const double s = 0.5 * ticks[t];
            const double dtD = k2 * std::fabs(fr.D[v1.m] * ticks[t]);
            const V3 x1a = shifted(v1.x, v1.m, -s), x1b = shifted(v1.x, v1.m, s), x2a = shifted(v2.x, v2.m, -s);
            for (uint32_t l = 0; l < nl; ++l) {
                const cplx a1 = g(il, l, x1a, steps) * f3(il, l, v1.m, x1a, x2a) / p2;
                const cplx a2 = g(il, nl + l, x1b, steps) * f3(il, nl + l, v1.m, x1a, x2a);
                const V3 q = sub(scale(0.5, add(inDir(l, v1.x), inDir(nl + l, v1.x))), d12);
                out.A[(size_t)t * nl + l] = a1 * std::conj(a2) * std::exp(-dtD * geo.dot(q, q));
            }
        }
        // coherent backscattering (NextEventEstimation.cuh:1237-1378): the first vertex's u1 and
        // u2 terms of either source, kept apart for the mixed pairings; the view side carries the
        // path weight like the illumination side (Gv0 *= 1/p, :991-1000)
        std::vector<cplx> fa1, fa2, fc1, fc2;
        if (cbs) {
            fa1.assign((size_t)nt * nl, cplx(0)); fa2.assign((size_t)nt * nl, cplx(0));
            fc1.assign((size_t)nt * nv, cplx(0)); fc2.assign((size_t)nt * nv, cplx(0));
            out.Ar.assign((size_t)nt * nv, cplx(0));
            out.Br.assign((size_t)nt * nl, cplx(0));
            out.X.assign((size_t)nt * nv * nl, cplx(0));
            for (uint32_t t = 0; t < nt; ++t) {
                const double s = 0.5 * ticks[t];
                const double dtD = k2 * std::fabs(fr.D[v1.m] * ticks[t]);
                const V3 x1a = shifted(v1.x, v1.m, -s), x1b = shifted(v1.x, v1.m, s), x2a = shifted(v2.x, v2.m, -s);
                for (uint32_t l = 0; l < nl; ++l) {
                    fa1[t * nl + l] = g(il, l, x1a, steps) * f3(il, l, v1.m, x1a, x2a) / p2;
                    fa2[t * nl + l] = g(il, nl + l, x1b, steps) * f3(il, nl + l, v1.m, x1a, x2a);
                }
                for (uint32_t v = 0; v < nv; ++v) {
                    fc1[t * nv + v] = g(vw, v, x1a, steps) * f3(vw, v, v1.m, x1a, x2a) / p2;
                    fc2[t * nv + v] = g(vw, nv + v, x1b, steps) * f3(vw, nv + v, v1.m, x1a, x2a);
                    // CBS04's first-vertex half: Gv0 conj(u2Gv0) C4_v1v2
                    out.Ar[t * nv + v] = fc1[t * nv + v] * std::conj(fc2[t * nv + v]) *
                                         temporalT(vw, v, vw, nv + v, v1.x, v2.x, dtD);
                }
            }
        }
        std::vector<cplx> S2, S3;   // per-path sums over vertices of the mixed pairings' last halves
        if (cbs) { S2.assign((size_t)nt * nv * nl, cplx(0)); S3.assign((size_t)nt * nv * nl, cplx(0)); }
        double mt = 0, len = 0;   // inner momentum transfer sum_j D_j |q_j|^2, path length from x1
        for (size_t k = 1; k < vs.size(); ++k) {
            const Vertex& xk = vs[k];
            const Vertex& xp = vs[k - 1];
            len += geo.norm(sub(xk.x, xp.x));
            if (k >= 2) {
                const V3 q = sub(geo.normalize(sub(xp.x, vs[k - 2].x)), geo.normalize(sub(xk.x, xp.x)));
                mt += geo.dot(q, q) * fr.D[xp.m];
            }
            const V3 back = geo.normalize(sub(xp.x, xk.x));
            const cplx e(std::cos(2.0 * PI * len / lambda), std::sin(2.0 * PI * len / lambda));
            for (uint32_t t = 0; t < nt; ++t) {
                const double s = 0.5 * ticks[t];
                const double dtD = k2 * std::fabs(fr.D[xk.m] * ticks[t]);
                const double inner = std::exp(-k2 * std::fabs(ticks[t]) * mt);
                const V3 xka = shifted(xk.x, xk.m, -s), xkb = shifted(xk.x, xk.m, s), xpa = shifted(xp.x, xp.m, -s);
                for (uint32_t v = 0; v < nv; ++v) {
                    const cplx G1 = g(vw, v, xka, steps, true) * fvw(v, xk.m, xpa, xka) * e * inner;
                    const cplx G2 = g(vw, nv + v, xkb, steps, true) * fvw(nv + v, xk.m, xpa, xka) * e;
                    const V3 q = sub(scale(-0.5, add(outDir(v, xk.x), outDir(nv + v, xk.x))), back);
                    out.B[(size_t)t * nv + v] += G1 * std::conj(G2) * std::exp(-dtD * geo.dot(q, q));
                }
                if (!cbs) continue;
                // the reversed path's last connection: the illuminations at x_k with their f toward
                // x_{k-1}, u1 carrying the inner momentum transfer (pathContriobutionKernel)
                std::vector<cplx> gl1(nl), gl2(nl), gv1(nv), gv2(nv);
                for (uint32_t l = 0; l < nl; ++l) {
                    gl1[l] = g(il, l, xka, steps) * f3(il, l, xk.m, xka, xpa) * e * inner;
                    gl2[l] = g(il, nl + l, xkb, steps) * f3(il, nl + l, xk.m, xka, xpa) * e;
                    // CBS04's last half: Gl conj(u2Gl) C4_l1l2
                    out.Br[(size_t)t * nl + l] +=
                        gl1[l] * std::conj(gl2[l]) * temporalT(il, l, il, nl + l, xk.x, xp.x, dtD);
                }
                for (uint32_t v = 0; v < nv; ++v) {
                    gv1[v] = g(vw, v, xka, steps, true) * fvw(v, xk.m, xpa, xka) * e * inner;
                    gv2[v] = g(vw, nv + v, xkb, steps, true) * fvw(nv + v, xk.m, xpa, xka) * e;
                }
                for (uint32_t v = 0; v < nv; ++v)
                    for (uint32_t l = 0; l < nl; ++l) {
                        const size_t i = ((size_t)t * nv + v) * nl + l;
                        // CBS02: (v1 -> ... -> l1) conj(l2 -> ... -> v2), last connection C2_l1v2
                        S2[i] += gl1[l] * std::conj(gv2[v]) * temporalT(il, l, vw, nv + v, xk.x, xp.x, dtD);
                        // CBS03: (l1 -> ... -> v1) conj(v2 -> ... -> l2), last connection C3_v1l2
                        S3[i] += std::conj(gl2[l]) * gv1[v] * temporalT(vw, v, il, nl + l, xk.x, xp.x, dtD);
                    }
            }
        }
        if (cbs) {
            for (uint32_t t = 0; t < nt; ++t) {
                const double dtD = k2 * std::fabs(fr.D[v1.m] * ticks[t]);
                for (uint32_t v = 0; v < nv; ++v)
                    for (uint32_t l = 0; l < nl; ++l) {
                        const size_t i = ((size_t)t * nv + v) * nl + l;
                        // first connections C2_v1l2 and C3_l1v2 at x1 toward x2
                        const cplx f2 = std::conj(fa2[t * nl + l]) * fc1[t * nv + v] *
                                        temporalT(vw, v, il, nl + l, v1.x, v2.x, dtD);
                        const cplx f3v = fa1[t * nl + l] * std::conj(fc2[t * nv + v]) *
                                         temporalT(il, l, vw, nv + v, v1.x, v2.x, dtD);
                        out.X[i] = f2 * S2[i] + f3v * S3[i];
                    }
            }
        }
    }

    // Adds the path's contribution to U[t][v][l].
    void accumulate(const Terms& tm, std::vector<cplx>& U) const {
        uint32_t nl = L(), nv = V(), nt = T();
        for (uint32_t t = 0; t < nt; ++t)
            for (uint32_t v = 0; v < nv; ++v)
                for (uint32_t l = 0; l < nl; ++l) {
                    size_t i = ((size_t)t * nv + v) * nl + l;
                    cplx s = tm.S[i];
                    if (output == SSTMC_OUTPUT_CORRELATION) s = tm.S[(size_t)v * nl + l] * std::conj(s);
                    U[i] += s;
                    if (tm.multiple && !cbs) U[i] += tm.A[t * nl + l] * tm.B[t * nv + v];
                    // correlation CBS: v = 1/2 * (CBS01 + CBS02 + CBS03 + CBS04) for multiple
                    // scattering (NextEventEstimation.cuh:1380-1391)
                    else if (tm.multiple && acorr)
                        U[i] += (tm.A[t * nl + l] * tm.B[t * nv + v] + tm.Br[t * nl + l] * tm.Ar[t * nv + v] +
                                 tm.X[i]) * 0.5;
                    // field CBS: v = 1/sqrt(2) * (direct + reversed) for multiple scattering
                    // (NextEventEstimation.cuh:1385-1391, 1448-1454)
                    else if (tm.multiple && cbs)
                        U[i] += (tm.A[t * nl + l] * tm.B[t * nv + v] + tm.Br[t * nl + l] * tm.Ar[t * nv + v]) *
                                0.70710678118654752440;
                }
    }
};

// ------------------------------------------------------------------ config parsing
// Mirrors buildScattering/buildMedium/buildSource/buildSampler (MatlabInterface.cu:712-1372).
int build(const sstmc_config* c, Problem& pb) {
    if (!c || (c->dims != 2 && c->dims != 3)) return SSTMC_INVALID_CONFIG;
    int dims = (int)c->dims;
    pb.dims = dims;
    pb.geo.dims = dims;
    pb.lambda = c->simulation.lambda > 0 ? c->simulation.lambda : 1.0;
    pb.P = c->simulation.batchSize ? c->simulation.batchSize : 1024;
    pb.fullIter = c->simulation.fullIteration ? c->simulation.fullIteration : 1;
    pb.seed = c->simulation.seed;
    pb.output = c->tracer.output;
    pb.cbs = c->tracer.isCBS != 0;
    pb.tauCut = c->precision.tauCut;
    if (pb.cbs && pb.output == SSTMC_OUTPUT_CORRELATION) return SSTMC_NOT_SUPPORTED;
    // scattering: one cell -> materials 0 and 1 (MatlabInterface.cu:750-760, 878-881)
    uint32_t nm = c->scattering.materialNum;
    if (nm == 0 || nm > 10 || !c->scattering.type) return SSTMC_INVALID_CONFIG;
    uint32_t nmat = nm == 1 ? 2 : nm;
    pb.sc.dims = dims;
    pb.sc.type.assign(nmat, 0);
    pb.sc.g.assign(nmat, 0.0);
    pb.sc.tab.assign(nmat, Scatter::Table());
    pb.sc.vmf.assign(nmat, Scatter::Vmf());
    std::vector<size_t> toff(nm, 0), voff(nm, 0);
    for (uint32_t m = 0, acc = 0, accv = 0; m < nm; ++m) {
        toff[m] = acc;
        voff[m] = accv;
        if (c->scattering.tabularSize) acc += 2 * c->scattering.tabularSize[m];
        if (c->scattering.vmfMixtures) accv += c->scattering.vmfMixtures[m];
    }
    for (uint32_t m = 0; m < nmat; ++m) {
        uint32_t src = nm == 1 ? 0 : m;
        int ty = c->scattering.type[src];
        if (ty != SSTMC_SCATTERING_ISOTROPIC && ty != SSTMC_SCATTERING_HG && ty != SSTMC_SCATTERING_TABULAR &&
            ty != SSTMC_SCATTERING_VMF)
            return SSTMC_NOT_SUPPORTED;
        pb.sc.type[m] = ty;
        pb.sc.g[m] = (ty == SSTMC_SCATTERING_HG && c->scattering.g) ? c->scattering.g[src] : 0.0;
        if (ty == SSTMC_SCATTERING_TABULAR) {
            if (!c->scattering.tabular || !c->scattering.tabularSize || c->scattering.tabularSize[src] < 2)
                return SSTMC_INVALID_CONFIG;
            pb.sc.setTabular(m, c->scattering.tabular + toff[src], c->scattering.tabularSize[src]);
        }
        if (ty == SSTMC_SCATTERING_VMF) {
            const sstmc_scattering& q = c->scattering;
            if (!q.vmfMixtures || !q.vmfMu || !q.vmfC || !q.vmfAlpha || q.vmfMixtures[src] == 0)
                return SSTMC_INVALID_CONFIG;
            pb.sc.setVmf(m, q.vmfMu + voff[src], q.vmfC + voff[src], q.vmfAlpha + voff[src], q.vmfMixtures[src]);
        }
    }
    // tabular never counts as isotropic (Scattering.cuh:478-482); vMF when every |kappa| <= 1e-4
    pb.sc.allIsotropic = true;
    for (uint32_t m = 0; m < nmat; ++m)
        if (!(pb.sc.type[m] == SSTMC_SCATTERING_ISOTROPIC ||
              (pb.sc.type[m] == SSTMC_SCATTERING_HG && std::fabs(pb.sc.g[m]) < 0.0001) ||
              (pb.sc.type[m] == SSTMC_SCATTERING_VMF && pb.sc.vmf[m].iso)))
            pb.sc.allIsotropic = false;
    // medium
    const sstmc_medium& md = c->medium;
    Medium& med = pb.med;
    med.dims = dims;
    uint32_t sz[3] = {md.xAxisSize, md.yAxisSize, dims == 3 ? md.zAxisSize : 2u};
    const double* axp[3] = {md.xAxis, md.yAxis, md.zAxis};
    for (int a = 0; a < 3; ++a) {
        if (a < dims) {
            if (sz[a] < 2 || !axp[a]) return SSTMC_INVALID_CONFIG;
            med.ax[a].assign(axp[a], axp[a] + sz[a]);
            for (uint32_t i = 1; i < sz[a]; ++i)
                if (!(med.ax[a][i] > med.ax[a][i - 1])) return SSTMC_INVALID_CONFIG;
        } else {
            med.ax[a] = {0.0, 0.0};
        }
        med.bmin[a] = med.ax[a].front();
        med.bmax[a] = med.ax[a].back();
    }
    size_t ncell = (size_t)(sz[0] - 1) * (sz[1] - 1) * (dims == 3 ? sz[2] - 1 : 1);
    if (!md.materialGrid || !md.sigs || md.materialNum != nmat) return SSTMC_INVALID_CONFIG;
    med.grid.assign(md.materialGrid, md.materialGrid + ncell);
    for (uint32_t v : med.grid)
        if (v == 0 || v >= nmat) return SSTMC_INVALID_CONFIG;
    med.sigs.assign(md.sigs, md.sigs + nmat);
    med.siga.assign(nmat, 0.0);
    if (md.siga) med.siga.assign(md.siga, md.siga + nmat);
    med.maxSigt = 0;
    for (uint32_t m = 0; m < nmat; ++m) med.maxSigt = std::max(med.maxSigt, med.sigt(m));
    if (!(med.maxSigt > 0)) return SSTMC_INVALID_CONFIG;
    // sources; the TemporalCorrelationSampler appends illumination2 / view2 (same type, count and
    // area flag, MatlabInterface.cu:1313-1331)
    const sstmc_sampler& sp = c->sampler;
    if (sp.type != SSTMC_SAMPLER_TEMPORAL && sp.type != SSTMC_SAMPLER_TEMPORAL_CORRELATION) return SSTMC_NOT_SUPPORTED;
    pb.acorr = sp.type == SSTMC_SAMPLER_TEMPORAL_CORRELATION;
    if (pb.acorr && pb.output != SSTMC_OUTPUT_FIELD) return SSTMC_INVALID_CONFIG;
    const sstmc_source* srcs[2] = {&c->illumination, &c->view};
    const sstmc_source* srcs2[2] = {&c->illumination2, &c->view2};
    Source* dst[2] = {&pb.il, &pb.vw};
    for (int i = 0; i < 2; ++i) {
        Source& d = *dst[i];
        const sstmc_source& s0 = *srcs[i];
        d.type = s0.type;
        d.illum = i == 0;
        d.area = s0.isAreaSource != 0;
        d.pt.clear();
        for (int part = 0; part < (pb.acorr ? 2 : 1); ++part) {
            const sstmc_source& s = part == 0 ? s0 : *srcs2[i];
            if (part == 1 && s.type == SSTMC_SOURCE_NONE) return SSTMC_INVALID_CONFIG;
            if (part == 1 && (s.type != s0.type || (s.isAreaSource != 0) != d.area)) return SSTMC_NOT_SUPPORTED;
            if (part == 1 && s.num != s0.num) return SSTMC_INVALID_CONFIG;
            if ((s.type != SSTMC_SOURCE_FARFIELD && s.type != SSTMC_SOURCE_POINT &&
                 s.type != SSTMC_SOURCE_GAUSSIAN_BEAM) || s.num == 0 || !s.x || !s.y || (dims == 3 && !s.z))
                return SSTMC_INVALID_CONFIG;
            for (uint32_t k = 0; k < s.num; ++k) {
                V3 p(s.x[k], s.y[k], dims == 3 ? s.z[k] : 0.0);
                d.pt.push_back(s.type == SSTMC_SOURCE_FARFIELD ? pb.geo.normalize(p) : p);
            }
            if (s.type == SSTMC_SOURCE_GAUSSIAN_BEAM) {
                // GaussianBeam (MatlabInterface.cu:1153-1190; GaussianBeam.cuh:1026-1062): focal
                // points and beam directions as given, gamma_a = 1 / aperture^2, the aperture
                // normalisation vMFnormalizationLog(gamma_a) (2D: -log 2 pi - log I_0(gamma_a),
                // SpecialFunctions.cuh:202-204)
                if (!s.dirX || !s.dirY || (dims == 3 && !s.dirZ) || !(s.aperture > 0)) return SSTMC_INVALID_CONFIG;
                for (uint32_t k = 0; k < s.num; ++k)
                    d.dir.push_back(V3(s.dirX[k], s.dirY[k], dims == 3 ? s.dirZ[k] : 0.0));
                d.gammaA = 1.0 / (s.aperture * s.aperture);
                d.apNorm = dims == 3 ? Scatter::vmfNormLog(d.gammaA)
                                     : -std::log(2.0 * PI) - logI0(cplx(d.gammaA, 0.0)).real();
            }
        }
        if (d.type == SSTMC_SOURCE_FARFIELD && i == 0) pb.buildFarField(d);
    }
    if (pb.il.type == SSTMC_SOURCE_GAUSSIAN_BEAM || pb.vw.type == SSTMC_SOURCE_GAUSSIAN_BEAM) {
        // both roles GaussianBeam (fsGaussianBeam pairs two beams, GaussianBeam.cuh:2040-2096), vMF
        // lobes in every material (updateSampler, :1283-1290), the TemporalSampler without
        // coherent backscattering
        if (pb.il.type != pb.vw.type || pb.acorr || pb.cbs) return SSTMC_NOT_SUPPORTED;
        for (uint32_t m = 1; m < nmat; ++m)
            if (pb.sc.type[m] != SSTMC_SCATTERING_VMF) return SSTMC_NOT_SUPPORTED;
        // the mean view beam direction, the axis of the second-point proposal (:1176-1194); the
        // reference draws a random axis per batch when the mean vanishes, here the z axis
        V3 mv(0, 0, 0);
        for (const V3& dv : pb.vw.dir) mv = add(mv, scale(1.0 / pb.vw.n(), dv));
        const double r = pb.geo.norm(mv);
        pb.meanView = r < 0.0001 ? V3(0, 0, 1) : scale(1.0 / r, mv);
        if (dims == 3)
            if (int e = pb.buildGaussian(pb.il)) return e;
    }
    // sampler: time table (TemporalSampler.cuh:941-1132); the correlation sampler takes the lags
    // as they are (TemporalSamplerCorrelation.cuh:455-530)
    if (sp.tNum == 0 || !sp.t || !sp.D) return SSTMC_INVALID_CONFIG;
    Frames& fr = pb.fr;
    fr.nt = sp.tNum;
    fr.temporalAmp = sp.isTemporalAmplitude != 0;
    fr.D.assign(sp.D, sp.D + nmat);
    for (double dv : fr.D)
        if (dv < 0) return SSTMC_INVALID_CONFIG;
    fr.U.assign(nmat, V3());
    for (uint32_t m = 0; m < nmat; ++m)
        fr.U[m] = V3(sp.Ux ? sp.Ux[m] : 0.0, sp.Uy ? sp.Uy[m] : 0.0, (dims == 3 && sp.Uz) ? sp.Uz[m] : 0.0);
    if (pb.acorr) {
        pb.ticks.assign(sp.t, sp.t + sp.tNum);
        return SSTMC_NO_ERROR;
    }
    double mean = 0;
    for (uint32_t t = 0; t < fr.nt; ++t) mean += sp.t[t];
    mean /= fr.nt;
    std::vector<uint32_t> neg, pos;
    std::vector<int> used(fr.nt, 0);
    for (uint32_t t = 0; t < fr.nt; ++t) (sp.t[t] - mean < 0 ? neg : pos).push_back(t);
    auto order = [&](std::vector<uint32_t>& lst, bool negative) {
        std::vector<uint32_t> out;
        for (size_t i = 0; i < lst.size(); ++i) {
            bool first = true;
            double best = 0;
            uint32_t bi = 0;
            for (uint32_t t = 0; t < fr.nt; ++t) {
                if (used[t] || ((sp.t[t] - mean < 0) != negative)) continue;
                double dd = negative ? mean - sp.t[t] : sp.t[t] - mean;
                if (first || dd < best) { first = false; best = dd; bi = t; }
            }
            used[bi] = 1;
            out.push_back(bi);
        }
        lst = out;
    };
    order(neg, true);
    order(pos, false);
    fr.neg = (uint32_t)neg.size();
    fr.dt.assign(fr.nt, 0.0);
    int prev = -1;
    for (uint32_t k = 0; k < fr.nt; ++k) {
        int tick = (int)(k < fr.neg ? neg[k] : pos[k - fr.neg]);
        if (k == 0 || k == fr.neg) prev = -1;
        fr.dt[k] = prev < 0 ? sp.t[tick] - mean : sp.t[tick] - sp.t[prev];
        prev = tick;
    }
    return SSTMC_NO_ERROR;
}

}  // namespace

// ====================================================================== C API (tests only)
extern "C" {

// One rendering: fullIteration batches of batchSize paths, path ids 0.. in order
// (MatlabInterface.cu:1783-1801 for a single rendering). u: 2*L*V*T doubles, layout u[t][v][l].
int oracle_render(const sstmc_config* cfg, double* u, double* iter, uint64_t* vertices, int nthreads) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    uint64_t npaths = (uint64_t)pb.P * pb.fullIter;
    size_t usz = (size_t)pb.L() * pb.V() * pb.T();
#ifdef _OPENMP
    int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nth = 1;
    (void)nthreads;
#endif
    std::vector<std::vector<cplx>> Ut(nth, std::vector<cplx>(usz, cplx(0)));
    std::vector<uint64_t> miss(nth, 0), nvtx(nth, 0);
    std::vector<int> errs(nth, 0);
    if (pb.tauCut > 0) {
        // the relative attenuation cut needs each batch's reference before its terms: batch by
        // batch, transport of the whole batch, then its tauReference, then the terms
        std::vector<std::vector<Vertex>> batch(pb.P);
        std::vector<uint32_t> ms(pb.P);
        for (uint32_t bi = 0; bi < pb.fullIter; ++bi) {
            const uint64_t b0 = (uint64_t)bi * pb.P;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nth) schedule(dynamic, 16)
#endif
            for (int64_t i = 0; i < (int64_t)pb.P; ++i) {
                int tid = 0;
#ifdef _OPENMP
                tid = omp_get_thread_num();
#endif
                int rc = pb.transport(b0 + (uint64_t)i, batch[i], &ms[i]);
                if (rc) errs[tid] = rc;
            }
            for (int t = 0; t < nth; ++t)
                if (errs[t]) return errs[t];
            pb.tauRef = pb.tauReference(batch);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nth) schedule(static)
#endif
            for (int64_t i = 0; i < (int64_t)pb.P; ++i) {
                int tid = 0;
#ifdef _OPENMP
                tid = omp_get_thread_num();
#endif
                Problem::Terms tm;
                miss[tid] += ms[i];
                nvtx[tid] += batch[i].size();
                pb.terms(batch[i], tm, nullptr);
                pb.accumulate(tm, Ut[tid]);
            }
        }
    } else
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        std::vector<Vertex> vs;
        Problem::Terms tm;
        uint64_t chunk = (npaths + nth - 1) / nth;
        uint64_t b = chunk * tid, e = std::min<uint64_t>(npaths, b + chunk);
        for (uint64_t p = b; p < e; ++p) {
            uint32_t ms = 0;
            int rc = pb.transport(p, vs, &ms);
            if (rc) { errs[tid] = rc; break; }
            miss[tid] += ms;
            nvtx[tid] += vs.size();
            pb.terms(vs, tm, nullptr);
            pb.accumulate(tm, Ut[tid]);
        }
    }
    for (int t = 0; t < nth; ++t)
        if (errs[t]) return errs[t];
    uint64_t mtot = 0, vtot = 0;
    for (size_t i = 0; i < usz; ++i) {
        cplx s(0);
        for (int t = 0; t < nth; ++t) s += Ut[t][i];
        u[2 * i] = s.real();
        u[2 * i + 1] = s.imag();
    }
    for (int t = 0; t < nth; ++t) { mtot += miss[t]; vtot += nvtx[t]; }
    if (iter) *iter = (double)npaths + (double)mtot;
    if (vertices) *vertices = vtot;
    return SSTMC_NO_ERROR;
}

// Per-path transport + terms for path ids [begin, begin+n). Arrays as sstmc_debug_batch in
// include/sstmc.h; call first with vtx_pos == NULL to get the vertex count.
int oracle_paths(const sstmc_config* cfg, uint64_t begin, uint32_t n, uint64_t* nvtx_out,
                 uint32_t* path_k, uint32_t* path_misses, uint64_t* path_off,
                 double* vtx_pos, uint32_t* vtx_mat, double* vtx_rho, double* vtx_xt, double* vtx_len,
                 double* term_a, double* term_b, double* term_s) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    uint32_t nl = pb.L(), nv = pb.V(), nt = pb.T();
    std::vector<std::vector<Vertex>> all(n);
    std::vector<uint32_t> ms(n);
    std::vector<int> errs(n, 0);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) errs[i] = pb.transport(begin + i, all[i], &ms[i]);
    for (uint32_t i = 0; i < n; ++i)
        if (errs[i]) return errs[i];
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; ++i) tot += all[i].size();
    if (nvtx_out) *nvtx_out = tot;
    if (!vtx_pos) return SSTMC_NO_ERROR;
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        path_k[i] = (uint32_t)all[i].size();
        path_misses[i] = ms[i];
        path_off[i] = off;
        for (const Vertex& v : all[i]) {
            for (int a = 0; a < 3; ++a) vtx_pos[3 * off + a] = v.x[a];
            vtx_mat[off] = v.m;
            vtx_rho[off] = v.rho;
            for (uint32_t t = 0; t < nt; ++t) {
                for (int a = 0; a < 3; ++a) vtx_xt[(off * nt + t) * 3 + a] = v.xt[t][a];
                vtx_len[off * nt + t] = v.len[t];
            }
            ++off;
        }
    }
    path_off[n] = off;
    if (!term_a && !term_b && !term_s) return SSTMC_NO_ERROR;   // vertices only
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        Problem::Terms tm;
        pb.terms(all[i], tm, nullptr);
        for (size_t j = 0; j < tm.A.size(); ++j) {
            term_a[2 * ((size_t)i * nt * nl + j)] = tm.multiple ? tm.A[j].real() : 0.0;
            term_a[2 * ((size_t)i * nt * nl + j) + 1] = tm.multiple ? tm.A[j].imag() : 0.0;
        }
        for (size_t j = 0; j < tm.B.size(); ++j) {
            term_b[2 * ((size_t)i * nt * nv + j)] = tm.B[j].real();
            term_b[2 * ((size_t)i * nt * nv + j) + 1] = tm.B[j].imag();
        }
        if (term_s)
            for (size_t j = 0; j < tm.S.size(); ++j) {
                term_s[2 * ((size_t)i * nt * nv * nl + j)] = tm.S[j].real();
                term_s[2 * ((size_t)i * nt * nv * nl + j) + 1] = tm.S[j].imag();
            }
    }
    return SSTMC_NO_ERROR;
}

// NEE terms in FP64 for caller-supplied vertices (for example the GPU's FP32 transport), so
// that term parity can be checked on identical paths. Inputs as produced by oracle_paths /
// sstmc_debug_batch (positions as doubles).
int oracle_terms_from_vertices(const sstmc_config* cfg, uint32_t n, const uint32_t* path_k,
                               const uint64_t* path_off, const double* vtx_pos, const uint32_t* vtx_mat,
                               const double* vtx_rho, const double* vtx_xt, const double* vtx_len,
                               double* term_a, double* term_b, double* term_s, double* term_s_sum,
                               double tau_ref) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    uint32_t nl = pb.L(), nv = pb.V(), nt = pb.T();
    // the relative attenuation cut: the given reference (NaN: this batch's own tauReference)
    if (pb.tauCut > 0) {
        if (std::isnan(tau_ref)) {
            std::vector<std::vector<Vertex>> batch(n);
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t k = 0; k < path_k[i]; ++k) {
                    const uint64_t o = path_off[i] + k;
                    Vertex v;
                    v.x = V3(vtx_pos[3 * o], vtx_pos[3 * o + 1], vtx_pos[3 * o + 2]);
                    v.xt.assign(1, V3(vtx_xt[o * nt * 3], vtx_xt[o * nt * 3 + 1], vtx_xt[o * nt * 3 + 2]));
                    batch[i].push_back(v);
                }
            pb.tauRef = pb.tauReference(batch);
        } else {
            pb.tauRef = tau_ref;
        }
    }
    // term_s_sum: sum over the n paths of the single-scattering terms, [t][v][l] complex, for
    // problems whose per-path terms are too large to return (each thread sums a contiguous range
    // of paths; the partial sums are added in thread order, so the result depends only on n and
    // the thread count)
    const size_t nsum = (size_t)nt * nv * nl;
    int nthr = 1;
#ifdef _OPENMP
    if (term_s_sum) nthr = omp_get_max_threads();
#endif
    std::vector<std::vector<cplx>> part(term_s_sum ? nthr : 0);
#ifdef _OPENMP
#pragma omp parallel num_threads(nthr)
#endif
    {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    if (term_s_sum) part[tid].assign(nsum, cplx(0));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        std::vector<Vertex> vs(path_k[i]);
        for (uint32_t k = 0; k < path_k[i]; ++k) {
            uint64_t o = path_off[i] + k;
            Vertex& v = vs[k];
            v.x = V3(vtx_pos[3 * o], vtx_pos[3 * o + 1], vtx_pos[3 * o + 2]);
            v.m = vtx_mat[o];
            v.rho = vtx_rho[o];
            v.xt.resize(nt);
            v.len.resize(nt);
            for (uint32_t t = 0; t < nt; ++t) {
                v.xt[t] = V3(vtx_xt[(o * nt + t) * 3], vtx_xt[(o * nt + t) * 3 + 1], vtx_xt[(o * nt + t) * 3 + 2]);
                v.len[t] = vtx_len[o * nt + t];
            }
        }
        Problem::Terms tm;
        pb.terms(vs, tm, nullptr);
        for (size_t j = 0; j < tm.A.size(); ++j) {
            term_a[2 * ((size_t)i * nt * nl + j)] = tm.multiple ? tm.A[j].real() : 0.0;
            term_a[2 * ((size_t)i * nt * nl + j) + 1] = tm.multiple ? tm.A[j].imag() : 0.0;
        }
        for (size_t j = 0; j < tm.B.size(); ++j) {
            term_b[2 * ((size_t)i * nt * nv + j)] = tm.B[j].real();
            term_b[2 * ((size_t)i * nt * nv + j) + 1] = tm.B[j].imag();
        }
        if (term_s)
            for (size_t j = 0; j < tm.S.size(); ++j) {
                term_s[2 * ((size_t)i * nt * nv * nl + j)] = tm.S[j].real();
                term_s[2 * ((size_t)i * nt * nv * nl + j) + 1] = tm.S[j].imag();
            }
        if (term_s_sum) {
            // correlation output: the path's single scattering enters as S(0) conj S(t), as in
            // Problem::accumulate
            const bool corr = pb.output == SSTMC_OUTPUT_CORRELATION;
            for (size_t j = 0; j < tm.S.size() && j < nsum; ++j)
                part[tid][j] += corr ? tm.S[j % ((size_t)nv * nl)] * std::conj(tm.S[j]) : tm.S[j];
        }
    }
    }
    if (term_s_sum) {
        for (size_t j = 0; j < nsum; ++j) {
            cplx acc(0);
            for (int q = 0; q < nthr; ++q) acc += part[q][j];
            term_s_sum[2 * j] = acc.real();
            term_s_sum[2 * j + 1] = acc.imag();
        }
    }
    return SSTMC_NO_ERROR;
}

// Coherent-backscattering terms (Ar [n][t][v], Br [n][t][l], complex interleaved) for
// caller-supplied vertices, with the same inputs as oracle_terms_from_vertices. Zero rows for
// paths with fewer than two vertices.
int oracle_cbs_terms_from_vertices(const sstmc_config* cfg, uint32_t n, const uint32_t* path_k,
                                   const uint64_t* path_off, const double* vtx_pos, const uint32_t* vtx_mat,
                                   const double* vtx_rho, const double* vtx_xt, const double* vtx_len,
                                   double* term_ar, double* term_br) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.cbs) return SSTMC_INVALID_CONFIG;
    uint32_t nl = pb.L(), nv = pb.V(), nt = pb.T();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        std::vector<Vertex> vs(path_k[i]);
        for (uint32_t k = 0; k < path_k[i]; ++k) {
            uint64_t o = path_off[i] + k;
            Vertex& v = vs[k];
            v.x = V3(vtx_pos[3 * o], vtx_pos[3 * o + 1], vtx_pos[3 * o + 2]);
            v.m = vtx_mat[o];
            v.rho = vtx_rho[o];
            v.xt.resize(nt);
            v.len.resize(nt);
            for (uint32_t t = 0; t < nt; ++t) {
                v.xt[t] = V3(vtx_xt[(o * nt + t) * 3], vtx_xt[(o * nt + t) * 3 + 1], vtx_xt[(o * nt + t) * 3 + 2]);
                v.len[t] = vtx_len[o * nt + t];
            }
        }
        Problem::Terms tm;
        pb.terms(vs, tm, nullptr);
        for (size_t j = 0; j < (size_t)nt * nv; ++j) {
            cplx x = tm.multiple ? tm.Ar[j] : cplx(0);
            term_ar[2 * ((size_t)i * nt * nv + j)] = x.real();
            term_ar[2 * ((size_t)i * nt * nv + j) + 1] = x.imag();
        }
        for (size_t j = 0; j < (size_t)nt * nl; ++j) {
            cplx x = tm.multiple ? tm.Br[j] : cplx(0);
            term_br[2 * ((size_t)i * nt * nl + j)] = x.real();
            term_br[2 * ((size_t)i * nt * nl + j) + 1] = x.imag();
        }
    }
    return SSTMC_NO_ERROR;
}

// TemporalCorrelationSampler with coherent backscattering: the mixed pairings CBS02 + CBS03
// summed over the given paths, x_sum [t][v][l] complex interleaved (unscaled: the field takes
// half of it), with the same vertex inputs as oracle_terms_from_vertices.
int oracle_cbs_cross_from_vertices(const sstmc_config* cfg, uint32_t n, const uint32_t* path_k,
                                   const uint64_t* path_off, const double* vtx_pos, const uint32_t* vtx_mat,
                                   const double* vtx_rho, const double* vtx_xt, const double* vtx_len, double* x_sum) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.cbs || !pb.acorr) return SSTMC_INVALID_CONFIG;
    const uint32_t nl = pb.L(), nv = pb.V(), nt = pb.T();
    const size_t nx = (size_t)nt * nv * nl;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<std::vector<cplx>> part(nthr, std::vector<cplx>(nx, cplx(0)));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        std::vector<Vertex> vs(path_k[i]);
        for (uint32_t k = 0; k < path_k[i]; ++k) {
            const uint64_t o = path_off[i] + k;
            Vertex& v = vs[k];
            v.x = V3(vtx_pos[3 * o], vtx_pos[3 * o + 1], vtx_pos[3 * o + 2]);
            v.m = vtx_mat[o];
            v.rho = vtx_rho[o];
            v.xt.resize(nt);
            v.len.resize(nt);
            for (uint32_t t = 0; t < nt; ++t) {
                v.xt[t] = V3(vtx_xt[(o * nt + t) * 3], vtx_xt[(o * nt + t) * 3 + 1], vtx_xt[(o * nt + t) * 3 + 2]);
                v.len[t] = vtx_len[o * nt + t];
            }
        }
        if (vs.size() < 2) continue;
        Problem::Terms tm;
        pb.terms(vs, tm, nullptr);
        for (size_t j = 0; j < nx; ++j) part[tid][j] += tm.X[j];
    }
    for (size_t j = 0; j < nx; ++j) {
        cplx acc(0);
        for (int q = 0; q < nthr; ++q) acc += part[q][j];
        x_sum[2 * j] = acc.real();
        x_sum[2 * j + 1] = acc.imag();
    }
    return SSTMC_NO_ERROR;
}

// log I_0(z) of the 2D GaussianBeam forms, out = (re, im), for checks against a library routine.
int oracle_log_i0(double re, double im, double* out) {
    const cplx v = logI0(cplx(re, im));
    out[0] = v.real();
    out[1] = v.imag();
    return SSTMC_NO_ERROR;
}

// --- GaussianBeam pieces, for quadrature and sampling checks --------------------------------
// f of illumination (view = 0) or view (view = 1) beam i at x0 with the path continuing to / coming
// from p2; fs of illumination beam l and view beam v at x0; out = (re, im).
int oracle_gauss_f(const sstmc_config* cfg, int view, uint32_t i, const double* x0, const double* p2, double* out) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.gauss()) return SSTMC_INVALID_CONFIG;
    cplx f = pb.fGauss(view ? pb.vw : pb.il, i, V3(x0[0], x0[1], x0[2]), V3(p2[0], p2[1], p2[2]));
    out[0] = f.real();
    out[1] = f.imag();
    return SSTMC_NO_ERROR;
}
int oracle_gauss_fs(const sstmc_config* cfg, uint32_t l, uint32_t v, const double* x0, double* out) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.gauss()) return SSTMC_INVALID_CONFIG;
    cplx f = pb.fsGauss(l, v, V3(x0[0], x0[1], x0[2]));
    out[0] = f.real();
    out[1] = f.imag();
    return SSTMC_NO_ERROR;
}
// first-point probability of beam l at x1 (n points), second-point proposal density of beam l
// at x0 toward each x1 (n points)
int oracle_gauss_pr(const sstmc_config* cfg, uint32_t l, uint32_t n, const double* x1, double* out) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.gauss()) return SSTMC_INVALID_CONFIG;
    for (uint32_t i = 0; i < n; ++i) out[i] = pb.prGauss(l, V3(x1[3 * i], x1[3 * i + 1], x1[3 * i + 2]));
    return SSTMC_NO_ERROR;
}
int oracle_gauss_pdir(const sstmc_config* cfg, uint32_t l, const double* x0, uint32_t n, const double* x1,
                      double* out) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.gauss()) return SSTMC_INVALID_CONFIG;
    for (uint32_t i = 0; i < n; ++i)
        out[i] = pb.pdirGauss(l, V3(x0[0], x0[1], x0[2]), V3(x1[3 * i], x1[3 * i + 1], x1[3 * i + 2]));
    return SSTMC_NO_ERROR;
}
// n first points drawn from the beam profile (path ids 0..n-1, no rejection outside the medium),
// their beams; n second-point directions of beam l at x0
int oracle_gauss_sample_first(const sstmc_config* cfg, uint32_t n, double* pts, uint32_t* beams) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.gauss()) return SSTMC_INVALID_CONFIG;
    for (uint32_t i = 0; i < n; ++i) {
        Stream s(pb.seed, i, 0, 0);
        V3 x = pb.sampleGaussFirst(s, &beams[i]);
        for (int a = 0; a < 3; ++a) pts[3 * i + a] = x[a];
    }
    return SSTMC_NO_ERROR;
}
int oracle_gauss_sample_second(const sstmc_config* cfg, uint32_t l, const double* x0, uint32_t n, double* dirs) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (!pb.gauss()) return SSTMC_INVALID_CONFIG;
    for (uint32_t i = 0; i < n; ++i) {
        Stream s(pb.seed, i, 0x80000002u, 0);
        V3 d = pb.sampleGaussSecond(l, V3(x0[0], x0[1], x0[2]), s);
        for (int a = 0; a < 3; ++a) dirs[3 * i + a] = d[a];
    }
    return SSTMC_NO_ERROR;
}

// --- unit-level entry points, for closed-form and invariant checks -------------------------
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    Philox::block(ctr, key, out);
}

// Woodcock / DDA / scattering primitives on a built problem.
double oracle_optical_depth(const sstmc_config* cfg, const double* p, const double* d, uint64_t* steps) {
    Problem pb;
    if (build(cfg, pb)) return -1.0;
    return pb.med.opticalDepth(V3(p[0], p[1], pb.dims == 3 ? p[2] : 0), V3(d[0], d[1], pb.dims == 3 ? d[2] : 0), steps);
}

// Samples n free flights from origin o along d with path ids [0,n) (event 0); writes the
// distance to the sampled point and the returned material (0 = escaped/absorbed).
int oracle_free_flights(const sstmc_config* cfg, const double* o, const double* d, uint32_t n,
                        double* dist, int32_t* mat) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    V3 O(o[0], o[1], pb.dims == 3 ? o[2] : 0), D(d[0], d[1], pb.dims == 3 ? d[2] : 0);
    for (uint32_t i = 0; i < n; ++i) {
        Stream s(pb.seed, i, 0, 0);
        V3 x;
        mat[i] = pb.med.woodcock(O, D, s, &x);
        dist[i] = pb.geo.norm(sub(x, O));
    }
    return SSTMC_NO_ERROR;
}

// Samples n scattering directions for material m about prev (path ids [0,n), event 0).
int oracle_sample_directions(const sstmc_config* cfg, uint32_t m, const double* prev, uint32_t n, double* out) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    V3 pv(prev[0], prev[1], pb.dims == 3 ? prev[2] : 0);
    for (uint32_t i = 0; i < n; ++i) {
        Stream s(pb.seed, i, 0, 0);
        V3 d = pb.sc.sample(m, pv, s);
        for (int a = 0; a < 3; ++a) out[3 * i + a] = d[a];
    }
    return SSTMC_NO_ERROR;
}

double oracle_phase_amplitude(const sstmc_config* cfg, uint32_t m, double cosang, int pdf) {
    Problem pb;
    if (build(cfg, pb)) return -1.0;
    return pdf ? pb.sc.pdf(m, cosang) : pb.sc.amp(m, cosang).real();
}

// Complex amplitude of material m at cosine c: out[0] re, out[1] im.
int oracle_phase_amplitude_complex(const sstmc_config* cfg, uint32_t m, double cosang, double* out) {
    Problem pb;
    int rc = build(cfg, pb);
    if (rc) return rc;
    cplx f = pb.sc.amp(m, cosang);
    out[0] = f.real();
    out[1] = f.imag();
    return SSTMC_NO_ERROR;
}

// The tabular table of material m as built: f (2n doubles, interleaved), evaluation pdf (n),
// inverse-cdf angles (n). Returns n (0 when m is not tabular); NULL outputs are skipped.
uint32_t oracle_tabular_table(const sstmc_config* cfg, uint32_t m, double* f, double* pdf, double* icdf) {
    Problem pb;
    if (build(cfg, pb) || m >= pb.sc.tab.size()) return 0;
    const Scatter::Table& t = pb.sc.tab[m];
    for (size_t i = 0; i < t.f.size(); ++i) {
        if (f) { f[2 * i] = t.f[i].real(); f[2 * i + 1] = t.f[i].imag(); }
        if (pdf) pdf[i] = t.pdf[i];
        if (icdf) icdf[i] = t.icdf[i];
    }
    return (uint32_t)t.f.size();
}

// FarField sampling table of illumination l: tangent, perpT, perpS, minT, maxT, minS, maxS, prob.
int oracle_farfield_table(const sstmc_config* cfg, uint32_t l, double* out16) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    if (pb.il.type != SSTMC_SOURCE_FARFIELD || l >= pb.L()) return SSTMC_INVALID_CONFIG;
    for (int a = 0; a < 3; ++a) {
        out16[a] = pb.il.tangent[l][a];
        out16[3 + a] = pb.il.perpT[l][a];
        out16[6 + a] = pb.il.perpS[l][a];
    }
    out16[9] = pb.il.minT[l]; out16[10] = pb.il.maxT[l];
    out16[11] = pb.il.minS[l]; out16[12] = pb.il.maxS[l];
    out16[13] = pb.il.prob[l];
    return SSTMC_NO_ERROR;
}

// Frame chain table: dt per slot and the number of negative-chain slots.
int oracle_frame_table(const sstmc_config* cfg, double* dt, uint32_t* neg) {
    Problem pb;
    int err = build(cfg, pb);
    if (err) return err;
    for (uint32_t k = 0; k < pb.fr.nt; ++k) dt[k] = pb.fr.dt[k];
    *neg = pb.fr.neg;
    return SSTMC_NO_ERROR;
}

}  // extern "C"
