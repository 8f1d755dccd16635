// tracer.hip -- host implementation of the plugin classes and the per-batch driver.
#include "tracer.h"
#include "gauss.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace sstmc {

static const double kPi = 3.14159265358979323846;
static const double kEps = 1e-12;   // EPSILON of the FP64 reference (Simulation.cuh:36-40)

template <typename T>
int DevBuf<T>::alloc(size_t count) {
    release();
    if (count == 0) return SSTMC_NO_ERROR;
    if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess) {
        p = nullptr;
        return SSTMC_ALLOCATION_ERROR;
    }
    n = count;
    return SSTMC_NO_ERROR;
}
template <typename T>
int DevBuf<T>::upload(const T* host, size_t count) {
    int e = alloc(count);
    if (e) return e;
    if (count && hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return SSTMC_ALLOCATION_ERROR;
    return SSTMC_NO_ERROR;
}
template <typename T>
void DevBuf<T>::release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
}
template struct DevBuf<float>;
template struct DevBuf<double>;
template struct DevBuf<uint8_t>;
template struct DevBuf<uint32_t>;
template struct DevBuf<float4>;
template struct DevBuf<double4>;
template struct DevBuf<float2>;
template struct DevBuf<double2>;
template struct DevBuf<PhaseTables>;
template struct DevBuf<unsigned long long>;
template struct DevBuf<VertexRec>;
template struct DevBuf<FarRec>;
template struct DevBuf<BoundDirs>;

// ====================================================================== HeterogeneousMedium
HeterogeneousMedium::HeterogeneousMedium(int* err, int dims, const uint32_t* grid, const double* xAxis,
                                         uint32_t xAxisSize, const double* yAxis, uint32_t yAxisSize,
                                         const double* zAxis, uint32_t zAxisSize) {
    *err = SSTMC_NO_ERROR;
    v_.dims = dims;
    const double* ax[3] = {xAxis, yAxis, zAxis};
    uint32_t sz[3] = {xAxisSize, yAxisSize, dims == 3 ? zAxisSize : 2u};
    static const double flat[2] = {0.0, 0.0};
    for (int a = 0; a < 3; ++a) {
        const double* e = a < dims ? ax[a] : flat;
        uint32_t n = sz[a];
        v_.n[a] = n - 1;
        v_.nf[a] = (float)(n - 1);
        std::vector<float> f(n);
        for (uint32_t i = 0; i < n; ++i) f[i] = (float)e[i];
        if ((*err = ef_[a].upload(f.data(), n))) return;
        if ((*err = ed_[a].upload(e, n))) return;
        v_.ef[a] = ef_[a].p;
        v_.ed[a] = ed_[a].p;
        v_.bmind[a] = e[0];
        v_.bmaxd[a] = e[n - 1];
        v_.bminf[a] = (float)e[0];
        v_.bmaxf[a] = (float)e[n - 1];
        // equal spacing: the fast locate path computes the cell and corrects it against the edges
        double h = (e[n - 1] - e[0]) / (double)(n - 1);
        bool uni = n >= 2 && h > 0;
        for (uint32_t i = 0; i < n && uni; ++i)
            if (std::fabs(e[i] - (e[0] + h * i)) > 1e-9 * std::max(1.0, std::fabs(e[i]))) uni = false;
        v_.uniformAxis[a] = uni ? 1 : 0;
        v_.e0[a] = (float)e[0];
        v_.invH[a] = uni ? (float)(1.0 / h) : 0.f;
        v_.h[a] = (float)h;
    }
    v_.allUniform = v_.uniformAxis[0] && v_.uniformAxis[1] && (dims == 2 || v_.uniformAxis[2]) ? 1 : 0;
    size_t ncell = (size_t)v_.n[0] * v_.n[1] * v_.n[2];
    hostGrid_.resize(ncell);
    for (size_t i = 0; i < ncell; ++i) hostGrid_[i] = (uint8_t)grid[i];
    if ((*err = grid_.upload(hostGrid_.data(), ncell))) return;
    v_.grid = grid_.p;
}

int HeterogeneousMedium::setMaterial(const double* sigs, const double* siga, uint32_t materialNum) {
    nmat_ = materialNum;
    sigs_.assign(sigs, sigs + materialNum);
    siga_.assign(materialNum, 0.0);
    if (siga) siga_.assign(siga, siga + materialNum);
    // majorant over every material (HeterogeneousMedium::setMaterial, :820-860)
    double maj = 0;
    for (uint32_t m = 0; m < materialNum; ++m) maj = std::max(maj, sigs_[m] + siga_[m]);
    if (!(maj > 0)) return SSTMC_INVALID_CONFIG;
    for (uint32_t m = 0; m < (uint32_t)kMaxMaterials; ++m) {
        double st = m < materialNum ? sigs_[m] + siga_[m] : 0.0;
        double ss = m < materialNum ? sigs_[m] : 0.0;
        v_.sigtf[m] = (float)st;
        v_.sigtd[m] = st;
        v_.accept[m] = (float)(st / maj);
        v_.albedo[m] = st > 0 ? (float)(ss / st) : 0.f;
    }
    v_.invMaj = (float)(1.0 / maj);
    return SSTMC_NO_ERROR;
}

// ====================================================================== Scattering
Scattering::Scattering(int dims) { v_.dims = dims; }

int Scattering::setIsotropicScattering(uint32_t m) {
    if (m >= (uint32_t)kMaxMaterials) return SSTMC_INVALID_CONFIG;
    v_.type[m] = SSTMC_SCATTERING_ISOTROPIC;
    v_.g[m] = 0.f;
    v_.gd[m] = 0.0;
    defined_[m] = true;
    v_.allIsotropic = isIsotropicScattering() ? 1 : 0;
    return SSTMC_NO_ERROR;
}

int Scattering::setHenyeyGreensteinScattering(double g, uint32_t m) {
    if (m >= (uint32_t)kMaxMaterials || !(std::fabs(g) < 1.0)) return SSTMC_INVALID_CONFIG;
    v_.type[m] = SSTMC_SCATTERING_HG;
    v_.g[m] = (float)g;
    v_.gd[m] = g;
    defined_[m] = true;
    v_.allIsotropic = isIsotropicScattering() ? 1 : 0;
    return SSTMC_NO_ERROR;
}

// Tabular scattering (Scattering.cuh:272-429), on the host in FP64: the raw sampling pdf |f|^2
// (times sin(theta) in 3D, theta = i pi / (n-1)) is normalised to sum 1; the cross section
// (sum / n) 2 pi (times pi in 3D) normalises the amplitude; the inverse cdf of entry i is the
// first cdf index at or above i / (n-1) (lower_bound), scaled to the angle range; the evaluation
// pdf is |f|^2 (times sin(theta) in 3D) of the normalised amplitude.
int Scattering::setTabularScattering(const double* f, uint32_t n, uint32_t m) {
    if (m >= (uint32_t)kMaxMaterials || n < 2 || !f) return SSTMC_INVALID_CONFIG;
    const bool d2 = v_.dims == 2;
    std::vector<double> pdf(n), cdf(n);
    double sum = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double th = ((double)i / (double)(n - 1)) * kPi;
        pdf[i] = (f[2 * i] * f[2 * i] + f[2 * i + 1] * f[2 * i + 1]) * (d2 ? 1.0 : std::sin(th));
        sum += pdf[i];
    }
    if (!(sum > 0) || !std::isfinite(sum)) return SSTMC_INVALID_CONFIG;
    double acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        pdf[i] *= 1.0 / sum;
        acc += pdf[i];
        cdf[i] = acc;
    }
    const double cross = (sum / (double)n) * 2.0 * kPi * (d2 ? 1.0 : kPi);
    const double scale = std::sqrt(1.0 / cross);
    const double range = d2 ? 2.0 * kPi : kPi;
    tab_.n[m] = n;
    tab_.off[m] = (uint32_t)tabF_.size();
    rawOff_[m] = (uint32_t)tabRaw_.size();
    tabRaw_.insert(tabRaw_.end(), f, f + 2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        const double th = ((double)i / (double)(n - 1)) * kPi;
        const double re = f[2 * i] * scale, im = f[2 * i + 1] * scale;
        const double a = std::hypot(re, im);
        const double pe = (re * re + im * im) * (d2 ? 1.0 : std::sin(th));
        const double q = (double)i / (double)(n - 1);
        const size_t lb = (size_t)(std::lower_bound(cdf.begin(), cdf.end(), q) - cdf.begin());
        tabF_.push_back(make_float4((float)a, (float)re, (float)im, (float)pe));
        tabD_.push_back(make_double4(a, re, im, pe));
        tabPh_.push_back(a > 0 ? std::atan2(im, re) / (2.0 * kPi) : 0.0);
        tabIcdf_.push_back((float)((double)lb * (range / (double)(n - 1))));
    }
    v_.type[m] = SSTMC_SCATTERING_TABULAR;
    v_.g[m] = 0.f;
    v_.gd[m] = 0.0;
    defined_[m] = true;
    v_.allIsotropic = isIsotropicScattering() ? 1 : 0;
    return SSTMC_NO_ERROR;
}

// Scattering.cuh:635-817: per mixture kappa and C + log(alpha); per pair (i + M j) the summed
// concentration, its log normalisation, the normalised weight exp(C_i + C_j + log alpha_i alpha_j -
// lognorm) and the exclusive cumulative weight (thrust::exclusive_scan).
int Scattering::setVonMisesFisherScattering(const double* mu, const double* C, const double* alpha, uint32_t M,
                                            uint32_t m) {
    if (m >= (uint32_t)kMaxMaterials || M == 0 || !mu || !C || !alpha) return SSTMC_INVALID_CONFIG;
    for (uint32_t i = 0; i < M; ++i)
        if (!(alpha[i] > 0) || !std::isfinite(mu[i]) || !std::isfinite(C[i])) return SSTMC_INVALID_CONFIG;
    tab_.vmfM[m] = M;
    tab_.vmfOff[m] = (uint32_t)vmfKappa_.size();
    tab_.vmfPairOff[m] = (uint32_t)vmfPairMu_.size();
    vmfRawOff_[m] = (uint32_t)vmfRaw_.size();
    bool iso = true;
    for (uint32_t i = 0; i < M; ++i) {
        vmfKappa_.push_back(mu[i]);
        vmfC_.push_back(C[i] + std::log(alpha[i]));
        vmfRaw_.insert(vmfRaw_.end(), {mu[i], C[i], alpha[i]});
        if (std::fabs(mu[i]) > 0.0001) iso = false;
    }
    std::vector<double> w(M * M);
    double sum = 0;
    for (uint32_t k = 0; k < M * M; ++k) {
        const uint32_t i = k % M, j = k / M;
        const double s = mu[i] + mu[j];
        vmfPairMu_.push_back(s);
        vmfPairLogNorm_.push_back(vmfNormLog3(s));
        w[k] = std::exp(C[i] + std::log(alpha[i]) + C[j] + std::log(alpha[j]) - vmfNormLog3(s));
        sum += w[k];
    }
    if (!(sum > 0) || !std::isfinite(sum)) return SSTMC_INVALID_CONFIG;
    double acc = 0;
    for (uint32_t k = 0; k < M * M; ++k) {
        vmfPairPdf_.push_back(w[k] / sum);
        vmfPairCdf_.push_back(acc);
        acc += w[k] / sum;
    }
    vmfIso_[m] = iso;
    v_.type[m] = SSTMC_SCATTERING_VMF;
    v_.g[m] = 0.f;
    v_.gd[m] = 0.0;
    defined_[m] = true;
    v_.allIsotropic = isIsotropicScattering() ? 1 : 0;
    return SSTMC_NO_ERROR;
}

int Scattering::upload() {
    if (tabF_.empty() && vmfKappa_.empty()) return SSTMC_NO_ERROR;
    int e;
    if (!tabF_.empty()) {
        if ((e = dTabF_.upload(tabF_.data(), tabF_.size())) || (e = dTabD_.upload(tabD_.data(), tabD_.size())) ||
            (e = dTabPh_.upload(tabPh_.data(), tabPh_.size())) ||
            (e = dTabIcdf_.upload(tabIcdf_.data(), tabIcdf_.size())))
            return e;
        tab_.f = dTabF_.p;
        tab_.d = dTabD_.p;
        tab_.ph = dTabPh_.p;
        tab_.icdf = dTabIcdf_.p;
    }
    if (!vmfKappa_.empty()) {
        const size_t nm = vmfKappa_.size(), np = vmfPairMu_.size();
        std::vector<double> all;
        all.reserve(2 * nm + 4 * np);
        for (const std::vector<double>* v : {&vmfKappa_, &vmfC_, &vmfPairMu_, &vmfPairLogNorm_, &vmfPairPdf_, &vmfPairCdf_})
            all.insert(all.end(), v->begin(), v->end());
        if ((e = dVmf_.upload(all.data(), all.size()))) return e;
        tab_.vmfKappa = dVmf_.p;
        tab_.vmfC = dVmf_.p + nm;
        tab_.vmfPairMu = dVmf_.p + 2 * nm;
        tab_.vmfPairLogNorm = dVmf_.p + 2 * nm + np;
        tab_.vmfPairPdf = dVmf_.p + 2 * nm + 2 * np;
        tab_.vmfPairCdf = dVmf_.p + 2 * nm + 3 * np;
    }
    if ((e = dTab_.upload(&tab_, 1))) return e;
    v_.tab = dTab_.p;
    return SSTMC_NO_ERROR;
}

bool Scattering::isIsotropicScattering() const {
    for (int m = 0; m < kMaxMaterials; ++m) {
        if (!defined_[m]) continue;
        if (v_.type[m] == SSTMC_SCATTERING_TABULAR) return false;   // isIsotropicScatteringTabular
        if (v_.type[m] == SSTMC_SCATTERING_VMF && !vmfIso_[m]) return false;   // all |kappa| <= 1e-4
        if (v_.type[m] == SSTMC_SCATTERING_HG && !(std::fabs(v_.gd[m]) < 0.0001)) return false;
    }
    return true;
}

bool Scattering::isUniform(uint32_t n) const {
    for (uint32_t m = 2; m < n; ++m) {
        if (v_.type[m] != v_.type[1] || (v_.type[m] == SSTMC_SCATTERING_HG && v_.gd[m] != v_.gd[1])) return false;
        if (v_.type[m] == SSTMC_SCATTERING_TABULAR &&
            (tab_.n[m] != tab_.n[1] ||
             !std::equal(tabRaw_.begin() + rawOff_[m], tabRaw_.begin() + rawOff_[m] + 2 * (size_t)tab_.n[m],
                         tabRaw_.begin() + rawOff_[1])))
            return false;
        if (v_.type[m] == SSTMC_SCATTERING_VMF &&
            (tab_.vmfM[m] != tab_.vmfM[1] ||
             !std::equal(vmfRaw_.begin() + vmfRawOff_[m], vmfRaw_.begin() + vmfRawOff_[m] + 3 * (size_t)tab_.vmfM[m],
                         vmfRaw_.begin() + vmfRawOff_[1])))
            return false;
    }
    return true;
}

bool Scattering::isExactlyIsotropic(uint32_t n) const {
    for (uint32_t m = 1; m < n; ++m)
        if (!(v_.type[m] == SSTMC_SCATTERING_ISOTROPIC || (v_.type[m] == SSTMC_SCATTERING_HG && v_.gd[m] == 0.0) ||
              (v_.type[m] == SSTMC_SCATTERING_VMF && v_.dims == 2)))
            return false;
    return true;
}

void Scattering::amplitude(uint32_t m, double c, double* re, double* im) const {
    *im = 0.0;
    if (v_.type[m] == SSTMC_SCATTERING_ISOTROPIC || (v_.type[m] == SSTMC_SCATTERING_VMF && v_.dims == 2)) {
        *re = v_.dims == 3 ? 0.282094791773878143 : 0.398942280401432678;
        return;
    }
    if (v_.type[m] == SSTMC_SCATTERING_VMF) {
        double a = 0;
        for (uint32_t i = tab_.vmfOff[m]; i < tab_.vmfOff[m] + tab_.vmfM[m]; ++i) a += std::exp(c * vmfKappa_[i] + vmfC_[i]);
        *re = a;
        return;
    }
    if (v_.type[m] == SSTMC_SCATTERING_TABULAR) {
        const double4 e = tabD_[tab_.off[m] + tabularEntry(tab_.n[m], v_.dims, c)];
        *re = e.y;
        *im = e.z;
        return;
    }
    double g = v_.gd[m];
    double up = v_.dims == 3 ? (1.0 - g * g) / (4.0 * kPi) : (1.0 - g * g) / (2.0 * kPi);
    double den = std::fma(-2.0 * g, c, 1.0 + g * g);
    *re = std::sqrt(v_.dims == 3 ? up / (den * std::sqrt(den)) : up / den);
}

// ====================================================================== Sources
static void normalize(int dims, double* v) {
    double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + (dims == 3 ? v[2] * v[2] : 0.0));
    if (n == 0) {
        v[0] = 0; v[1] = dims == 3 ? 0 : 1; v[2] = dims == 3 ? 1 : 0;
        return;
    }
    v[0] /= n; v[1] /= n; v[2] = dims == 3 ? v[2] / n : 0.0;
}

// projectBoxToLine (MediumInterface.cuh:156-199): extent of the box corners projected on a line.
static void projectBox(const Medium& m, int dims, const double* p0, const double* v, double* tmin, double* tmax) {
    int nc = dims == 3 ? 8 : 4;
    for (int c = 0; c < nc; ++c) {
        double p[3] = {(c & 1) ? m.boxMax(0) : m.boxMin(0), (c & 2) ? m.boxMax(1) : m.boxMin(1),
                       (c & 4) ? m.boxMax(2) : m.boxMin(2)};
        double q[3] = {p[0] - p0[0], p[1] - p0[1], p[2] - p0[2]};
        double o[3] = {0, 0, 0};
        for (int a = 0; a < dims; ++a)
            for (int b = 0; b < dims; ++b) o[a] += v[a] * v[b] * q[b];
        double t = 0;
        for (int a = 0; a < dims; ++a) t += o[a] * v[a];
        if (c == 0 || t < *tmin) *tmin = t;
        if (c == 0 || t > *tmax) *tmax = t;
    }
}

FarField::FarField(int* err, const Medium& med, int dims, const double* x, const double* y, const double* z,
                   uint32_t n, bool illumination) {
    *err = SSTMC_NO_ERROR;
    v_.type = SSTMC_SOURCE_FARFIELD;
    v_.illum = illumination ? 1 : 0;
    v_.n = n;
    pts_.assign(3 * (size_t)n, 0.0);
    std::vector<float> pf(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        double d[3] = {x[i], y[i], dims == 3 ? z[i] : 0.0};
        normalize(dims, d);
        for (int a = 0; a < 3; ++a) { pts_[3 * i + a] = d[a]; pf[3 * i + a] = (float)d[a]; }
    }
    if ((*err = pd_.upload(pts_.data(), pts_.size()))) return;
    if ((*err = pf_.upload(pf.data(), pf.size()))) return;
    v_.pf = pf_.p;
    v_.pd = pd_.p;
    if (!illumination) return;
    // fillBuffers (FarField.cuh:71-127)
    double r[3], ctr[3], rmax = -1e300;
    for (int a = 0; a < 3; ++a) {
        r[a] = 0.5 * (med.boxMax(a) - med.boxMin(a));
        ctr[a] = 0.5 * (med.boxMax(a) + med.boxMin(a));
        if (a < dims) rmax = std::max(rmax, r[a]);
    }
    double rad = rmax + kEps;
    std::vector<float> tab((size_t)n * kFarFieldTable, 0.f);
    std::vector<float> probf(n);
    std::vector<double> probd(n);
    for (uint32_t i = 0; i < n; ++i) {
        const double* d = &pts_[3 * i];
        double opp[3] = {-d[0], -d[1], -d[2]};
        double tang[3] = {ctr[0] + rad * opp[0], ctr[1] + rad * opp[1], dims == 3 ? ctr[2] + rad * opp[2] : 0.0};
        double T[3] = {0, 0, 0}, S[3] = {0, 0, 0};
        double minT = 0, maxT = 0, minS = 0, maxS = 0, prob;
        if (dims == 3) {
            if (std::fabs(1.0 - std::fabs(opp[2])) <= kEps) {
                T[0] = 1; S[1] = 1;
            } else {
                T[0] = -opp[1]; T[1] = opp[0]; T[2] = 0;
                normalize(3, T);
                S[0] = T[1] * opp[2] - T[2] * opp[1];
                S[1] = T[2] * opp[0] - T[0] * opp[2];
                S[2] = T[0] * opp[1] - T[1] * opp[0];
            }
            projectBox(med, dims, tang, T, &minT, &maxT);
            projectBox(med, dims, tang, S, &minS, &maxS);
            prob = 1.0 / ((maxT - minT) * (maxS - minS));
        } else {
            T[0] = -opp[1]; T[1] = opp[0];
            projectBox(med, dims, tang, T, &minT, &maxT);
            prob = 1.0 / (maxT - minT);
        }
        float* t = &tab[(size_t)i * kFarFieldTable];
        for (int a = 0; a < 3; ++a) {
            t[a] = (float)tang[a];
            t[3 + a] = (float)T[a];
            t[6 + a] = (float)S[a];
            t[14 + a] = (float)(std::fabs(d[a]) < kEps ? 1e12 : 1.0 / d[a]);
            t[17 + a] = (float)d[a];
        }
        t[9] = (float)minT; t[10] = (float)maxT; t[11] = (float)minS; t[12] = (float)maxS;
        probf[i] = (float)prob;
        probd[i] = prob;
    }
    if ((*err = table_.upload(tab.data(), tab.size()))) return;
    if ((*err = probf_.upload(probf.data(), n))) return;
    if ((*err = probd_.upload(probd.data(), n))) return;
    v_.table = table_.p;
    v_.probf = probf_.p;
    v_.probd = probd_.p;
}

PointSource::PointSource(int* err, int dims, const double* x, const double* y, const double* z, uint32_t n,
                         bool areaSource, bool illumination) {
    *err = SSTMC_NO_ERROR;
    v_.type = SSTMC_SOURCE_POINT;
    v_.illum = illumination ? 1 : 0;
    v_.area = areaSource ? 1 : 0;
    v_.n = n;
    pts_.assign(3 * (size_t)n, 0.0);
    std::vector<float> pf(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        pts_[3 * i] = x[i];
        pts_[3 * i + 1] = y[i];
        pts_[3 * i + 2] = dims == 3 ? z[i] : 0.0;
        for (int a = 0; a < 3; ++a) pf[3 * i + a] = (float)pts_[3 * i + a];
    }
    if ((*err = pd_.upload(pts_.data(), pts_.size()))) return;
    if ((*err = pf_.upload(pf.data(), pf.size()))) return;
    v_.pf = pf_.p;
    v_.pd = pd_.p;
}

// ====================================================================== GaussianBeam
uint32_t Scattering::vmfLobes(uint32_t m, const double** kappa, const double** c) const {
    if (m >= (uint32_t)kMaxMaterials || v_.type[m] != SSTMC_SCATTERING_VMF) return 0;
    *kappa = vmfKappa_.data() + tab_.vmfOff[m];
    *c = vmfC_.data() + tab_.vmfOff[m];
    return tab_.vmfM[m];
}

GaussianBeam::GaussianBeam(int* err, const Medium& medium, const Scattering& scattering, const sstmc_source& s,
                           const sstmc_source* views, double lambda, bool illumination) {
    *err = SSTMC_NO_ERROR;
    const uint32_t n = s.num;
    const double PI = 3.14159265358979323846;
    v_.type = SSTMC_SOURCE_GAUSSIAN_BEAM;
    v_.illum = illumination ? 1 : 0;
    v_.n = n;
    const bool d3d = medium.dims() == 3;
    if (n == 0 || !s.dirX || !s.dirY || (d3d && (!s.dirZ || !s.z)) || !(s.aperture > 0) || !(lambda > 0)) {
        *err = SSTMC_INVALID_CONFIG;
        return;
    }
    pts_.assign(3 * (size_t)n, 0.0);
    std::vector<double> dir(3 * (size_t)n);
    std::vector<float> pf(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        pts_[3 * i] = s.x[i]; pts_[3 * i + 1] = s.y[i]; pts_[3 * i + 2] = d3d ? s.z[i] : 0.0;
        dir[3 * i] = s.dirX[i]; dir[3 * i + 1] = s.dirY[i]; dir[3 * i + 2] = d3d ? s.dirZ[i] : 0.0;
        for (int a = 0; a < 3; ++a) pf[3 * i + a] = (float)pts_[3 * i + a];
    }
    dirs_ = dir;
    if ((*err = pd_.upload(pts_.data(), pts_.size()))) return;
    if ((*err = pf_.upload(pf.data(), pf.size()))) return;
    v_.pf = pf_.p;
    v_.pd = pd_.p;
    GaussBeams& g = v_.g;
    g.gammaA = 1.0 / (s.aperture * s.aperture);
    // the aperture's vMF normalisation (GaussianBeam.cuh:1058-1064; 2D through I_0)
    g.apNorm = d3d ? vmfNormLog3(g.gammaA) : vmfNormLog2(g.gammaA);
    g.waveK = 2.0 * PI / lambda;
    g.focal = pd_.p;
    g.meanView = mkd(0.0, 0.0, 1.0);
    // views, and 2D illuminations (whose points are drawn as from point sources at the focal
    // points, GaussianBeam.cuh:1667-1683), need no sampling tables
    if (!illumination || !d3d) {
        if ((*err = gd_.upload(dir.data(), dir.size()))) return;
        g.dir = gd_.p;
        return;
    }
    // mean view beam direction (GaussianBeam.cuh:1176-1194); the z axis when it vanishes
    if (views && views->num > 0 && views->dirX && views->dirY && views->dirZ) {
        double mv[3] = {0, 0, 0};
        for (uint32_t i = 0; i < views->num; ++i) {
            const double w = 1.0 / (double)views->num;
            mv[0] = mv[0] + w * views->dirX[i];
            mv[1] = mv[1] + w * views->dirY[i];
            mv[2] = mv[2] + w * views->dirZ[i];
        }
        const double r = std::sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
        if (!(r < 0.0001)) g.meanView = mkd(mv[0] * (1.0 / r), mv[1] * (1.0 / r), mv[2] * (1.0 / r));
    }
    // updateSampler (GaussianBeam.cuh:1281-1613): zB = 1000 points along each beam across the box
    // (initSamplingTable, fillTimesInSamplingTable, :177-224), the projection of the box corners on
    // the beam giving its extent
    const uint32_t B = 1000;
    g.B = B;
    const MediumView& mv = medium.view();
    std::vector<double> zmin(n), zmax(n), zdt(n), xyb(6 * (size_t)n);
    std::vector<double> pts(3 * (size_t)n * (B + 1)), wd(3 * (size_t)n * (B + 1));
    for (uint32_t l = 0; l < n; ++l) {
        const double* f = &pts_[3 * l];
        const double* d = &dir[3 * l];
        double lo = 0, hi = 0;
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {(c & 1) ? mv.bmaxd[0] : mv.bmind[0], (c & 2) ? mv.bmaxd[1] : mv.bmind[1],
                                 (c & 4) ? mv.bmaxd[2] : mv.bmind[2]};
            double o[3];
            for (int a = 0; a < 3; ++a) {
                double acc = 0;
                for (int b = 0; b < 3; ++b) acc += d[a] * d[b] * (p[b] - f[b]);
                o[a] = acc;
            }
            const double t = o[0] * d[0] + o[1] * d[1] + o[2] * d[2];
            if (c == 0 || t < lo) lo = t;
            if (c == 0 || t > hi) hi = t;
        }
        zmin[l] = lo;
        zmax[l] = hi;
        zdt[l] = (hi - lo) / (double)B;
        // orthogonalBase (VectorType.cuh:293-311)
        double* b0 = &xyb[6 * l];
        double* b1 = &xyb[6 * l + 3];
        if (std::fabs(1.0 - std::fabs(d[2])) <= 1e-12) {
            b0[0] = 1; b0[1] = 0; b0[2] = 0;
            b1[0] = 0; b1[1] = 1; b1[2] = 0;
        } else {
            double a0 = -d[1], a1 = d[0];
            const double nn = std::sqrt(a0 * a0 + a1 * a1);
            a0 = a0 / nn;
            a1 = a1 / nn;
            b0[0] = a0; b0[1] = a1; b0[2] = 0.0;
            b1[0] = a1 * d[2] - 0.0 * d[1];
            b1[1] = 0.0 * d[0] - a0 * d[2];
            b1[2] = a0 * d[1] - a1 * d[0];
        }
        // point 0: the focal point (its material); points 1..B: the z samples
        const size_t o = (size_t)l * (B + 1);
        for (int a = 0; a < 3; ++a) { pts[3 * o + a] = f[a]; wd[3 * o + a] = d[a]; }
        for (uint32_t b = 0; b < B; ++b) {
            const double t = zmin[l] + (double)b * zdt[l];
            for (int a = 0; a < 3; ++a) {
                pts[3 * (o + 1 + b) + a] = f[a] + t * d[a];
                wd[3 * (o + 1 + b) + a] = d[a];
            }
        }
    }
    const uint32_t np = n * (B + 1);
    DevBuf<double> dp, dw, datt;
    DevBuf<uint32_t> dmat;
    if ((*err = dp.upload(pts.data(), pts.size())) || (*err = dw.upload(wd.data(), wd.size())) ||
        (*err = datt.alloc(np)) || (*err = dmat.alloc(np)))
        return;
    std::vector<double> att(np);
    std::vector<uint32_t> mat(np);
    if (launchGaussSetup(mv, dp.p, dw.p, np, datt.p, dmat.p, nullptr) != hipSuccess ||
        hipMemcpy(att.data(), datt.p, np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(mat.data(), dmat.p, np * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
        *err = SSTMC_DEVICE_ERROR;
        return;
    }
    // per beam: the focal material's mixture weights exp(C + log-normalisation(|mu|))
    // (alphaPdfIcdfKernel, :254-295, with its large-kappa term log(2 pi) * kappa kept as written)
    // and per mixture the z density attenuation / w(z)^2 (samplingPdfKernel, :297-324),
    // exclusive-scanned and normalised, the density scaled by B - 1 (:336-355)
    std::vector<uint32_t> focalMat(n), mixOff(n);
    std::vector<double> apdf, aicdf, zpdf, zicdf;
    const double k = g.waveK;
    for (uint32_t l = 0; l < n; ++l) {
        const size_t o = (size_t)l * (B + 1);
        const uint32_t fm = mat[o];
        const double* kap = nullptr;
        const double* cc = nullptr;
        const uint32_t M = fm ? scattering.vmfLobes(fm, &kap, &cc) : 0;
        if (M == 0) { *err = SSTMC_INVALID_CONFIG; return; }
        if (M > kGaussLobes) { *err = SSTMC_NOT_SUPPORTED; return; }
        focalMat[l] = fm;
        mixOff[l] = (uint32_t)apdf.size();
        const size_t a0 = apdf.size();
        double asum = 0;
        for (uint32_t i = 0; i < M; ++i) {
            const double ka = std::fabs(kap[i]);
            const double ln = ka < 1e-5 ? -std::log(4.0 * PI)
                                        : std::log(ka) - (ka > 10.0 ? std::log(2.0 * PI) * ka
                                                                    : std::log(4.0 * PI * std::sinh(ka)));
            apdf.push_back(std::exp(cc[i] + ln));
            asum += apdf.back();
        }
        double acc = 0;
        for (uint32_t i = 0; i < M; ++i) {
            apdf[a0 + i] /= asum;
            aicdf.push_back(acc);
            acc += apdf[a0 + i];
        }
        const size_t z0 = zpdf.size();
        zpdf.resize(z0 + (size_t)M * B);
        zicdf.resize(z0 + (size_t)M * B);
        const double* f = &pts_[3 * l];
        const double* d = &dir[3 * l];
        for (uint32_t b = 0; b < B; ++b) {
            const double* z = &pts[3 * (o + 1 + b)];
            const double dz = (z[0] - f[0]) * d[0] + (z[1] - f[1]) * d[1] + (z[2] - f[2]) * d[2];
            for (uint32_t i = 0; i < M; ++i) {
                const double wz2 = (g.gammaA + std::fabs(kap[i])) / (k * k) + dz * dz / g.gammaA;
                zpdf[z0 + (size_t)i * B + b] = att[o + 1 + b] / wz2;
            }
        }
        for (uint32_t i = 0; i < M; ++i) {
            double* zp = &zpdf[z0 + (size_t)i * B];
            double* zi = &zicdf[z0 + (size_t)i * B];
            double run = 0;
            for (uint32_t b = 0; b < B; ++b) { zi[b] = run; run += zp[b]; }
            const double sum = zp[B - 1] + zi[B - 1];
            for (uint32_t b = 0; b < B; ++b) {
                zp[b] *= (double)(B - 1) / sum;
                zi[b] /= sum;
            }
        }
    }
    for (uint32_t m = 0; m < medium.materialNum() && m < (uint32_t)kMaxMaterials; ++m)
        g.prDen[m] = medium.sigs(m) * (mv.bmaxd[2] - mv.bmind[2]);
    // one FP64 allocation: dir, zmin, zmax, zdt, xyb, apdf, aicdf, zpdf, zicdf; one for the ids
    std::vector<double> all;
    std::vector<size_t> at;
    for (const std::vector<double>* v : {&dir, &zmin, &zmax, &zdt, &xyb, &apdf, &aicdf, &zpdf, &zicdf}) {
        at.push_back(all.size());
        all.insert(all.end(), v->begin(), v->end());
    }
    std::vector<uint32_t> ids(focalMat);
    ids.insert(ids.end(), mixOff.begin(), mixOff.end());
    if ((*err = gd_.upload(all.data(), all.size())) || (*err = gu_.upload(ids.data(), ids.size()))) return;
    g.dir = gd_.p + at[0];
    g.zmin = gd_.p + at[1];
    g.zmax = gd_.p + at[2];
    g.zdt = gd_.p + at[3];
    g.xyb = gd_.p + at[4];
    g.apdf = gd_.p + at[5];
    g.aicdf = gd_.p + at[6];
    g.zpdf = gd_.p + at[7];
    g.zicdf = gd_.p + at[8];
    g.focalMat = gu_.p;
    g.mixOff = gu_.p + n;
}

// ====================================================================== TemporalSampler
TemporalSampler::TemporalSampler(int* err, const double* t, uint32_t nt, bool isTemporalAmplitude,
                                 bool correlation) {
    *err = SSTMC_NO_ERROR;
    v_.nt = nt;
    v_.temporalAmp = isTemporalAmplitude ? 1 : 0;
    D_.assign(kMaxMaterials, 0.0);
    U_.assign(3 * kMaxMaterials, 0.0);
    if (correlation) {
        v_.analytic = 1;
        v_.trivial = 1;
        if ((*err = ticksDev_.upload(t, nt))) return;
        v_.ticks = ticksDev_.p;
        return;
    }
    // time table of the TemporalSampler constructor (TemporalSampler.cuh:941-1132): frames below
    // the mean form the negative chain, the rest the positive chain, each ordered by distance to
    // the mean; a slot's step is its time minus the previous slot's (or minus the mean).
    double mean = 0;
    for (uint32_t i = 0; i < nt; ++i) mean += t[i];
    mean /= nt;
    std::vector<int> used(nt, 0);
    std::vector<uint32_t> order;
    uint32_t neg = 0;
    for (uint32_t i = 0; i < nt; ++i) neg += (t[i] - mean < 0) ? 1u : 0u;
    for (int pass = 0; pass < 2; ++pass) {
        bool negative = pass == 0;
        uint32_t cnt = negative ? neg : nt - neg;
        for (uint32_t c = 0; c < cnt; ++c) {
            bool first = true;
            double best = 0;
            uint32_t bi = 0;
            for (uint32_t i = 0; i < nt; ++i) {
                if (used[i] || ((t[i] - mean < 0) != negative)) continue;
                double dd = negative ? mean - t[i] : t[i] - mean;
                if (first || dd < best) { first = false; best = dd; bi = i; }
            }
            used[bi] = 1;
            order.push_back(bi);
        }
    }
    v_.neg = neg;
    dt_.assign(nt, 0.0);
    int prev = -1;
    for (uint32_t k = 0; k < nt; ++k) {
        if (k == 0 || k == neg) prev = -1;
        int tick = (int)order[k];
        dt_[k] = prev < 0 ? t[tick] - mean : t[tick] - t[prev];
        prev = tick;
    }
    std::vector<float> f(nt);
    for (uint32_t k = 0; k < nt; ++k) f[k] = (float)dt_[k];
    if ((*err = dtDev_.upload(f.data(), nt))) return;
    v_.dt = dtDev_.p;
}

int TemporalSampler::setMaterial(double D, double ux, double uy, double uz, uint32_t m) {
    if (m >= (uint32_t)kMaxMaterials || D < 0) return SSTMC_INVALID_CONFIG;
    D_[m] = D;
    U_[3 * m] = ux; U_[3 * m + 1] = uy; U_[3 * m + 2] = uz;
    v_.D[m] = (float)D;
    v_.U[m][0] = (float)ux; v_.U[m][1] = (float)uy; v_.U[m][2] = (float)uz;
    v_.Dd[m] = D;
    v_.Ud[m][0] = ux; v_.Ud[m][1] = uy; v_.Ud[m][2] = uz;
    return SSTMC_NO_ERROR;
}

bool TemporalSampler::moving(uint32_t materialNum) const {
    for (uint32_t m = 1; m < materialNum && m < (uint32_t)kMaxMaterials; ++m)
        if (U_[3 * m] != 0.0 || U_[3 * m + 1] != 0.0 || U_[3 * m + 2] != 0.0) return true;
    return false;
}

void TemporalSampler::finalize(uint32_t materialNum) {
    if (v_.analytic) return;   // the correlation sampler never displaces a vertex
    bool trivial = true;
    for (uint32_t k = 0; k < v_.nt; ++k)
        for (uint32_t m = 1; m < materialNum; ++m) {
            if (dt_[k] != 0.0 && (D_[m] != 0.0 || U_[3 * m] != 0.0 || U_[3 * m + 1] != 0.0 || U_[3 * m + 2] != 0.0))
                trivial = false;
        }
    v_.trivial = trivial ? 1 : 0;
}

// ====================================================================== NextEventEstimation
NextEventEstimation::NextEventEstimation(int* err, const Medium& medium, const Scattering& scattering,
                                         const Source& illumination, const Source& view,
                                         const TemporalSampler& sampler, const TracerOptions& opt)
    : medium_(medium), sc_(scattering), il_(illumination), vw_(view), sampler_(sampler), opt_(opt) {
    seed_ = opt.seed;
    dims_ = medium.dims();
    // the correlation sampler's sources hold the u1 sources then the u2 ones
    const bool acorr = sampler_.correlation();
    L_ = acorr ? il_.size() / 2 : il_.size();
    V_ = acorr ? vw_.size() / 2 : vw_.size();
    T_ = sampler_.frames();
    P_ = opt.batchSize;
    fp64_ = opt.terms == SSTMC_TERMS_FP64;
    contractKind_ = fp64_ ? kContractFp64
                          : (opt.contraction == SSTMC_CONTRACT_FP32 || opt.contraction == SSTMC_CONTRACT_FP64
                                 ? kContractFp32 : kContractBf16x3);
    // single scattering: separable when f does not depend on the first vertex
    // (threePointFunctionSingle, FarField.cuh:546-599 / PointSource.cuh:363-446)
    uint32_t nm = medium.materialNum();
    double iso = dims_ == 3 ? 0.282094791773878143 : 0.398942280401432678;
    singleMode_ = kSingleGeneric;
    bool needF = false;
    if (il_.type() == SSTMC_SOURCE_GAUSSIAN_BEAM || vw_.type() == SSTMC_SOURCE_GAUSSIAN_BEAM) {
        // both roles GaussianBeam (fsGaussianBeam pairs two beams), the TemporalSampler without
        // coherent backscattering
        if (il_.type() != vw_.type() || acorr || opt.cbs) {
            *err = SSTMC_NOT_SUPPORTED;
            return;
        }
        singleMode_ = kSingleGauss;
    } else if (il_.type() == SSTMC_SOURCE_POINT && sc_.isIsotropicScattering()) {
        singleMode_ = kSingleSeparable;
        fscalar_ = iso;
    } else if (il_.type() == SSTMC_SOURCE_FARFIELD && vw_.type() == SSTMC_SOURCE_FARFIELD && sc_.isUniform(nm)) {
        singleMode_ = kSingleSeparable;
        needF = true;
    } else if (sc_.isExactlyIsotropic(nm)) {
        singleMode_ = kSingleSeparable;
        fscalar_ = iso;
    }
    if (acorr) {
        // TemporalCorrelationSampler single scattering (NextEventEstimation.cuh:1016-1136): with a
        // scalar Fs (point illumination and view, isotropic scattering: Sampler::isScalarFs,
        // SamplerInterface.cuh:155, PointSource.cuh:62) it is Fs^2 a^T b; otherwise each (l, v, path)
        // takes Fs_u1 conj(Fs_u2) and the decorrelation at x1
        needF = false;
        fscalar_ = 1.0;
        singleMode_ = kSingleCorrelation;
        if (il_.type() == SSTMC_SOURCE_POINT && vw_.type() == SSTMC_SOURCE_POINT && sc_.isIsotropicScattering()) {
            singleMode_ = kSingleSeparable;
            fscalar_ = iso * iso;
        }
        const std::vector<double>& hv = vw_.hostPoints();
        sameViews_ = std::equal(hv.begin(), hv.begin() + 3 * (size_t)V_, hv.begin() + 3 * (size_t)V_);
        const std::vector<double>& hl = il_.hostPoints();
        sameIllums_ = std::equal(hl.begin(), hl.begin() + 3 * (size_t)L_, hl.begin() + 3 * (size_t)L_);
        // coherent backscattering: CBS01-CBS04 (NextEventEstimation.cuh:1138-1391)
        cbsCorr_ = opt.cbs != 0;
    } else if (opt_.output == SSTMC_OUTPUT_CORRELATION) {
        // S(0) conj S(t): a scalar or fixed F enters squared, a per-path f as f(0) conj f(t)
        if (singleMode_ != kSingleSeparable && singleMode_ != kSingleGeneric && singleMode_ != kSingleGauss) {
            *err = SSTMC_NOT_SUPPORTED;
            return;
        }
        fscalar_ *= fscalar_;
    }
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&streamT_, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&streamC_, hipStreamNonBlocking) != hipSuccess) {
        *err = SSTMC_DEVICE_ERROR;
        return;
    }
    for (BatchSet& bs : sets_)
        if (hipEventCreateWithFlags(&bs.transported, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&bs.traced, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&bs.released, hipEventDisableTiming) != hipSuccess) {
            *err = SSTMC_DEVICE_ERROR;
            return;
        }
    if ((*err = buildWalkGrid())) return;
    // the far-field view kernel (k_nee_view_far): 3D, FP32 terms, HG or isotropic materials,
    // equally spaced grid, field output of the TemporalSampler, far-field views (or far-field
    // illuminations for the reversed paths) of more than 32 sources, one wavefront each
    farRecs_ = dims_ == 3 && !fp64_ && !acorr && opt_.output == SSTMC_OUTPUT_FIELD && !sc_.view().tab &&
               mview_.walkFast && singleMode_ != kSingleGauss &&
               ((vw_.type() == SSTMC_SOURCE_FARFIELD && V_ > 32) ||
                (opt.cbs && il_.type() == SSTMC_SOURCE_FARFIELD && L_ > 32));
    if ((*err = setupFarBound())) return;
    viewFirst_ = farRecs_ && vw_.type() == SSTMC_SOURCE_FARFIELD && V_ > 32 && !opt.cbs &&
                 contractKind_ == kContractBf16x3 && singleMode_ == kSingleSeparable;
    nsets_ = pipelineSets(setBytes(P_, L_, V_, T_, fp64_, !sampler_.view().trivial, opt.cbs != 0, acorr));
    if ((*err = allocate())) return;
    if (needF) {
        std::vector<float2> ff((size_t)L_ * V_);
        std::vector<double2> fd((size_t)L_ * V_);
        const std::vector<double>& dl = il_.hostPoints();
        const std::vector<double>& dv = vw_.hostPoints();
        for (uint32_t v = 0; v < V_; ++v)
            for (uint32_t l = 0; l < L_; ++l) {
                double c = dl[3 * l] * dv[3 * v] + dl[3 * l + 1] * dv[3 * v + 1] + dl[3 * l + 2] * dv[3 * v + 2];
                double fr, fi;
                sc_.amplitude(1, c, &fr, &fi);
                fd[(size_t)v * L_ + l] = make_double2(fr, fi);
                ff[(size_t)v * L_ + l] = make_float2((float)fr, (float)fi);
            }
        if ((*err = Ff_.upload(ff.data(), ff.size()))) return;
        if ((*err = Fd_.upload(fd.data(), fd.size()))) return;
    }
    *err = SSTMC_NO_ERROR;
}

NextEventEstimation::~NextEventEstimation() {
    syncAll();
    dropGraphs();
    for (hipEvent_t e : ev_) hipEventDestroy(e);
    for (BatchSet& bs : sets_)
        for (hipEvent_t e : {bs.transported, bs.traced, bs.released})
            if (e) hipEventDestroy(e);
    if (cleared_) hipEventDestroy(cleared_);
    if (wsMulti_.ptr) hipFree(wsMulti_.ptr);
    if (wsSingle_.ptr) hipFree(wsSingle_.ptr);
    for (hipStream_t st : {stream_, streamT_, streamC_})
        if (st) hipStreamDestroy(st);
}

int NextEventEstimation::syncAll() const {
    int rc = SSTMC_NO_ERROR;
    for (hipStream_t st : {streamT_, stream_, streamC_})
        if (st && hipStreamSynchronize(st) != hipSuccess) rc = SSTMC_KERNEL_ERROR;
    return rc;
}

int NextEventEstimation::allocate() {
    int e;
    if ((e = status_.alloc(4)) || (e = totals_.alloc(kTotals)) || (e = overflow_.alloc((opt_.fullIteration + 31) / 32)) ||
        (e = stage_.q.alloc(2)) || (e = stepSlots_.alloc(2 * (size_t)kCountSlots * kCountStride)))
        return e;
    size_t elem = fp64_ ? 16 : 8;
    size_t na = (size_t)T_ * P_ * L_ * elem, nb = (size_t)T_ * P_ * V_ * elem;
    for (int s = 0; s < nsets_; ++s) {
        BatchSet& bs = sets_[s];
        if ((viewFirst_ && (e = bs.rowNz.alloc(P_))) || (e = bs.k.alloc(P_)) || (e = bs.misses.alloc(P_)) ||
            (e = bs.off.alloc(P_ + 1)) ||
            (e = bs.blockSums.alloc((P_ + 1023) / 1024 + 1)) || (e = bs.tauRef.alloc(1)))
            return e;
        if ((e = bs.A.alloc(na)) || (e = bs.a.alloc(na)) || (e = bs.B.alloc(nb)) || (e = bs.b.alloc(nb))) return e;
        if (opt_.cbs && ((e = bs.Ar.alloc(nb)) || (e = bs.Br.alloc(na)))) return e;
        bytes_ += (opt_.cbs ? 3 : 2) * (na + nb) + 20 * (size_t)P_;
        if (cbsCorr_) {
            if ((e = bs.F1l.alloc(na)) || (e = bs.F2l.alloc(na)) || (e = bs.F1v.alloc(nb)) || (e = bs.F2v.alloc(nb)))
                return e;
            bytes_ += 2 * (na + nb);
        }
    }
    if (cbsCorr_) {
        const size_t n = (size_t)cbsCrossSplits(L_, V_, T_, P_) * T_ * L_ * V_;
        if ((e = cbsPart_.alloc(n))) return e;
        bytes_ += 16 * n;
    }
    if ((e = U_.alloc(2 * (size_t)T_ * L_ * V_))) return e;
    bytes_ += 16 * (size_t)T_ * L_ * V_;
    return growPool((uint64_t)P_ * 32);
}

VertexBatch NextEventEstimation::vertexBatch(int set) const {
    const BatchSet& bs = sets_[set];
    VertexBatch vb{bs.k.p, bs.misses.p, bs.off.p, bs.xm.p, bs.rho.p, bs.delta.p, cap_, status_.p, bs.rec.p,
                   farRecs_ ? bs.far.p : nullptr, farRecs_ ? bs.live.p : nullptr};
    vb.rowNz = viewFirst_ ? bs.rowNz.p : nullptr;
    return vb;
}

// Every set's vertex pool holds `cap` vertices; the streams must be idle.
int NextEventEstimation::growPool(uint64_t cap) {
    int e;
    dropGraphs();   // they name the pools being replaced
    // vertex records per frame; the correlation sampler's vertices do not move: one record each
    const uint64_t recT = sampler_.correlation() ? 1 : T_;
    for (int s = 0; s < nsets_; ++s) {
        BatchSet& bs = sets_[s];
        if ((e = bs.xm.alloc(cap)) || (e = bs.rho.alloc(cap)) || (e = bs.rec.alloc(cap * recT))) return e;
        if (farRecs_ && ((e = bs.far.alloc(cap * recT)) || (e = bs.live.alloc(2 * cap * recT)))) return e;
        if (!sampler_.view().trivial) {
            if ((e = bs.delta.alloc(cap * T_ * 3))) return e;
        } else {
            bs.delta.release();
        }
    }
    uint64_t perVertex = 20 + (farRecs_ ? 136 : 32) * recT + (sampler_.view().trivial ? 0 : 12 * (uint64_t)T_);
    if (cbsCorr_) {
        // per-vertex u1 / u2 terms of every illumination and view, per lag
        const size_t elem = fp64_ ? 16 : 8;
        const size_t nl = (size_t)T_ * cap * L_ * elem, nv = (size_t)T_ * cap * V_ * elem;
        for (int s = 0; s < nsets_; ++s)
            if (BatchSet& bs = sets_[s]; (e = bs.Gl1.alloc(nl)) || (e = bs.Gl2.alloc(nl)) || (e = bs.Gv1.alloc(nv)) || (e = bs.Gv2.alloc(nv)))
                return e;
        perVertex += 2 * (uint64_t)T_ * (L_ + V_) * elem;
    }
    bytes_ += (uint64_t)nsets_ * (cap - cap_) * perVertex;
    // the transport's staging: the pool plus the chunk tails of every path (TransportStage)
    const uint64_t stageNeed = cap + (uint64_t)P_ * (kStageChunk - 1);
    const uint64_t nchunks = (stageNeed + kStageChunk - 1) / kStageChunk + 1;
    if (nchunks > 0xffffffffull / kStageChunk) return SSTMC_ALLOCATION_ERROR;
    if (nchunks > stage_.nchunks) {
        const uint64_t nv = nchunks * kStageChunk;
        if ((e = stage_.xm.alloc(nv)) || (e = stage_.rho.alloc(nv)) || (e = stage_.chunkSlot.alloc(nchunks)) ||
            (e = stage_.chunkJ0.alloc(nchunks)))
            return e;
        if (!sampler_.view().trivial) {
            if ((e = stage_.delta.alloc(nv * T_ * 3))) return e;
        }
        const uint64_t perStaged = 20 + (sampler_.view().trivial ? 0 : 12 * (uint64_t)T_);
        bytes_ += (nchunks - stage_.nchunks) * (kStageChunk * perStaged + 8);
        stage_.nchunks = (uint32_t)nchunks;
    }
    cap_ = cap;
    return SSTMC_NO_ERROR;
}

TransportStage NextEventEstimation::transportStage() const {
    return TransportStage{stage_.xm.p, stage_.rho.p, stage_.delta.p, stage_.chunkSlot.p, stage_.chunkJ0.p,
                          stage_.nchunks, stage_.q.p};
}

int NextEventEstimation::contractSingle(uint32_t npaths, double* U, int set, hipStream_t st) {
    FramesView fr = sampler_.view();
    VertexBatch vb = vertexBatch(set);
    const BatchSet& bs = sets_[set];
    hipError_t he;
    if (singleMode_ == kSingleSeparable) {
        he = launchContract(contractKind_, bs.a.p, bs.b.p, npaths, L_, V_, T_, Ff_.p, Fd_.p,
                            opt_.output == SSTMC_OUTPUT_CORRELATION ? 1 : 0, fscalar_, U, &wsSingle_, st);
    } else if (singleMode_ == kSingleGauss) {
        const std::vector<double>& vd = static_cast<const GaussianBeam&>(vw_).directions();
        bool uniformDir = true;
        for (size_t i = 3; i < vd.size() && uniformDir; ++i) uniformDir = vd[i] == vd[i % 3];
        he = launchSingleGauss(dims_, fp64_ ? 1 : 0, uniformDir ? 1 : 0, opt_.output == SSTMC_OUTPUT_CORRELATION ? 1 : 0,
                               mview_, sc_.view(), il_.view(), vw_.view(), fr, vb, bs.a.p, bs.b.p, npaths, U,
                               &wsSingle_, st);
    } else if (singleMode_ == kSingleCorrelation) {
        const bool uniformFar = il_.type() == SSTMC_SOURCE_FARFIELD && vw_.type() == SSTMC_SOURCE_FARFIELD &&
                                sc_.isUniform(medium_.materialNum());
        he = launchSingleCorrelation(dims_, fp64_ ? 1 : 0, sc_.view(), il_.view(), vw_.view(), fr, vb, bs.a.p, bs.b.p,
                                     npaths, opt_.lambda, uniformFar ? 1 : 0, U, &wsSingle_, st);
    } else {
        he = launchSingleGeneric(dims_, fp64_ ? 1 : 0, sc_.view(), il_.view(), vw_.view(), fr, vb, bs.a.p, bs.b.p,
                                 npaths, opt_.output == SSTMC_OUTPUT_CORRELATION ? 1 : 0, U, &wsSingle_, st);
    }
    return he == hipSuccess ? SSTMC_NO_ERROR : SSTMC_KERNEL_ERROR;
}

// One batch in three stages: transport (free flight, vertex offsets, totals, then the cut
// reference, vertex records and walk bounds), NEE (first-vertex terms, view terms) and
// contraction (multiple then single scattering into U). Pipelined, each stage runs on its own stream: the transport of a set waits
// until the set's previous batch has been contracted, the NEE waits for its transport, the
// contraction for its NEE; batches take the sets in turn, and the contractions keep
// batch order on their one stream, so U is summed in the same order as without the pipeline.
int NextEventEstimation::runBatch(uint64_t pathBegin, uint32_t npaths, double* U, bool timed, uint32_t b, int set,
                                  bool pipelined) {
    const MediumView& m = mview_;
    const ScatterView& sc = sc_.view();
    const SourceView& il = il_.view();
    const SourceView& vw = vw_.view();
    const FramesView& fr = sampler_.view();
    BatchSet& bs = sets_[set];
    VertexBatch vb = vertexBatch(set);
    const hipStream_t sT = pipelined ? streamT_ : stream_, sN = stream_, sC = pipelined ? streamC_ : stream_;
    hipEvent_t* e = timed ? &ev_[(size_t)b * kEventsPerBatch] : nullptr;
    lastSet_ = set;

    NeeParams prm;
    prm.lambda = opt_.lambda;
    prm.output = opt_.output;
    prm.singleMode = singleMode_;
    // walks are exact by default (FP32: exp(-105) rounds to zero, so walks stop at 210); with
    // precision.tauCut the view-side terms take the batch-relative cut (k_tau_ref)
    prm.exactCutF = 210.0f;
    prm.tauRef = opt_.tauCut > 0 ? bs.tauRef.p : nullptr;
    prm.tauCutRel = (float)opt_.tauCut;
    prm.npaths = npaths;
    prm.viewFirst = viewFirst_ && pipelined ? 1u : 0u;   // debug batches form every term
    prm.sameViews = sameViews_ ? 1u : 0u;
    prm.sameIllums = sameIllums_ ? 1u : 0u;
    NeeOut out{bs.A.p, bs.B.p, bs.a.p, bs.b.p, bs.Ar.p, bs.Br.p, stepSlots_.p};
    if (cbsCorr_) {
        out.F1l = bs.F1l.p;
        out.F2l = bs.F2l.p;
        out.F1v = bs.F1v.p;
        out.F2v = bs.F2v.p;
        out.Gl1 = bs.Gl1.p;
        out.Gl2 = bs.Gl2.p;
        out.Gv1 = bs.Gv1.p;
        out.Gv2 = bs.Gv2.p;
        out.gcap = cap_;
    }
    const FarBound* fb = farBoundOn_ ? &farBound_ : nullptr;

    // ---- transport, then what NEE needs of the vertices alone (the cut's reference and the
    // per-vertex records), so that the NEE stream bounds the walks and forms the terms
    if (bs.used && hipStreamWaitEvent(sT, bs.released, 0) != hipSuccess) return SSTMC_DEVICE_ERROR;
    bs.used = true;
    if (e) hipEventRecord(e[0], sT);
    const TransportStage ts = transportStage();
    if (launchTransport(dims_, m, sc, il, fr, vb, ts, seed_, pathBegin, npaths, sT) != hipSuccess ||
        launchScan(bs.k.p, bs.off.p, bs.blockSums.p, npaths, sT) != hipSuccess ||
        launchTransportPack(vb, ts, npaths, fr.nt, fr.trivial ? 0 : 1, sT) != hipSuccess ||
        launchBatchTotals(bs.misses.p, bs.off.p, npaths, totals_.p, cap_, overflow_.p, b, bs.tauRef.p, ts, status_.p,
                          sT) != hipSuccess ||
        launchNee(dims_, fp64_ ? 1 : 0, m, sc, il, vw, fr, vb, prm, out, sT, nullptr, fb, kNeeVertexPart) != hipSuccess)
        return SSTMC_KERNEL_ERROR;
    if (e) hipEventRecord(e[1], sT);
    if (hipEventRecord(bs.transported, sT) != hipSuccess) return SSTMC_DEVICE_ERROR;

    // ---- NEE terms
    if (pipelined && hipStreamWaitEvent(sN, bs.transported, 0) != hipSuccess) return SSTMC_DEVICE_ERROR;
    if (e) hipEventRecord(e[6], sN);
    if (launchNee(dims_, fp64_ ? 1 : 0, m, sc, il, vw, fr, vb, prm, out, sN, e ? e[5] : nullptr, fb, kNeeTermPart) !=
        hipSuccess)
        return SSTMC_KERNEL_ERROR;
    if (e) hipEventRecord(e[2], sN);
    if (hipEventRecord(bs.traced, sN) != hipSuccess) return SSTMC_DEVICE_ERROR;

    // ---- contraction: multiple scattering U += sum_p A^T B; with coherent backscattering the
    // reversed paths add sum_p Br^T Ar and both are scaled by 1/sqrt(2)
    // (NextEventEstimation.cuh:1385-1391,1448-1454); under the correlation sampler CBS01 and CBS04
    // are these two products, the mixed pairings CBS02 + CBS03 follow, and all four take 1/2;
    // then single scattering
    if (pipelined && hipStreamWaitEvent(sC, bs.traced, 0) != hipSuccess) return SSTMC_DEVICE_ERROR;
    if (e) hipEventRecord(e[7], sC);
    int rc = contractStage(npaths, U, set, sC, e ? e[3] : nullptr);
    if (rc) return rc;
    if (e) hipEventRecord(e[4], sC);
    if (hipEventRecord(bs.released, sC) != hipSuccess) return SSTMC_DEVICE_ERROR;
    return SSTMC_NO_ERROR;
}

// The multiple-scattering contraction of a batch: U += msScale (sum_p A^T B [+ sum_p Br^T Ar]),
// and under the correlation sampler with coherent backscattering the mixed pairings.
int NextEventEstimation::contractMulti(uint32_t npaths, double* U, int set, hipStream_t sC) {
    const BatchSet& bs = sets_[set];
    const double msScale = opt_.cbs ? (cbsCorr_ ? 0.5 : 0.70710678118654752440) : 1.0;
    if (launchContract(contractKind_, bs.A.p, bs.B.p, npaths, L_, V_, T_, nullptr, nullptr, 0, msScale, U, &wsMulti_,
                       sC) != hipSuccess)
        return SSTMC_KERNEL_ERROR;
    if (opt_.cbs && launchContract(contractKind_, bs.Br.p, bs.Ar.p, npaths, L_, V_, T_, nullptr, nullptr, 0, msScale,
                                   U, &wsMulti_, sC) != hipSuccess)
        return SSTMC_KERNEL_ERROR;
    if (cbsCorr_) {
        const CbsCross cc{bs.F1l.p, bs.F2l.p, bs.F1v.p, bs.F2v.p, bs.Gv1.p, bs.Gv2.p, bs.Gl1.p, bs.Gl2.p, cap_};
        if (launchCbsCross(dims_, fp64_ ? 1 : 0, il_.view(), vw_.view(), sampler_.view(), vertexBatch(set), cc, npaths,
                           opt_.lambda, U, cbsPart_.p, cbsCrossSplits(L_, V_, T_, P_), sC) != hipSuccess)
            return SSTMC_KERNEL_ERROR;
    }
    return SSTMC_NO_ERROR;
}

void NextEventEstimation::dropGraphs() {
    for (StageGraph& g : graphs_) {
        if (g.multi) hipGraphExecDestroy(g.multi);
        if (g.single) hipGraphExecDestroy(g.single);
    }
    graphs_.clear();
}

// Captures the launches `run` makes on st (no allocation, copy or synchronisation among them) as
// an executable graph; null when capture or instantiation fails.
template <typename F>
static hipGraphExec_t captureStage(hipStream_t st, F run) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    const int rc = run();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc || e != hipSuccess || !graph) {
        if (graph) hipGraphDestroy(graph);
        hipGetLastError();
        return nullptr;
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
    hipGraphDestroy(graph);
    hipGetLastError();
    return exec;
}

// The contraction stage of one batch on sC: multiple scattering, `mid` (if set) recorded, then
// single scattering. With opt.graphs the first batch of each (set, batch size, destination) runs
// eagerly, which sizes the workspaces, and is then captured; later batches of that key replay
// the two graphs: the same kernels with the same arguments in the same order, hence the same
// bits. A reallocated workspace or vertex pool drops the graphs.
int NextEventEstimation::contractStage(uint32_t npaths, double* U, int set, hipStream_t sC, hipEvent_t mid) {
    if (opt_.graphs) {
        for (size_t i = 0; i < graphs_.size(); ++i) {
            StageGraph& g = graphs_[i];
            if (g.set != set || g.npaths != npaths || g.U != U) continue;
            if (g.wsMulti != wsMulti_.ptr || g.wsSingle != wsSingle_.ptr) {
                dropGraphs();
                break;
            }
            if (hipGraphLaunch(g.multi, sC) != hipSuccess) return SSTMC_KERNEL_ERROR;
            if (mid) hipEventRecord(mid, sC);
            if (hipGraphLaunch(g.single, sC) != hipSuccess) return SSTMC_KERNEL_ERROR;
            ++graphReplays_;
            return SSTMC_NO_ERROR;
        }
    }
    int rc = contractMulti(npaths, U, set, sC);
    if (rc) return rc;
    if (mid) hipEventRecord(mid, sC);
    if ((rc = contractSingle(npaths, U, set, sC))) return rc;
    if (opt_.graphs) {
        StageGraph g;
        g.set = set;
        g.npaths = npaths;
        g.U = U;
        g.multi = captureStage(sC, [&] { return contractMulti(npaths, U, set, sC); });
        g.single = g.multi ? captureStage(sC, [&] { return contractSingle(npaths, U, set, sC); }) : nullptr;
        if (g.multi && g.single) {
            g.wsMulti = wsMulti_.ptr;
            g.wsSingle = wsSingle_.ptr;
            graphs_.push_back(g);
        } else {
            if (g.multi) hipGraphExecDestroy(g.multi);
            opt_.graphs = 0;   // capture unavailable here: stay eager
        }
    }
    return SSTMC_NO_ERROR;
}

uint64_t NextEventEstimation::setBytes(uint32_t P, uint32_t L, uint32_t V, uint32_t T, bool fp64, bool frames,
                                       bool cbs, bool acorr) {
    const uint64_t elem = fp64 ? 16 : 8;
    uint64_t terms = (cbs ? 3ull : 2ull) * T * P * (uint64_t)(L + V) * elem;
    // vertex records: 32 bytes, and 104 more for the far-field view kernel's (FP32 terms: record and masks)
    uint64_t pool = (uint64_t)P * 32 * (20 + (fp64 ? 32ull : 136ull) * T + (frames ? 12ull * T : 0));
    if (cbs && acorr) {
        // correlation CBS: first-vertex u1 / u2 terms and per-vertex u1 / u2 terms
        terms += 2ull * T * P * (uint64_t)(L + V) * elem;
        pool += (uint64_t)P * 32 * 2ull * T * (L + V) * elem;
    }
    return terms + pool + 24ull * P;
}

// Boxes of half a cell along each axis and pyramid queries of up to 5 x 5 blocks for renderings of
// 2^21 paths and up (128^3 cells, 16 blocks of views: about 0.2 s to build, and as tight as the
// bound at each point, farLowerBound), else whole cells and queries of up to 9 x 9 blocks (an
// eighth of the boxes, about 60 ms).
void NextEventEstimation::farTableResolution(uint64_t paths, uint32_t* sub, uint32_t* span) {
    const bool fine = paths >= (1ull << 21);
    *sub = fine ? 2u : 1u;
    *span = fine ? 4u : 8u;
}

uint64_t NextEventEstimation::footprint(uint32_t P, uint32_t L, uint32_t V, uint32_t T, bool fp64, bool frames,
                                        uint64_t cells, bool cbs, bool acorr, uint64_t paths) {
    const uint64_t elem = fp64 ? 16 : 8;
    // the pipeline's batch sets: terms, per-path arrays and vertex pools
    const uint64_t set = setBytes(P, L, V, T, fp64, frames, cbs, acorr);
    uint64_t field = 16ull * T * L * V;
    uint64_t ws = 2ull * 8 * T * L * V * elem;   // split-K partials, two workspaces
    if (!fp64) ws += 2ull * 8 * ((uint64_t)P + 64) * (L + V + 256);   // and their bf16 hi/lo planes
    uint64_t grids = 2 * cells;                  // material grid + packed walk grid (<= 4 bits a cell)
    // the far-field view kernel's bound tables: 4 bytes a box per block of 64 sources, at most 32
    if (!fp64) {
        uint32_t sub, span;
        farTableResolution(paths, &sub, &span);
        const uint64_t boxes = cells * sub * sub * sub;
        grids += 4 * boxes * std::min<uint64_t>(32, (V + 63) / 64);
        if (cbs) grids += 4 * boxes * std::min<uint64_t>(32, (L + 63) / 64);
    }
    if (cbs && acorr) field += 16ull * cbsCrossSplits(L, V, T, P) * T * L * V;   // mixed-pairing slices
    // the transport's staging: a pool of 32 vertices a path plus the chunk tails
    const uint64_t stage = (uint64_t)P * (32 + kStageChunk) * (20 + (frames ? 12ull * T : 0) + 1);
    return (uint64_t)pipelineSets(set) * set + field + ws + grids + stage + (1ull << 20);
}

// The NEE walk grid: per cell 4 x the index of its material in the palette of the materials
// present (the byte offset of its sigt in the kernels' LDS table), packed 1 bit a cell for one
// material, 4 bits for up to four and 8 bits for more, with the axes
// ordered so that the axis the view walks mostly run along is fastest. A walk then meets its next
// cell in the same 4-byte word for runs of up to 32 / bits cells along that axis and re-reads
// memory only when it leaves the word; the 64 walks of a wavefront, which fan out from one
// vertex, then issue few reads and touch few cache lines per step. Padding of one slowest-axis
// layer (rounded up to a word) on both sides keeps the cell one step past any face inside the
// allocation. Grids whose bit offsets do not fit 32 bits, or with unequal spacing, use the
// general walk. Built once per fast axis and kept by the medium, so tracers (requests) that share
// a medium share its walk grids.
int Medium::walkGrid(int fast, MediumView* mv) const {
    std::lock_guard<std::mutex> lk(walkMu_);
    WalkCache& wc = walk_[fast];
    if (!wc.built) {
        const int rc = buildWalk(fast, &wc);
        if (rc) return rc;
    }
    for (int i = 0; i < kWalkPalette; ++i) mv->walkSig[i] = wc.sig[i];
    mv->walkFast = wc.fast;
    mv->walkGrid = wc.buf.p;
    mv->walkWords = wc.words;
    mv->walkOrigin = wc.origin;
    mv->walkBits = wc.bits;
    for (int a = 0; a < 3; ++a) mv->walkStride[a] = wc.stride[a];
    return SSTMC_NO_ERROR;
}

int Medium::buildWalk(int fast, WalkCache* out) const {
    WalkCache& wc = *out;
    const int dims = v_.dims;
    const uint32_t* n = v_.n;
    int order[3] = {fast, -1, -1}, k = 1;
    for (int a = 0; a < 3; ++a)
        if (a != fast) order[k++] = a;
    uint32_t S[3] = {0, 0, 0}, mul = 1;
    for (int j = 0; j < 3; ++j) {
        S[order[j]] = order[j] < dims ? mul : 0u;
        if (order[j] < dims) mul *= n[order[j]];
    }
    // palette of the materials present, and the narrowest field that indexes it
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    const std::vector<uint8_t>& g = hostGrid();
    bool present[256] = {};
    for (size_t i = 0; i < ncell; ++i) present[g[i]] = true;
    uint32_t index[256] = {}, npal = 0;
    for (int i = 0; i < kWalkPalette; ++i) wc.sig[i] = 0.f;
    for (int id = 0; id < 256; ++id)
        if (present[id]) {
            if (id >= kMaxMaterials || npal >= (uint32_t)kWalkPalette) return SSTMC_INVALID_CONFIG;
            index[id] = npal;
            wc.sig[npal++] = v_.sigtf[id];
        }
    // fields hold 4 x palette index (the sigt byte offset): 0 fits one bit, 0..12 four, 0..60 eight
    const uint32_t bits = npal <= 1 ? 1u : (npal <= 4 ? 4u : 8u);
    // bit offsets: one layer of the slowest axis of padding (whole words) before and after
    const int slow = dims == 3 ? order[2] : order[1];
    const uint64_t pad = ((uint64_t)S[slow] * bits + 31) & ~31ull;
    const uint64_t total = (pad + (uint64_t)ncell * bits + pad + 31) & ~31ull;
    wc.fast = 0;
    if (v_.allUniform && total < 0xffffffffull) {
        std::vector<uint32_t> wg(total / 32, 0u);
        for (uint32_t i2 = 0; i2 < n[2]; ++i2)
            for (uint32_t i1 = 0; i1 < n[1]; ++i1)
                for (uint32_t i0 = 0; i0 < n[0]; ++i0) {
                    const uint64_t bit =
                        pad + ((uint64_t)i0 * S[0] + (uint64_t)i1 * S[1] + (uint64_t)i2 * S[2]) * bits;
                    wg[bit >> 5] |= (4u * index[g[i0 + (size_t)n[0] * (i1 + (size_t)n[1] * i2)]]) << (bit & 31);
                }
        int e = wc.buf.upload(wg.data(), wg.size());
        if (e) return e;
        wc.words = (uint32_t)(total / 32);
        wc.origin = (uint32_t)pad;
        wc.bits = bits;
        for (int a = 0; a < 3; ++a) wc.stride[a] = S[a] * bits;
        wc.fast = 1;
    }
    wc.built = true;
    return SSTMC_NO_ERROR;
}

uint64_t Medium::walkBytes() const {
    std::lock_guard<std::mutex> lk(walkMu_);
    uint64_t b = 0;
    for (const WalkCache& wc : walk_) b += wc.buf.n * sizeof(uint32_t);
    for (const PyrCache& pc : pyr_) b += pc.buf.n * sizeof(float);
    return b;
}

int Medium::minPyramid(int axis, MinPyramid* out) const {
    *out = MinPyramid{};
    if (v_.dims != 3 || !v_.allUniform || axis < 0 || axis > 2) return SSTMC_NO_ERROR;
    std::lock_guard<std::mutex> lk(walkMu_);
    PyrCache& pc = pyr_[axis];
    if (!pc.built) {
        MinPyramid mp;
        const std::vector<float> h = minPyramidHost(hostGrid().data(), v_.n, v_.sigtf, axis, &mp);
        if (!h.empty() && h.size() * sizeof(float) <= (1ull << 30)) {
            const int e = pc.buf.upload(h.data(), h.size());
            if (e) return e;
            mp.p = pc.buf.p;
            pc.view = mp;
        }
        pc.built = true;
    }
    *out = pc.view;
    return SSTMC_NO_ERROR;
}

// The far-field view kernel's lower bounds: per launch (views; reversed-path illuminations) the
// axis the walks all advance along, its sign and the lateral slopes over the walks' FP32
// directions (widened by a relative 1e-6), and the medium's pyramid along that axis. A set whose
// walks do not all advance along one axis by at least 0.1 of their length gets no bound. The
// kernel takes its sources in a Z-order of their lateral slopes (compact cones of 64); the bound of
// each block of 64 about the set's axis is tabulated per cell of the medium once, here.
int NextEventEstimation::setupFarBound() {
    farBound_ = FarBound{};
    farBoundOn_ = false;
    if (!farRecs_) return SSTMC_NO_ERROR;
    uint32_t sub, span;
    farTableResolution(paths(), &sub, &span);
    const uint64_t boxes = (uint64_t)mview_.n[0] * mview_.n[1] * mview_.n[2] * sub * sub * sub;
    for (int role = 0; role < 2; ++role) {
        const Source& src = role == 0 ? vw_ : il_;
        if (src.type() != SSTMC_SOURCE_FARFIELD || (role == 1 && !opt_.cbs)) continue;
        // the kernel's FP32 directions; walks along the view direction, against the illumination's
        const std::vector<double>& hp = src.hostPoints();
        std::vector<float> df(hp.size());
        for (size_t i = 0; i < hp.size(); ++i) df[i] = (float)hp[i];
        int axis = -1;
        const double sg = role == 0 ? 1.0 : -1.0;
        const BoundDirs bd = boundDirsHost(df.data(), df.size() / 3, sg, &axis);
        if (!bd.valid) continue;
        MinPyramid mp;
        const int e = medium_.minPyramid(axis, &mp);
        if (e) return e;
        if (mp.axis < 0) continue;
        const size_t n = df.size() / 3;
        const std::vector<uint32_t> perm = farOrderHost(df.data(), n, sg, axis);
        std::vector<BoundDirs> blocks((n + 63) / 64);
        for (size_t b = 0; b < blocks.size(); ++b) {
            const size_t m = std::min<size_t>(64, n - 64 * b);
            blocks[b] = boundDirsAxisHost(df.data(), perm.data() + 64 * b, m, sg, axis, bd.sgn);
        }
        const uint32_t cols = (uint32_t)std::min<size_t>(blocks.size(), 32);
        DevBuf<BoundDirs> blockDirs;
        int e2;
        if ((e2 = farPerm_[role].upload(perm.data(), perm.size())) ||
            (e2 = blockDirs.upload(blocks.data(), blocks.size())) || (e2 = farTable_[role].alloc(boxes * cols)))
            return e2;
        if (launchFarTable(mview_, mp, blockDirs.p, (uint32_t)blocks.size(), cols, sub, span, farTable_[role].p,
                           stream_) != hipSuccess ||
            hipStreamSynchronize(stream_) != hipSuccess)
            return SSTMC_KERNEL_ERROR;
        bytes_ += 4 * boxes * cols + 4 * perm.size();
        farBound_.perm[role] = farPerm_[role].p;
        farBound_.table[role] = farTable_[role].p;
        farBound_.blocks[role] = (uint32_t)blocks.size();
        farBound_.cols[role] = cols;
        farBound_.sub = sub;
        farBoundOn_ = true;
    }
    return SSTMC_NO_ERROR;
}

int NextEventEstimation::buildWalkGrid() {
    mview_ = medium_.view();
    // dominant axis of the view walks: far field, mean |direction component|; points, mean
    // |component| of the unit vector from the box centre to the sensor
    double w[3] = {0, 0, 0}, c[3];
    for (int a = 0; a < 3; ++a) c[a] = 0.5 * (medium_.boxMin(a) + medium_.boxMax(a));
    const std::vector<double>& pv = vw_.hostPoints();
    for (size_t i = 0; i + 2 < pv.size(); i += 3) {
        double q[3] = {pv[i], pv[i + 1], dims_ == 3 ? pv[i + 2] : 0.0};
        if (vw_.type() != SSTMC_SOURCE_FARFIELD)
            for (int a = 0; a < 3; ++a) q[a] -= a < dims_ ? c[a] : q[a];
        const double r = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (r > 0)
            for (int a = 0; a < dims_; ++a) w[a] += std::fabs(q[a]) / r;
    }
    int fast = 0;
    for (int a = 1; a < dims_; ++a)
        if (w[a] > w[fast]) fast = a;
    return medium_.walkGrid(fast, &mview_);
}

// Sizes the vertex pool from the exact vertex count of the first batch (its transport and scan,
// read back once; the counts are exact even where the staging is too small to hold the vertices), so that the batches queued after it do not overflow a pool sized by guess and
// have to run again. The margin covers batch-to-batch variation; a batch that still overflows is
// re-run by finish().
int NextEventEstimation::sizePool() {
    if (poolSized_) return SSTMC_NO_ERROR;
    if (syncAll()) return SSTMC_KERNEL_ERROR;
    const FramesView& fr = sampler_.view();
    VertexBatch vb = vertexBatch(0);
    BatchSet& bs = sets_[0];
    if (launchTransport(dims_, mview_, sc_.view(), il_.view(), fr, vb, transportStage(), seed_, 0, P_, stream_) !=
            hipSuccess ||
        launchScan(bs.k.p, bs.off.p, bs.blockSums.p, P_, stream_) != hipSuccess)
        return SSTMC_KERNEL_ERROR;
    uint32_t need = 0;
    if (hipMemcpyAsync(&need, bs.off.p + P_, sizeof(need), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
        return SSTMC_DEVICE_ERROR;
    const uint64_t want = (uint64_t)need + need / 4 + 1024;
    if (want > cap_) {
        int rc = growPool(want);
        if (rc) return rc;
    }
    poolSized_ = true;
    return SSTMC_NO_ERROR;
}

// The batch totals and the slotted voxel-step counters, cleared on st.
int NextEventEstimation::zeroCounters(hipStream_t st) {
    if (hipMemsetAsync(totals_.p, 0, kTotals * sizeof(unsigned long long), st) != hipSuccess ||
        hipMemsetAsync(stepSlots_.p, 0, stepSlots_.n * sizeof(unsigned long long), st) != hipSuccess)
        return SSTMC_DEVICE_ERROR;
    return SSTMC_NO_ERROR;
}

// The voxel steps of k_nee and of the view kernels: the sums of their counter slots (the streams
// must be idle).
void NextEventEstimation::stepCounts(uint64_t* nee, uint64_t* view) const {
    std::vector<unsigned long long> h(stepSlots_.n, 0ull);
    *nee = *view = 0;
    if (hipMemcpy(h.data(), stepSlots_.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess)
        return;
    const size_t set = (size_t)kCountSlots * kCountStride;
    for (size_t i = 0; i < set; i += kCountStride) {
        *nee += h[i];
        *view += h[set + i];
    }
}

int NextEventEstimation::begin() {
    int rc = sizePool();
    if (rc) return rc;
    graphReplays_ = 0;
    uint32_t nb = opt_.fullIteration;
    if (evBatches_ < nb) {
        for (hipEvent_t ev : ev_) hipEventDestroy(ev);
        ev_.assign((size_t)nb * kEventsPerBatch, nullptr);
        for (auto& ev : ev_)
            if (hipEventCreate(&ev) != hipSuccess) return SSTMC_DEVICE_ERROR;
        evBatches_ = nb;
    }
    if (hipMemsetAsync(status_.p, 0, 4 * sizeof(uint32_t), stream_) != hipSuccess) return SSTMC_DEVICE_ERROR;
    if (zeroCounters(stream_)) return SSTMC_DEVICE_ERROR;
    if (hipMemsetAsync(overflow_.p, 0, overflow_.n * sizeof(uint32_t), stream_) != hipSuccess)
        return SSTMC_DEVICE_ERROR;
    if (launchZeroU(U_.p, U_.n, stream_) != hipSuccess) return SSTMC_KERNEL_ERROR;
    return orderAfterNee();
}

// The transport and contraction streams wait for the work already queued on the NEE stream (the
// clears of the totals, flags and U).
int NextEventEstimation::orderAfterNee() {
    if (!cleared_ && hipEventCreateWithFlags(&cleared_, hipEventDisableTiming) != hipSuccess) return SSTMC_DEVICE_ERROR;
    if (hipEventRecord(cleared_, stream_) != hipSuccess || hipStreamWaitEvent(streamT_, cleared_, 0) != hipSuccess ||
        hipStreamWaitEvent(streamC_, cleared_, 0) != hipSuccess)
        return SSTMC_DEVICE_ERROR;
    return SSTMC_NO_ERROR;
}

int NextEventEstimation::enqueueBatch(uint32_t b) {
    if (b >= opt_.fullIteration) return SSTMC_INVALID_CONFIG;
    return runBatch((uint64_t)b * P_, P_, U_.p, true, b, setOf(b), true);
}

int NextEventEstimation::enqueue() {
    int rc = begin();
    if (rc) return rc;
    for (uint32_t b = 0; b < opt_.fullIteration; ++b)
        if ((rc = enqueueBatch(b))) return rc;
    return SSTMC_NO_ERROR;
}

// A batch whose vertices overflowed the pool added nothing to U and was not counted (every
// kernel after the transport sees the overflow and writes zero terms); it is re-run here after the
// pool has grown to the largest batch seen plus a margin. The re-run adds its share to U after the
// other batches rather than in batch order; pool sizes depend on the config alone, so a config
// still gives the same bits every time, alone or in a serving mix.
int NextEventEstimation::finish(double* u, bool deviceDst) {
    uint32_t st[4];
    unsigned long long tot[kTotals];
    for (int round = 0;; ++round) {
        if (syncAll()) return SSTMC_KERNEL_ERROR;
        if (hipMemcpy(st, status_.p, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) return SSTMC_DEVICE_ERROR;
        if (hipMemcpy(tot, totals_.p, sizeof(tot), hipMemcpyDeviceToHost) != hipSuccess) return SSTMC_DEVICE_ERROR;
        if (st[1]) return SSTMC_NO_HIT;
        if (!(st[0] & 1u)) break;
        // the pool now holds the largest batch with margin, so a re-run batch fits; a second
        // overflow would mean a batch's vertex count changed, which it cannot
        if (round == 2) return SSTMC_ALLOCATION_ERROR;
        std::vector<uint32_t> bits(overflow_.n);
        if (hipMemcpy(bits.data(), overflow_.p, bits.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return SSTMC_DEVICE_ERROR;
        const uint64_t want = std::max<uint64_t>(cap_ + cap_ / 2, (uint64_t)tot[3] + tot[3] / 4 + 1024);
        int rc = growPool(want);
        if (rc) return rc;
        if (hipMemsetAsync(status_.p, 0, sizeof(uint32_t), stream_) != hipSuccess ||
            hipMemsetAsync(overflow_.p, 0, overflow_.n * sizeof(uint32_t), stream_) != hipSuccess)
            return SSTMC_DEVICE_ERROR;
        if ((rc = orderAfterNee())) return rc;
        for (uint32_t b = 0; b < opt_.fullIteration; ++b)
            if (bits[b >> 5] & (1u << (b & 31)))
                if ((rc = runBatch((uint64_t)b * P_, P_, U_.p, true, b, setOf(b), true))) return rc;
    }
    zeroContributionPaths_ = tot[0];
    photonsBudget_ = tot[1];
    if (u && hipMemcpy(u, U_.p, U_.n * sizeof(double), deviceDst ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost) !=
                 hipSuccess)
        return SSTMC_DEVICE_ERROR;
    // stage timing from the per-batch events
    sstmc_timing t{};
    for (uint32_t b = 0; b < opt_.fullIteration; ++b) {
        float ms;
        hipEvent_t* e = &ev_[(size_t)b * kEventsPerBatch];
        if (hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess) t.transport_ms += ms;
        if (hipEventElapsedTime(&ms, e[6], e[2]) == hipSuccess) t.nee_ms += ms;
        // the multiple-scattering view kernel: after k_nee, or before it with viewFirst
        if (hipEventElapsedTime(&ms, viewFirst_ ? e[6] : e[5], viewFirst_ ? e[5] : e[2]) == hipSuccess)
            t.nee_view_ms += ms;
        if (hipEventElapsedTime(&ms, e[7], e[3]) == hipSuccess) t.contract_ms += ms;
        if (hipEventElapsedTime(&ms, e[3], e[4]) == hipSuccess) t.single_ms += ms;
    }
    float ms;
    if (opt_.fullIteration && hipEventElapsedTime(&ms, ev_[0], ev_[(size_t)(opt_.fullIteration - 1) * kEventsPerBatch + 4]) == hipSuccess)
        t.total_ms = ms;
    t.vertices = tot[1];
    uint64_t neeSteps = 0, viewSteps = 0;
    stepCounts(&neeSteps, &viewSteps);
    t.dda_steps = neeSteps + viewSteps;
    t.dda_steps_view = viewSteps;
    t.batches = opt_.fullIteration;
    t.graph_replays = graphReplays_;
    t.pipeline_sets = (uint32_t)nsets_;
    last_ = t;
    return SSTMC_NO_ERROR;
}

int NextEventEstimation::trace(double* u) {
    int rc = enqueue();
    return rc ? rc : finish(u);
}

void NextEventEstimation::timing(sstmc_timing* out) const { *out = last_; }

// Per-path terms of the last batch, from the [t][p][x] device layout to [p][t][x] complex doubles.
void NextEventEstimation::fetchTerms(const DevBuf<unsigned char>& src, uint32_t n, uint32_t width, double* dst) const {
    const size_t elem = fp64_ ? 16 : 8;
    std::vector<unsigned char> h((size_t)T_ * n * width * elem);
    hipMemcpy(h.data(), src.p, h.size(), hipMemcpyDeviceToHost);
    for (uint32_t t = 0; t < T_; ++t)
        for (uint32_t p = 0; p < n; ++p)
            for (uint32_t x = 0; x < width; ++x) {
                size_t si = ((size_t)t * n + p) * width + x;
                size_t di = (((size_t)p * T_ + t) * width + x) * 2;
                if (fp64_) {
                    const double* s = (const double*)h.data();
                    dst[di] = s[2 * si]; dst[di + 1] = s[2 * si + 1];
                } else {
                    const float* s = (const float*)h.data();
                    dst[di] = s[2 * si]; dst[di + 1] = s[2 * si + 1];
                }
            }
}

int NextEventEstimation::debugCbsTerms(uint32_t n, double* termAr, double* termBr) const {
    if (!opt_.cbs || n != lastDebugPaths_ || n == 0) return SSTMC_INVALID_CONFIG;
    if (termAr) fetchTerms(sets_[0].Ar, n, V_, termAr);
    if (termBr) fetchTerms(sets_[0].Br, n, L_, termBr);
    return SSTMC_NO_ERROR;
}

int NextEventEstimation::debugBatch(uint64_t pathBegin, uint32_t n, uint64_t* nvtxOut, uint32_t* pathK,
                                    uint32_t* pathMisses, uint64_t* pathOff, float* vtxPos, uint32_t* vtxMat,
                                    float* vtxRho, float* vtxXt, double* vtxLen, double* termA, double* termB,
                                    double* termSa, double* termSb, double* uBatch) {
    if (n == 0 || n > P_) return SSTMC_INVALID_CONFIG;
    DevBuf<double> Ub;
    int rc;
    if ((rc = Ub.alloc(U_.n))) return rc;
    if (syncAll()) return SSTMC_KERNEL_ERROR;
    const BatchSet& bs = sets_[0];
    for (int attempt = 0; attempt < 8; ++attempt) {
        hipMemsetAsync(status_.p, 0, 4 * sizeof(uint32_t), stream_);
        zeroCounters(stream_);
        hipMemsetAsync(overflow_.p, 0, overflow_.n * sizeof(uint32_t), stream_);
        launchZeroU(Ub.p, Ub.n, stream_);
        if ((rc = runBatch(pathBegin, n, Ub.p, false, 0, 0, false))) return rc;
        if (hipStreamSynchronize(stream_) != hipSuccess) return SSTMC_KERNEL_ERROR;
        uint32_t st[4];
        unsigned long long tot[kTotals];
        hipMemcpy(st, status_.p, sizeof(st), hipMemcpyDeviceToHost);
        hipMemcpy(tot, totals_.p, sizeof(tot), hipMemcpyDeviceToHost);
        if (st[1]) return SSTMC_NO_HIT;
        if (st[0] & 1u) {
            if ((rc = growPool(std::max<uint64_t>(cap_ * 2, tot[3] + tot[3] / 4 + 1024)))) return rc;
            continue;
        }
        std::vector<uint32_t> k(n), ms(n), off(n + 1);
        hipMemcpy(k.data(), bs.k.p, n * 4, hipMemcpyDeviceToHost);
        hipMemcpy(ms.data(), bs.misses.p, n * 4, hipMemcpyDeviceToHost);
        hipMemcpy(off.data(), bs.off.p, (n + 1) * 4, hipMemcpyDeviceToHost);
        uint64_t nv = off[n];
        if (nvtxOut) *nvtxOut = nv;
        if (!pathK) return SSTMC_NO_ERROR;
        for (uint32_t i = 0; i < n; ++i) { pathK[i] = k[i]; pathMisses[i] = ms[i]; pathOff[i] = off[i]; }
        pathOff[n] = nv;
        std::vector<float4> xm(nv);
        std::vector<float> rho(nv), dl;
        hipMemcpy(xm.data(), bs.xm.p, nv * sizeof(float4), hipMemcpyDeviceToHost);
        hipMemcpy(rho.data(), bs.rho.p, nv * 4, hipMemcpyDeviceToHost);
        const FramesView& fr = sampler_.view();
        if (!fr.trivial) {
            dl.resize(nv * T_ * 3);
            hipMemcpy(dl.data(), bs.delta.p, dl.size() * 4, hipMemcpyDeviceToHost);
        }
        for (uint64_t o = 0; o < nv; ++o) {
            vtxPos[3 * o] = xm[o].x; vtxPos[3 * o + 1] = xm[o].y; vtxPos[3 * o + 2] = xm[o].z;
            uint32_t mb;
            std::memcpy(&mb, &xm[o].w, 4);
            vtxMat[o] = mb;
            vtxRho[o] = rho[o];
            for (uint32_t t = 0; t < T_; ++t)
                for (int a = 0; a < 3; ++a)
                    vtxXt[(o * T_ + t) * 3 + a] = fr.trivial ? 0.f : dl[(o * T_ + t) * 3 + a];
        }
        // path length from x1 per frame, accumulated in FP64 from the FP32 transport positions
        // exactly as the NEE kernel does (vtxXt holds the per-frame displacement)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t t = 0; t < T_; ++t) {
                double acc = 0;
                for (uint32_t j = 0; j < k[i]; ++j) {
                    uint64_t o = off[i] + j;
                    if (j > 0) {
                        double s2 = 0;
                        for (int a = 0; a < dims_; ++a) {
                            double xa = (double)vtxPos[3 * o + a] + (double)vtxXt[(o * T_ + t) * 3 + a];
                            double xb = (double)vtxPos[3 * (o - 1) + a] + (double)vtxXt[((o - 1) * T_ + t) * 3 + a];
                            s2 += (xa - xb) * (xa - xb);
                        }
                        acc += std::sqrt(s2);
                    }
                    vtxLen[o * T_ + t] = acc;
                }
            }
        auto fetch = [&](const DevBuf<unsigned char>& src, uint32_t width, double* dst) {
            fetchTerms(src, n, width, dst);
        };
        lastDebugPaths_ = n;
        if (termA) fetch(bs.A, L_, termA);
        if (termB) fetch(bs.B, V_, termB);
        if (termSa) fetch(bs.a, L_, termSa);
        if (termSb) fetch(bs.b, V_, termSb);
        if (uBatch) hipMemcpy(uBatch, Ub.p, Ub.n * sizeof(double), hipMemcpyDeviceToHost);
        return SSTMC_NO_ERROR;
    }
    return SSTMC_ALLOCATION_ERROR;
}

int NextEventEstimation::tauRef(float* out) const {
    if (!(opt_.tauCut > 0)) return SSTMC_INVALID_CONFIG;
    if (syncAll()) return SSTMC_KERNEL_ERROR;
    return hipMemcpy(out, sets_[lastSet_].tauRef.p, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
               ? SSTMC_NO_ERROR
               : SSTMC_DEVICE_ERROR;
}

int NextEventEstimation::debugContract(uint32_t n, const double* a, const double* b, double* u) {
    if (n == 0) return SSTMC_INVALID_CONFIG;
    if (syncAll()) return SSTMC_KERNEL_ERROR;
    size_t na = (size_t)T_ * n * L_, nb = (size_t)T_ * n * V_;
    DevBuf<unsigned char> da, db;
    DevBuf<double> du;
    int rc;
    if ((rc = du.alloc(U_.n))) return rc;
    if (contractKind_ == kContractFp64) {
        std::vector<double> ha(2 * na), hb(2 * nb);
        for (size_t i = 0; i < 2 * na; ++i) ha[i] = a[i];
        for (size_t i = 0; i < 2 * nb; ++i) hb[i] = b[i];
        if ((rc = da.upload((const unsigned char*)ha.data(), ha.size() * 8))) return rc;
        if ((rc = db.upload((const unsigned char*)hb.data(), hb.size() * 8))) return rc;
    } else {
        std::vector<float> ha(2 * na), hb(2 * nb);
        for (size_t i = 0; i < 2 * na; ++i) ha[i] = (float)a[i];
        for (size_t i = 0; i < 2 * nb; ++i) hb[i] = (float)b[i];
        if ((rc = da.upload((const unsigned char*)ha.data(), ha.size() * 4))) return rc;
        if ((rc = db.upload((const unsigned char*)hb.data(), hb.size() * 4))) return rc;
    }
    launchZeroU(du.p, du.n, stream_);
    if (launchContract(contractKind_, da.p, db.p, n, L_, V_, T_, nullptr, nullptr, 0, 1.0, du.p, &wsMulti_, stream_) !=
        hipSuccess)
        return SSTMC_KERNEL_ERROR;
    if (hipStreamSynchronize(stream_) != hipSuccess) return SSTMC_KERNEL_ERROR;
    if (hipMemcpy(u, du.p, du.n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return SSTMC_DEVICE_ERROR;
    return SSTMC_NO_ERROR;
}

}  // namespace sstmc
