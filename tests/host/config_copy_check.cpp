// Host check of the serving loop's config deep copy (csrc/config_copy.h), no GPU:
//   * builds configs whose every pointer field is in use (tabular, vMF and HG materials; the
//     TemporalCorrelationSampler's illumination2 / view2; GaussianBeam directions), each array a
//     separate heap allocation;
//   * copies them, then overwrites the caller's arrays with garbage and frees them;
//   * checks that sstmc_validate accepts each copy and that every array of the copy holds the
//     values the caller passed (compared with an independently generated second set).
// Built with AddressSanitizer where available (a pointer left in the copy that still leads into a
// freed caller array is a heap-use-after-free) and run by tests/test_config_copy.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/sstmc.h"
#include "../../temporalspecklescode_amd/csrc/config_copy.h"

static int g_fail = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);          \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

// Heap arrays of a config, owned here so the test can poison and free them.
struct Owned {
    std::vector<double*> d;
    std::vector<uint32_t*> u;
    std::vector<int32_t*> i;
    std::vector<size_t> dn, un, in;
    double* D(std::vector<double> v) {
        double* p = new double[v.size()];
        std::memcpy(p, v.data(), v.size() * sizeof(double));
        d.push_back(p);
        dn.push_back(v.size());
        return p;
    }
    uint32_t* U(std::vector<uint32_t> v) {
        uint32_t* p = new uint32_t[v.size()];
        std::memcpy(p, v.data(), v.size() * sizeof(uint32_t));
        u.push_back(p);
        un.push_back(v.size());
        return p;
    }
    int32_t* I(std::vector<int32_t> v) {
        int32_t* p = new int32_t[v.size()];
        std::memcpy(p, v.data(), v.size() * sizeof(int32_t));
        i.push_back(p);
        in.push_back(v.size());
        return p;
    }
    void poisonAndFree() {
        for (size_t k = 0; k < d.size(); ++k) {
            for (size_t j = 0; j < dn[k]; ++j) d[k][j] = std::nan("");
            delete[] d[k];
        }
        for (size_t k = 0; k < u.size(); ++k) {
            for (size_t j = 0; j < un[k]; ++j) u[k][j] = 0xdeadbeefu;
            delete[] u[k];
        }
        for (size_t k = 0; k < i.size(); ++k) {
            for (size_t j = 0; j < in[k]; ++j) i[k][j] = -77;
            delete[] i[k];
        }
        d.clear(); u.clear(); i.clear();
    }
};

static std::vector<double> ramp(size_t n, double a, double b) {
    std::vector<double> v(n);
    for (size_t j = 0; j < n; ++j) v[j] = a + b * (double)j;
    return v;
}

// 3D 3x3x3 grid, materials 1..3 (tabular, vMF, HG), point sources; with `corr` the
// TemporalCorrelationSampler and illumination2 / view2, with `beam` GaussianBeam sources and vMF
// everywhere. Every array comes from `o`.
static sstmc_config makeConfig(Owned& o, bool corr, bool beam) {
    sstmc_config c{};
    c.dims = 3;
    c.simulation.batchSize = 64;
    c.simulation.fullIteration = 2;
    c.simulation.renderingsNum = 3;
    c.simulation.seed = 11;
    c.simulation.lambda = 0.75;
    c.medium.type = SSTMC_MEDIUM_HETEROGENEOUS;
    c.medium.xAxisSize = c.medium.yAxisSize = c.medium.zAxisSize = 4;
    c.medium.xAxis = o.D(ramp(4, -3.0, 2.0));
    c.medium.yAxis = o.D(ramp(4, -3.5, 2.5));
    c.medium.zAxis = o.D(ramp(4, -2.0, 1.5));
    std::vector<uint32_t> grid(27);
    for (size_t j = 0; j < 27; ++j) grid[j] = 1 + (uint32_t)((j * 7) % 3);
    c.medium.materialGrid = o.U(grid);
    c.medium.materialNum = 4;
    c.medium.sigs = o.D({0.0, 0.5, 1.0, 1.5});
    c.medium.siga = o.D({0.0, 0.01, 0.02, 0.03});
    c.scattering.materialNum = 4;
    if (beam) {
        c.scattering.type = o.I({SSTMC_SCATTERING_VMF, SSTMC_SCATTERING_VMF, SSTMC_SCATTERING_VMF, SSTMC_SCATTERING_VMF});
        c.scattering.g = o.D({0, 0, 0, 0});
        c.scattering.vmfMixtures = o.U({1, 2, 3, 1});
        c.scattering.vmfMu = o.D(ramp(7, 0.5, 1.25));
        c.scattering.vmfC = o.D(ramp(7, -1.0, -0.5));
        c.scattering.vmfAlpha = o.D(ramp(7, 0.2, 0.1));
    } else {
        c.scattering.type = o.I({SSTMC_SCATTERING_HG, SSTMC_SCATTERING_TABULAR, SSTMC_SCATTERING_VMF, SSTMC_SCATTERING_HG});
        c.scattering.g = o.D({0.3, 0.0, 0.0, 0.85});
        c.scattering.tabularSize = o.U({0, 5, 0, 0});
        c.scattering.tabular = o.D(ramp(10, 1.0, -0.05));
        c.scattering.vmfMixtures = o.U({0, 0, 2, 0});
        c.scattering.vmfMu = o.D({2.0, 7.5});
        c.scattering.vmfC = o.D({-1.5, -6.0});
        c.scattering.vmfAlpha = o.D({0.7, 0.3});
    }
    sstmc_source* src[2] = {&c.illumination, &c.view};
    for (int s = 0; s < 2; ++s) {
        src[s]->num = 3;
        src[s]->x = o.D(ramp(3, -1.0 + s, 0.5));
        src[s]->y = o.D(ramp(3, 0.25 * s, -0.75));
        src[s]->z = o.D(ramp(3, s ? 4.0 : -4.0, 0.125));
        if (beam) {
            src[s]->type = SSTMC_SOURCE_GAUSSIAN_BEAM;
            src[s]->dirX = o.D({0.0, 0.6, 0.0});
            src[s]->dirY = o.D({0.0, 0.0, 0.8});
            src[s]->dirZ = o.D({1.0, 0.8, 0.6});
            src[s]->aperture = 0.3 + s;
        } else {
            src[s]->type = SSTMC_SOURCE_POINT;
        }
    }
    c.sampler.type = corr ? SSTMC_SAMPLER_TEMPORAL_CORRELATION : SSTMC_SAMPLER_TEMPORAL;
    c.sampler.tNum = 3;
    c.sampler.t = o.D({0.0, 0.5, 1.5});
    c.sampler.D = o.D({0.0, 1e-3, 2e-3, 0.0});
    c.sampler.Ux = o.D({0.0, 0.01, 0.0, 0.0});
    c.sampler.Uy = o.D({0.0, 0.0, -0.02, 0.0});
    c.sampler.Uz = o.D({0.0, 0.0, 0.0, 0.03});
    c.sampler.isTemporalAmplitude = 1;
    c.tracer.type = SSTMC_TRACER_NEE;
    if (corr) {
        sstmc_source* s2[2] = {&c.illumination2, &c.view2};
        for (int s = 0; s < 2; ++s) {
            s2[s]->type = SSTMC_SOURCE_POINT;
            s2[s]->num = 3;
            s2[s]->x = o.D(ramp(3, 2.0 - s, -0.25));
            s2[s]->y = o.D(ramp(3, -0.5, 0.75 + s));
            s2[s]->z = o.D(ramp(3, s ? 4.5 : -4.5, -0.125));
        }
    }
    return c;
}

template <typename T>
static bool same(const T* a, const T* b, size_t n) {
    if (!a || !b) return a == b;
    for (size_t j = 0; j < n; ++j)
        if (!(a[j] == b[j])) return false;
    return true;
}

static void compareSource(const sstmc_source& a, const sstmc_source& b) {
    CHECK(a.type == b.type && a.num == b.num && a.isAreaSource == b.isAreaSource && a.aperture == b.aperture);
    CHECK(same(a.x, b.x, a.num) && same(a.y, b.y, a.num) && same(a.z, b.z, a.num));
    CHECK(same(a.dirX, b.dirX, a.num) && same(a.dirY, b.dirY, a.num) && same(a.dirZ, b.dirZ, a.num));
}

// Every array of `got` equals the one of `want` (a config built independently with the same
// values).
static void compare(const sstmc_config& got, const sstmc_config& want) {
    CHECK(got.dims == want.dims);
    CHECK(std::memcmp(&got.simulation, &want.simulation, sizeof(got.simulation)) == 0);
    const sstmc_medium &m = got.medium, &w = want.medium;
    CHECK(m.xAxisSize == w.xAxisSize && m.yAxisSize == w.yAxisSize && m.zAxisSize == w.zAxisSize);
    CHECK(same(m.xAxis, w.xAxis, m.xAxisSize) && same(m.yAxis, w.yAxis, m.yAxisSize) &&
          same(m.zAxis, w.zAxis, m.zAxisSize));
    CHECK(same(m.materialGrid, w.materialGrid, 27));
    CHECK(m.materialNum == w.materialNum && same(m.sigs, w.sigs, 4) && same(m.siga, w.siga, 4));
    const sstmc_scattering &s = got.scattering, &ws = want.scattering;
    CHECK(s.materialNum == ws.materialNum && same(s.type, ws.type, 4) && same(s.g, ws.g, 4));
    CHECK(same(s.tabularSize, ws.tabularSize, 4) && same(s.vmfMixtures, ws.vmfMixtures, 4));
    size_t nt = 0, nv = 0;
    for (int j = 0; j < 4; ++j) {
        if (ws.tabularSize) nt += 2 * ws.tabularSize[j];
        if (ws.vmfMixtures) nv += ws.vmfMixtures[j];
    }
    CHECK(same(s.tabular, ws.tabular, nt));
    CHECK(same(s.vmfMu, ws.vmfMu, nv) && same(s.vmfC, ws.vmfC, nv) && same(s.vmfAlpha, ws.vmfAlpha, nv));
    compareSource(got.illumination, want.illumination);
    compareSource(got.view, want.view);
    compareSource(got.illumination2, want.illumination2);
    compareSource(got.view2, want.view2);
    const sstmc_sampler &p = got.sampler, &wp = want.sampler;
    CHECK(p.type == wp.type && p.tNum == wp.tNum && p.isTemporalAmplitude == wp.isTemporalAmplitude);
    CHECK(same(p.t, wp.t, p.tNum) && same(p.D, wp.D, 4) && same(p.Ux, wp.Ux, 4) && same(p.Uy, wp.Uy, 4) &&
          same(p.Uz, wp.Uz, 4));
    CHECK(std::memcmp(&got.tracer, &want.tracer, sizeof(got.tracer)) == 0);
    CHECK(std::memcmp(&got.precision, &want.precision, sizeof(got.precision)) == 0);
}

int main() {
    const bool cases[3][2] = {{false, false}, {true, false}, {false, true}};
    for (const auto& cs : cases) {
        Owned caller, expected;
        sstmc_config c = makeConfig(caller, cs[0], cs[1]);
        const sstmc_config want = makeConfig(expected, cs[0], cs[1]);
        CHECK(sstmc_validate(&c) == SSTMC_NO_ERROR);
        sstmc::ConfigCopy copy(c);
        caller.poisonAndFree();
        CHECK(sstmc_validate(&copy.c) == SSTMC_NO_ERROR);
        compare(copy.c, want);
        expected.poisonAndFree();
    }
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
