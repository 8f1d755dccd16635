// capi.hip -- the extern "C" boundary declared in include/sstmc.h.
//
// sstmc_create follows mexFunction's construction order (MatlabInterface.cu:1671-1700):
// scattering -> medium -> illumination/view -> sampler -> tracer, with the reference's config
// validation (MatlabInterface.cu:712-1563) turned into SSTMC_INVALID_CONFIG returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>

#include "../../include/sstmc.h"
#include "tracer.h"

using namespace sstmc;

struct sstmc_handle {
    int device = 0;
    uint32_t renderings = 1;
    uint64_t seed = 0;
    std::unique_ptr<Scattering> scattering;
    std::shared_ptr<Medium> medium;   // shared with the serving loop's medium cache
    std::unique_ptr<Source> illumination, view;
    std::unique_ptr<TemporalSampler> sampler;
    std::unique_ptr<NextEventEstimation> tracer;
};

static int buildSource(const sstmc_source& s, const Medium& med, int dims, bool illum, std::unique_ptr<Source>* out,
                       const Scattering* sc = nullptr, const sstmc_source* views = nullptr, double lambda = 1.0) {
    if (s.num == 0 || !s.x || !s.y || (dims == 3 && !s.z)) return SSTMC_INVALID_CONFIG;
    int err = SSTMC_NO_ERROR;
    if (s.type == SSTMC_SOURCE_GAUSSIAN_BEAM && sc)
        out->reset(new GaussianBeam(&err, med, *sc, s, views, lambda, illum));
    else if (s.type == SSTMC_SOURCE_FARFIELD)
        out->reset(new FarField(&err, med, dims, s.x, s.y, s.z, s.num, illum));
    else if (s.type == SSTMC_SOURCE_POINT)
        out->reset(new PointSource(&err, dims, s.x, s.y, s.z, s.num, s.isAreaSource != 0, illum));
    else
        return SSTMC_INVALID_CONFIG;
    return err;
}

// The TemporalCorrelationSampler's sources: the u1 sources then the u2 ones in one source object
// (validated to share type, count and area flag), so that the first vertex is drawn from both sets
// and the kernels address u2 source i as entry n + i.
static int buildSourcePair(const sstmc_source& s1, const sstmc_source& s2, const Medium& med, int dims, bool illum,
                           std::unique_ptr<Source>* out) {
    const uint32_t n = s1.num;
    std::vector<double> x(2 * (size_t)n), y(2 * (size_t)n), z(2 * (size_t)n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        x[i] = s1.x[i]; x[n + i] = s2.x[i];
        y[i] = s1.y[i]; y[n + i] = s2.y[i];
        if (dims == 3) { z[i] = s1.z[i]; z[n + i] = s2.z[i]; }
    }
    sstmc_source both = s1;
    both.num = 2 * n;
    both.x = x.data();
    both.y = y.data();
    both.z = z.data();
    return buildSource(both, med, dims, illum, out);
}

// Host-only validation, before anything touches the device (MatlabInterface.cu:712-1563).
// checkGrid = false skips the per-cell checks of the material grid (a medium taken from the
// serving loop's cache, whose grid was checked when it was built).
static int validate(const sstmc_config* c, bool checkGrid = true) {
    if (!c) return SSTMC_INVALID_CONFIG;
    int dims = (int)c->dims;
    if (dims != 2 && dims != 3) return SSTMC_INVALID_CONFIG;
    if (c->tracer.type != SSTMC_TRACER_NEE && c->tracer.type != 0) return SSTMC_INVALID_CONFIG;
    // coherent backscattering: the field output (NextEventEstimation.cuh:1212-1235, 1385-1391,
    // 1432-1454) and the TemporalCorrelationSampler's correlation (CBS01-CBS04, :1138-1391) are
    // built; the TemporalSampler's per-sub-path correlation output has no reversed-path form
    if (c->tracer.isCBS && c->tracer.output == SSTMC_OUTPUT_CORRELATION) return SSTMC_NOT_SUPPORTED;
    if (c->tracer.output != SSTMC_OUTPUT_FIELD && c->tracer.output != SSTMC_OUTPUT_CORRELATION)
        return SSTMC_INVALID_CONFIG;
    const sstmc_scattering& sc = c->scattering;
    if (sc.materialNum == 0 || sc.materialNum > (uint32_t)kMaxMaterials || !sc.type) return SSTMC_INVALID_CONFIG;
    uint32_t nmat = sc.materialNum == 1 ? 2 : sc.materialNum;
    for (uint32_t m = 0; m < sc.materialNum; ++m) {
        if (sc.type[m] != SSTMC_SCATTERING_ISOTROPIC && sc.type[m] != SSTMC_SCATTERING_HG &&
            sc.type[m] != SSTMC_SCATTERING_TABULAR && sc.type[m] != SSTMC_SCATTERING_VMF)
            return SSTMC_NOT_SUPPORTED;
        if (sc.type[m] == SSTMC_SCATTERING_VMF &&
            (!sc.vmfMixtures || !sc.vmfMu || !sc.vmfC || !sc.vmfAlpha || sc.vmfMixtures[m] == 0))
            return SSTMC_INVALID_CONFIG;
        if (sc.type[m] == SSTMC_SCATTERING_HG && sc.g && !(std::fabs(sc.g[m]) < 1.0)) return SSTMC_INVALID_CONFIG;
        if (sc.type[m] == SSTMC_SCATTERING_TABULAR && (!sc.tabularSize || !sc.tabular || sc.tabularSize[m] < 2))
            return SSTMC_INVALID_CONFIG;
    }
    const sstmc_medium& md = c->medium;
    if (md.type != SSTMC_MEDIUM_HETEROGENEOUS) return SSTMC_NOT_SUPPORTED;
    const double* ax[3] = {md.xAxis, md.yAxis, md.zAxis};
    uint32_t sz[3] = {md.xAxisSize, md.yAxisSize, md.zAxisSize};
    size_t ncell = 1;
    for (int a = 0; a < dims; ++a) {
        if (!ax[a] || sz[a] < 2) return SSTMC_INVALID_CONFIG;
        for (uint32_t i = 1; i < sz[a]; ++i)
            if (!(ax[a][i] > ax[a][i - 1])) return SSTMC_INVALID_CONFIG;
        ncell *= sz[a] - 1;
    }
    if ((checkGrid && !md.materialGrid) || !md.sigs || md.materialNum != nmat) return SSTMC_INVALID_CONFIG;
    if (checkGrid)
        for (size_t i = 0; i < ncell; ++i)   // material 0 only outside the grid (MatlabInterface.cu:975-992)
            if (md.materialGrid[i] == 0 || md.materialGrid[i] >= nmat) return SSTMC_INVALID_CONFIG;
    double maj = 0;
    for (uint32_t m = 0; m < nmat; ++m) {
        if (!(md.sigs[m] >= 0) || (md.siga && !(md.siga[m] >= 0))) return SSTMC_INVALID_CONFIG;
        maj = std::max(maj, md.sigs[m] + (md.siga ? md.siga[m] : 0.0));
    }
    if (!(maj > 0)) return SSTMC_INVALID_CONFIG;
    const sstmc_source* srcs[2] = {&c->illumination, &c->view};
    for (const sstmc_source* s : srcs) {
        if (s->type != SSTMC_SOURCE_FARFIELD && s->type != SSTMC_SOURCE_POINT && s->type != SSTMC_SOURCE_GAUSSIAN_BEAM)
            return SSTMC_INVALID_CONFIG;
        if (s->num == 0 || !s->x || !s->y || (dims == 3 && !s->z)) return SSTMC_INVALID_CONFIG;
        // GaussianBeam (MatlabInterface.cu:1153-1190): one direction per focal point, aperture > 0
        if (s->type == SSTMC_SOURCE_GAUSSIAN_BEAM &&
            (!s->dirX || !s->dirY || (dims == 3 && !s->dirZ) || !(s->aperture > 0)))
            return SSTMC_INVALID_CONFIG;
    }
    if (c->illumination.type == SSTMC_SOURCE_GAUSSIAN_BEAM || c->view.type == SSTMC_SOURCE_GAUSSIAN_BEAM) {
        // built: GaussianBeam illumination and view (fsGaussianBeam pairs two beams,
        // GaussianBeam.cuh:2040-2096), every material a vMF mixture of at most 16 lobes
        // (updateSampler, :1283-1290), the TemporalSampler without coherent backscattering
        if (c->illumination.type != c->view.type || c->tracer.isCBS ||
            c->sampler.type == SSTMC_SAMPLER_TEMPORAL_CORRELATION)
            return SSTMC_NOT_SUPPORTED;
        for (uint32_t m = 0; m < sc.materialNum; ++m)
            if (sc.type[m] != SSTMC_SCATTERING_VMF || sc.vmfMixtures[m] > 16) return SSTMC_NOT_SUPPORTED;
    }
    const sstmc_sampler& sp = c->sampler;
    if (sp.type != SSTMC_SAMPLER_TEMPORAL && sp.type != SSTMC_SAMPLER_TEMPORAL_CORRELATION && sp.type != 0)
        return SSTMC_NOT_SUPPORTED;
    if (sp.tNum == 0 || !sp.t || !sp.D) return SSTMC_INVALID_CONFIG;
    for (uint32_t m = 0; m < nmat; ++m)
        if (!(sp.D[m] >= 0)) return SSTMC_INVALID_CONFIG;   // MatlabInterface.cu:1293-1297
    if (sp.type == SSTMC_SAMPLER_TEMPORAL_CORRELATION) {
        // the sampler's output is already a correlation (with isCBS: CBS01-CBS04,
        // NextEventEstimation.cuh:1138-1391)
        if (c->tracer.output != SSTMC_OUTPUT_FIELD) return SSTMC_INVALID_CONFIG;
        // illumination2 / view2: present, and as large as illumination / view (MatlabInterface.cu:
        // 1313-1331); the u1 and u2 sets share one source type and area flag here
        const sstmc_source* pair[2] = {&c->illumination2, &c->view2};
        for (int i = 0; i < 2; ++i) {
            const sstmc_source* s2 = pair[i];
            const sstmc_source* s1 = srcs[i];
            if (s2->type != SSTMC_SOURCE_FARFIELD && s2->type != SSTMC_SOURCE_POINT) return SSTMC_INVALID_CONFIG;
            if (s2->num != s1->num || !s2->x || !s2->y || (dims == 3 && !s2->z)) return SSTMC_INVALID_CONFIG;
            if (s2->type != s1->type || (s2->isAreaSource != 0) != (s1->isAreaSource != 0)) return SSTMC_NOT_SUPPORTED;
        }
        if ((uint64_t)c->illumination.num * 2 > 0xffffffffull || (uint64_t)c->view.num * 2 > 0xffffffffull)
            return SSTMC_INVALID_CONFIG;
    }
    if (c->simulation.lambda < 0 || !std::isfinite(c->simulation.lambda)) return SSTMC_INVALID_CONFIG;
    if (c->precision.terms != SSTMC_TERMS_FP32 && c->precision.terms != SSTMC_TERMS_FP64) return SSTMC_INVALID_CONFIG;
    if (c->precision.contraction < SSTMC_CONTRACT_BF16X3 || c->precision.contraction > SSTMC_CONTRACT_FP64)
        return SSTMC_INVALID_CONFIG;
    if (!(c->precision.tauCut >= 0) || !std::isfinite(c->precision.tauCut)) return SSTMC_INVALID_CONFIG;
    return SSTMC_NO_ERROR;
}

// HeterogeneousMedium of a validated config (MatlabInterface.cu:951-1056).
static int buildMedium(const sstmc_config* c, std::shared_ptr<Medium>* out) {
    const int dims = (int)c->dims;
    const sstmc_medium& md = c->medium;
    const uint32_t nmat = c->scattering.materialNum == 1 ? 2 : c->scattering.materialNum;
    int err = SSTMC_NO_ERROR;
    std::shared_ptr<Medium> m(new HeterogeneousMedium(&err, dims, md.materialGrid, md.xAxis, md.xAxisSize, md.yAxis,
                                                      md.yAxisSize, md.zAxis, md.zAxisSize));
    if (err) return err;
    if ((err = m->setMaterial(md.sigs, md.siga, nmat))) return err;
    *out = std::move(m);
    return SSTMC_NO_ERROR;
}

// sstmc_create; with `shared` the handle uses that medium (built from an identical medium
// config) instead of building one, and the config's material grid is not read.
static int create(const sstmc_config* c, sstmc_handle** out, std::shared_ptr<Medium> shared = nullptr) {
    if (!c || !out) return SSTMC_INVALID_CONFIG;
    *out = nullptr;
    int vrc = validate(c, !shared);
    if (vrc) return vrc;
    int dims = (int)c->dims;
    std::unique_ptr<sstmc_handle> h(new sstmc_handle());
    h->device = (int)c->simulation.deviceNumber;
    if (hipSetDevice(h->device) != hipSuccess) return SSTMC_DEVICE_ERROR;
    h->renderings = c->simulation.renderingsNum ? c->simulation.renderingsNum : 1;
    h->seed = c->simulation.seed;

    // scattering (MatlabInterface.cu:712-888): a single cell also defines material 1
    const sstmc_scattering& sc = c->scattering;
    if (sc.materialNum == 0 || sc.materialNum > (uint32_t)kMaxMaterials || !sc.type) return SSTMC_INVALID_CONFIG;
    uint32_t nmat = sc.materialNum == 1 ? 2 : sc.materialNum;
    h->scattering.reset(new Scattering(dims));
    std::vector<size_t> tabOff(sc.materialNum, 0), vmfOff(sc.materialNum, 0);   // starts in the arrays
    for (uint32_t m = 0, acc = 0, accv = 0; m < sc.materialNum; ++m) {
        tabOff[m] = acc;
        vmfOff[m] = accv;
        if (sc.tabularSize) acc += 2 * sc.tabularSize[m];
        if (sc.vmfMixtures) accv += sc.vmfMixtures[m];
    }
    for (uint32_t m = 0; m < nmat; ++m) {
        uint32_t src = sc.materialNum == 1 ? 0 : m;
        int rc;
        if (sc.type[src] == SSTMC_SCATTERING_ISOTROPIC) rc = h->scattering->setIsotropicScattering(m);
        else if (sc.type[src] == SSTMC_SCATTERING_HG) rc = h->scattering->setHenyeyGreensteinScattering(sc.g ? sc.g[src] : 0.0, m);
        else if (sc.type[src] == SSTMC_SCATTERING_TABULAR)
            rc = h->scattering->setTabularScattering(sc.tabular + tabOff[src], sc.tabularSize[src], m);
        else if (sc.type[src] == SSTMC_SCATTERING_VMF)
            rc = h->scattering->setVonMisesFisherScattering(sc.vmfMu + vmfOff[src], sc.vmfC + vmfOff[src],
                                                            sc.vmfAlpha + vmfOff[src], sc.vmfMixtures[src], m);
        else return SSTMC_NOT_SUPPORTED;
        if (rc) return rc;
    }
    if (int rc = h->scattering->upload()) return rc;

    // medium (MatlabInterface.cu:951-1056), validated above
    int err = SSTMC_NO_ERROR;
    if (shared) h->medium = std::move(shared);
    else if ((err = buildMedium(c, &h->medium))) return err;

    // sources (MatlabInterface.cu:1062-1215)
    const sstmc_sampler& sp = c->sampler;
    const bool correlation = sp.type == SSTMC_SAMPLER_TEMPORAL_CORRELATION;
    if (correlation) {
        if ((err = buildSourcePair(c->illumination, c->illumination2, *h->medium, dims, true, &h->illumination)))
            return err;
        if ((err = buildSourcePair(c->view, c->view2, *h->medium, dims, false, &h->view))) return err;
    } else {
        const double lam = c->simulation.lambda > 0 ? c->simulation.lambda : 1.0;
        if ((err = buildSource(c->illumination, *h->medium, dims, true, &h->illumination, h->scattering.get(), &c->view,
                               lam)))
            return err;
        if ((err = buildSource(c->view, *h->medium, dims, false, &h->view, h->scattering.get(), nullptr, lam)))
            return err;
    }

    // sampler (MatlabInterface.cu:1217-1368)
    if (sp.tNum == 0 || !sp.t || !sp.D) return SSTMC_INVALID_CONFIG;
    h->sampler.reset(new TemporalSampler(&err, sp.t, sp.tNum, sp.isTemporalAmplitude != 0, correlation));
    if (err) return err;
    for (uint32_t m = 0; m < nmat; ++m) {
        if (sp.D[m] < 0) return SSTMC_INVALID_CONFIG;   // MatlabInterface.cu:1293-1297
        err = h->sampler->setMaterial(sp.D[m], sp.Ux ? sp.Ux[m] : 0.0, sp.Uy ? sp.Uy[m] : 0.0,
                                      (dims == 3 && sp.Uz) ? sp.Uz[m] : 0.0, m);
        if (err) return err;
    }
    h->sampler->finalize(nmat);

    // tracer (MatlabInterface.cu:1374-1457, defaults :1467, :1590-1594)
    TracerOptions opt;
    opt.batchSize = c->simulation.batchSize ? c->simulation.batchSize : 1024;
    opt.fullIteration = c->simulation.fullIteration ? c->simulation.fullIteration : 1;
    opt.seed = c->simulation.seed;
    opt.lambda = c->simulation.lambda > 0 ? c->simulation.lambda : 1.0;
    if (!(std::isfinite(opt.lambda))) return SSTMC_INVALID_CONFIG;
    opt.output = c->tracer.output;
    opt.cbs = c->tracer.isCBS ? 1 : 0;
    opt.terms = c->precision.terms;
    opt.contraction = c->precision.contraction;
    opt.tauCut = c->precision.tauCut;
    opt.graphs = c->simulation.graphs ? 1 : 0;
    h->tracer.reset(new NextEventEstimation(&err, *h->medium, *h->scattering, *h->illumination, *h->view,
                                            *h->sampler, opt));
    if (err) return err;
    *out = h.release();
    return SSTMC_NO_ERROR;
}

namespace sstmc {
NextEventEstimation* handleTracer(sstmc_handle* h) { return h ? h->tracer.get() : nullptr; }
int configValidate(const sstmc_config* c) { return validate(c); }
int mediumBuild(const sstmc_config* c, std::shared_ptr<Medium>* out) {
    int rc = validate(c);
    if (rc) return rc;
    if (hipSetDevice((int)c->simulation.deviceNumber) != hipSuccess) return SSTMC_DEVICE_ERROR;
    try {
        return buildMedium(c, out);
    } catch (...) {
        return SSTMC_ALLOCATION_ERROR;
    }
}
int createShared(const sstmc_config* c, std::shared_ptr<Medium> medium, sstmc_handle** out) {
    try {
        return create(c, out, std::move(medium));
    } catch (...) {
        return SSTMC_ALLOCATION_ERROR;
    }
}
}  // namespace sstmc

extern "C" {

int sstmc_validate(const sstmc_config* cfg) { return validate(cfg); }

int sstmc_create(const sstmc_config* cfg, sstmc_handle** out) {
    try {
        return create(cfg, out);
    } catch (...) {
        return SSTMC_ALLOCATION_ERROR;
    }
}

int sstmc_trace(sstmc_handle* h, double* u, double* iter) {
    if (!h || !u) return SSTMC_INVALID_CONFIG;
    hipSetDevice(h->device);
    int rc = h->tracer->trace(u);
    if (rc) return rc;
    // iter = batchSize * fullIteration + zeroContributionPaths (MatlabInterface.cu:1796)
    if (iter) *iter = (double)h->tracer->paths() + (double)h->tracer->getZeroContributionPaths();
    return SSTMC_NO_ERROR;
}

int sstmc_update_seed(sstmc_handle* h, uint64_t seed) {
    if (!h) return SSTMC_INVALID_CONFIG;
    h->seed = seed;
    h->tracer->updateSeed(seed);
    return SSTMC_NO_ERROR;
}

uint64_t sstmc_zero_contribution_paths(const sstmc_handle* h) { return h ? h->tracer->getZeroContributionPaths() : 0; }
uint64_t sstmc_photons_budget(const sstmc_handle* h) { return h ? h->tracer->getPhotonsBudget() : 0; }

int sstmc_sizes(const sstmc_handle* h, uint32_t* nl, uint32_t* nv, uint32_t* nt) {
    if (!h) return SSTMC_INVALID_CONFIG;
    if (nl) *nl = h->tracer->L();
    if (nv) *nv = h->tracer->V();
    if (nt) *nt = h->tracer->T();
    return SSTMC_NO_ERROR;
}

void sstmc_destroy(sstmc_handle* h) {
    if (!h) return;
    hipSetDevice(h->device);
    delete h;
}

const char* sstmc_error_string(int code) {
    switch (code) {
        case SSTMC_NO_ERROR: return "No error";
        case SSTMC_DEVICE_ERROR: return "Device error: the HIP device could not be used";
        case SSTMC_ALLOCATION_ERROR: return "Allocation error: not enough device memory, decrease batchSize";
        case SSTMC_KERNEL_ERROR: return "Kernel error: a HIP kernel failed";
        case SSTMC_BLAS_ERROR: return "BLAS error";
        case SSTMC_NOT_SUPPORTED: return "Not supported by this build";
        case SSTMC_MISC_ERROR: return "Miscellaneous error";
        case SSTMC_INVALID_CONFIG: return "Invalid config";
        case SSTMC_NO_HIT: return "No path reached the medium within the retry budget";
        case SSTMC_BUDGET_EXCEEDED: return "Request larger than the admission budget";
        case SSTMC_NOT_READY: return "Request still in flight";
        case SSTMC_SCRATCH_EXCEEDED:
            return "Contraction too large for the allocation-free entry point: one split's L x V partials exceed "
                   "its fixed 64 MiB scratch";
        default: return "Unknown error";
    }
}

int sstmc_render(const sstmc_config* cfg, double* u, double* iter) {
    sstmc_handle* h = nullptr;
    int rc = sstmc_create(cfg, &h);
    if (rc) return rc;
    uint32_t nl, nv, nt;
    sstmc_sizes(h, &nl, &nv, &nt);
    size_t stride = 2 * (size_t)nl * nv * nt;
    for (uint32_t i = 0; i < h->renderings; ++i) {
        sstmc_update_seed(h, cfg->simulation.seed + i);   // seed++ per rendering (MatlabInterface.cu:1799)
        rc = sstmc_trace(h, u + stride * i, iter ? iter + i : nullptr);
        if (rc) break;
    }
    sstmc_destroy(h);
    return rc;
}

int sstmc_last_timing(const sstmc_handle* h, sstmc_timing* out) {
    if (!h || !out) return SSTMC_INVALID_CONFIG;
    h->tracer->timing(out);
    return SSTMC_NO_ERROR;
}

int sstmc_debug_batch(sstmc_handle* h, uint64_t path_begin, uint32_t npaths, uint64_t* nvtx_out, uint32_t* path_k,
                      uint32_t* path_misses, uint64_t* path_off, float* vtx_pos, uint32_t* vtx_mat, float* vtx_rho,
                      float* vtx_dxt, double* vtx_len, double* term_a, double* term_b, double* term_sa,
                      double* term_sb, double* u_batch) {
    if (!h) return SSTMC_INVALID_CONFIG;
    hipSetDevice(h->device);
    return h->tracer->debugBatch(path_begin, npaths, nvtx_out, path_k, path_misses, path_off, vtx_pos, vtx_mat, vtx_rho,
                                 vtx_dxt, vtx_len, term_a, term_b, term_sa, term_sb, u_batch);
}

int sstmc_debug_cbs(sstmc_handle* h, uint32_t npaths, double* term_ar, double* term_br) {
    if (!h) return SSTMC_INVALID_CONFIG;
    hipSetDevice(h->device);
    return h->tracer->debugCbsTerms(npaths, term_ar, term_br);
}

int sstmc_debug_tau_ref(const sstmc_handle* h, float* tau_ref) {
    if (!h || !tau_ref) return SSTMC_INVALID_CONFIG;
    hipSetDevice(h->device);
    return h->tracer->tauRef(tau_ref);
}

int sstmc_debug_contract(sstmc_handle* h, uint32_t npaths, const double* a, const double* b, double* u) {
    if (!h || !a || !b || !u) return SSTMC_INVALID_CONFIG;
    hipSetDevice(h->device);
    return h->tracer->debugContract(npaths, a, b, u);
}

}  // extern "C"
