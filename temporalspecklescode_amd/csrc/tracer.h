// tracer.h -- host-side plugin split of the reference, MI355X-native underneath.
//
// The classes keep the reference's Medium / Scattering / Source / Sampler / Tracer interface
// split (MediumInterface.cuh:39-135, Scattering.cuh:32-125, SourceInterface.cuh:10-179,
// SamplerInterface.cuh:11-176, TracerInterface.cuh:9-50): constructors take host arrays and
// report through an `int* err` out-parameter, objects own their device memory. Instead of each
// class launching its own small kernels with a device sync after each (about 30 launches per
// bounce in NextEventEstimation::trace), each class exposes a device view (common.h) and the
// tracer runs a fixed sequence of fused kernels per batch of paths on its own HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/sstmc.h"
#include "common.h"
#include "kernels.h"

namespace sstmc {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t count);
    int upload(const T* host, size_t count);
    void release();
};

// ---------------------------------------------------------------- Medium
class Medium {
public:
    virtual ~Medium() = default;
    const MediumView& view() const { return v_; }
    int dims() const { return v_.dims; }
    double boxMin(int a) const { return v_.bmind[a]; }
    double boxMax(int a) const { return v_.bmaxd[a]; }
    // Medium::setMaterial (MediumInterface.cuh:386-397) for all materials at once.
    virtual int setMaterial(const double* sigs, const double* siga, uint32_t materialNum) = 0;
    uint32_t materialNum() const { return nmat_; }
    double sigs(uint32_t m) const { return sigs_[m]; }
    // material id per cell (x fastest) on the host
    virtual const std::vector<uint8_t>& hostGrid() const = 0;
    // The packed NEE walk grid with axis `fast` fastest, built on first use and kept (tracers
    // that share the medium share it): fills the walk fields of *mv, walkFast = 0 when the fast
    // walk does not apply. Thread-safe.
    int walkGrid(int fast, MediumView* mv) const;
    uint64_t walkBytes() const;
    // The per-layer least-sigt pyramid with layers along `axis` (3D, equally spaced grids), built on
    // first use and kept like the walk grids; axis = -1 in *out when it does not apply (or would
    // exceed 1 GiB). Thread-safe.
    int minPyramid(int axis, MinPyramid* out) const;

protected:
    MediumView v_{};
    uint32_t nmat_ = 0;
    std::vector<double> sigs_, siga_;

private:
    struct WalkCache {
        bool built = false;
        int fast = 0;
        DevBuf<uint32_t> buf;
        uint32_t words = 0, origin = 0, bits = 0, stride[3] = {0, 0, 0};
        float sig[kWalkPalette] = {};
    };
    int buildWalk(int fast, WalkCache* out) const;   // caller holds walkMu_
    mutable std::mutex walkMu_;
    mutable WalkCache walk_[3];
    struct PyrCache {
        bool built = false;
        DevBuf<float> buf;
        MinPyramid view;
    };
    mutable PyrCache pyr_[3];                        // guarded by walkMu_
};

// HeterogeneousMedium (HeterogeneousMedium.cuh:52-233): rectilinear grid of material ids.
class HeterogeneousMedium : public Medium {
public:
    HeterogeneousMedium(int* err, int dims, const uint32_t* materialGrid, const double* xAxis, uint32_t xAxisSize,
                        const double* yAxis, uint32_t yAxisSize, const double* zAxis, uint32_t zAxisSize);
    int setMaterial(const double* sigs, const double* siga, uint32_t materialNum) override;
    const std::vector<uint8_t>& hostGrid() const override { return hostGrid_; }

private:
    DevBuf<float> ef_[3];
    DevBuf<double> ed_[3];
    DevBuf<uint8_t> grid_;
    std::vector<uint8_t> hostGrid_;
};

// ---------------------------------------------------------------- Scattering
class Scattering {
public:
    explicit Scattering(int dims);
    int setIsotropicScattering(uint32_t materialNum);
    int setHenyeyGreensteinScattering(double g, uint32_t materialNum);
    // Scattering::setTabularScattering (Scattering.cuh:272-429): f holds n complex samples
    // (interleaved re, im) of the amplitude over the scattering angle.
    int setTabularScattering(const double* f, uint32_t n, uint32_t materialNum);
    // Scattering::setVonMisesFisherScattering (Scattering.cuh:635-817): M mixtures with
    // concentrations mu, log-scales C and weights alpha.
    int setVonMisesFisherScattering(const double* mu, const double* C, const double* alpha, uint32_t M,
                                    uint32_t materialNum);
    // copies the tabular tables to the device (after the last set*; no-op without tables)
    int upload();
    // Scattering::isIsotropicScattering (Scattering.cuh:1505-1527): HG with |g| < 1e-4 counts.
    bool isIsotropicScattering() const;
    // true when every material in [1, n) has the same phase function
    bool isUniform(uint32_t n) const;
    bool isExactlyIsotropic(uint32_t n) const;
    // amplitude (complex: tabular amplitudes carry a phase) at the scattering cosine
    void amplitude(uint32_t m, double cosang, double* re, double* im) const;
    const ScatterView& view() const { return v_; }
    // the vMF lobes of material m (kappa, C + log alpha); 0 when m is not a vMF mixture
    uint32_t vmfLobes(uint32_t m, const double** kappa, const double** c) const;

private:
    ScatterView v_{};
    bool defined_[kMaxMaterials] = {};
    // tabular tables of every material, host copies (entry order as on the device)
    std::vector<float4> tabF_;
    std::vector<double4> tabD_;
    std::vector<double> tabPh_, tabRaw_;   // tabRaw_: the input samples (uniformity test)
    std::vector<float> tabIcdf_;
    std::vector<double> vmfKappa_, vmfC_, vmfRaw_, vmfPairMu_, vmfPairLogNorm_, vmfPairPdf_, vmfPairCdf_;
    uint32_t vmfRawOff_[kMaxMaterials] = {};
    bool vmfIso_[kMaxMaterials] = {};
    DevBuf<double> dVmf_;   // the vMF arrays, one allocation
    uint32_t rawOff_[kMaxMaterials] = {};
    PhaseTables tab_{};               // sizes and offsets (host), device pointers after upload
    DevBuf<PhaseTables> dTab_;
    DevBuf<float4> dTabF_;
    DevBuf<double4> dTabD_;
    DevBuf<double> dTabPh_;
    DevBuf<float> dTabIcdf_;
};

// ---------------------------------------------------------------- Sources
class Source {
public:
    virtual ~Source() = default;
    const SourceView& view() const { return v_; }
    uint32_t size() const { return v_.n; }
    int type() const { return v_.type; }
    const std::vector<double>& hostPoints() const { return pts_; }

protected:
    SourceView v_{};
    std::vector<double> pts_;   // normalised directions or locations, n x 3 (host)
    DevBuf<float> pf_;
    DevBuf<double> pd_;
};

// FarField (FarField.cuh:8-396): plane waves. For an illumination the sampling table of
// fillBuffers (FarField.cuh:71-127) is built here on the host in FP64.
class FarField : public Source {
public:
    FarField(int* err, const Medium& medium, int dims, const double* x, const double* y, const double* z,
             uint32_t n, bool illumination);

private:
    DevBuf<float> table_, probf_;
    DevBuf<double> probd_;
};

// PointSource (PointSource.cuh:8-216).
class PointSource : public Source {
public:
    PointSource(int* err, int dims, const double* x, const double* y, const double* z, uint32_t n, bool areaSource,
                bool illumination);
};

// GaussianBeam (GaussianBeam.cuh:1026-1613), 3D: beams focused at the given points along the
// given directions with aperture a (gamma_a = 1 / a^2). An illumination also builds the sampling
// tables of updateSampler on the host in FP64 (the attenuation along each beam on the device), the
// axis of its second-point proposal being the mean of `views`' beam directions. Every material
// holding a focal point must be a vMF mixture (at most 16 lobes).
class GaussianBeam : public Source {
public:
    GaussianBeam(int* err, const Medium& medium, const Scattering& scattering, const sstmc_source& s,
                 const sstmc_source* views, double lambda, bool illumination);
    const std::vector<double>& directions() const { return dirs_; }   // beam directions, n x 3 (host)

private:
    std::vector<double> dirs_;
    DevBuf<double> gd_;
    DevBuf<uint32_t> gu_;
};

// ---------------------------------------------------------------- Sampler
// TemporalSampler (TemporalSampler.cuh:55-137, time table :941-1132), or with `correlation` the
// TemporalCorrelationSampler (TemporalSamplerCorrelation.cuh:41-106, 455-594): the ticks are
// then time lags, kept as they are, and every vertex stays at its sampled position.
class TemporalSampler {
public:
    TemporalSampler(int* err, const double* ticks, uint32_t ticksNum, bool isTemporalAmplitude,
                    bool correlation = false);
    int setMaterial(double D, double ux, double uy, double uz, uint32_t materialNum);
    void finalize(uint32_t materialNum);
    const FramesView& view() const { return v_; }
    uint32_t frames() const { return v_.nt; }
    bool correlation() const { return v_.analytic != 0; }
    // some material moves (U != 0): the correlation sampler's g and f then see shifted positions
    bool moving(uint32_t materialNum) const;

private:
    FramesView v_{};
    std::vector<double> dt_;
    DevBuf<float> dtDev_;
    DevBuf<double> ticksDev_;
    std::vector<double> D_, U_;
};

// ---------------------------------------------------------------- Tracer
struct TracerOptions {
    uint32_t batchSize;
    uint32_t fullIteration;
    uint64_t seed;
    double lambda;
    int output;        // SSTMC_OUTPUT_*
    int cbs;           // coherent backscattering (field output only)
    int terms;         // SSTMC_TERMS_*
    int contraction;   // SSTMC_CONTRACT_*
    double tauCut;     // precision.tauCut (0: exact)
    int graphs = 0;    // replay the contraction stage from captured HIP graphs
};

class Tracer {
public:
    virtual ~Tracer() = default;
    // Tracer::trace (TracerInterface.cuh:21): u is host memory, 2*Nl*Nv*Nt doubles.
    virtual int trace(double* u) = 0;
    void updateSeed(uint64_t seed) { seed_ = seed; }
    uint64_t getZeroContributionPaths() const { return zeroContributionPaths_; }
    uint64_t getPhotonsBudget() const { return photonsBudget_; }

protected:
    uint64_t seed_ = 0;
    uint64_t zeroContributionPaths_ = 0;
    uint64_t photonsBudget_ = 0;
};

// NextEventEstimation (NextEventEstimation.cuh:163-1637), batch driver.
class NextEventEstimation : public Tracer {
public:
    NextEventEstimation(int* err, const Medium& medium, const Scattering& scattering, const Source& illumination,
                        const Source& view, const TemporalSampler& sampler, const TracerOptions& opt);
    ~NextEventEstimation() override;
    int trace(double* u) override;
    // Asynchronous form used by the serving loop: begin() clears the accumulators, enqueueBatch()
    // queues one batch on the tracer's three streams (transport, NEE, contraction; consecutive
    // batches alternate between two buffer sets, so the transport of batch b + 1 and the
    // contraction of batch b - 1 overlap the NEE of batch b), finish() waits for the streams,
    // re-runs the batches whose vertices overflowed the pool (after growing it) and copies the
    // field to host or device memory.
    int begin();
    int enqueueBatch(uint32_t b);
    int enqueue();
    int finish(double* u, bool deviceDst = false);
    uint32_t batches() const { return opt_.fullIteration; }
    // device bytes a tracer with these sizes allocates (admission control); `paths` per rendering
    // (batch size times batches) sets the resolution of the far-field view kernel's bound tables
    static uint64_t footprint(uint32_t P, uint32_t L, uint32_t V, uint32_t T, bool fp64, bool frames, uint64_t cells,
                              bool cbs, bool acorr, uint64_t paths);
    // the resolution of those tables: sub x sub x sub boxes a cell, pyramid queries of up to
    // (span + 1)^2 blocks
    static void farTableResolution(uint64_t paths, uint32_t* sub, uint32_t* span);
    // device bytes of one batch set (terms, per-path arrays, a vertex pool of 32 vertices a path)
    static uint64_t setBytes(uint32_t P, uint32_t L, uint32_t V, uint32_t T, bool fp64, bool frames, bool cbs,
                             bool acorr);
    // batch sets of the pipeline: three while a set is small next to the GPU's memory, else two
    static int pipelineSets(uint64_t setBytes) { return setBytes <= (8ull << 30) ? 3 : 2; }
    int debugBatch(uint64_t pathBegin, uint32_t npaths, uint64_t* nvtxOut, uint32_t* pathK, uint32_t* pathMisses,
                   uint64_t* pathOff, float* vtxPos, uint32_t* vtxMat, float* vtxRho, float* vtxXt, double* vtxLen,
                   double* termA, double* termB, double* termSa, double* termSb, double* uBatch);
    int debugContract(uint32_t npaths, const double* a, const double* b, double* u);
    // the last batch's reference of the relative attenuation cut
    int tauRef(float* out) const;
    // coherent-backscattering terms (Ar [p][t][v], Br [p][t][l]) of the last debugBatch of n paths
    int debugCbsTerms(uint32_t n, double* termAr, double* termBr) const;
    void timing(sstmc_timing* out) const;
    hipStream_t stream() const { return stream_; }
    // the stream a batch's last stage (its contraction) runs on
    hipStream_t doneStream() const { return streamC_; }
    uint32_t L() const { return L_; }
    uint32_t V() const { return V_; }
    uint32_t T() const { return T_; }
    // device bytes held now: buffers, vertex pool, walk grid and split-K workspaces
    uint64_t deviceBytes() const { return bytes_ + wsMulti_.bytes + wsSingle_.bytes; }
    uint64_t paths() const { return (uint64_t)P_ * opt_.fullIteration; }

private:
    int allocate();
    int growPool(uint64_t cap);
    int sizePool();
    // Per-batch device state. Batch b uses set b mod nsets_: its per-path arrays, vertex pool,
    // term slabs and cut reference, so that batches can be in different stages at once.
    struct BatchSet {
        DevBuf<uint32_t> k, misses, off, blockSums;
        DevBuf<uint32_t> rowNz;               // VertexBatch::rowNz (viewFirst)
        DevBuf<float4> xm;
        DevBuf<float> rho, delta;
        DevBuf<VertexRec> rec;
        DevBuf<FarRec> far;                   // the far-field view kernel's records (farRecs_)
        DevBuf<uint32_t> live;                // and their live-block masks
        DevBuf<float> tauRef;                 // the batch's least view optical depth (relative cut)
        DevBuf<unsigned char> A, B, a, b;
        DevBuf<unsigned char> Ar, Br;         // coherent backscattering: reversed-path terms
        // coherent backscattering under the TemporalCorrelationSampler: first-vertex u1 / u2
        // terms [t][p][l|v] and per-vertex u1 / u2 terms [t][vertex][l|v] (NeeOut)
        DevBuf<unsigned char> F1l, F2l, F1v, F2v, Gl1, Gl2, Gv1, Gv2;
        hipEvent_t transported = nullptr;     // transport done (transport stream)
        hipEvent_t traced = nullptr;          // NEE done (NEE stream)
        hipEvent_t released = nullptr;        // contraction done: the set may be reused
        bool used = false;
    };
    VertexBatch vertexBatch(int set) const;
    int buildWalkGrid();
    void fetchTerms(const DevBuf<unsigned char>& src, uint32_t n, uint32_t width, double* dst) const;
    uint32_t lastDebugPaths_ = 0;
    int lastSet_ = 0;
    // pipelined: the three stages on their own streams, ordered by the set's events; otherwise
    // everything on stream_ (parity inspection)
    int runBatch(uint64_t pathBegin, uint32_t npaths, double* U, bool timed, uint32_t b, int set, bool pipelined);
    int contractSingle(uint32_t npaths, double* U, int set, hipStream_t st);
    int syncAll() const;
    int orderAfterNee();
    hipEvent_t cleared_ = nullptr;   // begin()'s clears on the NEE stream

    const Medium& medium_;
    MediumView mview_{};            // the medium's view plus the walk grid for this tracer's views
    const Scattering& sc_;
    const Source& il_;
    const Source& vw_;
    const TemporalSampler& sampler_;
    TracerOptions opt_;
    int dims_;
    uint32_t L_, V_, T_, P_;
    int singleMode_;
    double fscalar_ = 1.0;
    bool sameViews_ = false;        // correlation sampler: view2 holds the same sources as view
    bool sameIllums_ = false;       // correlation sampler: illumination2 holds the same sources
    bool cbsCorr_ = false;          // coherent backscattering under the correlation sampler
    bool farRecs_ = false;          // the far-field view kernel runs: FarRec per (vertex, frame)
    FarBound farBound_{};           // its source order and per-cell walk bounds (k_far_live)
    DevBuf<uint32_t> farPerm_[2];   // the kernel's source order per role (FarBound::perm)
    DevBuf<float> farTable_[2];     // FarBound::table per role
    bool farBoundOn_ = false;
    // the far-field view kernel runs before k_nee, which then skips the illumination terms of
    // paths whose view-side rows are zero (NeeParams::viewFirst): FP32 far-field views without
    // coherent backscattering, field output, bf16x3 contraction (the compacted one), separable
    // single scattering
    bool viewFirst_ = false;
    int setupFarBound();
    // the contraction stage captured as HIP graphs, per (set, batch size, destination): the
    // multiple-scattering part and the single-scattering part (timed apart); dropped whenever a
    // buffer they name is reallocated
    struct StageGraph {
        int set;
        uint32_t npaths;
        double* U;
        hipGraphExec_t multi = nullptr, single = nullptr;
        void *wsMulti = nullptr, *wsSingle = nullptr;
    };
    std::vector<StageGraph> graphs_;
    uint32_t graphReplays_ = 0;   // of the current rendering
    void dropGraphs();
    int contractStage(uint32_t npaths, double* U, int set, hipStream_t sC, hipEvent_t mid);
    int contractMulti(uint32_t npaths, double* U, int set, hipStream_t sC);
    DevBuf<double2> cbsPart_;       // its mixed pairings' split slices
    bool fp64_;
    int contractKind_;
    hipStream_t stream_ = nullptr;    // NEE (and every stage when not pipelined)
    hipStream_t streamT_ = nullptr;   // transport
    hipStream_t streamC_ = nullptr;   // contraction
    uint64_t bytes_ = 0;

    static constexpr int kMaxSets = 3;
    BatchSet sets_[kMaxSets];
    // staging of the single-pass transport, shared by the sets (every transport of a tracer runs
    // on one stream, or with the streams idle)
    struct Staging {
        DevBuf<float4> xm;
        DevBuf<float> rho, delta;
        DevBuf<uint32_t> chunkSlot, chunkJ0, q;
        uint32_t nchunks = 0;
    } stage_;
    TransportStage transportStage() const;
    int nsets_ = 2;                  // sets in use (pipelineSets)
    int setOf(uint32_t b) const { return (int)(b % (uint32_t)nsets_); }
    DevBuf<uint32_t> status_;
    DevBuf<uint32_t> overflow_;           // bit b: batch b's vertices did not fit the pool
    DevBuf<unsigned long long> stepSlots_;   // voxel steps: k_nee's counter set, the view kernels' (slotAdd)
    DevBuf<unsigned long long> totals_;   // [0] misses, [1] vertices, [3] largest batch
    uint64_t cap_ = 0;
    bool poolSized_ = false;   // sizePool ran: the pool holds a batch of this request with margin
    DevBuf<double> U_;
    DevBuf<float2> Ff_;             // single-scattering F (complex) when separable
    DevBuf<double2> Fd_;
    Workspace wsMulti_, wsSingle_;
    // per batch: 0 transport start, 1 transport end (transport stream); 6 NEE start, 5 between
    // k_nee and k_nee_view, 2 NEE end (NEE stream); 7 contraction start, 3 after the
    // multiple-scattering contraction, 4 after the single-scattering one (contraction stream)
    static constexpr size_t kEventsPerBatch = 8;
    // device totals: 0 missed attempts, 1 vertices, 3 largest batch (vertices); the voxel steps are
    // in stepSlots_
    static constexpr size_t kTotals = 5;
    int zeroCounters(hipStream_t st);
    void stepCounts(uint64_t* nee, uint64_t* view) const;
    std::vector<hipEvent_t> ev_;
    uint32_t evBatches_ = 0;
    sstmc_timing last_{};
};

}  // namespace sstmc
