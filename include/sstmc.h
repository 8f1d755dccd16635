/*
 * sstmc.h -- C ABI of the MI355X-native next-event-estimation speckle tracer.
 *
 * This is the drop-in boundary for the reference's `sstmc2D(config)` / `sstmc3D(config)`
 * MEX entry (MatlabInterface.cu:1565-1825) and for the C++ plugin split below it
 * (Tracer / Sampler / Source / Medium / Scattering, TracerInterface.cuh:9-50,
 * SamplerInterface.cuh:11-176, SourceInterface.cuh:10-179, MediumInterface.cuh:39-135,
 * Scattering.cuh:32-125).  Every argument is a plain pointer or size; nothing here names a
 * torch or HIP type, so a MEX shim, ctypes, cgo or JNI binding can call it directly.
 *
 * Conventions kept from the reference:
 *   - the config mirrors `config.simulation/medium/scattering/illumination/view/sampler/tracer`
 *     field by field (README.md:59-349, parsed at MatlabInterface.cu:274-1563);
 *   - material 0 is "outside the grid"; grid ids are 1..materialNum-1 (MatlabInterface.cu:975-992);
 *   - a single scattering cell is duplicated into material 1 (MatlabInterface.cu:750-760);
 *   - output u is complex, laid out u(l, v, t) with l fastest (MatlabInterface.cu:1731-1753),
 *     written here as interleaved (re, im) doubles; normalisation by sqrt(iter) (field) or
 *     iter (correlation) stays with the caller (README.md:365-373);
 *   - iter = batchSize * fullIteration + missed first-vertex samples (MatlabInterface.cu:1796);
 *   - error codes mirror ErrorType (Simulation.cuh:111-162) and are returned, never thrown.
 *
 * Deliberate differences (each also stated where it is implemented):
 *   - random numbers come from a counter-based Philox4x32-10 stream keyed by (seed, path id)
 *     instead of per-thread XORWOW states (TracerInterface.cuh:56-63), so every path is
 *     reproducible alone, in any batch and in any request mix;
 *   - no process-global device state: any number of handles may be live at once.
 */
#ifndef SSTMC_H
#define SSTMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- error codes */
/* Mirrors ErrorType (Simulation.cuh:111-162); extra codes are appended after MISC_ERROR. */
enum {
    SSTMC_NO_ERROR = 0,
    SSTMC_DEVICE_ERROR = 1,
    SSTMC_ALLOCATION_ERROR = 2,
    SSTMC_KERNEL_ERROR = 3,
    SSTMC_BLAS_ERROR = 4,
    SSTMC_NOT_SUPPORTED = 5,
    SSTMC_MISC_ERROR = 6,
    SSTMC_INVALID_CONFIG = 7,      /* config validation (MatlabInterface.cu mexErrMsgIdAndTxt sites) */
    SSTMC_NO_HIT = 8,              /* first-vertex sampling exceeded its retry budget */
    SSTMC_BUDGET_EXCEEDED = 9,     /* serving: request larger than the admission budget */
    SSTMC_NOT_READY = 10,          /* serving: request still in flight */
    SSTMC_SCRATCH_EXCEEDED = 11    /* sstmc_contract_noalloc: partials exceed its fixed scratch */
};

/* ---------------------------------------------------------------- enums */
enum { SSTMC_MEDIUM_HETEROGENEOUS = 1 };                 /* config.medium.type */
enum { SSTMC_SCATTERING_ISOTROPIC = 1,                    /* config.scattering.type{m} */
       SSTMC_SCATTERING_TABULAR = 2,
       SSTMC_SCATTERING_HG = 3,
       SSTMC_SCATTERING_VMF = 4 };                        /* values follow ScatteringNS::ScatteringType */
enum { SSTMC_SOURCE_NONE = 0,                             /* config.illumination/view.type */
       SSTMC_SOURCE_POINT = 1,
       SSTMC_SOURCE_FARFIELD = 2,
       SSTMC_SOURCE_GAUSSIAN_BEAM = 3 };                  /* 3D; both illumination and view */
enum { SSTMC_SAMPLER_TEMPORAL = 1,                        /* config.sampler.type: TemporalSampler */
       SSTMC_SAMPLER_TEMPORAL_CORRELATION = 2 };          /* TemporalCorrelationSampler */
enum { SSTMC_TRACER_NEE = 1 };                            /* config.tracer.type */
enum { SSTMC_OUTPUT_FIELD = 0,                            /* u(l,v,t): the field per frame */
       SSTMC_OUTPUT_CORRELATION = 1 };                    /* C(l,v,d) = sum u(t=0) conj(u(t=d)), per-sub-path estimator */
enum { SSTMC_TERMS_FP32 = 0,                              /* path terms (NEE attenuation, g, f) in FP32; */
       SSTMC_TERMS_FP64 = 1 };                            /* path length and phase are FP64 in both */
enum { SSTMC_CONTRACT_BF16X3 = 0,                         /* hi+lo split BF16 MFMA, FP32 accumulate */
       SSTMC_CONTRACT_FP32 = 1,                           /* plain FP32 FMA tiles */
       SSTMC_CONTRACT_FP64 = 2 };                         /* FP64 FMA tiles (accuracy reference) */

/* ---------------------------------------------------------------- config */
/* config.simulation (MatlabInterface.cu:1459-1563, 1590-1671). Zero means "reference default". */
typedef struct sstmc_simulation {
    uint32_t batchSize;        /* P, default 1024 (MatlabInterface.cu:1467) */
    uint32_t fullIteration;    /* default 1 */
    uint32_t renderingsNum;    /* default 1; rendering i uses seed + i (MatlabInterface.cu:1799) */
    uint32_t deviceNumber;     /* HIP device ordinal */
    uint64_t seed;             /* config.simulation.seed (MatlabInterface.cu:1649) */
    double lambda;             /* default 1.0; one wavelength (Nw = 1) */
    uint32_t graphs;           /* nonzero: each batch's contraction stage is captured once as a HIP
                                  graph and replayed for the later batches (same results) */
} sstmc_simulation;

/* config.medium, heterogeneous rectilinear grid (HeterogeneousMedium.cuh:107-233). */
typedef struct sstmc_medium {
    int32_t type;                  /* SSTMC_MEDIUM_HETEROGENEOUS */
    uint32_t xAxisSize, yAxisSize, zAxisSize;   /* number of bin edges; zAxisSize ignored in 2D */
    const double* xAxis;
    const double* yAxis;
    const double* zAxis;
    const uint32_t* materialGrid;  /* (nx-1)(ny-1)(nz-1) ids, x fastest */
    uint32_t materialNum;          /* entries of sigs/siga, including material 0 */
    const double* sigs;
    const double* siga;            /* may be NULL: all zero (MatlabInterface.cu:1019-1022) */
} sstmc_medium;

/* config.scattering (MatlabInterface.cu:712-888): one entry per material. */
typedef struct sstmc_scattering {
    uint32_t materialNum;
    const int32_t* type;           /* SSTMC_SCATTERING_ISOTROPIC, _TABULAR, _HG or _VMF */
    const double* g;               /* HG anisotropy per material (ignored otherwise) */
    /* Tabular amplitude functions (Scattering.cuh:272-429): tabularSize[m] complex samples of
     * the amplitude over the scattering angle, equally spaced on [0, pi] (3D) or [0, 2 pi] (2D),
     * for each material m with tabularSize[m] > 0, concatenated in material order as interleaved
     * (re, im) doubles. NULL when no material is tabular. */
    const uint32_t* tabularSize;
    const double* tabular;
    /* von Mises-Fisher mixtures (Scattering.cuh:613-884): vmfMixtures[m] mixtures for each
     * material m with vmfMixtures[m] > 0; their mixtureMu (kappa), mixtureC and mixtureAlpha
     * concatenated in material order. NULL when no material is vMF. In 2D a vMF material scatters
     * isotropically (Scattering.cuh:898-903, 1094-1099). */
    const uint32_t* vmfMixtures;
    const double* vmfMu;
    const double* vmfC;
    const double* vmfAlpha;
} sstmc_scattering;

/* config.illumination / config.view (MatlabInterface.cu:1062-1215). */
typedef struct sstmc_source {
    int32_t type;                  /* SSTMC_SOURCE_FARFIELD, _POINT or _GAUSSIAN_BEAM */
    uint32_t num;                  /* number of directions / locations / beams */
    const double* x;               /* direction.x, location.x or focalPoint.x */
    const double* y;
    const double* z;               /* ignored in 2D */
    int32_t isAreaSource;          /* PointSource only (PointSource.cuh:140-147) */
    /* GaussianBeam only (MatlabInterface.cu:1153-1190, GaussianBeam.cuh:1026-1196): the beam
     * directions focalDirection.{x,y,z} (unit vectors; dirZ ignored in 2D) and the aperture */
    const double* dirX;
    const double* dirY;
    const double* dirZ;
    double aperture;
} sstmc_source;

/* config.sampler (MatlabInterface.cu:1217-1368). TemporalSampler: t are frame times and u(l,v,t)
 * the field per frame. TemporalCorrelationSampler (TemporalSamplerCorrelation.cuh): t are time
 * lags tau and u(l,v,t) the analytic correlation sum u1(-tau/2) conj(u2(+tau/2)) between the
 * illumination/view pair and the illumination2/view2 pair. */
typedef struct sstmc_sampler {
    int32_t type;                  /* SSTMC_SAMPLER_TEMPORAL or SSTMC_SAMPLER_TEMPORAL_CORRELATION */
    uint32_t tNum;                 /* frames (or lags) Nt */
    const double* t;               /* frame times (or lags) */
    const double* D;               /* Brownian coefficient per material */
    const double* Ux;              /* linear velocity per material */
    const double* Uy;
    const double* Uz;
    int32_t isTemporalAmplitude;   /* TemporalSampler, default 1: g/f use displaced positions */
} sstmc_sampler;

/* config.tracer (MatlabInterface.cu:1374-1457). */
typedef struct sstmc_tracer {
    int32_t type;                  /* SSTMC_TRACER_NEE */
    int32_t isCBS;                 /* coherent backscattering: field CBS (TemporalSampler) or the
                                      correlation pairings CBS01-CBS04 (TemporalCorrelationSampler) */
    int32_t output;                /* SSTMC_OUTPUT_FIELD or SSTMC_OUTPUT_CORRELATION */
} sstmc_tracer;

/* Precision / performance policy of this implementation (no counterpart in the reference,
 * which is FP64 throughout, Simulation.cuh:22-24). Transport is always FP32. */
typedef struct sstmc_precision {
    int32_t terms;                 /* SSTMC_TERMS_FP32 (default) or SSTMC_TERMS_FP64 */
    int32_t contraction;           /* SSTMC_CONTRACT_BF16X3 (default), _FP32 or _FP64 */
    /* Error-budgeted, batch-relative attenuation cut: per batch, tau_ref is the least optical
     * depth from a sample of the batch's vertices to the views; a view-side term that carries no
     * path weight (b, B: g of one view at one vertex) whose optical depth exceeds tau_ref + tauCut
     * is dropped and its voxel walk stops there, so every dropped term is below exp(-tauCut / 2)
     * times the batch's least attenuated view term. Terms that carry the path weight (a, A, Ar),
     * the reversed paths' Br and the path probabilities stay exact. 0 (default) is exact: walks
     * run to the box (FP32 walks stop at 210, past which exp(-tau/2) is exactly 0 in FP32). The
     * reference walks every term to the box (HeterogeneousMedium.cuh:386-573). */
    double tauCut;
} sstmc_precision;

typedef struct sstmc_config {
    uint32_t dims;                 /* 2 -> sstmc2D, 3 -> sstmc3D */
    sstmc_simulation simulation;
    sstmc_medium medium;
    sstmc_scattering scattering;
    sstmc_source illumination;
    sstmc_source view;
    sstmc_sampler sampler;
    sstmc_tracer tracer;
    sstmc_precision precision;
    /* config.illumination2 / config.view2 (MatlabInterface.cu:1062-1090, 1313-1331): the u2
     * sources of the TemporalCorrelationSampler, of the same type, count and area flag as
     * illumination / view; type SSTMC_SOURCE_NONE (zero) when absent. */
    sstmc_source illumination2;
    sstmc_source view2;
} sstmc_config;

typedef struct sstmc_handle sstmc_handle;

/* ---------------------------------------------------------------- tracer handle */
/* Builds Simulation + Scattering + Medium + Sources + Sampler + Tracer, as mexFunction does
 * (MatlabInterface.cu:1671-1700). Replaces the constructors' `ErrorType* err` out-parameters. */
int sstmc_create(const sstmc_config* cfg, sstmc_handle** out);

/* Tracer::trace (TracerInterface.cuh:21, NextEventEstimation.cuh:1461-1637): one rendering of
 * fullIteration batches. u: 2*Nl*Nv*Nt doubles (host), iter: one double (may be NULL). */
int sstmc_trace(sstmc_handle* h, double* u, double* iter);

/* Tracer::updateSeed (TracerInterface.cuh:136-152). */
int sstmc_update_seed(sstmc_handle* h, uint64_t seed);

/* Tracer::getZeroContributionPaths / getPhotonsBudget (TracerInterface.cuh:23-24) of the
 * last trace: missed first-vertex samples, and the number of scattering vertices traced. */
uint64_t sstmc_zero_contribution_paths(const sstmc_handle* h);
uint64_t sstmc_photons_budget(const sstmc_handle* h);

/* Sizes of the handle's problem: Nl, Nv, Nt (sampler->getIllumination/View/SamplerSize). */
int sstmc_sizes(const sstmc_handle* h, uint32_t* nl, uint32_t* nv, uint32_t* nt);

void sstmc_destroy(sstmc_handle* h);

/* getErrorMsg (MatlabInterface.cu:24-272). Static string, never NULL. */
const char* sstmc_error_string(int code);

/* The config checks of sstmc_create alone (the MEX argument validation, MatlabInterface.cu:
 * 712-1563): returns the code sstmc_create would return for a config error, SSTMC_NO_ERROR for a
 * config it would accept. Touches no device; host entries use it to report errors early. */
int sstmc_validate(const sstmc_config* cfg);

/* The whole MEX call: `[u, iter] = sstmc{2,3}D(config)` (MatlabInterface.cu:1565-1825).
 * u: 2*Nl*Nv*Nt*renderingsNum doubles, iter: renderingsNum doubles. */
int sstmc_render(const sstmc_config* cfg, double* u, double* iter);

/* ---------------------------------------------------------------- stage timing */
/* Device time of each stage of the last trace, from HIP events on the tracer's stream, in ms.
 * Mirrors the TIME_REC TimeTracer report (NextEventEstimation.cuh:107-156). */
typedef struct sstmc_timing {
    double transport_ms;           /* free flight + direction sampling + frames (a), then the
                                      per-vertex NEE set-up (attenuation cut reference, vertex
                                      records, walk bounds) on the same stream */
    double nee_ms;                 /* NEE attenuation, g, f, probabilities, phase (b) */
    double contract_ms;            /* complex contraction (c) */
    double single_ms;              /* single-scattering contraction / fused kernel */
    double total_ms;               /* first event to last event */
    uint64_t vertices;             /* scattering vertices traced */
    uint64_t dda_steps;            /* voxel steps walked by the NEE attenuation */
    uint32_t batches;
    double nee_view_ms;            /* part of nee_ms: multiple-scattering view terms (the voxel-walk
                                      kernel that dominates nee_ms) */
    uint64_t dda_steps_view;       /* part of dda_steps walked by that kernel */
    uint32_t graph_replays;        /* batches whose contraction stage replayed captured graphs
                                      (simulation.graphs) */
    uint32_t pipeline_sets;        /* buffer sets the batch pipeline cycles through (batch b uses
                                      set b mod pipeline_sets) */
} sstmc_timing;
int sstmc_last_timing(const sstmc_handle* h, sstmc_timing* out);

/* ---------------------------------------------------------------- parity / inspection */
/* Traces one batch of P = npaths paths with global path ids [path_begin, path_begin+npaths)
 * and returns what the contraction consumes, for comparison with a CPU transcription.
 * Sizes are returned first by calling with all output pointers NULL (nvtx_out filled).
 *  path_k[P]        scattering vertices per path (x1..xK)
 *  path_misses[P]   missed first-vertex attempts
 *  path_off[P+1]    vertex offsets
 *  vtx_pos[nvtx*3]  central positions (FP32 transport)
 *  vtx_mat[nvtx]    material ids
 *  vtx_rho[nvtx]    random phase of the vertex, in [0,1)
 *  vtx_dxt[nvtx*Nt*3] displacement of the scatterer per frame (frame position = pos + dxt,
 *                     exact in FP64)
 *  vtx_len[nvtx*Nt]   path length from x1 per frame, FP64 (as accumulated by the NEE kernel)
 *  term_a[P*Nt*Nl*2]  multiple-scattering illumination terms A_p[l,t] (interleaved complex)
 *  term_b[P*Nt*Nv*2]  multiple-scattering view terms B_p[v,t]
 *  term_sa[P*Nt*Nl*2] single-scattering illumination factor g_l(x1) w1 exp(i 2 pi rho_1)
 *  term_sb[P*Nt*Nv*2] single-scattering view factor g_v(x1)
 *  u_batch[2*Nl*Nv*Nt] the batch's contribution to u (single + multiple, this handle's contraction)
 * Term layouts are [p][t][l] and [p][t][v]; any output pointer may be NULL. */
int sstmc_debug_batch(sstmc_handle* h, uint64_t path_begin, uint32_t npaths,
                      uint64_t* nvtx_out,
                      uint32_t* path_k, uint32_t* path_misses, uint64_t* path_off,
                      float* vtx_pos, uint32_t* vtx_mat, float* vtx_rho,
                      float* vtx_dxt, double* vtx_len,
                      double* term_a, double* term_b, double* term_sa, double* term_sb,
                      double* u_batch);

/* Coherent-backscattering terms of the last sstmc_debug_batch call (same npaths), for a handle
 * created with tracer.isCBS: the reversed path v -> x1 -> ... -> xK -> l
 * (NextEventEstimation.cuh:1212-1235, 1432-1443).
 *  term_ar[P*Nt*Nv*2]  view connected at x1: g_v(x1) f_v(x1 -> x2) w2 (zero when K < 2)
 *  term_br[P*Nt*Nl*2]  sum_{k>=2} g_l(x_k) f_l(x_k -> x_{k-1}) exp(i 2 pi (rho_k + L_k / lambda))
 * The multiple-scattering field is (sum_p A^T B + sum_p Br^T Ar) / sqrt(2).
 * Under the TemporalCorrelationSampler the two hold CBS04's halves, Gv0 conj(u2Gv0) C4_v1v2 at x1
 * (with the path weight) and sum_{k>=2} Gl conj(u2Gl) C4_l1l2 at x_k (NextEventEstimation.cuh:
 * 1332-1378); the multiple-scattering correlation is (sum_p A^T B + sum_p Br^T Ar + CBS02 +
 * CBS03) / 2, the mixed pairings CBS02 + CBS03 formed per (l, v) on the device. */
int sstmc_debug_cbs(sstmc_handle* h, uint32_t npaths, double* term_ar, double* term_br);

/* The reference optical depth of the relative attenuation cut (precision.tauCut) in the last
 * batch traced (sstmc_debug_batch or the last batch of sstmc_trace): the least view optical
 * depth its sample found; the cut dropped the view-side terms whose depth exceeded
 * tau_ref + tauCut. SSTMC_INVALID_CONFIG when the handle has no cut. */
int sstmc_debug_tau_ref(const sstmc_handle* h, float* tau_ref);

/* Contracts caller-supplied terms on the device with the handle's contraction precision:
 * u[t][v][l] += sum_p a[p][t][l] * b[p][t][v] (complex, interleaved doubles in and out).
 * Exposes kernel (c) alone for accuracy tests against an FP64 GEMM. */
int sstmc_debug_contract(sstmc_handle* h, uint32_t npaths, const double* a, const double* b,
                         double* u);

/* Allocation-free contraction. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It runs the contraction of sstmc_debug_contract, (c), on terms the caller already
 * holds on the current device and adds into the caller's device accumulator:
 *   u[t][v][l] += sum_p a[t][p][l] * b[t][p][v]   (complex, interleaved)
 * a, b: complex64 terms (interleaved floats) for SSTMC_CONTRACT_BF16X3 and _FP32, complex128 for
 * _FP64; u: complex128. Enqueued on `stream` (a hipStream_t; NULL: the default stream), not
 * synchronised. It allocates nothing: its split-K partials live in a fixed 64 MiB array per device
 * (held from the device's first call until the process ends); a problem whose L x V partials of
 * one split exceed it (L * V > 8 Mi complex64, 4 Mi complex128) returns SSTMC_SCRATCH_EXCEEDED,
 * and larger path counts take fewer split slices. Calls that overlap on one device share that
 * array and race: serialise them (one stream, or events). */
int sstmc_contract_noalloc(int32_t contraction, uint32_t npaths, uint32_t L, uint32_t V, uint32_t nt,
                           const void* a, const void* b, double* u, void* stream);

/* ---------------------------------------------------------------- serving */
/* Device-resident request loop with continuous batching (no counterpart in the reference,
 * whose `__constant__` globals allow one live simulation per process, Simulation.cuh:168-173).
 * Requests are admitted by HBM footprint; results stay on the device until fetched. */
typedef struct sstmc_server sstmc_server;
typedef struct sstmc_server_stats {
    uint64_t completed;
    double p50_ms, p99_ms, mean_ms;     /* submit -> completion latency */
    double paths_per_s;                 /* completed paths / busy wall time */
    double contributions_per_s;         /* completed paths*L*V*T / busy wall time */
    uint64_t hbm_in_use, hbm_budget;
} sstmc_server_stats;

int sstmc_server_create(uint32_t device, uint64_t hbm_budget_bytes, uint32_t max_inflight,
                        sstmc_server** out);
/* Builds the medium of cfg (grid, cross sections and its NEE walk grids) on the server's device
 * and keeps it in the server's medium cache, so that requests on the same medium (equal edges,
 * grid and cross sections) start without copying or building it. Requests also add the media
 * they build to the cache (at most 16 media; the least recently used one no request holds is
 * dropped). No counterpart in the reference. */
int sstmc_server_preload(sstmc_server* s, const sstmc_config* cfg);
/* Queues cfg: renderingsNum renderings of fullIteration batches each, rendering i at seed + i, as
 * sstmc_render does (MatlabInterface.cu:1797-1800); returns a request id. The config is copied
 * whole (every array it points into), so the caller may free or reuse its arrays on return. */
int sstmc_submit(sstmc_server* s, const sstmc_config* cfg, uint64_t* request_id);
/* Advances the loop without blocking; returns SSTMC_NO_ERROR when the request is done,
 * SSTMC_NOT_READY otherwise. */
int sstmc_poll(sstmc_server* s, uint64_t request_id);
/* Blocks until the request is done. */
int sstmc_wait(sstmc_server* s, uint64_t request_id);
/* Copies the finished result and releases its device memory. u: 2*Nl*Nv*Nt*renderingsNum doubles
 * laid out u(l, v, t, i) as sstmc_render writes them; iter (may be NULL): renderingsNum doubles. */
int sstmc_fetch(sstmc_server* s, uint64_t request_id, double* u, double* iter);
int sstmc_server_stats_get(sstmc_server* s, sstmc_server_stats* out);
void sstmc_server_destroy(sstmc_server* s);

#ifdef __cplusplus
}
#endif
#endif /* SSTMC_H */
