// kernels.h -- launch entry points of the HIP kernels (one translation unit per stage).
#pragma once

#include "common.h"

namespace sstmc {

// NEE outputs of one batch (element type float2 for FP32 terms, double2 for FP64 terms).
// Layouts are frame-major so each frame is a (P x L) / (P x V) row-major operand of the
// contraction: A[t][p][l], B[t][p][v], a[t][p][l], b[t][p][v].
// With the TemporalCorrelationSampler each term is already a u1 conj(u2) product at lag t
// (A: with 1 / p2 and the decorrelation at x1; B: summed over k >= 2 with the decorrelation at x_k
// and at the inner vertices; a: g_l1 conj(g_l2) / p1; b: g_v1 conj(g_v2)).
struct NeeOut {
    void* A;          // multiple-scattering illumination terms (or A(0) conj A(t) for correlation)
    void* B;          // multiple-scattering view terms (or sum_k G_k(0) conj G_k(t))
    void* a;          // single-scattering illumination terms g_l * w1 * exp(i 2 pi rho_1)
    void* b;          // single-scattering view terms g_v
    // coherent backscattering (null when off): the reversed path v -> x1 -> ... -> xK -> l
    void* Ar;         // [t][p][v] view connected at x1: g_v(x1) f_v(x1 -> x2) w2
    void* Br;         // [t][p][l] illumination connected at x_k, k >= 2, summed with the path phase
    unsigned long long* steps;   // voxel steps of the NEE walks (accumulated, slotAdd): k_nee adds to
                                 // the counter set at steps, the view kernels to the one after it
    // TemporalCorrelationSampler with coherent backscattering (null otherwise; Ar then holds
    // Gv0 conj(u2Gv0) C4_v1v2 and Br sum_k Gl conj(u2Gl) C4_l1l2): the first vertex's u1 and u2
    // terms of either source kept apart, F1l / F2l [t][p][l] (u1 with the path weight), F1v / F2v
    // [t][p][v]; and the u1 / u2 terms of every vertex k >= 1 without the decorrelation, views
    // Gv1 / Gv2 [t][vertex][v], illuminations Gl1 / Gl2 [t][vertex][l] (vertex < gcap, u1 without
    // the inner momentum transfer, neither with the path phase, which cancels in every pairing)
    void *F1l = nullptr, *F2l = nullptr, *F1v = nullptr, *F2v = nullptr;
    void *Gv1 = nullptr, *Gv2 = nullptr, *Gl1 = nullptr, *Gl2 = nullptr;
    uint64_t gcap = 0;
};

// How single scattering is contracted (decided on the host from the source/scattering types).
enum SingleMode : int {
    kSingleSeparable = 0,  // S = (a^T b) o F with F a fixed L x V matrix or scalar
    kSingleGeneric = 1,    // S_p[l,v] needs f(l -> x1 -> v) per path (fused kernel)
    kSingleCorrelation = 2, // TemporalCorrelationSampler: f1 conj(f2) times the decorrelation at x1
    kSingleGauss = 3       // GaussianBeam sources: fsGaussianBeam per (l, v, path)
};

struct NeeParams {
    double lambda;
    int output;            // 0 field, 1 correlation
    int singleMode;
    // Walks stop once tau passes a cut past which the term they form is exactly zero: FP32 210
    // (exp(-tau/2) is 0 in FP32 beyond it), FP64 never. With precision.tauCut = delta > 0 the
    // unweighted view-side terms (b, B: g of a view) use the relative cut tauRef[0] + delta, where
    // tauRef[0] is the batch's least view optical depth found by k_tau_ref; terms that carry the
    // path weight (a, A, Ar), the reversed path's Br and the path probabilities stay exact.
    float exactCutF;
    const float* tauRef;   // device scalar, or null (no relative cut)
    float tauCutRel;       // delta
    uint32_t npaths;
    uint32_t lpp;          // k_nee: lanes per path (power of two, 8..64), set by launchNee
    uint32_t lppView;      // k_nee_view: lanes per work unit (power of two, 8..64), set by launchNee
    uint32_t viewBlocks;   // view kernels: blocks of lppView sources per path and frame chunk
    uint32_t frameChunks;  // view kernels: frame chunks per path (field output: one frame each,
                           // correlation: up to 16 frames each)
    uint32_t sameViews;    // TemporalCorrelationSampler: view2 holds the same sources as view
    uint32_t sameIllums;   // TemporalCorrelationSampler: illumination2 holds the same sources
    // the far-field view kernel's view order (FarBound::perm[0]) when its records and live-block
    // masks exist (k_nee takes its single-scattering view walks in blocks of it), else null
    const uint32_t* farPerm = nullptr;
    // 1: the far-field view kernel runs before k_nee and flags the paths whose B row is nonzero
    // (VertexBatch::rowNz), and k_nee forms b first and leaves a and A unwritten for the paths whose
    // B and b rows are both zero (the compacted bf16x3 contraction reads no other rows)
    uint32_t viewFirst = 0;
};

// The single-pass transport of npaths path slots (zeroes ts.q, then k_transport): per-path
// vertex counts and misses into vb, the vertices into the staging; launchTransportPack copies them
// into vb's pool once vb.off holds the exclusive scan of the counts (launchScan).
hipError_t launchTransport(int dims, const MediumView& m, const ScatterView& sc, const SourceView& il,
                           const FramesView& fr, const VertexBatch& vb, const TransportStage& ts, uint64_t seed,
                           uint64_t pathBegin, uint32_t npaths, hipStream_t st);
hipError_t launchTransportPack(const VertexBatch& vb, const TransportStage& ts, uint32_t npaths, uint32_t nt,
                               int frames, hipStream_t st);
hipError_t launchScan(const uint32_t* k, uint32_t* off, uint32_t* blockSums, uint32_t n, hipStream_t st);
// Adds a batch's misses and vertices to acc[0], acc[1] and its vertex count to the maximum in
// acc[3]; a batch whose vertices exceed `cap` (or the staging) is flagged in the `overflow`
// bitmap and status[0] instead.
hipError_t launchBatchTotals(const uint32_t* misses, uint32_t* off, uint32_t n, unsigned long long* acc,
                             uint64_t cap, uint32_t* overflow, uint32_t batch, float* tauRef, const TransportStage& ts,
                             uint32_t* status, hipStream_t st);

// The NEE stage in two parts. kNeeVertexPart reads the batch's vertices only: k_tau_ref (when
// p.tauRef is set), k_vertex_prep (per-vertex records) and, with vb.far and a table in fb,
// k_far_live (which blocks of the far-field view kernel's sources each record's walks can reach
// inside the cut, from the cell bounds). kNeeTermPart forms the terms: k_nee (or k_nee_corr) for
// the first vertex, then the view kernels for vertices k >= 2; `between`, when not null, is
// recorded on `st` between the two.
enum : int { kNeeRefPart = 1, kNeeRecordPart = 2, kNeeVertexPart = 3, kNeeTermPart = 4, kNeeAll = 7 };
hipError_t launchNee(int dims, int fp64, const MediumView& m, const ScatterView& sc, const SourceView& il,
                     const SourceView& vw, const FramesView& fr, const VertexBatch& vb, const NeeParams& p,
                     const NeeOut& out, hipStream_t st, hipEvent_t between = nullptr, const FarBound* fb = nullptr,
                     int parts = kNeeAll);
// FarBound::table of one role: for every sub-box of m (sub x sub x sub a cell) and every column
// j < cols (at most 32, at most `blocks`), the least farBoxBound(span) over block j's directions
// (column cols - 1: blocks cols - 1 and up), on st.
hipError_t launchFarTable(const MediumView& m, const MinPyramid& mp, const BoundDirs* blockDirs, uint32_t blocks,
                          uint32_t cols, uint32_t sub, uint32_t span, float* table, hipStream_t st);

// TemporalCorrelationSampler with coherent backscattering: the mixed pairings CBS02 + CBS03 of
// every path with at least two vertices, from the NeeOut first-vertex (F*) and per-vertex (G*)
// terms, added to U[t][v][l] with the factor 1/2; deterministic (split over paths, slices summed
// in order into `part`, at most maxSplits slices of nt * L * V complex FP64).
struct CbsCross {
    const void *F1l, *F2l, *F1v, *F2v;
    const void *Gv1, *Gv2, *Gl1, *Gl2;
    uint64_t gcap;
};
hipError_t launchCbsCross(int dims, int fp64, const SourceView& il, const SourceView& vw, const FramesView& fr,
                          const VertexBatch& vb, const CbsCross& c, uint32_t npaths, double lambda, double* U,
                          double2* part, uint32_t maxSplits, hipStream_t st);
// Slices launchCbsCross uses for an L x V x nt field (its `part` holds that many).
uint32_t cbsCrossSplits(uint32_t L, uint32_t V, uint32_t nt, uint32_t npaths);

// Contraction U[t][v][l] += scale(l,v) * sum_p A[t][p][l] * B[t][p][v] for t < nt.
// U is complex FP64 (interleaved). F (optional, real, [v][l]) multiplies the sum elementwise,
// squared when fsq != 0; fscalar multiplies everything.
enum ContractKind : int { kContractBf16x3 = 0, kContractFp32 = 1, kContractFp64 = 2 };
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
hipError_t launchContract(int kind, const void* A, const void* B, uint32_t P, uint32_t L, uint32_t V, uint32_t nt,
                          const float2* F, const double2* Fd, int fsq, double fscalar, double* U, Workspace* ws,
                          hipStream_t st);

// Generic single scattering: U[t][v][l] += sum_p a[t][p][l] b[t][p][v] f_p(l, v, t) (field output),
// or with corr f_p(l, v, 0) conj f_p(l, v, t) (correlation output, a and b holding the lag products).
hipError_t launchSingleGeneric(int dims, int fp64, const ScatterView& sc, const SourceView& il,
                               const SourceView& vw, const FramesView& fr, const VertexBatch& vb, const void* a,
                               const void* b, uint32_t P, int corr, double* U, Workspace* ws, hipStream_t st);

// Single scattering of the TemporalCorrelationSampler (sources il / vw hold u1 then u2):
// U[t][v][l] += sum_p a[t][p][l] b[t][p][v] Fs_u1 conj(Fs_u2) exp(-k^2 |D tau_t| |q|^2).
// uniformFar: far-field illuminations and views and one phase function for every material.
hipError_t launchSingleCorrelation(int dims, int fp64, const ScatterView& sc, const SourceView& il,
                                   const SourceView& vw, const FramesView& fr, const VertexBatch& vb, const void* a,
                                   const void* b, uint32_t P, double lambda, int uniformFar, double* U,
                                   Workspace* ws, hipStream_t st);

// GaussianBeam set-up: per point i the material there (0 outside) and exp(-tau) along -dirs[i]
// (3D, FP64).
hipError_t launchGaussSetup(const MediumView& m, const double* pts, const double* dirs, uint32_t n, double* att,
                            uint32_t* mat, hipStream_t st);
// Single scattering between GaussianBeam illuminations and views:
// U[t][v][l] += sum_p a[t][p][l] b[t][p][v] fs(l, v, x1), or with corr fs(x1(0)) conj fs(x1(t)).
// uniformViewDir: every view beam has the direction of view 0 (the per-illumination part of fs is
// then shared by the views; field output).
hipError_t launchSingleGauss(int dims, int fp64, int uniformViewDir, int corr, const MediumView& m,
                             const ScatterView& sc, const SourceView& il, const SourceView& vw, const FramesView& fr,
                             const VertexBatch& vb, const void* a, const void* b, uint32_t P, double* U, Workspace* ws,
                             hipStream_t st);

// U += 0 helpers
hipError_t launchZeroU(double* U, size_t n, hipStream_t st);

}  // namespace sstmc
