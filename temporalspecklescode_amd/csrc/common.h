// common.h -- device-side types shared by the transport, NEE and contraction kernels.
//
// Layout decisions (MI355X-first, not a translation of the reference's AoS `TemporalPathPoint`,
// TemporalSampler.cuh:43-50):
//   * every per-request table is passed by value or by pointer in kernel arguments; there is no
//     `__constant__` global state (the reference keeps medium, scattering and sampler tables in
//     process-global `__constant__` symbols, MediumInterface.cuh:30-35, Scattering.cuh:15-29,
//     TemporalSampler.cuh:14-38), so any number of requests can be in flight on one GPU;
//   * the material grid is one byte per voxel (ids <= 9): 2 MiB for 128^3 stays resident in one
//     XCD's 4 MB L2, where a per-voxel FP32 sigt grid (8 MiB) would not; the NEE walks look sigt
//     up in a per-block LDS copy of the material table;
//   * vertices are written once by the transport kernel as SoA records and read by the NEE kernel
//     as wave-uniform broadcasts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace sstmc {

constexpr int kMaxMaterials = 10;   // MATERIAL_NUM (Simulation.cuh:32)
constexpr int kWalkPalette = 16;    // palette entries of the packed walk grid
constexpr uint32_t kMaxAttempts = 1u << 20;  // first-vertex retry budget per path slot

// ---------------------------------------------------------------- counter-based RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Stream spec shared with the CPU
// restatement in oracle/: key = (seed lo, seed hi); counter = (path lo, path hi, event,
// purpose << 24 | block). Events: a = first-vertex attempt a, 0x80000000|k = sampling of vertex
// k >= 2. Purposes: 0 transport, 1 frame displacement (slot s = block s), 2 random phase.
enum : uint32_t { kPurposeTransport = 0, kPurposeFrames = 1, kPurposePhase = 2 };
constexpr uint32_t kEventVertex = 0x80000000u;

__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Rng {
    uint32_t k0, k1, c0, c1, c2, c3;
    uint32_t buf[4];
    int pos;
    __host__ __device__ Rng(uint64_t seed, uint64_t path, uint32_t event, uint32_t purpose)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), c0((uint32_t)path), c1((uint32_t)(path >> 32)),
          c2(event), c3(purpose << 24), pos(4) {}
    __host__ __device__ void seek(uint32_t block) { c3 = (c3 & 0xff000000u) | block; pos = 4; }
    __host__ __device__ uint32_t word() {
        if (pos == 4) {
            philox4x32_10(c0, c1, c2, c3, k0, k1, buf);
            ++c3;
            pos = 0;
        }
        // selects rather than buf[pos]: a dynamic index would put buf in memory (scratch)
        const uint32_t w = pos == 0 ? buf[0] : (pos == 1 ? buf[1] : (pos == 2 ? buf[2] : buf[3]));
        ++pos;
        return w;
    }
    // ((w >> 8) + 0.5) * 2^-24: exactly representable in FP32, strictly inside (0, 1).
    __host__ __device__ float uniform() { return ((float)(word() >> 8) + 0.5f) * (1.0f / 16777216.0f); }
    // ceil(u * n) - 1 for that u (Simulation.cuh:306-309), in exact integer arithmetic.
    __host__ __device__ uint32_t index(uint32_t n) {
        return (uint32_t)(((2ull * (word() >> 8) + 1ull) * n) >> 25);
    }
};

// Hands every lane in `need` (a ballot taken by the calling lanes) its own consecutive index from
// *ctr with one atomic per wavefront; lanes outside `need` get 0.
__device__ inline uint32_t waveTake(uint32_t* ctr, uint64_t needMask, bool need) {
    const int leader = __builtin_ctzll(needMask);
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(needMask));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(needMask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)needMask, 0u));
    return need ? base + below : 0u;
}

// Diagnostic counters spread over kCountSlots lines of 128 bytes: a wavefront adds to the slot of
// its workgroup, so that a launch of a million wavefronts does not queue on one address
// (device-scope atomics to one address complete one at a time, ~10 ns each); the host sums the
// slots. A counter set is kCountSlots * kCountStride values.
constexpr uint32_t kCountSlots = 256, kCountStride = 16;
__device__ inline void slotAdd(unsigned long long* counter, unsigned long long v) {
    atomicAdd(counter + (size_t)(blockIdx.x & (kCountSlots - 1)) * kCountStride, v);
}

// ---------------------------------------------------------------- small vectors
struct f3 { float x, y, z; };
struct d3 { double x, y, z; };
__host__ __device__ inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
__host__ __device__ inline d3 mkd(double x, double y, double z) { d3 r = {x, y, z}; return r; }
__host__ __device__ inline float comp(const f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__host__ __device__ inline double comp(const d3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// ---------------------------------------------------------------- problem tables (device view)
// Medium: HeterogeneousMedium.cuh:17-49 + MediumInterface.cuh:8-35. Edges are kept in FP32 for
// transport and FP32 NEE, and in FP64 for the FP64 NEE variant.
struct MediumView {
    int dims;
    uint32_t n[3];                 // cells per axis (1 for the unused z in 2D)
    float nf[3];                   // n as float (the fast locate's clamp, read as a scalar operand)
    const float* ef[3];            // edges, n[a] + 1 each
    const double* ed[3];
    float bminf[3], bmaxf[3];
    double bmind[3], bmaxd[3];
    const uint8_t* grid;           // material id per cell, x fastest
    float sigtf[kMaxMaterials];
    double sigtd[kMaxMaterials];
    float accept[kMaxMaterials];   // sigt(m) / maxSigt  (Woodcock acceptance)
    float albedo[kMaxMaterials];   // sigs(m) / sigt(m)  (absorption test)
    float invMaj;                  // 1 / maxSigt
    int uniformAxis[3];            // 1 when the edges are equally spaced
    float e0[3], invH[3];          // first edge and inverse spacing for the fast locate
    float h[3];                    // spacing of an equally spaced axis
    int allUniform;                // every axis equally spaced
    // NEE walk grid (built per tracer), the material grid packed for the branch-free FP32 walk:
    // each cell holds 4 x its index in a palette of the materials present, i.e. the byte offset of
    // its sigt in walkSig (walkBits = 1 bit for one material, 4 bits for up to four, else 8),
    // cells in the order that puts the view walks' dominant axis fastest. Cell (i0, i1, i2) starts
    // at bit walkOrigin + i0*walkStride[0] + i1*walkStride[1] + i2*walkStride[2] (strides in bits)
    // of the little-endian word array walkGrid; walkOrigin padding bits (whole words, at least one
    // layer of the slowest axis) precede the cells and as many follow them.
    const uint32_t* walkGrid;
    uint32_t walkWords;            // 32-bit words of walkGrid
    uint32_t walkOrigin;
    uint32_t walkStride[3];
    uint32_t walkBits;
    int walkFast;                  // the fast walk applies (equal spacing; bit offsets fit 32 bits)
    float walkSig[kWalkPalette];   // sigt per palette index
};

// Scattering: Scattering.cuh:15-29, HG table :500-572.
// Tabulated phase functions, in device memory (kept out of ScatterView so that the kernels'
// by-value argument stays small).
// Tabular (Scattering.cuh:191-482): material m's n[m] entries start at off[m], equally spaced
// over the scattering angle on [0, pi] (3D) or [0, 2 pi] (2D); f is the amplitude normalised by
// the cross section, pdf the reference's evaluation pdf (|f|^2, times sin of the angle in 3D),
// icdf the inverse cdf's angle for entry i of a uniform draw.
// von Mises-Fisher mixtures (Scattering.cuh:613-884): material m's vmfM[m] mixtures start at
// vmfOff[m] (kappa and C + log alpha); its vmfM[m]^2 mixture pairs (i + M j) start at
// vmfPairOff[m] (kappa_i + kappa_j, the log normalisation of that sum, the pair's weight and the
// exclusive cumulative weight).
struct PhaseTables {
    uint32_t n[kMaxMaterials];
    uint32_t off[kMaxMaterials];
    const float4* f;               // (|f|, Re f, Im f, pdf) in FP32
    const double4* d;              // the same in FP64
    const double* ph;              // arg f / (2 pi): the amplitude's phase in cycles
    const float* icdf;
    uint32_t vmfM[kMaxMaterials];
    uint32_t vmfOff[kMaxMaterials];
    uint32_t vmfPairOff[kMaxMaterials];
    const double* vmfKappa;
    const double* vmfC;
    const double* vmfPairMu;
    const double* vmfPairLogNorm;
    const double* vmfPairPdf;
    const double* vmfPairCdf;
};

struct ScatterView {
    int dims;
    int allIsotropic;
    int type[kMaxMaterials];       // 1 isotropic, 2 tabular, 3 HG, 4 vMF mixture
    float g[kMaxMaterials];
    double gd[kMaxMaterials];
    const PhaseTables* tab;        // device tables when a material is tabular or vMF, else null
};

// Entry of cosine c in an n-entry tabular amplitude function: the nearest tabulated angle,
// acos with the reference's end clamps, rounded half to even and clamped to the table
// (Scattering.cuh:908-915; safeAcos and round_num, Simulation.cuh:274-297).
__host__ __device__ inline uint32_t tabularEntry(uint32_t n, int dims, double c) {
    const double a = c < -1.0 + 1e-12 ? 3.14159265358979323846 : (c > 1.0 - 1e-12 ? 0.0 : acos(c));
    const double x = a * (double)(n - 1) / (dims == 2 ? 6.28318530717958647692 : 3.14159265358979323846);
    const double r = rint(x);
    return r < 0 ? 0u : (r >= (double)n ? n - 1 : (uint32_t)r);
}

// 3D vMF mixture amplitude sum_i exp(c kappa_i + C_i + log alpha_i) (Scattering.cuh:932-945).
__host__ __device__ inline double vmfAmplitude(const PhaseTables& t, uint32_t m, double c) {
    double a = 0;
    for (uint32_t i = t.vmfOff[m], e = t.vmfOff[m] + t.vmfM[m]; i < e; ++i) a += exp(c * t.vmfKappa[i] + t.vmfC[i]);
    return a;
}
// 3D vMF mixture pdf sum_ij w_ij exp(c (kappa_i + kappa_j) + lognorm_ij) (Scattering.cuh:1003-1017).
__host__ __device__ inline double vmfPdf(const PhaseTables& t, uint32_t m, double c) {
    double p = 0;
    for (uint32_t i = t.vmfPairOff[m], e = t.vmfPairOff[m] + t.vmfM[m] * t.vmfM[m]; i < e; ++i)
        p += t.vmfPairPdf[i] * exp(c * t.vmfPairMu[i] + t.vmfPairLogNorm[i]);
    return p;
}
// binarySearchKernel (Simulation.cuh:325-362): the interval of a sorted n-entry array holding x.
__host__ __device__ inline uint32_t cdfSearch(const double* a, uint32_t n, double x) {
    uint32_t r = n - 1, l = 0;
    if (a[0] > x) return 0;
    if (a[r] < x) return r;
    while (l <= r) {
        const uint32_t m = l + (r - l) / 2;
        if (m == 0) return 0;
        if (a[m] == x) return m;
        if (a[m - 1] <= x && a[m] > x) return m - 1;
        if (a[m] < x) l = m + 1;
        else r = m - 1;
    }
    return 0;
}

// Sources: FarField (FarField.cuh:71-127) / PointSource (PointSource.cuh:176-216).
// FarField illumination sampling table per source, 20 floats: tangent[3] perpT[3] perpS[3]
// minT maxT minS maxS invDir[3] dir[3] -- precomputed on the host in FP64.
constexpr int kFarFieldTable = 20;
// GaussianBeam sources (GaussianBeam.cuh, 3D), FP64: focal points and beam directions,
// gammaA = 1 / aperture^2, apNorm = vMFnormalizationLog(gammaA) (:1053-1062) and the wave number
// 2 pi / lambda. An illumination also carries the sampling tables of updateSampler (:1281-1613),
// built on the host: per beam its focal material, the beam's extent [zmin, zmax] across the box
// in B bins of zdt and a transverse basis (xyb, 6 per beam); per mixture of the focal material
// (row mixOff[l] + i) its weight and exclusive cumulative weight (apdf, aicdf) and B z-densities
// with their exclusive cumulative sums (zpdf, zicdf); prDen[m] = sigs(m) times the box depth in z
// (the first-point probability's denominator) and meanView, the mean view beam direction (the
// axis of the second-point proposal).
struct GaussBeams {
    const double* focal;           // n x 3
    const double* dir;             // n x 3
    double gammaA, apNorm, waveK;
    d3 meanView;
    uint32_t B;
    const uint32_t* focalMat;      // [n]
    const uint32_t* mixOff;        // [n]
    const double* zmin;            // [n]
    const double* zmax;
    const double* zdt;
    const double* xyb;             // n x 6
    const double* apdf;            // [rows]
    const double* aicdf;
    const double* zpdf;            // [rows x B]
    const double* zicdf;
    double prDen[kMaxMaterials];
};

struct SourceView {
    int type;                      // 1 point, 2 far field, 3 Gaussian beam
    int illum;                     // 1 illumination, 0 view
    int area;                      // PointSource isAreaSource
    uint32_t n;
    const float* pf;               // directions (normalised) or locations, n x 3
    const double* pd;
    const float* table;            // FarField illumination sampling table (n x 20), else null
    const float* probf;            // FarField 1 / projected area, n
    const double* probd;
    GaussBeams g;                  // GaussianBeam (pf / pd then hold the focal points)
};

// Temporal frames: TemporalSampler.cuh:14-38 + the time table built in its ctor (:941-1132).
// With the TemporalCorrelationSampler (analytic != 0; TemporalSamplerCorrelation.cuh:14-31) the
// nt entries are time lags, no vertex is displaced (trivial) and the kernels apply the motion
// analytically from ticks, Dd and Ud.
struct FramesView {
    uint32_t nt;                   // output frames (or lags)
    uint32_t neg;                  // slots of the negative chain
    const float* dt;               // per chain slot
    int trivial;                   // every displacement is exactly zero (dt*U == 0 and D*dt == 0)
    int temporalAmp;               // isTemporalAmplitude
    float D[kMaxMaterials];
    float U[kMaxMaterials][3];
    int analytic;                  // TemporalCorrelationSampler
    const double* ticks;           // the lags tau (analytic), nt entries
    double Dd[kMaxMaterials];      // D and U in FP64 (analytic)
    double Ud[kMaxMaterials][3];
};

// One batch of vertices (SoA), written by the transport kernel.
// Per (vertex k >= 2, frame t) quantities the view terms need, derived once from the vertex list
// by k_vertex_prep (FP64): the unit vector from x_k^t back to x_{k-1}^t (positions displaced only
// when the amplitude follows the scatterer motion), and the phase base rho_k + L_k^t / lambda in
// cycles, L_k^t being the path length from x_1^t to x_k^t.
struct __align__(32) VertexRec {
    double bx, by, bz;
    double phase;
};

// Per (vertex k >= 2, frame t), for the far-field view kernel on equally spaced grids: everything
// of a term that depends on the vertex alone, so the kernel's per-(vertex, source) work is what
// depends on both. The walk's start: per axis the cell index b (as an exact float) and the offsets
// of the cell's low and high faces from the position (FP32 face minus FP32 position, as the walk
// forms them), the walk-grid bit offset of the cell (kFarOutside for a position outside the box,
// whose optical depth is 0); the material; the back direction rounded to FP32; the FP64 position
// and phase base of VertexRec.
// (The live-block masks of the far-field view kernel are kept apart, VertexBatch::live.)
constexpr uint32_t kFarOutside = 0xffffffffu;
struct __align__(32) FarRec {
    float dlo[3];
    uint32_t cell;
    float dhi[3];
    uint32_t mat;
    float b[3];
    float back[3];
    uint32_t pad[2];
    double px, py, pz;
    double phase;
};
static_assert(sizeof(FarRec) == 96, "FarRec is three 32-byte lines");

// Per-layer least sigt over aligned square blocks of cells, for lower bounds on optical depths
// (farBoxBound). Layers run along `axis`; each layer is a 2D grid over the two other axes (b, c in
// increasing order) reduced 2 x 2 per level, up to one block at the top level. Level L of layer l
// holds block (ib >> L, ic >> L) at p[off[L] + (l * nc[L] + (ic >> L)) * nb[L] + (ib >> L)]; FP32
// sigt as the walks use it.
constexpr int kPyrLevels = 16;
struct MinPyramid {
    const float* p = nullptr;
    int axis = -1;                 // -1: no pyramid (no bound)
    uint32_t nl = 0, levels = 0;
    uint64_t off[kPyrLevels] = {};
    uint32_t nb[kPyrLevels] = {}, nc[kPyrLevels] = {};
};
// A set of walk directions seen from the axis they all advance along: sign of that component and
// the range of each lateral component over |axis component| (the two other axes in order).
struct BoundDirs {
    int valid = 0;
    int sgn = 1;
    float rmin[2] = {0.f, 0.f}, rmax[2] = {0.f, 0.f};
};
// The bounds of the far-field view kernel's two launches: views (0) and, for coherent
// backscattering, the reversed paths' illuminations (1). Per role the kernel's sources are taken
// in the order perm (a Z-order over their lateral slopes, so a wavefront's 64 sources form a compact
// cone of directions), and table[box * cols + j] is a lower bound on the optical depth from any
// point of the box to the medium's boundary along every direction of block j, sources
// perm[64 j .. 64 j + 63] (farBoxBound; column 31 the least over blocks 31 and up), for the
// sub x sub x sub boxes of every cell (farSubBoxOf's index), built once per tracer.
struct FarBound {
    const uint32_t* perm[2] = {nullptr, nullptr};
    const float* table[2] = {nullptr, nullptr};
    uint32_t blocks[2] = {0u, 0u};
    uint32_t cols[2] = {0u, 0u};
    uint32_t sub = 1;
};

struct VertexBatch {
    uint32_t* k;                   // [P] scattering vertices per path
    uint32_t* misses;              // [P]
    uint32_t* off;                 // [P + 1] exclusive scan of k
    float4* xm;                    // [cap] position + material id (as float bits)
    float* rho;                    // [cap] random phase
    float* delta;                  // [cap * nt * 3] displacement per output frame (if non-trivial)
    uint64_t cap;
    uint32_t* status;              // [4]: 0 overflow flag, 1 no-hit flag
    VertexRec* rec;                // [cap * nt], record of vertex o, frame t at o * nt + t
    FarRec* far = nullptr;         // [cap * nt] like rec, when the far-field view kernel runs
    // [cap * nt * 2] with far: the live-block masks of record i at live[2 i + role], role 0 the
    // views and 1 the reversed paths' illuminations: bit b set when block b of the far-field view
    // kernel's sources (FarBound::perm) may hold a term inside the launch's cut, bit 31 standing for
    // every block from 31 on; all set when no bound applies (k_vertex_prep, k_far_live)
    uint32_t* live = nullptr;
    // [P] with NeeParams::viewFirst: 1 when the path's B row holds a nonzero term (cleared by
    // k_vertex_prep, set by k_nee_view_far)
    uint32_t* rowNz = nullptr;
};

// Staging of the single-pass transport. Lanes draw path slots from a queue and store each path's
// vertices in chunks of kStageChunk taken from a second counter as the path grows (chunk c holds
// vertices chunkJ0[c] .. chunkJ0[c] + kStageChunk - 1 of path slot chunkSlot[c]); once the
// exclusive scan of the per-path counts is known, k_transport_pack copies every staged vertex to
// off[slot] + j. With nchunks * kStageChunk >= cap + P * (kStageChunk - 1) the staging cannot run
// out while the batch fits the pool (each path leaves at most kStageChunk - 1 slots unused).
constexpr uint32_t kStageChunk = 16;
struct TransportStage {
    float4* xm;                    // [nchunks * kStageChunk] like VertexBatch::xm
    float* rho;                    // [nchunks * kStageChunk]
    float* delta;                  // [nchunks * kStageChunk * nt * 3] (if non-trivial frames)
    uint32_t* chunkSlot;           // [nchunks] path slot of the chunk
    uint32_t* chunkJ0;             // [nchunks] index within its path of the chunk's first vertex
    uint32_t nchunks;
    uint32_t* q;                   // [2]: next path slot, chunks taken (zeroed before each launch)
};

}  // namespace sstmc
