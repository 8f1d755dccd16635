// transport.hip -- (a) free-flight sampling of every path of a batch, FP32.
//
// Replaces the reference's per-bounce chain of launches (sapmleFirstPathPoints /
// sapmleNextPathPoints, NextEventEstimation.cuh:643-759, with their per-bounce compaction of live
// paths; TemporalSampler::sampleFirst / sampleSecond / sampleNext, TemporalSampler.cuh:1324-1462;
// dtshift_host, :778-937) by one pass of a persistent grid whose lanes take path slots from a
// queue and keep their wavefront busy (k_transport), an exclusive scan of the vertex counts, and a
// copy of the staged vertices to their offsets (k_transport_pack).
#include "common.h"
#include "gauss.h"
#include "geometry.h"
#include "kernels.h"

namespace sstmc {

// sampleFarFieldKernel (FarField.cuh:129-205) / PointSourceNS::sampleFirstPointKernel
// (PointSource.cuh:72-90): source point and direction of the first flight. A 2D GaussianBeam draws
// its first flight as a point source at its focal point does (sampleFirstBeam,
// GaussianBeam.cuh:754-773, then the medium's sample; the focal points are its locations).
template <int DIMS>
__device__ inline void sampleSource(const SourceView& il, const MediumView& m, Rng& rng, f3* x0, f3* dir) {
    uint32_t l = rng.index(il.n);
    if (il.type == 1 || il.type == 3) {
        *x0 = mk(il.pf[3 * l], il.pf[3 * l + 1], il.pf[3 * l + 2]);
        *dir = randomDirection<DIMS>(rng);
        return;
    }
    const float* t = il.table + (size_t)l * kFarFieldTable;
    float uT = rng.uniform();
    float sT = t[9] + uT * (t[10] - t[9]);
    float px = fmaf(sT, t[3], t[0]), py = fmaf(sT, t[4], t[1]), pz = fmaf(sT, t[5], t[2]);
    if (DIMS == 3) {
        float uS = rng.uniform();
        float sS = t[11] + uS * (t[12] - t[11]);
        px = fmaf(sS, t[6], px); py = fmaf(sS, t[7], py); pz = fmaf(sS, t[8], pz);
    }
    float d[3] = {t[17], t[18], t[19]};
    float inv[3] = {t[14], t[15], t[16]};
    float p[3] = {px, py, pz};
    float r[3], sh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r[a] = 0.5f * (m.bmaxf[a] - m.bminf[a]);
        sh[a] = p[a] - 0.5f * (m.bmaxf[a] + m.bminf[a]);
    }
    float bd = 0.f;
    bool found = false;
#pragma unroll
    for (int a = 0; a < DIMS; ++a) {
        int sgn = d[a] < 0.f ? 1 : -1;
        float dist = inv[a] * (r[a] * (float)sgn - sh[a]);
        bool ok = dist >= 0.0f;
#pragma unroll
        for (int b = 0; b < DIMS; ++b)
            if (b != a) ok = ok && fabsf(fmaf(d[b], dist, sh[b])) < r[b];
        if (ok && !found) { bd = dist; found = true; }
    }
    *x0 = mk(fmaf(bd, d[0], px), fmaf(bd, d[1], py), DIMS == 3 ? fmaf(bd, d[2], pz) : 0.f);
    *dir = mk(d[0], d[1], d[2]);
}

// dtShiftKernel + shiftFirstTemporalPoint + inclusive_scan_by_key + copyPointsFromBuffers*Temporal
// (TemporalSampler.cuh:376-647, 778-937): per output frame, the displacement of the scatterer from
// its central position, built as two chains anchored at the mean frame time.
template <int DIMS>
__device__ inline void frameDeltas(const FramesView& fr, uint64_t seed, uint64_t path, uint32_t event,
                                   uint32_t mat, float* out /* nt*3 */) {
    Rng rng(seed, path, event, kPurposeFrames);
    float D = fr.D[mat];
    float U0 = fr.U[mat][0], U1 = fr.U[mat][1], U2 = fr.U[mat][2];
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (uint32_t k = 0; k < fr.nt; ++k) {
        if (k == 0 || k == fr.neg) { cx = cy = cz = 0.f; }
        float dt = fr.dt[k];
        rng.seek(k);
        float u1 = rng.uniform(), u2 = rng.uniform();
        float nrm = sqrtf(-2.0f * __logf(u1)) * __cosf(6.28318530717958648f * u2);
        f3 w = randomDirection<DIMS>(rng);
        float dr = nrm * sqrtf(2.0f * (float)DIMS * D * fabsf(dt));
        cx += fmaf(dr, w.x, dt * U0);
        cy += fmaf(dr, w.y, dt * U1);
        if (DIMS == 3) cz += fmaf(dr, w.z, dt * U2);
        uint32_t t = k < fr.neg ? fr.neg - 1 - k : k;
        out[3 * t] = cx;
        out[3 * t + 1] = cy;
        out[3 * t + 2] = cz;
    }
}

constexpr uint32_t kNoSlot = 0xffffffffu;

// LDS copies of what a flight step reads besides the grid itself: the cell edges of each axis (when
// every axis has at most kLdsEdges - 1 cells) and the per-material acceptance and albedo. The
// lanes that trace a batch's longest paths run nearly alone at its end, so each flight step's
// memory round trips set the transport's duration: the edges (locate's correction) are one.
constexpr uint32_t kLdsEdges = 1025;
struct FlightLds {
    float e[3][kLdsEdges];
    float accept[kMaxMaterials], albedo[kMaxMaterials];
};
// locate<float> (same cells) with the edges read from LDS.
__device__ inline int locateLds(const MediumView& m, const FlightLds& f, int a, float x) {
    const float* e = f.e[a];
    const int n = (int)m.n[a];
    if (m.uniformAxis[a]) {
        const float g = (x - m.e0[a]) * m.invH[a];
        int b = g <= 0.f ? 0 : (g >= m.nf[a] ? n - 1 : (int)g);
        if (b > 0 && x < e[b]) --b;
        if (b < n - 1 && x >= e[b + 1]) ++b;
        return b;
    }
    if (x < e[0]) return 0;
    if (x >= e[n]) return n - 1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (e[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}
template <int DIMS>
__device__ inline uint32_t materialAtLds(const MediumView& m, const FlightLds& f, bool lds, const f3& x) {
    if (!lds) return materialAt<DIMS>(m, x);
    const int bx = locateLds(m, f, 0, x.x), by = locateLds(m, f, 1, x.y);
    const int bz = DIMS == 3 ? locateLds(m, f, 2, x.z) : 0;
    return m.grid[cellIndex(m, bx, by, bz)];
}
// A wavefront keeps stepping flights while at least this many lanes fly; below it, the lanes that
// wait on an event (a collision to store, a miss, a finished path) are served together.
constexpr uint32_t kFlyQuorum = 16;

// Single-pass transport over a queue of path slots. Every lane is a small state machine that
// advances one tentative Woodcock collision per flight step (woodcock's loop body, geometry.h)
// and, at an event step, resolves the flight (absorption test), stores the vertex in its staging
// chunk, samples the next direction, or finishes its path and takes the next slot from the queue.
// The random numbers of a path depend on (seed, path id, event) only, so the vertices are those
// of a thread that traces the path whole, in the same order; the lanes of a wavefront stay busy
// however unequal their paths' lengths (heavy-tailed: a few paths outlast the rest many times).
// GB: GaussianBeam illumination (3D), whose first point is drawn from the beam profile
// (sampleFirstBeam, GaussianBeam.cuh:357-418; a point outside the medium is a miss, no absorption
// test) and whose second direction comes from the beam's lobe-pair proposal
// (sampleSecondPoint, :1673-1748).
template <int DIMS, bool TAB, bool GB>
__global__ void __launch_bounds__(64) k_transport(MediumView m, ScatterView sc, SourceView il, FramesView fr,
                                                   VertexBatch vb, TransportStage ts, uint64_t seed,
                                                   uint64_t pathBegin, uint32_t npaths) {
    uint32_t slot = kNoSlot;       // the lane's path slot; kNoSlot: take one at the next event step
    bool dead = false;             // the queue ran dry for this lane
    bool flying = false;           // a flight is under way
    bool outside = false;          // the finished flight left the medium
    bool firstPhase = true;        // drawing the first vertex (attempts)
    bool chunkOk = false;
    uint64_t path = 0;
    uint32_t event = 0, misses = 0, kcount = 0, ev = 0, chunk = 0, beam = 0;
    uint32_t mat = 0, fmat = 0;
    f3 x = mk(0.f, 0.f, 0.f), dir = x, o = x, d = x, xx = x;
    float dist = 0.f;
    Rng rng(seed, 0, 0, kPurposeTransport);
    __shared__ FlightLds fl;
    const bool edgesLds = m.n[0] < kLdsEdges && m.n[1] < kLdsEdges && (DIMS == 2 || m.n[2] < kLdsEdges);
    for (uint32_t a = 0; a < (uint32_t)DIMS && edgesLds; ++a)
        for (uint32_t i = threadIdx.x; i <= m.n[a]; i += blockDim.x) fl.e[a][i] = m.ef[a][i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kMaxMaterials; i += blockDim.x) {
        fl.accept[i] = m.accept[i];
        fl.albedo[i] = m.albedo[i];
    }
    __syncthreads();

    for (;;) {
        const uint64_t live = __ballot(!dead);
        if (!live) break;
        const uint32_t nfly = (uint32_t)__popcll(__ballot(!dead && flying));
        if (nfly == (uint32_t)__popcll(live) || nfly >= kFlyQuorum) {
            // flight step: one tentative collision of every flying lane
            if (flying) {
                dist += -__logf(rng.uniform()) * m.invMaj;
                xx = mk(fmaf(dist, d.x, o.x), fmaf(dist, d.y, o.y), DIMS == 3 ? fmaf(dist, d.z, o.z) : 0.f);
                if (!inside<DIMS, float>(m, xx)) {
                    if (!pointsToward<DIMS>(m, xx, d)) { flying = false; outside = true; }
                } else {
                    fmat = materialAtLds<DIMS>(m, fl, edgesLds, xx);
                    if (rng.uniform() <= fl.accept[fmat]) { flying = false; outside = false; }
                }
            }
            continue;
        }
        // event step of every lane that is not flying
        if (dead || flying) continue;
        bool store = false, startSegment = false, startAttempt = false, needPath = slot == kNoSlot;
        if (!needPath) {
            // the flight's end: out of the medium, absorbed, or a scattering vertex
            const uint32_t nm = outside ? 0u : (rng.uniform() > fl.albedo[fmat] ? 0u : fmat);
            if (nm > 0) {
                x = xx;
                dir = d;
                mat = nm;
                store = true;
                startSegment = true;
                firstPhase = false;
            } else if (firstPhase) {
                ++misses;
                ++event;
                startAttempt = true;
            } else {
                vb.k[slot] = kcount;
                vb.misses[slot] = misses;
                slot = kNoSlot;
                needPath = true;
            }
        }
        // the next path slot for lanes without one (one queue atomic per wavefront)
        const uint64_t want = __ballot(needPath);
        if (want) {
            const uint32_t s = waveTake(ts.q, want, needPath);
            if (needPath) {
                if (s < npaths) {
                    slot = s;
                    path = pathBegin + s;
                    event = misses = kcount = 0;
                    firstPhase = true;
                    startAttempt = true;
                } else {
                    dead = true;
                }
            }
        }
        // first-vertex attempt (sapmleFirstPathPoints, NextEventEstimation.cuh:643-679)
        if (startAttempt) {
            bool giveUp = false;
            if constexpr (GB) {
                for (;;) {
                    if (event >= kMaxAttempts) { giveUp = true; break; }
                    rng = Rng(seed, path, event, kPurposeTransport);
                    const d3 xd = gbSampleFirst(il.g, *sc.tab, il.n, rng, &beam);
                    x = mk((float)xd.x, (float)xd.y, (float)xd.z);
                    dir = mk((float)il.g.dir[3 * beam], (float)il.g.dir[3 * beam + 1], (float)il.g.dir[3 * beam + 2]);
                    mat = gbMaterial(m, xd);
                    if (mat > 0) break;
                    ++misses;
                    ++event;
                }
                if (!giveUp) {
                    ev = event;
                    store = true;
                    startSegment = true;
                    firstPhase = false;
                }
            } else {
                if (event >= kMaxAttempts) {
                    giveUp = true;
                } else {
                    rng = Rng(seed, path, event, kPurposeTransport);
                    sampleSource<DIMS>(il, m, rng, &o, &d);
                    ev = event;
                    dist = 0.f;
                    flying = true;
                }
            }
            if (giveUp) {
                atomicOr(vb.status + 1, 1u);
                vb.k[slot] = 0;
                vb.misses[slot] = misses;
                slot = kNoSlot;
            }
        }
        // store vertex kcount + 1 in the path's staging chunk (a new chunk every kStageChunk)
        const bool needChunk = store && kcount % kStageChunk == 0;
        const uint64_t cm = __ballot(needChunk);
        if (cm) {
            const uint32_t c = waveTake(ts.q + 1, cm, needChunk);
            if (needChunk) {
                chunk = c;
                chunkOk = c < ts.nchunks;
                if (chunkOk) {
                    ts.chunkSlot[c] = slot;
                    ts.chunkJ0[c] = kcount;
                }
            }
        }
        if (store) {
            if (chunkOk) {
                const uint64_t so = (uint64_t)chunk * kStageChunk + kcount % kStageChunk;
                ts.xm[so] = make_float4(x.x, x.y, x.z, __uint_as_float(mat));
                ts.rho[so] = Rng(seed, path, ev, kPurposePhase).uniform();
                if (!fr.trivial) frameDeltas<DIMS>(fr, seed, path, ev, mat, ts.delta + so * fr.nt * 3);
            }
            ++kcount;
        }
        // the flight toward vertex kcount + 1 (sapmleNextPathPoints, NextEventEstimation.cuh:681-759)
        if (startSegment) {
            ev = kEventVertex | (kcount + 1);
            rng = Rng(seed, path, ev, kPurposeTransport);
            f3 nd;
            if constexpr (GB) {
                if (kcount == 1) nd = gbSampleSecond(il.g, *sc.tab, sc.allIsotropic != 0, beam, mkd(x.x, x.y, x.z), rng);
                else nd = sampleDirection<DIMS, TAB>(sc, mat, dir, rng);
            } else {
                nd = sampleDirection<DIMS, TAB>(sc, mat, dir, rng);
            }
            o = x;
            d = nd;
            dist = 0.f;
            flying = true;
        }
    }
}

// Copies the staged vertices to their places in the vertex pool (k_transport's staging).
__global__ void __launch_bounds__(256) k_transport_pack(VertexBatch vb, TransportStage ts, uint32_t npaths,
                                                        uint32_t nt, int frames) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t used = min(ts.q[1], ts.nchunks);
    if (i >= (uint64_t)used * kStageChunk || (uint64_t)vb.off[npaths] > vb.cap) return;
    const uint32_t c = (uint32_t)(i / kStageChunk);
    const uint32_t s = ts.chunkSlot[c];
    const uint32_t j = ts.chunkJ0[c] + (uint32_t)(i % kStageChunk);
    if (j >= vb.k[s]) return;
    const uint64_t o = (uint64_t)vb.off[s] + j;
    vb.xm[o] = ts.xm[i];
    vb.rho[o] = ts.rho[i];
    if (frames)
        for (uint32_t t = 0; t < 3 * nt; ++t) vb.delta[o * 3 * nt + t] = ts.delta[i * 3 * nt + t];
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------- exclusive scan of k (3 phases)
__global__ void __launch_bounds__(1024) k_scan_blocks(const uint32_t* in, uint32_t* out, uint32_t* blockSums,
                                                      uint32_t n) {
    __shared__ uint32_t s[1024];
    uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    uint32_t v = i < n ? in[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        uint32_t t = threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 1023) blockSums[blockIdx.x] = s[1023];
}

__global__ void __launch_bounds__(1024) k_scan_sums(uint32_t* blockSums, uint32_t nb, uint32_t* total) {
    __shared__ uint32_t s[1024];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? blockSums[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            uint32_t t = threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) blockSums[i] = s[threadIdx.x] - v + carry;
        uint32_t last = s[1023];
        __syncthreads();
        carry += last;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(1024) k_scan_add(uint32_t* out, const uint32_t* blockSums, uint32_t n) {
    uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += blockSums[blockIdx.x];
}

// Per-batch totals: misses, vertex count and the largest batch, in 64-bit counters (one block,
// launched in stream order after each batch, so the updates never race). A batch whose vertices
// did not fit the pool contributes nothing (every later kernel sees the overflow and writes zero
// terms): it is flagged in `overflow` (bit `batch`) and status[0] for the host to re-run after
// growing the pool, and its misses and vertices are counted by that re-run instead. Should the
// staging have run out (it cannot while the pool holds the batch, see TransportStage), off[n] is
// set past the pool so that the batch counts as overflowed everywhere.
__global__ void __launch_bounds__(256) k_batch_totals(const uint32_t* misses, uint32_t* off, uint32_t n,
                                                      unsigned long long* acc, uint64_t cap, uint32_t* overflow,
                                                      uint32_t batch, float* tauRef, TransportStage ts,
                                                      uint32_t* status) {
    if (tauRef && threadIdx.x == 0) *tauRef = 3.0e38f;   // the batch's k_tau_ref lowers it
    __shared__ unsigned long long s[256];
    unsigned long long v = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) v += misses[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t total = off[n];
        const bool staged = ts.q[1] <= ts.nchunks;
        if ((uint64_t)total <= cap && staged) {
            acc[0] += s[0];
            acc[1] += total;
        } else {
            overflow[batch >> 5] |= 1u << (batch & 31);
            atomicOr(status, 1u);
            if (!staged) off[n] = 0xffffffffu;
        }
        if (total > acc[3]) acc[3] = total;   // largest batch, sizes the vertex pool
    }
}

// Lanes of the transport's persistent grid: enough wavefronts to keep every CU's share of the
// batch's paths flowing, few enough that each lane traces several paths from the queue.
constexpr uint32_t kTransportWaves = 512;

hipError_t launchTransport(int dims, const MediumView& m, const ScatterView& sc, const SourceView& il,
                           const FramesView& fr, const VertexBatch& vb, const TransportStage& ts, uint64_t seed,
                           uint64_t pathBegin, uint32_t npaths, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ts.q, 2u);
    const uint32_t need = (npaths + 63) / 64, waves = need < kTransportWaves ? need : kTransportWaves;
    dim3 grid(waves), block(64);
    const bool tab = sc.tab != nullptr;
#define SSTMC_TRANSPORT(D, T, G) \
    hipLaunchKernelGGL((k_transport<D, T, G>), grid, block, 0, st, m, sc, il, fr, vb, ts, seed, pathBegin, npaths)
    if (il.type == 3 && dims == 3) {
        if (!tab) return hipErrorInvalidValue;
        SSTMC_TRANSPORT(3, true, true);
    } else if (dims == 2) { if (tab) SSTMC_TRANSPORT(2, true, false); else SSTMC_TRANSPORT(2, false, false); }
    else { if (tab) SSTMC_TRANSPORT(3, true, false); else SSTMC_TRANSPORT(3, false, false); }
#undef SSTMC_TRANSPORT
    return hipGetLastError();
}

hipError_t launchTransportPack(const VertexBatch& vb, const TransportStage& ts, uint32_t npaths, uint32_t nt,
                               int frames, hipStream_t st) {
    const uint64_t n = (uint64_t)ts.nchunks * kStageChunk;
    hipLaunchKernelGGL(k_transport_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vb, ts, npaths, nt,
                       frames);
    return hipGetLastError();
}

// GaussianBeam set-up (updateSampler's attenuation along the beam and getMaterialInfo of the focal
// points, GaussianBeam.cuh:1281-1613): per point the material there and exp(-tau) toward -dirs.
__global__ void __launch_bounds__(256) k_gauss_setup(MediumView m, const double* pts, const double* dirs, uint32_t n,
                                                     double* att, uint32_t* mat) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const d3 p = mkd(pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]);
    const d3 w = mkd(-dirs[3 * (size_t)i], -dirs[3 * (size_t)i + 1], -dirs[3 * (size_t)i + 2]);
    mat[i] = gbMaterial(m, p);
    att[i] = exp(-opticalDepth<3, double, d3>(m, m.sigtd, p, w, 1e300, nullptr));
}

hipError_t launchGaussSetup(const MediumView& m, const double* pts, const double* dirs, uint32_t n, double* att,
                            uint32_t* mat, hipStream_t st) {
    hipLaunchKernelGGL(k_gauss_setup, dim3((n + 255) / 256), dim3(256), 0, st, m, pts, dirs, n, att, mat);
    return hipGetLastError();
}

hipError_t launchScan(const uint32_t* k, uint32_t* off, uint32_t* blockSums, uint32_t n, hipStream_t st) {
    uint32_t nb = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(1024), 0, st, k, off, blockSums, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, blockSums, nb, off + n);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, st, off, blockSums, n);
    return hipGetLastError();
}

hipError_t launchBatchTotals(const uint32_t* misses, uint32_t* off, uint32_t n, unsigned long long* acc,
                             uint64_t cap, uint32_t* overflow, uint32_t batch, float* tauRef, const TransportStage& ts,
                             uint32_t* status, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_totals, dim3(1), dim3(256), 0, st, misses, off, n, acc, cap, overflow, batch, tauRef,
                       ts, status);
    return hipGetLastError();
}

}  // namespace sstmc
