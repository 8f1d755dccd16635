// transport.hip -- (a) free-flight sampling of every path of a batch, FP32.
//
// Replaces the reference's per-bounce chain of launches (sapmleFirstPathPoints /
// sapmleNextPathPoints, NextEventEstimation.cuh:643-759; TemporalSampler::sampleFirst /
// sampleSecond / sampleNext, TemporalSampler.cuh:1324-1462; dtshift_host, :778-937) by one
// thread per path slot that walks the whole path. The same kernel runs twice per batch: a count
// pass (vertices per path) and, after an exclusive scan, a write pass that stores the vertices at
// their offsets. The counter-based stream makes both passes draw identical numbers; the write pass
// never stores beyond the counted vertices and raises an overflow flag if the pool is too small.
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

// GB: GaussianBeam illumination (3D), whose first point is drawn from the beam profile
// (sampleFirstBeam, GaussianBeam.cuh:357-418; a point outside the medium is a miss, no absorption
// test) and whose second direction comes from the beam's lobe-pair proposal
// (sampleSecondPoint, :1673-1748).
template <int DIMS, bool TAB, bool GB>
__global__ void __launch_bounds__(256) k_transport(MediumView m, ScatterView sc, SourceView il, FramesView fr,
                                                   VertexBatch vb, uint64_t seed, uint64_t pathBegin,
                                                   uint32_t npaths, int write) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= npaths) return;
    uint64_t path = pathBegin + slot;
    uint32_t misses = 0;
    uint32_t event = 0;
    f3 x, dir;
    uint32_t mat = 0, beam = 0;
    for (;;) {
        if (event >= kMaxAttempts) {
            atomicOr(vb.status + 1, 1u);
            if (!write) { vb.k[slot] = 0; vb.misses[slot] = misses; }
            return;
        }
        Rng rng(seed, path, event, kPurposeTransport);
        if constexpr (GB) {
            const d3 xd = gbSampleFirst(il.g, *sc.tab, il.n, rng, &beam);
            x = mk((float)xd.x, (float)xd.y, (float)xd.z);
            dir = mk((float)il.g.dir[3 * beam], (float)il.g.dir[3 * beam + 1], (float)il.g.dir[3 * beam + 2]);
            mat = gbMaterial(m, xd);
        } else {
            f3 x0;
            sampleSource<DIMS>(il, m, rng, &x0, &dir);
            mat = woodcock<DIMS>(m, x0, dir, rng, &x);
        }
        if (mat > 0) break;
        ++misses;
        ++event;
    }
    uint32_t base = 0, limit = 0;
    bool store = false;
    if (write) {
        base = vb.off[slot];
        limit = vb.off[slot + 1] - base;
        store = (uint64_t)vb.off[npaths] <= vb.cap;
        if (!store && slot == 0) atomicOr(vb.status, 1u);
    }
    uint32_t kcount = 0;
    uint32_t ev = event;
    for (uint32_t k = 1;; ++k) {
        if (k > 1) {
            ev = kEventVertex | k;
            Rng rng(seed, path, ev, kPurposeTransport);
            f3 nd;
            if constexpr (GB) {
                if (k == 2) nd = gbSampleSecond(il.g, *sc.tab, sc.allIsotropic != 0, beam, mkd(x.x, x.y, x.z), rng);
                else nd = sampleDirection<DIMS, TAB>(sc, mat, dir, rng);
            } else {
                nd = sampleDirection<DIMS, TAB>(sc, mat, dir, rng);
            }
            f3 nx;
            uint32_t nm = woodcock<DIMS>(m, x, nd, rng, &nx);
            if (nm == 0) break;
            x = nx;
            dir = nd;
            mat = nm;
        }
        if (store && kcount < limit) {
            uint64_t o = (uint64_t)base + kcount;
            vb.xm[o] = make_float4(x.x, x.y, x.z, __uint_as_float(mat));
            vb.rho[o] = Rng(seed, path, ev, kPurposePhase).uniform();
            if (!fr.trivial) frameDeltas<DIMS>(fr, seed, path, ev, mat, vb.delta + o * fr.nt * 3);
        }
        ++kcount;
    }
    if (!write) {
        vb.k[slot] = kcount;
        vb.misses[slot] = misses;
    } else if (kcount != limit) {
        atomicOr(vb.status, 2u);
    }
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
// terms): it is flagged in `overflow` (bit `batch`) for the host to re-run after growing the pool,
// and its misses and vertices are counted by that re-run instead.
__global__ void __launch_bounds__(256) k_batch_totals(const uint32_t* misses, const uint32_t* off, uint32_t n,
                                                      unsigned long long* acc, uint64_t cap, uint32_t* overflow,
                                                      uint32_t batch, float* tauRef) {
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
        if ((uint64_t)off[n] <= cap) {
            acc[0] += s[0];
            acc[1] += off[n];
        } else {
            overflow[batch >> 5] |= 1u << (batch & 31);
        }
        if (off[n] > acc[3]) acc[3] = off[n];   // largest batch, sizes the vertex pool
    }
}

hipError_t launchTransport(int dims, const MediumView& m, const ScatterView& sc, const SourceView& il,
                           const FramesView& fr, const VertexBatch& vb, uint64_t seed, uint64_t pathBegin,
                           uint32_t npaths, int write, hipStream_t st) {
    dim3 grid((npaths + 255) / 256), block(256);
    const bool tab = sc.tab != nullptr;
#define SSTMC_TRANSPORT(D, T, G) \
    hipLaunchKernelGGL((k_transport<D, T, G>), grid, block, 0, st, m, sc, il, fr, vb, seed, pathBegin, npaths, write)
    if (il.type == 3 && dims == 3) {
        if (!tab) return hipErrorInvalidValue;
        SSTMC_TRANSPORT(3, true, true);
    } else if (dims == 2) { if (tab) SSTMC_TRANSPORT(2, true, false); else SSTMC_TRANSPORT(2, false, false); }
    else { if (tab) SSTMC_TRANSPORT(3, true, false); else SSTMC_TRANSPORT(3, false, false); }
#undef SSTMC_TRANSPORT
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

hipError_t launchBatchTotals(const uint32_t* misses, const uint32_t* off, uint32_t n, unsigned long long* acc,
                             uint64_t cap, uint32_t* overflow, uint32_t batch, float* tauRef, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_totals, dim3(1), dim3(256), 0, st, misses, off, n, acc, cap, overflow, batch, tauRef);
    return hipGetLastError();
}

}  // namespace sstmc
