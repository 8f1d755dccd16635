// contract.hip -- (c) the complex contraction U[t][v][l] += sum_p A[t][p][l] * B[t][p][v].
//
// The reference materialises the Nl x Nv x P outer products and reduces them with cublasZgemv
// (pointwiseMultipication + matrixMultipication, MatrixOperations.cuh:44-149, 444-550, called from
// NextEventEstimation.cuh:1134,1388,1451): ~48 bytes of HBM traffic per path-contribution. Here it
// is one dense GEMM per frame with K = paths:
//   * bf16x3: each FP32 operand is split into bf16 hi + lo (hi = rn(x), lo = rn(x - hi)) and the
//     product is hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16, FP32 accumulate. bf16 keeps
//     FP32's exponent range, so no per-path rescaling is needed. The complex product is the real
//     GEMM [Are Aim] x [[Bre Bim] [-Bim Bre]] with K' = 2P; the second operand is formed from the
//     (re, im) bf16 pairs in registers by a sign flip and a half-word rotate.
//   * fp32 / fp64: LDS-tiled FMA kernels, the accuracy references.
// Split-K partial sums go to a workspace and are reduced in a fixed order, so a result does not
// depend on scheduling: a request gives the same bits alone or in a mix.
#include <mutex>

#include "../../include/sstmc.h"
#include "common.h"
#include "gauss.h"
#include "kernels.h"

namespace sstmc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

// round-to-nearest-even FP32 -> bf16 bits
__device__ inline uint32_t bf16bits(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffff) ? 0x40u : 0u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline float bf16val(uint32_t h) { return __uint_as_float(h << 16); }

// ---------------------------------------------------------------- bf16x3 MFMA kernel
// The dispatcher hands workgroup n (x fastest, then y, then z) to XCD n % 8, and each XCD has its
// own L2. With the natural order XCD k would run l-tile k of every (v-tile, split), so every XCD
// would read all of B from memory. Renumbering gives XCD k a contiguous run of the logical order
// (whole splits of all tiles): the workgroups that run together on an XCD share one split's A
// and B chunks in its L2, and each operand is read from memory about once. Results do not depend
// on where a tile runs.
constexpr uint32_t kXcds = 8;
struct TileIdx { uint32_t x, y, z; };
__device__ inline TileIdx xcdTile() {
    const uint32_t gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
    uint32_t id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (n % kXcds == 0) id = (id % kXcds) * (n / kXcds) + id / kXcds;
    return TileIdx{id % gx, (id / gx) % gy, id / (gx * gy)};
}

constexpr int kTileL = 128, kTileV = 128, kChunk = 32;         // paths per LDS chunk
constexpr int kRowBytes = kChunk * 2 * 2 + 16;                  // k' = 2*paths bf16 + 16 B pad
constexpr int kRowWords = kRowBytes / 4;

__global__ void __launch_bounds__(256) k_contract_bf16x3(const float2* __restrict__ A, const float2* __restrict__ B,
                                                         uint32_t P, uint32_t L, uint32_t V, uint32_t pathsPerSplit,
                                                         float2* __restrict__ part) {
    __shared__ uint32_t sAh[kTileL * kRowWords], sAl[kTileL * kRowWords];
    __shared__ uint32_t sBh[kTileV * kRowWords], sBl[kTileV * kRowWords];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const TileIdx tile = xcdTile();
    const uint32_t l0 = tile.x * kTileL, v0 = tile.y * kTileV;
    const uint32_t split = tile.z;   // one frame per launch (the host offsets A, B and part)
    const uint32_t pBegin = split * pathsPerSplit;
    const uint32_t pEnd = min(P, pBegin + pathsPerSplit);
    const uint32_t wl = (wave & 1) * 64, wv = (wave >> 1) * 64;   // wave sub-tile origin
    f32x16 accRe[2][2], accIm[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) { accRe[i][j][r] = 0.f; accIm[i][j][r] = 0.f; }

    const uint32_t ld_row = tid & 127;        // l (or v) row this thread converts
    const uint32_t ld_grp = tid >> 7;         // path group parity
    for (uint32_t pc = pBegin; pc < pEnd; pc += kChunk) {
        // global -> registers -> bf16 hi/lo -> LDS; each thread covers 4 groups of 4 paths
#pragma unroll
        for (int gi = 0; gi < 4; ++gi) {
            const uint32_t g = ld_grp + 2 * gi;          // 0..7
            uint32_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t p = pc + 4 * g + q;
                float2 av = make_float2(0.f, 0.f), bv = make_float2(0.f, 0.f);
                if (p < pEnd) {
                    if (l0 + ld_row < L) av = A[(size_t)p * L + l0 + ld_row];
                    if (v0 + ld_row < V) bv = B[(size_t)p * V + v0 + ld_row];
                }
                uint32_t hr = bf16bits(av.x), hi = bf16bits(av.y);
                uint32_t lr = bf16bits(av.x - bf16val(hr)), li = bf16bits(av.y - bf16val(hi));
                ah[q] = hr | (hi << 16);
                al[q] = lr | (li << 16);
                hr = bf16bits(bv.x); hi = bf16bits(bv.y);
                lr = bf16bits(bv.x - bf16val(hr)); li = bf16bits(bv.y - bf16val(hi));
                bh[q] = hr | (hi << 16);
                bl[q] = lr | (li << 16);
            }
            const uint32_t w = ld_row * kRowWords + 4 * g;
            *(u32x4*)&sAh[w] = u32x4{ah[0], ah[1], ah[2], ah[3]};
            *(u32x4*)&sAl[w] = u32x4{al[0], al[1], al[2], al[3]};
            *(u32x4*)&sBh[w] = u32x4{bh[0], bh[1], bh[2], bh[3]};
            *(u32x4*)&sBl[w] = u32x4{bl[0], bl[1], bl[2], bl[3]};
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {      // 4 x (k' = 16) per chunk of 32 paths
            const uint32_t kw = ks * 8 + (lane >> 5) * 4;   // word offset of this lane's 8 k' values
            bf16x8 aH[2], aL[2], bReH[2], bReL[2], bImH[2], bImL[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const uint32_t row = wl + 32 * i + (lane & 31);
                u32x4 h = *(const u32x4*)&sAh[row * kRowWords + kw];
                u32x4 lo = *(const u32x4*)&sAl[row * kRowWords + kw];
                aH[i] = __builtin_bit_cast(bf16x8, h);
                aL[i] = __builtin_bit_cast(bf16x8, lo);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const uint32_t row = wv + 32 * j + (lane & 31);
                u32x4 h = *(const u32x4*)&sBh[row * kRowWords + kw];
                u32x4 lo = *(const u32x4*)&sBl[row * kRowWords + kw];
                u32x4 hre, lre, him, lim;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    hre[q] = h[q] ^ 0x80000000u;                 // (re, -im)
                    lre[q] = lo[q] ^ 0x80000000u;
                    him[q] = (h[q] >> 16) | (h[q] << 16);        // (im, re)
                    lim[q] = (lo[q] >> 16) | (lo[q] << 16);
                }
                bReH[j] = __builtin_bit_cast(bf16x8, hre);
                bReL[j] = __builtin_bit_cast(bf16x8, lre);
                bImH[j] = __builtin_bit_cast(bf16x8, him);
                bImL[j] = __builtin_bit_cast(bf16x8, lim);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accRe[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aL[i], bReH[j], accRe[i][j], 0, 0, 0);
                    accRe[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH[i], bReL[j], accRe[i][j], 0, 0, 0);
                    accRe[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH[i], bReH[j], accRe[i][j], 0, 0, 0);
                    accIm[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aL[i], bImH[j], accIm[i][j], 0, 0, 0);
                    accIm[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH[i], bImL[j], accIm[i][j], 0, 0, 0);
                    accIm[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH[i], bImH[j], accIm[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // C layout of a 32x32 f32 MFMA tile: lane holds column (lane % 32), rows
    // 8*(r/4) + 4*(lane/32) + r%4 for accumulator r.
    float2* out = part + (size_t)split * L * V;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t v = v0 + wv + 32 * j + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t l = l0 + wl + 32 * i + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
                if (l < L && v < V) out[(size_t)v * L + l] = make_float2(accRe[i][j][r], accIm[i][j][r]);
            }
        }
}

// ---------------------------------------------------------------- bf16x3 on pre-split operands
// The same product from operands split once per frame into bf16 hi/lo planes: hi = bf16(re) |
// bf16(im) << 16 per path, lo likewise for the residuals x - hi. A plane is stored chunk-major:
// chunk k (paths 32k .. 32k+31) holds `rows` rows (an l or a v each) of 32 paths, 128 bytes a row,
// so the 128 rows a GEMM tile reads for one chunk are one contiguous 16 KB run. Rows past L (V)
// and paths past P are zeros. The GEMM then loads 16-byte runs of 4 paths (8 lanes: one row's
// 128-byte chunk) with no conversion. Its 72 KB of LDS (one chunk of the four planes) and at most
// 256 registers a lane let two workgroups share a CU, so one's loads and barriers overlap the
// other's MFMAs (measured on the 65536 x 1024 x 1024 batch: 1.36 ms against 1.68 ms for one
// workgroup a CU with two LDS buffers and register prefetch). The MFMA sequence (chunks in order,
// 4 k-steps each, lo*hi, hi*lo, hi*hi per tile) is that of k_contract_bf16x3, so both give the
// same bits.
__device__ inline size_t planeIndex(uint32_t rows, uint32_t c, uint32_t p) {
    return ((size_t)(p / kChunk) * rows + c) * kChunk + (p % kChunk);
}

// Paths whose row of X (P rows of C complex values) holds a nonzero value: flag[p] = 1, else 0.
// One wavefront a path.
__global__ void __launch_bounds__(256) k_row_nonzero(const float2* __restrict__ X, uint32_t P, uint32_t C,
                                                     uint32_t* __restrict__ flag) {
    const uint32_t p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (p >= P) return;
    const float2* row = X + (size_t)p * C;
    bool any = false;
    for (uint32_t c = lane; c < C; c += 64) {
        const float2 x = row[c];
        any = any || x.x != 0.f || x.y != 0.f;
    }
    const uint64_t b = __ballot(any);
    if (lane == 0) flag[p] = b ? 1u : 0u;
}

// The flagged paths in increasing order (list[0 .. *count)), by one workgroup of 1024 threads:
// each thread counts a contiguous range of flags, an exclusive scan of the counts gives its first
// slot, and it writes its paths there. The list and the count depend on the flags only.
__global__ void __launch_bounds__(1024) k_compact_paths(const uint32_t* __restrict__ flag, uint32_t P,
                                                        uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    __shared__ uint32_t waveTot[16];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t per = (P + 1023) / 1024, b = min(P, t * per), e = min(P, b + per);
    uint32_t c = 0;
    for (uint32_t p = b; p < e; ++p) c += flag[p] ? 1u : 0u;
    // inclusive scan over the wavefront, then over the 16 wavefronts' totals
    uint32_t x = c;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) waveTot[w] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t j = 0; j < w; ++j) before += waveTot[j];
    uint32_t slot = before + x - c;
    for (uint32_t p = b; p < e; ++p)
        if (flag[p]) list[slot++] = p;
    if (t == 1023) *count = slot;
}

// One workgroup splits 64 paths x 64 rows: reads along the rows (coalesced float2), transposes
// through LDS, writes 2 chunks x 64 rows x 128 bytes per plane (with a path list: the listed paths
// in list order, zeros past the count). It also records, for each of its
// two (chunk, 64-row unit) blocks, whether any value in it is nonzero (nz[chunk * rows / 64 +
// unit]), and writes the planes of nonzero blocks only: the GEMM reads a block's planes only when
// its flag is set, so a zero block's planes may hold anything.
__global__ void __launch_bounds__(256) k_split_bf16(const float2* __restrict__ X, uint32_t P, uint32_t C,
                                                    uint32_t rows, uint32_t* __restrict__ hi,
                                                    uint32_t* __restrict__ lo, uint32_t* __restrict__ nz,
                                                    const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count) {
    __shared__ uint32_t th[64][65], tl[64][65];
    const uint32_t p0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // with a path list (k_compact_paths), plane path j holds path list[j] for j < *count
    const uint32_t n = list ? *count : P;
    // all 16 loads of a thread are issued before the first conversion
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t p = p0 + ty + 4 * k, c = c0 + tx;
        const uint32_t src = (list && p < n) ? list[p] : p;
        v[k] = (p < n && c < C) ? X[(size_t)src * C + c] : make_float2(0.f, 0.f);
    }
    // paths p0 + ty + 4 k: k < 8 in the first chunk, the rest in the second
    bool n0 = false, n1 = false;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const bool n = v[k].x != 0.f || v[k].y != 0.f;
        if (k < 8) n0 = n0 || n;
        else n1 = n1 || n;
    }
    const bool any0 = __syncthreads_or(n0) != 0, any1 = __syncthreads_or(n1) != 0;
    if (threadIdx.x == 0) {
        const uint32_t units = rows / 64, chunk = p0 / kChunk;
        nz[(size_t)chunk * units + c0 / 64] = any0 ? 1u : 0u;
        nz[(size_t)(chunk + 1) * units + c0 / 64] = any1 ? 1u : 0u;
    }
    if (!any0 && !any1) return;   // the whole workgroup: nothing to write
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t hr = bf16bits(v[k].x), hm = bf16bits(v[k].y);
        th[tx][ty + 4 * k] = hr | (hm << 16);
        tl[tx][ty + 4 * k] = bf16bits(v[k].x - bf16val(hr)) | (bf16bits(v[k].y - bf16val(hm)) << 16);
    }
    __syncthreads();
    // every (row, path) of the padded planes belongs to exactly one workgroup: no bounds checks
    if (tx < 32 ? !any0 : !any1) return;
#pragma unroll
    for (uint32_t j = ty; j < 64; j += 4) {
        const size_t o = planeIndex(rows, c0 + j, p0 + tx);
        hi[o] = th[j][tx];
        lo[o] = tl[j][tx];
    }
}

__global__ void __launch_bounds__(256, 2) k_contract_bf16x3_planes(const uint32_t* __restrict__ Ah,
                                                                const uint32_t* __restrict__ Al,
                                                                const uint32_t* __restrict__ Bh,
                                                                const uint32_t* __restrict__ Bl, uint32_t rowsA,
                                                                uint32_t rowsB,
                                                                uint32_t P, uint32_t L, uint32_t V,
                                                                uint32_t pathsPerSplit, const uint32_t* __restrict__ nzA,
                                                                const uint32_t* __restrict__ nzB,
                                                                const uint32_t* __restrict__ count,
                                                                float2* __restrict__ part) {
    __shared__ uint32_t sAh[kTileL * kRowWords], sAl[kTileL * kRowWords];
    __shared__ uint32_t sBh[kTileV * kRowWords], sBl[kTileV * kRowWords];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const TileIdx tile = xcdTile();
    const uint32_t l0 = tile.x * kTileL, v0 = tile.y * kTileV, split = tile.z;
    // with a device count (the planes hold the listed paths first), the count's paths split evenly
    // over the launch's slices
    uint32_t pps = pathsPerSplit;
    if (count) {
        P = *count;
        pps = ((P + gridDim.z - 1) / gridDim.z + kChunk - 1) / kChunk * kChunk;
    }
    const uint32_t pBegin = min(P, split * pps);
    const uint32_t pEnd = min(P, pBegin + pps);
    const uint32_t nChunks = (pEnd - pBegin + kChunk - 1) / kChunk;   // the planes hold zeros past P
    const uint32_t wl = (wave & 1) * 64, wv = (wave >> 1) * 64;
    f32x16 accRe[2][2], accIm[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) { accRe[i][j][r] = 0.f; accIm[i][j][r] = 0.f; }

    // loader: segment `seg` (4 paths = 16 bytes) of rows rbase + 32 q, q = 0..3, of all 4 planes
    const uint32_t seg = tid & 7, rbase = tid >> 3;
    auto stage = [&](uint32_t pc) {
        u32x4 rg[4][4];   // [plane][q]
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t ra = planeIndex(rowsA, l0 + rbase + 32 * q, pc + 4 * seg);
            const size_t rb = planeIndex(rowsB, v0 + rbase + 32 * q, pc + 4 * seg);
            rg[0][q] = *(const u32x4*)(Ah + ra);
            rg[1][q] = *(const u32x4*)(Al + ra);
            rg[2][q] = *(const u32x4*)(Bh + rb);
            rg[3][q] = *(const u32x4*)(Bl + rb);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t w = (rbase + 32 * q) * kRowWords + 4 * seg;
            *(u32x4*)&sAh[w] = rg[0][q];
            *(u32x4*)&sAl[w] = rg[1][q];
            *(u32x4*)&sBh[w] = rg[2][q];
            *(u32x4*)&sBl[w] = rg[3][q];
        }
    };
    // one k-step's operands as read from LDS: A hi/lo of the wave's 2 row tiles, B hi/lo of its 2
    // column tiles
    struct Frag { u32x4 ah[2], al[2], bh[2], bl[2]; };
    auto readFrag = [&](Frag& f, int ks) {
        const uint32_t kw = ks * 8 + (lane >> 5) * 4;   // word offset of this lane's 8 k' values
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const uint32_t ra = (wl + 32 * i + (lane & 31)) * kRowWords + kw;
            const uint32_t rb = (wv + 32 * i + (lane & 31)) * kRowWords + kw;
            f.ah[i] = *(const u32x4*)&sAh[ra];
            f.al[i] = *(const u32x4*)&sAl[ra];
            f.bh[i] = *(const u32x4*)&sBh[rb];
            f.bl[i] = *(const u32x4*)&sBl[rb];
        }
    };
    auto mfmas = [&](const Frag& f) {
        bf16x8 bReH[2], bReL[2], bImH[2], bImL[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            u32x4 hre, lre, him, lim;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                hre[q] = f.bh[j][q] ^ 0x80000000u;                       // (re, -im)
                lre[q] = f.bl[j][q] ^ 0x80000000u;
                him[q] = (f.bh[j][q] >> 16) | (f.bh[j][q] << 16);        // (im, re)
                lim[q] = (f.bl[j][q] >> 16) | (f.bl[j][q] << 16);
            }
            bReH[j] = __builtin_bit_cast(bf16x8, hre);
            bReL[j] = __builtin_bit_cast(bf16x8, lre);
            bImH[j] = __builtin_bit_cast(bf16x8, him);
            bImL[j] = __builtin_bit_cast(bf16x8, lim);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 aH = __builtin_bit_cast(bf16x8, f.ah[i]), aL = __builtin_bit_cast(bf16x8, f.al[i]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accRe[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aL, bReH[j], accRe[i][j], 0, 0, 0);
                accRe[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH, bReL[j], accRe[i][j], 0, 0, 0);
                accRe[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH, bReH[j], accRe[i][j], 0, 0, 0);
                accIm[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aL, bImH[j], accIm[i][j], 0, 0, 0);
                accIm[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH, bImL[j], accIm[i][j], 0, 0, 0);
                accIm[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aH, bImH[j], accIm[i][j], 0, 0, 0);
            }
        }
    };
    // a chunk whose A rows or B rows of this tile are all zero (k_split_bf16's flags) adds exact
    // zeros: skipped, with the same result
    const uint32_t unitsA = rowsA / 64, unitsB = rowsB / 64, ua = l0 / 64, ub = v0 / 64;
    static_assert(kTileL == 128 && kTileV == 128, "a tile spans two 64-row flag units");
    for (uint32_t c = 0; c < nChunks; ++c) {
        const uint32_t cg = pBegin / kChunk + c;
        const uint32_t* fa = nzA + (size_t)cg * unitsA + ua;
        const uint32_t* fb = nzB + (size_t)cg * unitsB + ub;
        if (!((fa[0] | fa[1]) && (fb[0] | fb[1]))) continue;
        stage(pBegin + c * kChunk);
        __syncthreads();
        // 4 k-steps (k' = 16) per chunk of 32 paths
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            Frag f;
            readFrag(f, ks);
            mfmas(f);
        }
        __syncthreads();
    }
    float2* out = part + (size_t)split * L * V;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t v = v0 + wv + 32 * j + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t l = l0 + wl + 32 * i + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
                if (l < L && v < V) out[(size_t)v * L + l] = make_float2(accRe[i][j][r], accIm[i][j][r]);
            }
        }
}

// ---------------------------------------------------------------- FMA reference kernels
template <typename T> struct C2T;
template <> struct C2T<float> { typedef float2 T; };
template <> struct C2T<double> { typedef double2 T; };

template <typename T>
__global__ void __launch_bounds__(256) k_contract_fma(const typename C2T<T>::T* __restrict__ A,
                                                      const typename C2T<T>::T* __restrict__ B, uint32_t P,
                                                      uint32_t L, uint32_t V, uint32_t pathsPerSplit,
                                                      typename C2T<T>::T* __restrict__ part) {
    typedef typename C2T<T>::T C;
    __shared__ C sA[16][64 + 1], sB[16][64 + 1];
    const uint32_t tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const uint32_t l0 = blockIdx.x * 64, v0 = blockIdx.y * 64;
    const uint32_t pBegin = blockIdx.z * pathsPerSplit, pEnd = min(P, pBegin + pathsPerSplit);
    T re[4][4], im[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) { re[i][j] = 0; im[i][j] = 0; }
    for (uint32_t pc = pBegin; pc < pEnd; pc += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t e = tid + 256 * q, kk = e >> 6, c = e & 63;
            const uint32_t p = pc + kk;
            C av, bv;
            av.x = av.y = 0; bv.x = bv.y = 0;
            if (p < pEnd && l0 + c < L) av = A[(size_t)p * L + l0 + c];
            if (p < pEnd && v0 + c < V) bv = B[(size_t)p * V + v0 + c];
            sA[kk][c] = av;
            sB[kk][c] = bv;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            C a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) { a[i] = sA[kk][tx + 16 * i]; b[i] = sB[kk][ty + 16 * i]; }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    re[i][j] = fma(a[i].x, b[j].x, fma(-a[i].y, b[j].y, re[i][j]));
                    im[i][j] = fma(a[i].x, b[j].y, fma(a[i].y, b[j].x, im[i][j]));
                }
        }
        __syncthreads();
    }
    C* out = part + (size_t)blockIdx.z * L * V;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t l = l0 + tx + 16 * i, v = v0 + ty + 16 * j;
            if (l < L && v < V) {
                C c;
                c.x = re[i][j];
                c.y = im[i][j];
                out[(size_t)v * L + l] = c;
            }
        }
}

// Fixed-order reduction of split partials into the FP64 accumulator, with the elementwise
// single-scattering factor F (complex; |F|^2 for correlations, S(0) conj S(t)) and a scalar.
template <typename C>
__global__ void __launch_bounds__(256) k_reduce(const C* __restrict__ part, uint32_t splits, uint32_t LV,
                                                const float2* __restrict__ F, const double2* __restrict__ Fd, int fsq,
                                                double fscalar, double* __restrict__ U) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= LV) return;
    double re = 0, im = 0;
    for (uint32_t s = 0; s < splits; ++s) {
        C c = part[(size_t)s * LV + i];
        re += (double)c.x;
        im += (double)c.y;
    }
    double fr = 1.0, fi = 0.0;
    if (Fd) { fr = Fd[i].x; fi = Fd[i].y; }
    else if (F) { fr = (double)F[i].x; fi = (double)F[i].y; }
    if (fsq) { fr = fr * fr + fi * fi; fi = 0.0; }
    U[2 * (size_t)i] += fscalar * (fr * re - fi * im);
    U[2 * (size_t)i + 1] += fscalar * (fr * im + fi * re);
}

static uint32_t chooseSplits(uint32_t tiles, uint32_t P, uint32_t minPaths) {
    uint32_t want = (1024 + tiles - 1) / tiles;
    uint32_t maxByP = (P + minPaths - 1) / minPaths;
    uint32_t s = want < maxByP ? want : maxByP;
    return s < 1 ? 1 : s;
}

// Split-K plan: `splits` slices of `pps` paths each (whole LDS chunks), at most maxSplits; for
// bf16x3 with `planes`, also the bf16 hi/lo operand planes (rows padded to the tile, paths to
// ldp) that follow the partials in the workspace.
struct SplitPlan {
    uint32_t splits, pps;
    size_t bytes;        // partials: splits x L x V complex
    uint32_t ldp = 0;    // planes: padded path count (0: no planes, the GEMM converts as it loads)
    uint32_t rowsA = 0, rowsB = 0;
    size_t planeOffset = 0, planeBytes = 0;
    size_t total() const { return planeBytes ? planeOffset + planeBytes : bytes; }
};
static SplitPlan planSplits(int kind, uint32_t P, uint32_t L, uint32_t V, uint32_t maxSplits, bool planes = false) {
    const uint32_t tl = kind == kContractBf16x3 ? kTileL : 64, tv = kind == kContractBf16x3 ? kTileV : 64;
    const uint32_t tiles = ((L + tl - 1) / tl) * ((V + tv - 1) / tv);
    uint32_t splits = chooseSplits(tiles, P, kind == kContractBf16x3 ? 512 : 256);
    if (splits > maxSplits) splits = maxSplits;
    if (splits < 1) splits = 1;
    uint32_t pps = (P + splits - 1) / splits;
    if (kind == kContractBf16x3) pps = (pps + kChunk - 1) / kChunk * kChunk;
    else pps = (pps + 15) / 16 * 16;
    splits = (P + pps - 1) / pps;
    if (splits < 1) splits = 1;
    const size_t elem = kind == kContractFp64 ? 16 : 8;
    SplitPlan plan{splits, pps, (size_t)splits * L * V * elem};
    if (planes && kind == kContractBf16x3) {
        plan.ldp = (P + 63) / 64 * 64;
        plan.rowsA = (L + kTileL - 1) / kTileL * kTileL;
        plan.rowsB = (V + kTileV - 1) / kTileV * kTileV;
        plan.planeOffset = (plan.bytes + 255) / 256 * 256;
        // hi and lo planes of both operands, the nonzero flags of their (chunk, 64-row) blocks, then
        // the path flags, the path list and its count
        plan.planeBytes = 2ull * ((size_t)plan.rowsA + plan.rowsB) * plan.ldp * sizeof(uint32_t) +
                          (size_t)(plan.ldp / kChunk) * ((plan.rowsA + plan.rowsB) / 64) * sizeof(uint32_t) +
                          (2ull * plan.ldp + 64) * sizeof(uint32_t);
    }
    return plan;
}

static hipError_t launchContractInto(int kind, const void* A, const void* B, uint32_t P, uint32_t L, uint32_t V,
                                     uint32_t nt, const float2* F, const double2* Fd, int fsq, double fscalar,
                                     double* U, const SplitPlan& plan, void* partials, hipStream_t st);

// The split-K workspace belongs to the caller (one per tracer), grown here when too small.
hipError_t launchContract(int kind, const void* A, const void* B, uint32_t P, uint32_t L, uint32_t V, uint32_t nt,
                          const float2* F, const double2* Fd, int fsq, double fscalar, double* U, Workspace* ws,
                          hipStream_t st) {
    const SplitPlan plan = planSplits(kind, P, L, V, 0xffffffffu, true);
    if (ws->bytes < plan.total()) {
        if (ws->ptr) {
            hipStreamSynchronize(st);
            hipFree(ws->ptr);
        }
        ws->ptr = nullptr;
        ws->bytes = 0;
        hipError_t e = hipMalloc(&ws->ptr, plan.total());
        if (e != hipSuccess) return e;
        ws->bytes = plan.total();
    }
    return launchContractInto(kind, A, B, P, L, V, nt, F, Fd, fsq, fscalar, U, plan, ws->ptr, st);
}

static hipError_t launchContractInto(int kind, const void* A, const void* B, uint32_t P, uint32_t L, uint32_t V,
                                     uint32_t nt, const float2* F, const double2* Fd, int fsq, double fscalar,
                                     double* U, const SplitPlan& plan, void* partials, hipStream_t st) {
    const uint32_t tl = kind == kContractBf16x3 ? kTileL : 64, tv = kind == kContractBf16x3 ? kTileV : 64;
    const uint32_t splits = plan.splits, pps = plan.pps;
    const uint32_t LV = L * V;
    for (uint32_t t = 0; t < nt; ++t) {
        dim3 grid((L + tl - 1) / tl, (V + tv - 1) / tv, splits);
        if (kind == kContractBf16x3 && plan.ldp) {
            const float2* a = (const float2*)A + (size_t)t * P * L;
            const float2* b = (const float2*)B + (size_t)t * P * V;
            uint32_t* ah = (uint32_t*)((char*)partials + plan.planeOffset);
            uint32_t* al = ah + (size_t)plan.rowsA * plan.ldp;
            uint32_t* bh = al + (size_t)plan.rowsA * plan.ldp;
            uint32_t* bl = bh + (size_t)plan.rowsB * plan.ldp;
            uint32_t* nza = bl + (size_t)plan.rowsB * plan.ldp;
            uint32_t* nzb = nza + (size_t)(plan.ldp / kChunk) * (plan.rowsA / 64);
            uint32_t* flag = nzb + (size_t)(plan.ldp / kChunk) * (plan.rowsB / 64);
            uint32_t* list = flag + plan.ldp;
            uint32_t* count = list + plan.ldp;
            // the paths whose B row is all zero add nothing: only the others are split and
            // multiplied, in path order (the view-side terms of most paths are cut to zero in thick
            // media)
            hipLaunchKernelGGL(k_row_nonzero, dim3((P + 3) / 4), dim3(256), 0, st, b, P, V, flag);
            hipLaunchKernelGGL(k_compact_paths, dim3(1), dim3(1024), 0, st, flag, P, list, count);
            hipLaunchKernelGGL(k_split_bf16, dim3(plan.ldp / 64, plan.rowsA / 64), dim3(256), 0, st, a, P, L,
                               plan.rowsA, ah, al, nza, list, count);
            hipLaunchKernelGGL(k_split_bf16, dim3(plan.ldp / 64, plan.rowsB / 64), dim3(256), 0, st, b, P, V,
                               plan.rowsB, bh, bl, nzb, list, count);
            hipLaunchKernelGGL(k_contract_bf16x3_planes, grid, dim3(256), 0, st, ah, al, bh, bl, plan.rowsA,
                               plan.rowsB, P, L, V, pps, nza, nzb, count, (float2*)partials);
            hipLaunchKernelGGL(k_reduce<float2>, dim3((LV + 255) / 256), dim3(256), 0, st, (const float2*)partials,
                               splits, LV, F, Fd, fsq, fscalar, U + 2 * (size_t)t * LV);
        } else if (kind == kContractBf16x3) {
            const float2* a = (const float2*)A + (size_t)t * P * L;
            const float2* b = (const float2*)B + (size_t)t * P * V;
            hipLaunchKernelGGL(k_contract_bf16x3, grid, dim3(256), 0, st, a, b, P, L, V, pps, (float2*)partials);
            hipLaunchKernelGGL(k_reduce<float2>, dim3((LV + 255) / 256), dim3(256), 0, st, (const float2*)partials,
                               splits, LV, F, Fd, fsq, fscalar, U + 2 * (size_t)t * LV);
        } else if (kind == kContractFp32) {
            const float2* a = (const float2*)A + (size_t)t * P * L;
            const float2* b = (const float2*)B + (size_t)t * P * V;
            hipLaunchKernelGGL(k_contract_fma<float>, grid, dim3(256), 0, st, a, b, P, L, V, pps, (float2*)partials);
            hipLaunchKernelGGL(k_reduce<float2>, dim3((LV + 255) / 256), dim3(256), 0, st, (const float2*)partials,
                               splits, LV, F, Fd, fsq, fscalar, U + 2 * (size_t)t * LV);
        } else {
            const double2* a = (const double2*)A + (size_t)t * P * L;
            const double2* b = (const double2*)B + (size_t)t * P * V;
            hipLaunchKernelGGL(k_contract_fma<double>, grid, dim3(256), 0, st, a, b, P, L, V, pps, (double2*)partials);
            hipLaunchKernelGGL(k_reduce<double2>, dim3((LV + 255) / 256), dim3(256), 0, st, (const double2*)partials,
                               splits, LV, F, Fd, fsq, fscalar, U + 2 * (size_t)t * LV);
        }
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- allocation-free contraction
// sstmc_contract_noalloc keeps its split-K partials in this fixed array instead of a workspace it
// allocates: each device the library runs on holds one copy (64 MiB) until the process ends, and
// every call on a device uses that device's copy. The kernels write every partial they later
// reduce, so nothing in it needs clearing between calls; calls that overlap on one device race.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on the current device, taken once per device with it current
// (on the host the array's name is a stand-in, not a device address).
constexpr int kMaxDevices = 64;
static std::mutex g_scratchMu;
static void* g_scratchPtr[kMaxDevices] = {};

static hipError_t scratchAddress(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return hipSuccess;
}

}  // namespace sstmc

extern "C" int sstmc_contract_noalloc(int32_t contraction, uint32_t npaths, uint32_t L, uint32_t V, uint32_t nt,
                                      const void* a, const void* b, double* u, void* stream) {
    using namespace sstmc;
    if (npaths == 0 || L == 0 || V == 0 || nt == 0 || !a || !b || !u) return SSTMC_INVALID_CONFIG;
    int kind;
    if (contraction == SSTMC_CONTRACT_BF16X3) kind = kContractBf16x3;
    else if (contraction == SSTMC_CONTRACT_FP32) kind = kContractFp32;
    else if (contraction == SSTMC_CONTRACT_FP64) kind = kContractFp64;
    else return SSTMC_INVALID_CONFIG;
    // one split's partials (L x V complex) must fit; more splits take what the array holds
    const size_t slice = (size_t)L * V * (kind == kContractFp64 ? 16u : 8u);
    if (slice > kScratchBytes) return SSTMC_SCRATCH_EXCEEDED;
    const SplitPlan plan = planSplits(kind, npaths, L, V, (uint32_t)(kScratchBytes / slice));
    if (plan.bytes > kScratchBytes) return SSTMC_SCRATCH_EXCEEDED;
    void* ptr = nullptr;
    if (scratchAddress(&ptr) != hipSuccess) return SSTMC_DEVICE_ERROR;
    const hipError_t e = launchContractInto(kind, a, b, npaths, L, V, nt, nullptr, nullptr, 0, 1.0, u, plan, ptr,
                                            (hipStream_t)stream);
    return e == hipSuccess ? SSTMC_NO_ERROR : SSTMC_KERNEL_ERROR;
}

namespace sstmc {

// ---------------------------------------------------------------- generic single scattering
// S_p[l,v,t] = a_p[l,t] b_p[v,t] f(l -> x1^t -> v) when f depends on the first vertex (point
// sources with anisotropic scattering, or per-material phase functions): threePointFunctionSingle
// (FarField.cuh:546-599, PointSource.cuh:363-446) fused with the sum over paths. Block = 16 x 16
// (l, v) outputs; per chunk of 16 paths the unit directions and terms are staged in LDS.
template <typename R> struct SingleT;
template <> struct SingleT<float> {
    typedef float2 C;
    __device__ static float sq(float x) { return sqrtf(x); }
    __device__ static float ex(float x) { return __expf(x); }
};
template <> struct SingleT<double> {
    typedef double2 C;
    __device__ static double sq(double x) { return sqrt(x); }
    __device__ static double ex(double x) { return exp(x); }
};

// Single-scattering amplitude at cosine c (complex: a tabular amplitude carries a phase).
template <int DIMS, typename R>
__device__ inline typename SingleT<R>::C singleAmp(const ScatterView& s, uint32_t m, R c) {
    typename SingleT<R>::C f;
    f.y = 0;
    if (s.type[m] == 1) {
        f.x = DIMS == 3 ? (R)0.282094791773878143 : (R)0.398942280401432678;
    } else if (s.type[m] == 2) {
        const PhaseTables& t = *s.tab;
        const uint32_t i = t.off[m] + tabularEntry(t.n[m], DIMS, (double)c);
        if constexpr (sizeof(R) == 4) { f.x = t.f[i].y; f.y = t.f[i].z; }
        else { f.x = t.d[i].y; f.y = t.d[i].z; }
    } else if (s.type[m] == 4) {
        f.x = DIMS == 3 ? (R)vmfAmplitude(*s.tab, m, (double)c) : (R)0.398942280401432678;
    } else {
        R g = (R)s.gd[m];
        R up = DIMS == 3 ? ((R)1 - g * g) / (R)12.566370614359172954 : ((R)1 - g * g) / (R)6.283185307179586477;
        R den = fma((R)-2 * g, c, (R)1 + g * g);
        f.x = SingleT<R>::sq(DIMS == 3 ? up / (den * SingleT<R>::sq(den)) : up / den);
    }
    return f;
}

template <int DIMS, typename R>
__global__ void __launch_bounds__(256) k_single_generic(ScatterView sc, SourceView il, SourceView vw, FramesView fr,
                                                        VertexBatch vb, const typename SingleT<R>::C* __restrict__ a,
                                                        const typename SingleT<R>::C* __restrict__ b, uint32_t P,
                                                        uint32_t pathsPerSplit, uint32_t splits,
                                                        typename SingleT<R>::C* __restrict__ part, int corr) {
    typedef typename SingleT<R>::C C;
    __shared__ R su[16][16][3], sw[16][16][3], su0[16][16][3], sw0[16][16][3];
    __shared__ C sa[16][16], sb[16][16];
    __shared__ uint32_t sm[16];
    const uint32_t L = il.n, V = vw.n;
    const uint32_t tl = threadIdx.x & 15, tq = threadIdx.x >> 4;
    const uint32_t split = blockIdx.z % splits, t = blockIdx.z / splits;
    const uint32_t l = blockIdx.x * 16 + tl, v = blockIdx.y * 16 + tq;
    const uint32_t pBegin = split * pathsPerSplit, pEnd = min(P, pBegin + pathsPerSplit);
    const bool iso = il.type == 1 && sc.allIsotropic;
    R re = 0, im = 0;
    for (uint32_t pc = pBegin; pc < pEnd; pc += 16) {
        const uint32_t p = pc + tq;
        const uint32_t ls = blockIdx.x * 16 + tl, vs = blockIdx.y * 16 + tl;
        C av, bv;
        av.x = av.y = 0; bv.x = bv.y = 0;
        double x1[3] = {0, 0, 0}, x10[3] = {0, 0, 0};
        uint32_t m1 = 0;
        bool ok = p < pEnd && vb.k[p] > 0 && (uint64_t)vb.off[P] <= vb.cap;
        if (ok) {
            const uint64_t o = vb.off[p];
            const float4 xm = vb.xm[o];
            m1 = __float_as_uint(xm.w);
            x1[0] = xm.x; x1[1] = xm.y; x1[2] = xm.z;
            x10[0] = x1[0]; x10[1] = x1[1]; x10[2] = x1[2];
            if (fr.temporalAmp && !fr.trivial) {
                const float* d = vb.delta + (o * fr.nt + t) * 3;
                const float* d0 = vb.delta + o * fr.nt * 3;
                x1[0] += (double)d[0]; x1[1] += (double)d[1]; x1[2] += (double)d[2];
                x10[0] += (double)d0[0]; x10[1] += (double)d0[1]; x10[2] += (double)d0[2];
            }
            if (ls < L) av = a[((size_t)t * P + p) * L + ls];
            if (vs < V) bv = b[((size_t)t * P + p) * V + vs];
        }
        // incoming direction of illumination ls at x1, outgoing direction toward view vs (frame t;
        // with the correlation output also at frame 0)
        for (int f = 0; f < (corr ? 2 : 1); ++f) {
            const double* xp = f == 0 ? x1 : x10;
            double u[3] = {0, 0, 1}, w[3] = {0, 0, 1};
            if (ls < L) {
                if (il.type == 2) { u[0] = il.pd[3 * ls]; u[1] = il.pd[3 * ls + 1]; u[2] = il.pd[3 * ls + 2]; }
                else {
                    double q[3] = {xp[0] - il.pd[3 * ls], xp[1] - il.pd[3 * ls + 1],
                                   DIMS == 3 ? xp[2] - il.pd[3 * ls + 2] : 0.0};
                    double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                    if (n > 0) { u[0] = q[0] / n; u[1] = q[1] / n; u[2] = q[2] / n; }
                }
            }
            if (vs < V) {
                if (vw.type == 2) { w[0] = vw.pd[3 * vs]; w[1] = vw.pd[3 * vs + 1]; w[2] = vw.pd[3 * vs + 2]; }
                else {
                    double q[3] = {vw.pd[3 * vs] - xp[0], vw.pd[3 * vs + 1] - xp[1],
                                   DIMS == 3 ? vw.pd[3 * vs + 2] - xp[2] : 0.0};
                    double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                    if (n > 0) { w[0] = q[0] / n; w[1] = q[1] / n; w[2] = q[2] / n; }
                }
            }
            if (DIMS == 2) { u[2] = 0; w[2] = 0; }
            R(*du)[3] = f == 0 ? su[tq] : su0[tq];
            R(*dw)[3] = f == 0 ? sw[tq] : sw0[tq];
            du[tl][0] = (R)u[0]; du[tl][1] = (R)u[1]; du[tl][2] = (R)u[2];
            dw[tl][0] = (R)w[0]; dw[tl][1] = (R)w[1]; dw[tl][2] = (R)w[2];
        }
        sa[tq][tl] = av;
        sb[tq][tl] = bv;
        if (tl == 0) sm[tq] = m1;
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            C x = sa[q][tl], y = sb[q][tq];
            R c = su[q][tl][0] * sw[q][tq][0] + su[q][tl][1] * sw[q][tq][1] + su[q][tl][2] * sw[q][tq][2];
            C f;
            if (iso) { f.x = DIMS == 3 ? (R)0.282094791773878143 : (R)0.398942280401432678; f.y = 0; }
            else f = singleAmp<DIMS, R>(sc, sm[q], c);
            if (corr) {   // S(0) conj S(t): the amplitude at frame 0 times the conjugate at frame t
                const R c0 = su0[q][tl][0] * sw0[q][tq][0] + su0[q][tl][1] * sw0[q][tq][1] + su0[q][tl][2] * sw0[q][tq][2];
                C f0;
                if (iso) { f0.x = f.x; f0.y = 0; }
                else f0 = singleAmp<DIMS, R>(sc, sm[q], c0);
                const R fr0 = f0.x * f.x + f0.y * f.y, fi0 = f0.y * f.x - f0.x * f.y;
                f.x = fr0;
                f.y = fi0;
            }
            R pr = x.x * y.x - x.y * y.y, pi = x.x * y.y + x.y * y.x;
            re = fma(f.x, pr, fma(-f.y, pi, re));
            im = fma(f.x, pi, fma(f.y, pr, im));
        }
        __syncthreads();
    }
    if (l < L && v < V) {
        C c;
        c.x = re;
        c.y = im;
        part[((size_t)blockIdx.z) * L * V + (size_t)v * L + l] = c;
    }
}

hipError_t launchSingleGeneric(int dims, int fp64, const ScatterView& sc, const SourceView& il,
                               const SourceView& vw, const FramesView& fr, const VertexBatch& vb, const void* a,
                               const void* b, uint32_t P, int corr, double* U, Workspace* ws, hipStream_t st) {
    const uint32_t L = il.n, V = vw.n, NT = fr.nt;
    uint32_t tiles = ((L + 15) / 16) * ((V + 15) / 16);
    uint32_t splits = chooseSplits(tiles, P, 256);
    uint32_t pps = ((P + splits - 1) / splits + 15) / 16 * 16;
    splits = (P + pps - 1) / pps;
    if (splits < 1) splits = 1;
    size_t elem = fp64 ? 16 : 8;
    size_t need = (size_t)splits * NT * L * V * elem;
    if (ws->bytes < need) {
        if (ws->ptr) {
            hipStreamSynchronize(st);
            hipFree(ws->ptr);
        }
        ws->ptr = nullptr;
        ws->bytes = 0;
        hipError_t e = hipMalloc(&ws->ptr, need);
        if (e != hipSuccess) return e;
        ws->bytes = need;
    }
    dim3 grid((L + 15) / 16, (V + 15) / 16, splits * NT);
    const uint32_t LV = L * V;
    if (fp64) {
        if (dims == 2)
            hipLaunchKernelGGL((k_single_generic<2, double>), grid, dim3(256), 0, st, sc, il, vw, fr, vb,
                               (const double2*)a, (const double2*)b, P, pps, splits, (double2*)ws->ptr, corr);
        else
            hipLaunchKernelGGL((k_single_generic<3, double>), grid, dim3(256), 0, st, sc, il, vw, fr, vb,
                               (const double2*)a, (const double2*)b, P, pps, splits, (double2*)ws->ptr, corr);
        for (uint32_t t = 0; t < NT; ++t)
            hipLaunchKernelGGL(k_reduce<double2>, dim3((LV + 255) / 256), dim3(256), 0, st,
                               (const double2*)ws->ptr + (size_t)t * splits * LV, splits, LV, (const float2*)nullptr,
                               (const double2*)nullptr, 0, 1.0, U + 2 * (size_t)t * LV);
    } else {
        if (dims == 2)
            hipLaunchKernelGGL((k_single_generic<2, float>), grid, dim3(256), 0, st, sc, il, vw, fr, vb,
                               (const float2*)a, (const float2*)b, P, pps, splits, (float2*)ws->ptr, corr);
        else
            hipLaunchKernelGGL((k_single_generic<3, float>), grid, dim3(256), 0, st, sc, il, vw, fr, vb,
                               (const float2*)a, (const float2*)b, P, pps, splits, (float2*)ws->ptr, corr);
        for (uint32_t t = 0; t < NT; ++t)
            hipLaunchKernelGGL(k_reduce<float2>, dim3((LV + 255) / 256), dim3(256), 0, st,
                               (const float2*)ws->ptr + (size_t)t * splits * LV, splits, LV, (const float2*)nullptr,
                               (const double2*)nullptr, 0, 1.0, U + 2 * (size_t)t * LV);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- GaussianBeam single scattering
// S_p[l,v] = a_p[l] b_p[v] fs(l, v, x1) with fs the two beams' throughputs through the vMF lobes of
// x1's material (fsGaussianBeam, GaussianBeam.cuh:894-999), FP64 whatever the term precision.
// Block = 16 x 16 (l, v) outputs; per chunk of 16 paths the first vertices, their materials and
// the terms are staged in LDS. DIMS = 2: the I_0 form of fs.
template <int DIMS, typename R>
__global__ void __launch_bounds__(256) k_single_gauss(MediumView m, ScatterView sc, SourceView il, SourceView vw,
                                                      FramesView fr, VertexBatch vb,
                                                      const typename SingleT<R>::C* __restrict__ a,
                                                      const typename SingleT<R>::C* __restrict__ b, uint32_t P,
                                                      uint32_t pathsPerSplit, uint32_t splits,
                                                      typename SingleT<R>::C* __restrict__ part, int corr) {
    typedef typename SingleT<R>::C C;
    __shared__ double sx[16][3], sx0[16][3];
    __shared__ uint32_t sm[16], sm0[16];
    __shared__ C sa[16][16], sb[16][16];
    const uint32_t L = il.n, V = vw.n;
    const uint32_t tl = threadIdx.x & 15, tq = threadIdx.x >> 4;
    const uint32_t split = blockIdx.z % splits, t = blockIdx.z / splits;
    const uint32_t l = blockIdx.x * 16 + tl, v = blockIdx.y * 16 + tq;
    const uint32_t pBegin = split * pathsPerSplit, pEnd = min(P, pBegin + pathsPerSplit);
    const PhaseTables& tab = *sc.tab;
    double re = 0, im = 0;
    for (uint32_t pc = pBegin; pc < pEnd; pc += 16) {
        const uint32_t p = pc + tq;
        const uint32_t ls = blockIdx.x * 16 + tl, vs = blockIdx.y * 16 + tl;
        C av, bv;
        av.x = av.y = 0; bv.x = bv.y = 0;
        const bool ok = p < pEnd && vb.k[p] > 0 && (uint64_t)vb.off[P] <= vb.cap;
        if (ok) {
            const uint64_t o = vb.off[p];
            if (tl == 0) {
                const float4 xm = vb.xm[o];
                d3 x1 = mkd(xm.x, xm.y, xm.z), x10 = x1;
                if (fr.temporalAmp && !fr.trivial) {
                    const float* d = vb.delta + (o * fr.nt + t) * 3;
                    const float* d0 = vb.delta + o * fr.nt * 3;
                    x1 = mkd(x1.x + (double)d[0], x1.y + (double)d[1], x1.z + (double)d[2]);
                    x10 = mkd(x10.x + (double)d0[0], x10.y + (double)d0[1], x10.z + (double)d0[2]);
                }
                sx[tq][0] = x1.x; sx[tq][1] = x1.y; sx[tq][2] = x1.z;
                sm[tq] = gbMaterial<DIMS>(m, x1);
                sx0[tq][0] = x10.x; sx0[tq][1] = x10.y; sx0[tq][2] = x10.z;
                sm0[tq] = gbMaterial<DIMS>(m, x10);
            }
            if (ls < L) av = a[((size_t)t * P + p) * L + ls];
            if (vs < V) bv = b[((size_t)t * P + p) * V + vs];
        } else if (tl == 0) {
            sm[tq] = 0;
            sm0[tq] = 0;
        }
        sa[tq][tl] = av;
        sb[tq][tl] = bv;
        __syncthreads();
        if (l < L && v < V)
            for (int q = 0; q < 16; ++q) {
                if (sm[q] == 0) continue;
                Cd f = gbFs<DIMS>(il.g, vw.g, tab, l, v, sm[q], mkd(sx[q][0], sx[q][1], sx[q][2]), il.g.waveK);
                if (corr) {   // S(0) conj S(t): fs at frame 0 times the conjugate at frame t
                    const bool same = sx0[q][0] == sx[q][0] && sx0[q][1] == sx[q][1] && sx0[q][2] == sx[q][2];
                    const Cd f0 = same ? f
                                       : (sm0[q] ? gbFs<DIMS>(il.g, vw.g, tab, l, v, sm0[q], mkd(sx0[q][0], sx0[q][1], sx0[q][2]),
                                                        il.g.waveK)
                                                 : cd(0, 0));
                    f = cd(f0.re * f.re + f0.im * f.im, f0.im * f.re - f0.re * f.im);
                }
                const C x = sa[q][tl], y = sb[q][tq];
                const double pr = (double)x.x * y.x - (double)x.y * y.y, pi = (double)x.x * y.y + (double)x.y * y.x;
                re += f.re * pr - f.im * pi;
                im += f.re * pi + f.im * pr;
            }
        __syncthreads();
    }
    if (l < L && v < V) {
        C c;
        c.x = (R)re;
        c.y = (R)im;
        part[((size_t)blockIdx.z) * L * V + (size_t)v * L + l] = c;
    }
}

constexpr uint32_t kFsFrames = 8;   // frames sharing one fs evaluation in k_single_gauss_uv (F)

// The same sum when every view beam has one direction d_v (the confocal geometry): then the inner
// convolution of fsGaussianBeam (its factor gob, the axis w, C and the exponent c) depends on the
// illumination, the lobe and the path only. Per path, the 16 x M (illumination, lobe) records of the
// block are formed once in LDS (one per thread) with the real parts and unit phases of r + C and
// r - C, r = c + log nu; per (l, v, lobe) there remain sm = sqrt((gob aT + bT).(gob aT + bT)), one
// sine-cosine pair of Im sm, two real exponentials exp(Re(r +- C) +- Re sm) (each exponent whole, as
// the reference forms it, so that no factor overflows alone) and the division by C sm. A term is dropped when the real part of its exponent is at most -10,
// as complex2timesSinh drops it; the rare zero cases of C or sm take the reference's special forms.
template <typename R, uint32_t F>
__global__ void __launch_bounds__(256) k_single_gauss_uv(MediumView m, ScatterView sc, SourceView il, SourceView vw,
                                                         FramesView fr, VertexBatch vb,
                                                         const typename SingleT<R>::C* __restrict__ a,
                                                         const typename SingleT<R>::C* __restrict__ b, uint32_t P,
                                                         uint32_t pathsPerSplit, uint32_t splits,
                                                         typename SingleT<R>::C* __restrict__ part, uint32_t fpb) {
    typedef typename SingleT<R>::C C;
    __shared__ double2 sGob[kGaussLobes][16], sC[kGaussLobes][16], sUp[kGaussLobes][16], sUm[kGaussLobes][16],
        sR[kGaussLobes][16];
    __shared__ double sReP[kGaussLobes][16], sReM[kGaussLobes][16], sNu[kGaussLobes];
    __shared__ double sx[3];
    __shared__ uint32_t sm1;
    const uint32_t L = il.n, V = vw.n;
    const uint32_t tl = threadIdx.x & 15, tq = threadIdx.x >> 4;
    // frames t0 .. t0 + nf - 1 of this block share fs (fpb > 1: the first vertex does not move with
    // the frame); fpb = 1 takes the first vertex of frame t0
    const uint32_t split = blockIdx.z % splits, t0 = (blockIdx.z / splits) * fpb;
    const uint32_t nf = fr.nt - t0 < fpb ? fr.nt - t0 : fpb;
    const uint32_t l = blockIdx.x * 16 + tl, v = blockIdx.y * 16 + tq;
    const uint32_t pBegin = split * pathsPerSplit, pEnd = min(P, pBegin + pathsPerSplit);
    const PhaseTables& tab = *sc.tab;
    const GaussBeams& ga = il.g;
    const GaussBeams& gb = vw.g;
    const double k = ga.waveK;
    const d3 db = gbDir(gb, 0);
    const double tc = ga.apNorm + gb.apNorm + 2.0 * 1.8378770664093454836;
    const bool poolOk = (uint64_t)vb.off[P] <= vb.cap;
    double re[F], im[F];
#pragma unroll
    for (uint32_t j = 0; j < F; ++j) re[j] = im[j] = 0;
    for (uint32_t p = pBegin; p < pEnd; ++p) {
        if (!poolOk || vb.k[p] == 0) continue;   // uniform over the block
        const uint64_t o = vb.off[p];
        if (threadIdx.x == 0) {
            const float4 xm = vb.xm[o];
            d3 x1 = mkd(xm.x, xm.y, xm.z);
            if (fpb == 1 && fr.temporalAmp && !fr.trivial) {
                const float* d = vb.delta + (o * fr.nt + t0) * 3;
                x1 = mkd(x1.x + (double)d[0], x1.y + (double)d[1], x1.z + (double)d[2]);
            }
            sx[0] = x1.x; sx[1] = x1.y; sx[2] = x1.z;
            sm1 = gbMaterial<3>(m, x1);
        }
        __syncthreads();
        const uint32_t mat = sm1;
        const d3 x1 = mkd(sx[0], sx[1], sx[2]);
        const uint32_t lo = mat ? tab.vmfOff[mat] : 0u, M = mat ? tab.vmfM[mat] : 0u;
        // (illumination tl, lobe tq) record
        if (tq < M && l < L) {
            const CV aT = CV{scale3(ga.gammaA, gbDir(ga, l)), scale3(k, sub3(x1, gbFocal(ga, l)))};
            const double gs0 = tab.vmfKappa[lo + tq];
            const double lognu = tc + tab.vmfC[lo + tq];
            Cd gob = cd(0, 0), Cc, c = cd(0, 0);
            if (fabs(gs0) < 0.000001) {
                Cc = csqrtd(sumSq(aT));
            } else {
                gob = gs0 * rcsqrtd(sumSq(cfmaR(gs0, db, aT)));
                const d3 w = realMultV(gob, aT);
                const double ra = 1.0 / sqrt(dot3(w, w));
                Cc = csqrtd(sumSq(cfmaR(gs0 * ra, w, aT)));
                c = (-ra) * (gob * dotCV(aT, w));
            }
            const Cd r = c + cd(lognu);
            sGob[tq][tl] = make_double2(gob.re, gob.im);
            sC[tq][tl] = make_double2(Cc.re, Cc.im);
            sR[tq][tl] = make_double2(r.re, r.im);
            const Cd rp = r + Cc, rm = r - Cc;
            double sn, cs;
            sincos(rp.im, &sn, &cs);
            sUp[tq][tl] = make_double2(cs, sn);
            sincos(rm.im, &sn, &cs);
            sUm[tq][tl] = make_double2(cs, sn);
            sReP[tq][tl] = rp.re;
            sReM[tq][tl] = rm.re;
            if (tl == 0) sNu[tq] = lognu;
        }
        __syncthreads();
        if (l < L && v < V && M > 0) {
            const CV aT = CV{scale3(ga.gammaA, gbDir(ga, l)), scale3(k, sub3(x1, gbFocal(ga, l)))};
            const CV bT = CV{scale3(gb.gammaA, gbDir(gb, v)), scale3(k, sub3(gbFocal(gb, v), x1))};
            Cd fs = cd(0, 0);
            for (uint32_t j = 0; j < M; ++j) {
                const double gs0 = tab.vmfKappa[lo + j];
                Cd sm;
                if (fabs(gs0) < 0.000001) sm = csqrtd(sumSq(bT));
                else sm = csqrtd(sumSq(cfmaC(cd(sGob[j][tl].x, sGob[j][tl].y), aT, bT)));
                const Cd Cc = cd(sC[j][tl].x, sC[j][tl].y);
                const bool smZero = fabs(sm.re) <= 0.0001 && fabs(sm.im) <= 0.0001;
                const bool cZero = fabs(Cc.re) <= 0.0001 && fabs(Cc.im) <= 0.0001;
                if (!smZero && !cZero) {
                    // exp(r + C + sm) - exp(r - C - sm), each dropped at real part <= -10
                    double sn, cs;
                    sincos(sm.im, &sn, &cs);
                    Cd num = cd(0, 0);
                    const double ep = sReP[j][tl] + sm.re, em = sReM[j][tl] - sm.re;
                    if (ep > -10.0) num = exp(ep) * (cd(sUp[j][tl].x, sUp[j][tl].y) * cd(cs, sn));
                    if (em > -10.0) num = num - exp(em) * (cd(sUm[j][tl].x, sUm[j][tl].y) * cd(cs, -sn));
                    fs = fs + cdiv(num, Cc * sm);
                } else {
                    const Cd r = cd(sR[j][tl].x, sR[j][tl].y);
                    if (smZero && cZero) fs = fs + 4.0 * cexpd(r);
                    else if (smZero) fs = fs + cdiv(2.0 * twoSinhd(Cc, r), Cc);
                    else fs = fs + cdiv(2.0 * twoSinhd(sm, cd(sNu[j])), sm);
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < F; ++j) {
                if (j >= nf) break;
                const size_t t = t0 + j;
                const C x = a[(t * P + p) * L + l], y = b[(t * P + p) * V + v];
                const double pr = (double)x.x * y.x - (double)x.y * y.y, pi = (double)x.x * y.y + (double)x.y * y.x;
                re[j] += fs.re * pr - fs.im * pi;
                im[j] += fs.re * pi + fs.im * pr;
            }
        }
        __syncthreads();
    }
    if (l < L && v < V) {
#pragma unroll
        for (uint32_t j = 0; j < F; ++j) {
            if (j >= nf) break;
            C c;
            c.x = (R)re[j];
            c.y = (R)im[j];
            part[((size_t)(t0 + j) * splits + split) * L * V + (size_t)v * L + l] = c;
        }
    }
}

hipError_t launchSingleGauss(int dims, int fp64, int uniformViewDir, int corr, const MediumView& m,
                             const ScatterView& sc, const SourceView& il, const SourceView& vw, const FramesView& fr,
                             const VertexBatch& vb, const void* a, const void* b, uint32_t P, double* U, Workspace* ws,
                             hipStream_t st) {
    // the shared-direction kernel forms the 3D field fs only
    if (corr || dims == 2) uniformViewDir = 0;
    const uint32_t L = il.n, V = vw.n, NT = fr.nt;
    uint32_t tiles = ((L + 15) / 16) * ((V + 15) / 16);
    uint32_t splits = chooseSplits(tiles, P, 64);
    uint32_t pps = ((P + splits - 1) / splits + 15) / 16 * 16;
    splits = (P + pps - 1) / pps;
    if (splits < 1) splits = 1;
    size_t elem = fp64 ? 16 : 8;
    size_t need = (size_t)splits * NT * L * V * elem;
    if (ws->bytes < need) {
        if (ws->ptr) {
            hipStreamSynchronize(st);
            hipFree(ws->ptr);
        }
        ws->ptr = nullptr;
        ws->bytes = 0;
        hipError_t e = hipMalloc(&ws->ptr, need);
        if (e != hipSuccess) return e;
        ws->bytes = need;
    }
    dim3 grid((L + 15) / 16, (V + 15) / 16, splits * NT);
    // frames per block of the shared-direction kernel: all frames share fs unless the amplitude
    // follows a moving first vertex
    const uint32_t fpb = (fr.temporalAmp && !fr.trivial) ? 1u : (NT < kFsFrames ? NT : kFsFrames);
    const dim3 gridUv((L + 15) / 16, (V + 15) / 16, splits * ((NT + fpb - 1) / fpb));
    const uint32_t LV = L * V;
    if (fp64) {
        if (uniformViewDir && fpb == 1)
            hipLaunchKernelGGL((k_single_gauss_uv<double, 1>), gridUv, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                               (const double2*)a, (const double2*)b, P, pps, splits, (double2*)ws->ptr, fpb);
        else if (uniformViewDir)
            hipLaunchKernelGGL((k_single_gauss_uv<double, kFsFrames>), gridUv, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                               (const double2*)a, (const double2*)b, P, pps, splits, (double2*)ws->ptr, fpb);
        else if (dims == 2)
            hipLaunchKernelGGL((k_single_gauss<2, double>), grid, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                               (const double2*)a, (const double2*)b, P, pps, splits, (double2*)ws->ptr, corr);
        else hipLaunchKernelGGL((k_single_gauss<3, double>), grid, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                                (const double2*)a, (const double2*)b, P, pps, splits, (double2*)ws->ptr, corr);
        for (uint32_t t = 0; t < NT; ++t)
            hipLaunchKernelGGL(k_reduce<double2>, dim3((LV + 255) / 256), dim3(256), 0, st,
                               (const double2*)ws->ptr + (size_t)t * splits * LV, splits, LV, (const float2*)nullptr,
                               (const double2*)nullptr, 0, 1.0, U + 2 * (size_t)t * LV);
    } else {
        if (uniformViewDir && fpb == 1)
            hipLaunchKernelGGL((k_single_gauss_uv<float, 1>), gridUv, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                               (const float2*)a, (const float2*)b, P, pps, splits, (float2*)ws->ptr, fpb);
        else if (uniformViewDir)
            hipLaunchKernelGGL((k_single_gauss_uv<float, kFsFrames>), gridUv, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                               (const float2*)a, (const float2*)b, P, pps, splits, (float2*)ws->ptr, fpb);
        else if (dims == 2)
            hipLaunchKernelGGL((k_single_gauss<2, float>), grid, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                               (const float2*)a, (const float2*)b, P, pps, splits, (float2*)ws->ptr, corr);
        else hipLaunchKernelGGL((k_single_gauss<3, float>), grid, dim3(256), 0, st, m, sc, il, vw, fr, vb,
                                (const float2*)a, (const float2*)b, P, pps, splits, (float2*)ws->ptr, corr);
        for (uint32_t t = 0; t < NT; ++t)
            hipLaunchKernelGGL(k_reduce<float2>, dim3((LV + 255) / 256), dim3(256), 0, st,
                               (const float2*)ws->ptr + (size_t)t * splits * LV, splits, LV, (const float2*)nullptr,
                               (const double2*)nullptr, 0, 1.0, U + 2 * (size_t)t * LV);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- correlation sampler single scattering
// S_p[l,v,t] = a_p[l,t] b_p[v,t] Fs_u1(l, v) conj(Fs_u2(l, v)) exp(-k^2 |D tau_t| |q|^2) for the
// TemporalCorrelationSampler (NextEventEstimation.cuh:1016-1136): the sources hold u1 (i < L, V)
// then u2 (L + i, V + i); both amplitudes see x1 at -tau/2 (copyPointsToBuffersThreePoints,
// TemporalSamplerCorrelation.cuh:289-313, quirk Q3); q = (i_l1 + i_l2 - o_v1 - o_v2) / 2 from the
// unshifted x1, i the incoming and o the outgoing (toward the view) directions
// (temporalCorrelationSourceContributionSingle, SourceInterface.cuh:217-277). Block = 16 x 16 (l, v)
// outputs; per chunk of 16 paths the directions, terms and decorrelation rates are staged in LDS.
__device__ inline void unitTo(const double* src, const double x[3], bool toward, int dims, double out[3]) {
    // n(x - src) (toward = false) or n(src - x) (toward = true)
    double q[3] = {x[0] - src[0], x[1] - src[1], dims == 3 ? x[2] - src[2] : 0.0};
    if (toward) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; }
    const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if (n > 0) { out[0] = q[0] / n; out[1] = q[1] / n; out[2] = q[2] / n; }
    else { out[0] = 0; out[1] = 0; out[2] = 1; }
}

template <int DIMS, typename R>
__global__ void __launch_bounds__(256) k_single_corrsampler(ScatterView sc, SourceView il, SourceView vw,
                                                            FramesView fr, VertexBatch vb,
                                                            const typename SingleT<R>::C* __restrict__ a,
                                                            const typename SingleT<R>::C* __restrict__ b, uint32_t P,
                                                            uint32_t pathsPerSplit, uint32_t splits, double kw2,
                                                            typename SingleT<R>::C* __restrict__ part) {
    typedef typename SingleT<R>::C C;
    __shared__ R su1[16][16][3], su2[16][16][3], sw1[16][16][3], sw2[16][16][3], sh[16][16][3], sk[16][16][3];
    __shared__ C sa[16][16], sb[16][16];
    __shared__ uint32_t sm[16];
    __shared__ R sr[16];
    const uint32_t L = il.n / 2, V = vw.n / 2;
    const uint32_t tl = threadIdx.x & 15, tq = threadIdx.x >> 4;
    const uint32_t split = blockIdx.z % splits, t = blockIdx.z / splits;
    const uint32_t l = blockIdx.x * 16 + tl, v = blockIdx.y * 16 + tq;
    const uint32_t pBegin = split * pathsPerSplit, pEnd = min(P, pBegin + pathsPerSplit);
    const bool iso = il.type == 1 && sc.allIsotropic;
    const double tau = fr.ticks[t], s = 0.5 * tau;
    R re = 0, im = 0;
    for (uint32_t pc = pBegin; pc < pEnd; pc += 16) {
        const uint32_t p = pc + tq;
        const uint32_t ls = blockIdx.x * 16 + tl, vs = blockIdx.y * 16 + tl;
        C av, bv;
        av.x = av.y = 0; bv.x = bv.y = 0;
        double x1[3] = {0, 0, 0}, xa[3] = {0, 0, 0};
        uint32_t m1 = 0;
        double rate = 0;
        const bool ok = p < pEnd && vb.k[p] > 0 && (uint64_t)vb.off[P] <= vb.cap;
        if (ok) {
            const float4 xm = vb.xm[vb.off[p]];
            m1 = __float_as_uint(xm.w);
            x1[0] = xm.x; x1[1] = xm.y; x1[2] = xm.z;
            for (int c = 0; c < 3; ++c) xa[c] = x1[c] - s * fr.Ud[m1][c];
            rate = kw2 * fabs(fr.Dd[m1] * tau);
            if (ls < L) av = a[((size_t)t * P + p) * L + ls];
            if (vs < V) bv = b[((size_t)t * P + p) * V + vs];
        }
        // incoming directions of u1 / u2 illumination ls at x1 shifted (amplitudes) and their mean
        // at x1 (decorrelation); outgoing directions toward u1 / u2 view vs likewise
        double u1[3] = {0, 0, 1}, u2[3] = {0, 0, 1}, h[3] = {0, 0, 0};
        double w1[3] = {0, 0, 1}, w2[3] = {0, 0, 1}, g[3] = {0, 0, 0};
        if (ls < L) {
            const double* d1 = il.pd + 3 * (size_t)ls;
            const double* d2 = il.pd + 3 * (size_t)(L + ls);
            double c1[3], c2[3];
            if (il.type == 2) {
                for (int c = 0; c < 3; ++c) { u1[c] = c1[c] = d1[c]; u2[c] = c2[c] = d2[c]; }
            } else {
                unitTo(d1, xa, false, DIMS, u1); unitTo(d2, xa, false, DIMS, u2);
                unitTo(d1, x1, false, DIMS, c1); unitTo(d2, x1, false, DIMS, c2);
            }
            for (int c = 0; c < 3; ++c) h[c] = 0.5 * (c1[c] + c2[c]);
        }
        if (vs < V) {
            const double* d1 = vw.pd + 3 * (size_t)vs;
            const double* d2 = vw.pd + 3 * (size_t)(V + vs);
            double c1[3], c2[3];
            if (vw.type == 2) {
                for (int c = 0; c < 3; ++c) { w1[c] = c1[c] = d1[c]; w2[c] = c2[c] = d2[c]; }
            } else {
                unitTo(d1, xa, true, DIMS, w1); unitTo(d2, xa, true, DIMS, w2);
                unitTo(d1, x1, true, DIMS, c1); unitTo(d2, x1, true, DIMS, c2);
            }
            for (int c = 0; c < 3; ++c) g[c] = 0.5 * (c1[c] + c2[c]);
        }
        if (DIMS == 2) { u1[2] = u2[2] = w1[2] = w2[2] = h[2] = g[2] = 0; }
        for (int c = 0; c < 3; ++c) {
            su1[tq][tl][c] = (R)u1[c]; su2[tq][tl][c] = (R)u2[c]; sh[tq][tl][c] = (R)h[c];
            sw1[tq][tl][c] = (R)w1[c]; sw2[tq][tl][c] = (R)w2[c]; sk[tq][tl][c] = (R)g[c];
        }
        sa[tq][tl] = av;
        sb[tq][tl] = bv;
        if (tl == 0) { sm[tq] = m1; sr[tq] = (R)rate; }
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            const C x = sa[q][tl], y = sb[q][tq];
            C f1, f2;
            if (iso) {
                f1.x = f2.x = DIMS == 3 ? (R)0.282094791773878143 : (R)0.398942280401432678;
                f1.y = f2.y = 0;
            } else {
                const R c1 = su1[q][tl][0] * sw1[q][tq][0] + su1[q][tl][1] * sw1[q][tq][1] + su1[q][tl][2] * sw1[q][tq][2];
                const R c2 = su2[q][tl][0] * sw2[q][tq][0] + su2[q][tl][1] * sw2[q][tq][1] + su2[q][tl][2] * sw2[q][tq][2];
                f1 = singleAmp<DIMS, R>(sc, sm[q], c1);
                f2 = singleAmp<DIMS, R>(sc, sm[q], c2);
            }
            const R dx = sh[q][tl][0] - sk[q][tq][0], dy = sh[q][tl][1] - sk[q][tq][1], dz = sh[q][tl][2] - sk[q][tq][2];
            const R w = SingleT<R>::ex(-sr[q] * (dx * dx + dy * dy + dz * dz));
            // F = f1 conj(f2) w; S = a b F
            const R fr_ = (f1.x * f2.x + f1.y * f2.y) * w, fi = (f1.y * f2.x - f1.x * f2.y) * w;
            const R pr = x.x * y.x - x.y * y.y, pi = x.x * y.y + x.y * y.x;
            re = fma(fr_, pr, fma(-fi, pi, re));
            im = fma(fr_, pi, fma(fi, pr, im));
        }
        __syncthreads();
    }
    if (l < L && v < V) {
        C c;
        c.x = re;
        c.y = im;
        part[((size_t)blockIdx.z) * L * V + (size_t)v * L + l] = c;
    }
}

// Far-field sources and one phase function for every material: Fs_u1 conj(Fs_u2) and |q|^2 depend
// on (l, v) only and the decorrelation rate on the first vertex's material and the lag only, so each
// thread keeps its (l, v) constants, the block a per-material table of exp(-rate_m |q|^2), and the
// path loop is one complex multiply-accumulate per path.
template <int DIMS, typename R>
__global__ void __launch_bounds__(256) k_single_corrsampler_far(ScatterView sc, SourceView il, SourceView vw,
                                                                FramesView fr, VertexBatch vb,
                                                                const typename SingleT<R>::C* __restrict__ a,
                                                                const typename SingleT<R>::C* __restrict__ b,
                                                                uint32_t P, uint32_t pathsPerSplit, uint32_t splits,
                                                                double kw2, typename SingleT<R>::C* __restrict__ part) {
    typedef typename SingleT<R>::C C;
    __shared__ C sa[16][16], sb[16][16];
    __shared__ uint32_t sm[16];
    __shared__ R ex[kMaxMaterials][256];
    const uint32_t L = il.n / 2, V = vw.n / 2;
    const uint32_t tl = threadIdx.x & 15, tq = threadIdx.x >> 4;
    const uint32_t split = blockIdx.z % splits, t = blockIdx.z / splits;
    const uint32_t l = blockIdx.x * 16 + tl, v = blockIdx.y * 16 + tq;
    const uint32_t pBegin = split * pathsPerSplit, pEnd = min(P, pBegin + pathsPerSplit);
    const double tau = fr.ticks[t];
    C f0;
    f0.x = f0.y = 0;
    {
        const uint32_t ll = l < L ? l : 0, vv = v < V ? v : 0;
        const double* i1 = il.pd + 3 * (size_t)ll;
        const double* i2 = il.pd + 3 * (size_t)(L + ll);
        const double* o1 = vw.pd + 3 * (size_t)vv;
        const double* o2 = vw.pd + 3 * (size_t)(V + vv);
        const C f1 = singleAmp<DIMS, R>(sc, 1, (R)(i1[0] * o1[0] + i1[1] * o1[1] + i1[2] * o1[2]));
        const C f2 = singleAmp<DIMS, R>(sc, 1, (R)(i2[0] * o2[0] + i2[1] * o2[1] + i2[2] * o2[2]));
        f0.x = f1.x * f2.x + f1.y * f2.y;
        f0.y = f1.y * f2.x - f1.x * f2.y;
        double q2 = 0;
        for (int c = 0; c < DIMS; ++c) {
            const double d = 0.5 * (i1[c] + i2[c] - o1[c] - o2[c]);
            q2 += d * d;
        }
        for (int m = 0; m < kMaxMaterials; ++m)
            ex[m][threadIdx.x] = SingleT<R>::ex((R)(-kw2 * fabs(fr.Dd[m] * tau) * q2));
    }
    R re = 0, im = 0;
    for (uint32_t pc = pBegin; pc < pEnd; pc += 16) {
        const uint32_t p = pc + tq;
        const uint32_t ls = blockIdx.x * 16 + tl, vs = blockIdx.y * 16 + tl;
        C av, bv;
        av.x = av.y = 0; bv.x = bv.y = 0;
        uint32_t m1 = 0;
        if (p < pEnd && vb.k[p] > 0 && (uint64_t)vb.off[P] <= vb.cap) {
            m1 = __float_as_uint(vb.xm[vb.off[p]].w);
            if (ls < L) av = a[((size_t)t * P + p) * L + ls];
            if (vs < V) bv = b[((size_t)t * P + p) * V + vs];
        }
        sa[tq][tl] = av;
        sb[tq][tl] = bv;
        if (tl == 0) sm[tq] = m1;
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            const C x = sa[q][tl], y = sb[q][tq];
            const R w = ex[sm[q]][threadIdx.x];
            const R pr = x.x * y.x - x.y * y.y, pi = x.x * y.y + x.y * y.x;
            re = fma(w, pr, re);
            im = fma(w, pi, im);
        }
        __syncthreads();
    }
    if (l < L && v < V) {
        C c;
        c.x = f0.x * re - f0.y * im;
        c.y = f0.x * im + f0.y * re;
        part[((size_t)blockIdx.z) * L * V + (size_t)v * L + l] = c;
    }
}

hipError_t launchSingleCorrelation(int dims, int fp64, const ScatterView& sc, const SourceView& il,
                                   const SourceView& vw, const FramesView& fr, const VertexBatch& vb, const void* a,
                                   const void* b, uint32_t P, double lambda, int uniformFar, double* U, Workspace* ws,
                                   hipStream_t st) {
    const uint32_t L = il.n / 2, V = vw.n / 2, NT = fr.nt;
    uint32_t tiles = ((L + 15) / 16) * ((V + 15) / 16);
    uint32_t splits = chooseSplits(tiles, P, 256);
    uint32_t pps = ((P + splits - 1) / splits + 15) / 16 * 16;
    splits = (P + pps - 1) / pps;
    if (splits < 1) splits = 1;
    size_t elem = fp64 ? 16 : 8;
    size_t need = (size_t)splits * NT * L * V * elem;
    if (ws->bytes < need) {
        if (ws->ptr) {
            hipStreamSynchronize(st);
            hipFree(ws->ptr);
        }
        ws->ptr = nullptr;
        ws->bytes = 0;
        hipError_t e = hipMalloc(&ws->ptr, need);
        if (e != hipSuccess) return e;
        ws->bytes = need;
    }
    const double kw2 = 39.478417604357434 / (lambda * lambda);
    dim3 grid((L + 15) / 16, (V + 15) / 16, splits * NT);
    const uint32_t LV = L * V;
#define SSTMC_SINGLE_CS(D, R, C2)                                                                          \
    do {                                                                                                   \
        if (uniformFar)                                                                                    \
            hipLaunchKernelGGL((k_single_corrsampler_far<D, R>), grid, dim3(256), 0, st, sc, il, vw, fr, vb,  \
                               (const C2*)a, (const C2*)b, P, pps, splits, kw2, (C2*)ws->ptr);             \
        else                                                                                               \
            hipLaunchKernelGGL((k_single_corrsampler<D, R>), grid, dim3(256), 0, st, sc, il, vw, fr, vb,      \
                               (const C2*)a, (const C2*)b, P, pps, splits, kw2, (C2*)ws->ptr);             \
    } while (0)
    if (fp64) {
        if (dims == 2) SSTMC_SINGLE_CS(2, double, double2);
        else SSTMC_SINGLE_CS(3, double, double2);
        for (uint32_t t = 0; t < NT; ++t)
            hipLaunchKernelGGL(k_reduce<double2>, dim3((LV + 255) / 256), dim3(256), 0, st,
                               (const double2*)ws->ptr + (size_t)t * splits * LV, splits, LV, (const float2*)nullptr,
                               (const double2*)nullptr, 0, 1.0, U + 2 * (size_t)t * LV);
    } else {
        if (dims == 2) SSTMC_SINGLE_CS(2, float, float2);
        else SSTMC_SINGLE_CS(3, float, float2);
        for (uint32_t t = 0; t < NT; ++t)
            hipLaunchKernelGGL(k_reduce<float2>, dim3((LV + 255) / 256), dim3(256), 0, st,
                               (const float2*)ws->ptr + (size_t)t * splits * LV, splits, LV, (const float2*)nullptr,
                               (const double2*)nullptr, 0, 1.0, U + 2 * (size_t)t * LV);
    }
    #undef SSTMC_SINGLE_CS
    return hipGetLastError();
}

__global__ void k_zero(double* U, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) U[i] = 0.0;
}

hipError_t launchZeroU(double* U, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, n);
    return hipGetLastError();
}

}  // namespace sstmc
