// Host-side check of the FP32 fast voxel walk (opticalDepthFast) on random grids and rays:
//   * every 4-byte walk-grid word it reads lies inside the padded grid (SSTMC_WALK_INDEX_CHECK),
//     for walk grids with each axis as the fastest one and 1-, 4- and 8-bit cells;
//   * tau agrees with the general FP32 walk (opticalDepth<3, float>, per-crossing parameters from
//     the fixed origin) to 1e-5 relative, and with the FP64 general walk to 2e-3 absolute;
//   * the tau cut: either both walks stop past the cut or both agree as above.
// Rays start inside the box, on cell faces and corners, and leave through all six faces.
// Also the far-field view kernel's forms: the record-based start with per-lane direction factors
// (bit-identical to the walk), and the per-vertex lower bound that lets it skip a vertex (runBound).
// Built and run by tests/test_walk_host.py (hipcc, host code only; AddressSanitizer when usable).
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <cstdint>
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

static size_t g_walk_words = 0;   // words of the padded walk grid under test
static long g_bad_index = 0;
#define SSTMC_WALK_INDEX_CHECK(m, cell)                                                         \
    do {                                                                                        \
        if ((size_t)((cell) >> 5) >= g_walk_words) ++g_bad_index;                               \
    } while (0)

#include "../../temporalspecklescode_amd/csrc/common.h"
#include "../../temporalspecklescode_amd/csrc/geometry.h"

using namespace sstmc;

// Packed walk grid with axis `fast` fastest, the other axes in increasing order: 4 x palette
// index of each cell (palette = material ids 1..nmat, `bits` bits a cell), one slowest-axis layer of
// padding (rounded up to a word) on both sides. Strides and the origin in bits.
static std::vector<uint32_t> walkGrid(const std::vector<uint8_t>& grid, const uint32_t n[3], int dims, int fast,
                                      uint32_t bits, uint32_t S[3], uint32_t* origin) {
    int order[3] = {fast, -1, -1}, k = 1;
    for (int a = 0; a < 3; ++a)
        if (a != fast) order[k++] = a;
    uint32_t mul = 1;
    for (int j = 0; j < 3; ++j) {
        S[order[j]] = order[j] < dims ? mul : 0u;
        if (order[j] < dims) mul *= n[order[j]];
    }
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    const size_t p = ((size_t)S[dims == 3 ? order[2] : order[1]] * bits + 31) & ~(size_t)31;
    std::vector<uint32_t> wg(((p + ncell * bits + p + 31) & ~(size_t)31) / 32, 0u);
    for (uint32_t i2 = 0; i2 < n[2]; ++i2)
        for (uint32_t i1 = 0; i1 < n[1]; ++i1)
            for (uint32_t i0 = 0; i0 < n[0]; ++i0) {
                const size_t bit = p + ((size_t)i0 * S[0] + (size_t)i1 * S[1] + (size_t)i2 * S[2]) * bits;
                wg[bit >> 5] |= 4u * (uint32_t)(grid[i0 + (size_t)n[0] * (i1 + (size_t)n[1] * i2)] - 1) << (bit & 31);
            }
    for (int a = 0; a < 3; ++a) S[a] *= bits;
    *origin = (uint32_t)p;
    return wg;
}

template <int DIMS>
static int run(int n, double h, int trials, unsigned seed, int fast, int nmat) {
    std::mt19937 rng(seed);
    std::vector<float> ef(n + 1);
    std::vector<double> ed(n + 1);
    for (int i = 0; i <= n; ++i) { ed[i] = -0.5 * n * h + i * h; ef[i] = (float)ed[i]; }
    const int nz = DIMS == 3 ? n : 1;
    const size_t nc = (size_t)n * n * nz;
    std::vector<uint8_t> grid(nc);
    for (auto& g : grid) g = (uint8_t)(1 + rng() % nmat);
    const uint32_t nn[3] = {(uint32_t)n, (uint32_t)n, (uint32_t)nz};
    const uint32_t bits = nmat <= 1 ? 1u : (nmat <= 4 ? 4u : 8u);
    uint32_t S[3], origin;
    std::vector<uint32_t> wg = walkGrid(grid, nn, DIMS, fast, bits, S, &origin);
    g_walk_words = wg.size();
    const double sig[kMaxMaterials] = {0, 0.21, 1.01, 3.02, 5.05, 0.5, 2.2, 0.05, 7.5, 1.7};
    MediumView m{};
    m.dims = DIMS;
    for (int a = 0; a < 3; ++a) {
        m.n[a] = a < DIMS ? n : 1;
        m.nf[a] = (float)m.n[a];
        m.ef[a] = ef.data(); m.ed[a] = ed.data();
        m.bminf[a] = ef[0]; m.bmaxf[a] = ef[n]; m.bmind[a] = ed[0]; m.bmaxd[a] = ed[n];
        m.uniformAxis[a] = 1; m.e0[a] = ef[0]; m.invH[a] = (float)(1.0 / h); m.h[a] = (float)h;
    }
    m.allUniform = 1; m.grid = grid.data(); m.walkGrid = wg.data(); m.walkWords = (uint32_t)wg.size();
    m.walkOrigin = origin; m.walkBits = bits;
    m.walkFast = 1;
    for (int a = 0; a < 3; ++a) m.walkStride[a] = S[a];
    for (int i = 0; i < kMaxMaterials; ++i) { m.sigtf[i] = (float)sig[i]; m.sigtd[i] = sig[i]; }
    for (int i = 0; i < nmat; ++i) m.walkSig[i] = (float)sig[i + 1];   // palette index i = material i + 1
    std::uniform_real_distribution<double> U(ed[0], ed[n]), N(-1, 1);
    long bad = 0, laneBad = 0;
    double worstRel = 0, worstAbs64 = 0;
    for (int t = 0; t < trials; ++t) {
        double p[3] = {U(rng), U(rng), DIMS == 3 ? U(rng) : 0.0};
        if (t % 4 == 1) p[t % 3 % DIMS] = ed[rng() % (n + 1)];           // on a face
        if (t % 9 == 2) { p[0] = ed[rng() % (n + 1)]; p[1] = ed[rng() % (n + 1)]; }   // on an edge
        double d[3] = {N(rng), N(rng), DIMS == 3 ? N(rng) : 0.0};
        if (t % 6 == 3) { d[0] = d[1] = d[2] = 0; d[t % DIMS] = (t & 8) ? 1 : -1; }   // axis aligned
        if (t % 10 == 5) { d[0] = (t & 16) ? 1 : -1; d[1] = (t & 32) ? 1 : -1; d[2] = DIMS == 3 ? ((t & 64) ? 1 : -1) : 0; }
        double nn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (double& x : d) x /= nn;
        const f3 pf = mk((float)p[0], (float)p[1], (float)p[2]), df = mk((float)d[0], (float)d[1], (float)d[2]);
        const d3 pd = mkd(pf.x, pf.y, pf.z), dd = mkd(df.x, df.y, df.z);
        const float cut = (t % 5 == 0) ? 7.5f : 3.0e38f;
        uint32_t s0 = 0, s1 = 0, s2 = 0;
        const float fast = opticalDepthFast<DIMS>(m, m.walkSig, pf, df, cut, &s0);
        if (inside<DIMS, float>(m, pf)) {
            // the far-field view kernel's form (start from a per-vertex record, per-lane direction
            // factors) must give the same bits and the same step count
            const WalkStart ws = walkStart<DIMS>(m, pf);
            const WalkLane wl = walkLane<DIMS>(m, df);
            uint32_t s3 = 0;
            const float lane = walkFromStartLane<DIMS>(m, m.walkSig, ws.dlo, ws.dhi, ws.b, ws.cell, wl, cut, &s3);
            if (std::memcmp(&lane, &fast, sizeof(float)) != 0 || s3 != s0) {
                if (laneBad < 5) printf("  ray %d: lane form %.9g (%u steps) vs %.9g (%u steps)\n", t, lane, s3, fast, s0);
                ++laneBad;
            }
        }
        const float slow = opticalDepth<DIMS, float>(m, m.sigtf, pf, df, cut, &s1);
        const double ref = opticalDepth<DIMS, double>(m, m.sigtd, pd, dd, 1e300, &s2);
        if (cut < 1e30f && (fast > cut || slow > cut)) {
            if (!(fast > cut && slow > cut)) ++bad;
            continue;
        }
        const double rel = std::fabs((double)fast - slow) / std::max(1.0, (double)std::fabs(slow));
        const double abs64 = std::fabs((double)fast - ref);
        worstRel = std::max(worstRel, rel);
        worstAbs64 = std::max(worstAbs64, abs64);
        if (rel > 1e-5 || abs64 > 2e-3) {
            if (bad < 5) printf("  ray %d: fast %.9g slow %.9g fp64 %.12g\n", t, fast, slow, ref);
            ++bad;
        }
    }
    printf("DIMS %d n %d h %g fastest axis %d, %d materials (%u-bit cells): %d rays, %ld mismatches, %ld "
           "lane-form mismatches, %ld out-of-range words, max rel vs FP32 general %.2e, max abs vs FP64 %.2e\n", DIMS,
           n, h, fast, nmat, bits, trials, bad, laneBad, g_bad_index, worstRel, worstAbs64);
    return bad != 0 || laneBad != 0 || g_bad_index != 0;
}

// The far-field view kernel's lower bounds (farLowerBound over a MinPyramid, BoundDirs of a cone of
// walk directions, and the same bound per block of its source order, blocks of farOrderHost, at the
// point and for boxes holding it, farBoxBound): for every point and every direction of the set (block), the FP32 fast walk's
// tau is at least the bound, and wherever the bound passes farSkip(cut) the walk under that cut ends past
// it, so skipping the point's walks changes no term. Grids of smooth blobs (box-blurred noise cut
// into nmat levels) so that the bound is often large; cones of up to 30 degrees about random
// axes, some containing the axis-aligned direction, points inside, on faces and on edges.
static int runBound(int n, double h, int trials, unsigned seed, int fast, int nmat) {
    std::mt19937 rng(seed);
    std::vector<double> ed(n + 1);
    std::vector<float> ef(n + 1);
    for (int i = 0; i <= n; ++i) { ed[i] = -0.5 * n * h + i * h; ef[i] = (float)ed[i]; }
    const size_t nc = (size_t)n * n * n;
    std::vector<double> f(nc), f2(nc);
    std::uniform_real_distribution<double> U01(0, 1);
    for (auto& x : f) x = U01(rng);
    for (int pass = 0; pass < 3; ++pass)
        for (int a = 0; a < 3; ++a) {
            const size_t st = a == 0 ? 1 : (a == 1 ? (size_t)n : (size_t)n * n);
            for (size_t i = 0; i < nc; ++i) {
                const int c = (int)((i / st) % n);
                double s = f[i], w = 1;
                if (c > 0) { s += f[i - st]; ++w; }
                if (c + 1 < n) { s += f[i + st]; ++w; }
                f2[i] = s / w;
            }
            f.swap(f2);
        }
    std::vector<double> sorted(f);
    std::sort(sorted.begin(), sorted.end());
    std::vector<uint8_t> grid(nc);
    for (size_t i = 0; i < nc; ++i) {
        int m = 1;
        for (int k = 1; k < nmat; ++k) m += f[i] > sorted[k * nc / nmat] ? 1 : 0;
        grid[i] = (uint8_t)m;
    }
    const uint32_t nn[3] = {(uint32_t)n, (uint32_t)n, (uint32_t)n};
    const uint32_t bits = nmat <= 1 ? 1u : (nmat <= 4 ? 4u : 8u);
    uint32_t S[3], origin;
    std::vector<uint32_t> wg = walkGrid(grid, nn, 3, fast, bits, S, &origin);
    g_walk_words = wg.size();
    const double sig[kMaxMaterials] = {0, 0.21, 1.01, 3.02, 5.05, 0.5, 2.2, 0.05, 7.5, 1.7};
    MediumView m{};
    m.dims = 3;
    for (int a = 0; a < 3; ++a) {
        m.n[a] = n;
        m.nf[a] = (float)n;
        m.ef[a] = ef.data(); m.ed[a] = ed.data();
        m.bminf[a] = ef[0]; m.bmaxf[a] = ef[n]; m.bmind[a] = ed[0]; m.bmaxd[a] = ed[n];
        m.uniformAxis[a] = 1; m.e0[a] = ef[0]; m.invH[a] = (float)(1.0 / h); m.h[a] = (float)h;
    }
    m.allUniform = 1; m.grid = grid.data(); m.walkGrid = wg.data(); m.walkWords = (uint32_t)wg.size();
    m.walkOrigin = origin; m.walkBits = bits; m.walkFast = 1;
    for (int a = 0; a < 3; ++a) m.walkStride[a] = S[a];
    for (int i = 0; i < kMaxMaterials; ++i) { m.sigtf[i] = (float)sig[i]; m.sigtd[i] = sig[i]; }
    for (int i = 0; i < nmat; ++i) m.walkSig[i] = (float)sig[i + 1];
    MinPyramid pyr[3];
    std::vector<float> pyrData[3];
    for (int a = 0; a < 3; ++a) {
        pyrData[a] = minPyramidHost(grid.data(), nn, m.sigtf, a, &pyr[a]);
        pyr[a].p = pyrData[a].data();
    }
    std::uniform_real_distribution<double> U(ed[0], ed[n]), N(-1, 1);
    long bad = 0, tested = 0, skipped = 0, valid = 0, blockTested = 0, blockSkipped = 0, cellSkipped = 0;
    double tight = 0;
    const float cuts[4] = {2.f, 8.f, 24.f, 210.f};
    for (int t = 0; t < trials; ++t) {
        // a cone of 48 walk directions about a random axis (every 7th set about a grid axis)
        double c[3] = {N(rng), N(rng), N(rng)};
        if (t % 7 == 3) { c[0] = c[1] = c[2] = 0; c[t % 3] = (t & 8) ? 1 : -1; }
        double cn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (double& x : c) x /= cn;
        const double half = (t % 5 == 0 ? 0.0 : U01(rng) * 0.52);
        std::vector<float> dirs;
        for (int k = 0; k < 48; ++k) {
            double d[3];
            for (;;) {
                double r[3] = {N(rng), N(rng), N(rng)};
                const double dot = r[0] * c[0] + r[1] * c[1] + r[2] * c[2];
                for (int a = 0; a < 3; ++a) r[a] -= dot * c[a];
                const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                if (rn < 1e-9) continue;
                const double ang = (k == 0 ? 0.0 : half * U01(rng));
                for (int a = 0; a < 3; ++a) d[a] = std::cos(ang) * c[a] + std::sin(ang) * r[a] / rn;
                break;
            }
            for (int a = 0; a < 3; ++a) dirs.push_back((float)d[a]);
        }
        int axis = -1;
        const BoundDirs bd = boundDirsHost(dirs.data(), 48, 1.0, &axis);
        if (!bd.valid) continue;
        ++valid;
        const std::vector<uint32_t> order = farOrderHost(dirs.data(), 48, 1.0, axis);
        for (int q = 0; q < 8; ++q) {
            double p[3] = {U(rng), U(rng), U(rng)};
            if (q == 1) p[t % 3] = ed[rng() % (n + 1)];
            if (q == 2) { p[0] = ed[rng() % (n + 1)]; p[1] = ed[rng() % (n + 1)]; }
            const f3 pf = mk((float)p[0], (float)p[1], (float)p[2]);
            const float lb = farLowerBound(m, pyr[axis], bd, pf, 3.0e38f);
            float minTau = 3.0e38f;
            for (int k = 0; k < 48; ++k) {
                const f3 df = mk(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]);
                uint32_t s0 = 0;
                const float tau = opticalDepthFast<3>(m, m.walkSig, pf, df, 3.0e38f, &s0);
                minTau = tau < minTau ? tau : minTau;
                for (float cut : cuts)
                    if (lb > farSkip(cut)) {
                        uint32_t s1 = 0;
                        const float tc = opticalDepthFast<3>(m, m.walkSig, pf, df, cut, &s1);
                        if (!(tc > cut)) {
                            if (bad < 5) printf("  set %d point %d dir %d: bound %.7g, cut %g, walk %.7g\n", t, q, k, lb, cut, tc);
                            ++bad;
                        }
                    }
            }
            if (!(minTau >= lb)) {
                if (bad < 5) printf("  set %d point %d: bound %.9g above the least walk %.9g\n", t, q, lb, minTau);
                ++bad;
            }
            ++tested;
            if (lb > farSkip(24.f)) ++skipped;
            if (minTau > 0) tight += lb / minTau;
            // the per-block bound of the kernel's source order (blocks of 16 directions about the
            // whole set's axis), at the point and for boxes holding it (the tabulated form: whole
            // cells and halves, coarse and fine pyramid queries): each at most every walk of the block
            uint32_t cc[3];
            float blo[2][3], bhi[2][3];
            farSubBoxOf(m, pf, 1, cc, blo[0], bhi[0]);
            farSubBoxOf(m, pf, 2, cc, blo[1], bhi[1]);
            for (int a = 0; a < 3; ++a)
                if (!(blo[1][a] <= p[a] + 1e-5 * h && p[a] <= bhi[1][a] + 1e-5 * h)) {
                    if (bad < 5) printf("  set %d point %d: half box [%g, %g] misses %g\n", t, q, blo[1][a], bhi[1][a], p[a]);
                    ++bad;
                }
            for (size_t b = 0; b < 48; b += 16) {
                const BoundDirs bb = boundDirsAxisHost(dirs.data(), order.data() + b, 16, 1.0, axis, bd.sgn);
                const float lbb = farLowerBound(m, pyr[axis], bb, pf, 3.0e38f);
                const float lbc[3] = {farBoxBound(m, pyr[axis], bb, cc, blo[0], bhi[0], 1),
                                      farBoxBound(m, pyr[axis], bb, cc, blo[1], bhi[1], 1),
                                      farBoxBound(m, pyr[axis], bb, cc, blo[1], bhi[1], 4)};
                ++blockTested;
                if (lbb > farSkip(24.f)) ++blockSkipped;
                if (lbc[2] > farSkip(24.f)) ++cellSkipped;
                for (size_t k = b; k < b + 16; ++k) {
                    const uint32_t i = order[k];
                    const f3 df = mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
                    uint32_t s0 = 0;
                    const float tau = opticalDepthFast<3>(m, m.walkSig, pf, df, 3.0e38f, &s0);
                    float lmax = lbb;
                    for (float x : lbc) lmax = x > lmax ? x : lmax;
                    if (!(tau >= lmax)) {
                        if (bad < 5) printf("  set %d point %d block %zu dir %u: bounds %.9g / %.9g %.9g %.9g above the walk %.9g\n", t, q, b, i, lbb, lbc[0], lbc[1], lbc[2], tau);
                        ++bad;
                    }
                    for (float cut : cuts)
                        if (lmax > farSkip(cut)) {
                            uint32_t s1 = 0;
                            if (!(opticalDepthFast<3>(m, m.walkSig, pf, df, cut, &s1) > cut)) ++bad;
                        }
                }
            }
        }
    }
    printf("bound: n %d h %g fastest axis %d, %d materials: %ld direction sets, %ld points, %ld violations, "
           "skippable at cut 24: %.3f (blocks of 16: %.3f, by half box %.3f), mean bound / least walk %.3f\n", n, h, fast,
           nmat, valid, tested, bad, tested ? (double)skipped / tested : 0.0,
           blockTested ? (double)blockSkipped / blockTested : 0.0, blockTested ? (double)cellSkipped / blockTested : 0.0,
           tested ? tight / tested : 0.0);
    return bad != 0 || g_bad_index != 0 || tested == 0;
}

int main() {
    int fail = 0;
    fail |= runBound(48, 1.0, 600, 21, 2, 4);
    fail |= runBound(33, 0.37, 600, 22, 0, 9);
    fail |= runBound(40, 2.5, 600, 23, 1, 1);
    const int mats[3] = {1, 4, 9};   // 1-, 4- and 8-bit cells
    for (int fast = 0; fast < 3; ++fast) {
        fail |= run<3>(32, 1.0, 40000, 1 + fast, fast, mats[fast]);
        fail |= run<3>(7, 0.37, 40000, 4 + fast, fast, mats[(fast + 1) % 3]);
        fail |= run<3>(9, 0.1, 20000, 7 + fast, fast, mats[(fast + 2) % 3]);
    }
    for (int fast = 0; fast < 2; ++fast) {
        fail |= run<2>(64, 1.0, 40000, 10 + fast, fast, mats[fast]);
        fail |= run<2>(5, 0.25, 40000, 12 + fast, fast, mats[fast + 1]);
        fail |= run<2>(33, 0.5, 20000, 14 + fast, fast, mats[2 - fast]);
    }
    printf(fail ? "FAIL\n" : "OK\n");
    return fail;
}
