// sstmc_cli -- command-line host entry over the C ABI (include/sstmc.h): the MEX call
// `[u, iter] = sstmc{2,3}D(config)` (MatlabInterface.cu:1565-1825) with the MATLAB config struct
// given as JSON. Field names, defaults and error messages follow the MEX argument parsing
// (MatlabInterface.cu:274-1563) and the Python host mirror (temporalspecklescode_amd/config.py).
//
//   sstmc_cli [--dims 2|3] [--check] [--out u.npy] config.json
//
// `u` is written as a complex128 .npy array of shape (Nl, Nv, Nt, renderingsNum) in Fortran
// order (l fastest, as MATLAB holds it); `iter` and the sizes go to stdout as one JSON line.
// --check validates the config (sstmc_validate) without touching the device.
// Exit status: 0, or the sstmc error code (config errors found while reading the JSON: 7).
//
// JSON layout (MATLAB struct field names):
//   simulation: batchSize, fullIteration, renderingsNum, deviceNumber, seed, lambda, graphs
//   medium: type "Heterogeneous", xAxis, yAxis[, zAxis], materialGrid (flat, x fastest, or nested
//           [x][y][z]) or materialGridFile (raw bytes, x fastest; materialGridType "uint8"|"uint32"),
//           sigs, siga
//   scattering: type ("Isotropic" | "Tabular" | "HenyeyGreenstein" | "HG" | "VonMisesFisher" |
//               "vMF", or a list), amplitudeFunction (HG: g; Tabular: {"re": [...], "im": [...]},
//               [[re, im], ...] or a real list; vMF: {"mixtureMu", "mixtureC", "mixtureAlpha"};
//               per material)
//   illumination / view: type "FarField" (direction{x,y[,z]}) | "PointSource" (location{...}),
//                        isAreaSource | "GaussianBeam" (focalPoint{x,y,z}, focalDirection{x,y,z},
//                        aperture; 3D)
//   sampler: type "TemporalSampler" | "TemporalCorrelationSampler", t, D, U{x,y,z},
//            isTemporalAmplitude; the correlation sampler also takes illumination2 and view2
//   tracer: type "nee", isCBS, output "field" | "correlation"
//   precision (this implementation): terms "fp32"|"fp64", contraction "bf16x3"|"fp32"|"fp64",
//   tauCut (optical depth past which an NEE term is dropped; 0, the default, is exact)
//   dims (optional; else --dims; else 3 when medium.zAxis is present)
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sstmc.h"
#include "json.h"

using sstmc_json::Value;

namespace {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Owns the arrays an sstmc_config points into.
struct HostConfig {
    sstmc_config c{};
    std::vector<std::vector<double>> d;
    std::vector<std::vector<uint32_t>> u;
    std::vector<std::vector<int32_t>> i;
    const double* keep(std::vector<double> v) { d.push_back(std::move(v)); return d.back().data(); }
};

double num(const Value& v, const char* what) {
    if (v.kind == Value::Number) return v.num;
    if (v.kind == Value::Bool) return v.b ? 1.0 : 0.0;
    if (v.kind == Value::Array && v.arr.size() == 1) return num(v.arr[0], what);
    throw ConfigError(std::string(what) + " must be a number");
}
double numOr(const Value& o, const char* k, double dflt) { return o.has(k) ? num(o.at(k), k) : dflt; }
bool flag(const Value& o, const char* k, bool dflt) { return o.has(k) ? num(o.at(k), k) != 0 : dflt; }
std::string str(const Value& o, const char* k, const std::string& dflt) {
    if (!o.has(k)) return dflt;
    const Value& v = o.at(k);
    if (v.kind != Value::String) throw ConfigError(std::string(k) + " must be a string");
    return v.str;
}
std::vector<double> nums(const Value& v) {
    try {
        return v.numbers();
    } catch (const std::runtime_error& e) {
        throw ConfigError(e.what());
    }
}
// per-material vector: a scalar is broadcast
std::vector<double> perMaterial(const Value& o, const char* k, size_t n, double dflt) {
    if (!o.has(k)) return std::vector<double>(n, dflt);
    std::vector<double> v = nums(o.at(k));
    if (v.size() == 1) return std::vector<double>(n, v[0]);
    if (v.size() != n) throw ConfigError(std::string(k) + " must hold " + std::to_string(n) + " elements");
    return v;
}

void source(HostConfig& h, sstmc_source& dst, const Value& src, uint32_t dims, const char* name) {
    const std::string ty = str(src, "type", "");
    const Value* loc;
    if (ty == "PointSource") {
        dst.type = SSTMC_SOURCE_POINT;
        loc = &src.at("location");
    } else if (ty == "FarField") {
        dst.type = SSTMC_SOURCE_FARFIELD;
        loc = &src.at("direction");
    } else if (ty == "GaussianBeam") {
        // MatlabInterface.cu:1153-1190: focalPoint, focalDirection (one per focal point), aperture
        if (dims != 3) throw ConfigError("GaussianBeam sources are supported in 3D");
        dst.type = SSTMC_SOURCE_GAUSSIAN_BEAM;
        loc = &src.at("focalPoint");
        const Value& fd = src.at("focalDirection");
        const size_t n = nums(loc->at("x")).size();
        const char* keys[3] = {"x", "y", "z"};
        const double** dirs[3] = {&dst.dirX, &dst.dirY, &dst.dirZ};
        for (int a = 0; a < 3; ++a) {
            std::vector<double> v = nums(fd.at(keys[a]));
            if (v.size() != n)
                throw ConfigError(std::string(name) + " focalDirection must hold one direction per focal point");
            *dirs[a] = h.keep(v);
        }
        dst.aperture = num(src.at("aperture"), "aperture");
        if (!(dst.aperture > 0)) throw ConfigError(std::string(name) + " aperture must be positive");
    } else {
        throw ConfigError(std::string("Unsupported ") + name + " source type " + ty +
                          ". Available types are PointSource, FarField and GaussianBeam");
    }
    std::vector<double> xs = nums(loc->at("x")), ys = nums(loc->at("y"));
    if (xs.size() != ys.size()) throw ConfigError(std::string(name) + " x and y must have the same size");
    dst.num = (uint32_t)xs.size();
    dst.x = h.keep(xs);
    dst.y = h.keep(ys);
    if (dims == 3) {
        std::vector<double> zs = nums(loc->at("z"));
        if (zs.size() != xs.size()) throw ConfigError(std::string(name) + " x and z must have the same size");
        dst.z = h.keep(zs);
    }
    dst.isAreaSource = flag(src, "isAreaSource", false) ? 1 : 0;
}

std::vector<uint32_t> readGridFile(const std::string& path, const std::string& type, size_t ncell) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ConfigError("cannot open materialGridFile " + path);
    std::vector<uint32_t> g(ncell);
    if (type == "uint32") {
        f.read(reinterpret_cast<char*>(g.data()), (std::streamsize)(ncell * 4));
    } else if (type == "uint8") {
        std::vector<uint8_t> b(ncell);
        f.read(reinterpret_cast<char*>(b.data()), (std::streamsize)ncell);
        for (size_t k = 0; k < ncell; ++k) g[k] = b[k];
    } else {
        throw ConfigError("materialGridType must be uint8 or uint32");
    }
    if (!f) throw ConfigError("materialGridFile " + path + " holds fewer than " + std::to_string(ncell) + " cells");
    return g;
}

void fill(HostConfig& h, const Value& cfg, uint32_t dims) {
    sstmc_config& s = h.c;
    s.dims = dims;
    static const Value empty = [] { Value v; v.kind = Value::Object; return v; }();
    const Value& sim = cfg.has("simulation") ? cfg.at("simulation") : empty;
    const double bs = numOr(sim, "batchSize", 1024);
    if (bs <= 0) throw ConfigError("Batch size must be a positive integer.");
    s.simulation.batchSize = (uint32_t)bs;
    const double fi = numOr(sim, "fullIteration", 1);
    if (fi <= 0) throw ConfigError("Full iteraion type must be non-complex positive double.");
    s.simulation.fullIteration = (uint32_t)fi;
    const double rn = numOr(sim, "renderingsNum", 1);
    if (rn <= 0) throw ConfigError("Renderings num type must be non-complex positive double.");
    s.simulation.renderingsNum = (uint32_t)rn;
    s.simulation.deviceNumber = (uint32_t)numOr(sim, "deviceNumber", 0);
    s.simulation.seed = (uint64_t)std::fabs(numOr(sim, "seed", 0));
    if (sim.has("lambda") && nums(sim.at("lambda")).size() != 1)
        throw ConfigError("only one wavelength is supported (Nw = 1)");
    s.simulation.lambda = numOr(sim, "lambda", 1.0);
    s.simulation.graphs = numOr(sim, "graphs", 0) != 0 ? 1u : 0u;

    // scattering (MatlabInterface.cu:712-888)
    const Value& sc = cfg.at("scattering");
    std::vector<std::string> types;
    const Value& tv = sc.at("type");
    if (tv.kind == Value::String) types.push_back(tv.str);
    else for (const Value& e : tv.arr) types.push_back(e.str);
    // amplitudeFunction per material: one entry per material, or one value for all
    std::vector<const Value*> ampv(types.size(), nullptr);
    if (sc.has("amplitudeFunction")) {
        const Value& av = sc.at("amplitudeFunction");
        const bool perMaterial = av.kind == Value::Array && av.arr.size() == types.size() &&
                                 (types.size() > 1 || av.arr[0].kind != Value::Number);
        for (size_t k = 0; k < types.size(); ++k) ampv[k] = perMaterial ? &av.arr[k] : &av;
    }
    std::vector<int32_t> tcode;
    std::vector<double> g, tab, vmu, vc, valpha;
    std::vector<uint32_t> tsize, vnum;
    for (size_t k = 0; k < types.size(); ++k) {
        const std::string mk = std::to_string(k);
        vnum.push_back(0);
        if (types[k] == "Isotropic") {
            tcode.push_back(SSTMC_SCATTERING_ISOTROPIC);
            g.push_back(0.0);
            tsize.push_back(0);
        } else if (types[k] == "HenyeyGreenstein" || types[k] == "HG") {
            tcode.push_back(SSTMC_SCATTERING_HG);
            g.push_back(ampv[k] ? nums(*ampv[k])[0] : 0.0);
            tsize.push_back(0);
        } else if (types[k] == "VonMisesFisher" || types[k] == "vMF") {
            // mixture of von Mises-Fisher lobes (MatlabInterface.cu:827-866)
            std::vector<double> mu, cc, al;
            if (ampv[k] && ampv[k]->kind == Value::Object) {
                if (ampv[k]->has("mixtureMu")) mu = nums(ampv[k]->at("mixtureMu"));
                if (ampv[k]->has("mixtureC")) cc = nums(ampv[k]->at("mixtureC"));
                if (ampv[k]->has("mixtureAlpha")) al = nums(ampv[k]->at("mixtureAlpha"));
            }
            if (mu.empty() || cc.size() != mu.size() || al.size() != mu.size())
                throw ConfigError("vMF amplitude function of material " + mk +
                                  " needs mixtureMu, mixtureC and mixtureAlpha of equal, non-zero length");
            for (size_t i = 0; i < mu.size(); ++i)
                if (!(al[i] > 0) || !std::isfinite(mu[i]) || !std::isfinite(cc[i]))
                    throw ConfigError("vMF amplitude function of material " + mk +
                                      ": mixtureAlpha must be positive and mixtureMu, mixtureC finite");
            tcode.push_back(SSTMC_SCATTERING_VMF);
            g.push_back(0.0);
            tsize.push_back(0);
            vnum.back() = (uint32_t)mu.size();
            vmu.insert(vmu.end(), mu.begin(), mu.end());
            vc.insert(vc.end(), cc.begin(), cc.end());
            valpha.insert(valpha.end(), al.begin(), al.end());
        } else if (types[k] == "Tabular") {
            // complex amplitude samples over the scattering angle (Scattering.cuh:272-429)
            std::vector<double> re, im;
            if (ampv[k] && ampv[k]->kind == Value::Object) {
                re = nums(ampv[k]->at("re"));
                im = ampv[k]->has("im") ? nums(ampv[k]->at("im")) : std::vector<double>(re.size(), 0.0);
            } else if (ampv[k] && ampv[k]->kind == Value::Array && !ampv[k]->arr.empty() &&
                       ampv[k]->arr[0].kind == Value::Array) {
                for (const Value& e : ampv[k]->arr) {
                    std::vector<double> c = nums(e);
                    re.push_back(c.at(0));
                    im.push_back(c.size() > 1 ? c[1] : 0.0);
                }
            } else if (ampv[k]) {
                re = nums(*ampv[k]);
                im.assign(re.size(), 0.0);
            }
            bool finite = re.size() == im.size(), nonzero = false;
            for (size_t i = 0; finite && i < re.size(); ++i) {
                finite = std::isfinite(re[i]) && std::isfinite(im[i]);
                nonzero = nonzero || re[i] != 0.0 || im[i] != 0.0;
            }
            if (re.size() < 2 || !finite || !nonzero)
                throw ConfigError("Tabular amplitude function of material " + mk +
                                  " must hold at least two finite samples, not all zero");
            tcode.push_back(SSTMC_SCATTERING_TABULAR);
            g.push_back(0.0);
            tsize.push_back((uint32_t)re.size());
            for (size_t i = 0; i < re.size(); ++i) {
                tab.push_back(re[i]);
                tab.push_back(im[i]);
            }
        } else {
            throw ConfigError("Unsupported scattering type in material " + mk + ". Type entered: " + types[k] +
                              ". Available scattering types are Isotropic, Tabular, HenyeyGreenstein (HG) and "
                              "VonMisesFisher (vMF)");
        }
    }
    const uint32_t nsc = (uint32_t)tcode.size();
    const uint32_t nmat = nsc == 1 ? 2 : nsc;
    h.i.push_back(tcode);
    s.scattering.materialNum = nsc;
    s.scattering.type = h.i.back().data();
    s.scattering.g = h.keep(g);
    if (!tab.empty()) {
        h.u.push_back(tsize);
        s.scattering.tabularSize = h.u.back().data();
        s.scattering.tabular = h.keep(tab);
    }
    if (!vmu.empty()) {
        h.u.push_back(vnum);
        s.scattering.vmfMixtures = h.u.back().data();
        s.scattering.vmfMu = h.keep(vmu);
        s.scattering.vmfC = h.keep(vc);
        s.scattering.vmfAlpha = h.keep(valpha);
    }

    // medium (MatlabInterface.cu:890-1060)
    const Value& md = cfg.at("medium");
    if (str(md, "type", "") != "Heterogeneous")
        throw ConfigError("Unsupported medium type. Available medium type is Heterogeneous");
    s.medium.type = SSTMC_MEDIUM_HETEROGENEOUS;
    const char* axn[3] = {"xAxis", "yAxis", "zAxis"};
    size_t ncell = 1;
    for (uint32_t a = 0; a < dims; ++a) {
        std::vector<double> e = nums(md.at(axn[a]));
        if (e.size() < 2) throw ConfigError("medium axes must hold at least two increasing values");
        for (size_t k = 1; k < e.size(); ++k)
            if (!(e[k] > e[k - 1])) throw ConfigError("medium axes must hold at least two increasing values");
        ncell *= e.size() - 1;
        const double* p = h.keep(e);
        if (a == 0) { s.medium.xAxis = p; s.medium.xAxisSize = (uint32_t)e.size(); }
        if (a == 1) { s.medium.yAxis = p; s.medium.yAxisSize = (uint32_t)e.size(); }
        if (a == 2) { s.medium.zAxis = p; s.medium.zAxisSize = (uint32_t)e.size(); }
    }
    std::vector<uint32_t> grid;
    if (md.has("materialGridFile")) {
        grid = readGridFile(str(md, "materialGridFile", ""), str(md, "materialGridType", "uint8"), ncell);
    } else {
        std::vector<double> gv = nums(md.at("materialGrid"));
        if (gv.size() != ncell) throw ConfigError("materialGrid must hold " + std::to_string(ncell) + " elements");
        grid.resize(ncell);
        for (size_t k = 0; k < ncell; ++k) grid[k] = (uint32_t)gv[k];
    }
    uint32_t gmax = 0;
    for (uint32_t m : grid) {
        if (m == 0)
            throw ConfigError("Material 0 is only allowed outside the grid. Allowed material numbers are from 1 to " +
                              std::to_string(nmat - 1) + ".");
        gmax = m > gmax ? m : gmax;
    }
    if (gmax >= nmat)
        throw ConfigError("Material " + std::to_string(gmax) + " is not defined. Allowed material numbers are from 1 to " +
                          std::to_string(nmat - 1) + ".");
    h.u.push_back(grid);
    s.medium.materialGrid = h.u.back().data();
    std::vector<double> sigs = nums(md.at("sigs"));
    if (sigs.size() != nmat) throw ConfigError("sigs must hold " + std::to_string(nmat) + " elements");
    s.medium.sigs = h.keep(sigs);
    s.medium.materialNum = nmat;
    if (md.has("siga") && md.at("siga").kind != Value::Null) {
        std::vector<double> siga = nums(md.at("siga"));
        if (siga.size() != nmat) throw ConfigError("siga must hold " + std::to_string(nmat) + " elements");
        s.medium.siga = h.keep(siga);
    }

    // sources (MatlabInterface.cu:1062-1215)
    source(h, s.illumination, cfg.at("illumination"), dims, "illumination");
    source(h, s.view, cfg.at("view"), dims, "view");

    // sampler (MatlabInterface.cu:1217-1368)
    const Value& sp = cfg.at("sampler");
    const std::string stype = str(sp, "type", "TemporalSampler");
    if (stype == "TemporalSampler") {
        s.sampler.type = SSTMC_SAMPLER_TEMPORAL;
    } else if (stype == "TemporalCorrelationSampler") {
        s.sampler.type = SSTMC_SAMPLER_TEMPORAL_CORRELATION;
        // MatlabInterface.cu:1313-1331
        if (!cfg.has("illumination2") || !cfg.has("view2"))
            throw ConfigError("For correlation sampler, source.illuminationsHandle2 and source.viewsHandle2 must be "
                              "entered");
        source(h, s.illumination2, cfg.at("illumination2"), dims, "illumination2");
        source(h, s.view2, cfg.at("view2"), dims, "view2");
        if (s.illumination2.num != s.illumination.num)
            throw ConfigError("Illumination 1 in size of " + std::to_string(s.illumination.num) +
                              ", Illumination 2 in size of " + std::to_string(s.illumination2.num) +
                              ". Both illuminations must be at the same size");
        if (s.view2.num != s.view.num)
            throw ConfigError("View 1 in size of " + std::to_string(s.view.num) + ", View 2 in size of " +
                              std::to_string(s.view2.num) + ". Both views must be at the same size");
    } else {
        throw ConfigError("Unsupported sapmler type. Available sampler types are TemporalSampler and "
                          "TemporalCorrelationSampler");
    }
    std::vector<double> t = sp.has("t") ? nums(sp.at("t")) : std::vector<double>{0.0};
    s.sampler.tNum = (uint32_t)t.size();
    s.sampler.t = h.keep(t);
    std::vector<double> D = sp.has("D") ? nums(sp.at("D")) : std::vector<double>(nmat, 0.0);
    if (D.size() != nmat) throw ConfigError("D must hold " + std::to_string(nmat) + " elements");
    for (double x : D)
        if (x < 0) throw ConfigError("D values in temporal sampler must be non-negative");
    s.sampler.D = h.keep(D);
    Value U;
    U.kind = Value::Object;
    if (sp.has("U")) U = sp.at("U");
    s.sampler.Ux = h.keep(perMaterial(U, "x", nmat, 0.0));
    s.sampler.Uy = h.keep(perMaterial(U, "y", nmat, 0.0));
    s.sampler.Uz = h.keep(perMaterial(U, "z", nmat, 0.0));
    s.sampler.isTemporalAmplitude = flag(sp, "isTemporalAmplitude", true) ? 1 : 0;

    static const Value nee = [] { Value v; v.kind = Value::Object; return v; }();
    const Value& tr = cfg.has("tracer") ? cfg.at("tracer") : nee;
    if (str(tr, "type", "nee") != "nee") throw ConfigError("Unsupported tracer type. Available tracer type is nee");
    s.tracer.type = SSTMC_TRACER_NEE;
    s.tracer.isCBS = flag(tr, "isCBS", false) ? 1 : 0;
    const std::string out = str(tr, "output", "field");
    if (out == "field") s.tracer.output = SSTMC_OUTPUT_FIELD;
    else if (out == "correlation") s.tracer.output = SSTMC_OUTPUT_CORRELATION;
    else throw ConfigError("tracer.output must be field or correlation");

    const Value& pr = cfg.has("precision") ? cfg.at("precision") : nee;
    const std::string terms = str(pr, "terms", "fp32"), con = str(pr, "contraction", "bf16x3");
    if (terms == "fp32") s.precision.terms = SSTMC_TERMS_FP32;
    else if (terms == "fp64") s.precision.terms = SSTMC_TERMS_FP64;
    else throw ConfigError("precision.terms must be fp32 or fp64");
    if (con == "bf16x3") s.precision.contraction = SSTMC_CONTRACT_BF16X3;
    else if (con == "fp32") s.precision.contraction = SSTMC_CONTRACT_FP32;
    else if (con == "fp64") s.precision.contraction = SSTMC_CONTRACT_FP64;
    else throw ConfigError("precision.contraction must be bf16x3, fp32 or fp64");
    s.precision.tauCut = numOr(pr, "tauCut", 0.0);
}

// .npy v1.0, complex128, Fortran order
bool writeNpy(const std::string& path, const double* data, const size_t shape[4]) {
    std::string hdr = "{'descr': '<c16', 'fortran_order': True, 'shape': (" + std::to_string(shape[0]) + ", " +
                      std::to_string(shape[1]) + ", " + std::to_string(shape[2]) + ", " + std::to_string(shape[3]) +
                      "), }";
    size_t total = 10 + hdr.size() + 1;
    hdr += std::string((64 - total % 64) % 64, ' ') + "\n";
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write(reinterpret_cast<const char*>(magic), 8);
    const uint16_t len = (uint16_t)hdr.size();
    const unsigned char lenb[2] = {(unsigned char)(len & 0xff), (unsigned char)(len >> 8)};
    f.write(reinterpret_cast<const char*>(lenb), 2);
    f.write(hdr.data(), (std::streamsize)hdr.size());
    const size_t n = 2 * shape[0] * shape[1] * shape[2] * shape[3];
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(double)));
    return (bool)f;
}

int usage() {
    std::fprintf(stderr, "usage: sstmc_cli [--dims 2|3] [--check] [--out u.npy] config.json\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    std::string cfgPath, outPath;
    int dimsArg = 0;
    bool check = false;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--dims") && a + 1 < argc) dimsArg = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--out") && a + 1 < argc) outPath = argv[++a];
        else if (!std::strcmp(argv[a], "--check")) check = true;
        else if (argv[a][0] == '-') return usage();
        else cfgPath = argv[a];
    }
    if (cfgPath.empty()) return usage();
    HostConfig h;
    try {
        std::ifstream f(cfgPath);
        if (!f) throw ConfigError("cannot open " + cfgPath);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        Value cfg = sstmc_json::Parser(text).parse();
        uint32_t dims = (uint32_t)dimsArg;
        if (!dims && cfg.has("dims")) dims = (uint32_t)num(cfg.at("dims"), "dims");
        if (!dims) dims = cfg.has("medium") && cfg.at("medium").has("zAxis") ? 3 : 2;
        if (dims != 2 && dims != 3) throw ConfigError("dims must be 2 or 3");
        fill(h, cfg, dims);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "sstmc_cli: %s\n", e.what());
        return SSTMC_INVALID_CONFIG;
    }
    const sstmc_config& c = h.c;
    const size_t shape[4] = {c.illumination.num, c.view.num, c.sampler.tNum, c.simulation.renderingsNum};
    int rc = sstmc_validate(&c);
    if (rc != SSTMC_NO_ERROR) {
        std::fprintf(stderr, "sstmc_cli: %s\n", sstmc_error_string(rc));
        return rc;
    }
    if (check) {
        std::printf("{\"valid\": true, \"dims\": %u, \"shape\": [%zu, %zu, %zu, %zu]}\n", c.dims, shape[0], shape[1],
                    shape[2], shape[3]);
        return 0;
    }
    std::vector<double> u(2 * shape[0] * shape[1] * shape[2] * shape[3]);
    std::vector<double> iter(shape[3]);
    const auto t0 = std::chrono::steady_clock::now();
    rc = sstmc_render(&c, u.data(), iter.data());
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != SSTMC_NO_ERROR) {
        std::fprintf(stderr, "sstmc_cli: %s\n", sstmc_error_string(rc));
        return rc;
    }
    if (!outPath.empty() && !writeNpy(outPath, u.data(), shape)) {
        std::fprintf(stderr, "sstmc_cli: cannot write %s\n", outPath.c_str());
        return SSTMC_MISC_ERROR;
    }
    std::printf("{\"dims\": %u, \"shape\": [%zu, %zu, %zu, %zu], \"seconds\": %.6f, \"iter\": [", c.dims, shape[0],
                shape[1], shape[2], shape[3], secs);
    for (size_t r = 0; r < shape[3]; ++r) std::printf(r ? ", %.17g" : "%.17g", iter[r]);
    std::printf("]}\n");
    return 0;
}
