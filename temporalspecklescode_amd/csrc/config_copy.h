// config_copy.h -- deep copy of an sstmc_config (host code only, no HIP).
//
// The serving loop builds a request's tracer on its worker thread, after sstmc_submit has
// returned and the caller may have freed or reused every array its config points into. A
// ConfigCopy owns a copy of each of those arrays and a config whose pointers lead into them:
// the medium's axes, grid and cross sections; every scattering array (types, HG g, tabular sizes
// and samples, vMF mixture counts and lobes); the x / y / z and GaussianBeam direction arrays of
// illumination, view, illumination2 and view2; the sampler's ticks, D and velocities. Sizes come
// from the config itself, so it must have passed sstmc_validate.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/sstmc.h"

namespace sstmc {

struct ConfigCopy {
    sstmc_config c{};

    // withGrid = false leaves the material grid out (c.medium.materialGrid = NULL): a request whose
    // medium the server already holds does not need it
    explicit ConfigCopy(const sstmc_config& in, bool withGrid = true) : c(in) {
        const int dims = (int)in.dims;
        const sstmc_medium& m = in.medium;
        c.medium.xAxis = keep(ax_[0], m.xAxis, m.xAxisSize);
        c.medium.yAxis = keep(ax_[1], m.yAxis, m.yAxisSize);
        c.medium.zAxis = dims == 3 ? keep(ax_[2], m.zAxis, m.zAxisSize) : nullptr;
        const size_t ncell = (size_t)(m.xAxisSize - 1) * (m.yAxisSize - 1) * (dims == 3 ? m.zAxisSize - 1 : 1);
        c.medium.materialGrid = withGrid ? keep(grid_, m.materialGrid, ncell) : nullptr;
        c.medium.sigs = keep(sigs_, m.sigs, m.materialNum);
        c.medium.siga = keep(siga_, m.siga, m.materialNum);

        const sstmc_scattering& s = in.scattering;
        const uint32_t ns = s.materialNum;
        c.scattering.type = keep(stype_, s.type, ns);
        c.scattering.g = keep(g_, s.g, ns);
        size_t ntab = 0, nvmf = 0;
        for (uint32_t i = 0; i < ns; ++i) {
            if (s.tabularSize) ntab += 2 * (size_t)s.tabularSize[i];
            if (s.vmfMixtures) nvmf += s.vmfMixtures[i];
        }
        c.scattering.tabularSize = keep(tabSize_, s.tabularSize, ns);
        c.scattering.tabular = keep(tab_, s.tabular, ntab);
        c.scattering.vmfMixtures = keep(vmfM_, s.vmfMixtures, ns);
        c.scattering.vmfMu = keep(vmf_[0], s.vmfMu, nvmf);
        c.scattering.vmfC = keep(vmf_[1], s.vmfC, nvmf);
        c.scattering.vmfAlpha = keep(vmf_[2], s.vmfAlpha, nvmf);

        const sstmc_source* src[4] = {&in.illumination, &in.view, &in.illumination2, &in.view2};
        sstmc_source* dst[4] = {&c.illumination, &c.view, &c.illumination2, &c.view2};
        for (int i = 0; i < 4; ++i) {
            const uint32_t n = src[i]->num;
            dst[i]->x = keep(pts_[i][0], src[i]->x, n);
            dst[i]->y = keep(pts_[i][1], src[i]->y, n);
            dst[i]->z = dims == 3 ? keep(pts_[i][2], src[i]->z, n) : nullptr;
            dst[i]->dirX = keep(dirs_[i][0], src[i]->dirX, n);
            dst[i]->dirY = keep(dirs_[i][1], src[i]->dirY, n);
            dst[i]->dirZ = keep(dirs_[i][2], src[i]->dirZ, n);
        }

        const uint32_t nmat = m.materialNum;
        c.sampler.t = keep(t_, in.sampler.t, in.sampler.tNum);
        c.sampler.D = keep(D_, in.sampler.D, nmat);
        c.sampler.Ux = keep(U_[0], in.sampler.Ux, nmat);
        c.sampler.Uy = keep(U_[1], in.sampler.Uy, nmat);
        c.sampler.Uz = dims == 3 ? keep(U_[2], in.sampler.Uz, nmat) : nullptr;
    }
    ConfigCopy(const ConfigCopy&) = delete;
    ConfigCopy& operator=(const ConfigCopy&) = delete;

private:
    template <typename T>
    static const T* keep(std::vector<T>& dst, const T* p, size_t n) {
        if (!p || n == 0) return nullptr;
        dst.assign(p, p + n);
        return dst.data();
    }
    std::vector<double> ax_[3], sigs_, siga_, g_, tab_, vmf_[3], pts_[4][3], dirs_[4][3], t_, D_, U_[3];
    std::vector<uint32_t> grid_, tabSize_, vmfM_;
    std::vector<int32_t> stype_;
};

}  // namespace sstmc
