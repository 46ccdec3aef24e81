// A C++ caller of include/alsvinn_hip.h (no Python in between): advances a
// 2D Euler state through the per-call Volume entry points in the order the
// reference's Simulator::incrementSolution uses them
// (alsfvm/src/simulator/Simulator.cpp:191-214) -- for each RK3 substep
// Integrator::performSubstep then Boundary::applyBoundaryConditions, then the
// swap of the first and last volume and the constraint check.
//
//   ref_step <in.bin> <out.bin> <nx> <ny> <steps>
//
// in.bin: the ghosted state, 4 arrays of (nx+4)*(ny+4) doubles (x fastest),
// ghosts already filled.  out.bin: the dt of every step, then the ghosted
// state after the last step.  Exit status 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <vector>

#include "../../include/alsvinn_hip.h"

namespace {

bool ok(avh_status s, const char* what) {
    if (s == AVH_OK) return true;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, (int)s, avh_last_error());
    return false;
}

// A reference Volume in device memory: one array per conserved variable.
struct DeviceVolume {
    avh_volume v{};
    size_t cells = 0;
    DeviceVolume(int nv, int nx, int ny, int g) {
        v.nvars = nv;
        v.n[0] = nx;
        v.n[1] = ny;
        v.n[2] = 1;
        v.ghosts[0] = g;
        v.ghosts[1] = g;
        v.ghosts[2] = 0;
        cells = (size_t)(nx + 2 * g) * (ny + 2 * g);
        for (int i = 0; i < nv; ++i) {
            void* p = nullptr;
            if (hipMalloc(&p, cells * sizeof(double)) != hipSuccess || hipMemset(p, 0, cells * sizeof(double)) != hipSuccess) {
                std::fprintf(stderr, "device allocation failed\n");
                std::exit(2);
            }
            v.var[i] = (double*)p;
        }
    }
    ~DeviceVolume() {
        for (int i = 0; i < v.nvars; ++i) (void)hipFree(v.var[i]);
    }
    DeviceVolume(const DeviceVolume&) = delete;
    DeviceVolume& operator=(const DeviceVolume&) = delete;
    bool upload(const double* host) {
        for (int i = 0; i < v.nvars; ++i)
            if (hipMemcpy(v.var[i], host + i * cells, cells * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                return false;
        return true;
    }
    bool download(double* host) const {
        for (int i = 0; i < v.nvars; ++i)
            if (hipMemcpy(host + i * cells, v.var[i], cells * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                return false;
        return true;
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: ref_step <in.bin> <out.bin> <nx> <ny> <steps>\n");
        return 2;
    }
    const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), steps = std::atoi(argv[5]);
    const int nv = 4, g = 2;
    avh_config cfg{};
    cfg.equation = AVH_EQ_EULER2;
    cfg.reconstruction = AVH_RECON_WENO2;
    cfg.flux = AVH_FLUX_HLL3;
    cfg.integrator = AVH_INT_RUNGEKUTTA3;
    cfg.boundary = AVH_BC_PERIODIC;
    cfg.n[0] = nx;
    cfg.n[1] = ny;
    cfg.n[2] = 1;
    cfg.upper[0] = cfg.upper[1] = cfg.upper[2] = 1.0;
    cfg.gamma = 1.4;
    cfg.cfl = 0.4;

    const size_t cells = (size_t)(nx + 2 * g) * (ny + 2 * g);
    std::vector<double> host((size_t)nv * cells);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::fclose(f);

    avh_ctx* ctx = nullptr;
    if (!ok(avh_ctx_create(0, &ctx), "avh_ctx_create")) return 1;
    int rc = 0;
    {
        // Simulator's volumes: U0 .. U3 (alsfvm/src/simulator/Simulator.cpp:59-63)
        std::vector<DeviceVolume*> vols;
        for (int i = 0; i < 4; ++i) vols.push_back(new DeviceVolume(nv, nx, ny, g));
        if (!vols[0]->upload(host.data())) rc = 1;
        std::vector<double> dts;
        for (int step = 0; step < steps && rc == 0; ++step) {
            double dt = 0.0;
            for (int s = 0; s < 3 && rc == 0; ++s) {
                const avh_volume* inputs[3] = {&vols[0]->v, &vols[1]->v, &vols[2]->v};
                double used = 0.0;
                if (!ok(avh_substep(ctx, &cfg, inputs, s, &vols[s + 1]->v, dt,
                                    std::numeric_limits<double>::infinity(), &used, nullptr),
                        "avh_substep") ||
                    !ok(avh_apply_boundary(ctx, AVH_BC_PERIODIC, &vols[s + 1]->v, nullptr), "avh_apply_boundary")) {
                    rc = 1;
                    break;
                }
                dt = used;
            }
            if (rc) break;
            std::swap(vols[0], vols[3]);
            int obeys = 0;
            if (!ok(avh_obeys_constraints(ctx, &cfg, &vols[0]->v, &obeys, nullptr), "avh_obeys_constraints")) {
                rc = 1;
                break;
            }
            if (!obeys) {
                std::fprintf(stderr, "constraint violation at step %d\n", step);
                rc = 3;
                break;
            }
            dts.push_back(dt);
        }
        if (rc == 0 && vols[0]->download(host.data())) {
            FILE* o = std::fopen(argv[2], "wb");
            if (!o || std::fwrite(dts.data(), sizeof(double), dts.size(), o) != dts.size() ||
                std::fwrite(host.data(), sizeof(double), host.size(), o) != host.size()) {
                std::fprintf(stderr, "cannot write %s\n", argv[2]);
                rc = 2;
            }
            if (o) std::fclose(o);
        } else if (rc == 0) {
            rc = 1;
        }
        for (auto* v : vols) delete v;
    }
    avh_ctx_destroy(ctx);
    return rc;
}
