// PoissonBackend over the C ABI of libsuperplace_hip.so
// (sp_hip_poisson_create / _solve / _name / _destroy).
#include "hip_poisson.h"

#include <stdexcept>
#include <string>

#include "superplace_hip.h"

namespace sp {
namespace {

void check(int status, const char* what) {
    if (status != SP_HIP_OK)
        throw std::runtime_error(std::string(what) + ": " + sp_hip_last_error());
}

class HipPoissonBackend final : public PoissonBackend {
public:
    HipPoissonBackend(sp_hip_poisson* handle, int nx, int ny) : h_(handle), nx_(nx), ny_(ny) {}
    ~HipPoissonBackend() override { sp_hip_poisson_destroy(h_); }

    // The spans' checks of CpuPoissonBackend::solve (cpu_poisson.cpp:220-223).
    void solve(BinGridSpan density, BinGridSpan phi, BinGridSpan field_x,
               BinGridSpan field_y) override {
        for (const BinGridSpan* s : {&density, &phi, &field_x, &field_y})
            if (!s->valid() || s->nx != nx_ || s->ny != ny_)
                throw std::runtime_error("HipPoissonBackend::solve: span dimension mismatch");
        check(sp_hip_poisson_solve(h_, density.data, phi.data, field_x.data, field_y.data),
              "HipPoissonBackend::solve");
    }

    const char* name() const override { return sp_hip_poisson_name(h_); }

private:
    sp_hip_poisson* h_;
    int nx_, ny_;
};

}  // namespace

std::unique_ptr<PoissonBackend> makeHipPoissonBackend(const BackendConfig& cfg) {
    sp_hip_poisson* h = nullptr;
    const int status = sp_hip_poisson_create(cfg.nx, cfg.ny, cfg.step_x, cfg.step_y, 0, &h);
    if (status == SP_HIP_ENODEV) return nullptr;
    check(status, "makeHipPoissonBackend");
    return std::make_unique<HipPoissonBackend>(h, cfg.nx, cfg.ny);
}

}  // namespace sp
