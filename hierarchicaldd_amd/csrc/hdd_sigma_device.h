// Device-side sigma (interface Schur complement) multigrid.
#pragma once
#include <memory>
#include "hdd_device.h"
#include "hdd_sigma.h"

namespace hdd {

struct SigmaDevice {
    int levels = 0;
    int n_if0 = 0;
    float* b0 = nullptr;   // level-0 interface rhs
    float* u0 = nullptr;   // level-0 interface solution
    ~SigmaDevice();
};

std::unique_ptr<SigmaDevice> make_sigma_device(const SigmaHost& host, int smoothing, int n_mg);
void sigma_vcycles(SigmaDevice& S, hipStream_t s);
void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const float* z, const float* r, hipStream_t s);
void launch_interface_scatter(SigmaDevice& S, const int* if_index, float* z, hipStream_t s);
void sigma_vcycle_host(SigmaDevice& S, const float* b, float* u, hipStream_t s);
void sigma_smooth_host(SigmaDevice& S, int level, const float* b, float* u, hipStream_t s);

}  // namespace hdd
