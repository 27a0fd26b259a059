#include "hdd_sigma_device.h"
namespace hdd {
SigmaDevice::~SigmaDevice() {}
std::unique_ptr<SigmaDevice> make_sigma_device(const SigmaHost&, int, int) { return std::unique_ptr<SigmaDevice>(new SigmaDevice()); }
void sigma_vcycles(SigmaDevice&, hipStream_t) {}
void launch_interface_gather(SigmaDevice&, const DevLevel&, const int*, const float*, const float*, hipStream_t) {}
void launch_interface_scatter(SigmaDevice&, const int*, float*, hipStream_t) {}
void sigma_vcycle_host(SigmaDevice&, const float*, float*, hipStream_t) { throw std::runtime_error("sigma not built"); }
void sigma_smooth_host(SigmaDevice&, int, const float*, float*, hipStream_t) { throw std::runtime_error("sigma not built"); }
void generate_interface_operators(const PagedLevel&, Int3, Int3, int levels, SigmaHost& out) { out.levels = 0; }
}
