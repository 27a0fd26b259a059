// Product runtime: the host facade that mirrors the reference's
// SPGrid_Domain_Decomposition_Solver (DDS/SPGrid_Domain_Decomposition_Solver.h:24-198)
// and runs every hot-path stage on the GPU.
#pragma once
#include <memory>
#include <vector>
#include "hdd_device.h"
#include "hdd_setup.h"
#include "hdd_sigma.h"

namespace hdd {

struct SolverParams {
    Int3 size, sub;
    int vlevels = 3, sigma_levels = 3, nV = 1, int_s = 1, bnd_s = 3, sig_s = 3, sig_mg = 1;
};

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() {}
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    ~DBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t count) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count) { hdd_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc"); hdd_check(hipMemset(p, 0, count * sizeof(T)), "hipMemset"); }
    }
    void upload(const T* h, size_t count) { alloc(count); if (count) hdd_check(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice), "upload"); }
    void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
    std::vector<T> download() const {
        std::vector<T> v(n);
        if (n) hdd_check(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "download");
        return v;
    }
};

struct LevelBuffers {
    PagedLevel host;
    DBuf<uint32_t> flags;
    DBuf<int> nbr, base, lut;
    DBuf<float> u[2], b, r;
    int cur = 0;  // which u buffer holds the current iterate
    DevLevel view() const;
    size_t cells() const { return (size_t)(host.npages() + 1) * kPageCells; }
};

class Solver {
public:
    SolverParams P;
    std::vector<LevelBuffers> lv;   // V-cycle levels; lv[0] is the main grid
    DBuf<float> x, bvec, q, s;      // second-allocator channels (x, b, q, s)
    DBuf<double> dpart, dscal;
    DBuf<float> fpart, scal;        // scal: rho, alpha, beta, rho_old, max|r|
    SigmaHost sigma_host;
    std::unique_ptr<SigmaDevice> sigma;
    DBuf<int> if_index;             // level-0 cell index of each interface dof
    hipStream_t stream = nullptr;
    double setup_seconds = 0;

    // host_only: run the host setup and skip every device allocation (introspection)
    Solver(const SolverParams& p, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, const float* b,
           bool host_only = false);
    bool host_only = false;
    ~Solver();

    // level-0 channel access in page order (c: 0=z 1=r 2=k 3=x 4=b 5=q 6=s)
    float* channel(int c);
    void vcycle();
    void smooth0(int mode);
    void apply_preconditioner();  // z <- M^{-1} r
    int pcg(double tol_abs, int max_it, std::vector<double>& hist, double* seconds);

private:
    void build_levels(int n_pages, const uint64_t* page_offsets, const uint32_t* flags);
    void upload_levels();
    void sweeps(int l, int count, bool interior, bool boundary);
    void interface_gather();
    void interface_scatter();
};

}  // namespace hdd
