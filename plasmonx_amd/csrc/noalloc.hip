// pxh_matvec_noalloc: the FP64 on-the-fly product of pxh_matvec for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Inputs and outputs are
// the caller's device buffers; the product's own working memory (system
// list, per-system parameters, split partials, t = T x) lives in one
// file-scope device array per GPU, so the call allocates nothing.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>
#include <string>

#include "pxh_internal.h"

namespace {

// Working memory of pxh_matvec_noalloc.  Each GPU the library runs on holds
// its own copy until the process ends.  Host code never uses this name as a
// device address: kernels take the address scratch_setup looked up for the
// device (g_scratch_ptr).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// list[k] = k and par[k] = p for the n_vec vectors of the call.
__global__ void k_fill_systems(int n_vec, pxh::SysParam p, int *__restrict__ list, pxh::SysParam *__restrict__ par) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_vec) return;
    list[k] = k;
    par[k] = p;
}

}  // namespace

namespace pxh {

// Device address of this GPU's scratch array, looked up once per device with
// that device current (called from pxh_ctx_create).
void scratch_setup(int device) {
    PXH_CHECK(device >= 0 && device < kMaxDevices, PXH_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    PXH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PXH_CHECK(p != nullptr, PXH_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = static_cast<unsigned char *>(p);
}

}  // namespace pxh

extern "C" int pxh_matvec_noalloc(pxh_ctx *c, double freq_au, int i_freq, int n_vec, const double *d_x,
                                  double *d_y) {
    if (!c) return PXH_EINVAL;
    try {
        PXH_HIP(hipSetDevice(c->device));
        pxh::DeviceShared gate;
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU, PXH_EINVAL,
                  "pxh_matvec_noalloc: no atomistic target set (BEM products are not covered)");
        PXH_CHECK(n_vec > 0 && d_x && d_y, PXH_EINVAL, "pxh_matvec_noalloc: bad vectors");
        pxh::scratch_setup(c->device);  // no-op once the device's address is known
        unsigned char *base = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            base = g_scratch_ptr[c->device];
        }
        const size_t nv = (size_t)c->atoms.n_var();
        const size_t o_list = 0;
        const size_t o_par = o_list + align256(sizeof(int) * (size_t)n_vec);
        const size_t o_part = o_par + align256(sizeof(pxh::SysParam) * (size_t)n_vec);
        const size_t o_t = o_part + align256(sizeof(double2) * pxh::fp64_product_partials(c, n_vec));
        const size_t need = o_t + align256(sizeof(double2) * nv * (size_t)n_vec);
        if (need > kScratchBytes)
            throw pxh::Error{PXH_ENOMEM, "pxh_matvec_noalloc: " + std::to_string(n_vec) + " vectors of " +
                                             std::to_string(nv) + " unknowns need " + std::to_string(need >> 20) +
                                             " MiB of working memory, more than the 64 MiB scratch; use "
                                             "pxh_matvec or fewer vectors per call"};
        int *d_list = reinterpret_cast<int *>(base + o_list);
        pxh::SysParam *d_par = reinterpret_cast<pxh::SysParam *>(base + o_par);
        double2 *d_part = reinterpret_cast<double2 *>(base + o_part);
        double2 *d_t = reinterpret_cast<double2 *>(base + o_t);
        k_fill_systems<<<(n_vec + 63) / 64, 64, 0, c->stream>>>(n_vec, pxh::make_sysparam(c, freq_au, i_freq), d_list,
                                                               d_par);
        PXH_HIP(hipGetLastError());
        pxh::apply_atoms_fp64(c, n_vec, d_list, d_par, reinterpret_cast<const double2 *>(d_x),
                              reinterpret_cast<double2 *>(d_y), (int)nv, d_part, d_t);
        PXH_HIP(hipStreamSynchronize(c->stream));
        return PXH_OK;
    } catch (const pxh::Error &e) {
        c->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        c->last_error = "host allocation failed";
        return PXH_ENOMEM;
    }
}
