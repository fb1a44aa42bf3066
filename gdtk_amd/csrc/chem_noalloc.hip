// chem_noalloc.hip -- the chemistry step of gdtk_hip_chem_step_device with its work buffers
// in a fixed device array of this library instead of memory the library allocates, for
// callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The array is one per GPU (each device holds its own copy of a __device__ variable) and
// lives until the process ends.  Every call on a GPU uses that GPU's copy, so calls that
// overlap on one GPU race; the step clears what must start at zero (queue counters and
// statistics) with a fill kernel, and every other buffer is written before it is read.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gdtk_hip.h"
#include "chem_host.hpp"
#include "host_common.hpp"

namespace gdtk {

static_assert(GDTK_CHEM_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");

// The work region of gdtk_hip_chem_step_device_noalloc.  In host code its name stands for
// no device address: the address is taken per device with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[GDTK_CHEM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of the current device's scratch array, looked up once per device
// with that device current.
void* scratch_region() {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) throw Error(GDTK_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error(GDTK_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

}  // namespace gdtk

using namespace gdtk;

extern "C" {

int gdtk_hip_chem_step_device_noalloc(int chem, int n, const double* rho, const double* T, const double* u,
                                      const double* p, const double* massf, const double* dt_chem, double dt,
                                      double* T_out, double* u_out, double* p_out, double* massf_out,
                                      double* dt_chem_out, int* status, void* stream) {
    return guarded([&] {
        Chem& c = get_chem(chem);
        if (n < 0) throw Error(GDTK_ERR_ARG, "negative cell count");
        if (n > 0 && (!rho || !T || !u || !p || !massf || !dt_chem || !T_out || !u_out || !p_out || !massf_out ||
                      !dt_chem_out || !status))
            throw Error(GDTK_ERR_ARG, "null array");
        const size_t need = c.work_bytes(n);
        if (need > (size_t)GDTK_CHEM_NOALLOC_SCRATCH_BYTES)
            throw Error(GDTK_ERR_ARG, "gdtk_hip_chem_step_device_noalloc: " + std::to_string(n) + " cells need " +
                                          std::to_string(need) + " bytes of work space, more than the fixed " +
                                          std::to_string((size_t)GDTK_CHEM_NOALLOC_SCRATCH_BYTES) +
                                          " (gdtk_hip_chem_noalloc_max_cells gives the limit)");
        const ChemWork w = c.carve_work(n, scratch_region());
        c.step_device(n, dt, rho, T, u, p, massf, dt_chem, T_out, u_out, p_out, massf_out, dt_chem_out, status,
                      static_cast<hipStream_t>(stream), &w);
        return GDTK_OK;
    });
}

int gdtk_hip_chem_noalloc_max_cells(int chem) {
    return guarded([&] {
        Chem& c = get_chem(chem);
        int lo = 0, hi = 1 << 30;
        while (lo < hi) {
            const int mid = lo + (hi - lo + 1) / 2;
            if (c.work_bytes(mid) <= (size_t)GDTK_CHEM_NOALLOC_SCRATCH_BYTES) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    });
}

}  // extern "C"
