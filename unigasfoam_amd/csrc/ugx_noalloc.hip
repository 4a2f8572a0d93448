// ugx_noalloc.hip - ugx_case_fill_noalloc: the meshFill of ugx_case_fill
// with its scratch in a fixed device array, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// The scratch of a fill is its parameter block (number densities, per-axis
// temperatures, drift velocity), the simulator count of every (cell,
// species) pair and the exclusive scan of those counts.  ugx_case_fill
// allocates them per call; here they live in `scratch`, a file-scope device
// array of kScratchWords 32-bit words (16 MiB) that each GPU holds for the
// life of the process.  Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is
// handed to the kernels.  Every call on one GPU shares the array: calls that
// overlap on a GPU (from several threads or streams) race on it.  Nothing of
// it needs to start at zero: the count kernel writes every count (and the
// zero that ends the scan) and the scan writes every offset.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ugx_world.hpp"

using namespace ugx;

// 4 Mi words = 16 MiB per GPU
constexpr uint32_t kScratchWords = 4u << 20;
constexpr uint32_t kParamWords = 64;                     // kFillParams doubles, padded to 256 B
static_assert(kFillParams * 2 <= (int)kParamWords, "the parameter block fits its slot");
// largest (cells x species) + 1 of a fill: counts and offsets, each padded to 64 words
constexpr uint32_t kMaxFillEntries = ((kScratchWords - kParamWords) / 2u) & ~63u;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on `device`, looked up on the first call there.
int scratch_address(int device, uint32_t** out)
{
    static std::mutex mu;
    static uint32_t* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return set_err(UGX_ERR_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipGetDevice failed");
        if (hipSetDevice(device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice(%d) failed", device);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return set_err(UGX_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
        addr[device] = static_cast<uint32_t*>(p);
    }
    *out = addr[device];
    return UGX_OK;
}

}  // namespace

extern "C" int ugx_case_fill_noalloc(ugx_case* c, const ugx_fill_desc* fill)
{
    try {
        if (!c || !c->world) return set_err(UGX_ERR_ARG, "null case");
        if (c->batch) return set_err(UGX_ERR_STATE, "a batch case is filled at ugx_batch_admit");
        if (!fill) return set_err(UGX_ERR_ARG, "null fill descriptor");
        for (int d = 0; d < 3; ++d) if (!(fill->T[d] > 0.0)) return set_err(UGX_ERR_ARG, "fill temperature must be > 0");
        World* w = c->world;
        if (hipSetDevice(w->device) != hipSuccess) return set_err(UGX_ERR_DEVICE, "hipSetDevice failed");
        const MeshGeom& g = w->specs_host[c->index].geom;
        const uint64_t entries = (uint64_t)g.nc * (uint64_t)w->cases[c->index].ns + 1u;
        if (entries > kMaxFillEntries)
            return set_err(UGX_ERR_CAPACITY,
                           "allocation-free fill: %llu (cell, species) entries exceed the %u of its %u-word scratch; "
                           "use ugx_case_fill", (unsigned long long)entries, kMaxFillEntries, kScratchWords);
        uint32_t* base = nullptr;
        if (int rc = scratch_address(w->device, &base)) return rc;
        const uint64_t padded = (entries + 63u) & ~63ull;
        World::FillScratch fs{reinterpret_cast<double*>(base), base + kParamWords, base + kParamWords + padded};
        return w->fill(c->index, fill, g, true, &fs);
    } catch (...) {
        return set_err(UGX_ERR_STATE, "unexpected exception");
    }
}
