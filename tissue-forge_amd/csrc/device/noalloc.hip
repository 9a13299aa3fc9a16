// The no-allocation force pass: engine_force for a host engine with every
// device buffer carved from one fixed array per GPU instead of hipMalloc.
//
// The array (scratch) is a file-scope device variable: each GPU the library
// runs on holds its own copy, at its own address, from the module's load
// until the process ends.  Host code never uses the array's name as a device
// address (on the host it names a stand-in); the address of each GPU's copy
// comes from hipGetSymbolAddress on the first call on that GPU and is kept
// per device.  Every call on one GPU shares that copy and finds the last
// call's contents in it: the upload writes or clears every buffer the pass
// reads before the first kernel runs.  Overlapping calls on one GPU would
// race on the array, so a per-device lock holds each call until its GPU work
// has drained.
#include <mutex>

#include "ensemble.h"

namespace tfg {

namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(64) << 20;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

void *g_scratch[kScratchDevices] = {};
std::mutex g_scratch_lock[kScratchDevices];

}  // namespace

size_t scratch_bytes() { return kScratchBytes; }

int nonbond_eval_no_alloc(Engine *e, int precision, double *f, Energies *en) {
    if (e->parts.empty()) return set_error(TFG_E_INVALIDARG, "engine has no particles");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(TFG_E_DEVICE, "no HIP device available");
    int dev = -1;
    TFG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return set_error(TFG_E_RANGE, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> hold(g_scratch_lock[dev]);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !ptr)
            return set_error(TFG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        g_scratch[dev] = ptr;
    }
    return ensemble_force_arena(e, precision, g_scratch[dev], kScratchBytes, f, en);
}

}  // namespace tfg
