// mdas_integrate_noalloc: mdas_integrate's reconstruction on a device array
// the library holds for the life of the process, for callers that cannot
// allocate (code under hipGraph capture, tight serving loops, harnesses that
// forbid hipMalloc).  It calls no hipMalloc, hipFree or other allocating
// routine; the kernels are mdas_integrate's, given pointers into the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "mdas_handles.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[MDAS_NOALLOC_SCRATCH_BYTES];

const int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};  // the array's device address on each device

// The array's address on the current device, looked up once per device (each
// device holds its own copy at its own address).  Host code uses only this
// address, never the array's name.
hipError_t scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

long long up256(long long b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int mdas_integrate_noalloc(mdas_handle h, float* pixels, int w, int hgt, void* stream) {
    if (!h || !pixels) return mdas::set_error(MDAS_ERR_INVALID, "null argument");
    mdas::Grid& g = h->g;
    if (w <= 0 || hgt <= 0 || w > 65535 || hgt > 65535 || (long long)w * hgt > 0x7fffffffLL)
        return mdas::set_error(MDAS_ERR_INVALID, "image size out of range");
    if (!g.initialized) return mdas::set_error(MDAS_ERR_STATE, "InitialSampling has not been called");
    const long long need = mdas::recon_bytes_needed(g, w, hgt);
    if (need < 0 || need > (long long)MDAS_NOALLOC_SCRATCH_BYTES)
        return mdas::set_error(MDAS_ERR_CAPACITY,
                               "reconstruction needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                   std::to_string((long long)MDAS_NOALLOC_SCRATCH_BYTES) +
                                   " the allocation-free path holds (use mdas_integrate)");
    void* base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess)
        return mdas::set_error(MDAS_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    // the same carving as mdas::recon_bytes_needed: histograms, tile ranges,
    // leaf records, then the two key buffers at the key bound
    const long long L = std::max<long long>(g.host.n_leaves, 1);
    const long long tiles = mdas::recon_tiles_cap(g, w, hgt);
    const long long kb = mdas::recon_key_bound(g, w, hgt);
    unsigned char* p = static_cast<unsigned char*>(base);
    mdas::ReconBufs rb;
    rb.rs_hist = reinterpret_cast<unsigned*>(p);
    p += up256(sizeof(unsigned) * mdas::kSortHistWords);
    rb.tile_start = reinterpret_cast<int*>(p);
    p += up256(sizeof(int) * tiles);
    rb.tile_end = reinterpret_cast<int*>(p);
    p += up256(sizeof(int) * tiles);
    rb.tiles_cap = tiles;
    rb.rrec = reinterpret_cast<float4*>(p);
    p += up256(sizeof(float4) * 2 * L);
    rb.rrec_cap = L;
    rb.keys = reinterpret_cast<unsigned long long*>(p);
    p += up256(sizeof(unsigned long long) * kb);
    rb.keys_alt = reinterpret_cast<unsigned long long*>(p);
    rb.keys_cap = kb;
    // no timing events (creating them would allocate)
    const bool timing = g.timing;
    g.timing = false;
    e = mdas::recon_on_buffers(g, rb, pixels, w, hgt, (hipStream_t)stream);
    g.timing = timing;
    if (e != hipSuccess) return mdas::set_error(MDAS_ERR_HIP, std::string("reconstruction: ") + hipGetErrorString(e));
    return MDAS_OK;
}
