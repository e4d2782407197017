// pic_noalloc.hip -- pic_deposit_noalloc: the stand-alone deposit
// (pic_deposit) for callers that may not allocate device memory.
//
// pic_deposit stages the caller's seven host arrays (old and new positions,
// weights) in a device buffer it allocates per call, plus a 16-byte counter
// of skipped particles.  This variant stages them in one file-scope device
// array instead, sized for PIC_NOALLOC_MAX_PARTICLES particles, and runs the
// same deposit kernel.  Each GPU the library runs on holds its own copy of
// the array (3.5 MiB) until the process ends.  Host code never passes the
// array's name to a kernel or a copy: the name is a host-side stand-in, not
// a device address.  The device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept per device.
//
// Every call on one GPU shares that GPU's array; each call overwrites it
// and clears its counter before use, and leaves its contents behind.  Calls
// that overlap on one GPU (two host threads, or two contexts on one device
// driven concurrently) race on it.
#include <mutex>

#include "pic_internal.h"

using namespace pic;

namespace {

constexpr int64_t kMaxN = PIC_NOALLOC_MAX_PARTICLES;
constexpr int64_t kArrayBytes = kMaxN * 8;            // one double per particle
constexpr int64_t kCounterOffset = 7 * kArrayBytes;   // skipped-particle counter (16 bytes)
constexpr int64_t kScratchBytes = kCounterOffset + 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // device address of 'scratch' per device (null: not looked up)

// The device address of this device's scratch array (the caller has made
// c->device current).
int scratch_ptr(pic_ctx* c, unsigned char** out) {
    if (c->device < 0 || c->device >= kMaxDevices)
        return set_err(c, PIC_ERR_ARG, "pic_deposit_noalloc: device %d out of range", c->device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[c->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(c, PIC_ERR_HIP, "pic_deposit_noalloc: hipGetSymbolAddress failed: %s",
                           hipGetErrorString(e));
        g_ptr[c->device] = p;
    }
    *out = (unsigned char*)g_ptr[c->device];
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_deposit_noalloc(pic_ctx* c, int s, int64_t n, const double* x_old, const double* y_old,
                        const double* z_old, const double* x, const double* y, const double* z,
                        const double* w, int64_t* skipped) {
    if (!c || s < 0 || s >= PIC_MAX_SPECIES || !c->sp[s].declared || n < 0) return PIC_ERR_ARG;
    if (n > 0 && (!x_old || !y_old || !z_old || !x || !y || !z || !w)) return PIC_ERR_ARG;
    if (n > kMaxN)
        return set_err(c, PIC_ERR_CAPACITY,
                       "pic_deposit_noalloc: %lld particles exceed its staging array of %lld "
                       "(PIC_NOALLOC_MAX_PARTICLES); use pic_deposit or split the call",
                       (long long)n, (long long)kMaxN);
    if (c->fp32) return set_err(c, PIC_ERR_STATE, "pic_deposit_noalloc is an FP64 entry point");
    Species& S = c->sp[s];
    if (S.immobile) return set_err(c, PIC_ERR_STATE, "species %d is immobile (no charge array)", s);
    if (skipped) *skipped = 0;
    if (n == 0) return PIC_OK;
    PIC_HIP(c, hipSetDevice(c->device));
    unsigned char* base = nullptr;
    int r = scratch_ptr(c, &base);
    if (r) return r;
    const double* src[7] = {x_old, y_old, z_old, x, y, z, w};
    const double* d[7];
    for (int i = 0; i < 7; i++) {
        d[i] = (const double*)(base + i * kArrayBytes);
        PIC_HIP(c, hipMemcpyAsync((void*)d[i], src[i], n * 8, hipMemcpyHostToDevice, c->stream));
    }
    int64_t* d_skip = (int64_t*)(base + kCounterOffset);
    if ((r = launch_fill(c, d_skip, 16))) return r;  // the previous call's count is left there
    if ((r = launch_deposit(c, s, n, d, d[6], d_skip))) return r;
    int64_t h = 0;
    PIC_HIP(c, hipMemcpyAsync(&h, d_skip, 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    if (skipped) *skipped = h;
    return PIC_OK;
}

}  // extern "C"
