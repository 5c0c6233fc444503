// bk_noalloc.hip -- the allocation-free sampling entry point.
//
// bk_sample_noalloc runs the kernels of bk_sample (FP64 block sums of |psi|^2,
// one-workgroup exclusive scan, one wave per shot) on a caller-held device
// state, with device random numbers and a device output.  Its only scratch,
// the block sums and their prefixes, lives in the file-scope array below, so
// the call makes no device allocation: it serves code that cannot call
// hipMalloc (work under graph capture, serving loops, harnesses that forbid
// allocation).
//
// Host code never passes `scratch` itself to a kernel: in host code the name
// is a stand-in, not a device address.  The first call on each device looks
// up that device's copy with hipGetSymbolAddress and keeps the pointer; every
// later use goes through it.
#include "bk_ctx.hpp"

#include <mutex>

namespace {

constexpr int kMaxDevices = 64;
constexpr int kNoAllocMaxN = BK_MAX_N;
// block sums (2^(n-12)) followed by the exclusive prefixes (2^(n-12) + 1)
constexpr size_t kMaxBlocks = size_t(1) << (kNoAllocMaxN - bk::kSampleBlockBits);
constexpr size_t kScratchDoubles = 2 * kMaxBlocks + 32;  // 16 MiB + 256 B

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

using namespace bkc;

// Device address of this device's scratch array, looked up once per device.
static int scratch_for_current_device(double** out)
{
    int dev = -1;
    HIPCALL(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BK_E_INVALID, "bk_sample_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        HIPCALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(BK_E_HIP, "bk_sample_noalloc: hipGetSymbolAddress returned a null address");
        g_ptr[dev] = p;
    }
    *out = static_cast<double*>(g_ptr[dev]);
    return BK_OK;
}

static int sample_noalloc_impl(int n, int precision, const void* d_amps, const double* d_randnums, int64_t shots,
                               int64_t* d_out_idx, void* hip_stream)
{
    if (n < 1 || n > kNoAllocMaxN)
        return fail(BK_E_INVALID, "bk_sample_noalloc: n = %d outside 1..%d (the static scratch holds %zu doubles)", n,
                    kNoAllocMaxN, kScratchDoubles);
    if (precision != BK_C64 && precision != BK_C128)
        return fail(BK_E_INVALID, "bk_sample_noalloc: precision must be BK_C64 or BK_C128");
    if (shots < 0 || shots > (int64_t(1) << 30)) return fail(BK_E_INVALID, "bk_sample_noalloc: shots outside 0..2^30");
    if (!d_amps || (shots > 0 && (!d_randnums || !d_out_idx)))
        return fail(BK_E_INVALID, "bk_sample_noalloc: null pointer");
    if (shots == 0) return BK_OK;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    int dev = -1;
    HIPCALL(hipGetDevice(&dev));
    if (st) {
        hipDevice_t sdev = -1;
        HIPCALL(hipStreamGetDevice(st, &sdev));
        if (sdev != dev)
            return fail(BK_E_INVALID, "bk_sample_noalloc: stream belongs to device %d, current device is %d", sdev, dev);
    }
    double* scr = nullptr;
    int rc = scratch_for_current_device(&scr);
    if (rc) return rc;
    const size_t nb = n > bk::kSampleBlockBits ? size_t(1) << (n - bk::kSampleBlockBits) : 0;
    HIPCALL(bk::launch_sample(d_amps, n, precision, d_randnums, static_cast<int>(shots), d_out_idx, scr, scr + nb,
                              st));
    return BK_OK;
}

extern "C" int bk_sample_noalloc(int n, int precision, const void* d_amps, const double* d_randnums, int64_t shots,
                                 int64_t* d_out_idx, void* hip_stream)
{
    return guard_abi([&] { return sample_noalloc_impl(n, precision, d_amps, d_randnums, shots, d_out_idx, hip_stream); });
}
