// gfe_noalloc.hip -- squarings that allocate nothing (gfe_square_noalloc, include/gfe.h).
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation): the caller owns the digits, and everything else a squaring
// needs -- the transform work space, the twiddle tables, the normalisation carries and the CRT
// constants -- lives in `scratch`, one static device array per device, held until the process ends.
// The squarings run the engine's own passes (Engine::bind, the engine's default plan), so the
// digits equal those of gfe_square_dup on an engine of the same (b, n).
//
// Every call rebuilds the twiddles and the CRT constants for its (b, n) with two small kernels (no
// host staging, no memset, no copy), so a call never depends on what an earlier call left in the
// array; calls that overlap on one device race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/gfe.h"
#include "gfe_engine.h"

namespace gfe {

namespace {

constexpr size_t NA_N = (size_t)1 << GFE_NOALLOC_MAX_N;
constexpr size_t NA_WS = 3 * NA_N * 4;      // rns x N residues
constexpr size_t NA_TW = 3 * NA_N * 4;      // rns x N twiddles
constexpr size_t NA_CARRY = NA_N;           // N / 8 int64, as the engine's carry buffer
constexpr size_t NA_CRT = 256;
static_assert(sizeof(CrtConst) <= NA_CRT, "CRT constants fit their slot");
static_assert(NA_WS + NA_TW + NA_CARRY + NA_CRT == GFE_NOALLOC_BYTES, "the documented size is the array's");

// The library's scratch array.  Host code never uses this name as an address: the device address
// of each device's copy comes from hipGetSymbolAddress (scratch_on).
static __device__ __align__(256) unsigned char scratch[GFE_NOALLOC_BYTES];

constexpr int NA_MAX_DEV = 64;
std::mutex g_na_mu;
void *g_na_ptr[NA_MAX_DEV] = {};

// Device address of `scratch` on the current device `dev`, looked up once per device.
int scratch_on(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(GFE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_ptr[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_na_ptr[dev]);
    return GFE_OK;
}

struct TwArgs {
    uint32_t *tw;
    uint32_t p[3], q[3], psim[3], one[3];  // prime, p^-1 mod 2^32, psi 2^32 mod p, 2^32 mod p
    int n, rns;
};

__device__ __forceinline__ uint32_t mont_full(uint32_t a, uint32_t b, uint32_t p, uint32_t q) {
    const uint64_t t = (uint64_t)a * b;
    const uint32_t m = (uint32_t)t * q;
    const uint32_t hi = (uint32_t)(t >> 32), mh = __umulhi(m, p);
    return hi >= mh ? hi - mh : hi - mh + p;
}

// Psi[j] = psi^bitrev_n(j) 2^32 mod p for prime k = blockIdx.y (the host tables of the engine,
// make_twiddles): square-and-multiply in Montgomery form, every product fully reduced.
__global__ void __launch_bounds__(256) k_na_twiddles(TwArgs A) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    const uint32_t N = 1u << A.n;
    if (j >= N || (int)k >= A.rns) return;
    const uint32_t e = __brev(j) >> (32 - A.n);
    const uint32_t p = A.p[k], q = A.q[k];
    uint32_t acc = A.one[k], base = A.psim[k];
    for (uint32_t x = e; x != 0; x >>= 1) {
        if (x & 1u) acc = mont_full(acc, base, p, q);
        base = mont_full(base, base, p, q);
    }
    A.tw[(size_t)k * N + j] = acc;
}

__global__ void k_na_crt(CrtConst c, CrtConst *dst) {
    if (threadIdx.x == 0) *dst = c;
}

}  // namespace

}  // namespace gfe

extern "C" int gfe_square_noalloc(uint32_t b, uint32_t n, int device, int32_t *digits, uint32_t count,
                                  uint64_t dup_bits, void *stream) {
    using namespace gfe;
    if (n < 12 || n > GFE_NOALLOC_MAX_N)
        return set_error(GFE_EINVAL, "square_noalloc: n must be in [12, " + std::to_string(GFE_NOALLOC_MAX_N) +
                                         "] (the scratch array's limit)");
    if (b < 1024 || b > 2000000000u) return set_error(GFE_EINVAL, "square_noalloc: b must be in [1024, 2000000000]");
    if (!digits) return set_error(GFE_EINVAL, "square_noalloc: null digits");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(GFE_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev || device >= NA_MAX_DEV) return set_error(GFE_ENODEV, "device index out of range");
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, digits) != hipSuccess || attr.type != hipMemoryTypeDevice ||
        attr.device != device) {
        (void)hipGetLastError();
        return set_error(GFE_EINVAL, "square_noalloc: digits must be device memory of the given device");
    }
    if (hipSetDevice(device) != hipSuccess) return set_error(GFE_EHIP, "hipSetDevice failed");
    unsigned char *base = nullptr;
    int rc = scratch_on(device, &base);
    if (rc != GFE_OK) return rc;
    if (count == 0) return GFE_OK;
    engine_set_lds_attributes();

    const size_t N = (size_t)1 << n;
    uint32_t *ws = reinterpret_cast<uint32_t *>(base);
    uint32_t *tw = reinterpret_cast<uint32_t *>(base + NA_WS);
    int64_t *carry = reinterpret_cast<int64_t *>(base + NA_WS + NA_TW);
    CrtConst *crt = reinterpret_cast<CrtConst *>(base + NA_WS + NA_TW + NA_CARRY);
    hipStream_t s = static_cast<hipStream_t>(stream);

    Engine e;
    e.bind(b, n, device, s, digits, ws, tw, carry, crt, default_plan(n));
    TwArgs T;
    T.tw = tw; T.n = (int)n; T.rns = e.rns;
    for (int k = 0; k < 3; ++k) {
        T.p[k] = prime_of(e.is32, k);
        T.q[k] = prime_qinv(T.p[k]);
        T.psim[k] = psi_mont(e.is32, k, n);
        T.one[k] = mulw_const(1, T.p[k]);
    }
    hipLaunchKernelGGL(k_na_twiddles, dim3((unsigned)(N / 256), (unsigned)e.rns, 1), dim3(256), 0, s, T);
    hipLaunchKernelGGL(k_na_crt, dim3(1), dim3(64), 0, s, make_crt(e.rns, e.is32, n, b), crt);
    for (uint32_t i = 0; i < count; ++i) e.square((int)((dup_bits >> (i % 64)) & 1));
    e.flush_carry();
    return e.check("square_noalloc");
}
