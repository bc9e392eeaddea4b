// Utility kernels: on-device synthetic input.
//
// gen_symbols fills bytes [0, n) with i.i.d. symbols lo + (r & (span - 1))
// (span a power of two <= 256), r taken from a counter-based hash: byte i is
// byte (i % 8) of splitmix64(seed ^ ((i / 8) * 0x9E3779B97F4A7C15)).
// splitmix64 is the published finaliser of Steele, Lea and Flood (2014).
// The same function on the host reproduces any byte range, so inputs far
// larger than host memory can be generated in HBM and checked piecewise.

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint32_t sym4(uint32_t h, uint32_t lo4, uint32_t mask4) { return lo4 + (h & mask4); }

}  // namespace

// n16: number of 16-byte vectors to write (the buffer holds 16 * n16 bytes).
extern "C" __global__ void __launch_bounds__(256)
gen_symbols(uint4* __restrict__ out, int64_t n16, uint64_t seed, uint32_t lo, uint32_t span)
{
    const uint32_t lo4 = lo * 0x01010101u;
    const uint32_t mask4 = (span - 1u) * 0x01010101u;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n16; v += stride) {
        const uint64_t a = splitmix64(seed ^ ((uint64_t)(2 * v) * 0x9E3779B97F4A7C15ull));
        const uint64_t b = splitmix64(seed ^ ((uint64_t)(2 * v + 1) * 0x9E3779B97F4A7C15ull));
        out[v] = make_uint4(sym4((uint32_t)a, lo4, mask4), sym4((uint32_t)(a >> 32), lo4, mask4),
                            sym4((uint32_t)b, lo4, mask4), sym4((uint32_t)(b >> 32), lo4, mask4));
    }
}

// Copies n bytes with the shader path instead of a copy engine, so a short
// copy never queues behind long ones on the engines. dst and src are 16-byte
// aligned; src may be pinned host memory mapped into the device's address
// space. Grid-stride over 16-byte vectors, then the tail bytes.
extern "C" __global__ void __launch_bounds__(256)
copy_bytes(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int64_t n)
{
    const int64_t n16 = n >> 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (int64_t v = t; v < n16; v += stride) d4[v] = s4[v];
    if (t < (n & 15)) dst[(n16 << 4) + t] = src[(n16 << 4) + t];
}
