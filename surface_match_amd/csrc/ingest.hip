// ingest.hip -- a host cloud becomes a device cloud: removeNan
// (util.cpp:86-105) and computeBoundingBox (util.cpp:178-207) on the GPU.
//
// matchScene copies the caller's cloud, keeps the indices of the points
// whose coordinates (and normal, when the cloud has normals) are finite, and
// takes the box of those points when the caller gave none (ppf.cpp:153,
// 170-176).  Here the AoS arrays are copied to HBM once and three kernels
// do the rest in two passes over the points:
//
//   ingest_kernel   one thread per point: AoS -> SoA originals, the
//                   validity bit (a wave ballot, kept per wave), the block's
//                   valid count and the block's six extreme keys;
//   summary_kernel  one workgroup: exclusive scan of the block counts, the
//                   six extremes over all blocks, the box read back from
//                   the winning points;
//   compact_kernel  one thread per point: its rank among the valid points
//                   from the wave masks -> the ascending valid index list.
//
// The box follows the reference's sequential rule exactly: starting from
// the first valid point, a coordinate replaces the maximum only if it is
// greater and the minimum only if it is smaller, so each bound is the
// earliest valid point holding the extreme value (+0 and -0 compare equal;
// the bits of the earliest one are kept).  Keys order (value, index) so that
// a min/max over them picks exactly that point.
#include "ppf_dev.h"

#include <cstring>

namespace ppfd {

namespace {

constexpr int kIngestBlock = 256;
constexpr int kWaves       = kIngestBlock / 64;

__device__ __forceinline__ bool finite3(float a, float b, float c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

// monotone map of a float to uint32 with -0 == +0 (NaN never reaches here)
__device__ __forceinline__ uint32_t order_key(float v) {
    const uint32_t u = __float_as_uint(v == 0.0f ? 0.0f : v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ unsigned long long wave_min(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v                          = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ unsigned long long wave_max(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v                          = w > v ? w : v;
    }
    return v;
}

// keys[b * 6 + a]: block b's min key of axis a (a < 3) = (order_key << 32) |
// index, max key of axis a - 3 = (order_key << 32) | ~index; an empty block
// holds ~0 / 0
template <bool NORMALS>
__global__ void __launch_bounds__(kIngestBlock)
    ingest_kernel(const float *__restrict__ aosP, const float *__restrict__ aosN, int n,
                  float *ox, float *oy, float *oz, float *onx, float *ony, float *onz,
                  unsigned long long *masks, int *blockCnt, unsigned long long *keys) {
    __shared__ unsigned long long sKey[kWaves][6];
    __shared__ int                sCnt[kWaves];
    const int i    = blockIdx.x * kIngestBlock + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool      ok   = false;
    float     p[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        p[0] = aosP[3 * (size_t)i];
        p[1] = aosP[3 * (size_t)i + 1];
        p[2] = aosP[3 * (size_t)i + 2];
        ox[i] = p[0];
        oy[i] = p[1];
        oz[i] = p[2];
        ok    = finite3(p[0], p[1], p[2]);
        if (NORMALS) {
            const float a = aosN[3 * (size_t)i], b = aosN[3 * (size_t)i + 1],
                        c = aosN[3 * (size_t)i + 2];
            onx[i] = a;
            ony[i] = b;
            onz[i] = c;
            ok     = ok && finite3(a, b, c);
        }
    }
    const unsigned long long m = __ballot(ok);
    if (lane == 0) {
        masks[(size_t)blockIdx.x * kWaves + wave] = m;
        sCnt[wave]                                = __popcll(m);
    }
    for (int a = 0; a < 3; a++) {
        const unsigned long long k = ok ? (unsigned long long)order_key(p[a]) << 32 : 0ull;
        const unsigned long long lo = wave_min(ok ? (k | (uint32_t)i) : ~0ull);
        const unsigned long long hi = wave_max(ok ? (k | (uint32_t)~(uint32_t)i) : 0ull);
        if (lane == 0) {
            sKey[wave][a]     = lo;
            sKey[wave][3 + a] = hi;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int          a = threadIdx.x;
        unsigned long long v = sKey[0][a];
        for (int w = 1; w < kWaves; w++)
            v = a < 3 ? (sKey[w][a] < v ? sKey[w][a] : v) : (sKey[w][a] > v ? sKey[w][a] : v);
        keys[(size_t)blockIdx.x * 6 + a] = v;
    }
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < kWaves; w++)
            c += sCnt[w];
        blockCnt[blockIdx.x] = c;
    }
}

// info[0] = valid count; info[1..6] = box bits (min xyz, max xyz) when the
// count is positive
constexpr int kSummaryBlock = 1024;
__global__ void __launch_bounds__(kSummaryBlock)
    summary_kernel(const int *blockCnt, const unsigned long long *keys, int nb,
                   const float *aosP, int *blockOff, uint32_t *info) {
    __shared__ int                sScan[kSummaryBlock];
    __shared__ unsigned long long sKey[kSummaryBlock / 64][6];
    const int                     t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int                           carry = 0;
    for (int base = 0; base < nb; base += kSummaryBlock) {
        const int b = base + t;
        const int v = b < nb ? blockCnt[b] : 0;
        sScan[t]    = v;
        __syncthreads();
        // Hillis-Steele inclusive scan over the chunk
        for (int o = 1; o < kSummaryBlock; o <<= 1) {
            const int add = t >= o ? sScan[t - o] : 0;
            __syncthreads();
            sScan[t] += add;
            __syncthreads();
        }
        if (b < nb)
            blockOff[b] = carry + sScan[t] - v;
        carry += sScan[kSummaryBlock - 1];
        __syncthreads();
    }
    unsigned long long acc[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    for (int b = t; b < nb; b += kSummaryBlock)
        for (int a = 0; a < 6; a++) {
            const unsigned long long k = keys[(size_t)b * 6 + a];
            acc[a] = a < 3 ? (k < acc[a] ? k : acc[a]) : (k > acc[a] ? k : acc[a]);
        }
    for (int a = 0; a < 6; a++) {
        acc[a] = a < 3 ? wave_min(acc[a]) : wave_max(acc[a]);
        if (lane == 0)
            sKey[wave][a] = acc[a];
    }
    __syncthreads();
    if (t == 0) {
        info[0] = (uint32_t)carry;
        for (int a = 0; a < 6; a++) {
            unsigned long long v = sKey[0][a];
            for (int w = 1; w < kSummaryBlock / 64; w++)
                v = a < 3 ? (sKey[w][a] < v ? sKey[w][a] : v) : (sKey[w][a] > v ? sKey[w][a] : v);
            uint32_t idx = (uint32_t)v;
            if (a >= 3)
                idx = ~idx;
            // the winning point's own coordinate (its bits, -0 included)
            info[1 + a] = carry > 0 ? __float_as_uint(aosP[3 * (size_t)idx + (a % 3)]) : 0u;
        }
    }
}

__global__ void __launch_bounds__(kIngestBlock)
    compact_kernel(const unsigned long long *masks, const int *blockOff, int n, int *valid) {
    const int i = blockIdx.x * kIngestBlock + threadIdx.x;
    if (i >= n)
        return;
    const int                lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m    = masks[(size_t)blockIdx.x * kWaves + wave];
    if (!((m >> lane) & 1ull))
        return;
    int before = 0;
    for (int w = 0; w < wave; w++)
        before += __popcll(masks[(size_t)blockIdx.x * kWaves + w]);
    const unsigned long long below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    valid[blockOff[blockIdx.x] + before + __popcll(below)] = i;
}

}  // namespace

IngestScratch::~IngestScratch() {
    if (info)
        (void)hipHostFree(info);
}

int ingest_cloud(IngestScratch &is, const float *aosP, const float *aosN, int n, float *ox,
                 float *oy, float *oz, float *onx, float *ony, float *onz, int *valid,
                 float ext[6], hipStream_t s) {
    if (n <= 0) {
        for (int a = 0; a < 6; a++)
            ext[a] = 0.0f;
        return 0;
    }
    if (!is.info)
        PPF_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&is.info), 8 * sizeof(uint32_t)));
    const int           nb       = div_up(n, kIngestBlock);
    unsigned long long *masks    = is.masks.as<unsigned long long>((size_t)nb * kWaves);
    int                *blockCnt = is.blockCnt.as<int>(nb);
    int                *blockOff = is.blockOff.as<int>(nb);
    unsigned long long *keys     = is.keys.as<unsigned long long>((size_t)nb * 6);
    uint32_t           *info     = is.devInfo.as<uint32_t>(8);
    if (aosN)
        ingest_kernel<true><<<nb, kIngestBlock, 0, s>>>(aosP, aosN, n, ox, oy, oz, onx, ony, onz,
                                                        masks, blockCnt, keys);
    else
        ingest_kernel<false><<<nb, kIngestBlock, 0, s>>>(aosP, nullptr, n, ox, oy, oz, nullptr,
                                                         nullptr, nullptr, masks, blockCnt, keys);
    summary_kernel<<<1, kSummaryBlock, 0, s>>>(blockCnt, keys, nb, aosP, blockOff, info);
    compact_kernel<<<nb, kIngestBlock, 0, s>>>(masks, blockOff, n, valid);
    PPF_HIP_CHECK(hipMemcpyAsync(is.info, info, 7 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    const int nValid = (int)is.info[0];
    for (int a = 0; a < 6; a++) {
        uint32_t u = is.info[1 + a];
        std::memcpy(&ext[a], &u, sizeof u);
    }
    return nValid;
}

}  // namespace ppfd
