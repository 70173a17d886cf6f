// table.hip -- the trained model's PPF hash table on gfx950.
//
// trainModel (ppf.cpp:105-146) pushes (i, alpha_ij) into bucket key_ij for
// every ordered pair i != j of the sampled model, under an OpenMP critical
// section.  Here every pair is one thread: its key and alpha land in a
// record array in (i, j) order, invalid records are compacted out, and a
// stable radix sort by key groups the buckets while keeping the
// single-thread insertion order (i ascending, j ascending) inside each one.
#include "ppf_dev.h"

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>

namespace ppfd {

__global__ void frames_kernel(int n, const float *px, const float *py, const float *pz,
                              const float *nx, const float *ny, const float *nz, float *frames) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float rt[12];
    ppfc::frame(px[i], py[i], pz[i], nx[i], ny[i], nz[i], rt);
    for (int k = 0; k < 12; k++)
        frames[(size_t)i * 12 + k] = rt[k];
}

struct PairArgs {
    int          n;
    const float *px, *py, *pz, *nx, *ny, *nz;
    const float *frames;
    float        angleStep, distStep;
    int          tailStart;  // partners j >= tailStart use the scalar-tail formula
    uint32_t    *keys;
    float       *alphas;
    char        *flags;
};

// record t = i * n + j
__global__ void __launch_bounds__(256) pairs_kernel(PairArgs a) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long nn = (long long)a.n * a.n;
    if (t >= nn)
        return;
    int i = (int)(t / a.n), j = (int)(t % a.n);
    float n1x = a.nx[i], n1y = a.ny[i], n1z = a.nz[i];
    bool  ok  = (i != j) && isfinite(n1x) && isfinite(n1y) && isfinite(n1z);
    float al  = 0.0f;
    uint32_t key = 0;
    if (ok) {
        const float *rt = a.frames + (size_t)i * 12;
        float        r[12];
#pragma unroll
        for (int k = 0; k < 12; k++)
            r[k] = rt[k];
        al = ppfc::alpha_of(r, a.px[j], a.py[j], a.pz[j]);
        ok = !ppfc::is_nan(al);
        if (ok)
            key = ppfc::ppf_key(a.px[i], a.py[i], a.pz[i], n1x, n1y, n1z, a.px[j], a.py[j],
                                a.pz[j], a.nx[j], a.ny[j], a.nz[j], a.angleStep, a.distStep,
                                j >= a.tailStart);
    }
    a.keys[t]   = key;
    a.alphas[t] = al;
    a.flags[t]  = ok ? 1 : 0;
}

__global__ void gather_keys_kernel(const int *idx, const uint32_t *keys, int m, uint32_t *out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m)
        out[e] = keys[idx[e]];
}

__global__ void finish_entries_kernel(const int *idxSorted, const float *alphas, int m, int n,
                                      uint32_t *refs, float *al) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m)
        return;
    int t   = idxSorted[e];
    refs[e] = (uint32_t)(t / n);
    al[e]   = alphas[t];
}

__global__ void iota_ll_kernel(int *o, long long m) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m)
        o[t] = (int)t;
}

void train_table(TrainScratch &ts, int n, const float *px, const float *py, const float *pz,
                 const float *nx, const float *ny, const float *nz, float angleStep,
                 float distStep, int refTail, std::vector<uint32_t> &keys, std::vector<int> &offs,
                 std::vector<uint32_t> &refs, std::vector<float> &alphas, hipStream_t s) {
    keys.clear();
    refs.clear();
    alphas.clear();
    offs.assign(1, 0);
    if (n <= 1)
        return;
    long long nn = (long long)n * n;
    if (nn > (1ll << 31) - 1)
        throw std::range_error("sampled model too large for the pair table");
    float *frames = ts.tmp.as<float>((size_t)n * 12);
    frames_kernel<<<div_up(n, 256), 256, 0, s>>>(n, px, py, pz, nx, ny, nz, frames);
    PairArgs a;
    a.n  = n;
    a.px = px;
    a.py = py;
    a.pz = pz;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.frames    = frames;
    a.angleStep = angleStep;
    a.distStep  = distStep;
    a.tailStart = refTail ? n - n % 8 : n;
    a.keys   = ts.keys.as<uint32_t>(nn);
    a.alphas = ts.alphas.as<float>(nn);
    a.flags  = ts.flags.as<char>(nn);
    pairs_kernel<<<div_up(nn, 256), 256, 0, s>>>(a);

    int *iota = ts.idx2.as<int>(nn);
    iota_ll_kernel<<<div_up(nn, 256), 256, 0, s>>>(iota, nn);
    int *sel  = ts.idx.as<int>(nn);
    int *cntD = ts.cnt.as<int>(4);
    DevBuf cubTmp;
    size_t tb = 0;
    PPF_HIP_CHECK(rocprim::select(nullptr, tb, iota, a.flags, sel, cntD, (int)nn, s));
    PPF_HIP_CHECK(rocprim::select(cubTmp.reserve(tb), tb, iota, a.flags, sel, cntD, (int)nn, s));
    int m = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(&m, cntD, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    if (m == 0)
        return;
    uint32_t *k1 = ts.keys2.as<uint32_t>(m);
    gather_keys_kernel<<<div_up(m, 256), 256, 0, s>>>(sel, a.keys, m, k1);
    // reuse the (now free) key array as the sorted-key output
    uint32_t *k2   = a.keys;
    int      *idx2 = iota;
    DevBuf    cubTmp2;
    tb = 0;
    PPF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, k1, k2, sel, idx2, m, 0, 32, s));
    PPF_HIP_CHECK(
        rocprim::radix_sort_pairs(cubTmp2.reserve(tb), tb, k1, k2, sel, idx2, m, 0, 32, s));
    uint32_t *uk  = ts.ukeys.as<uint32_t>(m);
    int      *uc  = ts.ucount.as<int>(m);
    int      *nrD = ts.cnt.as<int>(4);
    DevBuf    cubTmp3;
    tb = 0;
    PPF_HIP_CHECK(rocprim::run_length_encode(nullptr, tb, k2, m, uk, uc, nrD, s));
    PPF_HIP_CHECK(
        rocprim::run_length_encode(cubTmp3.reserve(tb), tb, k2, m, uk, uc, nrD, s));
    uint32_t *rf = ts.refs.as<uint32_t>(m);
    float    *al = reinterpret_cast<float *>(k1);  // k1 no longer needed
    finish_entries_kernel<<<div_up(m, 256), 256, 0, s>>>(idx2, a.alphas, m, n, rf, al);
    int K = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(&K, nrD, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    keys.resize(K);
    std::vector<int> cnt(K);
    refs.resize(m);
    alphas.resize(m);
    PPF_HIP_CHECK(hipMemcpyAsync(keys.data(), uk, sizeof(uint32_t) * K, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(cnt.data(), uc, sizeof(int) * K, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(refs.data(), rf, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(alphas.data(), al, sizeof(float) * m, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    offs.resize(K + 1);
    offs[0] = 0;
    for (int k = 0; k < K; k++)
        offs[k + 1] = offs[k] + cnt[k];
}

// ---------------------------------------------------------------- pair features (tests)

__global__ void pair_features_kernel(float p1x, float p1y, float p1z, float n1x, float n1y,
                                     float n1z, int n, const float *px, const float *py,
                                     const float *pz, const float *nx, const float *ny,
                                     const float *nz, float angleStep, float distStep, int tail,
                                     uint32_t *keys, float *alphas) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    float rt[12];
    ppfc::frame(p1x, p1y, p1z, n1x, n1y, n1z, rt);
    alphas[j] = ppfc::alpha_of(rt, px[j], py[j], pz[j]);
    keys[j]   = ppfc::ppf_key(p1x, p1y, p1z, n1x, n1y, n1z, px[j], py[j], pz[j], nx[j], ny[j],
                              nz[j], angleStep, distStep, tail != 0);
}

void pair_features(const float *q, int n, const float *px, const float *py, const float *pz,
                   const float *nx, const float *ny, const float *nz, float angleStep,
                   float distStep, int tail, uint32_t *keys, float *alphas, hipStream_t s) {
    if (n <= 0)
        return;
    pair_features_kernel<<<div_up(n, 256), 256, 0, s>>>(q[0], q[1], q[2], q[3], q[4], q[5], n, px,
                                                         py, pz, nx, ny, nz, angleStep, distStep,
                                                         tail, keys, alphas);
    PPF_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- upload

static void upload_soa(DevBuf &bx, DevBuf &by, DevBuf &bz, const float *aos, int n,
                       hipStream_t s) {
    std::vector<float> x(n), y(n), z(n);
    for (int i = 0; i < n; i++) {
        x[i] = aos[3 * i];
        y[i] = aos[3 * i + 1];
        z[i] = aos[3 * i + 2];
    }
    int m = std::max(n, 1);
    PPF_HIP_CHECK(hipMemcpyAsync(bx.as<float>(m), x.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemcpyAsync(by.as<float>(m), y.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemcpyAsync(bz.as<float>(m), z.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
}

// Pre-filters for the pair stage.  A scene pair's key is murmur(q1, q2, q3,
// q4) of its quantized features; angles lie in [0, pi] and the distance of
// an in-sphere pair below the model diameter, so every finite quantized
// value of a scene pair is at most Qa = ceil(pi / angleStep) (angles) or
// Qd = ceil(1.001 D / distStep) (distance).  That tuple space is small, so
// every tuple is hashed here and those whose key is in the table are
// projected onto (q4), (q3, q4) and (q1, q3, q4).  A pair whose partial tuple
// is in none of them cannot hit the table, and the pair kernel drops it
// before computing the remaining features.  Values outside the ranges (the
// NaN code of a degenerate pair) map to an index that always passes.
void build_key_filter(DevModel &m, const std::vector<uint32_t> &keys, hipStream_t s) {
    m.filterA = m.filterD = 0;
    if (keys.empty())
        return;
    const int Qa = (int)std::ceil(ppfc::kPi / m.angleStep) + 1;
    const int Qd = (int)std::ceil(m.diameter * 1.001f / m.distStep) + 1;
    const int A = Qa + 2, Dd = Qd + 2;  // values 0 .. Q, then "other"
    // the pair kernel packs quantized values in 10 bits
    if (Qa > 1000 || Qd > 1000 || (double)(Qa + 1) * (Qa + 1) * (Qa + 1) * (Qd + 1) > 3.4e7)
        return;
    std::vector<uint32_t> sorted(keys);
    std::sort(sorted.begin(), sorted.end());
    const size_t n4 = Dd, n34 = (size_t)A * Dd, n134 = (size_t)A * A * Dd;
    const size_t w4 = (n4 + 31) / 32, w34 = (n34 + 31) / 32, w134 = (n134 + 31) / 32;
    std::vector<uint32_t> bits(w4 + w34 + w134, 0u);
    auto set = [&](size_t base, size_t i) { bits[base + i / 32] |= 1u << (i % 32); };
    for (int q1 = 0; q1 <= Qa; q1++)
        for (int q2 = 0; q2 <= Qa; q2++)
            for (int q3 = 0; q3 <= Qa; q3++)
                for (int q4 = 0; q4 <= Qd; q4++) {
                    const uint32_t k = ppfc::murmur_key4((uint32_t)q1, (uint32_t)q2, (uint32_t)q3,
                                                         (uint32_t)q4);
                    if (!std::binary_search(sorted.begin(), sorted.end(), k))
                        continue;
                    set(0, (size_t)q4);
                    set(w4, (size_t)q3 * Dd + q4);
                    set(w4 + w34, ((size_t)q1 * A + q3) * Dd + q4);
                }
    // any "other" index passes
    set(0, (size_t)(Dd - 1));
    for (int q3 = 0; q3 < A; q3++)
        for (int q4 = 0; q4 < Dd; q4++) {
            if (q3 == A - 1 || q4 == Dd - 1)
                set(w4, (size_t)q3 * Dd + q4);
            for (int q1 = 0; q1 < A; q1++)
                if (q1 == A - 1 || q3 == A - 1 || q4 == Dd - 1)
                    set(w4 + w34, ((size_t)q1 * A + q3) * Dd + q4);
        }
    PPF_HIP_CHECK(hipMemcpyAsync(m.keyFilter.as<uint32_t>(bits.size()), bits.data(),
                                 sizeof(uint32_t) * bits.size(), hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    m.filterA = A;
    m.filterD = Dd;
}

void model_upload(DevModel &m, int nSampled, const float *sp, const float *sn, int nRe,
                  const float *rp, const float *rn, const std::vector<uint32_t> &keys,
                  const std::vector<int> &offs, const std::vector<uint32_t> &refs,
                  const std::vector<float> &alphas, int angleNum, float diameter,
                  float distStepRel, hipStream_t s) {
    m.nSampled   = nSampled;
    m.nReSampled = nRe;
    m.angleNum   = angleNum;
    m.diameter   = diameter;
    m.angleStep  = ppfc::kTwoPi / (float)angleNum;       // ppf.cpp:184
    m.distStep   = diameter * distStepRel;               // ppf.cpp:185
    upload_soa(m.sx, m.sy, m.sz, sp, nSampled, s);
    upload_soa(m.snx, m.sny, m.snz, sn, nSampled, s);
    upload_soa(m.rx, m.ry, m.rz, rp, nRe, s);
    upload_soa(m.rnx, m.rny, m.rnz, rn, nRe, s);
    float *fr = m.frames.as<float>((size_t)std::max(nSampled, 1) * 12);
    if (nSampled > 0)
        frames_kernel<<<div_up(nSampled, 256), 256, 0, s>>>(
            nSampled, m.sx.get<float>(), m.sy.get<float>(), m.sz.get<float>(), m.snx.get<float>(),
            m.sny.get<float>(), m.snz.get<float>(), fr);
    // entries
    int K = (int)keys.size(), E = (int)refs.size();
    m.nKeys    = K;
    m.nEntries = E;
    std::vector<uint2> ent(E);
    for (int e = 0; e < E; e++) {
        uint32_t ab;
        memcpy(&ab, &alphas[e], 4);
        ent[e] = make_uint2(refs[e], ab);
    }
    PPF_HIP_CHECK(hipMemcpyAsync(m.entries.as<uint2>(std::max(E, 1)), ent.data(), sizeof(uint2) * E,
                                 hipMemcpyHostToDevice, s));
    // the voting copy: refs premultiplied by angleNum, every bucket starting
    // on an 8-entry (64-byte) boundary so a segment of 8 entries is four
    // aligned 16-byte loads; padding entries name the spare accumulator row
    // past the last reference (voted into, never read)
    std::vector<uint32_t> vstart(std::max(K, 1) + 1, 0);
    for (int k = 0; k < K; k++)
        vstart[k + 1] = vstart[k] + (uint32_t)((offs[k + 1] - offs[k] + 7) & ~7);
    const size_t      VE = std::max<size_t>(vstart[K], 8);
    if (VE >= (size_t(1) << 31))  // the voting kernels index entries with 32-bit offsets
        throw std::runtime_error("model table too large for the voting kernels");
    std::vector<uint2> vent(VE, make_uint2((uint32_t)nSampled * (uint32_t)angleNum, 0u));
    for (int k = 0; k < K; k++)
        for (int e = offs[k]; e < offs[k + 1]; e++)
            vent[vstart[k] + (e - offs[k])] = make_uint2(refs[e] * (uint32_t)angleNum, ent[e].y);
    PPF_HIP_CHECK(hipMemcpyAsync(m.ventries.as<uint2>(VE), vent.data(), sizeof(uint2) * VE,
                                 hipMemcpyHostToDevice, s));
    // open-addressing key table, load factor <= 1/2; an empty slot has len 0
    uint32_t T = 64;
    while (T < 2u * (uint32_t)std::max(K, 1))
        T <<= 1;
    std::vector<uint4> slots(T, make_uint4(0, 0, 0, 0));
    for (int k = 0; k < K; k++) {
        uint32_t h = keys[k] & (T - 1);
        while (slots[h].z != 0)
            h = (h + 1) & (T - 1);
        slots[h] = make_uint4(keys[k], vstart[k], (uint32_t)(offs[k + 1] - offs[k]), (uint32_t)k);
    }
    m.tmask = T - 1;
    // buckets by index, and whether every bucket lists its references in
    // ascending order (true for tables built here: insertion order is i
    // ascending; a loaded table may differ)
    std::vector<uint2> bk(std::max(K, 1));
    m.refsSorted = 1;
    m.maxMult    = 1;
    std::vector<uint32_t> seen((size_t)std::max(nSampled, 1), 0u), mult((size_t)std::max(nSampled, 1), 0u);
    for (int k = 0; k < K; k++) {
        bk[k] = make_uint2(vstart[k], (uint32_t)(offs[k + 1] - offs[k]));
        for (int e = offs[k] + 1; e < offs[k + 1]; e++)
            if (refs[e] < refs[e - 1])
                m.refsSorted = 0;
        for (int e = offs[k]; e < offs[k + 1]; e++) {
            const uint32_t r = refs[e] < (uint32_t)nSampled ? refs[e] : 0u;
            if (seen[r] != (uint32_t)k + 1u) {
                seen[r] = (uint32_t)k + 1u;
                mult[r] = 0u;
            }
            m.maxMult = std::max(m.maxMult, ++mult[r]);
        }
    }
    PPF_HIP_CHECK(hipMemcpyAsync(m.buckets.as<uint2>(bk.size()), bk.data(), sizeof(uint2) * bk.size(),
                                 hipMemcpyHostToDevice, s));
    m.generation++;
    PPF_HIP_CHECK(hipMemcpyAsync(m.slots.as<uint4>(T), slots.data(), sizeof(uint4) * T,
                                 hipMemcpyHostToDevice, s));
    build_key_filter(m, keys, s);
    // pose angle of each bin: (float)((double)(fl(2pi) j / maxIdx) - pi), ppf.cpp:402
    std::vector<float> cs(2 * angleNum);
    for (int j = 0; j < angleNum; j++) {
        float  a0 = ppfc::kTwoPi * (float)j / (float)(angleNum - 1);
        float  a  = (float)((double)a0 - 3.14159265358979323846);
        cs[2 * j]     = (float)std::cos((double)a);
        cs[2 * j + 1] = (float)std::sin((double)a);
    }
    PPF_HIP_CHECK(hipMemcpyAsync(m.cosSin.as<float>(2 * angleNum), cs.data(),
                                 sizeof(float) * 2 * angleNum, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace ppfd
