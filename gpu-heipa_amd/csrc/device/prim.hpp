// Device primitives for wave64 / CDNA4: wave and block scans, device-wide
// tile scans driven by functors (used for ordered compaction and prefix sums),
// fill kernels, and a stable LSD radix sort of (key, value) pairs.
//
// Fills are kernels, never hipMemsetAsync, so every sequence here stays safe to
// capture in a HIP graph.
#pragma once

#include "ctx.hpp"

namespace heipa {

constexpr int WAVE = 64;

__device__ __forceinline__ u32 lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ u64 lanemask_lt() { return (1ull << lane_id()) - 1ull; }

template<typename T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    const u32 lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T y = __shfl_up(x, o, WAVE);
        if (lane >= (u32) o) x += y;
    }
    return x;
}

template<typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}

template<typename T>
__device__ __forceinline__ T wave_max(T x) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        T y = __shfl_xor(x, o, WAVE);
        x = y > x ? y : x;
    }
    return x;
}

// Block-wide exclusive scan; `lds` needs BLOCK/64 entries.  Returns the
// exclusive prefix of v and the block total in `total`.
template<int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *lds, T &total) {
    constexpr int NW = BLOCK / WAVE;
    const u32 lane = lane_id(), wid = threadIdx.x / WAVE;
    T x = wave_inclusive_scan(v);
    if (lane == WAVE - 1) lds[wid] = x;
    __syncthreads();
    if (wid == 0) {
        T t = lane < (u32) NW ? lds[lane] : T(0);
        t = wave_inclusive_scan(t);
        if (lane < (u32) NW) lds[lane] = t;
    }
    __syncthreads();
    const T wave_prefix = wid > 0 ? lds[wid - 1] : T(0);
    total = lds[NW - 1];
    __syncthreads();
    return x - v + wave_prefix;
}

template<int BLOCK, typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    constexpr int NW = BLOCK / WAVE;
    const u32 lane = lane_id(), wid = threadIdx.x / WAVE;
    T x = wave_sum(v);
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    T t = T(0);
#pragma unroll
    for (int w = 0; w < NW; ++w) t += lds[w];
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------- fills
template<typename T>
__global__ void fill_kernel(T *p, u64 n, T v) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64) gridDim.x * blockDim.x) p[i] = v;
}

template<typename T>
inline void fill(Ctx &c, T *p, u64 n, T v) {
    if (n == 0) return;
    launch(c, "fill", (double) n * sizeof(T), dim3(grid_for(c, n, 256 * 4)), dim3(256), 0, fill_kernel<T>, p, n, v);
}

template<typename T>
__global__ void iota_kernel(T *p, u64 n) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64) gridDim.x * blockDim.x) p[i] = (T) i;
}

// ---------------------------------------------------------------- tile scans
// A scan over n items whose values are produced by `src(i)` and whose
// exclusive prefixes are consumed by `dst(i, prefix)`.  Three launches for
// large n (tile sums, scan of tile sums, tile scan), one for small n.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template<typename T, typename Src>
__global__ __launch_bounds__(SCAN_BLOCK) void scan_tile_sums_kernel(Src src, u64 n, T *tile_sums) {
    __shared__ T lds[SCAN_BLOCK / WAVE];
    const u64 base = (u64) blockIdx.x * SCAN_TILE;
    T s = T(0);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        u64 i = base + (u64) j * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += src(i);
    }
    s = block_sum<SCAN_BLOCK>(s, lds);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

// exclusive scan of tile sums in place by one block; total to *total
template<typename T>
__global__ __launch_bounds__(1024) void scan_tile_prefix_kernel(T *tile_sums, u32 tiles, T *total) {
    __shared__ T lds[1024 / WAVE];
    T carry = T(0);
    for (u32 b = 0; b < tiles; b += 1024) {
        u32 i = b + threadIdx.x;
        T v = i < tiles ? tile_sums[i] : T(0);
        T t;
        T e = block_exclusive_scan<1024>(v, lds, t);
        if (i < tiles) tile_sums[i] = carry + e;
        carry += t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// tile scan with offsets: values are staged through LDS so loads stay
// coalesced while each thread scans a contiguous run of SCAN_ITEMS.
template<typename T, typename Src, typename Dst>
__global__ __launch_bounds__(SCAN_BLOCK) void scan_tiles_kernel(Src src, Dst dst, u64 n, const T *tile_prefix,
                                                                 T init) {
    __shared__ T stage[SCAN_TILE];
    __shared__ T lds[SCAN_BLOCK / WAVE];
    const u64 base = (u64) blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        u64 i = base + (u64) j * SCAN_BLOCK + threadIdx.x;
        stage[j * SCAN_BLOCK + threadIdx.x] = i < n ? src(i) : T(0);
    }
    __syncthreads();
    T vals[SCAN_ITEMS];
    T run = T(0);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        vals[j] = stage[threadIdx.x * SCAN_ITEMS + j];
        run += vals[j];
    }
    T tot;
    T excl = block_exclusive_scan<SCAN_BLOCK>(run, lds, tot);
    T off = (tile_prefix ? tile_prefix[blockIdx.x] : T(0)) + init + excl;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        u64 i = base + (u64) threadIdx.x * SCAN_ITEMS + j;
        if (i < n) dst(i, off);
        off += vals[j];
    }
}

// single-block scan for small n; writes the total to *total
template<typename T, typename Src, typename Dst>
__global__ __launch_bounds__(1024) void scan_small_kernel(Src src, Dst dst, u64 n, T *total, T init) {
    __shared__ T lds[1024 / WAVE];
    T carry = init;
    for (u64 b = 0; b < n; b += 1024) {
        u64 i = b + threadIdx.x;
        T v = i < n ? src(i) : T(0);
        T t;
        T e = block_exclusive_scan<1024>(v, lds, t);
        if (i < n) dst(i, carry + e);
        carry += t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// Device-wide exclusive scan; *total_dev (device pointer, may be null)
// receives the sum of all items.  `init` is added to every prefix (the total
// excludes it).
template<typename T, typename Src, typename Dst>
inline void scan(Ctx &c, const char *name, Src src, Dst dst, u64 n, T *total_dev, T init = T(0),
                 double bytes = 0.0) {
    if (n == 0) {
        if (total_dev) fill<T>(c, total_dev, 1, T(0));
        return;
    }
    if (n <= (u64) 8 * SCAN_TILE) {
        launch(c, name, bytes, dim3(1), dim3(1024), 0, scan_small_kernel<T, Src, Dst>, src, dst, n, total_dev, init);
        return;
    }
    BackScope bs(c.arena);
    const u32 tiles = div_up(n, SCAN_TILE);
    T *sums = bs.get<T>(tiles);
    launch(c, name, bytes * 0.5, dim3(tiles), dim3(SCAN_BLOCK), 0, scan_tile_sums_kernel<T, Src>, src, n, sums);
    launch(c, name, 0.0, dim3(1), dim3(1024), 0, scan_tile_prefix_kernel<T>, sums, tiles, total_dev);
    launch(c, name, bytes * 0.5, dim3(tiles), dim3(SCAN_BLOCK), 0, scan_tiles_kernel<T, Src, Dst>, src, dst, n,
           (const T *) sums, init);
}

// functors for common scans
template<typename T>
struct LoadArr {
    const T *p;
    __device__ T operator()(u64 i) const { return p[i]; }
};
template<typename T>
struct StoreArr {
    T *p;
    __device__ void operator()(u64 i, T v) const { p[i] = v; }
};

// exclusive prefix sum out[i] = sum_{j<i} in[i] for i in [0, n]; out has n+1
// entries (the last is the total); in and out may alias.
template<typename T>
struct LoadArrPad {
    const T *p;
    u64 n;
    __device__ T operator()(u64 i) const { return i < n ? p[i] : T(0); }
};
template<typename T>
inline void exclusive_sum(Ctx &c, const char *name, const T *in, T *out, u64 n) {
    scan<T>(c, name, LoadArrPad<T>{in, n}, StoreArr<T>{out}, n + 1, (T *) nullptr, T(0),
            (double) (2 * n + 1) * sizeof(T));
}

// ---------------------------------------------------------------- radix sort
// Stable LSD radix sort of (key, value) pairs, 8 bits per pass over the low
// `bits` bits of the key.  Keys K = u32 or u64, values u32.  Ping-pongs
// between (keys, vals) and (keys_tmp, vals_tmp); the result ends in the
// returned buffers.
constexpr int RS_BLOCK = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;  // 4096 pairs per tile

template<typename K>
__global__ __launch_bounds__(RS_BLOCK) void radix_hist_kernel(const K *keys, u64 n, u32 shift, u32 tiles,
                                                               u32 *counts) {
    __shared__ u32 hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const u64 base = (u64) blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ROUNDS; ++r) {
        u64 i = base + (u64) r * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&hist[(u32) (keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(u64) threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

template<typename K>
__global__ __launch_bounds__(RS_BLOCK) void radix_scatter_kernel(const K *keys, const u32 *vals, K *okeys, u32 *ovals,
                                                                  u64 n, u32 shift, u32 tiles, const u32 *offsets) {
    constexpr int NW = RS_BLOCK / WAVE;
    __shared__ u32 running[256];
    __shared__ u32 wcount[NW][256];
    __shared__ u32 goff[256];
    goff[threadIdx.x] = offsets[(u64) threadIdx.x * tiles + blockIdx.x];
    running[threadIdx.x] = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) wcount[w][threadIdx.x] = 0;
    __syncthreads();
    const u64 base = (u64) blockIdx.x * RS_TILE;
    const u32 wid = threadIdx.x / WAVE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const u64 i = base + (u64) r * RS_BLOCK + threadIdx.x;
        const bool valid = i < n;
        K key = valid ? keys[i] : K(0);
        u32 val = valid ? vals[i] : 0u;
        const u32 d = (u32) (key >> shift) & 255u;
        // lanes of this wave holding the same digit
        u64 mask = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const u64 bal = __ballot(bit);
            mask &= bit ? bal : ~bal;
        }
        const u32 rank = (u32) __popcll(mask & lanemask_lt());
        const bool leader = valid && (mask & lanemask_lt()) == 0;
        if (leader) wcount[wid][d] = (u32) __popcll(mask);
        __syncthreads();
        if (valid) {
            u32 pos = running[d] + rank;
            for (u32 w = 0; w < wid; ++w) pos += wcount[w][d];
            okeys[goff[d] + pos] = key;
            ovals[goff[d] + pos] = val;
        }
        __syncthreads();
        {
            u32 t = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                t += wcount[w][threadIdx.x];
                wcount[w][threadIdx.x] = 0;
            }
            running[threadIdx.x] += t;
        }
        __syncthreads();
    }
}

// Sorts n pairs by the low `bits` bits of the keys.  Returns true if the result
// is in (keys_tmp, vals_tmp), false if in (keys, vals).
template<typename K>
inline bool radix_sort_pairs(Ctx &c, K *keys, u32 *vals, K *keys_tmp, u32 *vals_tmp, u64 n, u32 bits) {
    if (n <= 1 || bits == 0) return false;
    BackScope bs(c.arena);
    const u32 tiles = div_up(n, RS_TILE);
    u32 *counts = bs.get<u32>((u64) 256 * tiles + 1);
    bool flip = false;
    for (u32 shift = 0; shift < bits; shift += 8) {
        const K *ik = flip ? keys_tmp : keys;
        const u32 *iv = flip ? vals_tmp : vals;
        K *ok = flip ? keys : keys_tmp;
        u32 *ov = flip ? vals : vals_tmp;
        launch(c, "radix_hist", (double) n * sizeof(K), dim3(tiles), dim3(RS_BLOCK), 0, radix_hist_kernel<K>, ik, n,
               shift, tiles, counts);
        exclusive_sum<u32>(c, "radix_offsets", counts, counts, (u64) 256 * tiles);
        launch(c, "radix_scatter", (double) n * 2 * (sizeof(K) + 4), dim3(tiles), dim3(RS_BLOCK), 0,
               radix_scatter_kernel<K>, ik, iv, ok, ov, n, shift, tiles, (const u32 *) counts);
        flip = !flip;
    }
    return flip;
}

inline u32 bits_for(u64 max_value) {
    u32 b = 0;
    while (b < 64 && (max_value >> b) != 0) ++b;
    return b;
}

} // namespace heipa
