// Kernels of the output stage: event build with crop and readout, key passes, a stable LSD radix sort,
// gather with the window cut, and the C4 columns.
#include <hip/hip_runtime.h>

#include "output_kernels.h"

namespace bsx {

// Exclusive scan of int32 counts into int64 offsets (defined with the scene kernels).
size_t scan_scratch_count(size_t n);
hipError_t launch_exclusive_scan(const int32_t* counts, size_t n, int64_t* out, int64_t* scratch, hipStream_t stream);

namespace {

// S1 class of a sensor label's two class bits (signal 0, noise 2, leak 3, hot 4; reference: label_s1,
// src/sensor/output_order.h:23-26).
__device__ __forceinline__ uint8_t label_s1(uint8_t c) {
  const uint8_t c3 = c & 3;
  return c3 == 0 ? 0 : static_cast<uint8_t>(c3 + 1);
}

constexpr uint64_t kBias = uint64_t{1} << 63;
constexpr int kOutBlock = 256;  // threads per block of the per-event kernels

__global__ void build_kernel(const RawEventDev* raw, size_t n, uint32_t width, Readout r, int crop, CropBox box,
                             OutEventDev* out, unsigned long long* scratch) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool keep = false;
  OutEventDev o{};
  if (i < n) {
    const RawEventDev e = raw[i];
    const int32_t x = static_cast<int32_t>(e.pixel % width), y = static_cast<int32_t>(e.pixel / width);
    keep = !crop || (x >= box.x0 && x < box.x1 && y >= box.y0 && y < box.y1);
    o.t_cross = e.t_cross;
    o.t_reported = reported_time(r, e.t_cross);
    o.x = static_cast<uint16_t>(x);
    o.y = static_cast<uint16_t>(y);
    o.flags = e.flags;
    o.p = e.polarity;
    o.label = label_s1(e.label);
  }
  // block-aggregated slots, one atomic per block (the order of the kept events does not matter: the
  // output key is unique)
  __shared__ unsigned int wave_kept[kOutBlock / 64];
  __shared__ unsigned long long block_base;
  const unsigned long long mask = __ballot(keep);
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) wave_kept[wave] = static_cast<unsigned int>(__popcll(mask));
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long total = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) total += wave_kept[w];
    block_base = total != 0 ? atomicAdd(scratch, total) : 0;
  }
  __syncthreads();
  unsigned long long base = block_base;
  for (int w = 0; w < wave; ++w) base += wave_kept[w];
  if (keep) out[base + __popcll(mask & ((1ull << lane) - 1ull))] = o;
}

// The extremes of t_cross and t_reported (biased to unsigned order) over ev[0 .. n_held + scratch[0]):
// a grid-stride loop over a bounded grid, reduced in the wave and then in the block, so that each block
// makes one set of four atomics.
__global__ void minmax_kernel(const OutEventDev* ev, size_t n_held, size_t n_max, unsigned long long* scratch) {
  __shared__ unsigned long long part[4][kOutBlock / 64];
  const size_t n = vmin(n_held + static_cast<size_t>(scratch[0]), n_max);
  unsigned long long a = ~0ull, b = 0ull, c = ~0ull, d = 0ull;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const unsigned long long tc = static_cast<uint64_t>(ev[i].t_cross) ^ kBias;
    const unsigned long long tr = static_cast<uint64_t>(ev[i].t_reported) ^ kBias;
    a = min(a, tc);
    b = max(b, tc);
    c = min(c, tr);
    d = max(d, tr);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a = min(a, static_cast<unsigned long long>(__shfl_xor(a, off)));
    b = max(b, static_cast<unsigned long long>(__shfl_xor(b, off)));
    c = min(c, static_cast<unsigned long long>(__shfl_xor(c, off)));
    d = max(d, static_cast<unsigned long long>(__shfl_xor(d, off)));
  }
  const int wave = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    part[0][wave] = a;
    part[1][wave] = b;
    part[2][wave] = c;
    part[3][wave] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
      a = min(a, part[0][w]);
      b = max(b, part[1][w]);
      c = min(c, part[2][w]);
      d = max(d, part[3][w]);
    }
    if (a != ~0ull) {
      atomicMin(scratch + 1, a);
      atomicMax(scratch + 2, b);
      atomicMin(scratch + 3, c);
      atomicMax(scratch + 4, d);
    }
  }
}

__global__ void keys_kernel(const OutEventDev* ev, const uint32_t* perm, uint32_t* perm_init, size_t n, int pass,
                            int y_bits, int64_t tc_min, int64_t tr_min, uint64_t* keys) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t e = static_cast<uint32_t>(i);
  if (perm != nullptr) e = perm[i];
  else if (perm_init != nullptr) perm_init[i] = e;
  const OutEventDev o = ev[e];
  keys[i] = pass == 0   ? static_cast<uint64_t>(o.t_cross - tc_min)
            : pass == 1 ? (static_cast<uint64_t>(o.x) << y_bits) | o.y
                        : static_cast<uint64_t>(o.t_reported - tr_min);
}

__global__ void gather_kernel(const OutEventDev* ev, const uint32_t* perm, size_t n, int64_t cut, OutEventDev* sorted,
                              unsigned long long* scratch) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool ready = false;
  if (i < n) {
    const OutEventDev o = ev[perm[i]];
    sorted[i] = o;
    ready = o.t_reported < cut;
  }
  __shared__ unsigned int wave_ready[kOutBlock / 64];
  const unsigned long long m = __ballot(ready);
  if (__lane_id() == 0) wave_ready[threadIdx.x >> 6] = static_cast<unsigned int>(__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long total = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) total += wave_ready[w];
    if (total != 0) atomicAdd(scratch + 5, total);
  }
}

__global__ void columns_kernel(const OutEventDev* ev, size_t n, int64_t t_offset_us, uint16_t* x, uint16_t* y,
                               uint8_t* p, int64_t* t_us, uint16_t* t_ns_rem, int64_t* t_rep, int64_t* t_cross,
                               uint8_t* label, uint16_t* flags) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const OutEventDev o = ev[i];
  const int64_t us = floor_div64(o.t_reported, 1000);
  x[i] = o.x;
  y[i] = o.y;
  p[i] = o.p;
  t_us[i] = us - t_offset_us;
  t_ns_rem[i] = static_cast<uint16_t>(o.t_reported - 1000 * us);
  t_rep[i] = o.t_reported;
  t_cross[i] = o.t_cross;
  label[i] = o.label;
  flags[i] = o.flags;
}

// The compact 16-byte records of ev[0 .. n) (one thread per event; 16-byte stores).
__global__ void pack_kernel(const OutEventDev* ev, size_t n, StreamEventDev* dst) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const OutEventDev o = ev[i];
  dst[i] = StreamEventDev{o.t_reported, o.x, o.y, o.p, o.label, o.flags};
}

// The radix sort: 8-bit digits, least significant first. A block owns a tile of kSortTile consecutive
// items (kSortItems per thread, item k * 256 + thread of the tile at step k, so that the block walks its
// tile in order). Per digit: every tile's digit histogram, written digit-major (hist[d * tiles + tile]);
// the exclusive scan of that array, which gives each (digit, tile) its first output slot with every
// smaller digit, and every earlier tile's equal digits, before it; then the scatter, in which each tile
// ranks its items stably (the wave's equal digits by an 8-ballot match, the earlier waves' through LDS
// counts) and writes them from its slots on.
constexpr int kSortBits = 8;
constexpr int kSortBins = 1 << kSortBits;
constexpr int kSortItems = 8;
constexpr int kSortTile = 256 * kSortItems;

__global__ __launch_bounds__(256) void digit_hist_kernel(const uint64_t* keys, size_t n, int shift, size_t tiles,
                                                         int32_t* hist) {
  __shared__ uint32_t h[kSortBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = static_cast<size_t>(blockIdx.x) * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const size_t i = base + static_cast<size_t>(k) * 256 + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kSortBins - 1)], 1u);
  }
  __syncthreads();
  hist[static_cast<size_t>(threadIdx.x) * tiles + blockIdx.x] = static_cast<int32_t>(h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void digit_scatter_kernel(const uint64_t* keys, const uint32_t* vals, size_t n,
                                                            int shift, size_t tiles, const int64_t* offsets,
                                                            uint64_t* keys_out, uint32_t* vals_out) {
  __shared__ uint32_t base[kSortBins];          // this tile's next free slot of each digit
  __shared__ uint32_t wave_count[4][kSortBins];  // the current step's items of each digit, per wave
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  base[threadIdx.x] = static_cast<uint32_t>(offsets[static_cast<size_t>(threadIdx.x) * tiles + blockIdx.x]);
  for (int w = 0; w < 4; ++w) wave_count[w][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const size_t tile = static_cast<size_t>(blockIdx.x) * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const size_t i = tile + static_cast<size_t>(k) * 256 + threadIdx.x;
    const bool valid = i < n;
    uint64_t key = 0;
    uint32_t val = 0, d = 0;
    if (valid) {
      key = keys[i];
      val = vals[i];
      d = static_cast<uint32_t>((key >> shift) & (kSortBins - 1));
    }
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < kSortBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t set = __ballot(bit);
      peers &= bit ? set : ~set;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & below));
    const bool leader = valid && rank == 0;
    if (leader) wave_count[wave][d] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t slot = base[d] + rank;
      for (int w = 0; w < wave; ++w) slot += wave_count[w][d];
      keys_out[slot] = key;
      vals_out[slot] = val;
    }
    __syncthreads();
    base[threadIdx.x] += wave_count[0][threadIdx.x] + wave_count[1][threadIdx.x] + wave_count[2][threadIdx.x] +
                         wave_count[3][threadIdx.x];
    __syncthreads();
    if (leader) wave_count[wave][d] = 0;
  }
}

inline unsigned grid(size_t n) { return static_cast<unsigned>((n + kOutBlock - 1) / kOutBlock); }

}  // namespace

hipError_t launch_out_build(const RawEventDev* raw, size_t n, uint32_t width, const Readout& r, bool crop,
                            const CropBox& box, OutEventDev* out, unsigned long long* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  build_kernel<<<grid(n), kOutBlock, 0, stream>>>(raw, n, width, r, crop ? 1 : 0, box, out, scratch);
  return hipGetLastError();
}

hipError_t launch_out_minmax(const OutEventDev* ev, size_t n_held, size_t n_max, unsigned long long* scratch,
                             hipStream_t stream) {
  if (n_max == 0) return hipSuccess;
  minmax_kernel<<<vmin(grid(n_max), 1024u), kOutBlock, 0, stream>>>(ev, n_held, n_max, scratch);
  return hipGetLastError();
}

hipError_t launch_out_keys(const OutEventDev* ev, const uint32_t* perm, uint32_t* perm_init, size_t n, int pass,
                           int y_bits, int64_t tc_min, int64_t tr_min, uint64_t* keys, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  keys_kernel<<<grid(n), 256, 0, stream>>>(ev, perm, perm_init, n, pass, y_bits, tc_min, tr_min, keys);
  return hipGetLastError();
}

hipError_t launch_out_sort(uint64_t* keys, uint64_t* keys_out, uint32_t* perm, uint32_t* perm_out, size_t n, int bits,
                           DeviceBuffer& temp, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (n >= (size_t{1} << 32) || bits < 1 || bits > 64) return hipErrorInvalidValue;
  const size_t tiles = (n + kSortTile - 1) / kSortTile;
  const size_t cells = tiles * kSortBins;
  const size_t scan_words = scan_scratch_count(cells);
  uint8_t* t = static_cast<uint8_t*>(temp.ensure(cells * 4 + cells * 8 + scan_words * 8));
  int32_t* hist = reinterpret_cast<int32_t*>(t);
  int64_t* offsets = reinterpret_cast<int64_t*>(t + cells * 4);
  int64_t* scan = reinterpret_cast<int64_t*>(t + cells * 12);
  uint64_t* k_src = keys;
  uint64_t* k_dst = keys_out;
  uint32_t* v_src = perm;
  uint32_t* v_dst = perm_out;
  for (int shift = 0; shift < bits; shift += kSortBits) {
    digit_hist_kernel<<<static_cast<unsigned>(tiles), 256, 0, stream>>>(k_src, n, shift, tiles, hist);
    hipError_t e = launch_exclusive_scan(hist, cells, offsets, scan, stream);
    if (e != hipSuccess) return e;
    digit_scatter_kernel<<<static_cast<unsigned>(tiles), 256, 0, stream>>>(k_src, v_src, n, shift, tiles, offsets,
                                                                           k_dst, v_dst);
    std::swap(k_src, k_dst);
    std::swap(v_src, v_dst);
  }
  if (v_src != perm_out) {  // an even number of digits: the sorted pairs are in the input buffers
    hipError_t e = hipMemcpyAsync(keys_out, keys, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(perm_out, perm, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t launch_out_gather(const OutEventDev* ev, const uint32_t* perm, size_t n, int64_t cut, OutEventDev* sorted,
                             unsigned long long* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  gather_kernel<<<grid(n), kOutBlock, 0, stream>>>(ev, perm, n, cut, sorted, scratch);
  return hipGetLastError();
}

hipError_t launch_out_columns(const OutEventDev* ev, size_t n, int64_t t_offset_us, uint16_t* x, uint16_t* y,
                              uint8_t* p, int64_t* t_us, uint16_t* t_ns_rem, int64_t* t_rep, int64_t* t_cross,
                              uint8_t* label, uint16_t* flags, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  columns_kernel<<<grid(n), 256, 0, stream>>>(ev, n, t_offset_us, x, y, p, t_us, t_ns_rem, t_rep, t_cross, label,
                                              flags);
  return hipGetLastError();
}

hipError_t launch_out_pack(const OutEventDev* ev, size_t n, StreamEventDev* dst, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  pack_kernel<<<grid(n), 256, 0, stream>>>(ev, n, dst);
  return hipGetLastError();
}

}  // namespace bsx
