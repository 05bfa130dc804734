// Device-wide scan and stable radix sort used by the tile binner.
//
// Both work on element counts that live in device memory (the binner only
// learns how many (tile, quad-step) pairs a frame has on the device), so the
// grids are sized for a capacity and blocks past the live count exit at once.
// Nothing here synchronizes with the host.
#include "primitives.h"

namespace rf {

static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_ITEMS = 16;
static constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;  // 4096

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of one block's SCAN_THREADS values; returns the block total
// through *total (all threads).
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* total) {
    __shared__ uint64_t s_wave[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t incl = wave_incl_scan_u64(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        uint64_t t = s_wave[w];
        if (w < wave) base += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return base + incl - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                               const uint64_t* n_ptr, uint64_t n_fixed,
                                                               uint64_t* __restrict__ block_sums) {
    const uint64_t n = n_ptr ? *n_ptr : n_fixed;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    uint64_t s = 0;
    if (base + SCAN_CHUNK <= n) {
        // whole chunk: 16-byte loads, each wave covering 1 KiB per step
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS / 4; i++) {
            const uint4 q = reinterpret_cast<const uint4*>(in + base)[i * SCAN_THREADS + threadIdx.x];
            s += (uint64_t)q.x + q.y + q.z + q.w;
        }
    } else if (base < n) {
        for (int i = 0; i < SCAN_ITEMS; i++) {
            uint64_t idx = base + (uint64_t)i * SCAN_THREADS + threadIdx.x;
            if (idx < n) s += in[idx];
        }
    }
    uint64_t total;
    block_excl_scan_u64(s, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block: exclusive scan of the block sums in place; writes the grand
// total to *total_out.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(uint64_t* __restrict__ sums, int nblocks,
                                                             uint64_t* __restrict__ total_out) {
    // each thread scans SCAN_ITEMS consecutive sums serially, then the
    // block scans the per-thread totals: one pass per SCAN_CHUNK sums
    uint64_t carry = 0;
    for (int start = 0; start < nblocks; start += SCAN_CHUNK) {
        const int first = start + threadIdx.x * SCAN_ITEMS;
        uint64_t v[SCAN_ITEMS];
        uint64_t t = 0;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            v[i] = first + i < nblocks ? sums[first + i] : 0;
            t += v[i];
        }
        uint64_t total;
        uint64_t run = carry + block_excl_scan_u64(t, &total);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            if (first + i < nblocks) sums[first + i] = run;
            run += v[i];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t* __restrict__ in,
                                                              const uint64_t* n_ptr, uint64_t n_fixed,
                                                              const uint64_t* __restrict__ block_offs,
                                                              uint64_t* __restrict__ out) {
    const uint64_t n = n_ptr ? *n_ptr : n_fixed;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    if (base >= n) return;
    // each thread owns SCAN_ITEMS consecutive elements, read and written as
    // 16-byte vectors where the whole run lies below n (both arrays are
    // 16-byte aligned, and so is every run: SCAN_ITEMS is a multiple of 4)
    const uint64_t first = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
    const bool whole = first + SCAN_ITEMS <= n;
    uint32_t v[SCAN_ITEMS];
    uint64_t s = 0;
    if (whole) {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS / 4; i++) {
            const uint4 q = reinterpret_cast<const uint4*>(in + first)[i];
            v[4 * i] = q.x;
            v[4 * i + 1] = q.y;
            v[4 * i + 2] = q.z;
            v[4 * i + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            const uint64_t idx = first + i;
            v[i] = idx < n ? in[idx] : 0u;
        }
    }
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) s += v[i];
    uint64_t total;
    uint64_t run = block_offs[blockIdx.x] + block_excl_scan_u64(s, &total);
    if (whole) {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS / 2; i++) {
            const uint64_t a = run;
            run += v[2 * i];
            const uint64_t b = run;
            run += v[2 * i + 1];
            reinterpret_cast<ulonglong2*>(out + first)[i] = make_ulonglong2(a, b);
        }
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            const uint64_t idx = first + i;
            if (idx < n) out[idx] = run;
            run += v[i];
        }
    }
}

size_t scan_workspace_bytes(uint64_t capacity) {
    uint64_t nb = (capacity + SCAN_CHUNK - 1) / SCAN_CHUNK;
    return align_up(nb * sizeof(uint64_t) + 64, 256);
}

int scan_exclusive_u32(const uint32_t* in, uint64_t* out, const uint64_t* n_dev, uint64_t capacity,
                       uint64_t* total_dev, void* ws, hipStream_t s) {
    if (((uintptr_t)in | (uintptr_t)out) & 15)
        return set_error(RF_ERR_INVALID, "scan_exclusive_u32: arrays must be 16-byte aligned");
    int nb = (int)((capacity + SCAN_CHUNK - 1) / SCAN_CHUNK);
    if (nb == 0) nb = 1;
    uint64_t* sums = (uint64_t*)ws;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n_dev, capacity, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, sums, nb, total_dev);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n_dev, capacity, sums, out);
    RF_LAUNCH_CHECK("scan_exclusive_u32");
    return RF_OK;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits
// ---------------------------------------------------------------------------

static constexpr int RS_THREADS = 256;
static constexpr int RS_ITEMS = 16;
static constexpr int RS_CHUNK = RS_THREADS * RS_ITEMS;  // 4096

// Digit counts of one 4096-item chunk.  Keys are read four at a time (the
// key array and every chunk start are 16-byte aligned) and counted in one LDS
// histogram per wave: neighbouring pairs often share a digit, and one shared
// histogram would serialize their atomics.
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                            const uint64_t* __restrict__ n_ptr,
                                                            int shift, int nblocks,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_hist[RS_THREADS / 64][256];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; w++) s_hist[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t n = *n_ptr;
    const uint64_t base = (uint64_t)blockIdx.x * RS_CHUNK;
    if (base < n) {
        for (int i = 0; i < RS_ITEMS / 4; i++) {
            const uint64_t idx = base + 4 * ((uint64_t)i * RS_THREADS + threadIdx.x);
            if (idx + 3 < n) {
                const uint4 k = *reinterpret_cast<const uint4*>(keys + idx);
                atomicAdd(&s_hist[wave][(k.x >> shift) & 255u], 1u);
                atomicAdd(&s_hist[wave][(k.y >> shift) & 255u], 1u);
                atomicAdd(&s_hist[wave][(k.z >> shift) & 255u], 1u);
                atomicAdd(&s_hist[wave][(k.w >> shift) & 255u], 1u);
            } else {
                for (uint64_t j = idx; j < n && j < idx + 4; j++)
                    atomicAdd(&s_hist[wave][(keys[j] >> shift) & 255u], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; w++) c += s_hist[w][threadIdx.x];
    // digit-major: all blocks of digit 0, then digit 1, ...
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = c;
}

// Scatter of one 4096-item chunk.  Wave w owns the contiguous quarter
// base + 1024 w ... base + 1024 w + 1023 and ranks it round by round (64
// items a round, ranks within a round from eight ballots over the digit bits)
// against its own digit counts in LDS, so the ranking synchronizes at wave
// level only.  The four waves' counts then give each wave its place inside
// each digit's run of the chunk, waves in order, so the sort stays stable;
// the items are placed in LDS in digit order and written out from there, so
// consecutive threads store consecutive addresses of one digit's run.  The
// chunk's digit counts come from k_radix_hist (the same items, so the same
// counts).
static constexpr int RS_WAVES = RS_THREADS / 64;
static constexpr int RS_WAVE_ITEMS = RS_CHUNK / RS_WAVES;  // 1024
static_assert(RS_WAVE_ITEMS == 64 * RS_ITEMS, "one round of 64 items per thread item");

__device__ __forceinline__ void rs_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint64_t* __restrict__ n_ptr, int shift, int nblocks,
    const uint32_t* __restrict__ hist, const uint64_t* __restrict__ digit_offs) {
    __shared__ uint32_t s_keys[RS_CHUNK], s_vals[RS_CHUNK];
    __shared__ uint32_t s_loc[256];             // first LDS slot of each digit
    __shared__ uint32_t s_wcnt[RS_WAVES][256];  // a wave's items per digit, then its offset in the run
    __shared__ uint32_t s_wtot[RS_WAVES];
    __shared__ uint64_t s_doff[256];            // output offset of each digit's run of this chunk
    const uint64_t n = *n_ptr;
    const uint64_t base = (uint64_t)blockIdx.x * RS_CHUNK;
    if (base >= n) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    // the wave's items, loaded up front so the rounds do not each wait for memory
    const uint64_t wbase = base + (uint64_t)wave * RS_WAVE_ITEMS;
    uint32_t kr[RS_ITEMS], vr[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint64_t idx = wbase + (uint64_t)r * 64 + lane;
        kr[r] = idx < n ? keys_in[idx] : 0u;
        vr[r] = idx < n ? vals_in[idx] : 0u;
    }
    // thread t owns digit t: the chunk's count, the run's output offset
    const uint32_t cnt = hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x];
    s_doff[threadIdx.x] = digit_offs[(uint64_t)threadIdx.x * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RS_WAVES; w++) s_wcnt[w][threadIdx.x] = 0;
    uint32_t inc = cnt;  // inclusive scan of the counts over the wave's 64 digits
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= off) inc += v;
    }
    if (lane == 63) s_wtot[wave] = inc;
    __syncthreads();
    uint32_t below = 0;
    for (int w = 0; w < wave; w++) below += s_wtot[w];
    s_loc[threadIdx.x] = below + inc - cnt;
    // ranks within the wave's quarter
    uint32_t rank[RS_ITEMS];
    uint32_t* wcnt = s_wcnt[wave];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const bool valid = wbase + (uint64_t)r * 64 + lane < n;
        const uint32_t d = (kr[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t in_round = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t before = wcnt[d];
        rank[r] = before + in_round;
        rs_wave_sync();  // every lane has read its digit's count
        if (valid && in_round == 0) wcnt[d] = before + (uint32_t)__popcll(peers);
        rs_wave_sync();
    }
    __syncthreads();
    {  // each wave's offset inside each digit's run, waves in order
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            const uint32_t c = s_wcnt[w][threadIdx.x];
            s_wcnt[w][threadIdx.x] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        if (wbase + (uint64_t)r * 64 + lane < n) {
            const uint32_t d = (kr[r] >> shift) & 255u;
            const uint32_t pos = s_loc[d] + wcnt[d] + rank[r];
            s_keys[pos] = kr[r];
            s_vals[pos] = vr[r];
        }
    }
    __syncthreads();
    const uint32_t items = (uint32_t)min((uint64_t)RS_CHUNK, n - base);
    for (uint32_t i = threadIdx.x; i < items; i += RS_THREADS) {
        const uint32_t key = s_keys[i];
        const uint32_t d = (key >> shift) & 255u;
        const uint64_t dst = s_doff[d] + (i - s_loc[d]);
        keys_out[dst] = key;
        vals_out[dst] = s_vals[i];
    }
}

static int radix_blocks(uint64_t capacity) {
    uint64_t nb = (capacity + RS_CHUNK - 1) / RS_CHUNK;
    return nb == 0 ? 1 : (int)nb;
}

size_t radix_workspace_bytes(uint64_t capacity) {
    uint64_t nb = (uint64_t)radix_blocks(capacity);
    uint64_t nh = 256 * nb;
    size_t b = 0;
    b += align_up(2 * capacity * sizeof(uint32_t), 256);  // ping-pong keys+values
    b += align_up(nh * sizeof(uint32_t), 256);           // histograms
    b += align_up(nh * sizeof(uint64_t), 256);           // digit offsets
    b += scan_workspace_bytes(nh);
    return b;
}

int radix_sort_pairs(uint32_t* keys, uint32_t* vals, const uint64_t* n_dev, uint64_t capacity,
                     int key_bits, void* ws, hipStream_t s) {
    const int nb = radix_blocks(capacity);
    const uint64_t nh = 256ull * nb;
    if (((uintptr_t)keys | (uintptr_t)ws) & 15)
        return set_error(RF_ERR_INVALID, "radix_sort_pairs: keys and workspace must be 16-byte aligned");
    char* p = (char*)ws;
    uint32_t* k2 = (uint32_t*)p;
    uint32_t* v2 = k2 + capacity;
    p += align_up(2 * capacity * sizeof(uint32_t), 256);
    uint32_t* hist = (uint32_t*)p;
    p += align_up(nh * sizeof(uint32_t), 256);
    uint64_t* offs = (uint64_t*)p;
    p += align_up(nh * sizeof(uint64_t), 256);
    void* scan_ws = p;
    int passes = (key_bits + 7) / 8;
    if (passes < 1) passes = 1;
    // an even number of passes leaves the result in the caller's arrays
    if (passes & 1) passes++;
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (int pass = 0; pass < passes; pass++) {
        int shift = 8 * pass;
        if (shift > 24) shift = 24;  // extra pass for parity re-sorts the top digit (stable)
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(RS_THREADS), 0, s, ka, n_dev, shift, nb, hist);
        int rc = scan_exclusive_u32(hist, offs, nullptr, nh, nullptr, scan_ws, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(RS_THREADS), 0, s, ka, va, kb, vb, n_dev,
                           shift, nb, hist, offs);
        RF_LAUNCH_CHECK("radix_sort_pairs");
        uint32_t* t = ka;
        ka = kb;
        kb = t;
        t = va;
        va = vb;
        vb = t;
    }
    return RF_OK;
}

}  // namespace rf
