// k_slicing.hip -- longitudinal slicing on the device (src/tasks/strongstrong/slicing.jl).
//
//   zlanes/zfold : _live_z_stats (slicing.jl:271-289): exact min/max, and the mean and sigma through
//                  the canonical 4096-lane fold (lane t sums z[t], z[t+4096], ... in order; the lanes
//                  are folded serially 0..4095), so the values are bit-identical to the CPU fold.
//   boundaries   : :normal_quantile (slicing.jl:405-429), :equal_area (104-156, histogram of
//                  ns*resolution bins), :equal_width (367-387), :specified (389-403).
//   classify     : s = clamp(searchsortedlast(boundaries, z), 1, ns) per particle (slicing.jl:431-460)
//                  plus a per-block slice histogram.
//   scan/scatter : stable counting sort: slice-major exclusive scan of the block histograms, then
//                  each particle's rank among same-slice particles of its block from wave ballots.
//                  Member lists come out ascending, as the reference's index lists.  The scatter also
//                  gathers the six coordinates into slice-contiguous state arrays.
//   centroid     : the same 4096-lane fold over each member list (slicing.jl:245-257, 523-559).
//   tile order   : members regrouped by transverse tile within each slice (no counterpart in the
//                  reference: a memory-layout choice for the field gathers; results do not depend on it).

#include "common.h"
#include "launch.h"

namespace octopic {

constexpr int kSliceBlock = 1024;     // particles per classify/scatter block
constexpr int kSliceThreads = 256;

// A block takes 64 lanes: wave 0 sums them in the fold's order from LDS while waves 1..15 stream the next
// kZChunk rows of them into the other half of a double buffer (one row of 64 lanes is 64 consecutive
// values), so the in-order sums do not wait on HBM.  Pass 1: the sums and z's min/max (block partials
// to mm_partial); pass 2: the sums of (z - mean)^2.
constexpr int kZWaves = 16;                                  // per block: one summing wave, 15 loading
constexpr int kZRowsPerLoader = 8;
constexpr int kZChunk = (kZWaves - 1) * kZRowsPerLoader;     // rows per buffer half
static_assert(kZFoldBlocks * 64 == kFoldLanes, "64 lanes per block");

__global__ __launch_bounds__(64 * kZWaves) void k_zlanes(const double *z, int64_t n, const SliceTable *tab, int pass,
                                                         double *lanes, double *mm_partial) {
    extern __shared__ double zbuf[];                          // [2][kZChunk][64]
    __shared__ double slo[kZWaves], shi[kZWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.x * 64 + lane;                     // this thread's lane of the fold
    const int64_t rows = (n + kFoldLanes - 1) / kFoldLanes;  // rows of the fold, the last one partial
    const int64_t nchunks = (rows + kZChunk - 1) / kZChunk;
    const double mu = pass == 2 ? tab->mean : 0.0;
    double lo = INFINITY, hi = -INFINITY;
    auto load = [&](int64_t c) {   // loader wave w: rows c*kZChunk + (w - 1) + 15 k of the chunk
        double v[kZRowsPerLoader];
#pragma unroll
        for (int k = 0; k < kZRowsPerLoader; ++k) {
            const int64_t row = c * kZChunk + (wave - 1) + (kZWaves - 1) * k;
            const int64_t i = row * kFoldLanes + t;
            v[k] = (row < rows && i < n) ? z[i] : 0.0;
            if (pass == 1 && row < rows && i < n) { lo = jmin(lo, v[k]); hi = jmax(hi, v[k]); }
        }
#pragma unroll
        for (int k = 0; k < kZRowsPerLoader; ++k)
            zbuf[((c & 1) * kZChunk + (wave - 1) + (kZWaves - 1) * k) * 64 + lane] = v[k];
    };
    double acc = 0.0;
    if (wave > 0) load(0);
    __syncthreads();
    for (int64_t c = 0; c < nchunks; ++c) {
        if (wave > 0) {
            if (c + 1 < nchunks) load(c + 1);
        } else {
            // this lane's rows of the chunk (the last row of the fold may miss it), read 16 at a time ahead
            // of the in-order adds
            const int64_t r0 = c * kZChunk;
            const int64_t mine = n > t ? (n - t + kFoldLanes - 1) / kFoldLanes : 0;
            const int rc = (int)((mine - r0) < kZChunk ? ((mine - r0) > 0 ? (mine - r0) : 0) : kZChunk);
            const double *b = zbuf + (c & 1) * kZChunk * 64 + lane;
            int r = 0;
            for (; r + 16 <= rc; r += 16) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = b[(r + u) * 64];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    if (pass == 2) v[u] = (v[u] - mu) * (v[u] - mu);
                    acc += v[u];
                }
            }
            for (; r < rc; ++r) {
                double v = b[r * 64];
                if (pass == 2) v = (v - mu) * (v - mu);
                acc += v;
            }
        }
        __syncthreads();   // the chunk is summed before its half is loaded again
    }
    if (wave == 0) lanes[t] = acc;
    if (pass == 1) {
        for (int o = 32; o > 0; o >>= 1) {
            lo = jmin(lo, __shfl_xor(lo, o));
            hi = jmax(hi, __shfl_xor(hi, o));
        }
        if (lane == 0) { slo[wave] = lo; shi[wave] = hi; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double l = slo[0], h = shi[0];
            for (int w = 1; w < kZWaves; ++w) { l = jmin(l, slo[w]); h = jmax(h, shi[w]); }
            mm_partial[2 * blockIdx.x] = l;
            mm_partial[2 * blockIdx.x + 1] = h;
        }
    }
}

// Serial lane fold (one lane) + boundaries.  method: OCTOPIC slicing method; q: normal quantiles
// sqrt(2)*erfinv(2s/ns - 1) or the sorted `specified` offsets, from the host.
// The 4096 lane partials are staged in LDS by the whole block, then folded serially in lane order.
__device__ __forceinline__ double fold_lanes_serial(const double *l) {
    // 32 values read ahead of their in-order adds, so the chain of adds does not wait on each LDS read
    double s = 0.0;
    for (int t = 0; t < kFoldLanes; t += 32) {
        double v[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) v[u] = l[t + u];
#pragma unroll
        for (int u = 0; u < 32; ++u) s += v[u];
    }
    return s;
}

__global__ __launch_bounds__(256) void k_zfold(const double *lanes, const double *mm_partial, int64_t n, SliceTable *tab,
                                               int pass, int method, int ns, const double *q, ErrorInfo *err) {
    __shared__ double sl[kFoldLanes];
    __shared__ double smm[2 * kZFoldBlocks];
    for (int t = threadIdx.x; t < kFoldLanes; t += blockDim.x) sl[t] = lanes[t];
    if (pass == 1)
        for (int t = threadIdx.x; t < 2 * kZFoldBlocks; t += blockDim.x) smm[t] = mm_partial[t];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double s = fold_lanes_serial(sl);
    if (pass == 1) {
        double lo = INFINITY, hi = -INFINITY;
        for (int b = 0; b < kZFoldBlocks; ++b) { lo = jmin(lo, smm[2 * b]); hi = jmax(hi, smm[2 * b + 1]); }
        tab->zmin = lo;
        tab->zmax = hi;
        tab->mean = s / (double)n;
        tab->ns = ns;
        return;
    }
    double sigma = sqrt(jmax(s / (double)n, 0.0));
    tab->sigma = sigma;
    const double zmin = tab->zmin, zmax = tab->zmax, mean = tab->mean;
    double *b = tab->boundary;
    if (method == 3) {   // normal_quantile
        if (sigma == 0.0) {
            for (int k = 0; k <= ns; ++k) b[k] = mean;
        } else {
            b[0] = zmin; b[ns] = zmax;
            for (int k = 1; k < ns; ++k) b[k] = jclamp(mean + sigma * q[k], b[0], b[ns]);
        }
    } else if (method == 4) {   // specified: q holds the positions; sort the internal boundaries
        double tmp[kMaxSlices];
        int m = ns - 1;
        for (int k = 0; k < m; ++k) tmp[k] = mean + q[k] * sigma;
        for (int a = 1; a < m; ++a) {   // stable insertion sort by isless
            double v = tmp[a];
            int c = a - 1;
            while (c >= 0 && jisless(v, tmp[c])) { tmp[c + 1] = tmp[c]; --c; }
            tmp[c + 1] = v;
        }
        b[0] = zmin; b[ns] = zmax;
        for (int k = 0; k < m; ++k) b[k + 1] = jclamp(tmp[k], b[0], b[ns]);
    } else if (method == 2) {   // equal_width
        b[0] = zmin; b[ns] = zmax;
        double step = (zmax - zmin) / (double)ns;
        for (int k = 1; k < ns; ++k) b[k] = zmin + (double)k * step;
    }
    // equal_area boundaries are produced by k_equal_area after the histogram, equal_count ones by
    // k_rank_table after the rank sort
    bool finite = true;
    if (method != 0 && method != 1)
        for (int k = 0; k <= ns; ++k) finite = finite && __builtin_isfinite(b[k]);
    finite = finite && __builtin_isfinite(mean) && __builtin_isfinite(zmin) && __builtin_isfinite(zmax);
    if (!finite) atomicCAS(&err->code, 0, 3);
    // equal_count membership is by rank: no degenerate special case (slicing.jl:330-365)
    tab->degenerate = method == 1 ? 0 : ((method == 0 || method == 2) ? (zmin == zmax) : (b[0] == b[ns]));
}

// :equal_area histogram (_threaded_histogram / _slice_bin, slicing.jl:172-176, 291-328), LDS-privatised.
__global__ __launch_bounds__(256) void k_equal_area_hist(const double *z, int64_t n, const SliceTable *tab, int bins,
                                                         unsigned long long *counts) {
    extern __shared__ unsigned int hcnt[];
    const double zmin = tab->zmin, zmax = tab->zmax;
    if (zmin == zmax) return;
    const double width = (zmax - zmin) / (double)bins;
    for (int k = threadIdx.x; k < bins; k += blockDim.x) hcnt[k] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = (z[i] - zmin) / width;
        if (!(d > -INFINITY && d < INFINITY)) continue;
        int bin = iclamp((int)floor(d) + 1, 1, bins);
        atomicAdd(&hcnt[bin - 1], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < bins; k += blockDim.x)
        if (hcnt[k]) atomicAdd(&counts[k], (unsigned long long)hcnt[k]);
}

// :equal_area boundaries (slicing.jl:104-156); the count buffer is reused for the cumulative
// fractions and reset to zero on exit.
__global__ void k_equal_area(SliceTable *tab, int64_t n, int ns, int bins, unsigned long long *counts, ErrorInfo *err) {
    if (threadIdx.x != 0) return;
    const double zmin = tab->zmin, zmax = tab->zmax;
    double *b = tab->boundary;
    if (zmin == zmax) {
        for (int k = 0; k <= ns; ++k) b[k] = zmin;
        tab->degenerate = 1;
    } else {
        const double width = (zmax - zmin) / (double)bins;
        b[0] = zmin; b[ns] = zmax;
        double *cumd = reinterpret_cast<double *>(counts);
        int64_t acc = 0;
        for (int k = 0; k < bins; ++k) {
            acc += (int64_t)counts[k];
            cumd[k] = (double)acc / (double)(n > 1 ? n : 1);
        }
        cumd[bins - 1] = 1.0;
        int current = 1;   // 1-based bin cursor
        for (int s = 1; s < ns; ++s) {
            double target = (double)s / (double)ns;
            while (current <= bins && cumd[current - 1] <= target) current += 1;
            double x1, x2, y1, y2;
            if (current <= 1) {
                x1 = b[0]; x2 = zmin + (1.0 - 0.5) * width; y1 = 0.0; y2 = cumd[0];
            } else if (current > bins) {
                x1 = zmin + ((double)bins - 0.5) * width; x2 = b[ns]; y1 = cumd[bins - 2]; y2 = 1.0;
            } else {
                x1 = zmin + ((double)(current - 1) - 0.5) * width;
                x2 = zmin + ((double)current - 0.5) * width;
                y1 = cumd[current - 2]; y2 = cumd[current - 1];
            }
            b[s] = (y2 == y1) ? (x1 + x2) / 2.0 : x2 * (target - y1) / (y2 - y1) + x1 * (target - y2) / (y1 - y2);
        }
        for (int k = 0; k < bins; ++k) counts[k] = 0ull;
        tab->degenerate = (b[0] == b[ns]);
    }
    bool finite = true;
    for (int k = 0; k <= ns; ++k) finite = finite && __builtin_isfinite(b[k]);
    if (!finite) atomicCAS(&err->code, 0, 3);
}

__device__ __forceinline__ int slice_of(double zi, const double *b, int ns) {
    // searchsortedlast(boundaries, z) over ns+1 entries, 1-based result; then clamp(., 1, ns)
    int lo = 0, hi = ns + 2;
    while (lo < hi - 1) {
        int m = lo + ((hi - lo) >> 1);
        if (jisless(zi, b[m - 1])) hi = m; else lo = m;
    }
    return iclamp(lo, 1, ns) - 1;
}

// Folds this block's (x min, x max, y min, y max) into the 4 bound slots (exact, any order).  Called by
// every thread of a kSliceThreads block.
__device__ __forceinline__ void fold_box(const double v[4], unsigned long long *box) {
    __shared__ unsigned long long red[kSliceThreads / 64][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 4; ++q) {
        unsigned long long u = bound_slot(v[q], (q & 1) == 0);
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long w = __shfl_xor(u, o);
            u = w > u ? w : u;
        }
        if (lane == 0) red[wave][q] = u;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long u = 0ull;
        for (int w = 0; w < kSliceThreads / 64; ++w) u = red[w][threadIdx.x] > u ? red[w][threadIdx.x] : u;
        if (u) atomicMax(&box[threadIdx.x], u);
    }
}

__global__ __launch_bounds__(kSliceThreads) void k_classify(const double *z, const double *x, const double *y, int64_t n,
                                                            const SliceTable *tab, uint16_t *sid, int *hist, int nblocks,
                                                            int by_bin, unsigned long long *box, const ErrorInfo *err) {
    __shared__ int cnt[kMaxSlices];
    __shared__ double b[kMaxSlices + 1];
    const int ns = tab->ns;
    const bool degen = tab->degenerate != 0;
    for (int k = threadIdx.x; k < ns; k += blockDim.x) cnt[k] = 0;
    for (int k = threadIdx.x; k <= ns; k += blockDim.x) b[k] = tab->boundary[k];
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSliceBlock;
    double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int r = 0; r < kSliceBlock / kSliceThreads; ++r) {
        int64_t i = base + r * kSliceThreads + threadIdx.x;
        if (i < n) {
            int s;
            if (degen) s = 0;
            else if (by_bin) {   // :equal_width membership is _slice_bin (slicing.jl:382-385)
                double d = (z[i] - b[0]) / ((b[ns] - b[0]) / (double)ns);
                s = (d > -INFINITY && d < INFINITY) ? iclamp((int)floor(d) + 1, 1, ns) - 1 : 0;
            } else s = slice_of(z[i], b, ns);
            sid[i] = (uint16_t)s;
            atomicAdd(&cnt[s], 1);
            if (box) {
                const double xi = x[i], yi = y[i];
                v[0] = jmin(v[0], xi); v[1] = jmax(v[1], xi); v[2] = jmin(v[2], yi); v[3] = jmax(v[3], yi);
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ns; k += blockDim.x) hist[(size_t)k * nblocks + blockIdx.x] = cnt[k];
    if (box) fold_box(v, box);   // the beam's transverse box, for the tile order below
}

// Exclusive scan of hist (slice-major, ns * nblocks) in place; writes the per-slice count/offset.
// Exclusive prefix sum of one int64 per thread over a block of up to 1024 threads (wave shuffles, then
// the wave totals), plus `base`.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t base) {
    __shared__ int64_t wave_tot[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x + 63) >> 6;
    int64_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    int64_t before = base;
    for (int w = 0; w < wave; ++w) before += wave_tot[w];
    (void)nw;
    __syncthreads();   // wave_tot may be reused by the caller's next scan
    return before + inc - v;
}

// Exclusive scan of `total` ints in place by one block: each wave takes a contiguous segment and scans it
// 64 values at a time (coalesced loads, shuffles, a running carry), after the segments' bases from a first
// pass of wave sums.  on(k, value) is called with each entry's exclusive prefix.
template <class On>
__device__ __forceinline__ void scan_ints_inplace(int *h, int64_t total, On on) {
    __shared__ int64_t seg_sum[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const int64_t seg = (total + nw - 1) / nw;
    const int64_t lo = wave * seg < total ? wave * seg : total, hi = lo + seg < total ? lo + seg : total;
    int64_t sum = 0;
    for (int64_t k0 = lo; k0 < hi; k0 += 4 * 64) {
        int v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t k = k0 + u * 64 + lane;
            v[u] = k < hi ? h[k] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) sum += v[u];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (lane == 0) seg_sum[wave] = sum;
    __syncthreads();
    int64_t base = 0;
    for (int w = 0; w < wave; ++w) base += seg_sum[w];
    for (int64_t k0 = lo; k0 < hi; k0 += 4 * 64) {
        int v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t k = k0 + u * 64 + lane;
            v[u] = k < hi ? h[k] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t k = k0 + u * 64 + lane;
            int64_t inc = v[u];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int64_t t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const int64_t ex = base + inc - v[u];
            if (k < hi) {
                h[k] = (int)ex;
                on(k, ex);
            }
            base += __shfl(inc, 63);
        }
    }
}

// Exclusive scan of hist (slice-major, ns * nblocks) in place; writes the per-slice offsets.
__global__ __launch_bounds__(1024) void k_scan(int *hist, int ns, int nblocks, SliceTable *tab) {
    scan_ints_inplace(hist, (int64_t)ns * nblocks, [&](int64_t k, int64_t ex) {
        if (k % nblocks == 0) tab->offset[k / nblocks] = ex;
    });
}

__global__ void k_counts(SliceTable *tab, int64_t n) {
    const int ns = tab->ns;
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        int64_t end = (s + 1 < ns) ? tab->offset[s + 1] : n;
        tab->count[s] = end - tab->offset[s];
    }
}

static_assert(kMaxSlices <= 256, "slice ids are ranked by their 8 low bits");
__global__ __launch_bounds__(kSliceThreads) void k_scatter(const double *const *beam, int64_t n, const uint16_t *sid,
                                                           const int *hist, int nblocks, int ns, double *const *st,
                                                           int64_t *perm, int32_t *inv) {
    __shared__ int run[kMaxSlices];
    __shared__ int wcnt[kSliceThreads / 64][kMaxSlices];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < ns; k += blockDim.x) run[k] = 0;
    const int64_t base = (int64_t)blockIdx.x * kSliceBlock;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSliceBlock / kSliceThreads; ++r) {
        for (int k = threadIdx.x; k < (kSliceThreads / 64) * kMaxSlices; k += blockDim.x) (&wcnt[0][0])[k] = 0;
        __syncthreads();
        const int64_t i = base + r * kSliceThreads + threadIdx.x;
        const bool valid = i < n;
        const int s = valid ? (int)sid[i] : 0;
        // the wave's lanes in the same slice: agree on every id bit (ids < kMaxSlices <= 256: 8 bits)
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long m = __ballot(valid && ((s >> bit) & 1));
            peers &= ((s >> bit) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][s] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = hist[(size_t)s * nblocks + blockIdx.x] + run[s] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][s];
            for (int c = 0; c < 6; ++c) st[c][pos] = beam[c][i];
            perm[pos] = i;
            if (inv) inv[i] = pos;   // coalesced: i is this thread's own index
        }
        __syncthreads();
        for (int k = threadIdx.x; k < ns; k += blockDim.x) {
            int add = 0;
            for (int w = 0; w < kSliceThreads / 64; ++w) add += wcnt[w][k];
            run[k] += add;
        }
        __syncthreads();
    }
}

// Lane partials of every slice's member z (slice-contiguous state); grid (16, ns).
__global__ __launch_bounds__(256) void k_centroid_lanes(const double *stz, const SliceTable *tab, double *lanes) {
    const int s = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t off = tab->offset[s], cnt = tab->count[s];
    double acc = 0.0;
    int64_t k = t;
    for (; k + 3 * kFoldLanes < cnt; k += 4 * kFoldLanes) {
        double v0 = stz[off + k], v1 = stz[off + k + kFoldLanes], v2 = stz[off + k + 2 * kFoldLanes],
               v3 = stz[off + k + 3 * kFoldLanes];
        acc += v0; acc += v1; acc += v2; acc += v3;
    }
    for (; k < cnt; k += kFoldLanes) acc += stz[off + k];
    lanes[(size_t)s * kFoldLanes + t] = acc;
}

// one block per slice
__global__ __launch_bounds__(256) void k_centroid_fold(const double *lanes, SliceTable *tab, int64_t n, int centroid) {
    __shared__ double sl[kFoldLanes];
    const int s = blockIdx.x;
    const int64_t cnt = tab->count[s];
    const bool fold = centroid && cnt > 0;
    if (fold) {
        for (int t = threadIdx.x; t < kFoldLanes; t += blockDim.x) sl[t] = lanes[(size_t)s * kFoldLanes + t];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    tab->weight[s] = (double)cnt / (double)n;
    if (fold) {
        tab->center[s] = fold_lanes_serial(sl) / (double)cnt;
    } else {
        tab->center[s] = (tab->boundary[s] + tab->boundary[s + 1]) / 2.0;
    }
}

// ---------------------------------------------------------------------- transverse tile order
// After the slice scatter and the centroid folds (which need each member list in ascending index
// order), the members of every slice are regrouped by transverse tile: T x T tiles over the beam's
// x-y box at slicing time, in Morton order of the tiles within the slice.  The field gathers of the kick then share
// cache lines across a wave.  Only the order inside a slice changes, and no later result depends on
// it: deposits are exact integer sums, bounds are exact min/max, kicks are per particle, and the
// final store goes through the permutation.
struct TileBox { double x0, y0, sx, sy; };

__device__ __forceinline__ TileBox tile_box(const unsigned long long *box, int T) {
    const double xmin = bound_value(box[0], true), xmax = bound_value(box[1], false);
    const double ymin = bound_value(box[2], true), ymax = bound_value(box[3], false);
    const double wx = xmax - xmin, wy = ymax - ymin;
    TileBox b;
    b.x0 = __builtin_isfinite(xmin) ? xmin : 0.0;
    b.y0 = __builtin_isfinite(ymin) ? ymin : 0.0;
    b.sx = (__builtin_isfinite(wx) && wx > 0.0) ? (double)T / wx : 0.0;
    b.sy = (__builtin_isfinite(wy) && wy > 0.0) ? (double)T / wy : 0.0;
    return b;
}
__device__ __forceinline__ int tile_axis(double u, int T) {
    return u >= 0.0 ? (u < (double)T ? (int)u : T - 1) : 0;   // NaN -> 0
}
// Tiles in Morton (Z) order, T a power of two: consecutive tiles form compact 2-D patches, so a run of
// consecutive members (one block's chunk) covers a small square of the mesh.
__device__ __forceinline__ unsigned spread_bits(unsigned v) {   // v < 2^8 -> bits at even positions
    v = (v | (v << 4)) & 0x0F0Fu;
    v = (v | (v << 2)) & 0x3333u;
    v = (v | (v << 1)) & 0x5555u;
    return v;
}
__device__ __forceinline__ int tile_of(double x, double y, const TileBox &b, int T) {
    return (int)((spread_bits((unsigned)tile_axis((x - b.x0) * b.sx, T)) << 1) |
                 spread_bits((unsigned)tile_axis((y - b.y0) * b.sy, T)));
}

constexpr int kTileThreads = 1024;
constexpr int kTileRunsMaxTiles = 4096;   // up to 64 x 64 tiles: the run-sorted scatter (LDS, 16-bit ids)

// grid (G, ns): per-block tile counts, thist[(s * TT + tile) * G + c]
__global__ __launch_bounds__(kTileThreads) void k_tile_hist(TileSortArgs a) {
    extern __shared__ int cnt[];
    const int s = blockIdx.y, c = blockIdx.x, T = a.T, TT = T * T, G = gridDim.x;
    const TileBox tb = tile_box(a.box, T);
    for (int t = threadIdx.x; t < TT; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    const int64_t off = a.tab->offset[s], n = a.tab->count[s];
    const double *X = a.src[0] + off, *Y = a.src[2] + off;
    for (int64_t base = (int64_t)c * kTileChunk; base < n; base += (int64_t)G * kTileChunk)
        for (int64_t k = base + threadIdx.x; k < n && k < base + kTileChunk; k += blockDim.x)
            atomicAdd(&cnt[tile_of(X[k], Y[k], tb, T)], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < TT; t += blockDim.x) a.thist[((size_t)s * TT + t) * G + c] = cnt[t];
}

// one block per slice: exclusive scan of the slice's TT * G counts, plus the slice offset
__global__ __launch_bounds__(1024) void k_tile_scan(TileSortArgs a, int G) {
    const int s = blockIdx.x;
    const int64_t total = (int64_t)a.T * a.T * G;
    int *h = a.thist + (size_t)s * total;
    const int64_t per = (total + blockDim.x - 1) / blockDim.x;
    const int64_t lo = threadIdx.x * per < total ? threadIdx.x * per : total;
    const int64_t hi = lo + per < total ? lo + per : total;
    int64_t sum = 0;
    for (int64_t k = lo; k < hi; ++k) sum += h[k];
    int64_t run = block_exclusive_scan(sum, a.tab->offset[s]);
    for (int64_t k = lo; k < hi; ++k) {
        const int v = h[k];
        h[k] = (int)run;
        run += v;
    }
}

// grid (G, ns): the same visits as k_tile_hist; a member's slot is its block's next free slot in its tile
__global__ __launch_bounds__(kTileThreads) void k_tile_scatter(TileSortArgs a) {
    extern __shared__ int cur[];
    const int s = blockIdx.y, c = blockIdx.x, T = a.T, TT = T * T, G = gridDim.x;
    const TileBox tb = tile_box(a.box, T);
    for (int t = threadIdx.x; t < TT; t += blockDim.x) cur[t] = a.thist[((size_t)s * TT + t) * G + c];
    __syncthreads();
    const int64_t off = a.tab->offset[s], n = a.tab->count[s];
    for (int64_t base = (int64_t)c * kTileChunk; base < n; base += (int64_t)G * kTileChunk)
        for (int64_t k = base + threadIdx.x; k < n && k < base + kTileChunk; k += blockDim.x) {
            const int64_t i = off + k;
            const double x = a.src[0][i], y = a.src[2][i];
            const int64_t pos = atomicAdd(&cur[tile_of(x, y, tb, T)], 1);
            a.dst[0][pos] = x;
            a.dst[1][pos] = a.src[1][i];
            a.dst[2][pos] = y;
            a.dst[3][pos] = a.src[3][i];
            a.dst[4][pos] = a.src[4][i];
            a.dst[5][pos] = a.src[5][i];
            const int64_t orig = a.perm_src[i];
            a.perm_dst[pos] = orig;
            a.inv[orig] = (int32_t)pos;
        }
}

// Exclusive scan of v over the block (kTileThreads threads, one value each); returns the block total.
__device__ __forceinline__ int block_exclusive_scan(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < kTileThreads / 64; ++w) { const int t = wsum[w]; wsum[w] = run; run += t; }
        wsum[kTileThreads / 64] = run;
    }
    __syncthreads();
    total = wsum[kTileThreads / 64];
    return wsum[wave] + incl - v;
}

// The same slots as k_tile_scatter, written in runs: each visit of kTileChunk members is first sorted
// by tile in LDS (local counting sort); then every coordinate array is read in member order
// (coalesced), placed at its sorted position in an LDS stage, and written back as the visit's tile
// runs to their consecutive slots, so both the reads and the writes coalesce.
// Dynamic LDS: cur, count, offset [TT] ints, the stage [kTileChunk] doubles, the sorted tiles
// [kTileChunk] ushort.
__global__ __launch_bounds__(kTileThreads) void k_tile_scatter_runs(TileSortArgs a) {
    extern __shared__ int smem[];
    const int s = blockIdx.y, c = blockIdx.x, T = a.T, TT = T * T, G = gridDim.x;
    int *cur = smem, *lcnt = smem + TT, *loff = smem + 2 * TT;
    double *stage = reinterpret_cast<double *>(smem + ((3 * TT + 1) & ~1));
    unsigned short *ltile = reinterpret_cast<unsigned short *>(stage + kTileChunk);
    __shared__ int wsum[kTileThreads / 64 + 1];
    constexpr int kPer = kTileChunk / kTileThreads;
    const TileBox tb = tile_box(a.box, T);
    for (int t = threadIdx.x; t < TT; t += blockDim.x) cur[t] = a.thist[((size_t)s * TT + t) * G + c];
    const int64_t off = a.tab->offset[s], n = a.tab->count[s];
    for (int64_t base = (int64_t)c * kTileChunk; base < n; base += (int64_t)G * kTileChunk) {
        const int cnt = (int)(n - base < kTileChunk ? n - base : kTileChunk);
        for (int t = threadIdx.x; t < TT; t += blockDim.x) lcnt[t] = 0;
        __syncthreads();
        int rt[kPer], lp[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int j = q * kTileThreads + threadIdx.x;
            rt[q] = -1;
            if (j < cnt) {
                const int64_t i = off + base + j;
                rt[q] = tile_of(a.src[0][i], a.src[2][i], tb, T);
                lp[q] = atomicAdd(&lcnt[rt[q]], 1);
            }
        }
        __syncthreads();
        // exclusive scan of the visit's tile counts, TT / kTileThreads consecutive tiles per thread
        {
            const int per = (TT + kTileThreads - 1) / kTileThreads;
            const int t0 = threadIdx.x * per;
            int mine = 0;
            for (int t = t0; t < t0 + per && t < TT; ++t) mine += lcnt[t];
            int total;
            int run = block_exclusive_scan(mine, wsum, total);
            for (int t = t0; t < t0 + per && t < TT; ++t) { loff[t] = run; run += lcnt[t]; }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (rt[q] >= 0) {
                lp[q] += loff[rt[q]];
                ltile[lp[q]] = (unsigned short)rt[q];
            }
        // arrays 0-5: coordinates; 6: the original indices (int64 through the same stage)
        for (int arr = 0; arr < 7; ++arr) {
            const double *src = arr < 6 ? a.src[arr] : reinterpret_cast<const double *>(a.perm_src);
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                if (rt[q] >= 0) stage[lp[q]] = src[off + base + q * kTileThreads + threadIdx.x];
            __syncthreads();
            for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
                const int t = ltile[p];
                const int64_t pos = (int64_t)cur[t] + (p - loff[t]);
                if (arr < 6) {
                    a.dst[arr][pos] = stage[p];
                } else {
                    const int64_t orig = reinterpret_cast<const int64_t *>(stage)[p];
                    a.perm_dst[pos] = orig;
                    a.inv[orig] = (int32_t)pos;
                }
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < TT; t += blockDim.x) cur[t] += lcnt[t];
        __syncthreads();
    }
}

hipError_t launch_tile_sort(const TileSortArgs &a, int G, hipStream_t s) {
    if (a.ns <= 0 || a.T <= 0 || G <= 0) return hipSuccess;
    const int TT = a.T * a.T;
    const size_t lds = (size_t)TT * sizeof(int);
    hipFuncSetAttribute((const void *)k_tile_hist, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_tile_hist, dim3(G, a.ns), dim3(kTileThreads), lds, s, a);
    hipLaunchKernelGGL(k_tile_scan, dim3(a.ns), dim3(1024), 0, s, a, G);
    if (TT <= kTileRunsMaxTiles) {   // tiles fit the LDS sort of a visit (16-bit tile and member ids)
        const size_t lds_runs = (size_t)((3 * TT + 1) & ~1) * sizeof(int) + (size_t)kTileChunk * sizeof(double) +
                                (size_t)kTileChunk * sizeof(unsigned short);
        hipFuncSetAttribute((const void *)k_tile_scatter_runs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_runs);
        hipLaunchKernelGGL(k_tile_scatter_runs, dim3(G, a.ns), dim3(kTileThreads), lds_runs, s, a);
    } else {
        hipFuncSetAttribute((const void *)k_tile_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_tile_scatter, dim3(G, a.ns), dim3(kTileThreads), lds, s, a);
    }
    return hipGetLastError();
}

// Tile-ordered states back into the beam through the inverse permutation: writes run in beam order
// (coalesced) and the reads gather.
__global__ __launch_bounds__(256) void k_store_gather(double *const *beam, int64_t n, double *const *st, const int32_t *inv,
                                                      const ErrorInfo *err, int with_z) {
    if (err->code != 0) return;
    double *X = beam[0], *PX = beam[1], *Y = beam[2], *PY = beam[3], *Z = beam[4], *PZ = beam[5];
    const double *SX = st[0], *SPX = st[1], *SY = st[2], *SPY = st[3], *SZ = st[4], *SPZ = st[5];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = inv[i];
        X[i] = SX[pos];
        PX[i] = SPX[pos];
        Y[i] = SY[pos];
        PY[i] = SPY[pos];
        if (with_z) Z[i] = SZ[pos];
        PZ[i] = SPZ[pos];
    }
}

hipError_t launch_store_gather(double *const *beam, int64_t n, double *const *st, const int32_t *inv, const ErrorInfo *err,
                               hipStream_t s, int with_z) {
    int g = (int)((n + 255) / 256);
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_store_gather, dim3(g), dim3(256), 0, s, beam, n, st, inv, err, with_z);
    return hipGetLastError();
}

// ---------------------------------------------------------------------- launcher
// ---------------------------------------------------------------- :equal_count (slicing.jl:330-365)
// Membership by rank: a stable sort of the order-preserving keys of z (ties keep ascending index, as
// the reference's stable sortperm; -0.0 below +0.0 as isless) gives the rank order; slice s holds ranks
// [floor(s n / ns), floor((s+1) n / ns)) and its members stay in rank order (the order the centroid
// lane fold sums them in).  Internal boundaries are midpoints of the neighbouring ranks' z.
__global__ void k_rank_keys(const double *z, int64_t n, unsigned long long *keys, int64_t *vals) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = order_key(z[i]);
        vals[i] = i;
    }
}

__global__ void k_rank_table(SliceTable *tab, const unsigned long long *sorted, int64_t n, int ns) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double *b = tab->boundary;
    b[0] = tab->zmin;
    b[ns] = tab->zmax;
    for (int s = 0; s < ns; ++s) {
        const int64_t lo = (int64_t)s * n / ns, hi = (int64_t)(s + 1) * n / ns;
        tab->offset[s] = lo;
        tab->count[s] = hi - lo;
    }
    for (int s = 1; s < ns; ++s) {
        const int64_t pos = (int64_t)s * n / ns;   // 1-based rank of the last member of slice s-1
        if (pos == 0) b[s] = key_value(sorted[0]);
        else if (pos == n) b[s] = key_value(sorted[n - 1]);
        else b[s] = (key_value(sorted[pos - 1]) + key_value(sorted[pos])) / 2.0;
    }
}

// states in rank order (st[c][r] = beam[c][perm[r]]) and the beam's transverse box
__global__ __launch_bounds__(kSliceThreads) void k_rank_gather(const double *const *beam, int64_t n, const int64_t *perm,
                                                               int32_t *inv, double *const *st, unsigned long long *box) {
    double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = perm[r];
        const double x = beam[0][i], y = beam[2][i];
        st[0][r] = x;
        st[1][r] = beam[1][i];
        st[2][r] = y;
        st[3][r] = beam[3][i];
        st[4][r] = beam[4][i];
        st[5][r] = beam[5][i];
        if (inv) inv[i] = (int32_t)r;
        v[0] = jmin(v[0], x); v[1] = jmax(v[1], x); v[2] = jmin(v[2], y); v[3] = jmax(v[3], y);
    }
    if (box) fold_box(v, box);
}

// The rank order itself: a stable LSD radix sort of the 64-bit keys with their indices, eight passes of
// 8-bit digits.  Each pass is a stable counting sort over kSliceBlock-element blocks: the blocks' digit
// histograms, their exclusive scan in digit-major order (a digit's elements of block b follow those of
// blocks < b), and a scatter that ranks a wave's elements within their digit by eight digit-bit ballots and
// the waves of a block in order, so every digit keeps the input order.
constexpr int kRadixBins = 256;

__global__ __launch_bounds__(kSliceThreads) void k_radix_hist(const unsigned long long *keys, int64_t n, int shift,
                                                              int *hist, int nblocks) {
    __shared__ int cnt[kRadixBins];
    for (int k = threadIdx.x; k < kRadixBins; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSliceBlock;
    for (int r = 0; r < kSliceBlock / kSliceThreads; ++r) {
        const int64_t i = base + r * kSliceThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(int)((keys[i] >> shift) & (kRadixBins - 1))], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kRadixBins; k += blockDim.x) hist[(size_t)k * nblocks + blockIdx.x] = cnt[k];
}

// exclusive scan of hist (digit-major, kRadixBins * nblocks) in place
__global__ __launch_bounds__(1024) void k_radix_scan(int *hist, int nblocks) {
    scan_ints_inplace(hist, (int64_t)kRadixBins * nblocks, [](int64_t, int64_t) {});
}

__global__ __launch_bounds__(kSliceThreads) void k_radix_scatter(const unsigned long long *kin, const int64_t *vin,
                                                                 int64_t n, int shift, const int *hist, int nblocks,
                                                                 unsigned long long *kout, int64_t *vout) {
    __shared__ int run[kRadixBins];
    __shared__ int wcnt[kSliceThreads / 64][kRadixBins];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = threadIdx.x; k < kRadixBins; k += blockDim.x) run[k] = 0;
    const int64_t base = (int64_t)blockIdx.x * kSliceBlock;
    for (int r = 0; r < kSliceBlock / kSliceThreads; ++r) {
        for (int k = threadIdx.x; k < (kSliceThreads / 64) * kRadixBins; k += blockDim.x) (&wcnt[0][0])[k] = 0;
        __syncthreads();
        const int64_t i = base + r * kSliceThreads + threadIdx.x;
        const bool valid = i < n;
        const unsigned long long key = valid ? kin[i] : 0ull;
        const int d = (int)((key >> shift) & (kRadixBins - 1));
        // the wave's lanes holding the same digit: agree with this lane on every digit bit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long m = __ballot(valid && ((d >> bit) & 1));
            peers &= ((d >> bit) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = hist[(size_t)d * nblocks + blockIdx.x] + run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kRadixBins; k += blockDim.x) {
            int add = 0;
            for (int w = 0; w < kSliceThreads / 64; ++w) add += wcnt[w][k];
            run[k] += add;
        }
        __syncthreads();
    }
}

hipError_t launch_slicing(const SlicingArgs &a, hipStream_t s) {
    const int nblocks = (int)((a.n + kSliceBlock - 1) / kSliceBlock);
    const size_t zlds = sizeof(double) * 2 * kZChunk * 64;
    hipFuncSetAttribute((const void *)k_zlanes, hipFuncAttributeMaxDynamicSharedMemorySize, (int)zlds);
    hipLaunchKernelGGL(k_zlanes, dim3(kZFoldBlocks), dim3(64 * kZWaves), zlds, s, a.z, a.n, a.tab, 1, a.lanes, a.mm_partial);
    hipLaunchKernelGGL(k_zfold, dim3(1), dim3(256), 0, s, a.lanes, a.mm_partial, a.n, a.tab, 1, a.method, a.ns, a.q, a.err);
    hipLaunchKernelGGL(k_zlanes, dim3(kZFoldBlocks), dim3(64 * kZWaves), zlds, s, a.z, a.n, a.tab, 2, a.lanes, a.mm_partial);
    hipLaunchKernelGGL(k_zfold, dim3(1), dim3(256), 0, s, a.lanes, a.mm_partial, a.n, a.tab, 2, a.method, a.ns, a.q, a.err);
    if (a.method == 0) {
        const int bins = a.ns * a.resolution;
        int g = (int)((a.n + 255) / 256);
        if (g > 4096) g = 4096;
        hipFuncSetAttribute((const void *)k_equal_area_hist, hipFuncAttributeMaxDynamicSharedMemorySize, bins * 4);
        hipLaunchKernelGGL(k_equal_area_hist, dim3(g), dim3(256), bins * 4, s, a.z, a.n, a.tab, bins, a.bin_counts);
        hipLaunchKernelGGL(k_equal_area, dim3(1), dim3(64), 0, s, a.tab, a.n, a.ns, bins, a.bin_counts, a.err);
    }
    if (a.method == 1) {
        int g = (int)((a.n + kSliceThreads - 1) / kSliceThreads);
        if (g > 4096) g = 4096;
        // keys and indices start in (rank_sorted, perm); the eight passes alternate with (rank_keys, rank_vals)
        // and end back in (rank_sorted, perm)
        hipLaunchKernelGGL(k_rank_keys, dim3(g), dim3(256), 0, s, a.z, a.n, a.rank_sorted, a.perm);
        for (int pass = 0; pass < 8; ++pass) {
            const bool even = (pass & 1) == 0;
            const unsigned long long *kin = even ? a.rank_sorted : a.rank_keys;
            const int64_t *vin = even ? a.perm : a.rank_vals;
            unsigned long long *kout = even ? a.rank_keys : a.rank_sorted;
            int64_t *vout = even ? a.rank_vals : a.perm;
            hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSliceThreads), 0, s, kin, a.n, 8 * pass, a.hist, nblocks);
            hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, a.hist, nblocks);
            hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSliceThreads), 0, s, kin, vin, a.n, 8 * pass, a.hist,
                               nblocks, kout, vout);
        }
        hipLaunchKernelGGL(k_rank_table, dim3(1), dim3(64), 0, s, a.tab, a.rank_sorted, a.n, a.ns);
        hipLaunchKernelGGL(k_rank_gather, dim3(g), dim3(kSliceThreads), 0, s, a.beam, a.n, a.perm, a.inv, a.st, a.box);
    } else {
        hipLaunchKernelGGL(k_classify, dim3(nblocks), dim3(kSliceThreads), 0, s, a.z, a.x, a.y, a.n, a.tab, a.sid, a.hist,
                           nblocks, a.method == 2 ? 1 : 0, a.box, a.err);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, a.hist, a.ns, nblocks, a.tab);
        hipLaunchKernelGGL(k_counts, dim3(1), dim3(256), 0, s, a.tab, a.n);
        hipLaunchKernelGGL(k_scatter, dim3(nblocks), dim3(kSliceThreads), 0, s, a.beam, a.n, a.sid, a.hist, nblocks, a.ns,
                           a.st, a.perm, a.inv);
    }
    hipLaunchKernelGGL(k_centroid_lanes, dim3(kFoldLanes / 256, a.ns), dim3(256), 0, s, a.st_z, a.tab, a.centroid_lanes);
    hipLaunchKernelGGL(k_centroid_fold, dim3(a.ns), dim3(256), 0, s, a.centroid_lanes, a.tab, a.n, a.centroid);
    return hipGetLastError();
}

int64_t slicing_hist_entries(int64_t n, int ns) { return ((n + kSliceBlock - 1) / kSliceBlock) * (int64_t)ns; }


}  // namespace octopic
