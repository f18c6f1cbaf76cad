// k_fft.hip -- the open-boundary Green convolution on the 2g-padded meshes, and the Green tables.
//
// Reference: _pic_solve_drifted_field_with_green_fft! (src/tasks/strongstrong/pic_cpu.jl:1965-1990)
//   phi = Re[ IFFT( FFT(Q) .* FFT(G) ) ][1:nx, 1:ny]
// with Q the charge zero-padded to (2nx, 2ny) and G the shifted Green table (pic_cpu.jl:2451-2509).
// MI355X form: Q and G are real, so every transform is a real-to-half-spectrum one (two real rows
// packed per complex FFT), the zero rows of Q are never transformed (row pass over ny rows only),
// and the rows above ny of the inverse are never produced:
//   rows_fwd   : ny real rows of Q (int64 fixed point, consumed and zeroed)   -> Qh[kx][y], kx <= nx
//   cols_conv  : per kx, FFT_y (zero-padded) * Gh[kx][ky], IFFT_y, keep y < ny -> Qh[kx][y]
//   rows_inv   : per row pair, Hermitian rebuild + IFFT_x, keep x < nx         -> phi[y][x]
//   field      : E = -grad(phi), 2nd/4th order (pic_cpu.jl:2531-2591)           -> table[x][y][3 per plane]
// Green spectra are built the same way from the full 2nx x 2ny real table (green_rows, green_cols).
#include <algorithm>

#include "common.h"
#include "fft.h"
#include "launch.h"

namespace octopic {

constexpr int kFFTLanes = 64;        // lanes per transform
constexpr int kFFTPerBlock = 4;      // transforms per workgroup
constexpr int kFFTThreads = kFFTLanes * kFFTPerBlock;

__device__ __forceinline__ bool failed(const ErrorInfo *err) { return *((volatile const int32_t *)&err->code) != 0; }

// ---------------------------------------------------------------------- Green kernel (pic_cpu.jl:2254-2272)
__device__ __forceinline__ double atan_ratio(double num, double den) {
    if (den == 0.0) {
        if (num == 0.0) return 0.0;
        return copysign(1.5707963267948966, num);
    }
    return atan(num / den);
}
__device__ __forceinline__ double kernel_integral(double x, double y) {
    double r2 = x * x + y * y;
    if (r2 == 0.0) return 0.0;
    return (log(r2) - 3.0) * x * y + atan_ratio(y, x) * x * x + atan_ratio(x, y) * y * y;
}
// _pic_green_columns! cell (i, j) of the (2nx, 2ny) table
__device__ __forceinline__ double green_cell(int i, int j, int nx, int ny, int integrated, const GreenJob &g) {
    int ii = i < nx ? i : i - 2 * nx;
    int jj = j < ny ? j : j - 2 * ny;
    double x = g.fx0 - g.sx0 + (double)ii * g.hx;
    double y = g.fy0 - g.sy0 + (double)jj * g.hy;
    double half_hx = g.hx / 2.0, half_hy = g.hy / 2.0;
    double hxihyi = -0.5 / (g.hx * g.hy);
    if (integrated || (x * x + y * y) == 0.0) {
        double val = kernel_integral(x + half_hx, y + half_hy);
        val += kernel_integral(x - half_hx, y - half_hy);
        val -= kernel_integral(x + half_hx, y - half_hy);
        val -= kernel_integral(x - half_hx, y + half_hy);
        return hxihyi * val;
    }
    return -0.5 * log(x * x + y * y);
}

// Unpack the FFT Z of (a + i b), a and b real, into A[k] (odd = 0) or B[k] (odd = 1), k in [0, P/2].
template <class C>
__device__ __forceinline__ C unpack_one(const C *Z, int P, int k, int odd) {
    using R = real_of<C>;
    const C z = Z[k];
    const C zc = cconj(Z[k == 0 ? 0 : P - k]);
    if (!odd) return cx<C>((R)0.5 * (z.x + zc.x), (R)0.5 * (z.y + zc.y));
    const C d = csub(z, zc);
    return cx<C>((R)0.5 * d.y, (R)-0.5 * d.x);   // d / (2i)
}

// Dynamic LDS of every transform kernel: the twiddle table (P entries), then per wave group two
// P-entry buffers a, b.
template <class C>
__device__ __forceinline__ C *group_buffer(C *lds, int P, int grp) { return lds + P + (size_t)grp * 2 * P; }

// Row-pass epilogue shared by the Green and charge row transforms: the block's 2*G rows (G transforms
// of two real rows each; ya0 + r, r = 2*grp + {0, 1}) of the x-half-spectrum, stored as
// out[k * kstride + ya0 + r] so that the 2*G consecutive rows of one k form one contiguous run.
template <int G, class C>
__device__ __forceinline__ void store_row_spectra(C *lds, const FFTPlanDev &plan, int nrows, int ya0, C *out,
                                                  int64_t kstride) {
    constexpr int R = 2 * G;
    const int P = plan.n, nk = P / 2 + 1;
    for (int t = threadIdx.x; t < nk * R; t += blockDim.x) {
        const int r = t % R, k = t / R;
        if (ya0 + r >= nrows) continue;
        C *ga = group_buffer(lds, P, r >> 1);
        const C *Z = fft_result(ga, ga + P, plan);
        out[(size_t)k * kstride + ya0 + r] = unpack_one(Z, P, k, r & 1);
    }
}

// The jobs of a Green launch that build (valid, slot in range), in job order, compacted into `list` by
// every block of the launch the same way (ballot per wave, then the waves in order); returns their
// number.  Called by all threads of the block; the list is visible to all of them on return.
constexpr int kMaxGreenJobs = 2048;
constexpr int kGreenSmallGrid = 64;   // blocks of a Green launch after a collide that built little
__device__ __forceinline__ int compact_green_jobs(const GreenJob *jobs, int njobs, int nslots, int *list, int *wcount) {
    int total = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    for (int base = 0; base < njobs; base += blockDim.x) {
        const int k = base + threadIdx.x;
        bool v = false;
        if (k < njobs) {
            const GreenJob &j = jobs[k];
            v = j.valid && j.slot >= 0 && j.slot < nslots;
        }
        const unsigned long long m = __ballot(v);
        if (lane == 0) wcount[wave] = __popcll(m);
        __syncthreads();
        int off = total, add = 0;
        for (int w = 0; w < nwaves; ++w) {
            if (w < wave) off += wcount[w];
            add += wcount[w];
        }
        if (v) list[off + __popcll(m & ((1ull << lane) - 1ull))] = k;
        __syncthreads();
        total += add;
    }
    return total;
}

// Green rows: rows (ya, ya+1) of the real table -> x-half-spectrum into slot[kx][y], y in [0, 2ny).
// Items are (building job, block of 2 * kFFTPerBlock rows); the blocks of the launch take them in
// turn, so a launch whose batch builds nothing is one pass of its blocks over the job flags.
__global__ __launch_bounds__(kFFTThreads) void k_green_rows(const GreenJob *jobs, int njobs, double2 *spec, int64_t slot_stride,
                                                            int nslots, int nx, int ny, int integrated, FFTPlanDev px,
                                                            const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    __shared__ int list[kMaxGreenJobs];
    __shared__ int wcount[kFFTThreads / 64];
    if (failed(err)) return;
    const int nbuild = compact_green_jobs(jobs, njobs, nslots, list, wcount);
    if (nbuild == 0) return;
    const int P1 = 2 * nx, P2 = 2 * ny;
    const int rblocks = (P2 + 2 * kFFTPerBlock - 1) / (2 * kFFTPerBlock);
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = group_buffer(lds, P1, grp);
    load_twiddles(px, lds);
    for (int item = blockIdx.x; item < nbuild * rblocks; item += gridDim.x) {
        const GreenJob job = jobs[list[item / rblocks]];
        const int ya0 = 2 * kFFTPerBlock * (item % rblocks), ya = ya0 + 2 * grp;
        for (int x = lane; x < P1; x += kFFTLanes) {
            double va = 0.0, vb = 0.0;
            if (ya < P2) {
                va = green_cell(x, ya, nx, ny, integrated, job);
                vb = green_cell(x, ya + 1, nx, ny, integrated, job);
            }
            a[x] = make_double2(va, vb);
        }
        __syncthreads();
        fft_lds<-1>(a, a + P1, px, lds, lane);
        __syncthreads();
        store_row_spectra<kFFTPerBlock>(lds, px, P2, ya0, spec + (size_t)job.slot * slot_stride, P2);
        __syncthreads();   // the item's buffers are read before the next item writes them
    }
}

// Green columns: in-place FFT along y of slot[kx][0..2ny); items are (building job, block of
// kFFTPerBlock columns).
// FP32 mode (spec32 set): each freshly built slot's FP64 half spectrum is also stored rounded to float, with a
// reference value Gref of the table taken out of every cell first (the DC term loses Gref * P1 * P2).
// A constant in G moves every potential of a plane by that constant times the plane's charge: it leaves
// E = -grad(phi) unchanged, and phi_L - phi_R too (the L, R and M planes of an interaction deposit the
// same particles on the same mesh), but it is the bulk of |phi|, and FP32 differences of large
// potentials lose digits of both kicks: the FP32 error of E is about eps * max|phi - Gref Q| / (h |E|).
// Gref is the table at cell (nx/4, ny/4), i.e. G a quarter of the mesh away from the source-field
// offset.  The potentials over the mesh lie between Q G(~sigma) near the source and Q G(~diagonal) at
// the far corners; G at a quarter mesh sits near the middle of that range in log distance, so
// max|phi - Gref Q| stays within about 1.4 Q for any offset (the cell (0, 0), which is the near field
// when the two meshes nearly coincide, leaves up to 4.6 Q).
__global__ __launch_bounds__(kFFTThreads) void k_green_cols(const GreenJob *jobs, int njobs, double2 *spec, float2 *spec32,
                                                            int64_t slot_stride, int nslots, int nx, int ny, int integrated,
                                                            FFTPlanDev py, const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    __shared__ int list[kMaxGreenJobs];
    __shared__ int wcount[kFFTThreads / 64];
    if (failed(err)) return;
    const int nbuild = compact_green_jobs(jobs, njobs, nslots, list, wcount);
    if (nbuild == 0) return;
    const int P2 = 2 * ny;
    const int cblocks = (nx + 1 + kFFTPerBlock - 1) / kFFTPerBlock;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = group_buffer(lds, P2, grp);
    load_twiddles(py, lds);
    for (int item = blockIdx.x; item < nbuild * cblocks; item += gridDim.x) {
        const GreenJob job = jobs[list[item / cblocks]];
        const int kx = (item % cblocks) * kFFTPerBlock + grp;
        const bool active = kx <= nx;
        double2 *col = spec + (size_t)job.slot * slot_stride + (size_t)(active ? kx : 0) * P2;
        for (int y = lane; y < P2; y += kFFTLanes) a[y] = active ? col[y] : make_double2(0.0, 0.0);
        __syncthreads();
        fft_lds<-1>(a, a + P2, py, lds, lane);
        if (active) {
            const double2 *Z = fft_result(a, a + P2, py);
            float2 *col32 = spec32 ? spec32 + (size_t)job.slot * slot_stride + (size_t)kx * P2 : nullptr;
            for (int y = lane; y < P2; y += kFFTLanes) {
                const double2 z = Z[y];
                col[y] = z;
                if (col32) {
                    double re = z.x;
                    if (kx == 0 && y == 0)
                        re -= green_cell(nx / 4, ny / 4, nx, ny, integrated, job) * (double)(2 * nx) * (double)(2 * ny);
                    col32[y] = make_float2((float)re, (float)z.y);
                }
            }
        }
        __syncthreads();   // the item's buffers are read before the next item writes them
    }
}

// Luminosity overlap (pic_cpu.jl:2731-2735), one of kLumBlocks blocks of pair k: the block sums q1.q2
// over a fixed cell range of the pair's two luminosity planes (resetting them to zero for the next
// batch) into its partial; the column launch that follows adds the kLumBlocks partials in block order
// (lum_fold), so the result does not depend on scheduling and no block waits on another (a device-wide
// release per block costs an L2 write-back on this chip).  Called by all threads of a block.
// The block's sum has a fixed shape whatever its thread count: kLumLanes lane sums (lane v takes cells
// c0 + v, c0 + v + kLumLanes, ...), then a fixed tree, so the partial does not depend on the transforms
// per workgroup of the launch it rides in.
constexpr int kLumLanes = 256;
__device__ __forceinline__ void lum_block(const LumArgs &a, int k, int b) {
    unsigned long long *q1 = a.charge + (size_t)(2 * k) * a.charge_stride;
    unsigned long long *q2 = a.charge + (size_t)(2 * k + 1) * a.charge_stride;
    const double inv = a.inv_scale;
    const int c0 = (int)((int64_t)a.ncells * b / kLumBlocks), c1 = (int)((int64_t)a.ncells * (b + 1) / kLumBlocks);
    __shared__ double red[kLumLanes];
    for (int v = threadIdx.x; v < kLumLanes; v += blockDim.x) {
        double acc = 0.0;
        for (int c = c0 + v; c < c1; c += kLumLanes) {
            const unsigned long long u1 = q1[c], u2 = q2[c];
            if (u1) q1[c] = 0ull;
            if (u2) q2[c] = 0ull;
            acc += ((double)u1 * inv) * ((double)u2 * inv);
        }
        red[v] = acc;
    }
    __syncthreads();
    for (int w = kLumLanes / 2; w > 0; w >>= 1) {
        for (int v = threadIdx.x; v < w; v += blockDim.x) red[v] += red[v + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.lum_partial[k * kLumBlocks + b] = red[0];
}

// The batch's per-pair luminosity from the partials of the row launch before (thread k: pair k).
__device__ __forceinline__ void lum_fold(const LumArgs &a) {
    for (int k = threadIdx.x; k < a.npairs; k += blockDim.x) {
        const double *part = a.lum_partial + k * kLumBlocks;
        double sum = 0.0;
        for (int q = 0; q < kLumBlocks; ++q) sum += part[q];
        const double *g = a.lumgeo + 3 * k;
        a.lum_parts[a.pairs[k].p] = sum * g[2] * g[0] * g[1];
    }
}

// Charge rows: int64 fixed-point rows (ya, ya+1) of plane -> x-half-spectrum Qh[plane][kx][y], y < ny.
// Each charge cell is read once and reset to zero for the next use of the buffer.  Block rows past the
// planes take the batch's luminosity overlap sums (lum.npairs pairs, kLumBlocks blocks each), which
// read only the luminosity planes the deposit has finished.  G transforms (2*G rows) per block.
template <class C, int G>
__global__ __launch_bounds__(kFFTLanes * G) void k_rows_fwd(const Plane *planes, unsigned long long *charge, int64_t charge_stride,
                                                          C *qh, int64_t qh_stride, int nx, int ny, double inv_scale,
                                                          FFTPlanDev px, const ErrorInfo *err, int nplanes, LumArgs lum) {
    using R = real_of<C>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *lds = reinterpret_cast<C *>(lds_raw);
    if ((int)blockIdx.y >= nplanes) {
        const int t = ((int)blockIdx.y - nplanes) * (int)gridDim.x + (int)blockIdx.x;
        if (t < lum.npairs * kLumBlocks) lum_block(lum, t / kLumBlocks, t % kLumBlocks);
        return;
    }
    if (failed(err)) return;
    const int plane = blockIdx.y;
    const int P1 = 2 * nx;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    C *a = group_buffer(lds, P1, grp);
    const int ya0 = 2 * G * blockIdx.x, ya = ya0 + 2 * grp;
    const bool active = ya < ny;
    const bool hasb = ya + 1 < ny;
    unsigned long long *q = charge + (size_t)plane * charge_stride;
    load_twiddles(px, lds);
    for (int x = lane; x < P1; x += kFFTLanes) {
        double va = 0.0, vb = 0.0;
        if (active && x < nx) {
            const unsigned long long ua = q[(size_t)ya * nx + x];
            va = (double)(long long)ua * inv_scale;   // signed: denoised planes may hold negative cells
            if (ua) q[(size_t)ya * nx + x] = 0ull;
            if (hasb) {
                const unsigned long long ub = q[(size_t)(ya + 1) * nx + x];
                vb = (double)(long long)ub * inv_scale;
                if (ub) q[(size_t)(ya + 1) * nx + x] = 0ull;
            }
        }
        a[x] = cx<C>((R)va, (R)vb);
    }
    __syncthreads();
    fft_lds<-1>(a, a + P1, px, lds, lane);
    __syncthreads();
    store_row_spectra<G>(lds, px, ny, ya0, qh + (size_t)plane * qh_stride, ny);
}

// Column pass: FFT_y of the zero-padded column, times the Green half spectrum, IFFT_y, keep y < ny.
// (One plane per block row: sharing a Green column across an interaction's planes in one block
// measured slower, the lost parallelism costing more than the saved spectrum reads.)  G columns per block.
template <class C, int G>
__global__ __launch_bounds__(kFFTLanes * G) void k_cols_conv(const Plane *planes, C *qh, int64_t qh_stride,
                                                           const C *gspec, int64_t slot_stride, int nx, int ny,
                                                           FFTPlanDev py, const ErrorInfo *err, LumArgs lum) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *lds = reinterpret_cast<C *>(lds_raw);
    if (failed(err)) return;
    if (blockIdx.x == 0 && blockIdx.y == 0 && lum.npairs > 0) lum_fold(lum);
    const int plane = blockIdx.y;
    const int P2 = 2 * ny;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    C *a = group_buffer(lds, P2, grp), *b = a + P2;
    const int kx = blockIdx.x * G + grp;
    const bool active = kx <= nx;
    C *col = qh + (size_t)plane * qh_stride + (size_t)(active ? kx : 0) * ny;
    load_twiddles(py, lds);
    for (int y = lane; y < P2; y += kFFTLanes) a[y] = (active && y < ny) ? col[y] : cx<C>(0, 0);
    __syncthreads();
    fft_lds<-1>(a, b, py, lds, lane);
    C *Z = fft_result(a, b, py), *W = Z == a ? b : a;
    const int slot = planes[plane].slot;
    const C *g = gspec + (size_t)(slot < 0 ? 0 : slot) * slot_stride + (size_t)(active ? kx : 0) * P2;
    for (int y = lane; y < P2; y += kFFTLanes) Z[y] = cmul(Z[y], g[y]);
    wave_lds_sync();
    fft_lds<+1>(Z, W, py, lds, lane);
    if (!active) return;
    const C *Rs = fft_result(Z, W, py);
    for (int y = lane; y < ny; y += kFFTLanes) col[y] = Rs[y];
}

// Row inverse: rows (ya, ya+1) from the half spectrum, IFFT_x, keep x < nx, scale 1/(P1 P2).  The
// block gathers its 2*G rows together: thread (kk, grp) reads rows ya, ya+1 of column kk (adjacent
// threads adjacent rows) and fills both kk and its Hermitian mirror P1 - kk.
template <class C, int G>
__global__ __launch_bounds__(kFFTLanes * G) void k_rows_inv(const Plane *planes, const C *qh, int64_t qh_stride,
                                                          real_of<C> *phi, int nx, int ny, FFTPlanDev px,
                                                          const ErrorInfo *err) {
    using R = real_of<C>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *lds = reinterpret_cast<C *>(lds_raw);
    if (failed(err)) return;
    const int plane = blockIdx.y;
    const int P1 = 2 * nx, P2 = 2 * ny;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    const int ya0 = 2 * G * blockIdx.x;
    const C *in = qh + (size_t)plane * qh_stride;
    load_twiddles(px, lds);
    for (int t = threadIdx.x; t < (nx + 1) * G; t += blockDim.x) {
        const int g = t % G, kk = t / G;
        const int y = ya0 + 2 * g;
        C A = cx<C>(0, 0), B = A;
        if (y < ny) A = in[(size_t)kk * ny + y];
        if (y + 1 < ny) B = in[(size_t)kk * ny + y + 1];
        C *ag = group_buffer(lds, P1, g);
        ag[kk] = cx<C>(A.x - B.y, A.y + B.x);   // A + i B
        if (kk > 0 && kk < nx) {
            A = cconj(A);
            B = cconj(B);
            ag[P1 - kk] = cx<C>(A.x - B.y, A.y + B.x);
        }
    }
    __syncthreads();
    C *a = group_buffer(lds, P1, grp);
    fft_lds<+1>(a, a + P1, px, lds, lane);
    const int ya = ya0 + 2 * grp;
    if (ya >= ny) return;
    const bool hasb = ya + 1 < ny;
    const C *Z = fft_result(a, a + P1, px);
    const R norm = (R)(1.0 / ((double)P1 * (double)P2));
    R *out = phi + (size_t)plane * nx * ny;
    for (int x = lane; x < nx; x += kFFTLanes) {
        out[(size_t)ya * nx + x] = Z[x].x * norm;
        if (hasb) out[(size_t)(ya + 1) * nx + x] = Z[x].y * norm;
    }
}

// E = -grad(phi) of every plane of an interaction into the interleaved kick table
// table[inter][x][y][3*q + {phi, Ex, Ey}] (pic_cpu.jl:2531-2591; h of the plane's deposit mesh).
// A block takes a kFieldTile x kFieldTile cell tile of one interaction: phi with a 2-cell halo is
// staged in LDS from contiguous rows of phi[y][x], and the tile's table entries are written back as
// one contiguous run per x.
constexpr int kFieldTile = 16;
constexpr int kFieldHalo = kFieldTile + 4;

// R = double (FP64 mode) or float (FP32 mode: phi, derivative and table in FP32).
template <class R>
__global__ __launch_bounds__(256) void k_field(const Inter *inters, const Plane *planes, const R *phi, R *table,
                                               int nx, int ny, int fourth, int npl, const ErrorInfo *err) {
    if (failed(err)) return;
    const int d = blockIdx.y;
    const Inter it = inters[d];
    if (!it.valid) return;
    __shared__ R ph[kFieldHalo][kFieldHalo + 1];                      // [y][x], origin (tx0 - 2, ty0 - 2)
    __shared__ R stage[kFieldTile * kFieldTile * 9];                  // [lx][ly][3q + c]
    const int ntx = (nx + kFieldTile - 1) / kFieldTile;
    const int tx0 = (blockIdx.x % ntx) * kFieldTile, ty0 = (blockIdx.x / ntx) * kFieldTile;
    const int C = 3 * npl;
    const int lx = threadIdx.x % kFieldTile, ly = threadIdx.x / kFieldTile;
    const int x = tx0 + lx, y = ty0 + ly;
    const bool inside = x < nx && y < ny;
    for (int q = 0; q < npl; ++q) {
        const bool live = q < it.nplanes;
        if (live) {
            const R *P = phi + (size_t)(it.plane0 + q) * nx * ny;   // [y][x]
            for (int k = threadIdx.x; k < kFieldHalo * kFieldHalo; k += blockDim.x) {
                const int hy = k / kFieldHalo, hx = k % kFieldHalo;
                const int gx = tx0 - 2 + hx, gy = ty0 - 2 + hy;
                ph[hy][hx] = (gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? P[(size_t)gy * nx + gx] : (R)0;
            }
        }
        __syncthreads();
        R v0 = 0, v1 = 0, v2 = 0;
        if (live && inside) {
            const Plane &pl = planes[it.plane0 + q];
            auto F = [&](int i, int j) { return ph[j - ty0 + 2][i - tx0 + 2]; };
            field_node<R>(F, x, y, nx, ny, fourth, (R)(1.0 / pl.hx), (R)(1.0 / pl.hy), v0, v1, v2);
        }
        R *sv = stage + (lx * kFieldTile + ly) * 9 + 3 * q;
        sv[0] = v0; sv[1] = v1; sv[2] = v2;
        __syncthreads();
    }
    // one contiguous run of kFieldTile * C doubles per x of the tile
    const int ylen = (ny - ty0) < kFieldTile ? (ny - ty0) : kFieldTile;
    const int run = ylen * C;
    for (int k = threadIdx.x; k < kFieldTile * run; k += blockDim.x) {
        const int tlx = k / run, rem = k - tlx * run;
        const int tly = rem / C, comp = rem - tly * C;
        const int gx = tx0 + tlx;
        if (gx >= nx) continue;
        table[(((size_t)d * nx + gx) * ny + ty0 + tly) * C + comp] = stage[(tlx * kFieldTile + tly) * 9 + comp];
    }
}

// ---------------------------------------------------------------------- launchers
static size_t fft_lds_bytes(int P, int G = kFFTPerBlock) { return (size_t)(1 + G * 2) * P * sizeof(double2); }

static hipError_t ensure_lds(const void *fn, size_t bytes) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Transforms per block of the charge-path passes: the requested 2, 4 or 8, halved while its dynamic LDS (the
// twiddles and two buffers per transform, of transform length P and element size `es`) would not leave
// 4 KB of the 160 KB for the kernels' static LDS.
static int fft_group(int requested, int P, size_t es) {
    int g = requested == 2 || requested == 8 ? requested : kFFTPerBlock;
    while (g > 2 && (size_t)(1 + 2 * g) * P * es > (size_t)156 * 1024) g /= 2;
    return g;
}

// `full` launches a block per item of every job (the cold case: every job builds); otherwise a small
// persistent grid takes the items of the few jobs that build, so batches of cache hits cost one short
// pass.  Results do not depend on the grid.
hipError_t launch_green_build(const GreenJob *jobs, int njobs, double2 *spec, float2 *spec32, int64_t slot_stride, int nslots,
                              int nx, int ny, int integrated, const FFTPlanDev &px, const FFTPlanDev &py, const ErrorInfo *err,
                              hipStream_t s, int full) {
    if (njobs <= 0) return hipSuccess;
    for (; njobs > kMaxGreenJobs; jobs += kMaxGreenJobs, njobs -= kMaxGreenJobs) {   // each launch compacts its own jobs
        const hipError_t e = launch_green_build(jobs, kMaxGreenJobs, spec, spec32, slot_stride, nslots, nx, ny, integrated, px,
                                                py, err, s, full);
        if (e != hipSuccess) return e;
    }
    const int P1 = 2 * nx, P2 = 2 * ny;
    ensure_lds((const void *)k_green_rows, fft_lds_bytes(P1));
    ensure_lds((const void *)k_green_cols, fft_lds_bytes(P2));
    const int rblocks = (P2 + 2 * kFFTPerBlock - 1) / (2 * kFFTPerBlock), cblocks = (nx + 1 + kFFTPerBlock - 1) / kFFTPerBlock;
    const int g1 = full ? njobs * rblocks : std::min(njobs * rblocks, kGreenSmallGrid);
    const int g2 = full ? njobs * cblocks : std::min(njobs * cblocks, kGreenSmallGrid);
    hipLaunchKernelGGL(k_green_rows, dim3(g1), dim3(kFFTThreads), fft_lds_bytes(P1), s, jobs, njobs, spec, slot_stride, nslots,
                       nx, ny, integrated, px, err);
    hipLaunchKernelGGL(k_green_cols, dim3(g2), dim3(kFFTThreads), fft_lds_bytes(P2), s, jobs, njobs, spec, spec32, slot_stride,
                       nslots, nx, ny, integrated, py, err);
    return hipGetLastError();
}

template <class C, int G>
static void rows_fwd_g(const Plane *planes, int nplanes, unsigned long long *charge, int64_t charge_stride, C *qh,
                       int64_t qh_stride, int nx, int ny, double inv_scale, const FFTPlanDev &px, const ErrorInfo *err,
                       const LumArgs &lum, hipStream_t s) {
    const size_t lds = (size_t)(1 + 2 * G) * (2 * nx) * sizeof(C);
    ensure_lds((const void *)k_rows_fwd<C, G>, lds);
    const int gx = (ny + 2 * G - 1) / (2 * G);
    const int lum_rows = (lum.npairs * kLumBlocks + gx - 1) / gx;
    dim3 g(gx, nplanes + lum_rows);
    hipLaunchKernelGGL((k_rows_fwd<C, G>), g, dim3(kFFTLanes * G), lds, s, planes, charge, charge_stride, qh, qh_stride, nx,
                       ny, inv_scale, px, err, nplanes, lum);
}

template <class C>
hipError_t launch_rows_fwd(const Plane *planes, int nplanes, unsigned long long *charge, int64_t charge_stride, C *qh,
                           int64_t qh_stride, int nx, int ny, double inv_scale, const FFTPlanDev &px, const ErrorInfo *err,
                           const LumArgs &lum, hipStream_t s, int group) {
    if (nplanes <= 0 && lum.npairs <= 0) return hipSuccess;
    switch (fft_group(group, 2 * nx, sizeof(C))) {
        case 2: rows_fwd_g<C, 2>(planes, nplanes, charge, charge_stride, qh, qh_stride, nx, ny, inv_scale, px, err, lum, s); break;
        case 8: rows_fwd_g<C, 8>(planes, nplanes, charge, charge_stride, qh, qh_stride, nx, ny, inv_scale, px, err, lum, s); break;
        default: rows_fwd_g<C, 4>(planes, nplanes, charge, charge_stride, qh, qh_stride, nx, ny, inv_scale, px, err, lum, s);
    }
    return hipGetLastError();
}

template <class C, int G>
static void cols_conv_g(const Plane *planes, int nplanes, C *qh, int64_t qh_stride, const C *gspec, int64_t slot_stride,
                        int nx, int ny, const FFTPlanDev &py, const ErrorInfo *err, const LumArgs &lum, hipStream_t s) {
    const size_t lds = (size_t)(1 + 2 * G) * (2 * ny) * sizeof(C);
    ensure_lds((const void *)k_cols_conv<C, G>, lds);
    dim3 g((nx + 1 + G - 1) / G, nplanes);
    hipLaunchKernelGGL((k_cols_conv<C, G>), g, dim3(kFFTLanes * G), lds, s, planes, qh, qh_stride, gspec, slot_stride, nx, ny,
                       py, err, lum);
}

template <class C>
hipError_t launch_cols_conv(const Plane *planes, int nplanes, C *qh, int64_t qh_stride, const C *gspec, int64_t slot_stride,
                            int nx, int ny, const FFTPlanDev &py, const ErrorInfo *err, const LumArgs &lum, hipStream_t s,
                            int group) {
    if (nplanes <= 0) return hipSuccess;
    switch (fft_group(group, 2 * ny, sizeof(C))) {
        case 2: cols_conv_g<C, 2>(planes, nplanes, qh, qh_stride, gspec, slot_stride, nx, ny, py, err, lum, s); break;
        case 8: cols_conv_g<C, 8>(planes, nplanes, qh, qh_stride, gspec, slot_stride, nx, ny, py, err, lum, s); break;
        default: cols_conv_g<C, 4>(planes, nplanes, qh, qh_stride, gspec, slot_stride, nx, ny, py, err, lum, s);
    }
    return hipGetLastError();
}

template <class C, int G>
static void rows_inv_g(const Plane *planes, int nplanes, const C *qh, int64_t qh_stride, real_of<C> *phi, int nx, int ny,
                       const FFTPlanDev &px, const ErrorInfo *err, hipStream_t s) {
    const size_t lds = (size_t)(1 + 2 * G) * (2 * nx) * sizeof(C);
    ensure_lds((const void *)k_rows_inv<C, G>, lds);
    dim3 g((ny + 2 * G - 1) / (2 * G), nplanes);
    hipLaunchKernelGGL((k_rows_inv<C, G>), g, dim3(kFFTLanes * G), lds, s, planes, qh, qh_stride, phi, nx, ny, px, err);
}

template <class C>
hipError_t launch_rows_inv(const Plane *planes, int nplanes, const C *qh, int64_t qh_stride, real_of<C> *phi, int nx,
                           int ny, const FFTPlanDev &px, const ErrorInfo *err, hipStream_t s, int group) {
    if (nplanes <= 0) return hipSuccess;
    switch (fft_group(group, 2 * nx, sizeof(C))) {
        case 2: rows_inv_g<C, 2>(planes, nplanes, qh, qh_stride, phi, nx, ny, px, err, s); break;
        case 8: rows_inv_g<C, 8>(planes, nplanes, qh, qh_stride, phi, nx, ny, px, err, s); break;
        default: rows_inv_g<C, 4>(planes, nplanes, qh, qh_stride, phi, nx, ny, px, err, s);
    }
    return hipGetLastError();
}

template <class R>
hipError_t launch_field(const Inter *inters, int ninters, const Plane *planes, const R *phi, R *table, int nx, int ny,
                        int fourth, int npl, const ErrorInfo *err, hipStream_t s) {
    if (ninters <= 0) return hipSuccess;
    const int ntiles = ((nx + kFieldTile - 1) / kFieldTile) * ((ny + kFieldTile - 1) / kFieldTile);
    hipLaunchKernelGGL(k_field<R>, dim3(ntiles, ninters), dim3(kFieldTile * kFieldTile), 0, s, inters, planes, phi, table,
                       nx, ny, fourth, npl, err);
    return hipGetLastError();
}

#define OCTOPIC_CONV_INSTANCES(C)                                                                                     \
    template hipError_t launch_rows_fwd<C>(const Plane *, int, unsigned long long *, int64_t, C *, int64_t, int, int,  \
                                           double, const FFTPlanDev &, const ErrorInfo *, const LumArgs &, hipStream_t, \
                                           int);                                                                       \
    template hipError_t launch_cols_conv<C>(const Plane *, int, C *, int64_t, const C *, int64_t, int, int,            \
                                            const FFTPlanDev &, const ErrorInfo *, const LumArgs &, hipStream_t, int); \
    template hipError_t launch_rows_inv<C>(const Plane *, int, const C *, int64_t, real_of<C> *, int, int,             \
                                           const FFTPlanDev &, const ErrorInfo *, hipStream_t, int);                   \
    template hipError_t launch_field<real_of<C>>(const Inter *, int, const Plane *, const real_of<C> *, real_of<C> *, \
                                                 int, int, int, int, const ErrorInfo *, hipStream_t);
OCTOPIC_CONV_INSTANCES(double2)
OCTOPIC_CONV_INSTANCES(float2)
#undef OCTOPIC_CONV_INSTANCES

}  // namespace octopic
