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

// Unpack the FFT Z of (a + i b), a and b real, into A[k] and B[k] for k in [0, P/2].
__device__ __forceinline__ void unpack_pair(const double2 *Z, int P, int k, double2 &A, double2 &B) {
    double2 z = Z[k];
    double2 zc = cconj(Z[k == 0 ? 0 : P - k]);
    A = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y + zc.y));
    double2 d = csub(z, zc);
    B = make_double2(0.5 * d.y, -0.5 * d.x);   // d / (2i)
}

// Green rows: rows (ya, ya+1) of the real table -> x-half-spectrum into slot[kx][y], y in [0, 2ny).
__global__ __launch_bounds__(kFFTThreads) void k_green_rows(const GreenJob *jobs, int njobs, double2 *spec, int64_t slot_stride,
                                                            int nslots, int nx, int ny, int integrated, FFTPlanDev px,
                                                            const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    if (failed(err)) return;
    const GreenJob job = jobs[blockIdx.y];
    if (!job.valid || job.slot < 0 || job.slot >= nslots) return;
    const int P1 = 2 * nx, P2 = 2 * ny;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = lds + (size_t)grp * 2 * P1;
    double2 *b = a + P1;
    const int ya = 2 * (blockIdx.x * kFFTPerBlock + grp);
    const bool active = ya < P2;
    for (int x = lane; x < P1; x += kFFTLanes) {
        double va = 0.0, vb = 0.0;
        if (active) {
            va = green_cell(x, ya, nx, ny, integrated, job);
            vb = green_cell(x, ya + 1, nx, ny, integrated, job);
        }
        a[x] = make_double2(va, vb);
    }
    __syncthreads();
    double2 *Z = fft_lds<-1>(a, b, px, lane, kFFTLanes);
    if (!active) return;
    double2 *out = spec + (size_t)job.slot * slot_stride;
    for (int k = lane; k <= nx; k += kFFTLanes) {
        double2 A, B;
        unpack_pair(Z, P1, k, A, B);
        out[(size_t)k * P2 + ya] = A;
        out[(size_t)k * P2 + ya + 1] = B;
    }
}

// Green columns: in-place FFT along y of slot[kx][0..2ny).
__global__ __launch_bounds__(kFFTThreads) void k_green_cols(const GreenJob *jobs, int njobs, double2 *spec, int64_t slot_stride,
                                                            int nslots, int nx, int ny, FFTPlanDev py, const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    if (failed(err)) return;
    const GreenJob job = jobs[blockIdx.y];
    if (!job.valid || job.slot < 0 || job.slot >= nslots) return;
    const int P2 = 2 * ny;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = lds + (size_t)grp * 2 * P2;
    double2 *b = a + P2;
    const int kx = blockIdx.x * kFFTPerBlock + grp;
    const bool active = kx <= nx;
    double2 *col = spec + (size_t)job.slot * slot_stride + (size_t)(active ? kx : 0) * P2;
    for (int y = lane; y < P2; y += kFFTLanes) a[y] = active ? col[y] : make_double2(0.0, 0.0);
    __syncthreads();
    double2 *Z = fft_lds<-1>(a, b, py, lane, kFFTLanes);
    if (!active) return;
    for (int y = lane; y < P2; y += kFFTLanes) col[y] = Z[y];
}

// Charge rows: int64 fixed-point rows (ya, ya+1) of plane -> x-half-spectrum Qh[plane][kx][y], y < ny.
// Each charge cell is read once and reset to zero for the next use of the buffer.
__global__ __launch_bounds__(kFFTThreads) void k_rows_fwd(const Plane *planes, unsigned long long *charge, int64_t charge_stride,
                                                          double2 *qh, int64_t qh_stride, int nx, int ny, double inv_scale,
                                                          FFTPlanDev px, const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    if (failed(err)) return;
    const int plane = blockIdx.y;
    const int P1 = 2 * nx;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = lds + (size_t)grp * 2 * P1;
    double2 *b = a + P1;
    const int ya = 2 * (blockIdx.x * kFFTPerBlock + grp);
    const bool active = ya < ny;
    const bool hasb = ya + 1 < ny;
    unsigned long long *q = charge + (size_t)plane * charge_stride;
    for (int x = lane; x < P1; x += kFFTLanes) {
        double va = 0.0, vb = 0.0;
        if (active && x < nx) {
            unsigned long long ua = q[(size_t)ya * nx + x];
            va = (double)ua * inv_scale;
            if (ua) q[(size_t)ya * nx + x] = 0ull;
            if (hasb) {
                unsigned long long ub = q[(size_t)(ya + 1) * nx + x];
                vb = (double)ub * inv_scale;
                if (ub) q[(size_t)(ya + 1) * nx + x] = 0ull;
            }
        }
        a[x] = make_double2(va, vb);
    }
    __syncthreads();
    double2 *Z = fft_lds<-1>(a, b, px, lane, kFFTLanes);
    if (!active) return;
    double2 *out = qh + (size_t)plane * qh_stride;
    for (int k = lane; k <= nx; k += kFFTLanes) {
        double2 A, B;
        unpack_pair(Z, P1, k, A, B);
        out[(size_t)k * ny + ya] = A;
        if (hasb) out[(size_t)k * ny + ya + 1] = B;
    }
}

// Column pass: FFT_y of the zero-padded column, times the Green half spectrum, IFFT_y, keep y < ny.
__global__ __launch_bounds__(kFFTThreads) void k_cols_conv(const Plane *planes, double2 *qh, int64_t qh_stride,
                                                           const double2 *gspec, int64_t slot_stride, int nx, int ny,
                                                           FFTPlanDev py, const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    if (failed(err)) return;
    const int plane = blockIdx.y;
    const int P2 = 2 * ny;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = lds + (size_t)grp * 2 * P2;
    double2 *b = a + P2;
    const int kx = blockIdx.x * kFFTPerBlock + grp;
    const bool active = kx <= nx;
    double2 *col = qh + (size_t)plane * qh_stride + (size_t)(active ? kx : 0) * ny;
    for (int y = lane; y < P2; y += kFFTLanes) a[y] = (active && y < ny) ? col[y] : make_double2(0.0, 0.0);
    __syncthreads();
    double2 *Z = fft_lds<-1>(a, b, py, lane, kFFTLanes);
    double2 *W = (Z == a) ? b : a;
    const int slot = planes[plane].slot;
    const double2 *g = gspec + (size_t)(slot < 0 ? 0 : slot) * slot_stride + (size_t)(active ? kx : 0) * P2;
    for (int y = lane; y < P2; y += kFFTLanes) Z[y] = cmul(Z[y], g[y]);
    __syncthreads();
    double2 *R = fft_lds<+1>(Z, W, py, lane, kFFTLanes);
    if (!active) return;
    for (int y = lane; y < ny; y += kFFTLanes) col[y] = R[y];
}

// Row inverse: rows (ya, ya+1) from the half spectrum, IFFT_x, keep x < nx, scale 1/(P1 P2).
__global__ __launch_bounds__(kFFTThreads) void k_rows_inv(const Plane *planes, const double2 *qh, int64_t qh_stride, double *phi,
                                                          int nx, int ny, FFTPlanDev px, const ErrorInfo *err) {
    extern __shared__ double2 lds[];
    if (failed(err)) return;
    const int plane = blockIdx.y;
    const int P1 = 2 * nx, P2 = 2 * ny;
    const int grp = threadIdx.x / kFFTLanes, lane = threadIdx.x % kFFTLanes;
    double2 *a = lds + (size_t)grp * 2 * P1;
    double2 *b = a + P1;
    const int ya = 2 * (blockIdx.x * kFFTPerBlock + grp);
    const bool active = ya < ny;
    const bool hasb = ya + 1 < ny;
    const double2 *in = qh + (size_t)plane * qh_stride;
    for (int k = lane; k < P1; k += kFFTLanes) {
        double2 z = make_double2(0.0, 0.0);
        if (active) {
            int kk = k <= nx ? k : P1 - k;
            double2 A = in[(size_t)kk * ny + ya];
            double2 B = hasb ? in[(size_t)kk * ny + ya + 1] : make_double2(0.0, 0.0);
            if (k > nx) { A = cconj(A); B = cconj(B); }
            z = make_double2(A.x - B.y, A.y + B.x);   // A + i B
        }
        a[k] = z;
    }
    __syncthreads();
    double2 *Z = fft_lds<+1>(a, b, px, lane, kFFTLanes);
    if (!active) return;
    const double norm = 1.0 / ((double)P1 * (double)P2);
    double *out = phi + (size_t)plane * nx * ny;
    for (int x = lane; x < nx; x += kFFTLanes) {
        out[(size_t)ya * nx + x] = Z[x].x * norm;
        if (hasb) out[(size_t)(ya + 1) * nx + x] = Z[x].y * norm;
    }
}

// E = -grad(phi) of every plane of an interaction into the interleaved kick table
// table[inter][x][y][3*q + {phi, Ex, Ey}] (pic_cpu.jl:2531-2591; h of the plane's deposit mesh).
__global__ __launch_bounds__(256) void k_field(const Inter *inters, const Plane *planes, const double *phi, double *table,
                                               int nx, int ny, int fourth, int npl, const ErrorInfo *err) {
    if (failed(err)) return;
    const int d = blockIdx.y;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nx * ny) return;
    const int x = cell / ny, y = cell % ny;     // table order: x outer, y inner
    const Inter it = inters[d];
    if (!it.valid) return;
    double *out = table + ((size_t)d * nx * ny + cell) * (3 * npl);
    const double c4 = 1.0 / 12.0;
    for (int q = 0; q < it.nplanes; ++q) {
        const Plane &pl = planes[it.plane0 + q];
        const double *P = phi + (size_t)(it.plane0 + q) * nx * ny;   // [y][x]
        auto F = [&](int i, int j) { return P[(size_t)j * nx + i]; };
        const double hxi = 1.0 / pl.hx, hyi = 1.0 / pl.hy;
        double ex, ey;
        if (y == 0) ey = hyi * (1.5 * F(x, 0) - 2.0 * F(x, 1) + 0.5 * F(x, 2));
        else if (y == ny - 1) ey = hyi * (-1.5 * F(x, ny - 1) + 2.0 * F(x, ny - 2) - 0.5 * F(x, ny - 3));
        else if (fourth && (y == 1 || y == ny - 2)) ey = 0.5 * hyi * (F(x, y - 1) - F(x, y + 1));
        else if (fourth) ey = c4 * hyi * ((F(x, y + 2) - F(x, y - 2)) + 8.0 * (F(x, y - 1) - F(x, y + 1)));
        else ey = 0.5 * hyi * (F(x, y - 1) - F(x, y + 1));
        if (x == 0) ex = hxi * (1.5 * F(0, y) - 2.0 * F(1, y) + 0.5 * F(2, y));
        else if (x == nx - 1) ex = hxi * (-1.5 * F(nx - 1, y) + 2.0 * F(nx - 2, y) - 0.5 * F(nx - 3, y));
        else if (fourth && (x == 1 || x == nx - 2)) ex = 0.5 * hxi * (F(x - 1, y) - F(x + 1, y));
        else if (fourth) ex = c4 * hxi * ((F(x + 2, y) - F(x - 2, y)) + 8.0 * (F(x - 1, y) - F(x + 1, y)));
        else ex = 0.5 * hxi * (F(x - 1, y) - F(x + 1, y));
        out[3 * q + 0] = F(x, y);
        out[3 * q + 1] = ex;
        out[3 * q + 2] = ey;
    }
}

// ---------------------------------------------------------------------- launchers
static size_t fft_lds_bytes(int P) { return (size_t)kFFTPerBlock * 2 * P * sizeof(double2); }

static hipError_t ensure_lds(const void *fn, size_t bytes) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_green_build(const GreenJob *jobs, int njobs, double2 *spec, int64_t slot_stride, int nslots, int nx, int ny,
                              int integrated, const FFTPlanDev &px, const FFTPlanDev &py, const ErrorInfo *err,
                              hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    const int P1 = 2 * nx, P2 = 2 * ny;
    ensure_lds((const void *)k_green_rows, fft_lds_bytes(P1));
    ensure_lds((const void *)k_green_cols, fft_lds_bytes(P2));
    dim3 g1((P2 / 2 + kFFTPerBlock - 1) / kFFTPerBlock, njobs);
    hipLaunchKernelGGL(k_green_rows, g1, dim3(kFFTThreads), fft_lds_bytes(P1), s, jobs, njobs, spec, slot_stride, nslots, nx,
                       ny, integrated, px, err);
    dim3 g2((nx + 1 + kFFTPerBlock - 1) / kFFTPerBlock, njobs);
    hipLaunchKernelGGL(k_green_cols, g2, dim3(kFFTThreads), fft_lds_bytes(P2), s, jobs, njobs, spec, slot_stride, nslots, nx,
                       ny, py, err);
    return hipGetLastError();
}

hipError_t launch_rows_fwd(const Plane *planes, int nplanes, unsigned long long *charge, int64_t charge_stride, double2 *qh,
                           int64_t qh_stride, int nx, int ny, double inv_scale, const FFTPlanDev &px, const ErrorInfo *err,
                           hipStream_t s) {
    if (nplanes <= 0) return hipSuccess;
    ensure_lds((const void *)k_rows_fwd, fft_lds_bytes(2 * nx));
    dim3 g(((ny + 1) / 2 + kFFTPerBlock - 1) / kFFTPerBlock, nplanes);
    hipLaunchKernelGGL(k_rows_fwd, g, dim3(kFFTThreads), fft_lds_bytes(2 * nx), s, planes, charge, charge_stride, qh,
                       qh_stride, nx, ny, inv_scale, px, err);
    return hipGetLastError();
}

hipError_t launch_cols_conv(const Plane *planes, int nplanes, double2 *qh, int64_t qh_stride, const double2 *gspec,
                            int64_t slot_stride, int nx, int ny, const FFTPlanDev &py, const ErrorInfo *err, hipStream_t s) {
    if (nplanes <= 0) return hipSuccess;
    ensure_lds((const void *)k_cols_conv, fft_lds_bytes(2 * ny));
    dim3 g((nx + 1 + kFFTPerBlock - 1) / kFFTPerBlock, nplanes);
    hipLaunchKernelGGL(k_cols_conv, g, dim3(kFFTThreads), fft_lds_bytes(2 * ny), s, planes, qh, qh_stride, gspec, slot_stride,
                       nx, ny, py, err);
    return hipGetLastError();
}

hipError_t launch_rows_inv(const Plane *planes, int nplanes, const double2 *qh, int64_t qh_stride, double *phi, int nx,
                           int ny, const FFTPlanDev &px, const ErrorInfo *err, hipStream_t s) {
    if (nplanes <= 0) return hipSuccess;
    ensure_lds((const void *)k_rows_inv, fft_lds_bytes(2 * nx));
    dim3 g(((ny + 1) / 2 + kFFTPerBlock - 1) / kFFTPerBlock, nplanes);
    hipLaunchKernelGGL(k_rows_inv, g, dim3(kFFTThreads), fft_lds_bytes(2 * nx), s, planes, qh, qh_stride, phi, nx, ny, px, err);
    return hipGetLastError();
}

hipError_t launch_field(const Inter *inters, int ninters, const Plane *planes, const double *phi, double *table, int nx,
                        int ny, int fourth, int npl, const ErrorInfo *err, hipStream_t s) {
    if (ninters <= 0) return hipSuccess;
    dim3 g((nx * ny + 255) / 256, ninters);
    hipLaunchKernelGGL(k_field, g, dim3(256), 0, s, inters, planes, phi, table, nx, ny, fourth, npl, err);
    return hipGetLastError();
}

}  // namespace octopic
