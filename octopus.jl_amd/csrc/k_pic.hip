// k_pic.hip -- particle-streaming stages of one wavefront batch of slice pairs.
//
//   bounds    : one pass over each slice of a pair (pic_cpu.jl:961-1035, 2700-2706): the slice's
//               source box at the field-slice boundary drifts sL/sR, its field box at the per-particle
//               drift 1/2(z - c_other), and its luminosity box at sM.  Exact min/max, any order.
//   geometry  : one lane per pair: the meshes (_pic_interaction_grids, pic_cpu.jl:1674-1718), the
//               slice-pair Green-cache decision (_pic_slice_pair_green!, 1729-1808) against the
//               cache geometry resident in HBM, the luminosity mesh (2700-2717), and the plane,
//               interaction and Green-build descriptors for the later stages.  No host round trip.
//   deposit   : CIC/TSC charge (pic_cpu.jl:2015-2163) into an LDS-privatised fixed-point (2^-k)
//               uint64 copy of the mesh, flushed by 64-bit integer atomics: integer sums are
//               associative, so the charge is bit-identical for every launch geometry.
//   kick      : forward virtual drift, field gather (x outer, y inner), px/py/pz kick, back drift
//               and the pz slingshot terms, fused (pic_cpu.jl:1218-1273, 2596-2680, 1448-1538).
//   lum       : per-pair overlap sum q1.q2 over the full (lnx+1)(lny+1) node extent (2731-2735).
#include "common.h"
#include "launch.h"

namespace octopic {

__device__ __forceinline__ bool failed(const ErrorInfo *err) { return *((volatile const int32_t *)&err->code) != 0; }

__device__ __forceinline__ void set_error(ErrorInfo *err, int code, int p, int dir, int stage, double a, double b) {
    if (atomicCAS(&err->code, 0, code) == 0) {
        err->p = p; err->dir = dir; err->stage = stage; err->a = a; err->b = b;
    }
}

// ---------------------------------------------------------------------------------- bounds
// Each segment (pair, side) owns 12 uint64 slots: q even = min, q odd = max of
//   0-3  source box at the field-slice boundary drifts sL, sR      (x min/max, y min/max)
//   4-7  field box at the per-particle drift 1/2(z - c_other)
//   8-11 luminosity box at sM
// A max slot holds order_key(v); a min slot holds ~order_key(v); NaN is all-ones in both and so wins
// every max, as NaN wins Julia's min/max.  Zero is the empty value.  Blocks fold their values in with
// 64-bit atomicMax (exact and order-free), and the geometry stage reads each slot and zeroes it.
constexpr int kBoundsThreads = 256;

__device__ __forceinline__ double take_bound(unsigned long long *slot, bool is_min) {
    const double v = bound_value(*slot, is_min);
    *slot = 0ull;
    return v;
}

template <bool SRCFLD, bool LUM>
__global__ __launch_bounds__(kBoundsThreads) void k_bounds(BoundsArgs a) {
    if (failed(a.err)) return;
    const int seg = blockIdx.y, k = seg >> 1, side = seg & 1;
    const BatchPair pr = a.pairs[k];
    const SliceTable *own = side == 0 ? a.tab1 : a.tab2;
    const SliceTable *oth = side == 0 ? a.tab2 : a.tab1;
    const int so = side == 0 ? pr.i : pr.j, sx = side == 0 ? pr.j : pr.i;
    const double c_own = own->center[so], c_oth = oth->center[sx];
    const double sL = 0.5 * (c_own - oth->boundary[sx]);
    const double sR = 0.5 * (c_own - oth->boundary[sx + 1]);
    const double sM = 0.5 * (c_own - c_oth);
    const int64_t off = own->offset[so], n = own->count[so];
    const int64_t lo = n * blockIdx.x / a.nchunks, hi = n * (blockIdx.x + 1) / a.nchunks;
    const double *X = a.st.c[side][0] + off, *PX = a.st.c[side][1] + off, *Y = a.st.c[side][2] + off,
                 *PY = a.st.c[side][3] + off, *Z = a.st.c[side][4] + off;
    double v[12];
    for (int q = 0; q < 12; q += 2) { v[q] = INFINITY; v[q + 1] = -INFINITY; }
    auto visit = [&](double x, double px, double y, double py, double z) {
        if (SRCFLD) {
            const double xl = x + px * sL, yl = y + py * sL, xr = x + px * sR, yr = y + py * sR;
            v[0] = jmin(jmin(v[0], xl), xr); v[1] = jmax(jmax(v[1], xl), xr);
            v[2] = jmin(jmin(v[2], yl), yr); v[3] = jmax(jmax(v[3], yl), yr);
            const double s = 0.5 * (z - c_oth);
            const double xf = x + s * px, yf = y + s * py;
            v[4] = jmin(v[4], xf); v[5] = jmax(v[5], xf); v[6] = jmin(v[6], yf); v[7] = jmax(v[7], yf);
        }
        if (LUM) {
            const double xv = x + px * sM, yv = y + py * sM;
            v[8] = jmin(v[8], xv); v[9] = jmax(v[9], xv); v[10] = jmin(v[10], yv); v[11] = jmax(v[11], yv);
        }
    };
    // two particles per step, their loads issued together (exact min/max: any order)
    const int64_t step = blockDim.x;
    int64_t t = lo + threadIdx.x;
    for (; t + step < hi; t += 2 * step) {
        const double x0 = X[t], px0 = PX[t], y0 = Y[t], py0 = PY[t], z0 = SRCFLD ? Z[t] : 0.0;
        const double x1 = X[t + step], px1 = PX[t + step], y1 = Y[t + step], py1 = PY[t + step],
                     z1 = SRCFLD ? Z[t + step] : 0.0;
        visit(x0, px0, y0, py0, z0);
        visit(x1, px1, y1, py1, z1);
    }
    if (t < hi) visit(X[t], PX[t], Y[t], PY[t], SRCFLD ? Z[t] : 0.0);
    constexpr int q0 = SRCFLD ? 0 : 8, q1 = LUM ? 12 : 8;
    constexpr int kWaves = kBoundsThreads / 64;
    __shared__ unsigned long long red[kWaves][12];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = q0; q < q1; ++q) {
        unsigned long long u = bound_slot(v[q], (q & 1) == 0);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long w = __shfl_xor(u, o);
            u = w > u ? w : u;
        }
        if (lane == 0) red[wave][q] = u;
    }
    __syncthreads();
    const int q = q0 + (int)threadIdx.x;
    if (q < q1) {
        unsigned long long u = red[0][q];
        for (int w = 1; w < kWaves; ++w) u = red[w][q] > u ? red[w][q] : u;
        if (u) atomicMax(&a.keys[(size_t)seg * kBoundSlots + q], u);
        // this chunk's own source box (sL and sR drifts) and field box, for the LDS windows of the
        // deposit and kick blocks that take the same particle range
        if (SRCFLD && q < 8) a.chunkbox[((size_t)seg * a.nchunks + blockIdx.x) * 8 + q] = bound_value(u, (q & 1) == 0);
    }
}

hipError_t launch_bounds(const BoundsArgs &a, hipStream_t s) {
    if (a.npairs <= 0 || !(a.srcfld || a.lum)) return hipSuccess;
    const dim3 g(a.nchunks, 2 * a.npairs), b(kBoundsThreads);
    if (a.srcfld && a.lum) hipLaunchKernelGGL((k_bounds<true, true>), g, b, 0, s, a);
    else if (a.srcfld) hipLaunchKernelGGL((k_bounds<true, false>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_bounds<false, true>), g, b, 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------- geometry
// Lane (k, dir) takes the slots it consumes: the source box (0-3) of the source side and the field box
// (4-7) of the field side; lane dir 1 also takes both luminosity boxes (8-11).  Every slot of the batch
// is thereby zeroed exactly once, before any early return.
// All slots are read before any is zeroed, so the loads are in flight together.
__device__ __forceinline__ void take_boxes(unsigned long long *keys, int k, int dir, double vs[4], double vf[4],
                                           double l1[4], double l2[4]) {
    unsigned long long *s0 = keys + (size_t)(2 * k) * kBoundSlots, *s1 = s0 + kBoundSlots;
    unsigned long long *src = dir == 1 ? s0 : s1, *fld = dir == 1 ? s1 : s0;
    unsigned long long u[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        u[q] = src[q];
        u[4 + q] = fld[4 + q];
        u[8 + q] = dir == 1 ? s0[8 + q] : 0ull;
        u[12 + q] = dir == 1 ? s1[8 + q] : 0ull;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool mn = (q & 1) == 0;
        vs[q] = bound_value(u[q], mn);
        vf[q] = bound_value(u[4 + q], mn);
        l1[q] = bound_value(u[8 + q], mn);
        l2[q] = bound_value(u[12 + q], mn);
        src[q] = 0ull;
        fld[4 + q] = 0ull;
        if (dir == 1) { s0[8 + q] = 0ull; s1[8 + q] = 0ull; }
    }
}

__device__ __forceinline__ bool resolve_extent(double xspan, double yspan, double fx, double fy, double &w, double &h) {
    w = jmax(xspan, fx);
    h = jmax(yspan, fy);
    return __builtin_isfinite(w) && __builtin_isfinite(h) && w > 0.0 && h > 0.0;
}

__device__ __forceinline__ void align_origins(int gtype, double s0, double f0, double h, double &so, double &fo) {
    const double f1 = s0 / h - floor(s0 / h);
    const double f2 = f0 / h - floor(f0 / h);
    double t = (f2 - f1) / 2.0;
    if (gtype == OCTOPIC_GREEN_STANDARD) {
        t += t > 0.0 ? -0.25 : 0.25;
    } else {
        if (t > 0.5) t -= 0.5;
        else if (t < -0.5) t += 0.5;
    }
    const double shift = t * h;
    so = s0 + shift;
    fo = f0 - shift;
}

__device__ __forceinline__ double quantize_extent(double w, double q) {
    if (!(q > 0.0 && __builtin_isfinite(w) && w > 0.0)) return w;
    return exp2(ceil(log2(w) / q) * q);
}

// _pic_interaction_grids (pic_cpu.jl:1674-1718).  Returns false for a degenerate extent.
__device__ __forceinline__ bool interaction_grids(const GeomOpts &o, double sxmin, double sxmax, double symin, double symax, double fxmin,
                                  double fxmax, double fymin, double fymax, Grid &sg, Grid &fg) {
    double width, height;
    if (!resolve_extent(jmax(sxmax - sxmin, fxmax - fxmin), jmax(symax - symin, fymax - fymin), o.min_x, o.min_y, width,
                        height))
        return false;
    const double tx = width / (double)(o.nx - 4);
    const double ty = height / (double)(o.ny - 4);
    width += 3.0 * tx;
    height += 3.0 * ty;
    if (o.quantize > 0.0) {
        width = quantize_extent(width, o.quantize);
        height = quantize_extent(height, o.quantize);
        const double hx = width / (double)(o.nx - 1);
        const double hy = height / (double)(o.ny - 1);
        double sx0 = rint(((sxmin + sxmax) / 2.0 - width / 2.0) / hx) * hx;
        double sy0 = rint(((symin + symax) / 2.0 - height / 2.0) / hy) * hy;
        double fx0 = rint(((fxmin + fxmax) / 2.0 - width / 2.0) / hx) * hx;
        double fy0 = rint(((fymin + fymax) / 2.0 - height / 2.0) / hy) * hy;
        align_origins(o.green_type, sx0, fx0, hx, sx0, fx0);
        align_origins(o.green_type, sy0, fy0, hy, sy0, fy0);
        sg = Grid{sx0, sy0, width, height};
        fg = Grid{fx0, fy0, width, height};
        return true;
    }
    double sx0 = sxmin - 1.5 * tx, sy0 = symin - 1.5 * ty;
    double fx0 = fxmin - 1.5 * tx, fy0 = fymin - 1.5 * ty;
    align_origins(o.green_type, sx0, fx0, tx, sx0, fx0);
    align_origins(o.green_type, sy0, fy0, ty, sy0, fy0);
    sg = Grid{sx0, sy0, width, height};
    fg = Grid{fx0, fy0, width, height};
    return true;
}

__device__ __forceinline__ bool size_usable(const Grid &c, const Grid &r, double min_ratio) {
    if (!(c.width >= r.width)) return false;
    if (!(c.height >= r.height)) return false;
    if (!(r.width >= min_ratio * c.width)) return false;
    if (!(r.height >= min_ratio * c.height)) return false;
    return true;
}
__device__ __forceinline__ bool covers(const GeomOpts &o, const Grid &g, double xmin, double xmax, double ymin, double ymax) {
    const double hx = g.width / (double)(o.nx - 1);
    const double hy = g.height / (double)(o.ny - 1);
    if (!(g.x0 + kTemplateMargin * hx <= xmin)) return false;
    if (!(g.x0 + g.width - kTemplateMargin * hx >= xmax)) return false;
    if (!(g.y0 + kTemplateMargin * hy <= ymin)) return false;
    if (!(g.y0 + g.height - kTemplateMargin * hy >= ymax)) return false;
    return true;
}
__device__ __forceinline__ Grid expand_grid(const Grid &g, double factor) {
    if (factor <= 1.0) return g;
    const double nw = g.width * factor, nh = g.height * factor;
    return Grid{g.x0 - (nw - g.width) / 2.0, g.y0 - (nh - g.height) / 2.0, nw, nh};
}

__device__ __forceinline__ void fill_record(octopic_interaction_record &r, int p, int i, int j, int dir, int event,
                                            const Grid &sg, const Grid &fg) {
    r.p = p; r.i = i; r.j = j; r.dir = dir; r.cache_event = event; r.reserved = 0;
    r.source_grid[0] = sg.x0; r.source_grid[1] = sg.y0; r.source_grid[2] = sg.width; r.source_grid[3] = sg.height;
    r.field_grid[0] = fg.x0; r.field_grid[1] = fg.y0; r.field_grid[2] = fg.width; r.field_grid[3] = fg.height;
}

__device__ __forceinline__ void set_plane(Plane &pl, int inter, int kind, int beam, int n1, int n2, int slot, int64_t off, int64_t n,
                          double s, double x0, double y0, double hx, double hy) {
    pl.inter = inter; pl.kind = kind; pl.beam = beam; pl.n1 = n1; pl.n2 = n2; pl.slot = slot; pl.count_box = 0;
    pl.reserved = 0; pl.off = off; pl.n = n; pl.s = s; pl.x0 = x0; pl.y0 = y0;
    pl.hx = hx; pl.hy = hy; pl.hxi = 1.0 / hx; pl.hyi = 1.0 / hy;
    pl.bx1 = 0.0; pl.by1 = 0.0;
}

// Luminosity mesh of one pair (pic_cpu.jl:2700-2717) and its two deposit planes.  l1, l2 are the
// luminosity boxes (x min, x max, y min, y max) of slice i of beam 1 and slice j of beam 2 at sM.
__device__ __forceinline__ bool lum_geometry(const GeomOpts &o, const double *l1, const double *l2, int k, int p,
                                          const SliceTable *t1, const SliceTable *t2, int i, int j, Plane *planes,
                                          int lum_plane0, double *lumgeo, ErrorInfo *err) {
    double xmin = jmin(l1[0], l2[0]), xmax = jmax(l1[1], l2[1]);
    double ymin = jmin(l1[2], l2[2]), ymax = jmax(l1[3], l2[3]);
    double width, height;
    if (!resolve_extent(xmax - xmin, ymax - ymin, o.min_x, o.min_y, width, height)) {
        set_error(err, OCTOPIC_ERR_DEGENERATE_EXTENT, p, 0, 2, xmax - xmin, ymax - ymin);
        return false;
    }
    const double tx = width / ((double)o.lum_nx - 1.1);
    const double ty = height / ((double)o.lum_ny - 1.1);
    width += 0.1 * tx;
    height += 0.1 * ty;
    xmin -= 0.05 * tx;
    ymin -= 0.05 * ty;
    const double hx = width / (double)(o.lum_nx - 1);
    const double hy = height / (double)(o.lum_ny - 1);
    const double c1 = t1->center[i], c2 = t2->center[j];
    set_plane(planes[lum_plane0 + 2 * k], k, 1, 0, o.lum_nx + 1, o.lum_ny + 1, -1, t1->offset[i], t1->count[i],
              0.5 * (c1 - c2), xmin, ymin, hx, hy);
    set_plane(planes[lum_plane0 + 2 * k + 1], k, 1, 1, o.lum_nx + 1, o.lum_ny + 1, -1, t2->offset[j], t2->count[j],
              0.5 * (c2 - c1), xmin, ymin, hx, hy);
    lumgeo[3 * k + 0] = 1.0 / hx;
    lumgeo[3 * k + 1] = 1.0 / hy;
    lumgeo[3 * k + 2] = o.klum;
    return true;
}

// Every descriptor an interaction owns is made inert before anything can return early, so that no
// later stage of the batch can read a stale or uninitialised descriptor.  Lane dir 1 of a pair also
// owns the pair's two luminosity planes.
__device__ __forceinline__ void neutralise_inter(int k, int dir, int npairs, const GeomOpts &o, Inter *inters, Plane *planes,
                                 GreenJob *jobs) {
    const int d = 2 * k + dir - 1;
    inters[d].valid = 0;
    inters[d].fld_n = 0;
    inters[d].nplanes = 0;
    inters[d].gauss = 0;
    if (jobs) { jobs[d].valid = 0; jobs[d].slot = -1; }
    for (int q = 0; q < o.npl; ++q) {
        Plane &pl = planes[d * o.npl + q];
        pl.n = 0; pl.off = 0; pl.beam = 0; pl.slot = -1; pl.count_box = 0;
    }
    if (dir == 1) {
        const int lum0 = 2 * npairs * o.npl;
        for (int side = 0; side < 2; ++side) {
            Plane &pl = planes[lum0 + 2 * k + side];
            pl.n = 0; pl.off = 0; pl.beam = 0; pl.slot = -1; pl.count_box = 0;
        }
    }
}

__device__ __forceinline__ void fill_inter(Inter &it, const GeomOpts &o, int p, int i, int j, int dir, const SliceTable *ts,
                           const SliceTable *tf, int si, int fi) {
    const double c_src = ts->center[si];
    it.p = p; it.i = i; it.j = j; it.dir = dir;
    it.src_beam = dir == 1 ? 0 : 1;
    it.fld_beam = dir == 1 ? 1 : 0;
    it.src_off = ts->offset[si]; it.src_n = ts->count[si];
    it.fld_off = tf->offset[fi]; it.fld_n = tf->count[fi];
    it.c_src = c_src; it.c_fld = tf->center[fi];
    it.lb = tf->boundary[fi]; it.rb = tf->boundary[fi + 1];
    it.sL = 0.5 * (c_src - it.lb);
    it.sR = 0.5 * (c_src - it.rb);
    it.sM = 0.5 * (c_src - it.c_fld);
    it.kick_scale = 2.0 * (dir == 1 ? o.kbb2 : o.kbb1);
}

// The source slice's Gaussian at the two field-slice boundary drifts (_gpic_source_moments,
// _gpic_drifted_gaussian, _gpic_boundary, gaussian_pic.jl:374-523, uncoupled) from the moment partials of
// interaction d, folded in chunk order.  Returns whether the control variate is defined
// (_gpic_control_variate_mode, :525-541): at least two members and positive finite widths at both.
__device__ bool gpic_gaussians(const GeometryArgs &a, int d, int beam, int64_t off, int64_t n, double sL, double sR,
                               GaussBnd g[2], double &mpx, double &mpy) {
    if (n < 1) return false;
    double sm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < a.o.mom_chunks; ++c)
        for (int j = 0; j < 10; ++j) sm[j] += a.mom_part[((size_t)d * a.o.mom_chunks + c) * 10 + j];
    const double x0 = a.st.c[beam][0][off], px0 = a.st.c[beam][1][off], y0 = a.st.c[beam][2][off], py0 = a.st.c[beam][3][off];
    const double invn = 1.0 / (double)n;
    const double dmx = sm[0] * invn, dmpx = sm[1] * invn, dmy = sm[2] * invn, dmpy = sm[3] * invn;
    const double mx = x0 + dmx, my = y0 + dmy;
    mpx = px0 + dmpx;
    mpy = py0 + dmpy;
    // _shifted_second_moment / _shifted_cross_moment (slicing.jl:664-668): muladd(-a, b, sum * invn)
    const double varx = jmax(fma(-dmx, dmx, sm[4] * invn), 0.0), varpx = jmax(fma(-dmpx, dmpx, sm[5] * invn), 0.0);
    const double vary = jmax(fma(-dmy, dmy, sm[6] * invn), 0.0), varpy = jmax(fma(-dmpy, dmpy, sm[7] * invn), 0.0);
    const double cxpx = fma(-dmx, dmpx, sm[8] * invn), cypy = fma(-dmy, dmpy, sm[9] * invn);
    bool ok = n >= 2;
    for (int b = 0; b < 2; ++b) {
        const double s = b == 0 ? sL : sR;
        g[b].mux = mx + mpx * s;
        g[b].muy = my + mpy * s;
        g[b].sigx = sqrt(jmax(varx + 2.0 * s * cxpx + s * s * varpx, 0.0));
        g[b].sigy = sqrt(jmax(vary + 2.0 * s * cypy + s * s * varpy, 0.0));
        g[b].rx = 2.0 * (cxpx + s * varpx);
        g[b].ry = 2.0 * (cypy + s * varpy);
        ok = ok && __builtin_isfinite(g[b].sigx) && __builtin_isfinite(g[b].sigy) && g[b].sigx > 0.0 && g[b].sigy > 0.0;
    }
    return ok;
}

// one lane per (pair k, direction): 2*npairs lanes
__global__ void k_geometry(GeometryArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= 2 * a.npairs) return;
    const int k = lane >> 1, dir = (lane & 1) + 1, d = lane;
    const BatchPair pr = a.pairs[k];     // loads first: the stores below may alias them for the compiler
    double vs[4], vf[4], l1[4], l2[4];   // source box, field box, luminosity boxes
    take_boxes(a.keys, k, dir, vs, vf, l1, l2);
    const bool fail = failed(a.err);
    neutralise_inter(k, dir, a.npairs, a.o, a.inters, a.planes, a.jobs);
    if (fail) return;
    const GeomOpts &o = a.o;
    const int i = pr.i, j = pr.j, p = pr.p;
    const SliceTable *ts = dir == 1 ? a.tab1 : a.tab2, *tf = dir == 1 ? a.tab2 : a.tab1;
    const int si = dir == 1 ? i : j, fi = dir == 1 ? j : i;
    Inter &it = a.inters[d];
    if (!finite4(vs[0], vs[1], vs[2], vs[3])) { set_error(a.err, OCTOPIC_ERR_NONFINITE, p, dir, 0, 0, 0); return; }
    if (!finite4(vf[0], vf[1], vf[2], vf[3])) { set_error(a.err, OCTOPIC_ERR_NONFINITE, p, dir, 1, 0, 0); return; }
    // Gaussian-subtracted solver: the source mesh also holds +-margin_sigma of the subtracted Gaussian
    // (gaussian_pic.jl:687-699); without a defined Gaussian the interaction is plain PIC
    bool gauss = false;
    GaussBnd gb[2];
    double gmpx = 0.0, gmpy = 0.0;
    if (o.gauss) {
        const double c_src = ts->center[si];
        const double sL = 0.5 * (c_src - tf->boundary[fi]), sR = 0.5 * (c_src - tf->boundary[fi + 1]);
        gauss = gpic_gaussians(a, d, dir == 1 ? 0 : 1, ts->offset[si], ts->count[si], sL, sR, gb, gmpx, gmpy);
        if (gauss && o.margin > 0.0) {
            const double m = o.margin;
            vs[0] = jmin(vs[0], jmin(gb[0].mux - m * gb[0].sigx, gb[1].mux - m * gb[1].sigx));
            vs[1] = jmax(vs[1], jmax(gb[0].mux + m * gb[0].sigx, gb[1].mux + m * gb[1].sigx));
            vs[2] = jmin(vs[2], jmin(gb[0].muy - m * gb[0].sigy, gb[1].muy - m * gb[1].sigy));
            vs[3] = jmax(vs[3], jmax(gb[0].muy + m * gb[0].sigy, gb[1].muy + m * gb[1].sigy));
            if (!finite4(vs[0], vs[1], vs[2], vs[3])) { set_error(a.err, OCTOPIC_ERR_NONFINITE, p, dir, 0, 0, 0); return; }
        }
    }
    Grid sg, fg;
    if (!interaction_grids(o, vs[0], vs[1], vs[2], vs[3], vf[0], vf[1], vf[2], vf[3], sg, fg)) {
        set_error(a.err, OCTOPIC_ERR_DEGENERATE_EXTENT, p, dir, 3, jmax(vs[1] - vs[0], vf[1] - vf[0]),
                  jmax(vs[3] - vs[2], vf[3] - vf[2]));
        return;
    }
    int event = 0, slot, build;
    if (o.green_cache == OCTOPIC_CACHE_SLICE_PAIR) {
        const int key = (i * o.ns2 + j) * 2 + (dir - 1);
        CacheGeo &e = a.cache[key];
        bool usable = false;
        if (e.valid)
            usable = size_usable(e.sg, sg, o.min_ratio) && size_usable(e.fg, fg, o.min_ratio) &&
                     covers(o, e.sg, vs[0], vs[1], vs[2], vs[3]) && covers(o, e.fg, vf[0], vf[1], vf[2], vf[3]);
        slot = key;
        if (usable) {
            event = 1;
            build = 0;
            sg = e.sg;
            fg = e.fg;
            atomicAdd(&a.counters[0], 1ull);
        } else {
            Grid es = expand_grid(sg, 1.0 + o.growth);
            Grid ef = expand_grid(fg, 1.0 + o.growth);
            if (o.green_type != OCTOPIC_GREEN_INTEGRATED) {   // _pic_realign_expanded_grids (pic_cpu.jl:1849-1858)
                const double hx = es.width / (double)(o.nx - 1);
                const double hy = es.height / (double)(o.ny - 1);
                double sx0, fx0, sy0, fy0;
                align_origins(o.green_type, es.x0, ef.x0, hx, sx0, fx0);
                align_origins(o.green_type, es.y0, ef.y0, hy, sy0, fy0);
                es.x0 = sx0; es.y0 = sy0; ef.x0 = fx0; ef.y0 = fy0;
            }
            event = e.valid ? 3 : 2;
            atomicAdd(&a.counters[e.valid ? 2 : 1], 1ull);
            e.valid = 1;
            e.epoch = o.epoch;
            e.sg = es;
            e.fg = ef;
            sg = es;
            fg = ef;
            build = 1;
        }
    } else {
        slot = o.temp_slot0 + d;
        build = 1;
    }
    const double hx = sg.width / (double)(o.nx - 1);
    const double hy = sg.height / (double)(o.ny - 1);
    GreenJob &job = a.jobs[d];
    job.slot = slot;
    job.valid = build;
    job.fx0 = fg.x0; job.fy0 = fg.y0; job.sx0 = sg.x0; job.sy0 = sg.y0; job.hx = hx; job.hy = hy;
    if (build) atomicAdd(&a.counters[3], 1ull);
    fill_inter(it, o, p, i, j, dir, ts, tf, si, fi);
    it.sg = sg; it.fg = fg; it.fgR = fg;
    if (gauss) {
        it.g[0] = gb[0];
        it.g[1] = gb[1];
        it.gmpx = gmpx;
        it.gmpy = gmpy;
    }
    it.plane0 = d * o.npl;
    it.nplanes = o.npl;
    it.event = event;
    it.count_field = 0;
    const int src_beam = it.src_beam;
    set_plane(a.planes[d * o.npl + 0], d, 0, src_beam, o.nx, o.ny, slot, it.src_off, it.src_n, it.sL, sg.x0, sg.y0, hx, hy);
    set_plane(a.planes[d * o.npl + 1], d, 0, src_beam, o.nx, o.ny, slot, it.src_off, it.src_n, it.sR, sg.x0, sg.y0, hx, hy);
    if (o.npl == 3)
        set_plane(a.planes[d * o.npl + 2], d, 0, src_beam, o.nx, o.ny, slot, it.src_off, it.src_n, 0.5 * (it.sL + it.sR),
                  sg.x0, sg.y0, hx, hy);
    fill_record(a.records[2 * p + dir - 1], p, i, j, dir, event, sg, fg);
    it.gauss = gauss ? 1 : 0;
    it.valid = 1;
    // after this lane's own checks, which see every x, px, y, py of both slices
    if (dir == 1 && o.compute_lum)
        lum_geometry(o, l1, l2, k, p, a.tab1, a.tab2, i, j, a.planes, 2 * a.npairs * o.npl, a.lumgeo, a.err);
}

hipError_t launch_geometry(const GeometryArgs &a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_geometry, dim3((2 * a.npairs + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

// A collide that fails leaves no entry it built: the Green kernels stop once an error is set, so an
// entry committed by its geometry stage may have no spectrum (or an older mesh's).
__global__ void k_cache_invalidate(CacheGeo *cache, int nkeys, int epoch) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nkeys && cache[k].epoch == epoch) cache[k].valid = 0;
}

hipError_t launch_cache_invalidate(CacheGeo *cache, int nkeys, int epoch, hipStream_t s) {
    if (!cache || nkeys <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cache_invalidate, dim3((nkeys + 255) / 256), dim3(256), 0, s, cache, nkeys, epoch);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------- deposit
constexpr int kDepositThreads = 1024;

// The block keeps an LDS copy of the stencil window of its chunks' source box (window-local rows),
// zeroed before and flushed after the deposit: every position it deposits lies in that box (the sM and
// luminosity drifts lie between sL and sR, and positions are x + px*s, monotonic in s), the first
// stencil node is monotonic in the position, and positions off the mesh have zero weights.  The LDS
// holds up to lds_cells cells (the whole mesh up to 143 x 143 nodes); a block whose window is larger
// (larger meshes only) adds straight into the plane in HBM.  Integer sums are associative, so the
// charge is bit-identical either way.
// W = double, or float in the FP32 mode (the weight product in FP32, then the same fixed point).
// One work item: chunk group `grp` of plane `plane` (called by every thread of the block).
template <int METHOD, class W>
__device__ __forceinline__ void deposit_item(const DepositArgs &a, unsigned long long *tile, int plane, int grp) {
    const Plane pl = a.planes[plane];
    // the item takes the chunks [grp*group, (grp+1)*group) of the batch's chunking
    const int c0 = grp * a.group, c1 = min(c0 + a.group, a.chunks);
    const int64_t lo = pl.n * c0 / a.chunks, hi = pl.n * c1 / a.chunks;
    if (lo >= hi) return;
    const int n1 = pl.n1, n2 = pl.n2;
    int wx0 = 0, wx1 = n1 - 1, wy0 = 0, wy1 = n2 - 1;
    if (a.chunkbox) {
        const int seg = 2 * (pl.kind == 0 ? pl.inter / 2 : pl.inter) + pl.beam;
        double b[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
        for (int c = c0; c < c1; ++c) {
            const double *bc = a.chunkbox + ((size_t)seg * a.chunks + c) * 8;
            b[0] = jmin(b[0], bc[0]); b[1] = jmax(b[1], bc[1]); b[2] = jmin(b[2], bc[2]); b[3] = jmax(b[3], bc[3]);
        }
        stencil_window<METHOD>((b[0] - pl.x0) * pl.hxi, (b[1] - pl.x0) * pl.hxi, n1, wx0, wx1);
        stencil_window<METHOD>((b[2] - pl.y0) * pl.hyi, (b[3] - pl.y0) * pl.hyi, n2, wy0, wy1);
    }
    const int wnx = wx1 - wx0 + 1, wcells = wnx * (wy1 - wy0 + 1);
    const bool in_lds = wcells <= a.lds_cells;   // block-uniform
    if (in_lds) {
        for (int c = threadIdx.x; c < wcells; c += blockDim.x) tile[c] = 0ull;
        __syncthreads();
    }
    unsigned long long *out = a.charge + (size_t)plane * a.charge_stride;
    const double *X = a.st.c[pl.beam][0] + pl.off, *PX = a.st.c[pl.beam][1] + pl.off, *Y = a.st.c[pl.beam][2] + pl.off,
                 *PY = a.st.c[pl.beam][3] + pl.off;
    const double s = pl.s, x0 = pl.x0, y0 = pl.y0, hxi = pl.hxi, hyi = pl.hyi, scale = a.scale;
    constexpr int K = METHOD == 0 ? 2 : 3;
    int dropped = 0;
    for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        const double xd = X[t] + PX[t] * s;
        const double yd = Y[t] + PY[t] * s;
        if (pl.count_box && !(__builtin_isfinite(xd) && x0 <= xd && xd <= pl.bx1 && __builtin_isfinite(yd) && y0 <= yd &&
                              yd <= pl.by1))
            ++dropped;
        int bx, by;
        double wx[3], wy[3];
        stencil<METHOD>((xd - x0) * hxi, n1, bx, wx);
        stencil<METHOD>((yd - y0) * hyi, n2, by, wy);
        // nonzero weights only for in-range positions, whose stencils lie inside the window
#pragma unroll
        for (int m = 0; m < K; ++m)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                // weights lie in [0, 1] and scale <= 2^40, so the product is in fixed_point's range
                unsigned long long v;
                if constexpr (sizeof(W) == sizeof(double)) v = fixed_point_product(wx[m] * scale, wy[q]);
                else v = fixed_point((double)((float)wx[m] * (float)wy[q] * (float)scale));
                // a zero weight adds nothing (off-mesh positions sit at node 0, possibly outside the window)
                if (!v) continue;
                if (in_lds) atomicAdd(&tile[(by + q - wy0) * wnx + (bx + m - wx0)], v);
                else atomicAdd(&out[(by + q) * n1 + bx + m], v);
            }
    }
    if (pl.count_box) {
        for (int o = 32; o > 0; o >>= 1) dropped += __shfl_down(dropped, o);
        if ((threadIdx.x & 63) == 0 && dropped) atomicAdd(&a.dropped[0], (unsigned long long)dropped);
    }
    if (in_lds) {
        __syncthreads();
        for (int c = threadIdx.x; c < wcells; c += blockDim.x) {
            const int r = c / wnx;
            const unsigned long long v = tile[c];
            if (v) atomicAdd(&out[(wy0 + r) * n1 + wx0 + (c - r * wnx)], v);
        }
    }
}

// One work item for the `np` field planes of one interaction (planes plane0 .. plane0 + np - 1: the
// same source particles on the same mesh at different drifts): each particle's x, px, y, py are read
// once and deposited into np LDS windows of the same geometry (the box of the sL and sR positions holds
// every drift between).  When the np windows do not fit, the planes go one at a time through
// deposit_item.  The charge is the same integer sum either way.
template <int METHOD, class W>
__device__ __forceinline__ void deposit_item_multi(const DepositArgs &a, unsigned long long *tile, int plane0, int np, int grp) {
    const Plane pl = a.planes[plane0];
    // the interaction's luminosity plane (a.lum_plane0 + interaction), on its own mesh, when folded in
    const int lplane = a.lum_plane0 >= 0 ? a.lum_plane0 + plane0 / np : -1;
    const int c0 = grp * a.group, c1 = min(c0 + a.group, a.chunks);
    const int64_t lo = pl.n * c0 / a.chunks, hi = pl.n * c1 / a.chunks;
    if (lo >= hi) return;
    const int n1 = pl.n1, n2 = pl.n2;
    int wx0 = 0, wx1 = n1 - 1, wy0 = 0, wy1 = n2 - 1;
    if (a.chunkbox) {
        const int seg = 2 * (pl.inter / 2) + pl.beam;
        double b[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
        for (int c = c0; c < c1; ++c) {
            const double *bc = a.chunkbox + ((size_t)seg * a.chunks + c) * 8;
            b[0] = jmin(b[0], bc[0]); b[1] = jmax(b[1], bc[1]); b[2] = jmin(b[2], bc[2]); b[3] = jmax(b[3], bc[3]);
        }
        stencil_window<METHOD>((b[0] - pl.x0) * pl.hxi, (b[1] - pl.x0) * pl.hxi, n1, wx0, wx1);
        stencil_window<METHOD>((b[2] - pl.y0) * pl.hyi, (b[3] - pl.y0) * pl.hyi, n2, wy0, wy1);
    }
    const int wnx = wx1 - wx0 + 1, wcells = wnx * (wy1 - wy0 + 1);
    // luminosity window: the same chunk box (its sM drift lies between sL and sR) on the luminosity mesh
    Plane lp{};
    int lx0 = 0, ly0 = 0, lnx = 0, lny = 0, lcells = 0;
    if (lplane >= 0) {
        lp = a.planes[lplane];
        int lx1 = lp.n1 - 1, ly1 = lp.n2 - 1;
        lx0 = 0; ly0 = 0;
        if (a.chunkbox) {
            const int seg = 2 * (pl.inter / 2) + pl.beam;
            double b[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
            for (int c = c0; c < c1; ++c) {
                const double *bc = a.chunkbox + ((size_t)seg * a.chunks + c) * 8;
                b[0] = jmin(b[0], bc[0]); b[1] = jmax(b[1], bc[1]); b[2] = jmin(b[2], bc[2]); b[3] = jmax(b[3], bc[3]);
            }
            stencil_window<METHOD>((b[0] - lp.x0) * lp.hxi, (b[1] - lp.x0) * lp.hxi, lp.n1, lx0, lx1);
            stencil_window<METHOD>((b[2] - lp.y0) * lp.hyi, (b[3] - lp.y0) * lp.hyi, lp.n2, ly0, ly1);
        }
        lnx = lx1 - lx0 + 1;
        lny = ly1 - ly0 + 1;
        lcells = lnx * lny;
    }
    if (np * wcells + lcells > a.lds_cells) {   // block-uniform
        for (int p = 0; p < np; ++p) {
            deposit_item<METHOD, W>(a, tile, plane0 + p, grp);
            __syncthreads();   // the plane's window is flushed before the next zeroes it
        }
        if (lplane >= 0) {
            deposit_item<METHOD, double>(a, tile, lplane, grp);   // luminosity deposits stay FP64
            __syncthreads();
        }
        return;
    }
    for (int c = threadIdx.x; c < np * wcells + lcells; c += blockDim.x) tile[c] = 0ull;
    __syncthreads();
    const double *X = a.st.c[pl.beam][0] + pl.off, *PX = a.st.c[pl.beam][1] + pl.off, *Y = a.st.c[pl.beam][2] + pl.off,
                 *PY = a.st.c[pl.beam][3] + pl.off;
    const double x0 = pl.x0, y0 = pl.y0, hxi = pl.hxi, hyi = pl.hyi, scale = a.scale;
    double sp[3];
    int cb[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        sp[p] = p < np ? a.planes[plane0 + p].s : 0.0;
        cb[p] = p < np ? a.planes[plane0 + p].count_box : 0;
    }
    constexpr int K = METHOD == 0 ? 2 : 3;
    int dropped[3] = {0, 0, 0}, ldropped = 0;
    for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        const double xs = X[t], pxs = PX[t], ys = Y[t], pys = PY[t];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            if (p >= np) break;
            const double xd = xs + pxs * sp[p];
            const double yd = ys + pys * sp[p];
            if (cb[p] && !(__builtin_isfinite(xd) && x0 <= xd && xd <= pl.bx1 && __builtin_isfinite(yd) && y0 <= yd &&
                           yd <= pl.by1))
                ++dropped[p];
            int bx, by;
            double wx[3], wy[3];
            stencil<METHOD>((xd - x0) * hxi, n1, bx, wx);
            stencil<METHOD>((yd - y0) * hyi, n2, by, wy);
            unsigned long long *tp = tile + p * wcells;
#pragma unroll
            for (int m = 0; m < K; ++m)
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    unsigned long long v;
                    if constexpr (sizeof(W) == sizeof(double)) v = fixed_point_product(wx[m] * scale, wy[q]);
                    else v = fixed_point((double)((float)wx[m] * (float)wy[q] * (float)scale));
                    if (v) atomicAdd(&tp[(by + q - wy0) * wnx + (bx + m - wx0)], v);
                }
        }
        if (lplane >= 0) {   // the luminosity plane at sM, FP64 weights
            const double xd = xs + pxs * lp.s;
            const double yd = ys + pys * lp.s;
            if (lp.count_box && !(__builtin_isfinite(xd) && lp.x0 <= xd && xd <= lp.bx1 && __builtin_isfinite(yd) &&
                                  lp.y0 <= yd && yd <= lp.by1))
                ++ldropped;
            int bx, by;
            double wx[3], wy[3];
            stencil<METHOD>((xd - lp.x0) * lp.hxi, lp.n1, bx, wx);
            stencil<METHOD>((yd - lp.y0) * lp.hyi, lp.n2, by, wy);
            unsigned long long *tl = tile + np * wcells;
            unsigned long long *lout = a.charge + (size_t)lplane * a.charge_stride;
#pragma unroll
            for (int m = 0; m < K; ++m)
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const unsigned long long v = fixed_point_product(wx[m] * scale, wy[q]);
                    if (!v) continue;
                    // the window is the chunk box's, which holds the sM positions when the field slice's
                    // centre lies between its boundaries; a node outside it adds straight into HBM
                    const int lxi = bx + m - lx0, lyi = by + q - ly0;
                    if ((unsigned)lxi < (unsigned)lnx && (unsigned)lyi < (unsigned)lny) atomicAdd(&tl[lyi * lnx + lxi], v);
                    else atomicAdd(&lout[(by + q) * lp.n1 + bx + m], v);
                }
        }
    }
    if (lplane >= 0 && lp.count_box) {
        int d = ldropped;
        for (int o = 32; o > 0; o >>= 1) d += __shfl_down(d, o);
        if ((threadIdx.x & 63) == 0 && d) atomicAdd(&a.dropped[0], (unsigned long long)d);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (p >= np || !cb[p]) continue;
        int d = dropped[p];
        for (int o = 32; o > 0; o >>= 1) d += __shfl_down(d, o);
        if ((threadIdx.x & 63) == 0 && d) atomicAdd(&a.dropped[0], (unsigned long long)d);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < np * wcells; c += blockDim.x) {
        const int p = c / wcells, cc = c - p * wcells, r = cc / wnx;
        const unsigned long long v = tile[c];
        if (v) atomicAdd(&a.charge[(size_t)(plane0 + p) * a.charge_stride + (wy0 + r) * n1 + wx0 + (cc - r * wnx)], v);
    }
    for (int c = threadIdx.x; c < lcells; c += blockDim.x) {
        const int r = c / lnx;
        const unsigned long long v = tile[np * wcells + c];
        if (v) atomicAdd(&a.charge[(size_t)lplane * a.charge_stride + (ly0 + r) * lp.n1 + lx0 + (c - r * lnx)], v);
    }
}

// Items are (plane, chunk group), or (interaction, chunk group) when a.multi > 1 (its a.multi field
// planes together).  Persistent blocks (as many as fit the CUs) take items from a per-launch atomic
// counter, the larger source slices first, until it passes the item count.
template <int METHOD, class W>
__global__ __launch_bounds__(kDepositThreads) void k_deposit(DepositArgs a) {
    extern __shared__ unsigned long long tile[];
    if (failed(a.err)) return;
    const int units = a.multi > 1 ? a.nplanes / a.multi : a.nplanes;
    const int ngroups = (a.chunks + a.group - 1) / a.group, total = ngroups * units;
    __shared__ int next_item;
    for (;;) {
        if (threadIdx.x == 0) next_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = next_item;
        __syncthreads();   // every thread has read next_item before thread 0 takes the next one
        if (item >= total) break;
        const int pos = item / ngroups;
        if (a.multi > 1)   // the planes of one interaction are consecutive in the launch order
            deposit_item_multi<METHOD, W>(a, tile, a.order[pos * a.multi], a.multi, item - pos * ngroups);
        else
            deposit_item<METHOD, W>(a, tile, a.order[pos], item - pos * ngroups);
        __syncthreads();   // the item's LDS window is flushed before the next item zeroes it
    }
}

hipError_t launch_deposit(DepositArgs a, hipStream_t s) {
    if (a.nplanes <= 0) return hipSuccess;
    a.lds_cells = a.ncells < kDepositMaxLdsCells ? a.ncells : kDepositMaxLdsCells;
    const size_t lds = (size_t)a.lds_cells * sizeof(unsigned long long);
    if (!a.queue || !a.order) return hipErrorInvalidValue;
    if (a.multi > 3 || (a.multi > 1 && a.nplanes % a.multi != 0)) return hipErrorInvalidValue;
    const int total = (a.chunks + a.group - 1) / a.group * (a.multi > 1 ? a.nplanes / a.multi : a.nplanes);
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
#define OCTOPIC_DEPOSIT(M, W)                                                                                  \
    do {                                                                                                       \
        hipFuncSetAttribute((const void *)k_deposit<M, W>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        int per_cu = 0;                                                                                        \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_deposit<M, W>, kDepositThreads, lds) !=      \
                hipSuccess || per_cu < 1)                                                                      \
            per_cu = 1;                                                                                        \
        hipLaunchKernelGGL((k_deposit<M, W>), dim3(std::min(total, cus * per_cu)), dim3(kDepositThreads), lds, s, a); \
    } while (0)
    if (a.method == 0) {
        if (a.fp32) OCTOPIC_DEPOSIT(0, float); else OCTOPIC_DEPOSIT(0, double);
    } else {
        if (a.fp32) OCTOPIC_DEPOSIT(1, float); else OCTOPIC_DEPOSIT(1, double);
    }
#undef OCTOPIC_DEPOSIT
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------- kick
constexpr int kKickThreads = 512;
constexpr int kKickLdsBytes = 72 * 1024;   // LDS window of the field table per block (2 blocks per CU)

// LDS window layout: each node's C values padded to a whole number of 16-byte vectors (double2 /
// float4), rows of wny nodes plus one vector, so a node is read with 16-byte LDS loads and vertically
// adjacent nodes start in different banks.
template <class R> __device__ __forceinline__ int node_stride(int C) { return sizeof(R) == 8 ? (C + 1) & ~1 : (C + 3) & ~3; }
template <class R> __device__ __forceinline__ int row_pad() { return 16 / (int)sizeof(R); }

template <int MODE>
__device__ __forceinline__ void load_node(const double *e, double *v) {
    const double2 *p = reinterpret_cast<const double2 *>(e);
#pragma unroll
    for (int i = 0; i < (MODE == 1 ? 5 : 3); ++i) { const double2 t = p[i]; v[2 * i] = t.x; v[2 * i + 1] = t.y; }
}
template <int MODE>
__device__ __forceinline__ void load_node(const float *e, float *v) {
    const float4 *p = reinterpret_cast<const float4 *>(e);
#pragma unroll
    for (int i = 0; i < (MODE == 1 ? 3 : 2); ++i) {
        const float4 t = p[i]; v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
    }
}

// Stages table rows [x0, x1] x columns [y0, y1] in the window when they fit (all threads of the block
// call it; the caller barriers before the gathers).  Returns the row stride in elements, 0 if not.
template <class R>
__device__ __forceinline__ int stage_window(R *win, const R *tab, int ny, int C, int x0, int x1, int y0, int y1) {
    const int CL = node_stride<R>(C), wny = y1 - y0 + 1, rs = wny * CL + row_pad<R>();
    if ((x1 - x0 + 1) * rs * (int)sizeof(R) > kKickLdsBytes) return 0;
    const int run = wny * C;
    for (int k = threadIdx.x; k < (x1 - x0 + 1) * run; k += blockDim.x) {
        const int r = k / run, e = k - r * run, node = e / C, c = e - node * C;
        win[r * rs + node * CL + c] = tab[((size_t)(x0 + r) * ny + y0) * C + e];
    }
    return rs;
}

// Gather of one particle's field from the interleaved table rows starting at e0 (row = one x, rows
// `rs` doubles apart, C doubles per cell), for the linear (MODE 0) or quadratic (MODE 1) slice
// interpolation (pic_cpu.jl:1218-1273, 2596-2680).
template <int METHOD, int MODE, class R>
__device__ __forceinline__ void gather_field(const R *e0, int rs, int C, const R wx[3], const R wy[3], R zL, R &Kx, R &Ky,
                                             R &Kz) {
    constexpr int K = METHOD == 0 ? 2 : 3;
    if (MODE == 0) {
        const R zR = (R)1 - zL;
#pragma unroll
        for (int m = 0; m < K; ++m)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const R *e = e0 + m * rs + q * C;
                const R w = wx[m] * wy[q];
                Kx += w * (zL * e[1] + zR * e[4]);
                Ky += w * (zL * e[2] + zR * e[5]);
                Kz += w * (e[0] - e[3]);
            }
    } else {
        const R tt = (R)1 - zL;
        const R t2 = tt * tt;
        const R aL = (R)2 * t2 - (R)3 * tt + (R)1, aM = (R)4 * tt - (R)4 * t2, aR = (R)2 * t2 - tt;
        const R bL = (R)3 - (R)4 * tt, bM = (R)8 * tt - (R)4, bR = (R)1 - (R)4 * tt;
#pragma unroll
        for (int m = 0; m < K; ++m)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const R *e = e0 + m * rs + q * C;   // L = 0..2, R = 3..5, M = 6..8
                const R w = wx[m] * wy[q];
                Kx += w * (aL * e[1] + aM * e[7] + aR * e[4]);
                Ky += w * (aL * e[2] + aM * e[8] + aR * e[5]);
                Kz += w * (bL * e[0] + bM * e[6] + bR * e[3]);
            }
    }
}

// The same gather from the LDS window (node stride CL, rows rs elements apart), 16-byte loads per node.
template <int METHOD, int MODE, class R>
__device__ __forceinline__ void gather_field_lds(const R *e0, int rs, int CL, const R wx[3], const R wy[3], R zL, R &Kx,
                                                 R &Ky, R &Kz) {
    constexpr int K = METHOD == 0 ? 2 : 3;
    R v[12];
    if (MODE == 0) {
        const R zR = (R)1 - zL;
#pragma unroll
        for (int m = 0; m < K; ++m)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                load_node<MODE>(e0 + m * rs + q * CL, v);
                const R w = wx[m] * wy[q];
                Kx += w * (zL * v[1] + zR * v[4]);
                Ky += w * (zL * v[2] + zR * v[5]);
                Kz += w * (v[0] - v[3]);
            }
    } else {
        const R tt = (R)1 - zL;
        const R t2 = tt * tt;
        const R aL = (R)2 * t2 - (R)3 * tt + (R)1, aM = (R)4 * tt - (R)4 * t2, aR = (R)2 * t2 - tt;
        const R bL = (R)3 - (R)4 * tt, bM = (R)8 * tt - (R)4, bR = (R)1 - (R)4 * tt;
#pragma unroll
        for (int m = 0; m < K; ++m)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                load_node<MODE>(e0 + m * rs + q * CL, v);
                const R w = wx[m] * wy[q];
                Kx += w * (aL * v[1] + aM * v[7] + aR * v[4]);
                Ky += w * (aL * v[2] + aM * v[8] + aR * v[5]);
                Kz += w * (bL * v[0] + bM * v[6] + bR * v[3]);
            }
    }
}

// One work item: part `sub` of chunk `chunk` of interaction d's field slice (called by every thread
// of the block).  In the slice-pair modes the table cells of the chunk's stencil window (from the
// chunk's field box) are staged in LDS when they fit; a particle whose stencil leaves the window reads
// the table in HBM.  R = double (FP64 mode) or float (FP32 mode: table, weights and gather sums in
// FP32; the kick is applied to the FP64 coordinates).
template <int METHOD, int MODE, class R>
__device__ __forceinline__ void kick_item(const KickArgs &a, R *win, int d, int chunk, int sub) {
    const Inter it = a.inters[d];
    if (!it.valid) return;
    const int64_t n = it.fld_n;
    const int64_t clo = n * chunk / a.chunks, chi = n * (chunk + 1) / a.chunks;
    const int64_t lo = clo + (chi - clo) * sub / a.subdiv, hi = clo + (chi - clo) * (sub + 1) / a.subdiv;
    const int nx = a.nx, ny = a.ny, C = 3 * a.npl;
    const R *tab = static_cast<const R *>(a.table) + (size_t)d * nx * ny * C;
    double *X = a.st.c[it.fld_beam][0] + it.fld_off, *PX = a.st.c[it.fld_beam][1] + it.fld_off;
    double *Y = a.st.c[it.fld_beam][2] + it.fld_off, *PY = a.st.c[it.fld_beam][3] + it.fld_off;
    const double *Z = a.st.c[it.fld_beam][4] + it.fld_off;
    double *PZ = a.st.c[it.fld_beam][5] + it.fld_off;
    // _slice_interpolation_parameters (pic_cpu.jl:1540-1548)
    double hzi, zbias;
    {
        const double denom = it.rb - it.lb;
        if (!__builtin_isfinite(denom) || denom == 0.0) { hzi = 0.0; zbias = 0.5; }
        else { hzi = 1.0 / denom; zbias = it.rb * hzi; }
    }
    const double hx = it.fg.width / (double)(nx - 1), hy = it.fg.height / (double)(ny - 1);
    const double hxR = it.fgR.width / (double)(nx - 1), hyR = it.fgR.height / (double)(ny - 1);
    const double c = it.c_src, ks = it.kick_scale;
    const int lk = a.lkick;
    constexpr int K = METHOD == 0 ? 2 : 3;
    // LDS window (uniform per block)
    int wx0 = 0, wy0 = 0, wnx = 0, wny = 0, wrs = 0;   // window origin, extent, row stride (elements)
    if (MODE != 2 && a.chunkbox && lo < hi && a.pass_windows != 2) {
        const double *b = a.chunkbox + ((size_t)(2 * (d >> 1) + it.fld_beam) * a.chunks + chunk) * 8 + 4;
        int x0, x1, y0, y1;
        stencil_window<METHOD>((b[0] - it.fg.x0) / hx, (b[1] - it.fg.x0) / hx, nx, x0, x1);
        stencil_window<METHOD>((b[2] - it.fg.y0) / hy, (b[3] - it.fg.y0) / hy, ny, y0, y1);
        wrs = stage_window<R>(win, tab, ny, C, x0, x1, y0, y1);
        if (wrs) {
            wx0 = x0; wy0 = y0; wnx = x1 - x0 + 1; wny = y1 - y0 + 1;
            __syncthreads();
        }
    }
    int dropped = 0;
    const double fx1 = it.fg.x0 + it.fg.width, fy1 = it.fg.y0 + it.fg.height;
    // A chunk whose window does not fit stages a window per pass instead: the stencil nodes of the
    // pass's particles (consecutive in tile order, so compact up to the drift) when they fit.
    const bool per_pass = a.pass_windows && MODE != 2 && a.chunkbox && lo < hi && wnx == 0;
    __shared__ int pass_box[kKickThreads / 64][4];
    for (int64_t base = lo; base < hi; base += blockDim.x) {   // block-uniform trip count
        const int64_t t = base + threadIdx.x;
        const bool act = t < hi;
        double x = 0.0, px = 0.0, y = 0.0, py = 0.0, z = 0.0, pz = 0.0;
        if (act) { x = X[t]; px = PX[t]; y = Y[t]; py = PY[t]; z = Z[t]; pz = PZ[t]; }
        const double s = 0.5 * (z - c);
        x += s * px;
        y += s * py;
        if (lk) pz -= 0.25 * (px * px + py * py);
        if (act && it.count_field && !(__builtin_isfinite(x) && it.fg.x0 <= x && x <= fx1 && __builtin_isfinite(y) &&
                                       it.fg.y0 <= y && y <= fy1))
            ++dropped;
        const double zLd = jclamp(-z * hzi + zbias, 0.0, 1.0);
        const R zL = (R)zLd;
        R Kx = 0, Ky = 0, Kz = 0;
        int bx, by;
        double wxd[3], wyd[3];
        const double ux = (x - it.fg.x0) / hx, uy = (y - it.fg.y0) / hy;
        stencil<METHOD>(ux, nx, bx, wxd);
        stencil<METHOD>(uy, ny, by, wyd);
        const R wx[3] = {(R)wxd[0], (R)wxd[1], (R)wxd[2]}, wy[3] = {(R)wyd[0], (R)wyd[1], (R)wyd[2]};
        if (per_pass) {
            // stencil-node box of the pass's in-range particles (out-of-range ones read node 0 with
            // zero weights, from HBM)
            const bool inr = act && 0.0 <= ux && ux <= (double)(nx - 1) && 0.0 <= uy && uy <= (double)(ny - 1);
            int b0 = inr ? bx : 0x7fffffff, b1 = inr ? bx + K - 1 : -1;
            int b2 = inr ? by : 0x7fffffff, b3 = inr ? by + K - 1 : -1;
            for (int o = 32; o > 0; o >>= 1) {
                b0 = min(b0, __shfl_xor(b0, o)); b1 = max(b1, __shfl_xor(b1, o));
                b2 = min(b2, __shfl_xor(b2, o)); b3 = max(b3, __shfl_xor(b3, o));
            }
            __syncthreads();   // the previous pass's gathers are done with the window and pass_box
            const int wv = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) { pass_box[wv][0] = b0; pass_box[wv][1] = b1; pass_box[wv][2] = b2; pass_box[wv][3] = b3; }
            __syncthreads();
            for (int v = 0; v < (int)(blockDim.x >> 6); ++v) {
                b0 = min(b0, pass_box[v][0]); b1 = max(b1, pass_box[v][1]);
                b2 = min(b2, pass_box[v][2]); b3 = max(b3, pass_box[v][3]);
            }
            wnx = 0;
            wrs = b1 >= 0 ? stage_window<R>(win, tab, ny, C, b0, b1, b2, b3) : 0;
            if (wrs) { wx0 = b0; wy0 = b2; wnx = b1 - b0 + 1; wny = b3 - b2 + 1; }
            __syncthreads();
        }
        if (!act) continue;
        if (MODE != 2) {
            constexpr int GM = MODE == 3 ? 0 : MODE;   // the add-back mode gathers as the linear mode
            if (bx >= wx0 && bx + K <= wx0 + wnx && by >= wy0 && by + K <= wy0 + wny)
                gather_field_lds<METHOD, GM, R>(win + (bx - wx0) * wrs + (by - wy0) * node_stride<R>(C), wrs,
                                                node_stride<R>(C), wx, wy, zL, Kx, Ky, Kz);
            else
                gather_field<METHOD, GM, R>(tab + ((size_t)bx * ny + by) * C, ny * C, C, wx, wy, zL, Kx, Ky, Kz);
        } else {
            // :node (pic_cpu.jl:1507-1530): L on its own mesh, R on its own mesh, phi_L - phi_Z on L's mesh
            R KxL = 0, KyL = 0, KxR = 0, KyR = 0;
#pragma unroll
            for (int m = 0; m < K; ++m)
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const R *e = tab + ((size_t)(bx + m) * ny + (by + q)) * C;
                    const R w = wx[m] * wy[q];
                    KxL += w * ((R)1 * e[1] + (R)0 * e[1]);
                    KyL += w * ((R)1 * e[2] + (R)0 * e[2]);
                    Kz += w * (e[0] - e[6]);
                }
            int bxr, byr;
            double wxrd[3], wyrd[3];
            stencil<METHOD>((x - it.fgR.x0) / hxR, nx, bxr, wxrd);
            stencil<METHOD>((y - it.fgR.y0) / hyR, ny, byr, wyrd);
            const R wxr[3] = {(R)wxrd[0], (R)wxrd[1], (R)wxrd[2]}, wyr[3] = {(R)wyrd[0], (R)wyrd[1], (R)wyrd[2]};
#pragma unroll
            for (int m = 0; m < K; ++m)
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const R *e = tab + ((size_t)(bxr + m) * ny + (byr + q)) * C;
                    const R w = wxr[m] * wyr[q];
                    KxR += w * ((R)1 * e[4] + (R)0 * e[4]);
                    KyR += w * ((R)1 * e[5] + (R)0 * e[5]);
                }
            const R zR = (R)1 - zL;
            Kx = zL * KxL + zR * KxR;
            Ky = zL * KyL + zR * KyR;
        }
        if (MODE == 3 && it.gauss) {
            // Gaussian-subtracted solver (gaussian_pic.jl:864-880): the residual's mesh kick plus the exact
            // field of the subtracted Gaussian, precomputed per particle at this same drifted position
            const double *ab = a.ab + 4 * (a.ab_base[it.fld_beam] + it.fld_off + t);
            px += ks * (double)Kx + ab[0];
            py += ks * (double)Ky + ab[1];
            if (lk) {
                pz += ks * (double)Kz * hzi;
                pz += ab[2];
                pz += ab[3];
            }
        } else {
            px += ks * (double)Kx;
            py += ks * (double)Ky;
            if (lk) pz += ks * (double)Kz * hzi;
        }
        const double s2 = 0.5 * (c - z);
        x += s2 * px;
        y += s2 * py;
        if (lk) pz += 0.25 * (px * px + py * py);
        X[t] = x; PX[t] = px; Y[t] = y; PY[t] = py; PZ[t] = pz;
    }
    if (it.count_field) {
        for (int o = 32; o > 0; o >>= 1) dropped += __shfl_down(dropped, o);
        if ((threadIdx.x & 63) == 0 && dropped) atomicAdd(&a.dropped[0], (unsigned long long)dropped);
    }
}

// Items are (interaction, chunk, part), one per block, in the launch order of the interactions (larger
// field slices first): the dispatcher hands blocks out in that order as CUs free up, which balances the
// launch without a work-queue counter (a shared atomic counter measured 12 % slower: its device-scope
// atomics serialise across the XCDs).  In the slice-pair modes the register budget is held to 4 waves
// per SIMD (two blocks per CU, as the LDS window allows; a few values spill); the node mode keeps its
// registers (its spills would cost more).
template <int METHOD, int MODE, class R>
__global__ __launch_bounds__(kKickThreads, (MODE == 2 ? 1 : 4)) void k_kick(KickArgs a) {
    extern __shared__ __align__(16) unsigned char win_raw[];
    R *win = reinterpret_cast<R *>(win_raw);
    if (failed(a.err)) return;
    const int per_inter = a.chunks * a.subdiv;
    const int item = blockIdx.x, dl = item / per_inter, rem = item - dl * per_inter;
    kick_item<METHOD, MODE, R>(a, win, a.order[dl], rem / a.subdiv, rem % a.subdiv);
}

hipError_t launch_kick(const KickArgs &a, hipStream_t s) {
    if (a.ninters <= 0) return hipSuccess;
    if (a.subdiv < 1 || !a.order) return hipErrorInvalidValue;
    const dim3 g(a.chunks * a.subdiv * a.ninters), b(kKickThreads);   // one block per item
    const size_t lds = (a.mode != 2 && a.chunkbox) ? kKickLdsBytes : 0;
#define OCTOPIC_KICK(M, MODE)                                                                                    \
    do {                                                                                                         \
        if (a.fp32) {                                                                                            \
            hipFuncSetAttribute((const void *)k_kick<M, MODE, float>, hipFuncAttributeMaxDynamicSharedMemorySize,  \
                                (int)lds);                                                                       \
            hipLaunchKernelGGL((k_kick<M, MODE, float>), g, b, lds, s, a);                                       \
        } else {                                                                                                 \
            hipFuncSetAttribute((const void *)k_kick<M, MODE, double>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)lds);                                                                       \
            hipLaunchKernelGGL((k_kick<M, MODE, double>), g, b, lds, s, a);                                      \
        }                                                                                                        \
    } while (0)
    if (a.mode == 3 && !a.ab) return hipErrorInvalidValue;
    if (a.method == 0) {
        if (a.mode == 0) OCTOPIC_KICK(0, 0);
        else if (a.mode == 1) OCTOPIC_KICK(0, 1);
        else if (a.mode == 2) OCTOPIC_KICK(0, 2);
        else OCTOPIC_KICK(0, 3);
    } else {
        if (a.mode == 0) OCTOPIC_KICK(1, 0);
        else if (a.mode == 1) OCTOPIC_KICK(1, 1);
        else if (a.mode == 2) OCTOPIC_KICK(1, 2);
        else OCTOPIC_KICK(1, 3);
    }
#undef OCTOPIC_KICK
    return hipGetLastError();
}

}  // namespace octopic

// ==================================================================================== :node meshes
// _pic_build_node_grids! / _pic_prebuild_node_caches! (pic_cpu.jl:1317-1419): at turn start, for every
// (source slice, direction), one mesh per node b of the field beam's slicing.  Node b's source box
// spans the source drifted to nodes b and b+1; its field box is the union of the two field slices
// adjacent to the node, each drifted to the source centre.  Built from the turn-start states, before
// any pair, as the reference requires for schedule independence.
namespace octopic {

constexpr int kNodeThreads = 256;

// Folds this wave's 4 extrema (x min, x max, y min, y max) into 4 bound slots with 64-bit atomicMax.
__device__ __forceinline__ void wave_fold_box(const double v[4], unsigned long long *slots) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        unsigned long long u = bound_slot(v[q], (q & 1) == 0);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long w = __shfl_xor(u, o);
            u = w > u ? w : u;
        }
        if ((threadIdx.x & 63) == 0 && u) atomicMax(&slots[q], u);
    }
}

constexpr int kNodeGroup = 4;   // node drifts (source bounds) or source slices (field bounds) per pass

// grid (nchunks, ns_src): source boxes at every node drift, folded into a.skeys[(i*nb + b)*4 + q];
// kNodeGroup drifts per pass over the chunk
__global__ __launch_bounds__(kNodeThreads) void k_node_src_bounds(NodeArgs a) {
    if (failed(a.err)) return;
    const int i = blockIdx.y;
    const int nbF = a.tab_fld->ns + 1;
    const double c = a.tab_src->center[i];
    const int64_t off = a.tab_src->offset[i], n = a.tab_src->count[i];
    const int64_t lo = n * blockIdx.x / a.nchunks, hi = n * (blockIdx.x + 1) / a.nchunks;
    const double *X = a.st.c[a.src_beam][0] + off, *PX = a.st.c[a.src_beam][1] + off;
    const double *Y = a.st.c[a.src_beam][2] + off, *PY = a.st.c[a.src_beam][3] + off;
    for (int b0 = 0; b0 < nbF; b0 += kNodeGroup) {
        double d[kNodeGroup], v[kNodeGroup][4];
#pragma unroll
        for (int g = 0; g < kNodeGroup; ++g) {
            d[g] = b0 + g < nbF ? 0.5 * (c - a.tab_fld->boundary[b0 + g]) : 0.0;
            v[g][0] = INFINITY; v[g][1] = -INFINITY; v[g][2] = INFINITY; v[g][3] = -INFINITY;
        }
        for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
            const double x = X[t], px = PX[t], y = Y[t], py = PY[t];
#pragma unroll
            for (int g = 0; g < kNodeGroup; ++g) {
                const double xv = x + px * d[g], yv = y + py * d[g];
                v[g][0] = jmin(v[g][0], xv); v[g][1] = jmax(v[g][1], xv);
                v[g][2] = jmin(v[g][2], yv); v[g][3] = jmax(v[g][3], yv);
            }
        }
        for (int g = 0; g < kNodeGroup && b0 + g < nbF; ++g) wave_fold_box(v[g], a.skeys + ((size_t)i * nbF + b0 + g) * 4);
    }
}

// grid (nchunks, ns_fld, ceil(ns_src / kNodeGroup)): field slice sl drifted to the centres of
// kNodeGroup source slices i, folded into a.fkeys[(i*ns_fld + sl)*4 + q]
__global__ __launch_bounds__(kNodeThreads) void k_node_fld_bounds(NodeArgs a) {
    if (failed(a.err)) return;
    const int sl = blockIdx.y, i0 = blockIdx.z * kNodeGroup;
    const int nsF = a.tab_fld->ns, nsS = a.tab_src->ns;
    const int64_t off = a.tab_fld->offset[sl], n = a.tab_fld->count[sl];
    const int64_t lo = n * blockIdx.x / a.nchunks, hi = n * (blockIdx.x + 1) / a.nchunks;
    const double *X = a.st.c[a.fld_beam][0] + off, *PX = a.st.c[a.fld_beam][1] + off;
    const double *Y = a.st.c[a.fld_beam][2] + off, *PY = a.st.c[a.fld_beam][3] + off;
    const double *Z = a.st.c[a.fld_beam][4] + off;
    double c[kNodeGroup], v[kNodeGroup][4];
#pragma unroll
    for (int g = 0; g < kNodeGroup; ++g) {
        c[g] = i0 + g < nsS ? a.tab_src->center[i0 + g] : 0.0;
        v[g][0] = INFINITY; v[g][1] = -INFINITY; v[g][2] = INFINITY; v[g][3] = -INFINITY;
    }
    for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        const double x = X[t], px = PX[t], y = Y[t], py = PY[t], z = Z[t];
#pragma unroll
        for (int g = 0; g < kNodeGroup; ++g) {
            const double sh = 0.5 * (z - c[g]);
            const double xv = x + sh * px, yv = y + sh * py;
            v[g][0] = jmin(v[g][0], xv); v[g][1] = jmax(v[g][1], xv);
            v[g][2] = jmin(v[g][2], yv); v[g][3] = jmax(v[g][3], yv);
        }
    }
    for (int g = 0; g < kNodeGroup && i0 + g < nsS; ++g) wave_fold_box(v[g], a.fkeys + ((size_t)(i0 + g) * nsF + sl) * 4);
}

__device__ __forceinline__ void read_box(const unsigned long long *slots, double v[4]) {
    for (int q = 0; q < 4; ++q) v[q] = bound_value(slots[q], (q & 1) == 0);
}

// one lane per (source slice i, node b)
__global__ void k_node_meshes(NodeArgs a) {
    const int nsS = a.tab_src->ns, nsF = a.tab_fld->ns, nbF = nsF + 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsS * nbF) return;
    a.jobs[k].valid = 0;
    a.nodes[k].valid = 0;
    if (failed(a.err)) return;
    const int i = k / nbF, b = k % nbF;
    CacheGeo &node = a.nodes[k];
    GreenJob &job = a.jobs[k];
    node.valid = 0;
    job.valid = 0;
    job.slot = a.slot0 + k;
    if (a.tab_src->count[i] == 0) return;
    const int bn = b + 1 < nbF ? b + 1 : nbF - 1;
    double s0[4], s1[4];
    read_box(a.skeys + ((size_t)i * nbF + b) * 4, s0);
    read_box(a.skeys + ((size_t)i * nbF + bn) * 4, s1);
    for (int q = 0; q < 4; ++q)
        if (s0[q] != s0[q] || s1[q] != s1[q]) { set_error(a.err, OCTOPIC_ERR_NONFINITE, -1, a.dir, 4, 0, 0); return; }
    const double sxmin = jmin(s0[0], s1[0]), sxmax = jmax(s0[1], s1[1]);
    const double symin = jmin(s0[2], s1[2]), symax = jmax(s0[3], s1[3]);
    double fxmin = INFINITY, fxmax = -INFINITY, fymin = INFINITY, fymax = -INFINITY;
    for (int adj = b - 1; adj <= b; ++adj) {
        if (adj < 0 || adj >= nsF) continue;
        double f[4];
        read_box(a.fkeys + ((size_t)i * nsF + adj) * 4, f);
        for (int q = 0; q < 4; ++q)
            if (f[q] != f[q]) { set_error(a.err, OCTOPIC_ERR_NONFINITE, -1, a.dir, 4, 0, 0); return; }
        if (!__builtin_isfinite(f[0])) continue;
        fxmin = jmin(fxmin, f[0]); fxmax = jmax(fxmax, f[1]);
        fymin = jmin(fymin, f[2]); fymax = jmax(fymax, f[3]);
    }
    if (!__builtin_isfinite(fxmin)) return;
    Grid sg, fg;
    if (!interaction_grids(a.o, sxmin, sxmax, symin, symax, fxmin, fxmax, fymin, fymax, sg, fg)) {
        set_error(a.err, OCTOPIC_ERR_DEGENERATE_EXTENT, -1, a.dir, 3, jmax(sxmax - sxmin, fxmax - fxmin),
                  jmax(symax - symin, fymax - fymin));
        return;
    }
    node.valid = 1;
    node.sg = sg;
    node.fg = fg;
    job.valid = 1;
    job.fx0 = fg.x0; job.fy0 = fg.y0; job.sx0 = sg.x0; job.sy0 = sg.y0;
    job.hx = sg.width / (double)(a.o.nx - 1);
    job.hy = sg.height / (double)(a.o.ny - 1);
    atomicAdd(&a.counters[3], 1ull);
}

hipError_t launch_node_prepare(const NodeArgs &a, int ns_src, int ns_fld, hipStream_t s) {
    hipLaunchKernelGGL(k_node_src_bounds, dim3(a.nchunks, ns_src), dim3(kNodeThreads), 0, s, a);
    hipLaunchKernelGGL(k_node_fld_bounds, dim3(a.nchunks, ns_fld, (ns_src + kNodeGroup - 1) / kNodeGroup),
                       dim3(kNodeThreads), 0, s, a);
    const int nn = ns_src * (ns_fld + 1);
    hipLaunchKernelGGL(k_node_meshes, dim3((nn + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

__device__ void set_node_plane(Plane &pl, int d, int beam, const GeomOpts &o, int slot, int64_t off, int64_t n, double s,
                               const Grid &g, int count) {
    const double hx = g.width / (double)(o.nx - 1), hy = g.height / (double)(o.ny - 1);
    set_plane(pl, d, 0, beam, o.nx, o.ny, slot, off, n, s, g.x0, g.y0, hx, hy);
    pl.count_box = count;
    pl.bx1 = g.x0 + g.width;
    pl.by1 = g.y0 + g.height;
}

// _pic_interaction_node! descriptors (pic_cpu.jl:1448-1538): one lane per (pair, direction).  A pair
// runs only when the node meshes of both its directions exist (pic_cpu.jl:1465-1470).
__global__ void k_node_geometry(NodeGeometryArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= 2 * a.npairs) return;
    const int k = lane >> 1, dir = (lane & 1) + 1, d = lane;
    neutralise_inter(k, dir, a.npairs, a.o, a.inters, a.planes, nullptr);
    double vs[4], vf[4], l1[4], l2[4];
    take_boxes(a.keys, k, dir, vs, vf, l1, l2);
    if (failed(a.err)) return;
    const GeomOpts &o = a.o;
    const BatchPair pr = a.pairs[k];
    const int i = pr.i, j = pr.j, p = pr.p;
    bool ready = true;
    for (int dd = 1; dd <= 2; ++dd) {
        const int si = dd == 1 ? i : j, fi = dd == 1 ? j : i;
        const int nbF = (dd == 1 ? a.tab2 : a.tab1)->ns + 1;
        const CacheGeo *nodes = dd == 1 ? a.nodes1 : a.nodes2;
        ready = ready && nodes[si * nbF + fi].valid && nodes[si * nbF + fi + 1].valid;
    }
    if (dir == 1) a.ran[p] = ready ? 1 : 0;
    if (!ready) return;
    const SliceTable *ts = dir == 1 ? a.tab1 : a.tab2, *tf = dir == 1 ? a.tab2 : a.tab1;
    const int si = dir == 1 ? i : j, fi = dir == 1 ? j : i;
    const int nbF = tf->ns + 1;
    const CacheGeo *nodes = dir == 1 ? a.nodes1 : a.nodes2;
    const int slot0 = dir == 1 ? a.slot1 : a.slot2;
    const int kl = si * nbF + fi, kr = si * nbF + fi + 1;
    const CacheGeo gL = nodes[kl], gR = nodes[kr];
    Inter &it = a.inters[d];
    fill_inter(it, o, p, i, j, dir, ts, tf, si, fi);
    it.sg = gL.sg; it.fg = gL.fg; it.fgR = gR.fg;
    it.plane0 = d * o.npl;
    it.nplanes = o.lkick ? 3 : 2;       // F_Z only with the longitudinal kick (pic_cpu.jl:1497-1503)
    it.event = 0;
    it.count_field = 1;
    const int sb = it.src_beam;
    set_node_plane(a.planes[d * o.npl + 0], d, sb, o, slot0 + kl, it.src_off, it.src_n, it.sL, gL.sg, 1);
    set_node_plane(a.planes[d * o.npl + 1], d, sb, o, slot0 + kr, it.src_off, it.src_n, it.sR, gR.sg, 1);
    set_node_plane(a.planes[d * o.npl + 2], d, sb, o, slot0 + kl, it.src_off, it.src_n, it.sR, gL.sg, 0);
    fill_record(a.records[2 * p + dir - 1], p, i, j, dir, 0, gL.sg, gL.fg);
    it.valid = 1;
    if (dir == 1 && o.compute_lum)
        lum_geometry(o, l1, l2, k, p, a.tab1, a.tab2, i, j, a.planes, 2 * a.npairs * o.npl, a.lumgeo, a.err);
}

hipError_t launch_node_geometry(const NodeGeometryArgs &a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_node_geometry, dim3((2 * a.npairs + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace octopic
