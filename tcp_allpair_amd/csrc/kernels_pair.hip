// All-pair real-space kernels: Ewald erfc for every pair within rc plus the
// electron-ion Deutsch term -exp(-r/lambda)/r (force.f90:87-178, 215-341).
//
// Shape (both precisions): one i particle per lane held in registers, 256 i per
// block (blockIdx.x), the j range split into chunks (blockIdx.y).  j is uniform
// across the wavefront, so its coordinates are scalar loads broadcast to all
// 64 lanes and no LDS staging is needed for them; the j loop runs over the
// electron part and then the ion part of the chunk so the pair species and the
// Deutsch branch are uniform per wavefront except in the one wavefront that
// straddles the electron/ion boundary of the i set.  Every ordered pair (i,j)
// is evaluated (no Newton's third law), so each i sums its own force without
// atomics and the per-(chunk,i) partials are summed in a fixed order later:
// FP64 results do not depend on scheduling.  RDF counts go to a per-block LDS
// histogram (3 x 512 uint32) flushed with one global 64-bit atomic per bin.
#include "tcp_internal.h"

namespace tcp {

namespace {

constexpr double kSqrtPi = 1.7724538509055159;  // data.f90:24

__device__ inline void hist_clear(unsigned* h)
{
    for (int b = threadIdx.x; b < 3 * kHistStride; b += blockDim.x) h[b] = 0u;
}

__device__ inline void hist_flush(const unsigned* h, unsigned long long* g, unsigned weight)
{
    for (int b = threadIdx.x; b < 3 * kHistStride; b += blockDim.x) {
        unsigned c = h[b];
        if (c) atomicAdd(g + b, (unsigned long long)c * weight);
    }
}

// ------------------------------------------------------------------ FP64 ---

template <bool kRdf>
__global__ __launch_bounds__(kPairBlock) void pair_f64_kernel(PairArgs a)
{
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    if (kRdf) {
        hist_clear(hist);
        __syncthreads();
    }
    const int i = blockIdx.x * kPairBlock + threadIdx.x;
    const bool live = i < a.n_i;
    const int ii = live ? i : a.n_i - 1;
    const double xi = a.pi.x[ii], yi = a.pi.y[ii], zi = a.pi.z[ii];
    const int si = ii >= a.ne_i;
    const int j0 = blockIdx.y * a.jchunk;
    const int j1 = min(a.n_j, j0 + a.jchunk);
    const double Lx = a.box[0], Ly = a.box[1], Lz = a.box[2];
    const double iLx = a.inv_box[0], iLy = a.inv_box[1], iLz = a.inv_box[2];
    const double alpha = a.alpha, a2 = a.alpha * a.alpha, rc2 = a.rc2;
    const double two_a_sqrtpi = 2.0 * a.alpha / kSqrtPi;
    const double inv_rs = 1.0 / a.rs;
    double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0;
    for (int sj = 0; sj < 2; ++sj) {
        const int lo = sj ? max(j0, a.ne_j) : j0;
        const int hi = sj ? j1 : min(j1, a.ne_j);
        const double qq = a.q[si] * a.q[sj];
        const double eqq = a.eps * qq;
        const bool ei = si != sj;
        const int ht = si + sj;  // 0 ee, 1 ei, 2 ii
        const double dxb = a.dx[ht];
#pragma unroll 2
        for (int j = lo; j < hi; ++j) {
            double xr, yr, zr, r2;
            {
                // minimum image and r^2 in the reference's operation order
                // (force.f90:89-94), without contraction, so that the cut-off
                // test and the RDF bin match the FP64 reference bit for bit
#pragma clang fp contract(off)
                const double dx = xi - a.pj.x[j];
                const double dy = yi - a.pj.y[j];
                const double dz = zi - a.pj.z[j];
                xr = dx - Lx * rint(dx * iLx);
                yr = dy - Ly * rint(dy * iLy);
                zr = dz - Lz * rint(dz * iLz);
                r2 = xr * xr + yr * yr + zr * zr;
            }
            if (r2 < rc2 && r2 > 0.0) {
                const double r = sqrt(r2);
                if (kRdf && live) {
                    const int b = (int)(r / dxb);
                    if (b < kNbin) atomicAdd(&hist[ht * kHistStride + b], 1u);
                }
                const double inv_r = 1.0 / r;
                const double inv_r2 = inv_r * inv_r;
                const double ec = erfc(alpha * r);
                const double g = two_a_sqrtpi * r * exp(-(a2 * r2));
                double num = g + ec, en = ec;
                if (ei) {
                    const double y = exp(-r * inv_rs);
                    num -= y * (1.0 + r * inv_rs);
                    en -= y;
                }
                const double df = eqq * num * inv_r2 * inv_r;
                u += qq * en * inv_r;
                fx += df * xr;
                fy += df * yr;
                fz += df * zr;
            }
        }
    }
    if (live) {
        const size_t n = a.n_i;
        double* pf = a.pf + (size_t)blockIdx.y * 3 * n;
        pf[i] = fx;
        pf[n + i] = fy;
        pf[2 * n + i] = fz;
        a.pu[(size_t)blockIdx.y * n + i] = u;
    }
    if (kRdf) {
        __syncthreads();
        hist_flush(hist, a.rdf, a.rdf_weight);
    }
}

// ----------------------------------------------------------------- MIXED ---
//
// Separations come from FixPos: the wrapped 32-bit difference of the fixed-point
// coordinates is the minimum-image separation exactly, and the FP32 residuals
// restore the FP64 position to FP32 rounding, dx = float(int(u_i - u_j)) s +
// (r_i - r_j): no rint and no cancellation even for pairs 0.01 A apart in a
// 50 A box.  erfc(x) = exp(-x^2) P(w), w = (x/2)/(1 + x/2), with P of degree 9
// and P(0) = 1 fitted here (weighted least squares plus Lawson iterations
// against exp(x^2) erfc(x) on [0, 4.5]; relative error 5.6e-9 in exact
// arithmetic, 3e-8 near x = 0 and below 7.2e-7 up to x = 3.2 once evaluated in
// FP32).  exp(-x^2) is shared with the 2x exp(-x^2)/sqrt(pi) force term.
// Electron-ion pairs closer than 2 lambda take a rare, divergent branch
// (close_ei) where the Deutsch and Ewald parts are evaluated without their
// mutual cancellation.
// Per-lane sums run in FP32 over tiles of 32 j and go to FP64 after each tile.

__device__ inline float erfcx_poly(float w)
{
    float p = 2.674967051e-02f;
    p = fmaf(p, w, -1.345596015e-01f);
    p = fmaf(p, w, 1.725731939e-01f);
    p = fmaf(p, w, 4.970539361e-03f);
    p = fmaf(p, w, 3.443385288e-02f);
    p = fmaf(p, w, -3.166950941e-01f);
    p = fmaf(p, w, -2.740879953e-01f);
    p = fmaf(p, w, 1.743201137e+00f);
    p = fmaf(p, w, -2.256757498e+00f);
    p = fmaf(p, w, 1.0f);
    return p;
}

// Close electron-ion pair (r < 2 lambda), where erfc + g and e^-s (1 + s) are
// both near 1 and cancel: num = B(s) - A(x), en = B(s) + s e^-s - erf(x) with
//   B(s)/s^2 by a degree-8 polynomial fitted here on s in [0, 2] (relative error
//            1.5e-9 in exact arithmetic, 3.3e-7 evaluated in FP32),
//   A(x)/x^3 = (2/sqrt(pi)) sum_n (-1)^(n+1) 2n/(n!(2n+1)) x^(2n-2), n = 1..8,
//   erf(x)/x = (2/sqrt(pi)) sum_n (-1)^n x^(2n)/(n!(2n+1)),        n = 0..7,
// the two Taylor series truncated for x < 0.6; larger x falls back to erfc.
__device__ inline void close_ei(float r, float s, float x, float y, float ec, float g,
                                float* num, float* en)
{
    float b = 9.974618998e-07f;
    b = fmaf(b, s, -1.720568391e-05f);
    b = fmaf(b, s, 1.642914247e-04f);
    b = fmaf(b, s, -1.179569318e-03f);
    b = fmaf(b, s, 6.936748278e-03f);
    b = fmaf(b, s, -3.333021883e-02f);
    b = fmaf(b, s, 1.249993556e-01f);
    b = fmaf(b, s, -3.333332808e-01f);
    b = fmaf(b, s, 4.999999993e-01f);
    const float B = b * s * s;
    const float z = x * x;
    float pa = -2.633938299e-05f;
    pa = fmaf(pa, z, 2.089591050e-04f);
    pa = fmaf(pa, z, -1.446639958e-03f);
    pa = fmaf(pa, z, 8.548327023e-03f);
    pa = fmaf(pa, z, -4.179182100e-02f);
    pa = fmaf(pa, z, 1.611970239e-01f);
    pa = fmaf(pa, z, -4.513516668e-01f);
    pa = fmaf(pa, z, 7.522527781e-01f);
    float pe = -1.492565036e-05f;
    pe = fmaf(pe, z, 1.205533298e-04f);
    pe = fmaf(pe, z, -8.548327023e-04f);
    pe = fmaf(pe, z, 5.223977625e-03f);
    pe = fmaf(pe, z, -2.686617065e-02f);
    pe = fmaf(pe, z, 1.128379167e-01f);
    pe = fmaf(pe, z, -3.761263890e-01f);
    pe = fmaf(pe, z, 1.128379167e+00f);
    const bool series = x < 0.6f;
    const float A = series ? pa * z * x : 1.0f - ec - g;
    const float erfx = series ? pe * x : 1.0f - ec;
    *num = B - A;
    *en = fmaf(s, y, B) - erfx;
    (void)r;
}

constexpr int kMixTile = 32;

template <bool kRdf>
__global__ __launch_bounds__(kPairBlock) void pair_mixed_kernel(PairArgs a)
{
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    if (kRdf) {
        hist_clear(hist);
        __syncthreads();
    }
    const int i = blockIdx.x * kPairBlock + threadIdx.x;
    const bool live = i < a.n_i;
    const int ii = live ? i : a.n_i - 1;
    const FixPos pi = a.ui[ii];
    const int si = ii >= a.ne_i;
    const int j0 = blockIdx.y * a.jchunk;
    const int j1 = min(a.n_j, j0 + a.jchunk);
    const float sx = (float)(a.box[0] * 0x1p-32);
    const float sy = (float)(a.box[1] * 0x1p-32);
    const float sz = (float)(a.box[2] * 0x1p-32);
    const float rc2 = (float)a.rc2;
    const float half_alpha = (float)(0.5 * a.alpha);
    const float na2l = (float)(-a.alpha * a.alpha * 1.4426950408889634);
    const float two_a_sqrtpi = (float)(2.0 * a.alpha / kSqrtPi);
    const float inv_rs = (float)(1.0 / a.rs);
    const float nl2rs = (float)(-1.4426950408889634 / a.rs);
    const float rclose = (float)(2.0 * a.rs);
    double fxd = 0.0, fyd = 0.0, fzd = 0.0, ud = 0.0;
    for (int sj = 0; sj < 2; ++sj) {
        const int lo = sj ? max(j0, a.ne_j) : j0;
        const int hi = sj ? j1 : min(j1, a.ne_j);
        const float qq = (float)(a.q[si] * a.q[sj]);
        const float eqq = (float)(a.eps * a.q[si] * a.q[sj]);
        const bool ei = si != sj;
        const int ht = si + sj;
        const float inv_dx = (float)a.inv_dx[ht];
        for (int jt = lo; jt < hi; jt += kMixTile) {
            const int je = min(hi, jt + kMixTile);
            float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f;
#pragma unroll 4
            for (int j = jt; j < je; ++j) {
                const FixPos pj = a.uj[j];
                const float dx = fmaf((float)(int)(pi.ux - pj.ux), sx, pi.rx - pj.rx);
                const float dy = fmaf((float)(int)(pi.uy - pj.uy), sy, pi.ry - pj.ry);
                const float dz = fmaf((float)(int)(pi.uz - pj.uz), sz, pi.rz - pj.rz);
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const bool in = (r2 < rc2) & (r2 > 0.f);
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                const float e2 = __builtin_amdgcn_exp2f(na2l * r2);
                const float px = half_alpha * r;
                const float t = __builtin_amdgcn_rcpf(px + 1.0f);
                const float ec = e2 * erfcx_poly(px * t);
                float num = fmaf(two_a_sqrtpi * r, e2, ec);
                float en = ec;
                if (ei) {
                    const float y = __builtin_amdgcn_exp2f(nl2rs * r);
                    const float sr = r * inv_rs;
                    num = fmaf(-y, sr + 1.0f, num);
                    en -= y;
                    if (in && r < rclose)
                        close_ei(r, sr, 2.0f * px, y, ec, two_a_sqrtpi * r * e2, &num, &en);
                }
                const float rinv2 = rinv * rinv;
                const float df = in ? eqq * num * rinv2 * rinv : 0.f;
                const float du = in ? qq * en * rinv : 0.f;
                fx = fmaf(df, dx, fx);
                fy = fmaf(df, dy, fy);
                fz = fmaf(df, dz, fz);
                u += du;
                if (kRdf && in && live) {
                    const int b = (int)(r * inv_dx);
                    if (b < kNbin) atomicAdd(&hist[ht * kHistStride + b], 1u);
                }
            }
            fxd += (double)fx;
            fyd += (double)fy;
            fzd += (double)fz;
            ud += (double)u;
        }
    }
    if (live) {
        const size_t n = a.n_i;
        double* pf = a.pf + (size_t)blockIdx.y * 3 * n;
        pf[i] = fxd;
        pf[n + i] = fyd;
        pf[2 * n + i] = fzd;
        a.pu[(size_t)blockIdx.y * n + i] = ud;
    }
    if (kRdf) {
        __syncthreads();
        hist_flush(hist, a.rdf, a.rdf_weight);
    }
}

}  // namespace

void launch_pair_f64(const PairArgs& a, int nchunk, hipStream_t s)
{
    dim3 grid((a.n_i + kPairBlock - 1) / kPairBlock, nchunk);
    if (a.rdf_weight)
        hipLaunchKernelGGL(pair_f64_kernel<true>, grid, dim3(kPairBlock), 0, s, a);
    else
        hipLaunchKernelGGL(pair_f64_kernel<false>, grid, dim3(kPairBlock), 0, s, a);
}

void launch_pair_mixed(const PairArgs& a, int nchunk, hipStream_t s)
{
    dim3 grid((a.n_i + kPairBlock - 1) / kPairBlock, nchunk);
    if (a.rdf_weight)
        hipLaunchKernelGGL(pair_mixed_kernel<true>, grid, dim3(kPairBlock), 0, s, a);
    else
        hipLaunchKernelGGL(pair_mixed_kernel<false>, grid, dim3(kPairBlock), 0, s, a);
}

}  // namespace tcp
