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
// Positions are 32-bit fixed-point fractions of the box (u = (x/L + 1/2) 2^32
// mod 2^32); the wrapped 32-bit difference of two of them is the minimum-image
// separation exactly, so the FP32 arithmetic starts from a separation with no
// cancellation error and no rint.  erfc(x) = exp(-x^2) P(t), t = 1/(1+0.4x),
// with P a degree-8 polynomial fitted here (weighted least squares plus Lawson
// iterations against exp(x^2) erfc(x), relative error 4.6e-9 on [0, 4.5],
// about 3e-7 once evaluated in FP32).  exp(-x^2) is shared with the
// 2x exp(-x^2)/sqrt(pi) term of the force.  Per-lane sums run in FP32 over
// tiles of 32 j and are added to FP64 accumulators after each tile.

constexpr float kLog2e = 1.4426950408889634f;

__device__ inline float erfcx_poly(float t)
{
    float p = 8.698318719e-02f;
    p = fmaf(p, t, -4.607766800e-01f);
    p = fmaf(p, t, 8.327167932e-01f);
    p = fmaf(p, t, -5.598961666e-01f);
    p = fmaf(p, t, 5.512020669e-01f);
    p = fmaf(p, t, 7.153115576e-02f);
    p = fmaf(p, t, 2.562803493e-01f);
    p = fmaf(p, t, 2.217379002e-01f);
    p = fmaf(p, t, 2.213986858e-04f);
    return p;
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
    const uint4 ui = a.ui[ii];
    const int si = ii >= a.ne_i;
    const int j0 = blockIdx.y * a.jchunk;
    const int j1 = min(a.n_j, j0 + a.jchunk);
    const float sx = (float)(a.box[0] * 0x1p-32);
    const float sy = (float)(a.box[1] * 0x1p-32);
    const float sz = (float)(a.box[2] * 0x1p-32);
    const float rc2 = (float)a.rc2;
    const float alpha = (float)a.alpha;
    const float na2l = (float)(-a.alpha * a.alpha * 1.4426950408889634);
    const float two_a_sqrtpi = (float)(2.0 * a.alpha / kSqrtPi);
    const float inv_rs = (float)(1.0 / a.rs);
    const float nl2rs = (float)(-1.4426950408889634 / a.rs);
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
                const uint4 uj = a.uj[j];
                const float dx = (float)(int)(ui.x - uj.x) * sx;
                const float dy = (float)(int)(ui.y - uj.y) * sy;
                const float dz = (float)(int)(ui.z - uj.z) * sz;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const bool in = (r2 < rc2) & (r2 > 0.f);
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                const float x = alpha * r;
                const float e2 = __builtin_amdgcn_exp2f(na2l * r2);
                const float t = __builtin_amdgcn_rcpf(fmaf(0.4f, x, 1.0f));
                const float ec = e2 * erfcx_poly(t);
                float num = fmaf(two_a_sqrtpi * r, e2, ec);
                float en = ec;
                if (ei) {
                    const float y = __builtin_amdgcn_exp2f(nl2rs * r);
                    num = fmaf(-y, fmaf(r, inv_rs, 1.0f), num);
                    en -= y;
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
