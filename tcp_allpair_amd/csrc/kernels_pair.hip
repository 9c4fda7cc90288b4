// All-pair real-space kernels: Ewald erfc for every pair within rc plus the
// electron-ion Deutsch term -exp(-r/lambda)/r (force.f90:87-178, 215-341),
// with the Kelbg form and e-e regularisation as options.
//
// Self interaction (Force_direct_self, every MD step): Newton's-third-law
// kernels, pair_mixed_n3_kernel (FP32 pair arithmetic), pair_mixed_n3_fine_kernel
// (the same on 4 x 4 tile blocks for small systems) and pair_f64_n3_kernel.
// Tiles of 64 particles; a persistent grid takes blocks of tile pairs of the
// upper triangle from a device queue; a wavefront holds its i tiles in
// registers and meets the block's j tiles by rotating the j data and their
// reaction sums one lane per step with DPP.  Each (block row, particle) partial
// is written once into its slot, so the fixed-order sum over slots
// (combine_kernel) does not depend on scheduling.
//
// External block (Force_direct_ext) and the adjoint: ordered-pair kernels with
// one i particle per lane, 256 i per block (blockIdx.x) and the j range split
// into chunks (blockIdx.y); j is uniform across the wavefront (scalar loads),
// and the per-(chunk, i) partials are summed in a fixed order later.
//
// RDF counts go to a per-block LDS histogram (3 x 512 uint32) flushed with one
// global 64-bit atomic per non-empty bin.
#include "f32_tables.h"
#include "f64_tables.h"
#include "tcp_internal.h"

namespace tcp {

namespace {

constexpr double kSqrtPi = 1.7724538509055159;  // data.f90:24
constexpr float kSqrtPiF = 1.7724538509055159f;

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

// One block: i block bx against j chunk by.  hist: the block's LDS histogram.
template <bool kRdf>
__device__ inline void f64_block(const PairArgs& a, int bx, int by, unsigned* hist)
{
    if (kRdf) {
        hist_clear(hist);
        __syncthreads();
    }
    const int i = bx * kPairBlock + threadIdx.x;
    const bool live = i < a.n_i;
    const int ii = live ? i : a.n_i - 1;
    const double xi = a.pi.x[ii], yi = a.pi.y[ii], zi = a.pi.z[ii];
    const int si = ii >= a.ne_i;
    const int j0 = by * a.jchunk;
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
        // regularised pair: electron-ion, and electron-electron with TCP_REG_EE
        const bool reg = si != sj || (si == 0 && sj == 0 && (a.flags & TCP_REG_EE));
        const bool kelbg = (a.flags & TCP_KELBG) != 0;
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
                if (reg) {
                    const double s = r * inv_rs;
                    if (kelbg) {  // R = -e^{-s^2} + sqrt(pi) s erfc(s), r dR/dr = sqrt(pi) s erfc(s)
                        const double y2 = exp(-(s * s));
                        num -= y2;
                        en += kSqrtPi * s * erfc(s) - y2;
                    } else {      // R = -e^{-s} (Deutsch, force.f90:136-139)
                        const double y = exp(-s);
                        num -= y * (1.0 + s);
                        en -= y;
                    }
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
        double* pf = a.pf + (size_t)by * 3 * n;
        pf[i] = fx;
        pf[n + i] = fy;
        pf[2 * n + i] = fz;
        a.pu[(size_t)by * n + i] = u;
    }
    if (kRdf) {
        __syncthreads();
        hist_flush(hist, a.rdf, a.rdf_weight);
    }
}

template <bool kRdf>
__global__ __launch_bounds__(kPairBlock) void pair_f64_kernel(PairArgs a)
{
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    f64_block<kRdf>(a, blockIdx.x, blockIdx.y, hist);
}

// --------------------------------------------------------------- ADJOINT ---
//
// Energy derivatives for fitting charges and lambda: the FP64 ordered-pair
// shape above with per-particle charges.  With R the regularisation and
// h = d(R/r)/dlambda:
//   Deutsch R = -e^{-s}:                          h = -e^{-s}/lambda^2,         dh/dr = e^{-s}/lambda^3
//   Kelbg   R = -e^{-s^2} + sqrt(pi) s erfc(s):   h = -sqrt(pi) erfc(s)/lambda^2, dh/dr = 2 e^{-s^2}/lambda^3
// (s = r/lambda), and dF_i/dlambda = -eps sum_j q_i q_j (dh/dr) r_ij / r.

__global__ __launch_bounds__(kPairBlock) void pair_adjoint_kernel(AdjArgs a)
{
    const int i = blockIdx.x * kPairBlock + threadIdx.x;
    const bool live = i < a.n;
    const int ii = live ? i : a.n - 1;
    const double xi = a.p.x[ii], yi = a.p.y[ii], zi = a.p.z[ii];
    const double qi = a.qp[ii];
    const int si = ii >= a.ne;
    const int j0 = blockIdx.y * a.jchunk;
    const int j1 = min(a.n, j0 + a.jchunk);
    const double alpha = a.alpha, a2 = alpha * alpha, two_a_sqrtpi = 2.0 * alpha / kSqrtPi;
    const double lam = a.rs, inv_l = 1.0 / lam, inv_l2 = inv_l * inv_l, inv_l3 = inv_l2 * inv_l;
    const bool kelbg = (a.flags & TCP_KELBG) != 0;
    double acc[kAdjQ];
#pragma unroll
    for (int k = 0; k < kAdjQ; ++k) acc[k] = 0.0;
    for (int sj = 0; sj < 2; ++sj) {
        const int lo = sj ? max(j0, a.ne) : j0;
        const int hi = sj ? j1 : min(j1, a.ne);
        const bool reg = si != sj || (si == 0 && sj == 0 && (a.flags & TCP_REG_EE));
        for (int j = lo; j < hi; ++j) {
            double xr, yr, zr, r2;
            {
#pragma clang fp contract(off)
                const double dx = xi - a.p.x[j];
                const double dy = yi - a.p.y[j];
                const double dz = zi - a.p.z[j];
                xr = dx - a.box[0] * rint(dx * a.inv_box[0]);
                yr = dy - a.box[1] * rint(dy * a.inv_box[1]);
                zr = dz - a.box[2] * rint(dz * a.inv_box[2]);
                r2 = xr * xr + yr * yr + zr * zr;
            }
            if (!(r2 < a.rc2 && r2 > 0.0)) continue;
            const double qj = a.qp[j];
            const double r = sqrt(r2), inv_r = 1.0 / r;
            const double ec = erfc(alpha * r);
            double num = two_a_sqrtpi * r * exp(-(a2 * r2)) + ec, en = ec, h = 0.0, hp = 0.0;
            if (reg) {
                const double s = r * inv_l;
                if (kelbg) {
                    const double y2 = exp(-(s * s)), es = erfc(s);
                    num -= y2;
                    en += kSqrtPi * s * es - y2;
                    h = -kSqrtPi * es * inv_l2;
                    hp = 2.0 * y2 * inv_l3;
                } else {
                    const double y = exp(-s);
                    num -= y * (1.0 + s);
                    en -= y;
                    h = -y * inv_l2;
                    hp = y * inv_l3;
                }
            }
            const double qq = qi * qj;
            const double df = a.eps * qq * num * inv_r * inv_r * inv_r;
            const double dg = -a.eps * qq * hp * inv_r;
            acc[0] += df * xr;
            acc[1] += df * yr;
            acc[2] += df * zr;
            acc[3] += qq * en * inv_r;
            acc[4] += qj * en * inv_r;
            acc[5] += qq * h;
            acc[6] += dg * xr;
            acc[7] += dg * yr;
            acc[8] += dg * zr;
        }
    }
    if (live) {
        double* out = a.part + (size_t)blockIdx.y * kAdjQ * a.n + i;
#pragma unroll
        for (int k = 0; k < kAdjQ; ++k) out[(size_t)k * a.n] = acc[k];
    }
}

__global__ __launch_bounds__(256) void adjoint_reduce_kernel(const double* part, int nchunk, int n,
                                                             double* out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t m = (size_t)kAdjQ * n;
    if (t >= m) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * m + t];
    out[t] = s;
}

// ----------------------------------------------------------------- MIXED ---
//
// Separations come from FixPos: the wrapped 32-bit difference of the fixed-point
// coordinates is the minimum-image separation exactly (no rint), dx =
// float(int(u_i - u_j)) s; for pairs closer than L 2^-8 the FP32 residuals
// restore the FP64 positions to FP32 rounding, dx = float(int(u_i - u_j)) s +
// (r_i - r_j), so there is no cancellation even for pairs 0.01 A apart in a
// 50 A box.  erfc(x) = exp(-x^2) erfcx(x), erfcx from the LDS table of
// kErfcxTab (erfcx_row), for the Ewald erfc(alpha r) and the Kelbg
// erfc(r/lambda) alike; exp(-x^2) is shared with the 2x exp(-x^2)/sqrt(pi)
// force term.
// Electron-ion pairs closer than 2 lambda take a rare, divergent branch
// (close_ei) where the Deutsch and Ewald parts are evaluated without their
// mutual cancellation.
// Per-lane sums run in FP32 over tiles of 32 j and go to FP64 after each tile.

static_assert(kErfcxRows > kErfcxPerUnit * kMixedErfcMax + 1,
              "the erfcx rows cover alpha r < alpha rc <= kMixedErfcMax with rounding to spare");

// The LDS copy of kErfcxTab of a mixed block (1.2 KB)
__device__ inline void load_erfcx(float4* xtab)
{
    for (int i = threadIdx.x; i < kErfcxRows; i += blockDim.x)
        xtab[i] = make_float4(kErfcxTab[i][0], kErfcxTab[i][1], kErfcxTab[i][2], kErfcxTab[i][3]);
}

// erfcx(x) = e^{x^2} erfc(x) from the LDS copy of kErfcxTab, given
// u = kErfcxPerUnit x >= 0 below kErfcxRows: row int(u) (the conversion is
// nearly free beside FMAs), t = fract(u), a cubic by three FMAs; 1.2e-7
// relative on [0, 4.5].  A NaN u (coincident particles) converts to row 0.
__device__ inline float erfcx_row(const float4* xtab, float u)
{
    const float4 c = xtab[(int)u];
    const float t = __builtin_amdgcn_fractf(u);
    return fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
}

// The row coordinate of the Kelbg erfc(s), s = r/lambda, which can run far
// past the table: clamped inside the last row.  Beyond s = 4.6 the factor
// e^{-s^2} it is taken with is below 6e-10, so the clamped row's error there
// is far below FP32 rounding of the pair's terms; a NaN s (coincident
// particles) gives the last row too.
__device__ inline float erfcx_u_clamped(float s)
{
    return fminf(s * (float)kErfcxPerUnit, (float)kErfcxRows - 1e-3f);
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

// Close pair with the Kelbg form (s = r/lambda < 1): num = B(s^2) - A(x) with
// B(z) = 1 - e^-z = z (1 - z/2 + z^2/6 - ...) (Taylor to z^9, relative error
// < 3e-8 for z < 1) and A as in close_ei; en = B - erf(x) + sqrt(pi) s erfc(s).
__device__ inline void close_kelbg(float s, float x, float erfc_s, float g, float ec, float* num,
                                   float* en)
{
    const float zs = s * s;
    float b = -2.755731922e-07f;                  // -1/10!
    b = fmaf(b, zs, 2.755731922e-06f);            // 1/9!
    b = fmaf(b, zs, -2.480158730e-05f);
    b = fmaf(b, zs, 1.984126984e-04f);
    b = fmaf(b, zs, -1.388888889e-03f);
    b = fmaf(b, zs, 8.333333333e-03f);
    b = fmaf(b, zs, -4.166666667e-02f);
    b = fmaf(b, zs, 1.666666667e-01f);
    b = fmaf(b, zs, -5.000000000e-01f);
    b = fmaf(b, zs, 1.0f);
    const float B = b * zs;
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
    *en = fmaf(kSqrtPiF * s, erfc_s, B - erfx);
}

constexpr int kMixTile = 32;

// Loop-invariant FP32 constants of the mixed pair arithmetic.
// One of three values by an index, as a select of values: the empty asm makes
// them values first, so the compiler cannot turn the select into a load from a
// selected address (which would keep the struct holding them in scratch).
__device__ inline float pick3(float a, float b, float c, int i)
{
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c));
    return i == 0 ? a : (i == 1 ? b : c);
}

struct MixConst {
    float sx, sy, sz;          // L_d 2^-32
    float rc2;
    float half_alpha;          // alpha/2
    float xrow;                // kErfcxPerUnit alpha: erfcx table row coordinate of r
    float na2l;                // -alpha^2 log2(e)
    float two_a_sqrtpi;        // 2 alpha/sqrt(pi)
    float inv_rs;              // 1/lambda
    float nl2rs;               // -log2(e)/lambda
    float rclose;              // 2 lambda (Deutsch) or lambda (Kelbg)
    float idx_ee, idx_ei, idx_ii;  // RDF bin widths (reciprocal) of e-e, e-i, i-i pairs
    float rfix2;               // below this r^2 the FP32 residuals join the separation
    bool kelbg;                // Kelbg instead of Deutsch regularisation
    bool rdf_clamp;            // a pair inside rc can land beyond the padded histogram row
};

__device__ inline MixConst mix_const(const PairArgs& a)
{
    MixConst k;
    k.sx = (float)(a.box[0] * 0x1p-32);
    k.sy = (float)(a.box[1] * 0x1p-32);
    k.sz = (float)(a.box[2] * 0x1p-32);
    k.rc2 = (float)a.rc2;
    k.half_alpha = (float)(0.5 * a.alpha);
    k.xrow = (float)(kErfcxPerUnit * a.alpha);
    k.na2l = (float)(-a.alpha * a.alpha * 1.4426950408889634);
    k.two_a_sqrtpi = (float)(2.0 * a.alpha / kSqrtPi);
    k.inv_rs = (float)(1.0 / a.rs);
    k.nl2rs = (float)(-1.4426950408889634 / a.rs);
    k.kelbg = (a.flags & TCP_KELBG) != 0;
    k.rclose = (float)((k.kelbg ? 1.0 : 2.0) * a.rs);
    k.idx_ee = (float)a.inv_dx[0];
    k.idx_ei = (float)a.inv_dx[1];
    k.idx_ii = (float)a.inv_dx[2];
    // the fixed-point separation alone is off by at most L 2^-33 per axis; beyond
    // r = L 2^-8 that is below FP32 rounding of the separation (2^-24 r)
    const double lmax = fmax(fmax(a.box[0], a.box[1]), a.box[2]);
    k.rfix2 = (float)(lmax * lmax * 0x1p-16);
    const double idx_max = fmax(fmax(a.inv_dx[0], a.inv_dx[1]), a.inv_dx[2]);
    k.rdf_clamp = sqrt(a.rc2) * idx_max >= (double)(kHistStride - 2);
    return k;
}

// One pair in FP32: separation d, scalar force df (eps q q included, 0 when
// the pair is masked out) and energy term du = q q (erfc - y)/r.
struct MixPair {
    float dx, dy, dz, r, df, du;
    bool in;
};

// The charge-free part of one pair: separation, r, 1/r and the numerator and
// energy factors num, en with df = eps qi qj num/r^3 and du = qi qj en/r.
struct MixRaw {
    float dx, dy, dz, r, rinv, num, en;
    bool in;
};

// kCheck: the pair may be a particle with itself (r = 0) or involve a padding
// lane (live false); without it only the cut-off masks the pair.
template <bool kKelbg, bool kCheck>
__device__ inline MixRaw mix_raw(const MixConst& k, const float4* xtab, const FixPos& pi,
                                 const FixPos& pj, bool reg, bool live)
{
    MixRaw o;
    o.dx = (float)(int)(pi.ux - pj.ux) * k.sx;
    o.dy = (float)(int)(pi.uy - pj.uy) * k.sy;
    o.dz = (float)(int)(pi.uz - pj.uz) * k.sz;
    float r2 = fmaf(o.dx, o.dx, fmaf(o.dy, o.dy, o.dz * o.dz));
    if (r2 < k.rfix2) {  // close pair (rare): add the FP32 residuals
        o.dx = fmaf((float)(int)(pi.ux - pj.ux), k.sx, pi.rx - pj.rx);
        o.dy = fmaf((float)(int)(pi.uy - pj.uy), k.sy, pi.ry - pj.ry);
        o.dz = fmaf((float)(int)(pi.uz - pj.uz), k.sz, pi.rz - pj.rz);
        r2 = fmaf(o.dx, o.dx, fmaf(o.dy, o.dy, o.dz * o.dz));
    }
    o.in = kCheck ? (r2 < k.rc2) & (r2 > 0.f) & live : r2 < k.rc2;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float e2 = __builtin_amdgcn_exp2f(k.na2l * r2);
    // the table row from r inside rc (alpha r < alpha rc <= 4.5), row 0 outside
    const float ec = e2 * erfcx_row(xtab, (o.in ? r : 0.f) * k.xrow);
    float num = fmaf(k.two_a_sqrtpi * r, e2, ec);
    float en = ec;
    if (reg) {
        const float sr = r * k.inv_rs;
        if (kKelbg) {
            const float y2 = __builtin_amdgcn_exp2f(-1.4426950408889634f * sr * sr);
            // erfc(s) = exp(-s^2) P(s/(2 + s)), s/(2 + s) = 1/(1 + 2 lambda/r)
            const float ecs = y2 * erfcx_row(xtab, erfcx_u_clamped(sr));
            num -= y2;
            en = fmaf(kSqrtPiF * sr, ecs, en - y2);
            if (o.in && r < k.rclose)
                close_kelbg(sr, 2.0f * k.half_alpha * r, ecs, k.two_a_sqrtpi * r * e2, ec, &num, &en);
        } else {
            const float y = __builtin_amdgcn_exp2f(k.nl2rs * r);
            num = fmaf(-y, sr + 1.0f, num);
            en -= y;
            if (o.in && r < k.rclose)
                close_ei(r, sr, 2.0f * k.half_alpha * r, y, ec, k.two_a_sqrtpi * r * e2, &num, &en);
        }
    }
    o.r = r;
    o.rinv = rinv;
    o.num = num;
    o.en = en;
    return o;
}

template <bool kKelbg, bool kCheck = true>
__device__ inline MixPair mix_pair(const MixConst& k, const float4* xtab, const FixPos& pi,
                                   const FixPos& pj, float qq, float eqq, bool reg, bool live)
{
    const MixRaw m = mix_raw<kKelbg, kCheck>(k, xtab, pi, pj, reg, live);
    MixPair o;
    o.dx = m.dx;
    o.dy = m.dy;
    o.dz = m.dz;
    o.in = m.in;
    o.r = m.r;
    const float rinv2 = m.rinv * m.rinv;
    o.df = m.in ? eqq * m.num * rinv2 * m.rinv : 0.f;
    o.du = m.in ? qq * m.en * m.rinv : 0.f;
    return o;
}

template <bool kRdf, bool kKelbg>
__global__ __launch_bounds__(kPairBlock) void pair_mixed_kernel(PairArgs a)
{
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    __shared__ float4 xtab[kErfcxRows];
    if (kRdf) hist_clear(hist);
    load_erfcx(xtab);
    __syncthreads();
    const MixConst k = mix_const(a);
    const int i = blockIdx.x * kPairBlock + threadIdx.x;
    const bool live = i < a.n_i;
    const int ii = live ? i : a.n_i - 1;
    const FixPos pi = a.ui[ii];
    const int si = ii >= a.ne_i;
    const int j0 = blockIdx.y * a.jchunk;
    const int j1 = min(a.n_j, j0 + a.jchunk);
    double fxd = 0.0, fyd = 0.0, fzd = 0.0, ud = 0.0;
    for (int sj = 0; sj < 2; ++sj) {
        const int lo = sj ? max(j0, a.ne_j) : j0;
        const int hi = sj ? j1 : min(j1, a.ne_j);
        const float qq = (float)(a.q[si] * a.q[sj]);
        const float eqq = (float)(a.eps * a.q[si] * a.q[sj]);
        const bool reg = si != sj || (si == 0 && sj == 0 && (a.flags & TCP_REG_EE));
        const int ht = si + sj;
        const float idx = pick3(k.idx_ee, k.idx_ei, k.idx_ii, ht);
        for (int jt = lo; jt < hi; jt += kMixTile) {
            const int je = min(hi, jt + kMixTile);
            float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f;
#pragma unroll 4
            for (int j = jt; j < je; ++j) {
                const MixPair o = mix_pair<kKelbg>(k, xtab, pi, a.uj[j], qq, eqq, reg, live);
                fx = fmaf(o.df, o.dx, fx);
                fy = fmaf(o.df, o.dy, fy);
                fz = fmaf(o.df, o.dz, fz);
                u += o.du;
                if (kRdf && o.in) {
                    const int b = (int)(o.r * idx);
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

// ------------------------------------------- MIXED, Newton's third law ---
//
// Self interaction of one particle set, each unordered pair once.  Particles
// form tiles of 64 (one wavefront); a block owns a kBlkTiles x kBlkTiles square
// of tile pairs (I, J) of the upper triangle (blockmap) with kN3Waves
// wavefronts, each taking two I tiles (8 x 8 by 4 waves: 32896 blocks at
// N = 131072, a short tail on 256 CUs).  For a tile pair the wavefront holds the 64 i in registers and
// the 64 j (FixPos, species and live flags) plus their reaction sums in
// registers as well, and rotates the j data and reactions one lane per step
// with DPP (wave_ror:1): after 64 steps every i has met every j once and each
// reaction sum is back in the lane of its j.  i forces add over a J tile in
// FP32 and across tiles in FP64; reactions go to an LDS FP64 accumulator per
// J particle.  Diagonal tile pairs (I == J) evaluate every ordered pair for the
// i side only, with half the energy and RDF weight 1 instead of 2.
// Partials: particle p of block-row b receives one (f, u) row per block-row
// slot s: from block (b, s) as an i when s > b, from block (s, b) as a j when
// s < b, and from block (b, b) for s == b, so the sum over slots in a fixed
// order is the full force and energy.

constexpr int kTile = 64;
constexpr int kBlkTiles = 8;
constexpr int kN3Waves = 4;

// wave_ror:1 -- every lane takes the value of the lane below it (lane 0 that of lane 63)
__device__ inline uint32_t rot_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x13C, 0xF, 0xF, false);
}

__device__ inline float rot_f(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x13C, 0xF, 0xF, false));
}

// One tile pair (64 i x 64 j) by 64 rotation steps.  kUni: both tiles hold
// one species and only live particles (all but the tiles at the electron/ion
// boundary and at the end), so species, charges and liveness are
// wave-uniform and the flags do not travel with j.
template <bool kRdf, bool kKelbg, bool kDiag, bool kUni>
__device__ inline void n3_tile(const MixConst& k, const PairArgs& a, unsigned* hist,
                               const float4* xtab, const FixPos& pi, int si, bool ilive,
                               int sj_uni, FixPos pj,
                               float& fxo, float& fyo, float& fzo, float& uo,
                               float& rx, float& ry, float& rz)
{
    const float qi = (float)a.q[si];
    const float eqi = (float)(a.eps * a.q[si]);
    // RDF bin widths for a j electron / a j ion, seen from this i
    const float idx0 = pick3(k.idx_ee, k.idx_ei, k.idx_ii, si);
    const float idx1 = pick3(k.idx_ee, k.idx_ei, k.idx_ii, si + 1);
    const float qj_uni = (float)a.q[sj_uni];
    const bool regee = (a.flags & TCP_REG_EE) != 0;
    const bool reg_uni = si != sj_uni || (regee && si == 0 && sj_uni == 0);
    float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f;
    rx = ry = rz = 0.f;
    for (int s = 0; s < kTile; ++s) {
        int sj;
        float qj;
        bool live, reg;
        if (kUni) {
            sj = sj_uni;
            qj = qj_uni;
            live = true;
            reg = reg_uni;
        } else {
            sj = (int)(pj.pad0 & 1u);
            qj = pj.pad1;  // the charge of j rides along with its coordinates
            live = ilive & ((pj.pad0 & 2u) != 0u);
            reg = si != sj || (regee && si == 0 && sj == 0);
        }
        const MixPair o = mix_pair<kKelbg>(k, xtab, pi, pj, qi * qj, eqi * qj, reg, live);
        const float gx = o.df * o.dx, gy = o.df * o.dy, gz = o.df * o.dz;
        fx += gx;
        fy += gy;
        fz += gz;
        if (kDiag) {
            u = fmaf(0.5f, o.du, u);
        } else {
            u += o.du;
            rx -= gx;
            ry -= gy;
            rz -= gz;
        }
        if (kRdf && o.in) {
            // bins past Ndist+1 land in the padding bins 501..511, never read
            const int b = min((int)(o.r * (sj ? idx1 : idx0)), kHistStride - 1);
            atomicAdd(&hist[(si + sj) * kHistStride + b], kDiag ? 1u : 2u);
        }
        pj.ux = rot_u(pj.ux);
        pj.uy = rot_u(pj.uy);
        pj.uz = rot_u(pj.uz);
        pj.rx = rot_f(pj.rx);
        pj.ry = rot_f(pj.ry);
        pj.rz = rot_f(pj.rz);
        if (!kUni) {
            pj.pad0 = rot_u(pj.pad0);
            pj.pad1 = rot_f(pj.pad1);
        }
        if (!kDiag) {
            rx = rot_f(rx);
            ry = rot_f(ry);
            rz = rot_f(rz);
        }
    }
    fxo = fx;
    fyo = fy;
    fzo = fz;
    uo = u;
}

// kNI i tiles against one j tile, off the diagonal, every tile single-species and
// fully live: each rotation step of the j data and reactions serves kNI pairs
// per lane, so the nine DPP moves per step are shared by kNI pair evaluations.
// The charges are uniform over the tile pair, so the sums run over the
// charge-free terms num/r^3 and en/r and take eps qi qj (forces, reactions)
// and qi qj (energy) once at the end.
// Constants of the uniform tile path in its working length unit u.  For a
// cubic box u = L 2^-32: the separation is the fixed-point difference itself
// converted to float, with no per-axis scaling, and every constant that meets
// a length carries the matching power of u (the FP32 sums of a tile pair take
// eps qi qj u^-2 and qi qj u^-1 in FP64 as they join the FP64 accumulators,
// so no FP32 rounding of those factors is common to all pairs).  Otherwise
// u = 1 and the
// separation is scaled per axis as in mix_raw.
struct MixUnits {
    float sx, sy, sz;        // fixed-point difference to u per axis (u = 1 only)
    float inv_u;             // FP32 residuals (physical) to u
    float rc2, rfix2;        // u^2
    float rnear2;            // u^2: max(rfix2, rclose^2)
    float half_alpha, two_a_sqrtpi, inv_rs, nl2rs;  // x u
    float na2l;              // x u^2
    float rclose;            // / u
    float idx_ee, idx_ei, idx_ii;  // RDF bin widths (reciprocal) x u
    float xrow;              // kErfcxPerUnit alpha: x u (erfcx table row coordinate of r)
    double f_scale, e_scale; // u^-2 (num d / r^3), u^-1 (en / r)
};

__device__ inline MixUnits mix_units(const PairArgs& a, bool cubic)
{
    const double u = cubic ? a.box[0] * 0x1p-32 : 1.0;
    MixUnits w;
    w.sx = (float)(a.box[0] * 0x1p-32);
    w.sy = (float)(a.box[1] * 0x1p-32);
    w.sz = (float)(a.box[2] * 0x1p-32);
    w.inv_u = (float)(1.0 / u);
    w.rc2 = (float)(a.rc2 / (u * u));
    const double lmax = fmax(fmax(a.box[0], a.box[1]), a.box[2]);
    w.rfix2 = (float)(lmax * lmax * 0x1p-16 / (u * u));
    w.half_alpha = (float)(0.5 * a.alpha * u);
    w.two_a_sqrtpi = (float)(2.0 * a.alpha / kSqrtPi * u);
    w.inv_rs = (float)(u / a.rs);
    w.nl2rs = (float)(-1.4426950408889634 * u / a.rs);
    w.na2l = (float)(-a.alpha * a.alpha * 1.4426950408889634 * u * u);
    const bool kelbg = (a.flags & TCP_KELBG) != 0;
    w.rclose = (float)((kelbg ? 1.0 : 2.0) * a.rs / u);
    w.rnear2 = fmaxf(w.rfix2, w.rclose * w.rclose);
    w.idx_ee = (float)(a.inv_dx[0] * u);
    w.idx_ei = (float)(a.inv_dx[1] * u);
    w.idx_ii = (float)(a.inv_dx[2] * u);
    w.xrow = (float)(kErfcxPerUnit * a.alpha * u);
    w.f_scale = 1.0 / (u * u);
    w.e_scale = 1.0 / u;
    return w;
}

// a 32-bit word of LDS addressed by its byte offset
typedef __attribute__((address_space(3))) uint32_t LdsU32;

// The three fixed-point coordinates of a j (12 bytes of its FixPos): one
// ds_read_b96, no register for the unused fourth word.
struct __align__(16) FixU3 {
    uint32_t x, y, z;
};

// One pair of the uniform tile path in the unit u of MixUnits: j's
// fixed-point coordinates in qu, its FP32 residuals (read only for a pair
// closer than L 2^-8) in q.  Returns the separation, r, the cut-off flag, and
// the masked charge-free sums' terms er = en/r and fs = num/r^3 (zero beyond
// rc).  With num = en + (num - en), num/r = en/r + k e2 + (regularisation
// term)/r, where k e2 = (2 alpha r/sqrt(pi)) e^{-alpha^2 r^2}/r needs no r:
// the Deutsch term -e^{-r/lambda} r/lambda and the Kelbg term
// -sqrt(pi) (r/lambda) erfc(r/lambda) give -y/lambda and -sqrt(pi) erfc/lambda.
// Close electron-ion pairs take num and en from the cancellation-free series.
struct MixUni {
    float dx, dy, dz, r, er, fs;
    bool in;
};

// The terms of one pair from its separation o.dx, o.dy, o.dz and r2 (unit u):
// o.r, o.in, o.er and o.fs.  kClose: a close electron-ion pair (r < rclose)
// takes num and en from the cancellation-free series.  r is taken as r2 times
// the masked 1/r, so it is 0 beyond rc, where every term is masked anyway, and
// inside rc alpha r < alpha rc <= 4.5 keeps the erfcx row in the table.
template <bool kKelbg, bool kReg, bool kClose>
__device__ inline void mix_terms(const MixUnits& w, const float4* xtab, float r2, MixUni& o)
{
    o.in = r2 < w.rc2;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rm = o.in ? rinv : 0.f;  // 1/r inside rc, else 0
    const float r = r2 * rm;             // r inside rc, else 0
    const float t = rm * rinv;           // 1/r^2 inside rc
    const float e2 = __builtin_amdgcn_exp2f(w.na2l * r2);
    const float ec = e2 * erfcx_row(xtab, r * w.xrow);
    float er, inner;  // en/r and num/r (inside rc)
    if (kReg) {
        const float sr = r * w.inv_rs;
        if (kKelbg) {
            const float y2 = __builtin_amdgcn_exp2f(-1.4426950408889634f * sr * sr);
            // erfc(s) = exp(-s^2) P(s/(2 + s)), s/(2 + s) = 1/(1 + 2 lambda/r)
            const float ecs = y2 * erfcx_row(xtab, erfcx_u_clamped(sr));
            er = fmaf(kSqrtPiF * sr, ecs, ec - y2) * rm;
            inner = fmaf(-kSqrtPiF * w.inv_rs, ecs, fmaf(w.two_a_sqrtpi, e2, er));
            if (kClose && o.in && r < w.rclose) {
                float num, en;
                close_kelbg(sr, 2.0f * w.half_alpha * r, ecs, w.two_a_sqrtpi * r * e2, ec, &num, &en);
                er = en * rinv;
                inner = num * rinv;
            }
        } else {
            const float y = __builtin_amdgcn_exp2f(w.nl2rs * r);
            er = (ec - y) * rm;
            inner = fmaf(-w.inv_rs, y, fmaf(w.two_a_sqrtpi, e2, er));
            if (kClose && o.in && r < w.rclose) {
                float num, en;
                close_ei(r, sr, 2.0f * w.half_alpha * r, y, ec, w.two_a_sqrtpi * r * e2, &num, &en);
                er = en * rinv;
                inner = num * rinv;
            }
        }
    } else {
        er = ec * rm;
        inner = fmaf(w.two_a_sqrtpi, e2, er);
    }
    o.r = r;
    o.er = er;
    o.fs = t * inner;
}

template <bool kKelbg, bool kCubic, bool kReg>
__device__ inline MixUni mix_uni(const MixUnits& w, const float4* xtab, const FixPos& pi, FixU3 qu,
                                 const FixPos& q)
{
    MixUni o;
    const float ix = (float)(int)(pi.ux - qu.x), iy = (float)(int)(pi.uy - qu.y),
                iz = (float)(int)(pi.uz - qu.z);
    o.dx = kCubic ? ix : ix * w.sx;
    o.dy = kCubic ? iy : iy * w.sy;
    o.dz = kCubic ? iz : iz * w.sz;
    const float r2 = fmaf(o.dx, o.dx, fmaf(o.dy, o.dy, o.dz * o.dz));
    mix_terms<kKelbg, kReg, false>(w, xtab, r2, o);
    // A close pair (rare) is evaluated again: with the FP32 residuals below
    // L 2^-8, where the 32-bit quantisation would exceed FP32 rounding, and,
    // for a regularised pair below rclose, with the series.  One test covers
    // both, so the common path carries a single compare for them.
    if (r2 < (kReg ? w.rnear2 : w.rfix2)) {
        o.dx = kCubic ? fmaf(pi.rx - q.rx, w.inv_u, ix) : fmaf(ix, w.sx, pi.rx - q.rx);
        o.dy = kCubic ? fmaf(pi.ry - q.ry, w.inv_u, iy) : fmaf(iy, w.sy, pi.ry - q.ry);
        o.dz = kCubic ? fmaf(pi.rz - q.rz, w.inv_u, iz) : fmaf(iz, w.sz, pi.rz - q.rz);
        mix_terms<kKelbg, kReg, kReg>(w, xtab, fmaf(o.dx, o.dx, fmaf(o.dy, o.dy, o.dz * o.dz)), o);
    }
    return o;
}

// kNI i tiles against one j tile, all live and of one species each: the
// charge-free sums of num d / r^3 (i side, and the reactions rx, ry, rz) and
// en / r in the unit of w; the caller applies the charges and powers of u.
template <bool kRdf, bool kKelbg, int kNI, bool kClamp, bool kCubic, bool kReg>
__device__ inline void n3_tile_multi(const MixUnits& w, const PairArgs& a, unsigned* hist,
                                     const float4* xtab, const FixPos (&pi)[kNI], int si, int sj,
                                     const FixPos* jpos,
                                     int tj, float (&fxo)[kNI], float (&fyo)[kNI],
                                     float (&fzo)[kNI], float (&uo)[kNI], float& rx, float& ry,
                                     float& rz)
{
    // The j of a lane comes from the block's LDS copy: lane l meets at step s
    // the j of lane l - s of the tile, whose byte offset lane l - 1 held at
    // step s - 1, so the offset rotates with the reactions (one DPP move) and
    // the next step's coordinates are read while this step's pairs compute.
    const char* jb = (const char*)jpos;
    uint32_t off = (uint32_t)(tj * kTile + (int)(threadIdx.x % kTile)) * (uint32_t)sizeof(FixPos);
    FixU3 qu = *(const FixU3*)(jb + off);
    const int ht = si + sj;
    const float idx = pick3(w.idx_ee, w.idx_ei, w.idx_ii, ht);
    unsigned* hrow = hist + (si + sj) * kHistStride;
    // LDS byte address of bin b of this row is hbias + 4 (2^23 + b) (mod 2^32)
    const uint32_t hbias = (uint32_t)(uintptr_t)(LdsU32*)hrow - 4u * 0x4B000000u;
    float fx[kNI], fy[kNI], fz[kNI], u[kNI];
#pragma unroll
    for (int n = 0; n < kNI; ++n) fx[n] = fy[n] = fz[n] = u[n] = 0.f;
    rx = ry = rz = 0.f;
#pragma unroll 2
    for (int s = 0; s < kTile; ++s) {  // by two: the carried offset and coordinates alternate registers
        const uint32_t noff = rot_u(off);
        const FixU3 nqu = *(const FixU3*)(jb + noff);
        const FixPos& q = *(const FixPos*)(jb + off);
#pragma unroll
        for (int n = 0; n < kNI; ++n) {
            const MixUni m = mix_uni<kKelbg, kCubic, kReg>(w, xtab, pi[n], qu, q);
            const float g = m.fs;
            fx[n] = fmaf(g, m.dx, fx[n]);
            fy[n] = fmaf(g, m.dy, fy[n]);
            fz[n] = fmaf(g, m.dz, fz[n]);
            u[n] += m.er;
            rx = fmaf(-g, m.dx, rx);
            ry = fmaf(-g, m.dy, ry);
            rz = fmaf(-g, m.dz, rz);
            // bins past Ndist+1 land in the padding bins 501..511, never read;
            // without kClamp every pair inside rc has its bin below 511.  The
            // bin floor(r idx) comes out of the mantissa: r idx + (2^23 - 1/2),
            // rounded once by the fma, is 2^23 + floor(r idx) unless r idx is an
            // integer (then possibly one below)
            if (kRdf && m.in) {
                const uint32_t t = __float_as_uint(fmaf(m.r, idx, 8388607.5f));
                const uint32_t b = kClamp ? min(t, 0x4B000000u + (kHistStride - 1)) : t;
                __hip_atomic_fetch_add((LdsU32*)(uintptr_t)(hbias + 4u * b), 2u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        off = noff;
        qu = nqu;
        rx = rot_f(rx);
        ry = rot_f(ry);
        rz = rot_f(rz);
    }
#pragma unroll
    for (int n = 0; n < kNI; ++n) {
        fxo[n] = fx[n];
        fyo[n] = fy[n];
        fzo[n] = fz[n];
        uo[n] = u[n];
    }
}

// Reaction staging of the blocks with one i tile per wave on 4 x 4 tiles:
// each wave's FP32 reaction sums for every j tile and the FP64 factor they
// take (eps qi qj u^-2 for a uniform tile pair, 1 otherwise), 12 KB of LDS.
template <int kBT, int kW>
struct N3Stage {
    static constexpr bool kOn = kBT == kW && kBT == 4;
    float r[kOn ? kW * kBT * 3 * kTile : 1];  // [wave][j tile][axis][lane]
    double scale[kOn ? kW * kBT : 1];         // [wave][j tile]
};

// One block pair (row br, column bc) of kBT x kBT tile pairs by kW wavefronts.
// jpos, racc, hist, stg: the block's LDS (kBT * kTile particles).  The sums
// of a block do not depend on the order in which its waves run, so the
// kernel's result is the same from run to run:
//  - with staging (N3Stage::kOn) each wave writes its reactions for every j
//    tile to its own slots, and the slots are summed in wave order at the end;
//  - otherwise the m j tiles of the block column are visited in kW turns of
//    s = ceil(m / kW) steps, wave w taking the j tile (t + s w) mod s kW at
//    step t, with a barrier between turns: within a turn two steps' tiles
//    differ by less than s and two waves' tiles by a nonzero multiple of s,
//    so each j tile's reactions (racc) take at most one wave's add per turn,
//    in an order the turns fix.  Every wave still meets every j tile once.
//  The diagonal block's i sums join racc after its last step.
template <bool kRdf, bool kKelbg, int kBT, int kW>
__device__ inline void n3_block(const PairArgs& a, int br, int bc, FixPos* jpos,
                                double (*racc)[kBT * kTile], unsigned* hist, const float4* xtab,
                                N3Stage<kBT, kW>& stg)
{
    constexpr int kBP = kBT * kTile;  // particles per block side
    static_assert(kBT % kW == 0, "the waves share the i tiles of a block equally");
    const bool dblk = br == bc;
    const int n = a.n_i, ne = a.ne_i;
    for (int p = threadIdx.x; p < kBP; p += blockDim.x) {
        const int jp = bc * kBP + p;
        FixPos f = a.ui[min(jp, n - 1)];
        f.pad0 = (jp >= ne ? 1u : 0u) | (jp < n ? 2u : 0u);
        f.pad1 = (float)a.q[jp >= ne];
        jpos[p] = f;
        racc[0][p] = racc[1][p] = racc[2][p] = racc[3][p] = 0.0;
    }
    if (kRdf) hist_clear(hist);
    __syncthreads();
    const MixConst k = mix_const(a);
    const bool cubic = a.box[0] == a.box[1] && a.box[1] == a.box[2];
    const MixUnits w = mix_units(a, cubic);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile), lane = threadIdx.x % kTile;
    const int ntiles = (n + kTile - 1) / kTile;
    // a tile holds one species and only live particles
    auto uniform_tile = [&](int T) {
        return (T * kTile >= ne || T * kTile + kTile - 1 < ne) && T * kTile + kTile - 1 < n;
    };
    constexpr int kNI = kBT / kW;  // i tiles per wave
    constexpr bool kStage = N3Stage<kBT, kW>::kOn;
    const int m = min(kBT, ntiles - bc * kBT);  // j tiles of this block column
    const int nturn = kStage ? 1 : kW;
    const int s = kStage ? kBT : (m + kW - 1) / kW;  // steps per turn
    const int L = s * nturn;                          // steps: m <= L <= kBT
    // j tile of this wave at step t (tiles past m are skipped)
    auto jtile = [&](int t) {
        const int x = kStage ? t : t + s * wave;
        return x >= L ? x - L : x;
    };
    if (kStage) {  // slots of tiles this wave does not meet stay zero
        for (int c = 0; c < kBT * 3; ++c) stg.r[(wave * kBT * 3 + c) * kTile + lane] = 0.f;
        if (lane < kBT) stg.scale[wave * kBT + lane] = 0.0;
    }
    // reactions of this wave's tile pair with j tile tj, times sc
    auto react = [&](int tj, float rx, float ry, float rz, double sc) {
        if (kStage) {
            float* r = stg.r + (wave * kBT + tj) * 3 * kTile + lane;
            r[0] = rx;
            r[kTile] = ry;
            r[2 * kTile] = rz;
            if (lane == 0) stg.scale[wave * kBT + tj] = sc;
        } else {
            const int q = tj * kTile + lane;
            atomicAdd(&racc[0][q], sc * (double)rx);
            atomicAdd(&racc[1][q], sc * (double)ry);
            atomicAdd(&racc[2][q], sc * (double)rz);
        }
    };
    if (!dblk) {
        // off the diagonal: all kNI i tiles of the wave walk the j tiles together
        FixPos pi[kNI];
        int ip[kNI];
        bool exists[kNI];
        double Fx[kNI], Fy[kNI], Fz[kNI], U[kNI];
        bool all_uni = true;
        const int s0 = (br * kBT + wave) * kTile >= ne;
#pragma unroll
        for (int h = 0; h < kNI; ++h) {
            const int I = br * kBT + wave + h * kW;
            exists[h] = I < ntiles;
            ip[h] = I * kTile + lane;
            pi[h] = a.ui[min(ip[h], n - 1)];
            all_uni = all_uni && exists[h] && uniform_tile(I) && ((I * kTile >= ne) == (s0 != 0));
            Fx[h] = Fy[h] = Fz[h] = U[h] = 0.0;
        }
        for (int t = 0; t < L; ++t) {
            if (t && t % s == 0) __syncthreads();  // the last turn's reaction adds are done
            const int tj = jtile(t);
            if (tj >= m) continue;
            const int J = bc * kBT + tj;
            const FixPos pj = jpos[tj * kTile + lane];
            const int sj = J * kTile >= ne;
            const bool juni = uniform_tile(J);
            if (all_uni && juni) {
                float fx[kNI], fy[kNI], fz[kNI], u[kNI], rx, ry, rz;
                // the tiles are single-species: regularised or not for the whole tile pair
                const bool reg = s0 != sj || ((a.flags & TCP_REG_EE) != 0 && s0 == 0 && sj == 0);
#define TCP_N3_MULTI(CLAMP, CUBIC, REG)                                                          \
    n3_tile_multi<kRdf, kKelbg, kNI, CLAMP, CUBIC, REG>(w, a, hist, xtab, pi, s0, sj, jpos, tj, fx, \
                                                        fy, fz, u, rx, ry, rz)
                if (cubic) {
                    if (k.rdf_clamp) {
                        if (reg) TCP_N3_MULTI(true, true, true);
                        else TCP_N3_MULTI(true, true, false);
                    } else {
                        if (reg) TCP_N3_MULTI(false, true, true);
                        else TCP_N3_MULTI(false, true, false);
                    }
                } else {
                    if (k.rdf_clamp) {
                        if (reg) TCP_N3_MULTI(true, false, true);
                        else TCP_N3_MULTI(true, false, false);
                    } else {
                        if (reg) TCP_N3_MULTI(false, false, true);
                        else TCP_N3_MULTI(false, false, false);
                    }
                }
#undef TCP_N3_MULTI
                // the charge-free FP32 sums take eps qi qj u^-2 and qi qj u^-1 in FP64
                const double qq = a.q[s0] * a.q[sj];
                const double sf = a.eps * qq * w.f_scale, su = qq * w.e_scale;
#pragma unroll
                for (int h = 0; h < kNI; ++h) {
                    Fx[h] += sf * (double)fx[h];
                    Fy[h] += sf * (double)fy[h];
                    Fz[h] += sf * (double)fz[h];
                    U[h] += su * (double)u[h];
                }
                react(tj, rx, ry, rz, sf);
            } else {
#pragma unroll
                for (int h = 0; h < kNI; ++h) {
                    if (!exists[h]) continue;
                    const int I = br * kBT + wave + h * kW;
                    const bool ilive = ip[h] < n;
                    const int si = ip[h] >= ne;
                    float fx, fy, fz, u, rx, ry, rz;
                    if (uniform_tile(I) && juni)
                        n3_tile<kRdf, kKelbg, false, true>(k, a, hist, xtab, pi[h], si, ilive, sj, pj, fx, fy, fz, u, rx, ry, rz);
                    else
                        n3_tile<kRdf, kKelbg, false, false>(k, a, hist, xtab, pi[h], si, ilive, sj, pj, fx, fy, fz, u, rx, ry, rz);
                    Fx[h] += (double)fx;
                    Fy[h] += (double)fy;
                    Fz[h] += (double)fz;
                    U[h] += (double)u;
                    react(tj, rx, ry, rz, 1.0);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < kNI; ++h) {
            if (!exists[h] || ip[h] >= n) continue;
            double* pf = a.pf + (size_t)bc * 3 * n;
            pf[ip[h]] = Fx[h];
            pf[n + ip[h]] = Fy[h];
            pf[2 * n + ip[h]] = Fz[h];
            a.pu[(size_t)bc * n + ip[h]] = U[h];
        }
    }
    if (dblk) {
        // the tile pairs (ti, tj >= ti) of the diagonal block, in the same
        // j-tile turns; each wave adds its i sums to its own tiles at the end
        FixPos pi[kNI];
        bool exists[kNI], iuni[kNI];
        double Fx[kNI], Fy[kNI], Fz[kNI], U[kNI];
#pragma unroll
        for (int h = 0; h < kNI; ++h) {
            const int I = br * kBT + wave + h * kW;
            exists[h] = I < ntiles;
            pi[h] = a.ui[min(I * kTile + lane, n - 1)];
            // the i tile holds one species and only live particles
            iuni[h] = (I * kTile >= ne || I * kTile + kTile - 1 < ne) && I * kTile + kTile - 1 < n;
            Fx[h] = Fy[h] = Fz[h] = U[h] = 0.0;
        }
        for (int t = 0; t < L; ++t) {
            if (t && t % s == 0) __syncthreads();  // the last turn's reaction adds are done
            const int tj = jtile(t);
            if (tj >= m) continue;
            const int J = bc * kBT + tj;
            const FixPos pj = jpos[tj * kTile + lane];
            const int sj_uni = J * kTile >= ne;
            const bool juni = (J * kTile >= ne || J * kTile + kTile - 1 < ne) && J * kTile + kTile - 1 < n;
#pragma unroll
            for (int h = 0; h < kNI; ++h) {
                const int ti = wave + h * kW;
                if (!exists[h] || tj < ti) continue;
                const int ip = (br * kBT + ti) * kTile + lane;
                const bool ilive = ip < n;
                const int si = ip >= ne;
                const bool uni = iuni[h] && juni;
                float fx, fy, fz, u, rx, ry, rz;
                if (tj == ti) {
                    if (uni)
                        n3_tile<kRdf, kKelbg, true, true>(k, a, hist, xtab, pi[h], si, ilive, sj_uni, pj, fx, fy, fz, u, rx, ry, rz);
                    else
                        n3_tile<kRdf, kKelbg, true, false>(k, a, hist, xtab, pi[h], si, ilive, sj_uni, pj, fx, fy, fz, u, rx, ry, rz);
                } else {
                    if (uni)
                        n3_tile<kRdf, kKelbg, false, true>(k, a, hist, xtab, pi[h], si, ilive, sj_uni, pj, fx, fy, fz, u, rx, ry, rz);
                    else
                        n3_tile<kRdf, kKelbg, false, false>(k, a, hist, xtab, pi[h], si, ilive, sj_uni, pj, fx, fy, fz, u, rx, ry, rz);
                    react(tj, rx, ry, rz, 1.0);
                }
                Fx[h] += (double)fx;
                Fy[h] += (double)fy;
                Fz[h] += (double)fz;
                U[h] += (double)u;
            }
        }
        if (!kStage) __syncthreads();  // every reaction of the block is in racc
#pragma unroll
        for (int h = 0; h < kNI; ++h) {
            if (!exists[h]) continue;
            const int q = (wave + h * kW) * kTile + lane;
            racc[0][q] += Fx[h];
            racc[1][q] += Fy[h];
            racc[2][q] += Fz[h];
            racc[3][q] += U[h];
        }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < kBP; p += blockDim.x) {
        const int jp = bc * kBP + p;
        if (jp >= n) continue;
        double rx = racc[0][p], ry = racc[1][p], rz = racc[2][p];
        if (kStage) {  // the staged reactions, in wave order
            const int tj = p / kTile, l = p % kTile;
#pragma unroll
            for (int v = 0; v < kW; ++v) {
                const double sc = stg.scale[v * kBT + tj];
                const float* r = stg.r + (v * kBT + tj) * 3 * kTile + l;
                rx += sc * (double)r[0];
                ry += sc * (double)r[kTile];
                rz += sc * (double)r[2 * kTile];
            }
        }
        double* pf = a.pf + (size_t)br * 3 * n;
        pf[jp] = rx;
        pf[n + jp] = ry;
        pf[2 * n + jp] = rz;
        if (dblk) a.pu[(size_t)br * n + jp] = racc[3][p];  // zero off the diagonal: not stored
    }
    if (kRdf) {
        __syncthreads();
        hist_flush(hist, a.rdf, 1u);
    }
}

__device__ inline int queue_next(int* queue, int* slot)
{
    __syncthreads();  // every thread has read the previous item, LDS is free again
    if (threadIdx.x == 0) *slot = atomicAdd(queue, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

__device__ inline void queue_done(int* queue)
{
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(queue + 1, 1) == (int)gridDim.x - 1) {
            atomicExch(queue, 0);
            atomicExch(queue + 1, 0);
        }
    }
}


// ------------------------------------------- FP64, Newton's third law ---
//
// The same tiling as the mixed kernel (tiles of 64, 8 x 8 block pairs of the
// upper triangle, reactions rotated by DPP, j read from LDS at a rotating
// offset), one i tile per wave (8 waves).  The pair arithmetic is the
// reference's formula (force.f90:87-178) in FP64: its minimum image and r^2
// without contraction, so the cut-off test and the RDF bins are the
// reference's bit for bit, with erfc and exp from the tables below, so each
// pair's force agrees with the reference's to a few ulp, summed in a different
// order.  Reactions of the waves meet through a per-wave LDS buffer summed in
// wave order after each j tile, the i sums of a diagonal block join once every
// j tile is done, so the result does not depend on scheduling.

constexpr int kF64Tiles = 8;   // tiles per block side
constexpr int kF64Waves = 4;   // waves per 8 x 8 block: two i tiles each
constexpr int kFineTiles = 4;  // tiles per block side of the small-system shape (a wave per i tile)

template <int kBT>
constexpr int f64_waves() { return kBT == kFineTiles ? kFineTiles : kF64Waves; }

__device__ inline double rot_d(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long lo = rot_u((uint32_t)b), hi = rot_u((uint32_t)(b >> 32));
    return __longlong_as_double((long long)((hi << 32) | lo));
}

// The LDS copy of kErfcTab by coefficient pairs: pair j (coefficients 2j and
// 2j + 1) of row k at [j][k], 16 bytes apart from row to row.  A lane's four
// 16-byte reads of its row then fall on bank groups spread by its row index,
// where rows 64 bytes apart would put every lane's read of one pair on a
// quarter of the banks.
constexpr int kErfcPairs = (kErfcDeg + 1) / 2;
static_assert((kErfcDeg + 1) % 2 == 0, "the erfc rows split into coefficient pairs");
struct __align__(16) ErfcPair {
    double lo, hi;  // coefficients 2j and 2j + 1
};

// erfc(x), x >= 0, from the LDS copy of kErfcTab, given xs = kErfcScale x
// (64 x), and in dp the derivative dP/dt of the interval's polynomial:
// erfc'(x) = kErfcScale dP/dt = -(2/sqrt(pi)) e^{-x^2}, so the Gaussian term of
// the Ewald force, (2x/sqrt(pi)) e^{-x^2} = -xs dP/dt, and Kelbg's e^{-s^2}
// need no exponential.  The interval is k = floor(xs/2) (the zero row kErfcN
// for x >= 6; xs is clamped before the conversion, so any finite x is safe) and
// t = xs - (2k + 1) in [-1, 1), both exact.  P and dP/dt come from one Horner
// pass (degree 7: 13 fused multiply-adds).
__device__ inline double erfc_tab(double xs, const ErfcPair* tab, double& dp)
{
    const int ki = (int)fmin(xs, (double)(2 * kErfcN + 1));
    const double t = xs - (double)(ki | 1);
    const ErfcPair* row = tab + (ki >> 1);
    ErfcPair c[kErfcPairs];
#pragma unroll
    for (int j = 0; j < kErfcPairs; ++j) c[j] = row[j * (kErfcN + 1)];
    double p = c[kErfcPairs - 1].hi;
    double d = p;
    p = fma(p, t, c[kErfcPairs - 1].lo);
#pragma unroll
    for (int j = kErfcDeg - 2; j >= 0; --j) {
        d = fma(d, t, p);
        p = fma(p, t, (j & 1) ? c[j >> 1].hi : c[j >> 1].lo);
    }
    dp = d;
    return p;
}

// e^y for y <= 0 by Tang's table-driven reduction (ACM TOMS 15, 1989):
// y = (256 n + j) ln2/256 + r with |r| <= ln2/512, e^y = 2^n 2^(j/256) e^r,
// 2^(j/256) from kExp2Tab (read through the vector cache: only the Deutsch
// form calls this, so the table stays out of LDS) and e^r - 1 by its Taylor polynomial
// to r^4 (remainder below r^5/120 < 4e-17); ln2/256 as a sum of two doubles
// keeps r to an ulp.  About 1.5 ulp; underflows to 0 past -745.
__device__ inline double exp_tab(double y, const double* e2tab)
{
    constexpr double k256_ln2 = 0x1.71547652b82fep+8;  // 256 / ln 2
    constexpr double ln2_256_hi = 0x1.62e42fefa39efp-9;  // ln 2 / 256
    constexpr double ln2_256_lo = 0x1.abc9e3b39803fp-64; // ln 2 / 256 - ln2_256_hi
    const double k = rint(y * k256_ln2);
    const double r = fma(-k, ln2_256_lo, fma(-k, ln2_256_hi, y));
    const int ki = (int)k;
    double q = fma(r, 1.0 / 24.0, 1.0 / 6.0);
    q = fma(q, r, 0.5);
    q = fma(q, r, 1.0);
    q *= r;  // e^r - 1
    const double t = e2tab[ki & (kExp2N - 1)];
    return ldexp(fma(t, q, t), ki >> kExp2Bits);
}

// 1/sqrt(r2) to about one ulp: the hardware estimate and two Newton steps
__device__ inline double rsq_f64(double r2)
{
    const double h = 0.5 * r2;
    double y = __builtin_amdgcn_rsq(r2);
    double e = fma(-(h * y), y, 0.5);
    y = fma(y, e, y);
    e = fma(-(h * y), y, 0.5);
    return fma(y, e, y);
}

// A j of the FP64 block's LDS copy: position and flags (species bit 0, live bit 1).
struct __align__(32) F64J {
    double x, y, z;
    int f, pad;
};

// One tile pair: lane i against the 64 j of the tile at jt (LDS).  Lane l meets
// at step s the j of lane l - s, whose byte offset lane l - 1 held at step
// s - 1: the offset rotates with the reactions (one DPP move) and the next j is
// read while this pair computes.  kDiag: I == J, ordered pairs for
// the i side only, half energy, RDF weight 1.  The cut-off test and the RDF bin
// are the reference's bit for bit: minimum image and r^2 without contraction,
// and bin = int(sqrt(r2) / dx) with both operations correctly rounded, which
// r * (1/dx) (r = r2 / sqrt(r2) to an ulp or two) decides unless its fraction
// lies within 1e-12 of a bin edge; there the exact sqrt and division decide.
// erfc(x) and the Gaussian term come from one row of the erfc table.
template <bool kRdf, bool kKelbg, bool kDiag>
__device__ inline void n3_tile_f64(const PairArgs& a, unsigned* hist, const ErfcPair* etab,
                                   const double* e2tab, double xi, double yi, double zi, int si,
                                   bool ilive, const F64J* jt, double& fxo, double& fyo,
                                   double& fzo, double& uo, double& rx, double& ry, double& rz)
{
    const char* jb = (const char*)jt;
    uint32_t off = (uint32_t)(threadIdx.x % kTile) * (uint32_t)sizeof(F64J);
    F64J q = *(const F64J*)(jb + off);
    const double Lx = a.box[0], Ly = a.box[1], Lz = a.box[2];
    const double iLx = a.inv_box[0], iLy = a.inv_box[1], iLz = a.inv_box[2];
    const double alpha_s = kErfcScale * a.alpha, rc2 = a.rc2;
    const double inv_rs = 1.0 / a.rs, inv_rs_s = kErfcScale / a.rs;
    // per-i constants for a j electron (0) / a j ion (1)
    const double qi = si ? a.q[1] : a.q[0];
    const double eqi = a.eps * qi;
    const double qj0 = a.q[0], qj1 = a.q[1];
    const double dx0 = si ? a.dx[1] : a.dx[0], dx1 = si ? a.dx[2] : a.dx[1];
    const double idx0 = si ? a.inv_dx[1] : a.inv_dx[0], idx1 = si ? a.inv_dx[2] : a.inv_dx[1];
    const bool reg0 = si != 0 || (a.flags & TCP_REG_EE) != 0, reg1 = si == 0;
    double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0;
    rx = ry = rz = 0.0;
#pragma unroll 2
    for (int s = 0; s < kTile; ++s) {  // by two: the carried offset and j alternate registers
        const uint32_t noff = rot_u(off);
        const F64J nq = *(const F64J*)(jb + noff);
        const int fj = q.f;
        const int sj = fj & 1;
        double xr, yr, zr, r2;
        {
#pragma clang fp contract(off)
            const double dx = xi - q.x;
            const double dy = yi - q.y;
            const double dz = zi - q.z;
            xr = dx - Lx * rint(dx * iLx);
            yr = dy - Ly * rint(dy * iLy);
            zr = dz - Lz * rint(dz * iLz);
            r2 = xr * xr + yr * yr + zr * zr;
        }
        if (r2 < rc2 && r2 > 0.0 && ilive && (fj & 2) != 0) {
            const double inv_r = rsq_f64(r2);
            const double r = r2 * inv_r;
            if (kRdf) {
                const double q = r * (sj ? idx1 : idx0);
                int b = (int)q;
                if (fabs(__builtin_amdgcn_fract(q) - 0.5) > 0.5 - 1e-12)
                    b = (int)(sqrt(r2) / (sj ? dx1 : dx0));
                if (b < kNbin) atomicAdd(&hist[(si + sj) * kHistStride + b], kDiag ? 1u : 2u);
            }
            const double qq = qi * (sj ? qj1 : qj0);
            const double inv_r2 = inv_r * inv_r;
            // num = erfc(x) + (2x/sqrt(pi)) e^{-x^2}, x = alpha r
            double d;
            const double xs = alpha_s * r;
            const double ec = erfc_tab(xs, etab, d);
            double num = fma(-xs, d, ec), en = ec;
            if (sj ? reg1 : reg0) {
                if (kKelbg) {
                    // R = -e^{-s^2} + sqrt(pi) s erfc(s), s = r/lambda
                    double ds;
                    const double ss = r * inv_rs_s;
                    const double es = erfc_tab(ss, etab, ds);
                    const double y2 = ds * (-0.5 * kErfcScale * kSqrtPi);
                    num -= y2;
                    en += fma((kSqrtPi / kErfcScale) * ss, es, -y2);
                } else {
                    const double sr = r * inv_rs;
                    // R = -e^{-s} (Deutsch, force.f90:136-139)
                    const double y = exp_tab(-sr, e2tab);
                    num -= y * (1.0 + sr);
                    en -= y;
                }
            }
            const double df = eqi * (sj ? qj1 : qj0) * num * inv_r2 * inv_r;
            u += (kDiag ? 0.5 : 1.0) * (qq * en * inv_r);
            fx += df * xr;
            fy += df * yr;
            fz += df * zr;
            if (!kDiag) {
                rx -= df * xr;
                ry -= df * yr;
                rz -= df * zr;
            }
        }
        off = noff;
        q = nq;
        if (!kDiag) {
            rx = rot_d(rx);
            ry = rot_d(ry);
            rz = rot_d(rz);
        }
    }
    fxo = fx;
    fyo = fy;
    fzo = fz;
    uo = u;
}

// n3_tile_f64 for a tile pair whose two tiles are full and hold one species
// each (si, sj wave-uniform): every per-species constant is a scalar, the
// regularisation branch is uniform, no live flag is read, and the sums are
// charge-free (eps qi qj and qi qj applied once at the end).  The minimum image
// is dx - L rint(dx/L) as one fma, which equals the separately rounded product
// and difference while L rint(dx/L) is exact (|dx| < 2.5 L; wrapped positions
// have |dx| <= L), so r^2, the cut-off and the RDF bins stay the reference's.
template <bool kRdf, bool kKelbg, bool kDiag>
__device__ inline void n3_tile_f64_uni(const PairArgs& a, unsigned* hist, const ErfcPair* etab,
                                       const double* e2tab, double xi, double yi, double zi,
                                       int si, int sj, const F64J* jt, double& fxo, double& fyo,
                                       double& fzo, double& uo, double& rx, double& ry, double& rz)
{
    const char* jb = (const char*)jt;
    uint32_t off = (uint32_t)(threadIdx.x % kTile) * (uint32_t)sizeof(F64J);
    double qx = *(const double*)(jb + off), qy = *(const double*)(jb + off + 8),
           qz = *(const double*)(jb + off + 16);
    const double Lx = a.box[0], Ly = a.box[1], Lz = a.box[2];
    const double iLx = a.inv_box[0], iLy = a.inv_box[1], iLz = a.inv_box[2];
    const double alpha_s = kErfcScale * a.alpha, rc2 = a.rc2;
    const double inv_rs = 1.0 / a.rs, inv_rs_s = kErfcScale / a.rs;
    const int ht = si + sj;
    const double dxb = a.dx[ht], idxb = a.inv_dx[ht];
    const bool reg = sj == 0 ? (si != 0 || (a.flags & TCP_REG_EE) != 0) : si == 0;
    unsigned* hrow = hist + ht * kHistStride;
    double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0;
    rx = ry = rz = 0.0;
#pragma unroll 2
    for (int s = 0; s < kTile; ++s) {  // by two: the carried offset and j alternate registers
        const uint32_t noff = rot_u(off);
        const double nx = *(const double*)(jb + noff), ny = *(const double*)(jb + noff + 8),
                     nz = *(const double*)(jb + noff + 16);
        double xr, yr, zr, r2;
        {
#pragma clang fp contract(off)
            const double dx = xi - qx;
            const double dy = yi - qy;
            const double dz = zi - qz;
            xr = fma(-Lx, rint(dx * iLx), dx);
            yr = fma(-Ly, rint(dy * iLy), dy);
            zr = fma(-Lz, rint(dz * iLz), dz);
            r2 = xr * xr + yr * yr + zr * zr;
        }
        if (r2 < rc2 && r2 > 0.0) {
            const double inv_r = rsq_f64(r2);
            const double r = r2 * inv_r;
            if (kRdf) {
                const double q = r * idxb;
                int b = (int)q;
                if (fabs(__builtin_amdgcn_fract(q) - 0.5) > 0.5 - 1e-12) b = (int)(sqrt(r2) / dxb);
                if (b < kNbin) atomicAdd(&hrow[b], kDiag ? 1u : 2u);
            }
            double d;
            const double xs = alpha_s * r;
            const double ec = erfc_tab(xs, etab, d);
            double num = fma(-xs, d, ec), en = ec;
            if (reg) {
                if (kKelbg) {
                    double ds;
                    const double ss = r * inv_rs_s;
                    const double es = erfc_tab(ss, etab, ds);
                    const double y2 = ds * (-0.5 * kErfcScale * kSqrtPi);
                    num -= y2;
                    en += fma((kSqrtPi / kErfcScale) * ss, es, -y2);
                } else {
                    const double sr = r * inv_rs;
                    const double y = exp_tab(-sr, e2tab);
                    num -= y * (1.0 + sr);
                    en -= y;
                }
            }
            const double df = num * (inv_r * inv_r) * inv_r;
            u = fma(en, inv_r, u);
            fx = fma(df, xr, fx);
            fy = fma(df, yr, fy);
            fz = fma(df, zr, fz);
            if (!kDiag) {
                rx = fma(-df, xr, rx);
                ry = fma(-df, yr, ry);
                rz = fma(-df, zr, rz);
            }
        }
        off = noff;
        qx = nx;
        qy = ny;
        qz = nz;
        if (!kDiag) {
            rx = rot_d(rx);
            ry = rot_d(ry);
            rz = rot_d(rz);
        }
    }
    const double qq = a.q[si] * a.q[sj];
    const double eqq = a.eps * qq;
    fxo = eqq * fx;
    fyo = eqq * fy;
    fzo = eqq * fz;
    uo = (kDiag ? 0.5 * qq : qq) * u;
    rx *= eqq;
    ry *= eqq;
    rz *= eqq;
}

// n3_tile_f64_uni for two i tiles of one species against one j tile, off the
// diagonal: each step's j coordinates and the rotation of the reactions serve
// both pairs, and each pair is the single-tile path's arithmetic.
template <bool kRdf, bool kKelbg>
__device__ inline void n3_tile_f64_uni2(const PairArgs& a, unsigned* hist, const ErfcPair* etab,
                                       const double* e2tab, const double (&xi)[2],
                                       const double (&yi)[2], const double (&zi)[2], int si,
                                       int sj, const F64J* jt, double (&fxo)[2],
                                       double (&fyo)[2], double (&fzo)[2], double (&uo)[2],
                                       double& rx, double& ry, double& rz)
{
    constexpr int kNI = 2;
    constexpr bool kDiag = false;
    const char* jb = (const char*)jt;
    uint32_t off = (uint32_t)(threadIdx.x % kTile) * (uint32_t)sizeof(F64J);
    double qx = *(const double*)(jb + off), qy = *(const double*)(jb + off + 8),
           qz = *(const double*)(jb + off + 16);
    const double Lx = a.box[0], Ly = a.box[1], Lz = a.box[2];
    const double iLx = a.inv_box[0], iLy = a.inv_box[1], iLz = a.inv_box[2];
    const double alpha_s = kErfcScale * a.alpha, rc2 = a.rc2;
    const double inv_rs = 1.0 / a.rs, inv_rs_s = kErfcScale / a.rs;
    const int ht = si + sj;
    const double dxb = a.dx[ht], idxb = a.inv_dx[ht];
    const bool reg = sj == 0 ? (si != 0 || (a.flags & TCP_REG_EE) != 0) : si == 0;
    unsigned* hrow = hist + ht * kHistStride;
    double fx[kNI], fy[kNI], fz[kNI], u[kNI];
#pragma unroll
    for (int h = 0; h < kNI; ++h) fx[h] = fy[h] = fz[h] = u[h] = 0.0;
    rx = ry = rz = 0.0;
#pragma unroll 2
    for (int s = 0; s < kTile; ++s) {  // by two: the carried offset and j alternate registers
        const uint32_t noff = rot_u(off);
        const double nx = *(const double*)(jb + noff), ny = *(const double*)(jb + noff + 8),
                     nz = *(const double*)(jb + noff + 16);
#pragma unroll
        for (int h = 0; h < kNI; ++h) {
            double xr, yr, zr, r2;
            {
#pragma clang fp contract(off)
                const double dx = xi[h] - qx;
                const double dy = yi[h] - qy;
                const double dz = zi[h] - qz;
                xr = fma(-Lx, rint(dx * iLx), dx);
                yr = fma(-Ly, rint(dy * iLy), dy);
                zr = fma(-Lz, rint(dz * iLz), dz);
                r2 = xr * xr + yr * yr + zr * zr;
            }
            if (r2 < rc2 && r2 > 0.0) {
                const double inv_r = rsq_f64(r2);
                const double r = r2 * inv_r;
                if (kRdf) {
                    const double q = r * idxb;
                    int b = (int)q;
                    if (fabs(__builtin_amdgcn_fract(q) - 0.5) > 0.5 - 1e-12) b = (int)(sqrt(r2) / dxb);
                    if (b < kNbin) atomicAdd(&hrow[b], kDiag ? 1u : 2u);
                }
                double d;
                const double xs = alpha_s * r;
                const double ec = erfc_tab(xs, etab, d);
                double num = fma(-xs, d, ec), en = ec;
                if (reg) {
                    if (kKelbg) {
                        double ds;
                        const double ss = r * inv_rs_s;
                        const double es = erfc_tab(ss, etab, ds);
                        const double y2 = ds * (-0.5 * kErfcScale * kSqrtPi);
                        num -= y2;
                        en += fma((kSqrtPi / kErfcScale) * ss, es, -y2);
                    } else {
                        const double sr = r * inv_rs;
                        const double y = exp_tab(-sr, e2tab);
                        num -= y * (1.0 + sr);
                        en -= y;
                    }
                }
                const double df = num * (inv_r * inv_r) * inv_r;
                u[h] = fma(en, inv_r, u[h]);
                fx[h] = fma(df, xr, fx[h]);
                fy[h] = fma(df, yr, fy[h]);
                fz[h] = fma(df, zr, fz[h]);
                if (!kDiag) {
                    rx = fma(-df, xr, rx);
                    ry = fma(-df, yr, ry);
                    rz = fma(-df, zr, rz);
                }
            }
        }
        off = noff;
        qx = nx;
        qy = ny;
        qz = nz;
        if (!kDiag) {
            rx = rot_d(rx);
            ry = rot_d(ry);
            rz = rot_d(rz);
        }
    }
    const double qq = a.q[si] * a.q[sj];
    const double eqq = a.eps * qq;
#pragma unroll
    for (int h = 0; h < kNI; ++h) {
        fxo[h] = eqq * fx[h];
        fyo[h] = eqq * fy[h];
        fzo[h] = eqq * fz[h];
        uo[h] = (kDiag ? 0.5 * qq : qq) * u[h];
    }
    rx *= eqq;
    ry *= eqq;
    rz *= eqq;
}

// LDS of one FP64 block of kBT x kBT tiles worked by kW waves: j positions and
// flags, the block's reaction sums (slot br), the per-wave reaction buffer, the
// RDF histogram and the erfc table (52 KiB for 8 x 8 tiles: three blocks per
// CU).
template <int kBT, int kW>
struct F64Lds {
    F64J jp[kBT * kTile];
    double racc[3][kBT * kTile];
    double wbuf[kW][3][kTile];
    unsigned hist[3 * kHistStride];
    ErfcPair etab[kErfcPairs * (kErfcN + 1)];
};

static_assert(sizeof(F64Lds<kF64Tiles, kF64Waves>) + 64 <= 160 * 1024 / 3,
              "three 8 x 8 FP64 blocks fit in one CU's LDS");

template <int kBT, int kW>
__device__ inline void load_tables(F64Lds<kBT, kW>& L)
{
    for (int i = threadIdx.x; i < kErfcPairs * (kErfcN + 1); i += blockDim.x) {
        const int j = i / (kErfcN + 1), k = i - j * (kErfcN + 1);
        L.etab[i] = ErfcPair{kErfcTab[k][2 * j], kErfcTab[k][2 * j + 1]};
    }
}

// kBT x kBT tiles by kW waves, kNI = kBT / kW i tiles per wave: wave w holds
// the block's i tiles w, w + kW, ..  For each j tile the wave's i tiles walk
// it together when the tile pairs are uniform and off the diagonal (every
// step's j read and reaction rotation serving kNI pairs), else one after the
// other, their reactions added in i-tile order; then the waves' reactions for
// the j tile are summed in wave order.  A diagonal block adds its i sums to
// those after the last j tile.  Every sum has a fixed order.
template <bool kRdf, bool kKelbg, int kBT, int kW>
__device__ inline void n3_block_f64(const PairArgs& a, int br, int bc, F64Lds<kBT, kW>& L)
{
    constexpr int kBP = kBT * kTile;
    constexpr int kNI = kBT / kW;
    static_assert(kBT % kW == 0 && kNI <= 2, "one or two i tiles per wave");
    const bool dblk = br == bc;
    const int n = a.n_i, ne = a.ne_i;
    for (int p = threadIdx.x; p < kBP; p += blockDim.x) {
        const int jp = bc * kBP + p;
        const int jc = min(jp, n - 1);
        L.jp[p] = F64J{a.pi.x[jc], a.pi.y[jc], a.pi.z[jc], (jp >= ne ? 1 : 0) | (jp < n ? 2 : 0), 0};
    }
    if (kRdf) hist_clear(L.hist);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile), lane = threadIdx.x % kTile;
    const int ntiles = (n + kTile - 1) / kTile;
    // a full tile of one species
    auto uniform = [&](int T) {
        const int p0 = T * kTile, p1 = p0 + kTile - 1;
        return p1 < n && (p0 >= ne || p1 < ne);
    };
    int tl[kNI], ip[kNI], si[kNI], siu[kNI];
    bool exists[kNI], ilive[kNI], iuni[kNI];
    double xi[kNI], yi[kNI], zi[kNI], Fx[kNI], Fy[kNI], Fz[kNI], U[kNI];
#pragma unroll
    for (int h = 0; h < kNI; ++h) {
        tl[h] = wave + h * kW;  // i tile within the block
        const int I = br * kBT + tl[h];
        exists[h] = I < ntiles;
        ip[h] = I * kTile + lane;
        ilive[h] = exists[h] && ip[h] < n;
        const int ic = min(ip[h], n - 1);
        xi[h] = a.pi.x[ic];
        yi[h] = a.pi.y[ic];
        zi[h] = a.pi.z[ic];
        si[h] = ip[h] >= ne;
        iuni[h] = uniform(I);
        siu[h] = I * kTile >= ne;
        Fx[h] = Fy[h] = Fz[h] = U[h] = 0.0;
    }
    for (int tj = 0; tj < kBT; ++tj) {
        const int J = bc * kBT + tj;
        const F64J* jt = L.jp + tj * kTile;
        const bool juni = J < ntiles && uniform(J);
        const int sju = J * kTile >= ne;
        double rx = 0.0, ry = 0.0, rz = 0.0;
        bool both = kNI == 2 && !dblk && juni;
#pragma unroll
        for (int h = 0; h < kNI; ++h) both = both && exists[h] && iuni[h] && siu[h] == siu[0];
        if (both) {
            double fx[2], fy[2], fz[2], u[2];
            const double x2[2] = {xi[0], xi[kNI - 1]}, y2[2] = {yi[0], yi[kNI - 1]},
                         z2[2] = {zi[0], zi[kNI - 1]};
            n3_tile_f64_uni2<kRdf, kKelbg>(a, L.hist, L.etab, kExp2Tab, x2, y2, z2, siu[0], sju, jt,
                                           fx, fy, fz, u, rx, ry, rz);
#pragma unroll
            for (int h = 0; h < kNI; ++h) {
                Fx[h] += fx[h];
                Fy[h] += fy[h];
                Fz[h] += fz[h];
                U[h] += u[h];
            }
        } else {
#pragma unroll
            for (int h = 0; h < kNI; ++h) {
                if (!(exists[h] && J < ntiles && (!dblk || tj >= tl[h]))) continue;
                double fx, fy, fz, u, hx, hy, hz;
                const bool diag = dblk && tj == tl[h];
                if (iuni[h] && juni) {
                    if (diag)
                        n3_tile_f64_uni<kRdf, kKelbg, true>(a, L.hist, L.etab, kExp2Tab, xi[h], yi[h],
                                                            zi[h], siu[h], sju, jt, fx, fy, fz, u, hx,
                                                            hy, hz);
                    else
                        n3_tile_f64_uni<kRdf, kKelbg, false>(a, L.hist, L.etab, kExp2Tab, xi[h], yi[h],
                                                             zi[h], siu[h], sju, jt, fx, fy, fz, u, hx,
                                                             hy, hz);
                } else if (diag)
                    n3_tile_f64<kRdf, kKelbg, true>(a, L.hist, L.etab, kExp2Tab, xi[h], yi[h], zi[h],
                                                    si[h], ilive[h], jt, fx, fy, fz, u, hx, hy, hz);
                else
                    n3_tile_f64<kRdf, kKelbg, false>(a, L.hist, L.etab, kExp2Tab, xi[h], yi[h], zi[h],
                                                     si[h], ilive[h], jt, fx, fy, fz, u, hx, hy, hz);
                Fx[h] += fx;
                Fy[h] += fy;
                Fz[h] += fz;
                U[h] += u;
                rx += hx;
                ry += hy;
                rz += hz;
            }
        }
        L.wbuf[wave][0][lane] = rx;
        L.wbuf[wave][1][lane] = ry;
        L.wbuf[wave][2][lane] = rz;
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * kTile; t += blockDim.x) {
            const int c = t / kTile, p = t % kTile;
            double s = 0.0;
            for (int w = 0; w < kW; ++w) s += L.wbuf[w][c][p];
            L.racc[c][tj * kTile + p] = s;
        }
        __syncthreads();
    }
    // the i sums: slot bc, or on the diagonal added to the reactions (slot br;
    // the energy slots below the diagonal hold zeros and are not stored)
#pragma unroll
    for (int h = 0; h < kNI; ++h) {
        if (!exists[h]) continue;
        if (dblk) {
            const int q = tl[h] * kTile + lane;
            L.racc[0][q] += Fx[h];
            L.racc[1][q] += Fy[h];
            L.racc[2][q] += Fz[h];
            if (ilive[h]) a.pu[(size_t)br * n + ip[h]] = U[h];
        } else if (ilive[h]) {
            double* pf = a.pf + (size_t)bc * 3 * n;
            pf[ip[h]] = Fx[h];
            pf[n + ip[h]] = Fy[h];
            pf[2 * n + ip[h]] = Fz[h];
            a.pu[(size_t)bc * n + ip[h]] = U[h];
        }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < kBP; p += blockDim.x) {
        const int jp = bc * kBP + p;
        if (jp >= n) continue;
        double* pf = a.pf + (size_t)br * 3 * n;
        pf[jp] = L.racc[0][p];
        pf[n + jp] = L.racc[1][p];
        pf[2 * n + jp] = L.racc[2][p];
    }
    if (kRdf) {
        __syncthreads();
        hist_flush(L.hist, a.rdf, 1u);
    }
}

// 8 x 8 tile blocks by four waves, two i tiles each (51 KB of LDS, at most 168
// VGPRs: three blocks per CU, three waves per SIMD, each with two independent
// pair chains)
template <bool kRdf, bool kKelbg>
__global__ __launch_bounds__(kF64Waves * kTile, 3) void pair_f64_n3_kernel(PairArgs a,
                                                                        const int2* blkmap,
                                                                        int nblk, int* queue)
{
    __shared__ F64Lds<kF64Tiles, kF64Waves> lds;
    __shared__ int slot;
    load_tables(lds);  // the first queue_next orders it before use
    for (;;) {
        const int it = queue_next(queue, &slot);
        if (it >= nblk) break;
        const int2 rc = blkmap[it];
        n3_block_f64<kRdf, kKelbg, kF64Tiles, kF64Waves>(a, rc.x, rc.y, lds);
    }
    queue_done(queue);
}

// The same over blocks of 4 x 4 tiles by four waves, one i tile each, for small
// systems (three blocks per CU, three waves per SIMD).
template <bool kRdf, bool kKelbg>
__global__ __launch_bounds__(kFineTiles * kTile, 3) void pair_f64_n3_fine_kernel(PairArgs a,
                                                                              const int2* blkmap,
                                                                              int nblk, int* queue)
{
    __shared__ F64Lds<kFineTiles, kFineTiles> lds;
    __shared__ int slot;
    load_tables(lds);  // the first queue_next orders it before use
    for (;;) {
        const int it = queue_next(queue, &slot);
        if (it >= nblk) break;
        const int2 rc = blkmap[it];
        n3_block_f64<kRdf, kKelbg, kFineTiles, kFineTiles>(a, rc.x, rc.y, lds);
    }
    queue_done(queue);
}

template <bool kRdf, bool kKelbg, int kBT>
__global__ __launch_bounds__(f64_waves<kBT>() * kTile, 3) void pair_f64_n3_batch_kernel(const BatchRep* reps,
                                                                        const PairItem* items,
                                                                        int nitems, int* queue)
{
    __shared__ F64Lds<kBT, f64_waves<kBT>()> lds;
    __shared__ int slot;
    load_tables(lds);  // the first queue_next orders it before use
    for (;;) {
        const int it = queue_next(queue, &slot);
        if (it >= nitems) break;
        const PairItem w = items[it];
        n3_block_f64<kRdf, kKelbg, kBT, f64_waves<kBT>()>(reps[w.rep].pa, w.a, w.b, lds);
    }
    queue_done(queue);
}

// Persistent: resident blocks take the block pairs of blkmap (full blocks
// first, the half-cost diagonal ones last) from a device queue.
template <bool kRdf, bool kKelbg>
__global__ __launch_bounds__(kN3Waves * kTile, 4) void pair_mixed_n3_kernel(PairArgs a,
                                                                         const int2* blkmap,
                                                                         int nblk, int* queue)
{
    __shared__ FixPos jpos[kBlkTiles * kTile];
    __shared__ double racc[4][kBlkTiles * kTile];
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    __shared__ N3Stage<kBlkTiles, kN3Waves> stg;
    __shared__ float4 xtab[kErfcxRows];
    __shared__ int slot;
    load_erfcx(xtab);  // the first queue_next orders it before use
    for (;;) {
        const int it = queue_next(queue, &slot);
        if (it >= nblk) break;
        const int2 rc = blkmap[it];
        n3_block<kRdf, kKelbg, kBlkTiles, kN3Waves>(a, rc.x, rc.y, jpos, racc, hist, xtab, stg);
    }
    queue_done(queue);
}

// The same over blocks of 4 x 4 tiles by four waves (one i tile each), for
// small systems, whose 8 x 8 tile blocks would leave most of the chip idle.
template <bool kRdf, bool kKelbg>
__global__ __launch_bounds__(kN3Waves * kTile, 4) void pair_mixed_n3_fine_kernel(PairArgs a,
                                                                              const int2* blkmap,
                                                                              int nblk, int* queue)
{
    constexpr int kBT = 4;
    __shared__ FixPos jpos[kBT * kTile];
    __shared__ double racc[4][kBT * kTile];
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    __shared__ N3Stage<kBT, kN3Waves> stg;
    __shared__ float4 xtab[kErfcxRows];
    __shared__ int slot;
    load_erfcx(xtab);  // the first queue_next orders it before use
    for (;;) {
        const int it = queue_next(queue, &slot);
        if (it >= nblk) break;
        const int2 rc = blkmap[it];
        n3_block<kRdf, kKelbg, kBT, kN3Waves>(a, rc.x, rc.y, jpos, racc, hist, xtab, stg);
    }
    queue_done(queue);
}

// ------------------------------------------------ persistent (batches) ---
//
// A fixed grid of resident blocks takes work items (replica, block pair) from
// a device queue in list order (the host sorts the list by decreasing cost),
// so replicas of very different sizes share one launch.  The item body is the
// single-context kernel's, so a replica's partial slots hold exactly what its
// own launch would write.  The last block to finish resets the queue, which
// keeps the launch free of memset nodes when it is captured in a graph.

template <bool kRdf, bool kKelbg, int kBT, int kW>
// four waves per SIMD (128 VGPRs) where two i tiles per wave fit in them, two
// where four do not
__global__ __launch_bounds__(kW * kTile, (kBT / kW) > 2 ? 2 : 4) void pair_mixed_n3_batch_kernel(const BatchRep* reps,
                                                                         const PairItem* items,
                                                                         int nitems, int* queue)
{
    __shared__ FixPos jpos[kBT * kTile];
    __shared__ double racc[4][kBT * kTile];
    __shared__ unsigned hist[kRdf ? 3 * kHistStride : 1];
    __shared__ N3Stage<kBT, kW> stg;
    __shared__ float4 xtab[kErfcxRows];
    __shared__ int slot;
    load_erfcx(xtab);  // the first queue_next orders it before use
    for (;;) {
        const int it = queue_next(queue, &slot);
        if (it >= nitems) break;
        const PairItem w = items[it];
        n3_block<kRdf, kKelbg, kBT, kW>(reps[w.rep].pa, w.a, w.b, jpos, racc, hist, xtab, stg);
    }
    queue_done(queue);
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

void launch_pair_adjoint(const AdjArgs& a, int nchunk, hipStream_t s)
{
    dim3 grid((a.n + kPairBlock - 1) / kPairBlock, nchunk);
    hipLaunchKernelGGL(pair_adjoint_kernel, grid, dim3(kPairBlock), 0, s, a);
}

void launch_adjoint_reduce(const double* part, int nchunk, int n, double* out, hipStream_t s)
{
    const size_t m = (size_t)kAdjQ * n;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                       part, nchunk, n, out);
}

namespace {

template <bool kRdf, bool kKelbg>
const void* mixed_n3_fn(int shape)
{
    return shape == kFineN3Shape ? (const void*)pair_mixed_n3_fine_kernel<kRdf, kKelbg>
                                 : (const void*)pair_mixed_n3_kernel<kRdf, kKelbg>;
}

}  // namespace

int pair_n3_occupancy(bool rdf, bool kelbg, int shape)
{
    const void* f = rdf ? (kelbg ? mixed_n3_fn<true, true>(shape) : mixed_n3_fn<true, false>(shape))
                        : (kelbg ? mixed_n3_fn<false, true>(shape) : mixed_n3_fn<false, false>(shape));
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kN3Waves * kTile, 0) != hipSuccess) nb = 1;
    return nb < 1 ? 1 : nb;
}

namespace {

template <bool kRdf, bool kKelbg>
const void* f64_n3_fn(int shape)
{
    return shape == kFineN3Shape ? (const void*)pair_f64_n3_fine_kernel<kRdf, kKelbg>
                                 : (const void*)pair_f64_n3_kernel<kRdf, kKelbg>;
}

int f64_threads(int shape) { return (shape == kFineN3Shape ? kFineTiles : kF64Waves) * kTile; }

}  // namespace

int pair_f64_n3_occupancy(bool rdf, bool kelbg, int shape)
{
    const void* f = rdf ? (kelbg ? f64_n3_fn<true, true>(shape) : f64_n3_fn<true, false>(shape))
                        : (kelbg ? f64_n3_fn<false, true>(shape) : f64_n3_fn<false, false>(shape));
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, f64_threads(shape), 0) != hipSuccess) nb = 1;
    return nb < 1 ? 1 : nb;
}

void launch_pair_f64_n3(const PairArgs& a, const int2* blkmap, int nblk, int* queue, int grid,
                        int shape, hipStream_t s)
{
    const dim3 g(grid), b(f64_threads(shape));
    const bool kel = (a.flags & TCP_KELBG) != 0;
    if (shape == kFineN3Shape) {
        if (a.rdf_weight) {
            if (kel) hipLaunchKernelGGL((pair_f64_n3_fine_kernel<true, true>), g, b, 0, s, a, blkmap, nblk, queue);
            else hipLaunchKernelGGL((pair_f64_n3_fine_kernel<true, false>), g, b, 0, s, a, blkmap, nblk, queue);
        } else {
            if (kel) hipLaunchKernelGGL((pair_f64_n3_fine_kernel<false, true>), g, b, 0, s, a, blkmap, nblk, queue);
            else hipLaunchKernelGGL((pair_f64_n3_fine_kernel<false, false>), g, b, 0, s, a, blkmap, nblk, queue);
        }
        return;
    }
    if (a.rdf_weight) {
        if (kel) hipLaunchKernelGGL((pair_f64_n3_kernel<true, true>), g, b, 0, s, a, blkmap, nblk, queue);
        else hipLaunchKernelGGL((pair_f64_n3_kernel<true, false>), g, b, 0, s, a, blkmap, nblk, queue);
    } else {
        if (kel) hipLaunchKernelGGL((pair_f64_n3_kernel<false, true>), g, b, 0, s, a, blkmap, nblk, queue);
        else hipLaunchKernelGGL((pair_f64_n3_kernel<false, false>), g, b, 0, s, a, blkmap, nblk, queue);
    }
}

void launch_pair_mixed_n3(const PairArgs& a, const int2* blkmap, int nblk, int* queue, int grid,
                          int shape, hipStream_t s)
{
    const dim3 g(grid), b(kN3Waves * kTile);
    const bool kel = (a.flags & TCP_KELBG) != 0;
    if (shape == kFineN3Shape) {
        if (a.rdf_weight) {
            if (kel) hipLaunchKernelGGL((pair_mixed_n3_fine_kernel<true, true>), g, b, 0, s, a, blkmap, nblk, queue);
            else hipLaunchKernelGGL((pair_mixed_n3_fine_kernel<true, false>), g, b, 0, s, a, blkmap, nblk, queue);
        } else {
            if (kel) hipLaunchKernelGGL((pair_mixed_n3_fine_kernel<false, true>), g, b, 0, s, a, blkmap, nblk, queue);
            else hipLaunchKernelGGL((pair_mixed_n3_fine_kernel<false, false>), g, b, 0, s, a, blkmap, nblk, queue);
        }
        return;
    }
    if (a.rdf_weight) {
        if (kel) hipLaunchKernelGGL((pair_mixed_n3_kernel<true, true>), g, b, 0, s, a, blkmap, nblk, queue);
        else hipLaunchKernelGGL((pair_mixed_n3_kernel<true, false>), g, b, 0, s, a, blkmap, nblk, queue);
    } else {
        if (kel) hipLaunchKernelGGL((pair_mixed_n3_kernel<false, true>), g, b, 0, s, a, blkmap, nblk, queue);
        else hipLaunchKernelGGL((pair_mixed_n3_kernel<false, false>), g, b, 0, s, a, blkmap, nblk, queue);
    }
}

void launch_pair_mixed(const PairArgs& a, int nchunk, hipStream_t s)
{
    const dim3 grid((a.n_i + kPairBlock - 1) / kPairBlock, nchunk), b(kPairBlock);
    const bool kel = (a.flags & TCP_KELBG) != 0;
    if (a.rdf_weight) {
        if (kel) hipLaunchKernelGGL((pair_mixed_kernel<true, true>), grid, b, 0, s, a);
        else hipLaunchKernelGGL((pair_mixed_kernel<true, false>), grid, b, 0, s, a);
    } else {
        if (kel) hipLaunchKernelGGL((pair_mixed_kernel<false, true>), grid, b, 0, s, a);
        else hipLaunchKernelGGL((pair_mixed_kernel<false, false>), grid, b, 0, s, a);
    }
}

N3Shape n3_shape(int shape)
{
    static const N3Shape shapes[kNumN3Shapes] = {{kBlkTiles, kN3Waves}, {8, 8}, {16, 8}, {4, 4},
                                                 {16, 4}, {8, 2}};
    return shapes[shape < 0 || shape >= kNumN3Shapes ? 0 : shape];
}

int n3_block_rows_shape(int n, int shape)
{
    const int ntiles = (n + kTile - 1) / kTile;
    const int bt = n3_shape(shape).tiles;
    return (ntiles + bt - 1) / bt;
}

namespace {

template <bool kRdf, bool kKelbg>
const void* n3_batch_fn(int shape)
{
    switch (shape) {
        case 1: return (const void*)pair_mixed_n3_batch_kernel<kRdf, kKelbg, 8, 8>;
        case 2: return (const void*)pair_mixed_n3_batch_kernel<kRdf, kKelbg, 16, 8>;
        case 3: return (const void*)pair_mixed_n3_batch_kernel<kRdf, kKelbg, 4, 4>;
        case 4: return (const void*)pair_mixed_n3_batch_kernel<kRdf, kKelbg, 16, 4>;
        case 5: return (const void*)pair_mixed_n3_batch_kernel<kRdf, kKelbg, 8, 2>;
        default: return (const void*)pair_mixed_n3_batch_kernel<kRdf, kKelbg, kBlkTiles, kN3Waves>;
    }
}

const void* n3_batch_kernel(int shape, bool rdf, bool kelbg)
{
    if (rdf) return kelbg ? n3_batch_fn<true, true>(shape) : n3_batch_fn<true, false>(shape);
    return kelbg ? n3_batch_fn<false, true>(shape) : n3_batch_fn<false, false>(shape);
}

}  // namespace

void launch_pair_mixed_n3_batch(const BatchRep* reps, const PairItem* items, int nitems,
                                int* queue, int shape, bool rdf, bool kelbg, int grid,
                                hipStream_t s)
{
    const dim3 g(grid), b(n3_shape(shape).waves * kTile);
    void* args[] = {(void*)&reps, (void*)&items, (void*)&nitems, (void*)&queue};
    (void)hipLaunchKernel(n3_batch_kernel(shape, rdf, kelbg), g, b, args, 0, s);
}

namespace {

template <int kBT>
void launch_f64_batch_bt(const BatchRep* reps, const PairItem* items, int nitems, int* queue,
                         bool rdf, bool kelbg, int grid, hipStream_t s)
{
    const dim3 g(grid), b(f64_waves<kBT>() * kTile);
    if (rdf) {
        if (kelbg) hipLaunchKernelGGL((pair_f64_n3_batch_kernel<true, true, kBT>), g, b, 0, s, reps, items, nitems, queue);
        else hipLaunchKernelGGL((pair_f64_n3_batch_kernel<true, false, kBT>), g, b, 0, s, reps, items, nitems, queue);
    } else {
        if (kelbg) hipLaunchKernelGGL((pair_f64_n3_batch_kernel<false, true, kBT>), g, b, 0, s, reps, items, nitems, queue);
        else hipLaunchKernelGGL((pair_f64_n3_batch_kernel<false, false, kBT>), g, b, 0, s, reps, items, nitems, queue);
    }
}

template <int kBT>
const void* f64_batch_fn(bool rdf, bool kelbg)
{
    return rdf ? (kelbg ? (const void*)pair_f64_n3_batch_kernel<true, true, kBT>
                        : (const void*)pair_f64_n3_batch_kernel<true, false, kBT>)
               : (kelbg ? (const void*)pair_f64_n3_batch_kernel<false, true, kBT>
                        : (const void*)pair_f64_n3_batch_kernel<false, false, kBT>);
}

}  // namespace

void launch_pair_f64_batch(const BatchRep* reps, const PairItem* items, int nitems, int* queue,
                           bool rdf, bool kelbg, int grid, int shape, hipStream_t s)
{
    if (shape == kFineN3Shape) launch_f64_batch_bt<kFineTiles>(reps, items, nitems, queue, rdf, kelbg, grid, s);
    else launch_f64_batch_bt<kF64Tiles>(reps, items, nitems, queue, rdf, kelbg, grid, s);
}

int pair_batch_occupancy(bool mixed, int shape, bool rdf, bool kelbg)
{
    int nb = 0;
    if (mixed) {
        const void* f = n3_batch_kernel(shape, rdf, kelbg);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, n3_shape(shape).waves * kTile, 0) !=
            hipSuccess)
            nb = 1;
    } else {
        const void* f = shape == kFineN3Shape ? f64_batch_fn<kFineTiles>(rdf, kelbg)
                                              : f64_batch_fn<kF64Tiles>(rdf, kelbg);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, f64_threads(shape), 0) != hipSuccess)
            nb = 1;
    }
    return nb < 1 ? 1 : nb;
}

}  // namespace tcp
