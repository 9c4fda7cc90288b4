// Smooth particle-mesh Ewald reciprocal part (Force_SPME4, force.f90:652-929,
// generalised to even B-spline order p) and the direct Ewald k-space sum.
//
// Pipeline per force evaluation, all FP64:
//   spread   particles counting-sorted by the 8^3 brick of their first grid
//            point; one block per brick adds its particles' centred order-p
//            B-spline weights into an LDS tile (brick plus halo) as 64-bit
//            fixed-point integers, and a pass over the covering tiles forms
//            the real grid Q (x fastest, the Fortran order of Q(ix,iy,iz)) in
//            a fixed order.  Grids whose sizes are not multiples of 8 take
//            one thread per particle with FP64 atomic adds into Q.
//   FFT      batched radix-4 (and one radix-2) Stockham transforms in LDS
//            along x, y, z; each block holds 1024 complex points (1024/K
//            lines) and the twiddles.
//            The x pass reads Q as real input and keeps the Hermitian half
//            kx <= K1/2; the z pass transforms, multiplies by gb = PSI |b|^2
//            (writing each block's energy partial gb |S|^2) and transforms
//            back; the inverse x pass restores the other half by conjugation
//            and writes only the real part C.
//   gather   one wavefront per brick of the spreading's binning: the brick's
//            (kBrick + p - 1)^3 points of C go to LDS once and each of its
//            particles reads its p^3 support there (one thread per particle
//            reading C where the grid is not binned); force from the spline
//            derivatives (force.f90:868-919).
// With C the convolution of Q with the real, even kernel IDFT(gb):
//   Um = eps/(2 pi V) sum_m gb(m) |S(m)|^2,  F_i = -eps/(pi V) sum dQ/dx_i C.
#include <algorithm>
#include <vector>

#include "tcp_internal.h"

namespace tcp {

namespace {

constexpr double kPi = 3.1415926535897931;  // data.f90:23

// Centred cardinal B-spline weights: w[i] belongs to grid point
// floor(u) - P/2 + 1 + i and equals M_P(f + P - 1 - i), f = frac(u);
// dw[i] is its derivative with respect to u.  Recursion of the cardinal
// spline: W_k[j] = M_k(f + j) = ((f + j) W_{k-1}[j] + (k - f - j) W_{k-1}[j-1])/(k-1).
template <int P>
__device__ inline void bspline(double f, double* w, double* dw)
{
    double W[P];
    W[0] = f;
    W[1] = 1.0 - f;
    // W holds M_k(f + j) in W[j] after step k; start at k = 2: M_2(f) = f, M_2(f+1) = 1-f
#pragma unroll
    for (int j = 2; j < P; ++j) W[j] = 0.0;
#pragma unroll
    for (int k = 3; k <= P; ++k) {
        if (k == P) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const double prev = j > 0 ? W[j - 1] : 0.0;
                dw[P - 1 - j] = W[j] - prev;
            }
        }
        const double div = 1.0 / (double)(k - 1);
#pragma unroll
        for (int j = k - 1; j >= 0; --j) {
            const double prev = j > 0 ? W[j - 1] : 0.0;
            W[j] = div * ((f + j) * W[j] + ((double)k - f - j) * prev);
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) w[P - 1 - j] = W[j];
}

__device__ inline double charge(const RecipArgs& a, int m)
{
    return a.qp ? a.qp[m] : a.q[m >= a.ne];
}

__device__ inline int wrap(int i, int K)
{
    i %= K;
    return i < 0 ? i + K : i;
}

template <int P>
__global__ __launch_bounds__(256) void spread_kernel(RecipArgs a)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const double q = charge(a, m);
    const double pos[3] = {a.x[m], a.y[m], a.z[m]};
    double w[3][P], dw[3][P];
    int base[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // u = K (x + L/2)/L, force.f90:700-702
        const double u = (double)a.K[d] * (pos[d] + a.box[d] * 0.5) / a.box[d];
        const double fl = floor(u);
        bspline<P>(u - fl, w[d], dw[d]);
        base[d] = (int)fl - P / 2 + 1;
    }
    const int K1 = a.K[0], K2 = a.K[1], K3 = a.K[2];
    for (int k = 0; k < P; ++k) {
        const int iz = wrap(base[2] + k, K3);
        const double wz = q * w[2][k];
        for (int j = 0; j < P; ++j) {
            const int iy = wrap(base[1] + j, K2);
            const double wyz = wz * w[1][j];
            double* row = a.Q + ((size_t)iz * K2 + iy) * K1;
#pragma unroll
            for (int i = 0; i < P; ++i) unsafeAtomicAdd(row + wrap(base[0] + i, K1), wyz * w[0][i]);
        }
    }
}

// Force and potential of particle m from the p^3 grid values C around it
// (force.f90:868-919); cval(ix, iy, iz) returns C at the support's grid point
// base + (ix, iy, iz), so the same sums serve the global grid and a brick's
// LDS copy.
template <int P, typename CVal>
__device__ inline void gather_one(const RecipArgs& a, int m, const double (&w)[3][P],
                                  const double (&dw)[3][P], CVal cval)
{
    double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
    // z planes one at a time (p^2 reads in flight, not p^3); the plane's
    // weights come from the front of wz/dwz, shifted after each plane, so no
    // register array is indexed by a loop variable
    double wz[P], dwz[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        wz[k] = w[2][k];
        dwz[k] = dw[2][k];
    }
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        const double w2 = wz[0], dw2 = dwz[0];
#pragma unroll
        for (int s = 0; s + 1 < P; ++s) {
            wz[s] = wz[s + 1];
            dwz[s] = dwz[s + 1];
        }
#pragma unroll
        for (int j = 0; j < P; ++j) {
            double cx = 0.0, c0 = 0.0;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const double c = cval(i, j, k);
                cx += dw[0][i] * c;
                c0 += w[0][i] * c;
            }
            sx += cx * w[1][j] * w2;
            sy += c0 * dw[1][j] * w2;
            sz += c0 * w[1][j] * dw2;
            sp += c0 * w[1][j] * w2;
        }
    }
    const double q = charge(a, m);
    const double V = a.box[0] * a.box[1] * a.box[2];
    const double pre = -a.eps * q / (kPi * V);
    if (a.pot) a.pot[m] = a.eps / (kPi * V) * sp;  // dUm/dq_m
    a.frx[m] = pre * sx * ((double)a.K[0] / a.box[0]);
    a.fry[m] = pre * sy * ((double)a.K[1] / a.box[1]);
    a.frz[m] = pre * sz * ((double)a.K[2] / a.box[2]);
}

// spline weights of particle m and the first grid point of its support
template <int P>
__device__ inline void spline_of(const RecipArgs& a, int m, double (&w)[3][P], double (&dw)[3][P],
                                 int (&base)[3])
{
    const double pos[3] = {a.x[m], a.y[m], a.z[m]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double u = (double)a.K[d] * (pos[d] + a.box[d] * 0.5) / a.box[d];
        const double fl = floor(u);
        bspline<P>(u - fl, w[d], dw[d]);
        base[d] = (int)fl - P / 2 + 1;
    }
}

// one thread per particle reading the grid (grids that are not brick-binned)
template <int P>
__global__ __launch_bounds__(256) void gather_kernel(RecipArgs a)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    double w[3][P], dw[3][P];
    int base[3];
    spline_of<P>(a, m, w, dw, base);
    const int K1 = a.K[0], K2 = a.K[1], K3 = a.K[2];
    gather_one<P>(a, m, w, dw, [&](int i, int j, int k) {
        return a.C[((size_t)wrap(base[2] + k, K3) * K2 + wrap(base[1] + j, K2)) * K1 +
                   wrap(base[0] + i, K1)];
    });
}

// One wavefront per brick: the (kBrick + p - 1)^3 grid points its particles'
// splines reach are copied to LDS once, and each particle of the brick (the
// binned order of the spreading) reads its support from there at offsets
// inside the tile.
template <int P>
__device__ inline void gather_brick_body(const RecipArgs& a, int bid)
{
    constexpr int E = kBrick + P - 1;
    __shared__ double tile[E * E * E];
    const int p0 = a.brick_start[bid], p1 = a.brick_start[bid + 1];
    if (p0 == p1) return;
    const int borg[3] = {(bid % a.nbrick[0]) * kBrick, ((bid / a.nbrick[0]) % a.nbrick[1]) * kBrick,
                         (bid / (a.nbrick[0] * a.nbrick[1])) * kBrick};
    const int K1 = a.K[0], K2 = a.K[1], K3 = a.K[2];
    // eight loads in flight per thread before their LDS writes
    constexpr int kGroup = 8;
    for (int t0 = 0; t0 < E * E * E; t0 += 64 * kGroup) {
        double c[kGroup];
#pragma unroll
        for (int r = 0; r < kGroup; ++r) {
            const int t = t0 + r * 64 + (int)threadIdx.x;
            c[r] = 0.0;
            if (t < E * E * E) {
                const int lx = t % E, ly = (t / E) % E, lz = t / (E * E);
                int gx = borg[0] + lx, gy = borg[1] + ly, gz = borg[2] + lz;  // < K + E <= 2K
                gx -= gx >= K1 ? K1 : 0;
                gy -= gy >= K2 ? K2 : 0;
                gz -= gz >= K3 ? K3 : 0;
                c[r] = a.C[((size_t)gz * K2 + gy) * K1 + gx];
            }
        }
#pragma unroll
        for (int r = 0; r < kGroup; ++r) {
            const int t = t0 + r * 64 + (int)threadIdx.x;
            if (t < E * E * E) tile[t] = c[r];
        }
    }
    __syncthreads();
    for (int t = p0 + (int)threadIdx.x; t < p1; t += blockDim.x) {
        const int m = a.sorted[t];
        double w[3][P], dw[3][P];
        int base[3];
        spline_of<P>(a, m, w, dw, base);
        // offset of the support inside the tile, in [0, kBrick) per axis
        const int o0 = wrap(base[0], K1) - borg[0], o1 = wrap(base[1], K2) - borg[1],
                  o2 = wrap(base[2], K3) - borg[2];
        const double* tb = tile + (o2 * E + o1) * E + o0;
        gather_one<P>(a, m, w, dw, [&](int i, int j, int k) { return tb[(k * E + j) * E + i]; });
    }
}

template <int P>
__global__ __launch_bounds__(64) void gather_brick_kernel(RecipArgs a)
{
    gather_brick_body<P>(a, blockIdx.x);
}

// ------------------------------------------------- brick-binned spreading ---
//
// Device-scope FP64 atomics resolve beyond the per-XCD L2, so spreading p^3
// atomics per particle straight into the grid moves ~32 B of memory traffic
// per update.  Instead particles are counting-sorted by the brick (kBrick^3
// grid points) that holds their first grid point; one block per brick spreads
// its particles into an LDS tile of (kBrick + p - 1)^3 points (the brick plus
// the halo its splines reach), writes the tile out, and a gather over the
// tiles that cover each grid point (up to 3 per axis) forms Q in a fixed
// order.  Inside a brick the contributions are added as 64-bit fixed-point
// integers scaled by 2^k with k chosen per brick so that (particles in the
// brick) x max|q| stays below 2^62: integer addition is associative, so the
// tile and hence Q do not depend on the order in which threads add, and the
// resolution 2^-k (2^-56 for 32 unit charges) is finer than FP64 rounding of
// the tile values.

__device__ inline int brick_axis(double u, int P, int K)
{
    const int b0 = wrap((int)floor(u) - P / 2 + 1, K);
    return b0 / kBrick;
}

template <int P>
__device__ inline void bin_count_body(const RecipArgs& a, int m)
{
    if (m >= a.n) return;
    const double pos[3] = {a.x[m], a.y[m], a.z[m]};
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double u = (double)a.K[d] * (pos[d] + a.box[d] * 0.5) / a.box[d];
        b[d] = brick_axis(u, P, a.K[d]);
    }
    const int id = (b[2] * a.nbrick[1] + b[1]) * a.nbrick[0] + b[0];
    a.brick_of[m] = id;
    atomicAdd(a.brick_count + id, 1);
}

template <int P>
__global__ __launch_bounds__(256) void bin_count_kernel(RecipArgs a)
{
    bin_count_body<P>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// exclusive scan of the brick counts (one block), cursors reset
__device__ inline void bin_scan_body(const RecipArgs& a, int nb)
{
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = min(nb, lo + per);
    int s = 0;
    for (int b = lo; b < hi; ++b) s += a.brick_count[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int b = lo; b < hi; ++b) {
        a.brick_start[b] = run;
        run += a.brick_count[b];
        a.brick_cursor[b] = 0;
    }
    if (threadIdx.x == 1023) a.brick_start[nb] = part[1023];
}

__global__ __launch_bounds__(1024) void bin_scan_kernel(RecipArgs a, int nb)
{
    bin_scan_body(a, nb);
}

__device__ inline void bin_scatter_body(const RecipArgs& a, int m)
{
    if (m >= a.n) return;
    const int id = a.brick_of[m];
    a.sorted[a.brick_start[id] + atomicAdd(a.brick_cursor + id, 1)] = m;
}

__global__ __launch_bounds__(256) void bin_scatter_kernel(RecipArgs a)
{
    bin_scatter_body(a, blockIdx.x * blockDim.x + threadIdx.x);
}

template <int P>
__device__ inline void spread_brick_body(const RecipArgs& a, int bid)
{
    constexpr int E = kBrick + P - 1;      // tile edge
    constexpr int kBatch = 64;             // particles staged per pass
    __shared__ unsigned long long tile[E * E * E];
    __shared__ double ws[kBatch][3][P];
    __shared__ int org[kBatch][3];
    __shared__ double qred[256];
    const int bx = bid % a.nbrick[0], by = (bid / a.nbrick[0]) % a.nbrick[1],
              bz = bid / (a.nbrick[0] * a.nbrick[1]);
    for (int t = threadIdx.x; t < E * E * E; t += blockDim.x) tile[t] = 0ull;
    const int p0 = a.brick_start[bid], p1 = a.brick_start[bid + 1];
    const int borg[3] = {bx * kBrick, by * kBrick, bz * kBrick};
    // fixed-point scale: (p1 - p0) max|q| < 2^(62 - k)
    double qm = 0.0;
    for (int t = p0 + threadIdx.x; t < p1; t += blockDim.x) qm = fmax(qm, fabs(charge(a, a.sorted[t])));
    qred[threadIdx.x] = qm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) qred[threadIdx.x] = fmax(qred[threadIdx.x], qred[threadIdx.x + w]);
        __syncthreads();
    }
    int e = 0;
    frexp(qred[0] * (double)(p1 - p0), &e);
    const int k = qred[0] > 0.0 ? min(1000, max(-1000, 62 - e)) : 0;
    for (int q0 = p0; q0 < p1; q0 += kBatch) {
        const int nq = min(kBatch, p1 - q0);
        __syncthreads();
        if ((int)threadIdx.x < nq) {
            const int m = a.sorted[q0 + threadIdx.x];
            const double q = charge(a, m);
            const double pos[3] = {a.x[m], a.y[m], a.z[m]};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double u = (double)a.K[d] * (pos[d] + a.box[d] * 0.5) / a.box[d];
                const double fl = floor(u);
                double w[P], dw[P];
                bspline<P>(u - fl, w, dw);
                // offset of the first grid point inside the tile, in [0, kBrick)
                org[threadIdx.x][d] = wrap((int)fl - P / 2 + 1, a.K[d]) - borg[d];
#pragma unroll
                for (int i = 0; i < P; ++i) ws[threadIdx.x][d][i] = d == 0 ? q * w[i] : w[i];
            }
        }
        __syncthreads();
        // one work item = one particle and one (j, k) row of its support
        for (int it = threadIdx.x; it < nq * P * P; it += blockDim.x) {
            const int qq = it / (P * P), jk = it % (P * P), j = jk % P, kk = jk / P;
            const double wyz = ws[qq][1][j] * ws[qq][2][kk];
            unsigned long long* row = tile + ((org[qq][2] + kk) * E + org[qq][1] + j) * E + org[qq][0];
#pragma unroll
            for (int i = 0; i < P; ++i)
                atomicAdd(row + i, (unsigned long long)llrint(ldexp(ws[qq][0][i] * wyz, k)));
        }
    }
    __syncthreads();
    double* out = a.tiles + (size_t)bid * (E * E * E);
    for (int t = threadIdx.x; t < E * E * E; t += blockDim.x)
        out[t] = ldexp((double)(long long)tile[t], -k);
}

template <int P>
__global__ __launch_bounds__(256) void spread_brick_kernel(RecipArgs a)
{
    spread_brick_body<P>(a, blockIdx.x);
}

// Q(g) = sum over the tiles covering g, in a fixed order (x fastest grid index)
template <int P>
__device__ inline void grid_from_tiles_body(const RecipArgs& a, size_t g)
{
    constexpr int E = kBrick + P - 1;
    const size_t G = (size_t)a.K[0] * a.K[1] * a.K[2];
    if (g >= G) return;
    const int gx = (int)(g % a.K[0]), gy = (int)((g / a.K[0]) % a.K[1]),
              gz = (int)(g / ((size_t)a.K[0] * a.K[1]));
    const int gg[3] = {gx, gy, gz};
    // tile (b', l) covers g when b' kBrick + l = g (mod K) and l < E: b' = b - c
    // for c = 0, 1, .. as long as o + c kBrick < E (up to 3 for order 12)
    int nb[3], bs[3][3], ls[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int b = gg[d] / kBrick, o = gg[d] % kBrick;
        nb[d] = 0;
        for (int c = 0; c < 3 && o + c * kBrick < E; ++c) {
            bs[d][c] = ((b - c) % a.nbrick[d] + a.nbrick[d]) % a.nbrick[d];
            ls[d][c] = o + c * kBrick;
            nb[d] = c + 1;
        }
    }
    double s = 0.0;
    for (int cz = 0; cz < nb[2]; ++cz)
        for (int cy = 0; cy < nb[1]; ++cy)
            for (int cx = 0; cx < nb[0]; ++cx) {
                const int bid = (bs[2][cz] * a.nbrick[1] + bs[1][cy]) * a.nbrick[0] + bs[0][cx];
                s += a.tiles[(size_t)bid * (E * E * E) + (ls[2][cz] * E + ls[1][cy]) * E + ls[0][cx]];
            }
    a.Q[g] = s;
}

template <int P>
__global__ __launch_bounds__(256) void grid_from_tiles_kernel(RecipArgs a)
{
    grid_from_tiles_body<P>(a, blockIdx.x * (size_t)blockDim.x + threadIdx.x);
}

// -------------------------------------------------------------------- FFT ---
//
// Q is real, so its transform is Hermitian, S(-m) = conj S(m), and gb is real
// and even.  The forward x pass keeps kx = 0..K1/2 only (H = K1/2 + 1 columns
// of the complex work grid A, x fastest), the y and z passes run on that half,
// and the inverse x pass restores kx > K1/2 by conjugation before writing the
// real part.  The z pass transforms a line, multiplies it by gb, adds
// c(kx) gb |S|^2 to the energy partial of its block (c = 1 for kx = 0 and
// K1/2, 2 for the columns that stand for their mirror too) and transforms it
// back, so the spectrum makes one round trip through memory less.

constexpr int kFftTile = 1024;   // complex points per block
constexpr int kFftThreads = 256;

enum { kFftFwdX = 0, kFftY = 1, kFftZConv = 2, kFftInvX = 3 };

// line l starts at (l / inner) outer + (l % inner) lstride; its point k is
// estride further on
struct FftAddr {
    int inner, outer, lstride, estride;
};

struct FftPass {
    const double* in_real;
    const double2* in;
    double2* out;
    double* out_real;
    const double2* tw;   // exp(+2 pi i m / K), m < K/2
    const double* gb;
    double* epart;
    int K, logK, lines, lines_per_block;
    int sign;            // -1 forward, +1 inverse (the y pass)
    FftAddr ia, oa;
    int H, K1, K2;       // half width and grid sizes for the z pass's gb and weights
};

__device__ inline size_t fft_base(const FftAddr& d, int l)
{
    return (size_t)(l / d.inner) * d.outer + (size_t)(l % d.inner) * d.lstride;
}

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// exp(-+2 pi i m / K) for 0 <= m < K from the half table exp(+2 pi i m / K),
// m < K/2 (the second half is its negative)
__device__ inline double2 twiddle(const double2* tws, int m, int half, bool inverse)
{
    double2 w = m < half ? tws[m] : make_double2(-tws[m - half].x, -tws[m - half].y);
    if (!inverse) w.y = -w.y;
    return w;
}

// Mixed-radix Stockham stages over the block's lines, ping-pong in buf (a
// radix-2 stage first when log2 K is odd, then radix-4 stages: four
// barriers for K = 128); src names the buffer that holds the result.
// Stage with span Ns and radix R: butterfly j of a line reads points
// j + r K/R, multiplies point r by w^(r (j mod Ns) K/(R Ns)), transforms the R
// values and writes them to (j - j mod Ns) R + j mod Ns + r Ns.
__device__ inline void fft_stages(double2 (*buf)[kFftTile], int& src, int K, int logK, int npts,
                                  const double2* tws, bool inverse)
{
    const int half = K >> 1, quarter = K >> 2;
    int Ns = 1, lNs = 0;
    if (logK & 1) {
        for (int b = threadIdx.x; b < (npts >> 1); b += kFftThreads) {
            const int lb = b >> (logK - 1), j = b & (half - 1);
            const double2* A = buf[src] + lb * K;
            double2* Bo = buf[src ^ 1] + lb * K;
            const double2 a0 = A[j], a1 = A[j + half];
            Bo[2 * j] = cadd(a0, a1);
            Bo[2 * j + 1] = csub(a0, a1);
        }
        src ^= 1;
        __syncthreads();
        Ns = 2;
        lNs = 1;
    }
    for (; Ns < K; Ns <<= 2, lNs += 2) {
        for (int b = threadIdx.x; b < (npts >> 2); b += kFftThreads) {
            const int lb = b >> (logK - 2), j = b & (quarter - 1);
            const int k = j & (Ns - 1);
            const int m = k << (logK - 2 - lNs);  // k K / (4 Ns)
            const double2* A = buf[src] + lb * K;
            double2* Bo = buf[src ^ 1] + lb * K;
            const double2 v0 = A[j];
            const double2 v1 = cmul(A[j + quarter], twiddle(tws, m, half, inverse));
            const double2 v2 = cmul(A[j + 2 * quarter], twiddle(tws, 2 * m, half, inverse));
            const double2 v3 = cmul(A[j + 3 * quarter], twiddle(tws, 3 * m, half, inverse));
            const double2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), d = csub(v1, v3);
            // times -i (forward) or +i (inverse)
            const double2 t3 = inverse ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
            const int o = ((j - k) << 2) + k;
            Bo[o] = cadd(t0, t2);
            Bo[o + Ns] = cadd(t1, t3);
            Bo[o + 2 * Ns] = csub(t0, t2);
            Bo[o + 3 * Ns] = csub(t1, t3);
        }
        src ^= 1;
        __syncthreads();
    }
}

template <int kMode>
__device__ inline void fft_pass_body(const FftPass& f, int blk)
{
    __shared__ double2 buf[2][kFftTile];
    __shared__ double2 tws[kFftTile / 2];
    const int K = f.K, B = f.lines_per_block, half = K >> 1;
    const int l0 = blk * B;
    const int npts = B * K;
    // x lines: the block's lines are one contiguous run (of H or K1 points each)
    constexpr bool kX = kMode == kFftFwdX || kMode == kFftInvX;
    for (int t = threadIdx.x; t < half; t += kFftThreads) tws[t] = f.tw[t];
    // all of a thread's loads in flight before its LDS writes
    constexpr int kPer = kFftTile / kFftThreads;
    const int nin = kMode == kFftInvX ? f.H : K;
    double2 v[kPer];
    int slot[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        v[r] = make_double2(0.0, 0.0);
        slot[r] = -1;
        if (e < B * nin) {
            int lb, k;
            if (kX) { lb = e / nin; k = e % nin; }
            else { k = e / B; lb = e % B; }
            const int l = l0 + lb;
            slot[r] = lb * K + k;
            if (l < f.lines) {
                const size_t g = fft_base(f.ia, l) + (size_t)k * f.ia.estride;
                if (kMode == kFftFwdX) v[r] = make_double2(f.in_real[g], 0.0);
                else v[r] = f.in[g];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r)
        if (slot[r] >= 0) buf[0][slot[r]] = v[r];
    __syncthreads();
    if (kMode == kFftInvX) {  // X(K1 - k) = conj X(k), k = 1 .. K1/2 - 1
        for (int e = threadIdx.x; e < B * (half - 1); e += kFftThreads) {
            const int lb = e / (half - 1), k = 1 + e % (half - 1);
            buf[0][lb * K + K - k] = conjd(buf[0][lb * K + k]);
        }
        __syncthreads();
    }
    int src = 0;
    fft_stages(buf, src, K, f.logK, npts, tws, kMode == kFftInvX || (kMode == kFftY && f.sign > 0));
    if (kMode == kFftZConv) {
        double esum = 0.0;
        for (int e = threadIdx.x; e < npts; e += kFftThreads) {
            const int lb = e / K, k = e % K, l = l0 + lb;
            if (l >= f.lines) continue;
            const int kx = l % f.H, ky = l / f.H;
            const double gbv = f.gb[((size_t)k * f.K2 + ky) * f.K1 + kx];
            const double2 v = buf[src][e];
            const double c = (kx == 0 || 2 * kx == f.K1) ? 1.0 : 2.0;
            esum += c * gbv * (v.x * v.x + v.y * v.y);
            buf[src][e] = make_double2(v.x * gbv, v.y * gbv);
        }
        // the block's energy partial, summed in a fixed order over the idle buffer
        double* red = (double*)buf[src ^ 1];
        red[threadIdx.x] = esum;
        __syncthreads();
        for (int w = kFftThreads / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) f.epart[blk] = red[0];
        __syncthreads();
        fft_stages(buf, src, K, f.logK, npts, tws, true);
    }
    const int nout = kMode == kFftFwdX ? f.H : K;
    for (int e = threadIdx.x; e < B * nout; e += kFftThreads) {
        int lb, k;
        if (kX) { lb = e / nout; k = e % nout; }
        else { k = e / B; lb = e % B; }
        const int l = l0 + lb;
        if (l >= f.lines) continue;
        const size_t g = fft_base(f.oa, l) + (size_t)k * f.oa.estride;
        const double2 v = buf[src][lb * K + k];
        if (kMode == kFftInvX) f.out_real[g] = v.x;
        else f.out[g] = v;
    }
}

template <int kMode>
__global__ __launch_bounds__(kFftThreads) void fft_pass_kernel(FftPass f)
{
    fft_pass_body<kMode>(f, blockIdx.x);
}

int ilog2(int k)
{
    int l = 0;
    while ((1 << l) < k) ++l;
    return l;
}

// The four passes over K = (K1, K2, K3); A holds H = K1/2 + 1 columns.
FftPass make_pass(const RecipArgs& a, int mode, int sign)
{
    FftPass f{};
    const int K1 = a.K[0], K2 = a.K[1], K3 = a.K[2], H = K1 / 2 + 1;
    const int axis = mode == kFftY ? 1 : (mode == kFftZConv ? 2 : 0);
    f.K = a.K[axis];
    f.logK = ilog2(f.K);
    f.lines_per_block = kFftTile / f.K;
    f.tw = a.tw[axis];
    f.sign = sign;
    f.gb = a.gb;
    f.epart = a.epart;
    f.H = H;
    f.K1 = K1;
    f.K2 = K2;
    if (axis == 0) {
        f.lines = K2 * K3;
        const FftAddr full{1, K1, 0, 1}, halfx{1, H, 0, 1};
        f.ia = mode == kFftFwdX ? full : halfx;
        f.oa = mode == kFftFwdX ? halfx : full;
        f.in_real = a.Q;
        f.in = a.A;
        f.out = a.A;
        f.out_real = a.C;
    } else if (axis == 1) {
        f.lines = H * K3;
        f.ia = f.oa = FftAddr{H, K2 * H, 1, H};
        f.in = a.A;
        f.out = a.A;
    } else {
        f.lines = H * K2;
        f.ia = f.oa = FftAddr{H * K2, 0, 1, H * K2};
        f.in = a.A;
        f.out = a.A;
    }
    return f;
}

template <int kMode>
void run_pass(const FftPass& f, hipStream_t s)
{
    const int blocks = (f.lines + f.lines_per_block - 1) / f.lines_per_block;
    hipLaunchKernelGGL(fft_pass_kernel<kMode>, dim3(blocks), dim3(kFftThreads), 0, s, f);
}

__global__ void fill_int_kernel(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void fill_kernel(double* p, size_t n, double v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------ direct k-space sum ---
//
// Only the half space of m is summed (linear index of the (2 kmax + 1)^3 cube
// above its centre, i.e. mz > 0, or mz = 0 and my > 0, or mz = my = 0 and
// mx > 0): S(-m) = conj S(m), so both the energy and the forces are twice the
// half-space sums.  Phases e^{2 pi i m x/L} come from one sincos per particle
// and axis and complex-multiplication recurrences.
//
// S(m): blocks of `chunk` particles write per-chunk partials of every half-space
// m (phase tables of the chunk in LDS, one thread per m), then one thread per
// m sums the chunks in order, so S is independent of scheduling.

__device__ inline double2 cexpi(double th)
{
    double sn, cs;
    sincos(th, &sn, &cs);
    return make_double2(cs, sn);
}

__device__ inline void decode_m(int mi, int kmax, int* mx, int* my, int* mz)
{
    const int nk = 2 * kmax + 1;
    *mx = mi % nk - kmax;
    *my = (mi / nk) % nk - kmax;
    *mz = mi / (nk * nk) - kmax;
}

// half-space rows (my, mz) of the row form, and how many groups of that many
// threads a 256-thread block holds (1 where the row form is not used)
__host__ __device__ inline int ewald_sk_rows(int kmax) { return 1 + kmax + kmax * (2 * kmax + 1); }
__host__ __device__ inline int ewald_sk_groups(int kmax)
{
    return kmax <= 8 ? 256 / ewald_sk_rows(kmax) : 1;
}

// ph: LDS table [3][kmax + 1][chunk] of q^(d == 0) e^{i m theta_d}, m = 0..kmax
__device__ inline void ewald_sk_part_body(const RecipArgs& a, int kmax, int chunk, int c,
                                          double2* ph)
{
    const int K1 = kmax + 1;
    for (int t = threadIdx.x; t < 3 * chunk; t += blockDim.x) {
        const int d = t / chunk, p = t % chunk, j = c * chunk + p;
        const double* xd = d == 0 ? a.x : (d == 1 ? a.y : a.z);
        double2 e = make_double2(j < a.n ? (d == 0 ? charge(a, j) : 1.0) : 0.0, 0.0);
        const double2 u = cexpi(2.0 * kPi * (j < a.n ? xd[j] : 0.0) / a.box[d]);
        for (int m = 0; m < K1; ++m) {
            ph[(d * K1 + m) * chunk + p] = e;
            e = cmul(e, u);
        }
    }
    __syncthreads();
    const int nk = 2 * kmax + 1, nh = (nk * nk * nk - 1) / 2;
    for (int h = threadIdx.x; h < nh; h += blockDim.x) {
        int mx, my, mz;
        decode_m(nh + 1 + h, kmax, &mx, &my, &mz);
        const double sx = mx < 0 ? -1.0 : 1.0, sy = my < 0 ? -1.0 : 1.0;
        const double2* X = ph + (0 * K1 + abs(mx)) * chunk;
        const double2* Y = ph + (1 * K1 + abs(my)) * chunk;
        const double2* Z = ph + (2 * K1 + mz) * chunk;
        double sr = 0.0, si = 0.0;
        for (int p = 0; p < chunk; ++p) {
            const double2 x = make_double2(X[p].x, sx * X[p].y);
            const double2 y = make_double2(Y[p].x, sy * Y[p].y);
            const double2 e = cmul(cmul(x, y), Z[p]);
            sr += e.x;
            si += e.y;
        }
        a.skpart[(size_t)c * nh + h] = make_double2(sr, si);
    }
}

// The same partials by rows: a row is the line (my, mz), walked along
// mx = -kmax..kmax (kLen = 2 kmax + 1 values, a compile-time count, so the
// walk has no per-m test) by the recurrence e <- e e^{i theta_x}; a particle
// costs a row two complex products to start and one per m, and the row's S(m)
// partials stay in registers.  Row 0 is (my, mz) = (0, 0), whose half-space
// part mx = 1..kmax is the only part stored; rows 1..kmax are mz = 0,
// my = 1..kmax; then mz = 1..kmax with my = -kmax..kmax.  The block's threads form ewald_sk_groups(kmax) groups of
// `rows` threads; group g walks its own contiguous share of the chunk's
// particles, and the groups' partials are joined in LDS in group order, so the
// sum is fixed by kmax and chunk.  ux: e^{i theta_x} of the chunk.
template <int kLen>
__device__ inline void ewald_sk_rows_body(const RecipArgs& a, int kmax, int chunk, int c,
                                          double2* ph, double2* ux)
{
    const int K1 = kmax + 1, nk = 2 * kmax + 1, nh = (nk * nk * nk - 1) / 2;
    for (int t = threadIdx.x; t < 3 * chunk; t += blockDim.x) {
        const int d = t / chunk, p = t % chunk, j = c * chunk + p;
        const double* xd = d == 0 ? a.x : (d == 1 ? a.y : a.z);
        double2 e = make_double2(j < a.n ? (d == 0 ? charge(a, j) : 1.0) : 0.0, 0.0);
        const double2 u = cexpi(2.0 * kPi * (j < a.n ? xd[j] : 0.0) / a.box[d]);
        if (d == 0) ux[p] = u;
        for (int m = 0; m < K1; ++m) {
            ph[(d * K1 + m) * chunk + p] = e;
            e = cmul(e, u);
        }
    }
    __syncthreads();
    const int rows = ewald_sk_rows(kmax), groups = ewald_sk_groups(kmax);
    const int g = threadIdx.x / rows, t = threadIdx.x - g * rows;
    const bool live = g < groups;
    int my, mz;
    if (t <= kmax) {
        mz = 0;
        my = t;
    } else {
        mz = 1 + (t - kmax - 1) / nk;
        my = (t - kmax - 1) % nk - kmax;
    }
    // every row walks mx = -kmax .. kmax, so the mx loop has one length for all
    // threads; of row (0, 0) only mx > 0 lies in the half space and is stored
    double2 acc[kLen];
#pragma unroll
    for (int k = 0; k < kLen; ++k) acc[k] = make_double2(0.0, 0.0);
    if (live) {
        const double2* Y = ph + (1 * K1 + abs(my)) * chunk;
        const double2* Z = ph + (2 * K1 + mz) * chunk;
        const double2* X = ph + (0 * K1 + kmax) * chunk;  // q u^kmax, conjugated: q u^-kmax
        const int p1 = (g + 1) * chunk / groups;
        for (int p = g * chunk / groups; p < p1; ++p) {
            const double2 y = my < 0 ? conjd(Y[p]) : Y[p];
            const double2 x = conjd(X[p]);
            double2 e = cmul(cmul(x, y), Z[p]);
            const double2 u = ux[p];
#pragma unroll
            for (int k = 0; k < kLen; ++k) {
                acc[k].x += e.x;
                acc[k].y += e.y;
                e = cmul(e, u);
            }
        }
    }
    const int h0 = ((mz + kmax) * nk + (my + kmax)) * nk - (nh + 1);  // mx = -kmax
    if (groups > 1) {
        // groups 1.. park their partials over the phase table, group 0 adds them in order
        __syncthreads();
        if (live && g > 0) {
#pragma unroll
            for (int k = 0; k < kLen; ++k)
                if (h0 + k >= 0) ph[(size_t)(g - 1) * nh + h0 + k] = acc[k];
        }
        __syncthreads();
        if (g == 0) {
            for (int o = 0; o < groups - 1; ++o) {
#pragma unroll
                for (int k = 0; k < kLen; ++k)
                    if (h0 + k >= 0) {
                        const double2 v = ph[(size_t)o * nh + h0 + k];
                        acc[k].x += v.x;
                        acc[k].y += v.y;
                    }
            }
        }
    }
    if (g == 0) {
#pragma unroll
        for (int k = 0; k < kLen; ++k)
            if (h0 + k >= 0) a.skpart[(size_t)c * nh + h0 + k] = acc[k];
    }
}

// S(m) and G(m) = exp(-pi^2 m^2/alpha^2)/m^2 of every half-space m into sk[3 mi ..]
__device__ inline void ewald_sk_reduce_body(const RecipArgs& a, int kmax, double alpha, int nchunk,
                                            double* sk, int h)
{
    const int nk = 2 * kmax + 1, nh = (nk * nk * nk - 1) / 2;
    if (h >= nh) return;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        const double2 v = a.skpart[(size_t)c * nh + h];
        sr += v.x;
        si += v.y;
    }
    const int mi = nh + 1 + h;
    int mx, my, mz;
    decode_m(mi, kmax, &mx, &my, &mz);
    const double kx = mx / a.box[0], ky = my / a.box[1], kz = mz / a.box[2];
    const double m2 = kx * kx + ky * ky + kz * kz;
    const double G = exp(-kPi * kPi * m2 / (alpha * alpha)) / m2;
    // G S (the force and potential sums take S only with G) and G |S|^2
    sk[3 * mi + 0] = G * sr;
    sk[3 * mi + 1] = G * si;
    sk[3 * mi + 2] = G * (sr * sr + si * si);
}

// Forces F_j = (2 eps q_j / V) sum_m G m/L Im(conj(S) e^{2 pi i m.r_j}) (all m),
// potential dUm/dq_j = eps/(pi V) sum_m G Re(conj(S) e^{2 pi i m.r_j}), and in
// block 0 the energy sum_m G |S|^2 into epart[0]; all as twice the half space.
// A block holds 64 particles and kKsumWaves waves; wave w walks its contiguous
// share of the half space for all 64 (phases by recurrences from the share's
// first m), and the shares' sums join in wave order, so the result does not
// depend on scheduling.  Every lane reads the same m at the same time, so the
// half-space table is staged in (dynamic) LDS when it fits (kStaged) and read
// from memory otherwise; the same LDS then carries the reductions.
constexpr int kSkLds = 2048;   // half-space entries staged in LDS (kmax <= 7)
constexpr int kKsumWaves = 8;  // shares of the half space per particle

// u^k for an integer k (|k| small): repeated products, u^-k = conj(u)^k on the unit circle
__device__ inline double2 cpowi(double2 u, int k)
{
    const double2 b = k < 0 ? conjd(u) : u;
    double2 e = make_double2(1.0, 0.0);
    for (int t = k < 0 ? -k : k; t > 0; --t) e = cmul(e, b);
    return e;
}

// Dynamic LDS of ewald_force_body, in doubles: the staged G S pairs, then
// reused for the reductions.
__host__ __device__ inline int ewald_force_lds(int kmax, bool staged)
{
    const int nk = 2 * kmax + 1, nh = (nk * nk * nk - 1) / 2;
    const int red = kKsumWaves * 4 * 64;
    return staged && 2 * nh > red ? 2 * nh : red;
}

template <bool kStaged>
__device__ inline void ewald_force_body(const RecipArgs& a, int kmax, const double* skg, int blk,
                                        double* lds)
{
    const int lane = threadIdx.x % 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int j = blk * 64 + lane;
    const int nk = 2 * kmax + 1, nh = (nk * nk * nk - 1) / 2;
    if (kStaged) {
        // G S of half-space entry mi at lds[2 (mi - nh - 1) ..]: one 16-byte read
        for (int t = threadIdx.x; t < 2 * nh; t += blockDim.x)
            lds[t] = skg[3 * (nh + 1) + 3 * (t >> 1) + (t & 1)];
        __syncthreads();
    }
    // half-space entry h = mi - nh - 1: G Re S and G Im S at sk[kS h], sk[kS h + 1]
    constexpr int kS = kStaged ? 2 : 3;
    const double* sk = kStaged ? lds : skg + 3 * (nh + 1);
    const int h0 = wave * nh / kKsumWaves, h1 = (wave + 1) * nh / kKsumWaves;
    double fx = 0.0, fy = 0.0, fz = 0.0, pt = 0.0;
    if (j < a.n && h0 < h1) {
        const double2 ux = cexpi(2.0 * kPi * a.x[j] / a.box[0]);
        const double2 uy = cexpi(2.0 * kPi * a.y[j] / a.box[1]);
        const double2 uz = cexpi(2.0 * kPi * a.z[j] / a.box[2]);
        const double2 xk = cpowi(ux, -kmax), yk = cpowi(uy, -kmax);  // row starts
        int mx, my, mz;
        decode_m(nh + 1 + h0, kmax, &mx, &my, &mz);
        double2 ez = cpowi(uz, mz);
        double2 eyz = cmul(cpowi(uy, my), ez);
        double2 e = cmul(eyz, cpowi(ux, mx));
        int h = h0;
        for (;;) {
            // the share's part of row (my, mz): mx up to kmax, branch-free
            const int hend = min(h1, h + kmax - mx + 1);
            double gx = 0.0, gi = 0.0, mxd = (double)mx;
#pragma unroll 2
            for (; h < hend; ++h) {
                const double A = sk[kS * h], B = sk[kS * h + 1];  // G S
                const double im = fma(A, e.y, -B * e.x);           // G Im(conj(S) e)
                gx = fma(im, mxd, gx);
                gi += im;
                pt = fma(A, e.x, fma(B, e.y, pt));                 // G Re(conj(S) e)
                e = cmul(e, ux);
                mxd += 1.0;
            }
            fx += gx;
            fy += gi * (double)my;
            fz += gi * (double)mz;
            if (h >= h1) break;
            // the next row's first m
            mx = -kmax;
            if (my < kmax) {
                ++my;
                eyz = cmul(eyz, uy);
            } else {
                my = -kmax;
                ++mz;
                ez = cmul(ez, uz);
                eyz = cmul(yk, ez);
            }
            e = cmul(eyz, xk);
        }
    }
    __syncthreads();  // every wave is done with the staged table
    double* red = lds;  // [kKsumWaves][4][64]
    red[(wave * 4 + 0) * 64 + lane] = fx;
    red[(wave * 4 + 1) * 64 + lane] = fy;
    red[(wave * 4 + 2) * 64 + lane] = fz;
    red[(wave * 4 + 3) * 64 + lane] = pt;
    __syncthreads();
    if (wave == 0 && j < a.n) {
        for (int w = 1; w < kKsumWaves; ++w) {
            fx += red[(w * 4 + 0) * 64 + lane];
            fy += red[(w * 4 + 1) * 64 + lane];
            fz += red[(w * 4 + 2) * 64 + lane];
            pt += red[(w * 4 + 3) * 64 + lane];
        }
        const double V = a.box[0] * a.box[1] * a.box[2];
        if (a.pot) a.pot[j] = 2.0 * a.eps / (kPi * V) * pt;
        const double pre = 4.0 * a.eps * charge(a, j) / V;
        a.frx[j] = pre * fx / a.box[0];
        a.fry[j] = pre * fy / a.box[1];
        a.frz[j] = pre * fz / a.box[2];
    }
    if (blk == 0) {
        __syncthreads();  // wave 0 has read the force shares
        double* esum = lds;  // [kKsumWaves * 64]
        double en = 0.0;
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            const int mi = nh + 1 + h;
            en += skg[3 * mi + 2];  // G |S|^2
        }
        esum[threadIdx.x] = en;
        __syncthreads();
        for (int w = kKsumWaves * 32; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) esum[threadIdx.x] += esum[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.epart[0] = 2.0 * esum[0];
    }
}

// the row form for kmax <= 8 (2 kmax + 1 <= 17 accumulators), per m otherwise
__device__ inline void ewald_sk_part_any(const RecipArgs& a, int kmax, int chunk, int c,
                                         double2* dyn)
{
    double2* ux = dyn + 3 * (kmax + 1) * chunk;
    switch (kmax) {  // the row length 2 kmax + 1 as a constant
        case 1: ewald_sk_rows_body<3>(a, kmax, chunk, c, dyn, ux); break;
        case 2: ewald_sk_rows_body<5>(a, kmax, chunk, c, dyn, ux); break;
        case 3: ewald_sk_rows_body<7>(a, kmax, chunk, c, dyn, ux); break;
        case 4: ewald_sk_rows_body<9>(a, kmax, chunk, c, dyn, ux); break;
        case 5: ewald_sk_rows_body<11>(a, kmax, chunk, c, dyn, ux); break;
        case 6: ewald_sk_rows_body<13>(a, kmax, chunk, c, dyn, ux); break;
        case 7: ewald_sk_rows_body<15>(a, kmax, chunk, c, dyn, ux); break;
        case 8: ewald_sk_rows_body<17>(a, kmax, chunk, c, dyn, ux); break;
        default: ewald_sk_part_body(a, kmax, chunk, c, dyn); break;
    }
}

__global__ __launch_bounds__(256) void ewald_sk_part_kernel(RecipArgs a, int kmax, int chunk)
{
    extern __shared__ double2 ph_dyn[];
    ewald_sk_part_any(a, kmax, chunk, blockIdx.x, ph_dyn);
}

__global__ __launch_bounds__(256) void ewald_sk_reduce_kernel(RecipArgs a, int kmax, double alpha,
                                                              int nchunk, double* sk)
{
    ewald_sk_reduce_body(a, kmax, alpha, nchunk, sk, blockIdx.x * 256 + threadIdx.x);
}

template <bool kStaged>
__global__ __launch_bounds__(kKsumWaves * 64) void ewald_force_kernel(RecipArgs a, int kmax,
                                                                      const double* sk)
{
    extern __shared__ double kf_dyn[];
    ewald_force_body<kStaged>(a, kmax, sk, blockIdx.x, kf_dyn);
}

// batched: blockIdx.y selects the replica (see kernels_md.hip)
__global__ __launch_bounds__(256) void ewald_sk_part_batch_kernel(const BatchRep* reps, int chunk)
{
    extern __shared__ double2 ph_dyn[];
    const BatchRep& r = reps[blockIdx.y];
    if (r.recip != TCP_RECIP_EWALD || (int)blockIdx.x >= (r.ra.n + chunk - 1) / chunk) return;
    ewald_sk_part_any(r.ra, r.kmax, chunk, blockIdx.x, ph_dyn);
}

__global__ __launch_bounds__(256) void ewald_sk_reduce_batch_kernel(const BatchRep* reps, int chunk)
{
    const BatchRep& r = reps[blockIdx.y];
    if (r.recip != TCP_RECIP_EWALD) return;
    ewald_sk_reduce_body(r.ra, r.kmax, r.alpha, (r.ra.n + chunk - 1) / chunk, r.sk,
                         blockIdx.x * 256 + threadIdx.x);
}

template <bool kStaged>
__global__ __launch_bounds__(kKsumWaves * 64) void ewald_force_batch_kernel(const BatchRep* reps)
{
    const BatchRep& r = reps[blockIdx.y];
    if (r.recip != TCP_RECIP_EWALD || (int)blockIdx.x >= (r.ra.n + 63) / 64) return;
    extern __shared__ double kf_dyn[];
    ewald_force_body<kStaged>(r.ra, r.kmax, r.sk, blockIdx.x, kf_dyn);
}

// ---------------------------------------- batched SPME (tcp_batch_*) ---
//
// The bodies of the binned single-context kernels over a replica table:
// blockIdx.y selects the replica, blocks past that replica's own count
// return at once, so each replica's spreading, FFT and gather are its
// standalone ones, bit for bit.  Every replica has a brick-binned grid and
// the batch's spline order.

__device__ inline int brick_total(const RecipArgs& a)
{
    return a.nbrick[0] * a.nbrick[1] * a.nbrick[2];
}

__global__ __launch_bounds__(256) void bin_clear_batch_kernel(const BatchRep* reps)
{
    const RecipArgs& a = reps[blockIdx.y].ra;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < brick_total(a)) a.brick_count[i] = 0;
}

template <int P>
__global__ __launch_bounds__(256) void bin_count_batch_kernel(const BatchRep* reps)
{
    bin_count_body<P>(reps[blockIdx.y].ra, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(1024) void bin_scan_batch_kernel(const BatchRep* reps)
{
    const RecipArgs& a = reps[blockIdx.x].ra;
    bin_scan_body(a, brick_total(a));
}

__global__ __launch_bounds__(256) void bin_scatter_batch_kernel(const BatchRep* reps)
{
    bin_scatter_body(reps[blockIdx.y].ra, blockIdx.x * blockDim.x + threadIdx.x);
}

template <int P>
__global__ __launch_bounds__(256) void spread_brick_batch_kernel(const BatchRep* reps)
{
    const RecipArgs& a = reps[blockIdx.y].ra;
    if ((int)blockIdx.x >= brick_total(a)) return;
    spread_brick_body<P>(a, blockIdx.x);
}

template <int P>
__global__ __launch_bounds__(256) void grid_from_tiles_batch_kernel(const BatchRep* reps)
{
    grid_from_tiles_body<P>(reps[blockIdx.y].ra, blockIdx.x * (size_t)blockDim.x + threadIdx.x);
}

template <int kMode>
__global__ __launch_bounds__(kFftThreads) void fft_pass_batch_kernel(const FftPass* passes)
{
    const FftPass& f = passes[blockIdx.y];
    if ((int)blockIdx.x >= (f.lines + f.lines_per_block - 1) / f.lines_per_block) return;
    fft_pass_body<kMode>(f, blockIdx.x);
}

template <int P>
__global__ __launch_bounds__(64) void gather_brick_batch_kernel(const BatchRep* reps)
{
    const RecipArgs& a = reps[blockIdx.y].ra;
    if ((int)blockIdx.x >= brick_total(a)) return;
    gather_brick_body<P>(a, blockIdx.x);
}

// the five passes of launch_spme_convolve, in its order
constexpr int kPassMode[kSpmeBatchPasses] = {kFftFwdX, kFftY, kFftZConv, kFftY, kFftInvX};
constexpr int kPassSign[kSpmeBatchPasses] = {-1, -1, -1, +1, +1};

template <int P>
void spme_batch_order(const BatchRep* reps, const SpmeBatchPlan& pl, hipStream_t s)
{
    const int nr = pl.nrep;
    const dim3 gp((pl.maxn + 255) / 256, nr), b256(256);
    const FftPass* passes = (const FftPass*)pl.passes;
    hipLaunchKernelGGL(bin_clear_batch_kernel, dim3((pl.maxnb + 255) / 256, nr), b256, 0, s, reps);
    hipLaunchKernelGGL(bin_count_batch_kernel<P>, gp, b256, 0, s, reps);
    hipLaunchKernelGGL(bin_scan_batch_kernel, dim3(nr), dim3(1024), 0, s, reps);
    hipLaunchKernelGGL(bin_scatter_batch_kernel, gp, b256, 0, s, reps);
    hipLaunchKernelGGL(spread_brick_batch_kernel<P>, dim3(pl.maxnb, nr), b256, 0, s, reps);
    hipLaunchKernelGGL(grid_from_tiles_batch_kernel<P>, dim3(pl.maxgblk, nr), b256, 0, s, reps);
    const dim3 bf(kFftThreads);
    hipLaunchKernelGGL(fft_pass_batch_kernel<kFftFwdX>, dim3(pl.pass_blocks[0], nr), bf, 0, s, passes);
    hipLaunchKernelGGL(fft_pass_batch_kernel<kFftY>, dim3(pl.pass_blocks[1], nr), bf, 0, s, passes + nr);
    hipLaunchKernelGGL(fft_pass_batch_kernel<kFftZConv>, dim3(pl.pass_blocks[2], nr), bf, 0, s,
                       passes + 2 * nr);
    hipLaunchKernelGGL(fft_pass_batch_kernel<kFftY>, dim3(pl.pass_blocks[3], nr), bf, 0, s,
                       passes + 3 * nr);
    hipLaunchKernelGGL(fft_pass_batch_kernel<kFftInvX>, dim3(pl.pass_blocks[4], nr), bf, 0, s,
                       passes + 4 * nr);
    hipLaunchKernelGGL(gather_brick_batch_kernel<P>, dim3(pl.maxnb, nr), dim3(64), 0, s, reps);
}

}  // namespace

int fft_epart_blocks(const int* K)
{
    const int lines = (K[0] / 2 + 1) * K[1], per = kFftTile / K[2];
    return (lines + per - 1) / per;
}

void launch_fill(double* p, size_t n, double v, hipStream_t s)
{
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
}

template <int P>
void spread_binned(const RecipArgs& a, hipStream_t s)
{
    const int nb = a.nbrick[0] * a.nbrick[1] * a.nbrick[2];
    const size_t G = (size_t)a.K[0] * a.K[1] * a.K[2];
    hipLaunchKernelGGL(fill_int_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, a.brick_count, nb);
    hipLaunchKernelGGL(bin_count_kernel<P>, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, s, a, nb);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(spread_brick_kernel<P>, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(grid_from_tiles_kernel<P>, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, a);
}

void launch_spme_spread(const RecipArgs& a, hipStream_t s)
{
    if (a.nbrick[0]) {
        switch (a.order) {
            case 4: spread_binned<4>(a, s); return;
            case 6: spread_binned<6>(a, s); return;
            case 8: spread_binned<8>(a, s); return;
            case 10: spread_binned<10>(a, s); return;
            default: spread_binned<12>(a, s); return;
        }
    }
    const size_t G = (size_t)a.K[0] * a.K[1] * a.K[2];
    launch_fill(a.Q, G, 0.0, s);
    dim3 grid((a.n + 255) / 256), blk(256);
    switch (a.order) {
        case 4: hipLaunchKernelGGL(spread_kernel<4>, grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL(spread_kernel<6>, grid, blk, 0, s, a); break;
        case 8: hipLaunchKernelGGL(spread_kernel<8>, grid, blk, 0, s, a); break;
        case 10: hipLaunchKernelGGL(spread_kernel<10>, grid, blk, 0, s, a); break;
        default: hipLaunchKernelGGL(spread_kernel<12>, grid, blk, 0, s, a); break;
    }
}

void launch_spme_gather(const RecipArgs& a, hipStream_t s)
{
    if (a.nbrick[0]) {
        const dim3 g(a.nbrick[0] * a.nbrick[1] * a.nbrick[2]), b(64);
        switch (a.order) {
            case 4: hipLaunchKernelGGL(gather_brick_kernel<4>, g, b, 0, s, a); break;
            case 6: hipLaunchKernelGGL(gather_brick_kernel<6>, g, b, 0, s, a); break;
            case 8: hipLaunchKernelGGL(gather_brick_kernel<8>, g, b, 0, s, a); break;
            case 10: hipLaunchKernelGGL(gather_brick_kernel<10>, g, b, 0, s, a); break;
            default: hipLaunchKernelGGL(gather_brick_kernel<12>, g, b, 0, s, a); break;
        }
        return;
    }
    dim3 grid((a.n + 255) / 256), blk(256);
    switch (a.order) {
        case 4: hipLaunchKernelGGL(gather_kernel<4>, grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL(gather_kernel<6>, grid, blk, 0, s, a); break;
        case 8: hipLaunchKernelGGL(gather_kernel<8>, grid, blk, 0, s, a); break;
        case 10: hipLaunchKernelGGL(gather_kernel<10>, grid, blk, 0, s, a); break;
        default: hipLaunchKernelGGL(gather_kernel<12>, grid, blk, 0, s, a); break;
    }
}

void launch_spme_convolve(const RecipArgs& a, hipStream_t s)
{
    run_pass<kFftFwdX>(make_pass(a, kFftFwdX, -1), s);
    run_pass<kFftY>(make_pass(a, kFftY, -1), s);
    run_pass<kFftZConv>(make_pass(a, kFftZConv, -1), s);
    run_pass<kFftY>(make_pass(a, kFftY, +1), s);
    run_pass<kFftInvX>(make_pass(a, kFftInvX, +1), s);
}

int ewald_chunk(int kmax)
{
    // LDS phase table 3 (kmax + 1) chunk double2 plus ux[chunk] within 64 KiB
    int c = 128;
    while (c > 16 && (size_t)(48 * (kmax + 1) + 16) * c > 65536) c /= 2;
    return c;
}

// dynamic LDS of the partial kernels: the phase table, or the parked group
// partials of the row form when they need more
static size_t ewald_part_lds(int kmax, int chunk)
{
    const size_t table = (size_t)(48 * (kmax + 1) + 16) * chunk;
    const size_t parked = (size_t)(ewald_sk_groups(kmax) - 1) * ewald_half(kmax) * sizeof(double2);
    return table > parked ? table : parked;
}

int ewald_half(int kmax)
{
    const int nk = 2 * kmax + 1;
    return (nk * nk * nk - 1) / 2;
}

void launch_ewald_kspace(const RecipArgs& a, int kmax, double alpha, double* sk, hipStream_t s)
{
    const int chunk = ewald_chunk(kmax), nchunk = (a.n + chunk - 1) / chunk;
    const size_t lds = ewald_part_lds(kmax, chunk);
    hipLaunchKernelGGL(ewald_sk_part_kernel, dim3(nchunk), dim3(256), lds, s, a, kmax, chunk);
    hipLaunchKernelGGL(ewald_sk_reduce_kernel, dim3((ewald_half(kmax) + 255) / 256), dim3(256), 0,
                       s, a, kmax, alpha, nchunk, sk);
    const dim3 fg((a.n + 63) / 64), fb(kKsumWaves * 64);
    const bool staged = ewald_half(kmax) <= kSkLds;
    const size_t flds = sizeof(double) * ewald_force_lds(kmax, staged);
    if (staged)
        hipLaunchKernelGGL(ewald_force_kernel<true>, fg, fb, flds, s, a, kmax, (const double*)sk);
    else
        hipLaunchKernelGGL(ewald_force_kernel<false>, fg, fb, flds, s, a, kmax, (const double*)sk);
}

void launch_ewald_batch(const BatchRep* reps, int nrep, int kmax, int maxn, hipStream_t s)
{
    const int chunk = ewald_chunk(kmax);
    const size_t lds = ewald_part_lds(kmax, chunk);
    hipLaunchKernelGGL(ewald_sk_part_batch_kernel, dim3((maxn + chunk - 1) / chunk, nrep), dim3(256),
                       lds, s, reps, chunk);
    hipLaunchKernelGGL(ewald_sk_reduce_batch_kernel, dim3((ewald_half(kmax) + 255) / 256, nrep),
                       dim3(256), 0, s, reps, chunk);
    const dim3 fg((maxn + 63) / 64, nrep), fb(kKsumWaves * 64);
    const bool staged = ewald_half(kmax) <= kSkLds;
    const size_t flds = sizeof(double) * ewald_force_lds(kmax, staged);
    if (staged)
        hipLaunchKernelGGL(ewald_force_batch_kernel<true>, fg, fb, flds, s, reps);
    else
        hipLaunchKernelGGL(ewald_force_batch_kernel<false>, fg, fb, flds, s, reps);
    launch_um_batch(reps, nrep, s);
}


int spme_batch_plan(const RecipArgs* ra, int nrep, SpmeBatchPlan* pl)
{
    *pl = SpmeBatchPlan{};
    pl->nrep = nrep;
    pl->order = ra[0].order;
    std::vector<FftPass> tab((size_t)kSpmeBatchPasses * nrep);
    for (int r = 0; r < nrep; ++r) {
        const RecipArgs& a = ra[r];
        const int nb = a.nbrick[0] * a.nbrick[1] * a.nbrick[2];
        const size_t G = (size_t)a.K[0] * a.K[1] * a.K[2];
        pl->maxn = std::max(pl->maxn, a.n);
        pl->maxnb = std::max(pl->maxnb, nb);
        pl->maxgblk = std::max(pl->maxgblk, (int)((G + 255) / 256));
        for (int k = 0; k < kSpmeBatchPasses; ++k) {
            const FftPass f = make_pass(a, kPassMode[k], kPassSign[k]);
            tab[(size_t)k * nrep + r] = f;
            pl->pass_blocks[k] =
                std::max(pl->pass_blocks[k], (f.lines + f.lines_per_block - 1) / f.lines_per_block);
        }
    }
    const size_t bytes = tab.size() * sizeof(FftPass);
    if (hipMalloc(&pl->passes, bytes) != hipSuccess) {
        pl->passes = nullptr;
        return TCP_ERR_HIP;
    }
    if (hipMemcpy(pl->passes, tab.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return TCP_ERR_HIP;
    return 0;
}

void spme_batch_release(SpmeBatchPlan* pl)
{
    if (pl->passes) (void)hipFree(pl->passes);
    pl->passes = nullptr;
}

void launch_spme_batch(const BatchRep* reps, const SpmeBatchPlan& pl, hipStream_t s)
{
    switch (pl.order) {
        case 4: spme_batch_order<4>(reps, pl, s); break;
        case 6: spme_batch_order<6>(reps, pl, s); break;
        case 8: spme_batch_order<8>(reps, pl, s); break;
        case 10: spme_batch_order<10>(reps, pl, s); break;
        default: spme_batch_order<12>(reps, pl, s); break;
    }
    launch_um_batch(reps, pl.nrep, s);
}

}  // namespace tcp
